// Controlled Y gate (reference include/qclab/qgates/CY.hpp).
#pragma once

#include <memory>

#include "qclab/qgates/QControlledGate2.hpp"
#include "qclab/qgates/PauliY.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class CY : public QControlledGate2<T> {
 public:
  using gate_type = QGate1<T>;

  CY() : QControlledGate2<T>(0), gate_(std::make_unique<PauliY<T>>(1)) {}
  CY(const int control, const int target, const int controlState = 1)
      : QControlledGate2<T>(control, controlState), gate_(std::make_unique<PauliY<T>>(target)) {
    assert(control >= 0 && target >= 0 && control != target);
  }
  CY(const CY<T>& o)
      : QControlledGate2<T>(o.control(), o.controlState()), gate_(std::make_unique<PauliY<T>>(o.target())) {}
  bool fixed() const override { return gate_->fixed(); }
  int qubit() const override { return std::min(this->control_, gate_->qubit()); }
  int target() const override { return gate_->qubit(); }
  const gate_type* gate() const override { return gate_.get(); }
  void setTarget(const int target) override {
    assert(target >= 0 && this->control() != target);
    gate_->setQubit(target);
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    return this->qasm_controlled(stream, offset, [&](int c, int t) { return qasmCY(c, t); });
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    qsv_gate g = this->record(QSV_GATE_CY, op, offset);

    out.push_back(g);
  }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const CY<T>*>(&other);
    return p && this->same_layout(*p);
  }
  std::unique_ptr<PauliY<T>> gate_;
};

}  // namespace qgates
}  // namespace qclab
