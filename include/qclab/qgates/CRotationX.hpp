// Controlled X rotation (reference include/qclab/qgates/CRotationX.hpp).
#pragma once

#include <memory>

#include "qclab/qgates/QControlledGate2.hpp"
#include "qclab/qgates/RotationX.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class CRotationX : public QControlledGate2<T> {
 public:
  using gate_type = QGate1<T>;
  using real_type = real_t<T>;
  using rotation_type = typename RotationX<T>::rotation_type;

  CRotationX() : QControlledGate2<T>(0), gate_(std::make_unique<RotationX<T>>(1, real_type(0))) {}
  CRotationX(const int control, const int target, const rotation_type& a, const int controlState = 1)
      : QControlledGate2<T>(control, controlState), gate_(std::make_unique<RotationX<T>>(target, a)) {
    assert(control >= 0 && target >= 0 && control != target);
  }
  CRotationX(const int control, const int target, const real_type theta, const int controlState = 1)
      : QControlledGate2<T>(control, controlState), gate_(std::make_unique<RotationX<T>>(target, theta)) {
    assert(control >= 0 && target >= 0 && control != target);
  }
  CRotationX(const int control, const int target, const real_type cos, const real_type sin,
     const int controlState = 1)
      : QControlledGate2<T>(control, controlState), gate_(std::make_unique<RotationX<T>>(target, cos, sin)) {
    assert(control >= 0 && target >= 0 && control != target);
  }
  CRotationX(const CRotationX<T>& o) : QControlledGate2<T>(o.control(), o.controlState()), gate_(std::make_unique<RotationX<T>>(*o.gate_)) {}

  const rotation_type& rotation() const { return gate_->rotation(); }
  real_type theta() const { return gate_->theta(); }
  real_type cos() const { return gate_->cos(); }
  real_type sin() const { return gate_->sin(); }
  void update(const rotation_type& a) { gate_->update(a); }
  void update(const real_type theta) { gate_->update(theta); }
  void update(const real_type cos, const real_type sin) { gate_->update(cos, sin); }
  bool fixed() const override { return gate_->fixed(); }
  int qubit() const override { return std::min(this->control_, gate_->qubit()); }
  int target() const override { return gate_->qubit(); }
  const gate_type* gate() const override { return gate_.get(); }
  void setTarget(const int target) override {
    assert(target >= 0 && this->control() != target);
    gate_->setQubit(target);
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    return this->qasm_controlled(stream, offset, [&](int c, int t) { return qasmCRx(c, t, theta()); });
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    qsv_gate g = this->record(QSV_GATE_CRX, op, offset);
    g.theta = double(theta());
    backend::set_cs(g, double(cos()), double(sin()));
    out.push_back(g);
  }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const CRotationX<T>*>(&other);
    return p && p->rotation() == rotation() && this->same_layout(*p);
  }
  std::unique_ptr<RotationX<T>> gate_;
};

}  // namespace qgates
}  // namespace qclab
