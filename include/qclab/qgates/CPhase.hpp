// Controlled phase gate (reference include/qclab/qgates/CPhase.hpp).
#pragma once

#include <memory>

#include "qclab/qgates/QControlledGate2.hpp"
#include "qclab/qgates/Phase.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class CPhase : public QControlledGate2<T> {
 public:
  using gate_type = QGate1<T>;
  using real_type = real_t<T>;
  using angle_type = typename Phase<T>::angle_type;

  CPhase() : QControlledGate2<T>(0), gate_(std::make_unique<Phase<T>>(1, real_type(0))) {}
  CPhase(const int control, const int target, const angle_type& a, const int controlState = 1)
      : QControlledGate2<T>(control, controlState), gate_(std::make_unique<Phase<T>>(target, a)) {
    assert(control >= 0 && target >= 0 && control != target);
  }
  CPhase(const int control, const int target, const real_type theta, const int controlState = 1)
      : QControlledGate2<T>(control, controlState), gate_(std::make_unique<Phase<T>>(target, theta)) {
    assert(control >= 0 && target >= 0 && control != target);
  }
  CPhase(const int control, const int target, const real_type cos, const real_type sin,
     const int controlState = 1)
      : QControlledGate2<T>(control, controlState), gate_(std::make_unique<Phase<T>>(target, cos, sin)) {
    assert(control >= 0 && target >= 0 && control != target);
  }
  CPhase(const CPhase<T>& o) : QControlledGate2<T>(o.control(), o.controlState()), gate_(std::make_unique<Phase<T>>(*o.gate_)) {}

  const angle_type& angle() const { return gate_->angle(); }
  real_type theta() const { return gate_->theta(); }
  real_type cos() const { return gate_->cos(); }
  real_type sin() const { return gate_->sin(); }
  void update(const angle_type& a) { gate_->update(a); }
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
    return this->qasm_controlled(stream, offset, [&](int c, int t) { return qasmCPhase(c, t, theta()); });
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    qsv_gate g = this->record(QSV_GATE_CP, op, offset);
    g.theta = double(theta());
    backend::set_cs(g, double(cos()), double(sin()));
    out.push_back(g);
  }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const CPhase<T>*>(&other);
    return p && p->angle() == angle() && this->same_layout(*p);
  }
  std::unique_ptr<Phase<T>> gate_;
};

}  // namespace qgates
}  // namespace qclab
