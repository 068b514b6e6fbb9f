// Phase gate diag(1, e^{i theta}) (reference include/qclab/qgates/Phase.hpp).
// The angle is stored as (cos theta, sin theta).
#pragma once

#include "qclab/QAdjustable.hpp"
#include "qclab/QAngle.hpp"
#include "qclab/qgates/QGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class Phase : public QGate1<T>, public QAdjustable {
 public:
  using real_type = real_t<T>;
  using angle_type = QAngle<real_type>;

  Phase() : QGate1<T>(0), QAdjustable(false), angle_() {}
  Phase(const angle_type& angle) : QGate1<T>(0), QAdjustable(false), angle_(angle) {}
  Phase(const real_type theta) : QGate1<T>(0), QAdjustable(false), angle_(theta) {}
  Phase(const real_type cos, const real_type sin) : QGate1<T>(0), QAdjustable(false), angle_(cos, sin) {}
  Phase(const int qubit, const angle_type& angle, const bool fixed = false)
      : QGate1<T>(qubit), QAdjustable(fixed), angle_(angle) {}
  Phase(const int qubit, const real_type theta, const bool fixed = false)
      : QGate1<T>(qubit), QAdjustable(fixed), angle_(theta) {}
  Phase(const int qubit, const real_type cos, const real_type sin, const bool fixed = false)
      : QGate1<T>(qubit), QAdjustable(fixed), angle_(cos, sin) {}

  bool fixed() const override { return QAdjustable::fixed(); }
  dense::SquareMatrix<T> matrix() const override {
    return dense::SquareMatrix<T>(T(1), T(0), T(0), T(cos(), sin()));
  }
  /// Written as rz(theta), as the reference does (equal up to global phase).
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmRz(this->qubit_ + offset, theta());
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    qsv_gate g = this->record(QSV_GATE_P, op, offset);
    g.theta = double(theta());
    backend::set_cs(g, double(cos()), double(sin()));
    out.push_back(g);
  }

  const angle_type& angle() const { return angle_; }
  real_type theta() const { return angle_.theta(); }
  real_type cos() const { return angle_.cos(); }
  real_type sin() const { return angle_.sin(); }
  void update(const angle_type& angle) { assert(!fixed()); angle_.update(angle); }
  void update(const real_type theta) { assert(!fixed()); angle_.update(theta); }
  void update(const real_type cos, const real_type sin) { assert(!fixed()); angle_.update(cos, sin); }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const Phase<T>*>(&other);
    return p && p->angle() == angle();
  }
  angle_type angle_;
};

}  // namespace qgates
}  // namespace qclab
