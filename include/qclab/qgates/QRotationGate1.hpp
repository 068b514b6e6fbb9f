// Base class of 1-qubit rotation gates
// (reference include/qclab/qgates/QRotationGate1.hpp).
#pragma once

#include "qclab/QAdjustable.hpp"
#include "qclab/QRotation.hpp"
#include "qclab/qgates/QGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class QRotationGate1 : public QGate1<T>, public QAdjustable {
 public:
  using real_type = real_t<T>;
  using rotation_type = QRotation<real_type>;

  QRotationGate1() : QGate1<T>(0), QAdjustable(false), rotation_() {}
  QRotationGate1(const int qubit, const rotation_type& rot, const bool fixed = false)
      : QGate1<T>(qubit), QAdjustable(fixed), rotation_(rot) {}
  QRotationGate1(const int qubit, const real_type theta, const bool fixed = false)
      : QGate1<T>(qubit), QAdjustable(fixed), rotation_(theta) {}
  QRotationGate1(const int qubit, const real_type cos, const real_type sin, const bool fixed = false)
      : QGate1<T>(qubit), QAdjustable(fixed), rotation_(cos, sin) {}

  bool fixed() const override { return QAdjustable::fixed(); }

  const rotation_type& rotation() const { return rotation_; }
  real_type theta() const { return rotation_.theta(); }
  real_type cos() const { return rotation_.cos(); }
  real_type sin() const { return rotation_.sin(); }
  void update(const rotation_type& rot) { assert(!fixed()); rotation_ = rot; }
  void update(const real_type theta) { assert(!fixed()); rotation_.update(theta); }
  void update(const real_type cos, const real_type sin) { assert(!fixed()); rotation_.update(cos, sin); }

 protected:
  // the stored (cos, sin) of theta/2 travel verbatim
  qsv_gate rotation_record(int kind, Op op, int offset) const {
    qsv_gate g = this->record(kind, op, offset);
    g.theta = double(theta());
    backend::set_cs(g, double(cos()), double(sin()));
    return g;
  }
  rotation_type rotation_;
};

}  // namespace qgates
}  // namespace qclab
