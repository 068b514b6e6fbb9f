// QRotation: a rotation parameter theta stored as the QAngle theta/2
// (reference include/qclab/QRotation.hpp).
#pragma once

#include "qclab/QAngle.hpp"

namespace qclab {

template <typename T>
class QRotation {
 public:
  using angle_type = QAngle<T>;
  QRotation() : angle_(T(0)) {}
  QRotation(const angle_type& angle) : angle_(angle) {}
  QRotation(const T theta) : angle_(theta / 2) {}
  QRotation(const T cos, const T sin) : angle_(cos, sin) {}

  const angle_type& angle() const { return angle_; }
  T theta() const { return 2 * angle_.theta(); }
  T cos() const { return angle_.cos(); }
  T sin() const { return angle_.sin(); }

  void update(const angle_type& a) { angle_.update(a); }
  void update(const T theta) { angle_.update(theta / 2); }
  void update(const T cos, const T sin) { angle_.update(cos, sin); }

  bool operator==(const QRotation<T>& o) const { return o.angle() == angle_; }
  bool operator!=(const QRotation<T>& o) const { return o.angle() != angle_; }
  QRotation<T>& operator*=(const QRotation<T>& r) { update(angle_ + r.angle()); return *this; }
  QRotation<T>& operator/=(const QRotation<T>& r) { update(angle_ - r.angle()); return *this; }
  friend QRotation<T> operator*(QRotation<T> l, const QRotation<T>& r) { return l *= r; }
  friend QRotation<T> operator/(QRotation<T> l, const QRotation<T>& r) { return l /= r; }
  QRotation<T> inv() const { return QRotation<T>(-angle_); }

 protected:
  angle_type angle_;
};

}  // namespace qclab
