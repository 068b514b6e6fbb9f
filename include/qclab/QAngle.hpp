// QAngle: an angle stored as (cos, sin) (reference include/qclab/QAngle.hpp).
#pragma once

#include <cmath>

namespace qclab {

template <typename T>
class QAngle {
 public:
  using value_type = T;
  QAngle() : cos_(1), sin_(0) {}
  QAngle(const T theta) : cos_(std::cos(theta)), sin_(std::sin(theta)) {}
  QAngle(const T cos, const T sin) : cos_(cos), sin_(sin) {}

  const T theta() const { return std::atan2(sin_, cos_); }
  const T cos() const { return cos_; }
  const T sin() const { return sin_; }

  void update(const QAngle<T>& a) { cos_ = a.cos(); sin_ = a.sin(); }
  void update(const T theta) { cos_ = std::cos(theta); sin_ = std::sin(theta); }
  void update(const T cos, const T sin) { cos_ = cos; sin_ = sin; }

  bool operator==(const QAngle<T>& o) const { return o.cos() == cos_ && o.sin() == sin_; }
  bool operator!=(const QAngle<T>& o) const { return !(*this == o); }

  // angle addition / subtraction through the sum formulas
  QAngle<T>& operator+=(const QAngle<T>& r) {
    const T c = cos_ * r.cos() - sin_ * r.sin(), s = sin_ * r.cos() + cos_ * r.sin();
    update(c, s);
    return *this;
  }
  QAngle<T>& operator-=(const QAngle<T>& r) {
    const T c = cos_ * r.cos() + sin_ * r.sin(), s = sin_ * r.cos() - cos_ * r.sin();
    update(c, s);
    return *this;
  }
  friend QAngle<T> operator+(QAngle<T> l, const QAngle<T>& r) { return l += r; }
  friend QAngle<T> operator-(QAngle<T> l, const QAngle<T>& r) { return l -= r; }
  QAngle<T> operator-() const { return QAngle<T>(cos_, -sin_); }

 protected:
  T cos_;
  T sin_;
};

}  // namespace qclab
