// Rotation about Y: [c, -s; s, c] with c, s = cos, sin(theta/2) (reference include/qclab/qgates/RotationY.hpp).
#pragma once

#include "qclab/qgates/QRotationGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class RotationY : public QRotationGate1<T> {
 public:
  using real_type = real_t<T>;
  using rotation_type = QRotation<real_type>;

  RotationY() : QRotationGate1<T>() {}
  RotationY(const rotation_type& rot) : QRotationGate1<T>(0, rot) {}
  RotationY(const real_type theta) : QRotationGate1<T>(0, theta) {}
  RotationY(const real_type cos, const real_type sin) : QRotationGate1<T>(0, cos, sin) {}
  RotationY(const int qubit, const rotation_type& rot, const bool fixed = false)
      : QRotationGate1<T>(qubit, rot, fixed) {}
  RotationY(const int qubit, const real_type theta, const bool fixed = false)
      : QRotationGate1<T>(qubit, theta, fixed) {}
  RotationY(const int qubit, const real_type cos, const real_type sin, const bool fixed = false)
      : QRotationGate1<T>(qubit, cos, sin, fixed) {}

  dense::SquareMatrix<T> matrix() const override {
    const real_type c = this->cos(), s = this->sin();
    return dense::SquareMatrix<T>(T(c), T(-s), T(s), T(c));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmRy(this->qubit_ + offset, this->theta());
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->rotation_record(QSV_GATE_RY, op, offset));
  }

  RotationY<T>& operator*=(const RotationY<T>& rhs) {
    assert(this->qubit() == rhs.qubit());
    this->rotation_ *= rhs.rotation();
    return *this;
  }
  RotationY<T>& operator/=(const RotationY<T>& rhs) {
    assert(this->qubit() == rhs.qubit());
    this->rotation_ /= rhs.rotation();
    return *this;
  }
  friend RotationY<T> operator*(RotationY<T> lhs, const RotationY<T>& rhs) { return lhs *= rhs; }
  friend RotationY<T> operator/(RotationY<T> lhs, const RotationY<T>& rhs) { return lhs /= rhs; }
  /// Inverse rotation; like the reference it is built on qubit 0, not fixed.
  RotationY<T> inv() const { return RotationY<T>(this->rotation_.inv()); }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const RotationY<T>*>(&other);
    return p && p->rotation() == this->rotation();
  }
};

}  // namespace qgates
}  // namespace qclab
