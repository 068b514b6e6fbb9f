// Rotation about X: [c, -is; -is, c] with c, s = cos, sin(theta/2) (reference include/qclab/qgates/RotationX.hpp).
#pragma once

#include "qclab/qgates/QRotationGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class RotationX : public QRotationGate1<T> {
 public:
  using real_type = real_t<T>;
  using rotation_type = QRotation<real_type>;

  RotationX() : QRotationGate1<T>() {}
  RotationX(const rotation_type& rot) : QRotationGate1<T>(0, rot) {}
  RotationX(const real_type theta) : QRotationGate1<T>(0, theta) {}
  RotationX(const real_type cos, const real_type sin) : QRotationGate1<T>(0, cos, sin) {}
  RotationX(const int qubit, const rotation_type& rot, const bool fixed = false)
      : QRotationGate1<T>(qubit, rot, fixed) {}
  RotationX(const int qubit, const real_type theta, const bool fixed = false)
      : QRotationGate1<T>(qubit, theta, fixed) {}
  RotationX(const int qubit, const real_type cos, const real_type sin, const bool fixed = false)
      : QRotationGate1<T>(qubit, cos, sin, fixed) {}

  dense::SquareMatrix<T> matrix() const override {
    const real_type c = this->cos(), s = this->sin();
    return dense::SquareMatrix<T>(T(c), T(0, -s), T(0, -s), T(c));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmRx(this->qubit_ + offset, this->theta());
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->rotation_record(QSV_GATE_RX, op, offset));
  }

  RotationX<T>& operator*=(const RotationX<T>& rhs) {
    assert(this->qubit() == rhs.qubit());
    this->rotation_ *= rhs.rotation();
    return *this;
  }
  RotationX<T>& operator/=(const RotationX<T>& rhs) {
    assert(this->qubit() == rhs.qubit());
    this->rotation_ /= rhs.rotation();
    return *this;
  }
  friend RotationX<T> operator*(RotationX<T> lhs, const RotationX<T>& rhs) { return lhs *= rhs; }
  friend RotationX<T> operator/(RotationX<T> lhs, const RotationX<T>& rhs) { return lhs /= rhs; }
  /// Inverse rotation; like the reference it is built on qubit 0, not fixed.
  RotationX<T> inv() const { return RotationX<T>(this->rotation_.inv()); }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const RotationX<T>*>(&other);
    return p && p->rotation() == this->rotation();
  }
};

}  // namespace qgates
}  // namespace qclab
