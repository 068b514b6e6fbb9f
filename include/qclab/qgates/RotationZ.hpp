// Rotation about Z: diag(c - is, c + is) with c, s = cos, sin(theta/2) (reference include/qclab/qgates/RotationZ.hpp).
#pragma once

#include "qclab/qgates/QRotationGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class RotationZ : public QRotationGate1<T> {
 public:
  using real_type = real_t<T>;
  using rotation_type = QRotation<real_type>;

  RotationZ() : QRotationGate1<T>() {}
  RotationZ(const rotation_type& rot) : QRotationGate1<T>(0, rot) {}
  RotationZ(const real_type theta) : QRotationGate1<T>(0, theta) {}
  RotationZ(const real_type cos, const real_type sin) : QRotationGate1<T>(0, cos, sin) {}
  RotationZ(const int qubit, const rotation_type& rot, const bool fixed = false)
      : QRotationGate1<T>(qubit, rot, fixed) {}
  RotationZ(const int qubit, const real_type theta, const bool fixed = false)
      : QRotationGate1<T>(qubit, theta, fixed) {}
  RotationZ(const int qubit, const real_type cos, const real_type sin, const bool fixed = false)
      : QRotationGate1<T>(qubit, cos, sin, fixed) {}

  dense::SquareMatrix<T> matrix() const override {
    const real_type c = this->cos(), s = this->sin();
    return dense::SquareMatrix<T>(T(c, -s), T(0), T(0), T(c, s));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmRz(this->qubit_ + offset, this->theta());
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->rotation_record(QSV_GATE_RZ, op, offset));
  }

  RotationZ<T>& operator*=(const RotationZ<T>& rhs) {
    assert(this->qubit() == rhs.qubit());
    this->rotation_ *= rhs.rotation();
    return *this;
  }
  RotationZ<T>& operator/=(const RotationZ<T>& rhs) {
    assert(this->qubit() == rhs.qubit());
    this->rotation_ /= rhs.rotation();
    return *this;
  }
  friend RotationZ<T> operator*(RotationZ<T> lhs, const RotationZ<T>& rhs) { return lhs *= rhs; }
  friend RotationZ<T> operator/(RotationZ<T> lhs, const RotationZ<T>& rhs) { return lhs /= rhs; }
  /// Inverse rotation; like the reference it is built on qubit 0, not fixed.
  RotationZ<T> inv() const { return RotationZ<T>(this->rotation_.inv()); }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const RotationZ<T>*>(&other);
    return p && p->rotation() == this->rotation();
  }
};

}  // namespace qgates
}  // namespace qclab
