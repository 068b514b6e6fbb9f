// YY rotation exp(-i theta/2 Y(x)Y) (reference include/qclab/qgates/RotationYY.hpp).
#pragma once

#include "qclab/qgates/QRotationGate2.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class RotationYY : public QRotationGate2<T> {
 public:
  using real_type = real_t<T>;
  using rotation_type = QRotation<real_type>;

  RotationYY() : QRotationGate2<T>() {}
  RotationYY(const rotation_type& rot) : QRotationGate2<T>(0, 1, rot) {}
  RotationYY(const real_type theta) : QRotationGate2<T>(0, 1, theta) {}
  RotationYY(const real_type cos, const real_type sin) : QRotationGate2<T>(0, 1, cos, sin) {}
  RotationYY(const int* qubits, const rotation_type& rot, const bool fixed = false)
      : QRotationGate2<T>(qubits, rot, fixed) {}
  RotationYY(const int* qubits, const real_type theta, const bool fixed = false)
      : QRotationGate2<T>(qubits, theta, fixed) {}
  RotationYY(const int* qubits, const real_type cos, const real_type sin, const bool fixed = false)
      : QRotationGate2<T>(qubits, cos, sin, fixed) {}
  RotationYY(const int qubit0, const int qubit1, const rotation_type& rot, const bool fixed = false)
      : QRotationGate2<T>(qubit0, qubit1, rot, fixed) {}
  RotationYY(const int qubit0, const int qubit1, const real_type theta, const bool fixed = false)
      : QRotationGate2<T>(qubit0, qubit1, theta, fixed) {}
  RotationYY(const int qubit0, const int qubit1, const real_type cos, const real_type sin,
     const bool fixed = false)
      : QRotationGate2<T>(qubit0, qubit1, cos, sin, fixed) {}

  dense::SquareMatrix<T> matrix() const override {
    const real_type c = this->cos(), s = this->sin();
    const T o(0);
    return dense::SquareMatrix<T>(T(c), o, o, T(0, s),  o, T(c), T(0, -s), o,  o, T(0, -s), T(c), o,  T(0, s), o, o, T(c));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmRyy(this->qubits_[0] + offset, this->qubits_[1] + offset, this->theta());
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->rotation_record(QSV_GATE_RYY, op, offset));
  }

  RotationYY<T>& operator*=(const RotationYY<T>& rhs) {
    assert(this->qubits() == rhs.qubits());
    this->rotation_ *= rhs.rotation();
    return *this;
  }
  RotationYY<T>& operator/=(const RotationYY<T>& rhs) {
    assert(this->qubits() == rhs.qubits());
    this->rotation_ /= rhs.rotation();
    return *this;
  }
  friend RotationYY<T> operator*(RotationYY<T> lhs, const RotationYY<T>& rhs) { return lhs *= rhs; }
  friend RotationYY<T> operator/(RotationYY<T> lhs, const RotationYY<T>& rhs) { return lhs /= rhs; }
  /// Inverse rotation; like the reference it is built on qubits 0 and 1.
  RotationYY<T> inv() const { return RotationYY<T>(this->rotation_.inv()); }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const RotationYY<T>*>(&other);
    return p && p->rotation() == this->rotation();
  }
};

}  // namespace qgates
}  // namespace qclab
