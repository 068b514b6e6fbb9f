// XX rotation exp(-i theta/2 X(x)X) (reference include/qclab/qgates/RotationXX.hpp).
#pragma once

#include "qclab/qgates/QRotationGate2.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class RotationXX : public QRotationGate2<T> {
 public:
  using real_type = real_t<T>;
  using rotation_type = QRotation<real_type>;

  RotationXX() : QRotationGate2<T>() {}
  RotationXX(const rotation_type& rot) : QRotationGate2<T>(0, 1, rot) {}
  RotationXX(const real_type theta) : QRotationGate2<T>(0, 1, theta) {}
  RotationXX(const real_type cos, const real_type sin) : QRotationGate2<T>(0, 1, cos, sin) {}
  RotationXX(const int* qubits, const rotation_type& rot, const bool fixed = false)
      : QRotationGate2<T>(qubits, rot, fixed) {}
  RotationXX(const int* qubits, const real_type theta, const bool fixed = false)
      : QRotationGate2<T>(qubits, theta, fixed) {}
  RotationXX(const int* qubits, const real_type cos, const real_type sin, const bool fixed = false)
      : QRotationGate2<T>(qubits, cos, sin, fixed) {}
  RotationXX(const int qubit0, const int qubit1, const rotation_type& rot, const bool fixed = false)
      : QRotationGate2<T>(qubit0, qubit1, rot, fixed) {}
  RotationXX(const int qubit0, const int qubit1, const real_type theta, const bool fixed = false)
      : QRotationGate2<T>(qubit0, qubit1, theta, fixed) {}
  RotationXX(const int qubit0, const int qubit1, const real_type cos, const real_type sin,
     const bool fixed = false)
      : QRotationGate2<T>(qubit0, qubit1, cos, sin, fixed) {}

  dense::SquareMatrix<T> matrix() const override {
    const real_type c = this->cos(), s = this->sin();
    const T o(0);
    return dense::SquareMatrix<T>(T(c), o, o, T(0, -s),  o, T(c), T(0, -s), o,  o, T(0, -s), T(c), o,  T(0, -s), o, o, T(c));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmRxx(this->qubits_[0] + offset, this->qubits_[1] + offset, this->theta());
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->rotation_record(QSV_GATE_RXX, op, offset));
  }

  RotationXX<T>& operator*=(const RotationXX<T>& rhs) {
    assert(this->qubits() == rhs.qubits());
    this->rotation_ *= rhs.rotation();
    return *this;
  }
  RotationXX<T>& operator/=(const RotationXX<T>& rhs) {
    assert(this->qubits() == rhs.qubits());
    this->rotation_ /= rhs.rotation();
    return *this;
  }
  friend RotationXX<T> operator*(RotationXX<T> lhs, const RotationXX<T>& rhs) { return lhs *= rhs; }
  friend RotationXX<T> operator/(RotationXX<T> lhs, const RotationXX<T>& rhs) { return lhs /= rhs; }
  /// Inverse rotation; like the reference it is built on qubits 0 and 1.
  RotationXX<T> inv() const { return RotationXX<T>(this->rotation_.inv()); }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const RotationXX<T>*>(&other);
    return p && p->rotation() == this->rotation();
  }
};

}  // namespace qgates
}  // namespace qclab
