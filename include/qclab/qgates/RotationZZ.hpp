// ZZ rotation exp(-i theta/2 Z(x)Z) (reference include/qclab/qgates/RotationZZ.hpp).
#pragma once

#include "qclab/qgates/QRotationGate2.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class RotationZZ : public QRotationGate2<T> {
 public:
  using real_type = real_t<T>;
  using rotation_type = QRotation<real_type>;

  RotationZZ() : QRotationGate2<T>() {}
  RotationZZ(const rotation_type& rot) : QRotationGate2<T>(0, 1, rot) {}
  RotationZZ(const real_type theta) : QRotationGate2<T>(0, 1, theta) {}
  RotationZZ(const real_type cos, const real_type sin) : QRotationGate2<T>(0, 1, cos, sin) {}
  RotationZZ(const int* qubits, const rotation_type& rot, const bool fixed = false)
      : QRotationGate2<T>(qubits, rot, fixed) {}
  RotationZZ(const int* qubits, const real_type theta, const bool fixed = false)
      : QRotationGate2<T>(qubits, theta, fixed) {}
  RotationZZ(const int* qubits, const real_type cos, const real_type sin, const bool fixed = false)
      : QRotationGate2<T>(qubits, cos, sin, fixed) {}
  RotationZZ(const int qubit0, const int qubit1, const rotation_type& rot, const bool fixed = false)
      : QRotationGate2<T>(qubit0, qubit1, rot, fixed) {}
  RotationZZ(const int qubit0, const int qubit1, const real_type theta, const bool fixed = false)
      : QRotationGate2<T>(qubit0, qubit1, theta, fixed) {}
  RotationZZ(const int qubit0, const int qubit1, const real_type cos, const real_type sin,
     const bool fixed = false)
      : QRotationGate2<T>(qubit0, qubit1, cos, sin, fixed) {}

  dense::SquareMatrix<T> matrix() const override {
    const real_type c = this->cos(), s = this->sin();
    const T o(0);
    return dense::SquareMatrix<T>(T(c, -s), o, o, o,  o, T(c, s), o, o,  o, o, T(c, s), o,  o, o, o, T(c, -s));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmRzz(this->qubits_[0] + offset, this->qubits_[1] + offset, this->theta());
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->rotation_record(QSV_GATE_RZZ, op, offset));
  }

  RotationZZ<T>& operator*=(const RotationZZ<T>& rhs) {
    assert(this->qubits() == rhs.qubits());
    this->rotation_ *= rhs.rotation();
    return *this;
  }
  RotationZZ<T>& operator/=(const RotationZZ<T>& rhs) {
    assert(this->qubits() == rhs.qubits());
    this->rotation_ /= rhs.rotation();
    return *this;
  }
  friend RotationZZ<T> operator*(RotationZZ<T> lhs, const RotationZZ<T>& rhs) { return lhs *= rhs; }
  friend RotationZZ<T> operator/(RotationZZ<T> lhs, const RotationZZ<T>& rhs) { return lhs /= rhs; }
  /// Inverse rotation; like the reference it is built on qubits 0 and 1.
  RotationZZ<T> inv() const { return RotationZZ<T>(this->rotation_.inv()); }

 protected:
  bool equals(const QObject<T>& other) const override {
    const auto* p = dynamic_cast<const RotationZZ<T>*>(&other);
    return p && p->rotation() == this->rotation();
  }
};

}  // namespace qgates
}  // namespace qclab
