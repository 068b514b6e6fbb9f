// Pauli-X gate (reference include/qclab/qgates/PauliX.hpp).
#pragma once

#include <cmath>

#include "qclab/qgates/QGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class PauliX : public QGate1<T> {
 public:
  PauliX() : QGate1<T>(0) {}
  PauliX(const int qubit) : QGate1<T>(qubit) {}

  bool fixed() const override { return true; }
  dense::SquareMatrix<T> matrix() const override {
    return dense::SquareMatrix<T>(T(0), T(1), T(1), T(0));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmX(this->qubit_ + offset);
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->record(QSV_GATE_X, op, offset));
  }

 protected:
  bool equals(const QObject<T>& other) const override {
    return dynamic_cast<const PauliX<T>*>(&other) != nullptr;
  }
};

}  // namespace qgates
}  // namespace qclab
