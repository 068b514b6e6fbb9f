// Pauli-Y gate (reference include/qclab/qgates/PauliY.hpp).
#pragma once

#include <cmath>

#include "qclab/qgates/QGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class PauliY : public QGate1<T> {
 public:
  PauliY() : QGate1<T>(0) {}
  PauliY(const int qubit) : QGate1<T>(qubit) {}

  bool fixed() const override { return true; }
  dense::SquareMatrix<T> matrix() const override {
    return dense::SquareMatrix<T>(T(0), T(0, -1), T(0, 1), T(0));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmY(this->qubit_ + offset);
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->record(QSV_GATE_Y, op, offset));
  }

 protected:
  bool equals(const QObject<T>& other) const override {
    return dynamic_cast<const PauliY<T>*>(&other) != nullptr;
  }
};

}  // namespace qgates
}  // namespace qclab
