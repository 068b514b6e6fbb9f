// Hadamard gate (reference include/qclab/qgates/Hadamard.hpp).
#pragma once

#include <cmath>

#include "qclab/qgates/QGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class Hadamard : public QGate1<T> {
 public:
  Hadamard() : QGate1<T>(0) {}
  Hadamard(const int qubit) : QGate1<T>(qubit) {}

  bool fixed() const override { return true; }
  dense::SquareMatrix<T> matrix() const override {
    using R = real_t<T>;
    return dense::SquareMatrix<T>(T(1 / std::sqrt(R(2))), T(1 / std::sqrt(R(2))), T(1 / std::sqrt(R(2))), T(-1 / std::sqrt(R(2))));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmH(this->qubit_ + offset);
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->record(QSV_GATE_H, op, offset));
  }

 protected:
  bool equals(const QObject<T>& other) const override {
    return dynamic_cast<const Hadamard<T>*>(&other) != nullptr;
  }
};

}  // namespace qgates
}  // namespace qclab
