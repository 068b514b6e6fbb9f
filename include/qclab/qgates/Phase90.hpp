// Phase gate with angle pi/2 (S gate) (reference include/qclab/qgates/Phase90.hpp).
#pragma once

#include <cmath>

#include "qclab/qgates/QGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class Phase90 : public QGate1<T> {
 public:
  Phase90() : QGate1<T>(0) {}
  Phase90(const int qubit) : QGate1<T>(qubit) {}

  bool fixed() const override { return true; }
  dense::SquareMatrix<T> matrix() const override {
    return dense::SquareMatrix<T>(T(1), T(0), T(0), T(0, 1));
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmS(this->qubit_ + offset);
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->record(QSV_GATE_S, op, offset));
  }

 protected:
  bool equals(const QObject<T>& other) const override {
    return dynamic_cast<const Phase90<T>*>(&other) != nullptr;
  }
};

}  // namespace qgates
}  // namespace qclab
