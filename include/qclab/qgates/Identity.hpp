// Identity gate (reference include/qclab/qgates/Identity.hpp).
#pragma once

#include <cmath>

#include "qclab/qgates/QGate1.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class Identity : public QGate1<T> {
 public:
  Identity() : QGate1<T>(0) {}
  Identity(const int qubit) : QGate1<T>(qubit) {}

  bool fixed() const override { return true; }
  dense::SquareMatrix<T> matrix() const override {
    return dense::eye<T>(2);
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    (void)stream; (void)offset;  // the identity writes no QASM
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(this->record(QSV_GATE_I, op, offset));
  }

 protected:
  bool equals(const QObject<T>& other) const override {
    return dynamic_cast<const Identity<T>*>(&other) != nullptr;
  }
};

}  // namespace qgates
}  // namespace qclab
