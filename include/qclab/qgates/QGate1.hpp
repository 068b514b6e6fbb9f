// Base class of 1-qubit gates (reference include/qclab/qgates/QGate1.hpp).
// A subclass that only defines matrix() runs on the engine as a dense 2x2
// gate (record kind QSV_GATE_U1); the built-in gates emit their own kinds.
#pragma once

#include <cassert>

#include "qclab/QObject.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class QGate1 : public QObject<T> {
 public:
  QGate1() : qubit_(0) {}
  QGate1(const int qubit) : qubit_(qubit) { assert(qubit >= 0); }

  int nbQubits() const override { return 1; }
  bool controlled() const override { return false; }
  int qubit() const override { return qubit_; }
  void setQubit(const int qubit) override { assert(qubit >= 0); qubit_ = qubit; }
  std::vector<int> qubits() const override { return {qubit_}; }
  void setQubits(const int* qubits) override { assert(qubits[0] >= 0); qubit_ = qubits[0]; }

  void print() const override {
    std::cout << "qubit: " << qubit() << std::endl;
    printMatrix(this->matrix());
  }

  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    qsv_gate g = backend::make(QSV_GATE_U1, qubit() + offset, 0, char(op));
    backend::set_mat(g, this->matrix());
    out.push_back(g);
  }

 protected:
  // a record of a built-in kind on this gate's qubit
  qsv_gate record(int kind, Op op, int offset) const {
    return backend::make(kind, qubit() + offset, 0, char(op));
  }
  int qubit_;
};

}  // namespace qgates
}  // namespace qclab
