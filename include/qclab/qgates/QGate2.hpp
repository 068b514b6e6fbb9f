// Base class of 2-qubit gates (reference include/qclab/qgates/QGate2.hpp).
// A subclass that only defines matrix() runs on the engine as a dense 4x4
// gate (record kind QSV_GATE_U2), applied to a vector with the same basis
// convention as the reference's QGate2::apply (src/qgates/QGate2.cpp).
#pragma once

#include <cassert>

#include "qclab/QObject.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class QGate2 : public QObject<T> {
 public:
  int nbQubits() const override { return 2; }
  void setQubit(const int) override { assert(false); }

  void print() const override {
    const auto q = this->qubits();
    std::cout << "qubits: " << q[0] << ", " << q[1] << std::endl;
    printMatrix(this->matrix());
  }

  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    const auto q = this->qubits();
    qsv_gate g = backend::make(QSV_GATE_U2, q[0] + offset, q[1] + offset, char(op));
    backend::set_mat(g, this->matrix());
    out.push_back(g);
  }
};

}  // namespace qgates
}  // namespace qclab
