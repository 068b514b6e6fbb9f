// iSWAP gate (reference include/qclab/qgates/iSWAP.hpp).
#pragma once

#include <algorithm>

#include "qclab/qgates/QGate2.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class iSWAP : public QGate2<T> {
 public:
  iSWAP() : qubits_{0, 1} {}
  iSWAP(const int qubit0, const int qubit1) { const int q[2] = {qubit0, qubit1}; setQubits(q); }
  iSWAP(const int* qubits) { setQubits(qubits); }

  bool fixed() const override { return true; }
  bool controlled() const override { return false; }
  int qubit() const override { return qubits_[0]; }
  std::vector<int> qubits() const override { return {qubits_[0], qubits_[1]}; }
  void setQubits(const int* qubits) override {
    assert(qubits[0] >= 0 && qubits[1] >= 0 && qubits[0] != qubits[1]);
    qubits_[0] = std::min(qubits[0], qubits[1]);
    qubits_[1] = std::max(qubits[0], qubits[1]);
  }
  dense::SquareMatrix<T> matrix() const override {
    const T o(0), l(1), i(0, 1);
    return dense::SquareMatrix<T>(l, o, o, o,  o, o, i, o,  o, i, o, o,  o, o, o, l);
  }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    stream << qasmiSWAP(qubits_[0] + offset, qubits_[1] + offset);
    return 0;
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    out.push_back(backend::make(QSV_GATE_ISWAP, qubits_[0] + offset, qubits_[1] + offset, char(op)));
  }

 protected:
  bool equals(const QObject<T>& other) const override {
    return dynamic_cast<const iSWAP<T>*>(&other) != nullptr;
  }
  int qubits_[2];
};

}  // namespace qgates
}  // namespace qclab
