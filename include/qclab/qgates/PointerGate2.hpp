// 2-qubit gate that refers to another gate with a qubit offset
// (reference include/qclab/qgates/PointerGate2.hpp).  The pointee is not owned.
#pragma once

#include "qclab/qgates/QGate2.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class PointerGate2 : public QGate2<T> {
 public:
  PointerGate2(QGate2<T>* gate, const int offset = 0) : offset_(offset), gate_(gate) {}

  bool fixed() const override { return gate_->fixed(); }
  bool controlled() const override { return gate_->controlled(); }
  int qubit() const override { return gate_->qubit() + offset_; }
  void setQubit(const int) override { assert(false); }
  std::vector<int> qubits() const override {
    std::vector<int> q = gate_->qubits();
    for (int& x : q) x += offset_;
    return q;
  }
  void setQubits(const int*) override { assert(false); }
  dense::SquareMatrix<T> matrix() const override { return gate_->matrix(); }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    return gate_->toQASM(stream, offset_ + offset);
  }
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    gate_->records(out, op, offset_ + offset);
  }

  int offset() const { return offset_; }
  void setOffset(const int offset) { offset_ = offset; }
  QGate2<T>* ptr() { return gate_; }
  const QGate2<T>* ptr() const { return gate_; }

 protected:
  bool equals(const QObject<T>& other) const override {
    if (const auto* p = dynamic_cast<const PointerGate2<T>*>(&other)) return *p->ptr() == *ptr();
    return other == *ptr();
  }
  int offset_;
  QGate2<T>* gate_;
};

}  // namespace qgates
}  // namespace qclab
