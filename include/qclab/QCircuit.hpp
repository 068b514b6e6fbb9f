// Quantum circuit: an ordered list of gates and sub-circuits
// (reference include/qclab/QCircuit.hpp).
//
// simulate()/apply() on a vector flatten the whole circuit -- nested
// circuits, qubit offsets, pointer gates and the reversed order of [Conj]Trans
// (QCircuit.hpp:130-143) -- into engine records and run them in ONE engine
// call: one upload, the fused schedule on the GPU, one download.
// simulate_device() does the same on a device-resident vector.
#pragma once

#include <algorithm>
#include <cassert>
#include <cmath>
#include <iterator>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "qclab/QAdjustable.hpp"
#include "qclab/QObject.hpp"
#include "qclab/io/QASMFile.hpp"

namespace qclab {

template <typename T, typename G = QObject<T>>
class QCircuit : public QObject<T>, public QAdjustable {
 public:
  using gate_type = G;
  using vector_type = std::vector<std::unique_ptr<gate_type>>;
  using size_type = typename vector_type::size_type;
  using reference = typename vector_type::reference;
  using const_reference = typename vector_type::const_reference;
  using iterator = typename vector_type::iterator;
  using const_iterator = typename vector_type::const_iterator;
  using reverse_iterator = typename vector_type::reverse_iterator;
  using const_reverse_iterator = typename vector_type::const_reverse_iterator;

  QCircuit(const int nbQubits) : nbQubits_(nbQubits), offset_(0) { assert(nbQubits > 0); }
  QCircuit(const int nbQubits, const int offset) : nbQubits_(nbQubits), offset_(offset) {
    assert(nbQubits > 0 && offset >= 0);
  }
  QCircuit(const int nbQubits, const int offset, const int nbGates)
      : nbQubits_(nbQubits), offset_(offset), gates_(nbGates) {
    assert(nbQubits > 0 && offset >= 0 && nbGates >= 0);
  }
  /// Circuit read from an OpenQASM 2.0 file (reference loader semantics).
  QCircuit(const std::string filename, const int offset = 0) : offset_(offset) {
    io::QASMFile qasm(filename);
    nbQubits_ = qasm.nbQubits();
    for (auto& gate : qasm.template gates<G>()) push_back(std::move(gate));
  }

  int nbQubits() const override { return nbQubits_; }
  bool fixed() const override { return QAdjustable::fixed(); }
  bool controlled() const override { return false; }
  int qubit() const override { return offset_; }
  void setQubit(const int) override { assert(false); }
  std::vector<int> qubits() const override {
    std::vector<int> v(nbQubits_);
    std::iota(v.begin(), v.end(), offset_);
    return v;
  }
  void setQubits(const int*) override { assert(false); }

  /// Unitary of the circuit (MSB kron convention): on the GPU from
  /// backend::matrix_gpu_min() qubits up, host arithmetic below.
  dense::SquareMatrix<T> matrix() const override {
    return nbQubits_ >= backend::matrix_gpu_min() ? matrix_gpu() : matrix_host();
  }
  /// The reference's construction: every gate's apply(Side::Right, ...) on
  /// an identity, in host arithmetic (QCircuit.hpp:121-127).
  dense::SquareMatrix<T> matrix_host() const {
    auto mat = dense::eye<T>(int64_t(1) << nbQubits_);
    for (auto it = begin(); it != end(); ++it) (*it)->apply(Side::Right, Op::NoTrans, nbQubits_, mat);
    return mat;
  }
  /// The same unitary built on the GPU: the identity as a 2k-qubit state,
  /// the gates applied to its row qubits (one engine call, one download).
  dense::SquareMatrix<T> matrix_gpu() const {
    dense::SquareMatrix<T> mat(int64_t(1) << nbQubits_);
    backend::circuit_matrix(mat.ptr(), nbQubits_, gpu_records(), backend::fuse_k());
    return mat;
  }
  /// Engine records of the gates as matrix() applies them: in order, with
  /// offset 0 (nested circuits add their own offsets).
  std::vector<qsv_gate> gpu_records() const {
    std::vector<qsv_gate> g;
    for (auto it = begin(); it != end(); ++it) (*it)->records(g, Op::NoTrans, 0);
    return g;
  }

  /// Applies the circuit (with `op`) to `vector` on the GPU.
  void apply(Op op, const int nbQubits, std::vector<T>& vector, const int offset = 0) const override {
    QObject<T>::apply(op, nbQubits, vector, offset);
  }
  void apply_device(Op op, const int nbQubits, T* vector, const int offset = 0) const override {
    QObject<T>::apply_device(op, nbQubits, vector, offset);
  }
  /// Simulates the circuit on `vector` (2^nbQubits amplitudes, in place).
  void simulate(std::vector<T>& vector) const { apply(Op::NoTrans, nbQubits_, vector); }
  /// Simulates the circuit on a device vector (hipMalloc'd, in place).
  void simulate_device(T* vector) const { apply_device(Op::NoTrans, nbQubits_, vector); }

  void apply(Side side, Op op, const int nbQubits, dense::SquareMatrix<T>& matrix,
             const int offset = 0) const override {
    // Left (matrix * op(C)) takes the gates last-first for NoTrans and
    // first-first for [Conj]Trans; Right (op(C) * matrix) the opposite.
    const bool forward = (op == Op::NoTrans) == (side == Side::Right);
    if (forward)
      for (auto it = begin(); it != end(); ++it) (*it)->apply(side, op, nbQubits, matrix, offset_ + offset);
    else
      for (auto it = rbegin(); it != rend(); ++it) (*it)->apply(side, op, nbQubits, matrix, offset_ + offset);
  }

  void print() const override { printMatrix(matrix()); }
  int toQASM(std::ostream& stream, const int offset = 0) const override {
    for (auto it = begin(); it != end(); ++it) {
      const int out = (*it)->toQASM(stream, offset_ + offset);
      if (out != 0) return out;
    }
    return 0;
  }

  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    if (op == Op::NoTrans)
      for (auto it = begin(); it != end(); ++it) (*it)->records(out, op, offset_ + offset);
    else
      for (auto it = rbegin(); it != rend(); ++it) (*it)->records(out, op, offset_ + offset);
  }

  int offset() const { return offset_; }
  void setOffset(const int offset) { offset_ = offset; }

  // element access
  reference at(const size_type pos) { return gates_.at(pos); }
  const_reference at(const size_type pos) const { return gates_.at(pos); }
  reference operator[](const size_type pos) { return gates_[pos]; }
  const_reference operator[](const size_type pos) const { return gates_[pos]; }
  reference front() { return gates_.front(); }
  const_reference front() const { return gates_.front(); }
  reference back() { return gates_.back(); }
  const_reference back() const { return gates_.back(); }

  // iterators
  iterator begin() { return gates_.begin(); }
  const_iterator begin() const { return gates_.begin(); }
  iterator end() { return gates_.end(); }
  const_iterator end() const { return gates_.end(); }
  reverse_iterator rbegin() { return gates_.rbegin(); }
  const_reverse_iterator rbegin() const { return gates_.rbegin(); }
  reverse_iterator rend() { return gates_.rend(); }
  const_reverse_iterator rend() const { return gates_.rend(); }

  // capacity
  bool empty() const { return gates_.empty(); }
  size_type nbGates() const { return gates_.size(); }
  size_type maxNbGates() const { return gates_.max_size(); }
  void reserve(const size_type new_cap) { gates_.reserve(new_cap); }
  size_type capacity() const { return gates_.capacity(); }
  void shrink_to_fit() { gates_.shrink_to_fit(); }

  // modifiers
  void clear() { gates_.clear(); }
  iterator insert(const_iterator pos, std::unique_ptr<gate_type> gate) {
    assert(canInsert(*gate));
    return gates_.insert(pos, std::move(gate));
  }
  template <typename InputIt>
  iterator insert(const_iterator pos, InputIt first, InputIt last) {
    const auto idx = std::distance(const_iterator(gates_.begin()), pos);
    for (InputIt it = first; it != last; ++it) assert(canInsert(**it));
    gates_.insert(gates_.begin() + idx, std::make_move_iterator(first), std::make_move_iterator(last));
    return gates_.begin() + idx;
  }
  iterator erase(const_iterator pos) { return gates_.erase(pos); }
  iterator erase(const_iterator first, const_iterator last) { return gates_.erase(first, last); }
  void push_back(std::unique_ptr<gate_type> gate) {
    assert(canInsert(*gate));
    gates_.push_back(std::move(gate));
  }
  void pop_back() { gates_.pop_back(); }
  void resize(const size_type count) { gates_.resize(count); }

  /// A gate fits if its highest qubit is below nbQubits.
  bool canInsert(const gate_type& gate) const {
    const auto q = gate.qubits();
    return *std::max_element(q.begin(), q.end()) < nbQubits_;
  }

 protected:
  bool equals(const QObject<T>& other) const override { return other.matrix() == matrix(); }

  int nbQubits_;
  int offset_;
  vector_type gates_;
};

/// Frobenius norm of the difference of two circuits' unitaries, computed
/// with a max-abs scale against overflow (reference QCircuit.hpp:406-430),
/// in host arithmetic.
template <typename C1, typename C2>
double nrmF_host(const C1& circuit1, const C2& circuit2) {
  assert(circuit1.nbQubits() == circuit2.nbQubits());
  const auto m1 = circuit1.matrix();
  const auto m2 = circuit2.matrix();
  double scale = 0.0;
  for (int64_t c = 0; c < m1.cols(); ++c)
    for (int64_t r = 0; r < m1.rows(); ++r) scale = std::max(scale, double(std::abs(m1(r, c) - m2(r, c))));
  if (scale == 0.0) scale = 1.0;
  double sum = 0.0;
  for (int64_t c = 0; c < m1.cols(); ++c)
    for (int64_t r = 0; r < m1.rows(); ++r) {
      const double re = std::real(m1(r, c) - m2(r, c)) / scale, im = std::imag(m1(r, c) - m2(r, c)) / scale;
      sum += re * re + im * im;
    }
  return scale * std::sqrt(sum);
}

template <typename C1, typename C2>
double nrmF(const C1& circuit1, const C2& circuit2) {
  return nrmF_host(circuit1, circuit2);
}

/// nrmF of two circuits: from backend::matrix_gpu_min() qubits up both
/// unitaries are built and reduced on the GPU, with the same scaling.
template <typename T, typename G1, typename G2>
double nrmF(const QCircuit<T, G1>& circuit1, const QCircuit<T, G2>& circuit2) {
  assert(circuit1.nbQubits() == circuit2.nbQubits());
  if (circuit1.nbQubits() < backend::matrix_gpu_min()) return nrmF_host(circuit1, circuit2);
  return backend::circuit_nrmf<T>(circuit1.nbQubits(), circuit1.gpu_records(), circuit2.gpu_records(),
                                  backend::fuse_k());
}

}  // namespace qclab
