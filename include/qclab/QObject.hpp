// Abstract base of gates and circuits (reference include/qclab/QObject.hpp).
// apply() runs on the MI355X engine: an object contributes engine records
// (records()) and the vector is simulated in one upload/run/download.
#pragma once

#include <ostream>
#include <vector>

#include "qclab/backend.hpp"
#include "qclab/dense/SquareMatrix.hpp"
#include "qclab/dense/kron.hpp"
#include "qclab/dense/transpose.hpp"
#include "qclab/qasm.hpp"
#include "qclab/util.hpp"

namespace qclab {

template <typename T>
class QObject {
 public:
  using value_type = T;

  virtual int nbQubits() const = 0;
  virtual bool fixed() const = 0;
  virtual bool controlled() const = 0;
  virtual int qubit() const = 0;
  virtual void setQubit(const int qubit) = 0;
  /// Qubits in ascending order.
  virtual std::vector<int> qubits() const = 0;
  virtual void setQubits(const int* qubits) = 0;
  virtual dense::SquareMatrix<T> matrix() const = 0;

  /// Applies this object (with `op`) to `vector` of `nbQubits` qubits, its
  /// qubits shifted by `offset` (MSB order: qubit 0 is the most significant).
  virtual void apply(Op op, const int nbQubits, std::vector<T>& vector,
                     const int offset = 0) const {
    std::vector<qsv_gate> g;
    records(g, op, offset);
    backend::run_host(vector, nbQubits, g, backend::fuse_k());
  }

  /// Same on a device vector (hipMalloc'd, 2^nbQubits elements).
  virtual void apply_device(Op op, const int nbQubits, T* vector, const int offset = 0) const {
    std::vector<qsv_gate> g;
    records(g, op, offset);
    backend::run_device(vector, nbQubits, g, backend::fuse_k());
  }

  /// matrix := matrix * op(G) (Left) or op(G) * matrix (Right), with G this
  /// object embedded in nbQubits qubits (host arithmetic).
  virtual void apply(Side side, Op op, const int nbQubits, dense::SquareMatrix<T>& matrix,
                     const int offset = 0) const {
    dense::SquareMatrix<T> g = matrix_in(nbQubits, offset);
    dense::operateInPlace(op, g);
    matrix = side == Side::Left ? matrix * g : g * matrix;
  }

  virtual void print() const { printMatrix(matrix()); }
  virtual int toQASM(std::ostream& stream, const int offset = 0) const = 0;

  bool operator==(const QObject<T>& other) const { return equals(other); }
  bool operator!=(const QObject<T>& other) const { return !(*this == other); }

  virtual ~QObject() noexcept = default;

  /// Engine records of this object applied with `op`, in application order.
  virtual void records(std::vector<qsv_gate>& out, Op op, int offset) const = 0;

  /// This object's unitary embedded in nbQubits qubits (qubits + offset).
  virtual dense::SquareMatrix<T> matrix_in(const int nbQubits, const int offset) const {
    const dense::SquareMatrix<T> m = matrix();
    std::vector<int> q = qubits();
    for (int& x : q) x += offset;
    const int k = int(q.size());
    const int64_t N = int64_t(1) << nbQubits;
    dense::SquareMatrix<T> G(N, T(0));
    auto local = [&](int64_t idx) {
      int64_t l = 0;
      for (int j = 0; j < k; ++j) l = (l << 1) | ((idx >> (nbQubits - 1 - q[j])) & 1);
      return l;
    };
    int64_t mask = 0;
    for (int j = 0; j < k; ++j) mask |= int64_t(1) << (nbQubits - 1 - q[j]);
    for (int64_t c = 0; c < N; ++c)
      for (int64_t lr = 0; lr < (int64_t(1) << k); ++lr) {
        int64_t r = c & ~mask;
        for (int j = 0; j < k; ++j)
          if (lr >> (k - 1 - j) & 1) r |= int64_t(1) << (nbQubits - 1 - q[j]);
        G(r, c) = m(lr, local(c));
      }
    return G;
  }

 protected:
  virtual bool equals(const QObject<T>& other) const = 0;
};

}  // namespace qclab
