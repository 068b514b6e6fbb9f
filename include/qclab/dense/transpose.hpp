// Op applied to small matrices (reference include/qclab/dense/transpose.hpp):
// NoTrans leaves A, Trans transposes it, ConjTrans conjugate-transposes it.
#pragma once

#include <utility>

#include "qclab/util.hpp"

namespace qclab {
namespace dense {

template <typename M>
void operateInPlace(Op op, M& A) {
  if (op == Op::NoTrans) return;
  const int64_t n = A.rows();
  if constexpr (is_complex_v<typename M::value_type>) {
    if (op == Op::ConjTrans)
      for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) A(i, j) = std::conj(A(i, j));
  }
  for (int64_t j = 0; j + 1 < n; ++j)
    for (int64_t i = j + 1; i < n; ++i) std::swap(A(i, j), A(j, i));
}

template <typename M> M operate(Op op, const M& A) { M B(A); operateInPlace(op, B); return B; }
template <typename M> void transInPlace(M& A) { operateInPlace(Op::Trans, A); }
template <typename M> M trans(const M& A) { return operate(Op::Trans, A); }
template <typename M> void conjTransInPlace(M& A) { operateInPlace(Op::ConjTrans, A); }
template <typename M> M conjTrans(const M& A) { return operate(Op::ConjTrans, A); }

}  // namespace dense
}  // namespace qclab
