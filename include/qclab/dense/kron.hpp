// Kronecker products and identities (reference include/qclab/dense/kron.hpp).
#pragma once

#include "qclab/dense/SquareMatrix.hpp"

namespace qclab {
namespace dense {

/// Identity of the given size.
template <typename T>
SquareMatrix<T> eye(const int64_t size) {
  SquareMatrix<T> I(size, T(0));
  for (int64_t i = 0; i < size; ++i) I(i, i) = T(1);
  return I;
}

/// kron(A, B): A's index is the more significant one.
template <typename T>
SquareMatrix<T> kron(const SquareMatrix<T>& A, const SquareMatrix<T>& B) {
  const int64_t a = A.size(), b = B.size();
  SquareMatrix<T> C(a * b, T(0));
  for (int64_t ja = 0; ja < a; ++ja)
    for (int64_t ia = 0; ia < a; ++ia)
      for (int64_t jb = 0; jb < b; ++jb)
        for (int64_t ib = 0; ib < b; ++ib) C(ia * b + ib, ja * b + jb) = A(ia, ja) * B(ib, jb);
  return C;
}

template <typename T, typename... Ms>
SquareMatrix<T> kron(const SquareMatrix<T>& A, const SquareMatrix<T>& B, const Ms&... rest) {
  return kron(kron(A, B), rest...);
}

}  // namespace dense
}  // namespace qclab
