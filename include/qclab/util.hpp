// Core enums and type helpers of the QCLAB++ API (include/qclab/util.hpp of
// the reference: Op, Side, real_t, is_complex, print helpers).
#pragma once

#include <cassert>
#include <complex>
#include <cstdio>
#include <iostream>
#include <type_traits>

namespace qclab {

/// Operation applied to a gate: the matrix, its transpose, or its adjoint.
enum class Op : char { NoTrans = 'N', Trans = 'T', ConjTrans = 'C' };

/// Side of a matrix product.
enum class Side : char { Left = 'L', Right = 'R' };

template <typename T> struct real_t_helper { using real_type = T; };
template <typename T> struct real_t_helper<std::complex<T>> { using real_type = T; };
template <typename T> using real_t = typename real_t_helper<T>::real_type;

template <typename T> struct is_complex : std::false_type {};
template <typename T> struct is_complex<std::complex<T>> : std::true_type {};
template <typename T> inline constexpr bool is_complex_v = is_complex<T>::value;

template <typename T>
void print(const T val) {
  if (val == T(0)) std::printf("         0");
  else std::printf("%10.4f", double(val));
}

template <typename T>
void print(const std::complex<T> val) {
  const double re = std::real(val), im = std::abs(std::imag(val));
  const char sign = std::imag(val) >= T(0) ? '+' : '-';
  if (std::abs(val) == T(0)) std::printf("                   0");
  else if (std::abs(re) == 0) std::printf("         0 %c %6.4fi", sign, im);
  else if (im == 0) std::printf("%10.4f %c      0i", re, sign);
  else std::printf("%10.4f %c %6.4fi", re, sign, im);
}

template <typename M>
void printMatrix(const M& mat) {
  for (int64_t r = 0; r < mat.rows(); ++r) {
    for (int64_t c = 0; c < mat.cols(); ++c) print(mat(r, c));
    std::cout << std::endl;
  }
}
template <typename M> void printMatrix2x2(const M& mat) { printMatrix(mat); }
template <typename M> void printMatrix4x4(const M& mat) { printMatrix(mat); }

}  // namespace qclab
