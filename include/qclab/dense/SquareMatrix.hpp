// Column-major square matrix of the QCLAB++ API
// (reference include/qclab/dense/SquareMatrix.hpp): element (i,j) at
// data[i + j*size]; the 2x2 and 4x4 constructors take row-order arguments.
#pragma once

#include <cassert>
#include <cstdint>
#include <vector>

namespace qclab {
namespace dense {

template <typename T>
class SquareMatrix {
 public:
  using value_type = T;
  using size_type = int64_t;

  SquareMatrix() : size_(0) {}
  explicit SquareMatrix(const int64_t size) : size_(size), data_(size * size) { assert(size > 0); }
  SquareMatrix(const int64_t size, const T value) : size_(size), data_(size * size, value) {
    assert(size > 0);
  }
  SquareMatrix(const T m00, const T m01, const T m10, const T m11) : size_(2), data_(4) {
    (*this)(0, 0) = m00; (*this)(0, 1) = m01;
    (*this)(1, 0) = m10; (*this)(1, 1) = m11;
  }
  SquareMatrix(const T m00, const T m01, const T m02, const T m03,
               const T m10, const T m11, const T m12, const T m13,
               const T m20, const T m21, const T m22, const T m23,
               const T m30, const T m31, const T m32, const T m33)
      : size_(4), data_(16) {
    const T v[16] = {m00, m01, m02, m03, m10, m11, m12, m13,
                     m20, m21, m22, m23, m30, m31, m32, m33};
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) (*this)(r, c) = v[4 * r + c];
  }

  int64_t size() const { return size_; }
  int64_t rows() const { return size_; }
  int64_t cols() const { return size_; }
  int64_t ld() const { return size_; }
  T* ptr() { return data_.data(); }
  const T* ptr() const { return data_.data(); }

  T& operator()(const int64_t i, const int64_t j) { return data_[i + j * size_]; }
  const T& operator()(const int64_t i, const int64_t j) const { return data_[i + j * size_]; }

  bool operator==(const SquareMatrix<T>& o) const { return size_ == o.size_ && data_ == o.data_; }
  bool operator!=(const SquareMatrix<T>& o) const { return !(*this == o); }

  SquareMatrix<T>& operator+=(const SquareMatrix<T>& o) {
    assert(o.size_ == size_);
    for (size_t k = 0; k < data_.size(); ++k) data_[k] += o.data_[k];
    return *this;
  }
  SquareMatrix<T>& operator-=(const SquareMatrix<T>& o) {
    assert(o.size_ == size_);
    for (size_t k = 0; k < data_.size(); ++k) data_[k] -= o.data_[k];
    return *this;
  }
  SquareMatrix<T>& operator*=(const SquareMatrix<T>& o) { return *this = *this * o; }

  friend SquareMatrix<T> operator+(SquareMatrix<T> a, const SquareMatrix<T>& b) { return a += b; }
  friend SquareMatrix<T> operator-(SquareMatrix<T> a, const SquareMatrix<T>& b) { return a -= b; }
  friend SquareMatrix<T> operator*(const SquareMatrix<T>& a, const SquareMatrix<T>& b) {
    assert(a.size_ == b.size_);
    SquareMatrix<T> c(a.size_, T(0));
    for (int64_t j = 0; j < a.size_; ++j)
      for (int64_t k = 0; k < a.size_; ++k) {
        const T bkj = b(k, j);
        for (int64_t i = 0; i < a.size_; ++i) c(i, j) += a(i, k) * bkj;
      }
    return c;
  }

  /// The trace.
  T trace() const {
    T t(0);
    for (int64_t i = 0; i < size_; ++i) t += (*this)(i, i);
    return t;
  }

 private:
  int64_t size_;
  std::vector<T> data_;
};

}  // namespace dense
}  // namespace qclab
