// Glue between the QCLAB++-compatible classes and the MI355X engine's C ABI
// (include/qsv.h).  Not part of the reference API.
#pragma once

#include <complex>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "qsv.h"

namespace qclab {
namespace backend {

/// Engine element type of T; only complex states run on the GPU.
template <typename T>
constexpr int dtype() {
  static_assert(std::is_same_v<T, std::complex<double>> || std::is_same_v<T, std::complex<float>>,
                "the MI355X engine simulates std::complex<float> and std::complex<double> states");
  return std::is_same_v<T, std::complex<double>> ? QSV_C128 : QSV_C64;
}

/// Engine failures (the reference asserts) surface as exceptions here.
class error : public std::runtime_error {
 public:
  error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(const int status) {
  if (status != QSV_OK) throw error(status, std::string("qclab MI355X engine: ") + qsv_last_error());
}

/// Compiled form used by QCircuit::simulate/apply: environment QCLAB_FUSE_K =
/// 0 (one plain sweep per gate), 1..5 (fused blocks of up to k qubits),
/// "tiled" (interpreted tiled passes), "jit" (compiled tiled passes),
/// "jitfixed" (compiled with the rotations' forms fixed: for circuits
/// simulated repeatedly with the same angles) or "auto" (the default:
/// compiled tiled passes from 22 qubits up, fused k = 5 below).
inline int fuse_k() {
  const char* e = std::getenv("QCLAB_FUSE_K");
  if (!e || std::string(e) == "auto") return QSV_AUTO;
  if (std::string(e) == "tiled") return QSV_TILED;
  if (std::string(e) == "jit") return QSV_TILED_JIT;
  if (std::string(e) == "jitfixed") return QSV_TILED_JIT_FIXED;
  const int k = std::atoi(e);
  return k < 0 ? 0 : (k > 5 ? 5 : k);
}

inline qsv_gate make(int kind, int q0, int q1 = 0, int op = QSV_OP_N) {
  qsv_gate g{};
  g.kind = kind;
  g.q[0] = q0;
  g.q[1] = q1;
  g.ctrl_state = 1;
  g.op = op;
  g.param = -1;
  return g;
}

inline void set_cs(qsv_gate& g, double c, double s) {
  g.has_cs = 1;
  g.cs[0] = c;
  g.cs[1] = s;
}

/// Column-major matrix into the record's interleaved mat[] field.
template <typename M>
void set_mat(qsv_gate& g, const M& m) {
  const int64_t n = m.size();
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i) {
      const std::complex<double> z(m(i, j));
      g.mat[2 * (i + j * n)] = z.real();
      g.mat[2 * (i + j * n) + 1] = z.imag();
    }
}

/// Run records on a host vector / a device pointer (one upload, one download).
template <typename T>
void run_host(std::vector<T>& v, int n, const std::vector<qsv_gate>& g, int fuse) {
  if (v.size() != (size_t(1) << n)) throw error(QSV_EINVAL, "qclab: vector size is not 2^nbQubits");
  if (g.empty()) return;
  check(qsv_host_simulate(v.data(), n, dtype<T>(), g.data(), g.size(), fuse));
}

template <typename T>
void run_device(T* v, int n, const std::vector<qsv_gate>& g, int fuse) {
  if (g.empty()) return;
  check(qsv_dev_simulate(v, n, dtype<T>(), g.data(), g.size(), fuse, nullptr));
}

/// QCircuit::matrix() and nrmF of two circuits run on the GPU from this many
/// qubits up (environment QCLAB_MATRIX_GPU_MIN, default 6); smaller unitaries
/// use host arithmetic, as the reference does for all sizes.
inline int matrix_gpu_min() {
  const char* e = std::getenv("QCLAB_MATRIX_GPU_MIN");
  return e ? std::atoi(e) : 6;
}

/// The 2^k x 2^k unitary of k-qubit records into column-major host memory.
template <typename T>
void circuit_matrix(T* out, int k, const std::vector<qsv_gate>& g, int fuse) {
  check(qsv_host_circuit_matrix(out, k, dtype<T>(), g.data(), g.size(), fuse));
}

/// nrmF of two k-qubit record lists, both unitaries built and reduced on the GPU.
template <typename T>
double circuit_nrmf(int k, const std::vector<qsv_gate>& g1, const std::vector<qsv_gate>& g2, int fuse) {
  double out = 0.0;
  check(qsv_circuit_nrmf(k, dtype<T>(), g1.data(), g1.size(), g2.data(), g2.size(), fuse, &out));
  return out;
}

}  // namespace backend
}  // namespace qclab
