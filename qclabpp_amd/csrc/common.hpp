// Shared host/device definitions of the MI355X state-vector engine.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <cstdio>
#include <string>

#include "qsv.h"

namespace qsv {

using cd = std::complex<double>;

// ---- error reporting -------------------------------------------------------
// Each C entry point returns a status; the message of the last failure is kept
// per host thread and handed out by qsv_last_error().
void set_error(const std::string& msg);
const char* get_error();

struct Status {
  int code;
};

#define QSV_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::qsv::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));   \
      return QSV_EHIP;                                                        \
    }                                                                         \
  } while (0)

#define QSV_TRY(expr)                                                         \
  do {                                                                        \
    int s_ = (expr);                                                          \
    if (s_ != QSV_OK) return s_;                                              \
  } while (0)

// A launch may hold fewer than 2^32 work-items (gridDim.x * blockDim.x), so
// sweeps over more blocks than this run as several launches over consecutive
// block ranges; each such kernel takes the index of its first block (blk0).
constexpr uint64_t kMaxLaunchBlocks = uint64_t(1) << 22;

#define QSV_LAUNCH_BLOCKS(KERNEL, NBLOCKS, THREADS, LDS, STREAM, ...)                           \
  for (uint64_t b0_ = 0, nb_ = (NBLOCKS); b0_ < nb_; b0_ += ::qsv::kMaxLaunchBlocks)           \
  hipLaunchKernelGGL(KERNEL, dim3(unsigned(nb_ - b0_ < ::qsv::kMaxLaunchBlocks ? nb_ - b0_       \
                                                                      : ::qsv::kMaxLaunchBlocks)), \
                     dim3(THREADS), LDS, STREAM, __VA_ARGS__, b0_)

inline int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

// ---- device complex helpers -------------------------------------------------
// Streaming load and store of one amplitude: a sweep reads and writes every
// amplitude once, so both carry the nontemporal hint (measured on gfx950 for
// an in-place 2x2 sweep: +3-5% bandwidth with both, none with stores alone).
typedef double qsv_v2d __attribute__((ext_vector_type(2)));
typedef float qsv_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_stream(double2* p, const double2 v) {
  __builtin_nontemporal_store(qsv_v2d{v.x, v.y}, reinterpret_cast<qsv_v2d*>(p));
}
__device__ __forceinline__ void st_stream(float2* p, const float2 v) {
  __builtin_nontemporal_store(qsv_v2f{v.x, v.y}, reinterpret_cast<qsv_v2f*>(p));
}
__device__ __forceinline__ double2 ld_stream(const double2* p) {
  const qsv_v2d v = __builtin_nontemporal_load(reinterpret_cast<const qsv_v2d*>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 ld_stream(const float2* p) {
  const qsv_v2f v = __builtin_nontemporal_load(reinterpret_cast<const qsv_v2f*>(p));
  return make_float2(v.x, v.y);
}

template <typename R> struct CT;
template <> struct CT<double> { using type = double2; };
template <> struct CT<float>  { using type = float2; };

template <typename C>
__host__ __device__ __forceinline__ C mk(decltype(C::x) re, decltype(C::x) im) {
  C r;
  r.x = re;
  r.y = im;
  return r;
}

// a * b
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// a * b + c
template <typename C>
__device__ __forceinline__ C cmad(C a, C b, C c) {
  return mk<C>(c.x + a.x * b.x - a.y * b.y, c.y + a.x * b.y + a.y * b.x);
}

// ---- lowered kernel operations ---------------------------------------------
// Every gate kind maps to one of these sweep classes (the classes of
// src/qgates/apply.hpp:258-371), with its Op already folded into the
// coefficients on the host.
enum KKind : int {
  K_NOP = 0,
  K_DENSE1,   // 2x2 on every pair                      (apply2)
  K_DIAG1,    // diag(d0,d1) on every amplitude         (apply2 with RotationZ)
  K_PHASE1,   // scale the |1> half                     (apply2b)
  K_CDENSE1,  // 2x2 on the control==cs half            (apply4, controlled)
  K_CPHASE,   // scale the |cs,1> quarter               (apply4b)
  K_SWAP,     // exchange |01>,|10>, scale both         (apply4bc)
  K_DENSE2,   // 4x4 on every quad                      (apply4, dense)
  K_DIAG2,    // diag(d00,d01,d10,d11) on every amplitude (diagonal 4x4)
  K_DENSEK,   // fused dense 2^k x 2^k block, k <= 5
  K_DIAGK,    // merged diagonal over k qubits (table of 2^k phases)
};

// Row-major coefficient storage: m[r*dim + c] multiplies input c into output r.
struct KOp {
  int kind = K_NOP;
  int nq = 0;          // number of qubits acted on
  int q[8] = {0};      // qubits (MSB numbering).  Controlled: q[0]=control,
                       // q[1]=target.  Dense2/diag2: q[0] < q[1].
                       // Fused: ascending.
  int cs = 1;          // control state
  cd m[16];            // DENSE1: 4, DIAG1: 2, PHASE1/CPHASE/SWAP: m[0],
                       // CDENSE1: 4, DENSE2: 16, DIAG2: 4
  int blk = -1;        // K_DENSEK / K_DIAGK: index of the coefficient block
  // adjoint support: Pauli generator of a rotation, -1 if none
  int param = -1;
  int gen = -1;        // gate kind of the generator owner (QSV_GATE_*)
  // Adjoint reverse sweep in compiled passes: when grow >= 0, the op's
  // generator is measured as result row grow right before the op:
  // G|y> = gph * (-1)^popc(y & gz) * [(y & gpm) == gpv] |y ^ gflip> over
  // address bits, and the row is gw_re * Re<lam|G|phi> + gw_im * Im<lam|G|phi>.
  int grow = -1;
  uint64_t gflip = 0, gz = 0, gpm = 0, gpv = 0;
  double gph_re = 1, gph_im = 0, gw_re = 0, gw_im = 0;
};

// Lower one gate record to a kernel op; validates qubits against n.
int lower_gate(const qsv_gate& g, int n, KOp& out);

// Number of bytes one plain pass of `op` must read and write
// (the algorithmic traffic of its sweep class).
double algorithmic_bytes(const KOp& op, int n, int elem_bytes);

}  // namespace qsv
