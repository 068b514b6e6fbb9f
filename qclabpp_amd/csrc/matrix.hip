// Unitary of a circuit on the GPU and the Frobenius distance of two
// unitaries: the matrix state's set-up and the nrmF reduction.
#include "matrix.hpp"

#include <algorithm>
#include <cmath>

namespace qsv {

int matrix_records(int k, const qsv_gate* gates, size_t ngates, std::vector<qsv_gate>& out) {
  if (k < 1 || 2 * k > 40) return fail(QSV_EINVAL, "matrix qubits must be in [1, 20]");
  if (ngates && !gates) return fail(QSV_EINVAL, "null gate array");
  out.assign(gates, gates + ngates);
  for (qsv_gate& g : out) {
    if (g.kind < 0 || g.kind >= QSV_GATE_COUNT) return fail(QSV_EINVAL, "unknown gate kind");
    const bool two = g.kind >= QSV_GATE_CX;
    for (int j = 0; j < (two ? 2 : 1); ++j) {
      if (g.q[j] < 0 || g.q[j] >= k) return fail(QSV_EQUBIT, "gate qubit outside the circuit");
      g.q[j] += k;
    }
    if (g.kind == QSV_GATE_U2) {
      // swap basis positions 1 and 2 of the column-major 4x4 (P M P); the
      // lowering swaps them back, so the kernels apply M in kron order
      double m[32];
      auto p = [](int i) { return i == 1 ? 2 : (i == 2 ? 1 : i); };
      for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
          m[2 * (r + 4 * c)] = g.mat[2 * (p(r) + 4 * p(c))];
          m[2 * (r + 4 * c) + 1] = g.mat[2 * (p(r) + 4 * p(c)) + 1];
        }
      std::copy(m, m + 32, g.mat);
    }
  }
  return QSV_OK;
}

namespace {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

// m[i] = 1 on the diagonal (i a multiple of 2^k + 1), 0 elsewhere
template <typename C>
__global__ __launch_bounds__(256) void k_set_identity(C* __restrict__ m, uint64_t count, uint64_t diag_step) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
    C z;
    z.x = i % diag_step == 0 ? 1 : 0;
    z.y = 0;
    __builtin_nontemporal_store(z.x, &m[i].x);
    __builtin_nontemporal_store(z.y, &m[i].y);
  }
}

// per-block max |a_i - b_i| (pass 1) or sum |a_i - b_i|^2 / scale^2 (pass 2)
template <typename C, bool kMax>
__global__ __launch_bounds__(kRedThreads) void k_diff_reduce(const C* __restrict__ a, const C* __restrict__ b,
                                                            uint64_t count, const double* __restrict__ scale,
                                                            double* __restrict__ partials) {
  __shared__ double sh[kRedThreads];
  const double inv = kMax ? 1.0 : 1.0 / *scale;
  double acc = 0;
  const uint64_t stride = uint64_t(gridDim.x) * kRedThreads;
  for (uint64_t i = uint64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < count; i += stride) {
    const double dx = double(a[i].x) - double(b[i].x), dy = double(a[i].y) - double(b[i].y);
    if (kMax) {
      acc = fmax(acc, sqrt(dx * dx + dy * dy));
    } else {
      const double rx = dx * inv, ry = dy * inv;
      acc += rx * rx + ry * ry;
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = kMax ? fmax(sh[threadIdx.x], sh[threadIdx.x + s])
                                                : sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// single-block fold of the max partials into the scale (1 where all are 0)
__global__ __launch_bounds__(kRedThreads) void k_fold_scale(const double* __restrict__ partials, int n,
                                                           double* __restrict__ scale) {
  __shared__ double sh[kRedThreads];
  double m = 0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) m = fmax(m, partials[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) scale[0] = sh[0] == 0.0 ? 1.0 : sh[0];
}

template <typename C>
void launch_nrmf(const void* a, const void* b, uint64_t count, double* work, hipStream_t st) {
  const C* pa = static_cast<const C*>(a);
  const C* pb = static_cast<const C*>(b);
  double* scale = work + kRedBlocks;
  hipLaunchKernelGGL((k_diff_reduce<C, true>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, pa, pb, count,
                     static_cast<const double*>(scale), work);
  hipLaunchKernelGGL(k_fold_scale, dim3(1), dim3(kRedThreads), 0, st, static_cast<const double*>(work),
                     kRedBlocks, scale);
  hipLaunchKernelGGL((k_diff_reduce<C, false>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, pa, pb, count,
                     static_cast<const double*>(scale), work);
}

}  // namespace

int launch_set_identity(void* m, int k, int dtype, hipStream_t st) {
  const uint64_t N = uint64_t(1) << k, count = N * N;
  const unsigned grid = unsigned(std::min<uint64_t>((count + 255) / 256, 8192));
  if (dtype == QSV_C128)
    hipLaunchKernelGGL(k_set_identity<double2>, dim3(grid), dim3(256), 0, st, static_cast<double2*>(m), count,
                       N + 1);
  else
    hipLaunchKernelGGL(k_set_identity<float2>, dim3(grid), dim3(256), 0, st, static_cast<float2*>(m), count,
                       N + 1);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

int dev_nrmf(const void* a, const void* b, uint64_t count, int dtype, double* out, hipStream_t st) {
  if (!a || !b || !out) return fail(QSV_EINVAL, "null argument");
  if (dtype != QSV_C64 && dtype != QSV_C128) return fail(QSV_EDTYPE, "dtype must be QSV_C64 or QSV_C128");
  if (!count) {
    *out = 0.0;
    return QSV_OK;
  }
  double* work = nullptr;  // kRedBlocks partials, then the scale
  if (hipMalloc(&work, (kRedBlocks + 1) * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(QSV_ENOMEM, "nrmF: device allocation failed");
  }
  if (dtype == QSV_C128) launch_nrmf<double2>(a, b, count, work, st);
  else launch_nrmf<float2>(a, b, count, work, st);
  int rc = hipGetLastError() == hipSuccess ? QSV_OK : fail(QSV_EHIP, "nrmF: kernel launch failed");
  std::vector<double> h(kRedBlocks + 1);
  if (rc == QSV_OK &&
      (hipMemcpyAsync(h.data(), work, h.size() * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
       hipStreamSynchronize(st) != hipSuccess))
    rc = fail(QSV_EHIP, "nrmF: copy of the partial sums failed");
  (void)hipFree(work);
  if (rc != QSV_OK) return rc;
  double sum = 0.0;
  for (int i = 0; i < kRedBlocks; ++i) sum += h[i];
  *out = h[kRedBlocks] * std::sqrt(sum);
  return QSV_OK;
}

}  // namespace qsv
