// C ABI of the engine (declared in include/qsv.h).
#include <algorithm>
#include <cstring>
#include <new>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"
#include "plain.hpp"
#include "fusion.hpp"
#include "program.hpp"
#include "matrix.hpp"

struct qsv_program {
  qsv::Program p;
};

namespace qsv {
namespace {

int check_state_args(const void* psi, int n, int dtype) {
  if (!psi) return fail(QSV_EINVAL, "null state pointer");
  if (dtype != QSV_C64 && dtype != QSV_C128)
    return fail(QSV_EDTYPE, "dtype must be QSV_C64 or QSV_C128");
  if (n < 1 || n > 40) return fail(QSV_EINVAL, "number of qubits must be in [1, 40]");
  return QSV_OK;
}

int elem_bytes(int dtype) { return dtype == QSV_C128 ? 16 : 8; }

int check_fuse_k(int fuse_k) {
  if ((fuse_k >= 0 && fuse_k <= kMaxFuse) || fuse_k == QSV_TILED || fuse_k == QSV_AUTO ||
      fuse_k == QSV_TILED_JIT || fuse_k == QSV_TILED_JIT_FIXED)
    return QSV_OK;
  return fail(QSV_EINVAL, "fuse_k must be in [0, 5], QSV_TILED, QSV_TILED_JIT, QSV_TILED_JIT_FIXED or QSV_AUTO");
}

int apply_lowered(const KOp& op, void* psi, int n, int dtype, hipStream_t st) {
  PlainTuning tu;
  return dtype == QSV_C128 ? launch_plain<double>(op, psi, n, st, tu)
                           : launch_plain<float>(op, psi, n, st, tu);
}

// Device buffer reused by the host-vector entry points.
struct Cache {
  std::mutex mu;
  void* ptr = nullptr;
  size_t bytes = 0;
} g_cache;

// Grows the cache to `bytes` (caller holds g_cache.mu).
int reserve_cache(size_t bytes) {
  if (g_cache.bytes >= bytes) return QSV_OK;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    (void)hipGetLastError();
    return fail(QSV_ENODEV, "no GPU visible to the HIP runtime");
  }
  if (g_cache.ptr) QSV_CHECK_HIP(hipFree(g_cache.ptr));
  g_cache.ptr = nullptr;
  g_cache.bytes = 0;
  if (hipMalloc(&g_cache.ptr, bytes) != hipSuccess) {
    g_cache.ptr = nullptr;
    (void)hipGetLastError();
    return fail(QSV_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  g_cache.bytes = bytes;
  return QSV_OK;
}

// Transfers between a pageable host vector and the device state go
// through two pinned staging buffers: the host-side copy of one chunk (split
// over several threads) overlaps the DMA of the other.  A plain hipMemcpy
// from pageable memory reached 22 GB/s for the two transfers of a 16 GiB
// state.  Guarded by g_cache.mu.
constexpr size_t kStageBytes = size_t(64) << 20;
constexpr int kStageThreads = 8;

struct Stage {
  void* buf[2] = {nullptr, nullptr};
  hipStream_t st = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
} g_stage;

int stage_init() {
  if (g_stage.st) return QSV_OK;
  for (int b = 0; b < 2; ++b) {
    if (hipHostMalloc(&g_stage.buf[b], kStageBytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return fail(QSV_ENOMEM, "pinned staging allocation failed");
    }
    QSV_CHECK_HIP(hipEventCreateWithFlags(&g_stage.ev[b], hipEventDisableTiming));
  }
  QSV_CHECK_HIP(hipStreamCreateWithFlags(&g_stage.st, hipStreamNonBlocking));
  return QSV_OK;
}

int stage_release() {
  for (int b = 0; b < 2; ++b) {
    if (g_stage.buf[b]) QSV_CHECK_HIP(hipHostFree(g_stage.buf[b]));
    if (g_stage.ev[b]) QSV_CHECK_HIP(hipEventDestroy(g_stage.ev[b]));
    g_stage.buf[b] = nullptr;
    g_stage.ev[b] = nullptr;
  }
  if (g_stage.st) QSV_CHECK_HIP(hipStreamDestroy(g_stage.st));
  g_stage.st = nullptr;
  return QSV_OK;
}

// memcpy split over threads (page-aligned slices)
void par_copy(void* dst, const void* src, size_t bytes) {
  const int nt = bytes >= (size_t(8) << 20) ? kStageThreads : 1;
  if (nt == 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t slice = (bytes / nt + 4095) & ~size_t(4095);
  std::thread th[kStageThreads];
  for (int t = 0; t < nt; ++t) {
    const size_t a = std::min(bytes, size_t(t) * slice), b = std::min(bytes, a + slice);
    th[t] = std::thread([=] { std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a); });
  }
  for (int t = 0; t < nt; ++t) th[t].join();
}

int upload(void* dev, const void* host, size_t bytes) {
  QSV_TRY(stage_init());
  for (size_t off = 0, k = 0; off < bytes; off += kStageBytes, ++k) {
    const size_t len = std::min(kStageBytes, bytes - off);
    const int b = int(k & 1);
    QSV_CHECK_HIP(hipEventSynchronize(g_stage.ev[b]));  // buffer b's previous DMA is done
    par_copy(g_stage.buf[b], static_cast<const char*>(host) + off, len);
    QSV_CHECK_HIP(hipMemcpyAsync(static_cast<char*>(dev) + off, g_stage.buf[b], len,
                                 hipMemcpyHostToDevice, g_stage.st));
    QSV_CHECK_HIP(hipEventRecord(g_stage.ev[b], g_stage.st));
  }
  QSV_CHECK_HIP(hipStreamSynchronize(g_stage.st));
  return QSV_OK;
}

int download(void* host, const void* dev, size_t bytes) {
  QSV_TRY(stage_init());
  const size_t nk = (bytes + kStageBytes - 1) / kStageBytes;
  auto issue = [&](size_t k) -> int {
    const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
    QSV_CHECK_HIP(hipMemcpyAsync(g_stage.buf[k & 1], static_cast<const char*>(dev) + off, len,
                                 hipMemcpyDeviceToHost, g_stage.st));
    QSV_CHECK_HIP(hipEventRecord(g_stage.ev[k & 1], g_stage.st));
    return QSV_OK;
  };
  for (size_t k = 0; k < nk && k < 2; ++k) QSV_TRY(issue(k));
  for (size_t k = 0; k < nk; ++k) {
    const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
    QSV_CHECK_HIP(hipEventSynchronize(g_stage.ev[k & 1]));
    par_copy(static_cast<char*>(host) + off, g_stage.buf[k & 1], len);
    if (k + 2 < nk) QSV_TRY(issue(k + 2));
  }
  return QSV_OK;
}

}  // namespace
}  // namespace qsv

using namespace qsv;

extern "C" {

const char* qsv_last_error(void) { return get_error(); }

const char* qsv_version(void) { return "qclabpp_amd 0.1 (HIP, gfx950)"; }

int qsv_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

int qsv_dev_set_basis(void* psi, int n, int dtype, uint64_t index, void* stream) {
  QSV_TRY(check_state_args(psi, n, dtype));
  if (index >= (uint64_t(1) << n)) return fail(QSV_EINVAL, "basis index out of range");
  auto st = static_cast<hipStream_t>(stream);
  return dtype == QSV_C128 ? launch_set_basis<double>(psi, n, index, st)
                           : launch_set_basis<float>(psi, n, index, st);
}

int qsv_dev_apply_gate(void* psi, int n, int dtype, const qsv_gate* gate, void* stream) {
  QSV_TRY(check_state_args(psi, n, dtype));
  if (!gate) return fail(QSV_EINVAL, "null gate");
  KOp op;
  QSV_TRY(lower_gate(*gate, n, op));
  return apply_lowered(op, psi, n, dtype, static_cast<hipStream_t>(stream));
}

int qsv_dev_apply_gates(void* psi, int n, int dtype, const qsv_gate* gates, size_t ngates,
                        void* stream) {
  QSV_TRY(check_state_args(psi, n, dtype));
  if (ngates && !gates) return fail(QSV_EINVAL, "null gate array");
  std::vector<KOp> ops(ngates);
  for (size_t i = 0; i < ngates; ++i) QSV_TRY(lower_gate(gates[i], n, ops[i]));
  for (const KOp& op : ops)
    QSV_TRY(apply_lowered(op, psi, n, dtype, static_cast<hipStream_t>(stream)));
  return QSV_OK;
}

int qsv_program_create(int n, int dtype, const qsv_gate* gates, size_t ngates, int fuse_k,
                       int diag_k, qsv_program** out) {
  if (!out) return fail(QSV_EINVAL, "null output pointer");
  *out = nullptr;
  if (dtype != QSV_C64 && dtype != QSV_C128) return fail(QSV_EDTYPE, "dtype must be QSV_C64 or QSV_C128");
  if (n < 1 || n > 40) return fail(QSV_EINVAL, "number of qubits must be in [1, 40]");
  if (ngates && !gates) return fail(QSV_EINVAL, "null gate array");
  QSV_TRY(check_fuse_k(fuse_k));
  if (diag_k < 0 || diag_k > kMaxDiag) return fail(QSV_EINVAL, "diag_k must be in [0, 10]");
  auto* prog = new (std::nothrow) qsv_program;
  if (!prog) return fail(QSV_ENOMEM, "host allocation failed");
  const int st = prog->p.build(n, dtype, gates, ngates, fuse_k, diag_k);
  if (st != QSV_OK) {
    delete prog;
    return st;
  }
  *out = prog;
  return QSV_OK;
}

int qsv_tiled_jit_compile_check(int n, int dtype, const qsv_gate* gates, size_t ngates,
                                size_t* code_bytes, int64_t* npasses, uint64_t* source_hash) {
  return qsv_tiled_jit_compile_check_form(n, dtype, gates, ngates, QSV_TILED_JIT, code_bytes, npasses,
                                          source_hash);
}

int qsv_tiled_jit_compile_check_form(int n, int dtype, const qsv_gate* gates, size_t ngates, int fuse_k,
                                     size_t* code_bytes, int64_t* npasses, uint64_t* source_hash) {
  if (fuse_k != QSV_TILED_JIT && fuse_k != QSV_TILED_JIT_FIXED)
    return fail(QSV_EINVAL, "fuse_k must be QSV_TILED_JIT or QSV_TILED_JIT_FIXED");
  if (dtype != QSV_C64 && dtype != QSV_C128) return fail(QSV_EDTYPE, "dtype must be QSV_C64 or QSV_C128");
  if (n < kQTileBits || n > 40) return fail(QSV_EINVAL, "tiled passes need n in [12, 40]");
  if (ngates && !gates) return fail(QSV_EINVAL, "null gate array");
  std::vector<KOp> lowered(ngates);
  for (size_t i = 0; i < ngates; ++i) QSV_TRY(lower_gate(gates[i], n, lowered[i]));
  QProgram qp;
  // as Program::build does on a device with room for the work buffer
  QSV_TRY(compile_qtile(lowered, n, dtype == QSV_C128 ? 3 : 4, qp,
                       jit_tile_bits(dtype, n, fuse_k == QSV_TILED_JIT_FIXED), true, true));
  qtile_scale_rotations(qp);  // as Program::build does
  qp.fixed_forms = fuse_k == QSV_TILED_JIT_FIXED;
  size_t total = 0;
  uint64_t h = 1469598103934665603ull;  // FNV-1a over every pass's source
  for (size_t i = 0; i < qp.passes.size(); ++i) {
    const std::string src = qtile_jit_source(qp, dtype == QSV_C128, i, i + 1);
    for (unsigned char ch : src) h = (h ^ ch) * 1099511628211ull;
    if (!code_bytes) continue;
    std::vector<char> code;
    QSV_TRY(qtile_jit_compile(src, code));
    total += code.size();
  }
  if (code_bytes) *code_bytes = total;
  if (source_hash) *source_hash = h;
  if (npasses) *npasses = int64_t(qp.passes.size());
  return QSV_OK;
}

int qsv_program_run(const qsv_program* prog, void* psi, void* stream) {
  if (!prog) return fail(QSV_EINVAL, "null program");
  QSV_TRY(check_state_args(psi, prog->p.n, prog->p.dtype));
  return prog->p.run(psi, static_cast<hipStream_t>(stream));
}

int qsv_program_info_get(const qsv_program* prog, qsv_program_info* out) {
  if (!prog || !out) return fail(QSV_EINVAL, "null argument");
  prog->p.info(out);
  return QSV_OK;
}

int qsv_program_op(const qsv_program* prog, int i, int* k, int* qubits, int* diag,
                   double* coeffs) {
  if (!prog) return fail(QSV_EINVAL, "null program");
  if (i < 0 || size_t(i) >= prog->p.ops.size()) return fail(QSV_EINVAL, "launch index out of range");
  const KOp& op = prog->p.ops[size_t(i)];
  std::vector<int> q(op.q, op.q + op.nq);
  std::sort(q.begin(), q.end());
  const int kk = int(q.size());
  if (k) *k = kk;
  if (qubits)
    for (int j = 0; j < kk; ++j) qubits[j] = q[j];
  const bool dg = is_diagonal(op);
  if (diag) *diag = dg ? 1 : 0;
  if (!coeffs) return QSV_OK;
  if (op.kind == K_DENSEK || op.kind == K_DIAGK) {
    const Block& b = prog->p.blocks[op.blk];
    for (size_t j = 0; j < b.m.size(); ++j) {
      coeffs[2 * j] = b.m[j].real();
      coeffs[2 * j + 1] = b.m[j].imag();
    }
    return QSV_OK;
  }
  KOp local = op;  // renumber to the op's own qubits
  for (int j = 0; j < op.nq; ++j)
    local.q[j] = int(std::lower_bound(q.begin(), q.end(), op.q[j]) - q.begin());
  const int D = 1 << kk;
  std::vector<cd> col(D);
  for (int c = 0; c < D; ++c) {
    std::fill(col.begin(), col.end(), cd(0.0, 0.0));
    col[c] = 1.0;
    apply_op_host(local, kk, col.data());
    if (dg) {
      coeffs[2 * c] = col[c].real();
      coeffs[2 * c + 1] = col[c].imag();
    } else {
      for (int r = 0; r < D; ++r) {
        coeffs[2 * (r * D + c)] = col[r].real();
        coeffs[2 * (r * D + c) + 1] = col[r].imag();
      }
    }
  }
  return QSV_OK;
}

int qsv_program_destroy(qsv_program* prog) {
  delete prog;
  return QSV_OK;
}

int qsv_host_apply_gates(void* host, int n, int dtype, const qsv_gate* gates, size_t ngates) {
  QSV_TRY(check_state_args(host, n, dtype));
  if (ngates && !gates) return fail(QSV_EINVAL, "null gate array");
  std::vector<KOp> ops(ngates);
  for (size_t i = 0; i < ngates; ++i) QSV_TRY(lower_gate(gates[i], n, ops[i]));
  const size_t bytes = (size_t(1) << n) * elem_bytes(dtype);
  std::lock_guard<std::mutex> lk(g_cache.mu);
  QSV_TRY(reserve_cache(bytes));
  QSV_TRY(upload(g_cache.ptr, host, bytes));
  for (const KOp& op : ops) QSV_TRY(apply_lowered(op, g_cache.ptr, n, dtype, nullptr));
  QSV_CHECK_HIP(hipStreamSynchronize(nullptr));
  QSV_TRY(download(host, g_cache.ptr, bytes));
  return QSV_OK;
}

int qsv_dev_simulate(void* psi, int n, int dtype, const qsv_gate* gates, size_t ngates,
                     int fuse_k, void* stream) {
  QSV_TRY(check_state_args(psi, n, dtype));
  if (ngates && !gates) return fail(QSV_EINVAL, "null gate array");
  QSV_TRY(check_fuse_k(fuse_k));
  Program prog;
  QSV_TRY(prog.build(n, dtype, gates, ngates, fuse_k, 0));
  auto st = static_cast<hipStream_t>(stream);
  QSV_TRY(prog.run(psi, st));
  QSV_CHECK_HIP(hipStreamSynchronize(st));  // the program's coefficients are freed on return
  return QSV_OK;
}

int qsv_host_simulate(void* host, int n, int dtype, const qsv_gate* gates, size_t ngates,
                      int fuse_k) {
  QSV_TRY(check_state_args(host, n, dtype));
  if (ngates && !gates) return fail(QSV_EINVAL, "null gate array");
  QSV_TRY(check_fuse_k(fuse_k));
  Program prog;
  QSV_TRY(prog.build(n, dtype, gates, ngates, fuse_k, 0));
  const size_t bytes = (size_t(1) << n) * elem_bytes(dtype);
  std::lock_guard<std::mutex> lk(g_cache.mu);
  QSV_TRY(reserve_cache(bytes));
  QSV_TRY(upload(g_cache.ptr, host, bytes));
  QSV_TRY(prog.run(g_cache.ptr, nullptr));
  QSV_CHECK_HIP(hipStreamSynchronize(nullptr));
  QSV_TRY(download(host, g_cache.ptr, bytes));
  return QSV_OK;
}

int qsv_dev_circuit_matrix(void* mat, int k, int dtype, const qsv_gate* gates, size_t ngates,
                           int fuse_k, void* stream) {
  if (!mat) return fail(QSV_EINVAL, "null matrix pointer");
  QSV_TRY(check_state_args(mat, 2 * k, dtype));
  std::vector<qsv_gate> recs;
  QSV_TRY(matrix_records(k, gates, ngates, recs));
  auto st = static_cast<hipStream_t>(stream);
  QSV_TRY(launch_set_identity(mat, k, dtype, st));
  if (recs.empty()) {
    QSV_CHECK_HIP(hipStreamSynchronize(st));
    return QSV_OK;
  }
  return qsv_dev_simulate(mat, 2 * k, dtype, recs.data(), recs.size(), fuse_k, stream);
}

int qsv_host_circuit_matrix(void* mat, int k, int dtype, const qsv_gate* gates, size_t ngates,
                            int fuse_k) {
  if (!mat) return fail(QSV_EINVAL, "null matrix pointer");
  QSV_TRY(check_state_args(mat, 2 * k, dtype));
  QSV_TRY(check_fuse_k(fuse_k));
  std::vector<qsv_gate> recs;
  QSV_TRY(matrix_records(k, gates, ngates, recs));
  Program prog;
  QSV_TRY(prog.build(2 * k, dtype, recs.data(), recs.size(), fuse_k, 0));
  const size_t bytes = (size_t(1) << (2 * k)) * elem_bytes(dtype);
  std::lock_guard<std::mutex> lk(g_cache.mu);
  QSV_TRY(reserve_cache(bytes));
  QSV_TRY(launch_set_identity(g_cache.ptr, k, dtype, nullptr));
  if (!recs.empty()) QSV_TRY(prog.run(g_cache.ptr, nullptr));
  QSV_CHECK_HIP(hipStreamSynchronize(nullptr));
  QSV_TRY(download(mat, g_cache.ptr, bytes));
  return QSV_OK;
}

int qsv_dev_nrmf(const void* a, const void* b, uint64_t count, int dtype, double* out, void* stream) {
  return dev_nrmf(a, b, count, dtype, out, static_cast<hipStream_t>(stream));
}

int qsv_circuit_nrmf(int k, int dtype, const qsv_gate* gates1, size_t ngates1, const qsv_gate* gates2,
                     size_t ngates2, int fuse_k, double* out) {
  if (!out) return fail(QSV_EINVAL, "null result pointer");
  if (dtype != QSV_C64 && dtype != QSV_C128) return fail(QSV_EDTYPE, "dtype must be QSV_C64 or QSV_C128");
  if (k < 1 || 2 * k > 40) return fail(QSV_EINVAL, "matrix qubits must be in [1, 20]");
  const size_t bytes = (size_t(1) << (2 * k)) * elem_bytes(dtype);
  void* m[2] = {nullptr, nullptr};
  for (int i = 0; i < 2; ++i)
    if (hipMalloc(&m[i], bytes) != hipSuccess) {
      (void)hipGetLastError();
      if (m[0]) (void)hipFree(m[0]);
      return fail(QSV_ENOMEM, "nrmF: device allocation of two " + std::to_string(bytes) + "-byte matrices failed");
    }
  int rc = qsv_dev_circuit_matrix(m[0], k, dtype, gates1, ngates1, fuse_k, nullptr);
  if (rc == QSV_OK) rc = qsv_dev_circuit_matrix(m[1], k, dtype, gates2, ngates2, fuse_k, nullptr);
  if (rc == QSV_OK) rc = dev_nrmf(m[0], m[1], uint64_t(1) << (2 * k), dtype, out, nullptr);
  (void)hipFree(m[0]);
  (void)hipFree(m[1]);
  return rc;
}

int qsv_cache_stats_get(qsv_cache_stats* out) {
  if (!out) return fail(QSV_EINVAL, "null argument");
  qtile_schedule_cache_stats(&out->schedule_hits, &out->schedule_misses);
  qtile_jit_cache_stats(&out->module_hits, &out->module_disk_hits, &out->module_compiles);
  return QSV_OK;
}

int qsv_release_cache(void) {
  std::lock_guard<std::mutex> lk(g_cache.mu);
  if (g_cache.ptr) QSV_CHECK_HIP(hipFree(g_cache.ptr));
  g_cache.ptr = nullptr;
  g_cache.bytes = 0;
  QSV_TRY(stage_release());
  return QSV_OK;
}

}  // extern "C"
