// C ABI: expectation values, Pauli-sum products and the adjoint gradient.
//
// Adjoint differentiation (Jones & Gacon, arXiv:2009.02823): with
// phi = U_N..U_1 psi0 and lam = H phi, sweep k = N..1:
//   dE/dtheta_k from <lam|G_k|phi>, then phi <- U_k^H phi, lam <- U_k^H lam,
// where the un-apply U_k^H is the gate applied with Op ConjTrans -- the
// reference's own inverse (QCircuit::apply with Op::ConjTrans reverses the
// gate order and conjugate-transposes each gate, include/qclab/QCircuit.hpp:137-141).
// States are recomputed by inversion instead of being stored: the sweep needs
// two state buffers whatever the depth.
//
// The reverse sweep runs as LDS tile passes (k_adjoint_tiles): the inverse
// gates, last first, are scheduled like a circuit, and within a pass each
// parametrised gate's inner product is taken over every tile just before the
// gate's inverse is applied there, so a pass over both states (4 S of HBM
// traffic) covers many gates.
//
// Rotations exp(-i theta P / 2) give dE/dtheta = Im<lam|P|phi>; Phase/CPhase
// diag(1, e^{i theta}) give dE/dtheta = -2 Im<lam|Pi_1|phi> (Pi_1 projects the
// target, and the control, onto the phased state).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <future>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "fusion.hpp"
#include "observables.hpp"
#include "plain.hpp"
#include "program.hpp"
#include "qtile.hpp"
#include "qtile_jit.hpp"
#include "tile.hpp"

namespace qsv {
namespace {

// Device scratch for reduction partials and term tables, reused between
// calls (callers are serialised by the mutex).
struct Scratch {
  std::mutex mu;
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return QSV_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, need) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return fail(QSV_ENOMEM, "scratch allocation of " + std::to_string(need) + " bytes failed");
    }
    bytes = need;
    return QSV_OK;
  }
} g_scratch;

// Tile bits of the adjoint's LDS passes: two tiles (phi and lam) of 32 KiB
// each, plus the row accumulators, leave room for two blocks per CU.
constexpr int kAdjTileBits128 = 11;
constexpr int kAdjTileBits64 = 12;

// A device copy of a host array, freed with the object.
template <typename T>
struct DevArray {
  T* p = nullptr;
  ~DevArray() {
    if (p) (void)hipFree(p);
  }
  int upload(const std::vector<T>& v) {
    if (v.empty()) return QSV_OK;
    if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return fail(QSV_ENOMEM, "device allocation of the adjoint pass tables failed");
    }
    QSV_CHECK_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return QSV_OK;
  }
};

int check_state(const void* psi, int n, int dtype) {
  if (!psi) return fail(QSV_EINVAL, "null state pointer");
  if (dtype != QSV_C64 && dtype != QSV_C128) return fail(QSV_EDTYPE, "dtype must be QSV_C64 or QSV_C128");
  if (n < 1 || n > 40) return fail(QSV_EINVAL, "number of qubits must be in [1, 40]");
  return QSV_OK;
}

int terms_to_ops(const qsv_pauli_term* terms, size_t nterms, int n, std::vector<PauliOp>& ops) {
  if (nterms && !terms) return fail(QSV_EINVAL, "null term array");
  ops.resize(nterms);
  for (size_t i = 0; i < nterms; ++i) QSV_TRY(pauli_from_term(terms[i], n, ops[i]));
  return QSV_OK;
}

template <typename R>
int inner(const void* a, const void* b, int n, const PauliOp& g, void* partials, hipStream_t st) {
  return launch_pauli_inner<R>(a, b, n, g, partials, st);
}

int inner_any(int dtype, const void* a, const void* b, int n, const PauliOp& g, void* partials,
              hipStream_t st) {
  return dtype == QSV_C128 ? inner<double>(a, b, n, g, partials, st)
                           : inner<float>(a, b, n, g, partials, st);
}

int plain_any(int dtype, const KOp& op, void* psi, int n, hipStream_t st) {
  PlainTuning tu;
  return dtype == QSV_C128 ? launch_plain<double>(op, psi, n, st, tu)
                           : launch_plain<float>(op, psi, n, st, tu);
}

// Generator of a differentiable gate and the weight turning z = <lam|G|phi>
// into dE/dtheta = wr*Re(z) + wi*Im(z).
bool generator(const qsv_gate& g, int n, PauliOp& G, double& wr, double& wi) {
  G = PauliOp();
  auto bit = [n](int q) { return uint64_t(1) << (n - 1 - q); };
  const uint64_t b0 = bit(g.q[0]);
  const uint64_t b1 = g.kind >= QSV_GATE_CX ? bit(g.q[1]) : 0;
  wr = 0;
  wi = 1;  // rotations: Im<lam|P|phi> (times the P's own phase below)
  switch (g.kind) {
    case QSV_GATE_RX: G.flip = b0; break;
    case QSV_GATE_RY: G.flip = b0; G.zmask = b0; G.phase_re = 0; G.phase_im = 1; break;
    case QSV_GATE_RZ: G.zmask = b0; break;
    case QSV_GATE_RXX: G.flip = b0 | b1; break;
    case QSV_GATE_RYY: G.flip = b0 | b1; G.zmask = b0 | b1; G.phase_re = -1; break;
    case QSV_GATE_RZZ: G.zmask = b0 | b1; break;
    case QSV_GATE_CRX: case QSV_GATE_CRY: case QSV_GATE_CRZ:
      G.pmask = b0;
      G.pval = g.ctrl_state ? b0 : 0;
      if (g.kind == QSV_GATE_CRX) G.flip = b1;
      if (g.kind == QSV_GATE_CRY) { G.flip = b1; G.zmask = b1; G.phase_re = 0; G.phase_im = 1; }
      if (g.kind == QSV_GATE_CRZ) G.zmask = b1;
      break;
    case QSV_GATE_P:
      G.pmask = b0; G.pval = b0; wi = -2;
      break;
    case QSV_GATE_CP:
      G.pmask = b0 | b1; G.pval = (g.ctrl_state ? b0 : 0) | b1; wi = -2;
      break;
    default:
      return false;
  }
  return true;
}

static_assert(kAdjPartStride == kRedBlocks, "compiled adjoint passes write reduction-row partials");

// Phase timing of the last adjoint call (HIP events on its stream).
struct AdjProfile {
  std::mutex mu;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int dev = -1;
  bool valid = false;
  int64_t reverse_launches = 0;
  int compiled = 0;
  int record(int i, hipStream_t st) {
    int d = 0;
    QSV_CHECK_HIP(hipGetDevice(&d));
    if (d != dev) {
      for (hipEvent_t& e : ev)
        if (e) (void)hipEventDestroy(e), e = nullptr;
      dev = d;
    }
    if (!ev[i]) QSV_CHECK_HIP(hipEventCreate(&ev[i]));
    QSV_CHECK_HIP(hipEventRecord(ev[i], st));
    return QSV_OK;
  }
} g_adjprof;

// Tile bits of the compiled reverse sweep: one shared LDS tile of 2^12
// amplitudes (64 KiB for complex128) that both states pass through.
constexpr int kAdjJitTileBits = 12;

// Which reverse sweep runs: compiled two-state passes where hiprtc loads
// and the state holds a tile (QSV_ADJOINT=interp forces the interpreted
// LDS passes, =jit requires the compiled ones).
int adjoint_form(int n) {
  const char* e = std::getenv("QSV_ADJOINT");
  const std::string f = e ? e : "auto";
  if (f == "interp") return 0;
  const bool ok = n >= kAdjJitTileBits && qtile_jit_available();
  if (f == "jit" && !ok) return -1;
  return ok ? 1 : 0;
}

// The reverse sweep as compiled passes over both states (the inverse gates,
// last first, scheduled as tiled passes; each parametrised gate's generator
// is measured just before its inverse, see Q_GEN).  Appends one result row
// per parametrised gate: its parameter, and the scale |lam|^2 the pass's
// scaled rotations took out of both states before it.  Row partials go to
// partials[(row0 + r) * kRedBlocks + block].
int adjoint_schedule(int n, int dtype, const std::vector<qsv_gate>& gs, QProgram& qp,
                     std::vector<int>& pass_row0, std::vector<int>& row_param,
                     std::vector<double>& row_scale) {
  std::vector<KOp> rev;
  rev.reserve(gs.size());
  for (size_t i = gs.size(); i-- > 0;) {
    qsv_gate c = gs[i];
    c.op = QSV_OP_C;
    KOp o;
    QSV_TRY(lower_gate(c, n, o));
    if (gs[i].param >= 0) {
      PauliOp G;
      double wr = 0, wi = 0;
      if (!generator(gs[i], n, G, wr, wi))
        return fail(QSV_EUNSUP, "gate kind " + std::to_string(gs[i].kind) + " is not differentiable");
      o.grow = int(i);
      o.gflip = G.flip;
      o.gz = G.zmask;
      o.gpm = G.pmask;
      o.gpv = G.pval;
      o.gph_re = G.phase_re;
      o.gph_im = G.phase_im;
      o.gw_re = wr;
      o.gw_im = wi;
    }
    rev.push_back(o);
  }
  QSV_TRY(compile_qtile(rev, n, dtype == QSV_C128 ? 3 : 4, qp, kAdjJitTileBits, true));
  qtile_scale_rotations(qp);
  // rows numbered pass by pass, so each pass writes a contiguous block
  pass_row0.assign(qp.passes.size(), 0);
  int rows = 0;
  for (size_t i = 0; i < qp.passes.size(); ++i) {
    const QPass& P = qp.passes[i];
    pass_row0[i] = rows;
    int local = 0;
    for (int o = qp.phases[P.phase_begin].op_begin; o < qp.phases[P.phase_end - 1].op_end; ++o) {
      QOp& x = qp.ops[o];
      if (x.kind != Q_GEN) continue;
      row_param.push_back(gs[size_t(x.a)].param);
      row_scale.push_back(x.m[0]);
      x.a = local++;
    }
    rows += local;
  }
  return QSV_OK;
}

// The host half of the compiled reverse sweep (schedule, op table and
// modules), built on a worker thread while the forward pass and H|phi>
// run on the device.  ok stays false when a pass holds more generator rows
// than LDS does (the caller then runs the interpreted sweep).
struct AdjReversePlan {
  QProgram qp;
  std::vector<int> pass_row0, row_param;
  std::vector<double> row_scale;
  QJit qj;
  bool ok = false;
  int rc = QSV_OK;
  std::string err;
};

int adjoint_reverse_prepare(int dev, int n, int dtype, const std::vector<qsv_gate>& gs,
                            AdjReversePlan& plan) {
  QSV_CHECK_HIP(hipSetDevice(dev));  // module loads go to the caller's device
  QSV_TRY(adjoint_schedule(n, dtype, gs, plan.qp, plan.pass_row0, plan.row_param, plan.row_scale));
  if (qtile_jit_adjoint_check(plan.qp) != QSV_OK) {
    plan.row_param.clear();
    plan.row_scale.clear();
    return QSV_OK;
  }
  QSV_TRY(qtile_jit_build_adjoint(plan.qp, dtype == QSV_C128, plan.qj));
  plan.ok = true;
  return QSV_OK;
}

int adjoint_reverse_jit(void* psi, void* lam, const AdjReversePlan& plan, double2* partials, int row0,
                        hipStream_t st, int64_t& launches) {
  DevArray<QOp> d_ops;
  QSV_TRY(d_ops.upload(plan.qp.ops));
  QSV_TRY(g_adjprof.record(2, st));
  for (size_t i = 0; i < plan.qp.passes.size(); ++i) {
    double* part = reinterpret_cast<double*>(partials + size_t(row0 + plan.pass_row0[i]) * kRedBlocks);
    QSV_TRY(qtile_jit_launch_adjoint(plan.qj, i, psi, lam, d_ops.p, part, st));
  }
  QSV_TRY(g_adjprof.record(3, st));
  launches = int64_t(plan.qp.passes.size());
  // the op table is freed with d_ops: finish before returning
  QSV_CHECK_HIP(hipStreamSynchronize(st));
  return QSV_OK;
}

// Device scratch an expectation needs for `rows` terms: per-term block
// partials, the reduced rows, and the diagonal-term table of one pass.
inline size_t expectation_scratch_bytes(size_t rows) {
  return rows * kRedBlocks * sizeof(double2) + rows * sizeof(double2) + 64 * sizeof(uint64_t);
}

// The expectation values with the caller's scratch (expectation_scratch_bytes
// at `scratch`, every byte of which is written before it is read).
template <typename R>
int expectation_core(const void* psi, int n, const std::vector<PauliOp>& ops, double* values,
                     void* scratch, hipStream_t st) {
  // diagonal, projector-free terms: 32 per pass; others: one pass each
  std::vector<int> diag, off;
  for (int i = 0; i < int(ops.size()); ++i)
    (ops[i].flip == 0 && ops[i].pmask == 0 ? diag : off).push_back(i);
  const size_t rows = ops.size();
  auto* partials = static_cast<double2*>(scratch);
  auto* red = partials + rows * kRedBlocks;
  void* tmp = red + rows;
  // rows are laid out in the order: diagonal chunks, then off-diagonal terms
  std::vector<int> row_of(ops.size());
  size_t r = 0;
  for (size_t c = 0; c < diag.size(); c += 32) {
    std::vector<PauliOp> chunk;
    for (size_t j = c; j < std::min(diag.size(), c + 32); ++j) {
      chunk.push_back(ops[diag[j]]);
      row_of[diag[j]] = int(r + (j - c));
    }
    QSV_TRY(launch_diag_expect<R>(psi, n, chunk, partials + r * kRedBlocks, tmp, st));
    r += chunk.size();
  }
  for (int i : off) {
    row_of[i] = int(r);
    QSV_TRY(launch_pauli_inner<R>(psi, psi, n, ops[i], partials + r * kRedBlocks, st));
    ++r;
  }
  QSV_TRY(launch_reduce_rows(partials, int(rows), red, st));
  std::vector<double2> host(rows);
  QSV_CHECK_HIP(hipMemcpyAsync(host.data(), red, rows * sizeof(double2), hipMemcpyDeviceToHost, st));
  QSV_CHECK_HIP(hipStreamSynchronize(st));
  for (size_t i = 0; i < ops.size(); ++i) {
    const double2 z = host[row_of[i]];
    // <P> = phase * z (real for a Hermitian P)
    values[i] = ops[i].phase_re * z.x - ops[i].phase_im * z.y;
  }
  return QSV_OK;
}

template <typename R>
int expectation_impl(const void* psi, int n, const std::vector<PauliOp>& ops, double* values,
                     hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_scratch.mu);
  QSV_TRY(g_scratch.ensure(expectation_scratch_bytes(ops.size())));
  return expectation_core<R>(psi, n, ops, values, g_scratch.p, st);
}

// ---- the allocation-free expectation ------------------------------------
// A fixed scratch in the library's own device memory (one copy per GPU,
// held until the process ends), for callers that cannot allocate.  Host
// code never uses the array's name as a device address: it takes the
// address hipGetSymbolAddress reports, once per device, with that device
// current.
constexpr int kNoAllocMaxTerms = 256;
constexpr size_t kNoAllocScratchBytes = (size_t(kNoAllocMaxTerms) * kRedBlocks + kNoAllocMaxTerms) * 16 + 64 * 8;
static_assert(kNoAllocScratchBytes == (size_t(kNoAllocMaxTerms) * kRedBlocks + kNoAllocMaxTerms) * sizeof(double2) +
                                          64 * sizeof(uint64_t),
              "scratch layout of expectation_core");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocScratch {
  std::mutex mu;
  void* ptr[64] = {};  // device address of `scratch`, per device
  int get(void** out) {
    int dev = 0;
    QSV_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(QSV_ENODEV, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return fail(QSV_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      }
      ptr[dev] = p;
    }
    *out = ptr[dev];
    return QSV_OK;
  }
} g_noalloc;

}  // namespace
}  // namespace qsv

using namespace qsv;

extern "C" {

int qsv_dev_expectation(const void* psi, int n, int dtype, const qsv_pauli_term* terms,
                        size_t nterms, double* values, double* total, void* stream) {
  QSV_TRY(check_state(psi, n, dtype));
  std::vector<PauliOp> ops;
  QSV_TRY(terms_to_ops(terms, nterms, n, ops));
  std::vector<double> v(nterms);
  auto st = static_cast<hipStream_t>(stream);
  if (nterms)
    QSV_TRY(dtype == QSV_C128 ? expectation_impl<double>(psi, n, ops, v.data(), st)
                              : expectation_impl<float>(psi, n, ops, v.data(), st));
  double s = 0;
  for (size_t i = 0; i < nterms; ++i) {
    if (values) values[i] = v[i];
    s += ops[i].coeff * v[i];
  }
  if (total) *total = s;
  return QSV_OK;
}

int qsv_dev_expectation_noalloc(const void* psi, int n, int dtype, const qsv_pauli_term* terms,
                                size_t nterms, double* values, double* total, void* stream) {
  QSV_TRY(check_state(psi, n, dtype));
  if (nterms > size_t(kNoAllocMaxTerms))
    return fail(QSV_EINVAL, "qsv_dev_expectation_noalloc takes at most " + std::to_string(kNoAllocMaxTerms) +
                                " terms (got " + std::to_string(nterms) + ")");
  std::vector<PauliOp> ops;
  QSV_TRY(terms_to_ops(terms, nterms, n, ops));
  void* scr = nullptr;
  QSV_TRY(g_noalloc.get(&scr));
  std::vector<double> v(nterms);
  auto st = static_cast<hipStream_t>(stream);
  if (nterms)
    QSV_TRY(dtype == QSV_C128 ? expectation_core<double>(psi, n, ops, v.data(), scr, st)
                              : expectation_core<float>(psi, n, ops, v.data(), scr, st));
  double s = 0;
  for (size_t i = 0; i < nterms; ++i) {
    if (values) values[i] = v[i];
    s += ops[i].coeff * v[i];
  }
  if (total) *total = s;
  return QSV_OK;
}

int qsv_dev_inner(const void* a, const void* b, int n, int dtype, double* re, double* im,
                  void* stream) {
  QSV_TRY(check_state(a, n, dtype));
  QSV_TRY(check_state(b, n, dtype));
  auto st = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_scratch.mu);
  QSV_TRY(g_scratch.ensure((kRedBlocks + 1) * sizeof(double2)));
  auto* partials = static_cast<double2*>(g_scratch.p);
  QSV_TRY(inner_any(dtype, a, b, n, PauliOp(), partials, st));
  QSV_TRY(launch_reduce_rows(partials, 1, partials + kRedBlocks, st));
  double2 z;
  QSV_CHECK_HIP(hipMemcpyAsync(&z, partials + kRedBlocks, sizeof(z), hipMemcpyDeviceToHost, st));
  QSV_CHECK_HIP(hipStreamSynchronize(st));
  if (re) *re = z.x;
  if (im) *im = z.y;
  return QSV_OK;
}

int qsv_dev_apply_pauli_sum(const void* in, void* out, int n, int dtype,
                            const qsv_pauli_term* terms, size_t nterms, void* stream) {
  QSV_TRY(check_state(in, n, dtype));
  QSV_TRY(check_state(out, n, dtype));
  if (in == out) return fail(QSV_EINVAL, "input and output must be different buffers");
  std::vector<PauliOp> ops;
  QSV_TRY(terms_to_ops(terms, nterms, n, ops));
  if (ops.empty()) return fail(QSV_EINVAL, "empty Pauli sum");
  if (ops.size() > size_t(kMaxPauliTerms)) return fail(QSV_EINVAL, "Pauli sum has too many terms");
  std::vector<DevPauli> dt;
  std::vector<int> gs;
  group_terms(ops, dt, gs);
  auto st = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_scratch.mu);
  const size_t tb = dt.size() * sizeof(DevPauli);
  QSV_TRY(g_scratch.ensure(tb + gs.size() * sizeof(int)));
  void* dterms = g_scratch.p;
  void* dgs = static_cast<unsigned char*>(g_scratch.p) + tb;
  QSV_CHECK_HIP(hipMemcpyAsync(dterms, dt.data(), tb, hipMemcpyHostToDevice, st));
  QSV_CHECK_HIP(hipMemcpyAsync(dgs, gs.data(), gs.size() * sizeof(int), hipMemcpyHostToDevice, st));
  const int ng = int(gs.size()) - 1;
  // Sums whose terms flip at most one qubit take the tiled passes once the
  // state is well beyond the 256 MB last-level cache (measured at n=30
  // complex128: 46 against 53 ms); a smaller state serves the per-flip-mask
  // gathers from that cache (n=24: 4.5 against 4.9 ms for the call).
  // QSV_PAULI_TILES=1 takes the tiled passes from 12 qubits up, =0 never.
  const size_t state_bytes = (size_t(1) << n) * (dtype == QSV_C128 ? 16 : 8);
  const char* tenv = std::getenv("QSV_PAULI_TILES");
  const bool tiles = tenv && tenv[0] == '1' ? n >= 12 : !(tenv && tenv[0] == '0') && state_bytes > (size_t(1) << 30);
  if (tiles && pauli_flips_single(dt, gs))
    QSV_TRY(dtype == QSV_C128 ? launch_pauli_apply_tiles<double>(in, out, n, dt, gs, dterms, dgs, st)
                              : launch_pauli_apply_tiles<float>(in, out, n, dt, gs, dterms, dgs, st));
  else
    QSV_TRY(dtype == QSV_C128
                ? launch_pauli_apply<double>(in, out, n, dterms, dgs, ng, int(dt.size()), st)
                : launch_pauli_apply<float>(in, out, n, dterms, dgs, ng, int(dt.size()), st));
  QSV_CHECK_HIP(hipStreamSynchronize(st));  // host term tables go out of scope
  return QSV_OK;
}

int qsv_adjoint_profile_get(qsv_adjoint_profile* out) {
  if (!out) return fail(QSV_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(g_adjprof.mu);
  if (!g_adjprof.valid) return fail(QSV_EINVAL, "no adjoint call to report");
  float ms[3];
  for (int i = 0; i < 3; ++i) QSV_CHECK_HIP(hipEventElapsedTime(&ms[i], g_adjprof.ev[i], g_adjprof.ev[i + 1]));
  out->forward_ms = ms[0];
  out->hamiltonian_ms = ms[1];
  out->reverse_ms = ms[2];
  out->reverse_launches = g_adjprof.reverse_launches;
  out->compiled = g_adjprof.compiled;
  return QSV_OK;
}

int qsv_adjoint_compile_check(int n, int dtype, const qsv_gate* gates, size_t ngates,
                              size_t* code_bytes, int64_t* npasses, int64_t* nrows) {
  if (dtype != QSV_C64 && dtype != QSV_C128) return fail(QSV_EDTYPE, "dtype must be QSV_C64 or QSV_C128");
  if (n < kAdjJitTileBits || n > 40) return fail(QSV_EINVAL, "compiled adjoint passes need n in [12, 40]");
  if (ngates && !gates) return fail(QSV_EINVAL, "null gate array");
  std::vector<qsv_gate> gs(gates, gates + ngates);
  for (const qsv_gate& g : gs)
    if (g.param >= 0 && g.op != QSV_OP_N) return fail(QSV_EUNSUP, "parametrised gates must use Op 'N'");
  QProgram qp;
  std::vector<int> pass_row0, rp;
  std::vector<double> rs;
  QSV_TRY(adjoint_schedule(n, dtype, gs, qp, pass_row0, rp, rs));
  QSV_TRY(qtile_jit_adjoint_check(qp));
  size_t total = 0;
  for (size_t i = 0; i < qp.passes.size(); ++i) {
    std::vector<char> code;
    QSV_TRY(qtile_jit_compile(qtile_jit_adjoint_source(qp, dtype == QSV_C128, i, i + 1), code));
    total += code.size();
  }
  if (code_bytes) *code_bytes = total;
  if (npasses) *npasses = int64_t(qp.passes.size());
  if (nrows) *nrows = int64_t(rp.size());
  return QSV_OK;
}

int qsv_dev_adjoint_gradient(void* psi, void* lam, int n, int dtype, const qsv_gate* gates,
                             size_t ngates, const double* theta, size_t nparams,
                             const qsv_pauli_term* terms, size_t nterms, int fuse_k,
                             double* energy, double* grad, void* stream) {
  QSV_TRY(check_state(psi, n, dtype));
  QSV_TRY(check_state(lam, n, dtype));
  if (psi == lam) return fail(QSV_EINVAL, "psi and lam must be different buffers");
  if (ngates && !gates) return fail(QSV_EINVAL, "null gate array");
  if (nparams && !grad) return fail(QSV_EINVAL, "null gradient output");
  auto st = static_cast<hipStream_t>(stream);
  // parameter overrides
  std::vector<qsv_gate> gs(gates, gates + ngates);
  for (auto& g : gs) {
    if (g.param >= 0) {
      if (size_t(g.param) >= nparams) return fail(QSV_EINVAL, "gate parameter index out of range");
      if (g.op != QSV_OP_N) return fail(QSV_EUNSUP, "parametrised gates must use Op 'N'");
      if (theta) {
        g.theta = theta[g.param];
        g.has_cs = 0;
      }
      // reject a gate without a generator before psi is touched
      PauliOp G;
      double a = 0, b = 0;
      if (!generator(g, n, G, a, b))
        return fail(QSV_EUNSUP, "gate kind " + std::to_string(g.kind) + " is not differentiable");
    }
  }
  const int form = adjoint_form(n);
  if (form < 0) return fail(QSV_EUNSUP, "QSV_ADJOINT=jit: compiled passes need hiprtc and n >= 12");
  // the compiled reverse sweep's host work overlaps everything up to its launch
  AdjReversePlan plan;
  std::future<void> prep;
  if (form == 1) {
    int dev = 0;
    QSV_CHECK_HIP(hipGetDevice(&dev));
    prep = std::async(std::launch::async, [&plan, &gs, dev, n, dtype] {
      plan.rc = adjoint_reverse_prepare(dev, n, dtype, gs, plan);
      if (plan.rc != QSV_OK) plan.err = get_error();
    });
  }
  // forward: phi = U psi0
  Program fwd;
  QSV_TRY(fwd.build(n, dtype, gs.data(), gs.size(), fuse_k, 0));
  std::lock_guard<std::mutex> plk(g_adjprof.mu);
  g_adjprof.valid = false;
  QSV_TRY(g_adjprof.record(0, st));
  QSV_TRY(fwd.run(psi, st));
  QSV_TRY(g_adjprof.record(1, st));
  // lam = H phi
  QSV_TRY(qsv_dev_apply_pauli_sum(psi, lam, n, dtype, terms, nterms, stream));
  // per-gate generators and un-apply ops
  std::vector<KOp> inv(gs.size());
  for (size_t k = 0; k < gs.size(); ++k) {
    qsv_gate c = gs[k];
    c.op = QSV_OP_C;
    QSV_TRY(lower_gate(c, n, inv[k]));
  }
  const size_t nrows = 1 + gs.size();  // energy row + at most one per gate
  std::lock_guard<std::mutex> lk(g_scratch.mu);
  QSV_TRY(g_scratch.ensure(nrows * (kRedBlocks + 1) * sizeof(double2)));
  auto* partials = static_cast<double2*>(g_scratch.p);
  auto* red = partials + nrows * kRedBlocks;
  // energy row: <phi|lam>
  QSV_TRY(inner_any(dtype, psi, lam, n, PauliOp(), partials, st));
  if (form == 1) {
    prep.get();
    if (plan.rc != QSV_OK) return fail(plan.rc, plan.err);
    if (plan.ok) {
      QSV_TRY(adjoint_reverse_jit(psi, lam, plan, partials, 1, st, g_adjprof.reverse_launches));
      g_adjprof.compiled = 1;
      g_adjprof.valid = true;
      const std::vector<int>& rp = plan.row_param;
      const std::vector<double>& rs = plan.row_scale;
      const int rows = 1 + int(rp.size());
      QSV_TRY(launch_reduce_rows(partials, rows, red, st));
      std::vector<double2> host(rows);
      QSV_CHECK_HIP(hipMemcpyAsync(host.data(), red, rows * sizeof(double2), hipMemcpyDeviceToHost, st));
      QSV_CHECK_HIP(hipStreamSynchronize(st));
      if (energy) *energy = host[0].x;
      for (size_t p = 0; p < nparams; ++p) grad[p] = 0;
      for (int r = 1; r < rows; ++r) grad[rp[r - 1]] += rs[r - 1] * host[r].x;
      return QSV_OK;
    }
  }
  int rows = 1;
  std::vector<int> row_param(1, -1);
  std::vector<double> row_wr(1, 0), row_wi(1, 0);
  // The reverse sweep in LDS tile passes: the inverse gates, last first,
  // scheduled like a batch circuit; within a pass each parametrised op
  // contributes <lam|G|phi> over every tile just before it is applied.
  std::vector<KOp> rev(gs.size());
  for (size_t k = 0; k < gs.size(); ++k) rev[gs.size() - 1 - k] = inv[k];
  std::vector<TPass> passes;
  std::vector<TOp> tops;
  std::vector<int> origin;
  if (!rev.empty())
    QSV_TRY(compile_tiled(rev, n, dtype == QSV_C128 ? kAdjTileBits128 : kAdjTileBits64, false, 0, passes, tops,
                          &origin));
  std::vector<TGen> gens(tops.size());
  std::vector<TPass> runs;        // passes split to at most kAdjRowsMax rows each
  std::vector<int> run_rows;      // rows of each run
  std::vector<int> run_row0;      // first global row of each run
  for (const TPass& P : passes) {
    auto split = [&](uint64_t mask, uint32_t& loc, uint64_t& out) {
      loc = 0;
      out = 0;
      for (int b = 0; b < n; ++b) {
        if (!(mask >> b & 1)) continue;
        int i = 0;
        while (i < P.t && P.tbits[i] != b) ++i;
        if (i < P.t) loc |= 1u << i;
        else out |= uint64_t(1) << b;
      }
    };
    TPass cur = P;
    int nr = 0;
    auto close = [&](int end) {
      cur.op_end = end;
      runs.push_back(cur);
      run_rows.push_back(nr);
      run_row0.push_back(rows - nr);
      cur.op_begin = end;
      nr = 0;
    };
    for (int j = P.op_begin; j < P.op_end; ++j) {
      const size_t kk = gs.size() - 1 - size_t(origin[j]);
      TGen& g = gens[j];
      g = TGen{};
      g.row = -1;
      if (gs[kk].param < 0) continue;
      PauliOp G;
      double a = 0, b = 0;
      if (!generator(gs[kk], n, G, a, b))
        return fail(QSV_EUNSUP, "gate kind " + std::to_string(gs[kk].kind) + " is not differentiable");
      if (nr == kAdjRowsMax) close(j);
      uint64_t fout = 0;
      split(G.flip, g.floc, fout);
      if (fout) return fail(QSV_EINVAL, "adjoint tiles: generator flip outside the tile");
      split(G.zmask, g.zloc, g.zout);
      split(G.pmask, g.ploc, g.pout);
      split(G.pval, g.pvloc, g.pvout);
      // the inner product fits inside the op's own loop when the flip stays
      // in the thread's pair or quad and the pairs a controlled op skips lie
      // outside the projector
      const TOp& t = tops[j];
      const uint32_t tgt = t.a >= 0 ? 1u << t.a : 0u;
      switch (t.kind) {
        case T_DENSE1:
          g.fused = g.floc == 0 || g.floc == tgt;
          break;
        case T_CDENSE1: {
          const bool proj = t.b >= 0 ? ((g.ploc >> t.b & 1u) && int(g.pvloc >> t.b & 1u) == t.cs)
                                     : ((g.pout >> ~t.b & 1u) && int(g.pvout >> ~t.b & 1u) == t.cs);
          g.fused = proj && (g.floc == 0 || g.floc == tgt);
          break;
        }
        case T_DIAG1:
        case T_DIAG2:
          g.fused = g.floc == 0;
          break;
        case T_DENSE2:
          g.fused = t.b >= 0 && (g.floc & ~(tgt | (1u << t.b))) == 0;
          break;
        default:
          g.fused = 0;
      }
      g.row = nr++;
      // fold the generator's own phase into the weights: z_true = phase * z
      row_param.push_back(gs[kk].param);
      row_wr.push_back(a * G.phase_re + b * G.phase_im);
      row_wi.push_back(-a * G.phase_im + b * G.phase_re);
      ++rows;
    }
    close(P.op_end);
  }
  // barriers between ops: needed where the thread that owns an element
  // changes (see TGen::sync), and before a separate inner-product sweep
  auto owner_map = [](const TOp& t) {  // -1: thread e % 256 owns element e
    constexpr int kOwnBits = 8;      // log2(kTileThreads)
    switch (t.kind) {
      case T_DENSE1:
      case T_CDENSE1:
        return t.a >= kOwnBits ? -1 : t.a;
      case T_DENSE2:
      case T_SWAP: {
        const int lo = std::min(t.a, t.b), hi = std::max(t.a, t.b);
        return lo >= kOwnBits ? -1 : 1000 + 64 * lo + hi;
      }
      default:
        return -1;
    }
  };
  for (const TPass& R : runs)
    for (int j = R.op_begin; j < R.op_end; ++j) {
      const bool last = j + 1 == R.op_end;
      const int next = last ? -1 : owner_map(tops[j + 1]);
      const bool sweep_next = !last && gens[j + 1].row >= 0 && !gens[j + 1].fused;
      gens[j].sync = owner_map(tops[j]) != next || sweep_next;
    }
  DevArray<TOp> d_tops;
  DevArray<TGen> d_gens;
  QSV_TRY(d_tops.upload(tops));
  QSV_TRY(d_gens.upload(gens));
  QSV_TRY(g_adjprof.record(2, st));
  for (size_t r = 0; r < runs.size(); ++r) {
    void* part = partials + size_t(run_row0[r]) * kRedBlocks;
    const TOp* o = d_tops.p + runs[r].op_begin;
    const TGen* g = d_gens.p + runs[r].op_begin;
    QSV_TRY(dtype == QSV_C128
                ? launch_adjoint_tiles<double>(psi, lam, runs[r], o, g, run_rows[r], part, kRedBlocks, st)
                : launch_adjoint_tiles<float>(psi, lam, runs[r], o, g, run_rows[r], part, kRedBlocks, st));
  }
  QSV_TRY(g_adjprof.record(3, st));
  g_adjprof.compiled = 0;
  g_adjprof.reverse_launches = int64_t(runs.size());
  g_adjprof.valid = true;
  QSV_TRY(launch_reduce_rows(partials, rows, red, st));
  std::vector<double2> host(rows);
  QSV_CHECK_HIP(hipMemcpyAsync(host.data(), red, rows * sizeof(double2), hipMemcpyDeviceToHost, st));
  QSV_CHECK_HIP(hipStreamSynchronize(st));
  if (energy) *energy = host[0].x;
  for (size_t p = 0; p < nparams; ++p) grad[p] = 0;
  for (int r = 1; r < rows; ++r) grad[row_param[r]] += row_wr[r] * host[r].x + row_wi[r] * host[r].y;
  return QSV_OK;
}

}  // extern "C"
