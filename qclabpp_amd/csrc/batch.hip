// Batched execution of many independent circuits of uneven size.
//
// States live in one packed device buffer (circuit c at amplitude offset
// sum_{c'<c} 2^n_c').  Every circuit is compiled into LDS tile passes
// (tile.cpp).  Circuits whose state fits one tile (n <= t_small) run whole in
// LDS in a single work item; larger ones run pass by pass over t_large-bit
// tiles.  One run = a prologue kernel, the small-circuit level on a second
// stream, and one launch per pass level of the large circuits, joined back --
// captured once into a hipGraph and replayed.  An autotuner times candidate
// (tile bits, fusion width) pairs on the caller's buffer and keeps the
// fastest.  QCLAB++ has no batching: each circuit is the
// QCircuit::simulate of |0..0> (include/qclab/QCircuit.hpp:164-166).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "fusion.hpp"
#include "tile.hpp"
#include "qtile.hpp"
#include "tile_launch.hpp"

namespace qsv {
namespace {

struct Level {
  size_t item_begin = 0;
  uint64_t nitems = 0;
  int tmax = 0;
  unsigned grid = 0;
};

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

template <typename T>
int upload(DevBuf& b, const std::vector<T>& v) {
  if (v.empty()) return QSV_OK;
  if (hipMalloc(&b.p, v.size() * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return fail(QSV_ENOMEM, "batch: device allocation failed");
  }
  QSV_CHECK_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return QSV_OK;
}

}  // namespace

struct Batch {
  int dtype = QSV_C64;
  int ncirc = 0;
  qsv_batch_opts opts{};
  std::vector<int> nq;
  std::vector<uint64_t> offsets;
  std::vector<TCircuit> circs;
  std::vector<TPass> passes;
  std::vector<TOp> ops;
  std::vector<TItem> items;
  // register-phase passes of the circuits above small_bits (opts.phases)
  std::vector<QPass> qpasses;
  std::vector<QPhase> qphases;
  std::vector<QOp> qops;
  std::vector<QPassInfo> qpinfo;
  std::vector<int32_t> large_phase_ids;  // circuits of the register-phase passes
  Level small;
  std::vector<Level> large;
  DevBuf d_circs, d_passes, d_ops, d_items, d_tickets, d_done, d_stamps;
  DevBuf d_qpasses, d_qphases, d_qops, d_qpinfo, d_zero_ids;
  // graph
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  void* graph_psi = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  std::mutex mu;

  ~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (side) (void)hipStreamDestroy(side);
  }

  int build(int dtype_, int n_circ, const int* nqubits, const qsv_gate* gates,
            const size_t* goff, const qsv_batch_opts& o);
  int enqueue(void* psi, hipStream_t st);
  int run(void* psi, hipStream_t st, bool use_graph);
};

static unsigned cus() {
  static int n = 0;
  if (!n) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) {
      (void)hipGetLastError();
      n = 256;
    }
  }
  return unsigned(n);
}

int Batch::build(int dtype_, int n_circ, const int* nqubits, const qsv_gate* gates,
                 const size_t* goff, const qsv_batch_opts& o) {
  dtype = dtype_;
  ncirc = n_circ;
  opts = o;
  const int eb = dtype == QSV_C128 ? 16 : 8;
  const int cap = dtype == QSV_C128 ? kTileMaxBits - 1 : kTileMaxBits;  // 128 KiB of LDS
  if (opts.tile_bits <= 0) opts.tile_bits = dtype == QSV_C128 ? 11 : 12;
  if (opts.small_bits <= 0) opts.small_bits = cap;
  if (opts.tile_bits < 6 || opts.tile_bits > cap) return fail(QSV_EINVAL, "tile_bits out of range");
  if (opts.small_bits < 1 || opts.small_bits > cap) return fail(QSV_EINVAL, "small_bits out of range");
  if (opts.fuse_k < 0 || opts.fuse_k > 2) return fail(QSV_EINVAL, "batch fuse_k must be 0, 1 or 2");
  if (opts.phases != 0 && opts.phases != 1) return fail(QSV_EINVAL, "batch phases must be 0 or 1");
  // register-phase passes take tiles of 2^kQTileBits: smaller circuits stay whole in LDS
  const int small_cut = opts.phases ? std::max(opts.small_bits, kQTileBits - 1) : opts.small_bits;
  nq.assign(nqubits, nqubits + ncirc);
  offsets.assign(ncirc + 1, 0);
  for (int c = 0; c < ncirc; ++c) {
    if (nq[c] < 1 || nq[c] > 34) return fail(QSV_EINVAL, "batch circuit qubit count out of range");
    offsets[c + 1] = offsets[c] + (uint64_t(1) << nq[c]);
  }
  std::vector<int> small_ids, large_ids;
  std::vector<int> first_pass(ncirc), npass(ncirc);
  for (int c = 0; c < ncirc; ++c) {
    const int n = nq[c];
    std::vector<KOp> lowered;
    for (size_t g = goff[c]; g < goff[c + 1]; ++g) {
      KOp k;
      QSV_TRY(lower_gate(gates[g], n, k));
      lowered.push_back(k);
    }
    const bool is_small = n <= small_cut;
    if (!is_small && opts.phases) {
      // passes and register phases of the interpreted tiled kernel, from the
      // unfused gates; descriptors appended to the batch's arrays
      QProgram qp;
      QSV_TRY(compile_qtile(lowered, n, dtype == QSV_C128 ? 3 : 4, qp, kQTileBits, false));
      qtile_scale_rotations(qp);
      const int pb = int(qpasses.size()), phb = int(qphases.size()), ob = int(qops.size());
      for (QPhase ph : qp.phases) {
        ph.op_begin += ob;
        ph.op_end += ob;
        qphases.push_back(ph);
      }
      const int np = int(qp.passes.size());
      for (int k = 0; k < np; ++k) {
        QPass P = qp.passes[k];
        P.phase_begin += phb;
        P.phase_end += phb;
        qpasses.push_back(P);
        qpinfo.push_back(QPassInfo{c, k == 0, k == np - 1, 0});
      }
      qops.insert(qops.end(), qp.ops.begin(), qp.ops.end());
      first_pass[c] = pb;
      npass[c] = np;
      circs.push_back(TCircuit{offsets[c], n, np});
      large_ids.push_back(c);
      continue;
    }
    std::vector<KOp> use;
    if (opts.fuse_k == 2) {
      FusionOptions fo;
      fo.max_k = 2;
      fo.diag_k = 2;
      std::vector<Block> blocks;
      fuse(lowered, n, fo, use, blocks);
    } else {
      use = lowered;
    }
    first_pass[c] = int(passes.size());
    QSV_TRY(compile_tiled(use, n, is_small ? n : opts.tile_bits, opts.fuse_k >= 1, c, passes, ops));
    npass[c] = int(passes.size()) - first_pass[c];
    circs.push_back(TCircuit{offsets[c], n, npass[c]});
    (is_small ? small_ids : large_ids).push_back(c);
  }
  // small level: one item per circuit, most work first
  auto cost = [&](int c) {
    const TPass& P = passes[first_pass[c]];
    return double(P.op_end - P.op_begin + 2) * std::ldexp(1.0, nq[c]);
  };
  std::sort(small_ids.begin(), small_ids.end(), [&](int a, int b) { return cost(a) > cost(b); });
  small.item_begin = items.size();
  for (int c : small_ids) {
    items.push_back(TItem{first_pass[c], 0, 0});
    small.tmax = std::max(small.tmax, nq[c]);
  }
  small.nitems = items.size() - small.item_begin;
  // large levels: pass k of every large circuit that has one
  int maxp = 0;
  for (int c : large_ids) maxp = std::max(maxp, npass[c]);
  // Register-phase passes start from |0...0>: before pass k only the
  // addresses whose set bits lie in `support` (the tile bits of passes
  // 0..k-1) can hold nonzero amplitudes.  A fill kernel zeroes the states
  // first, pass 0 computes only its tile 0 (the other tiles hold zeros
  // before and after it), and a later pass skips each tile whose base has a
  // bit outside the support: the tile holds zeros, and its gates (on tile
  // bits, or diagonal) keep them.
  std::vector<uint64_t> support(ncirc, 0);
  for (int k = 0; k < maxp; ++k) {
    Level L;
    L.item_begin = items.size();
    for (int c : large_ids) {
      if (k >= npass[c]) continue;
      const int pi = first_pass[c] + k;
      if (!opts.phases) {
        for (uint64_t t = 0; t < passes[pi].ntiles; ++t) items.push_back(TItem{pi, 0, t});
        L.tmax = std::max(L.tmax, passes[pi].t);
        continue;
      }
      const QPass& P = qpasses[pi];
      uint64_t tmask = 0;
      for (int i = 0; i < kQTileBits; ++i) tmask |= uint64_t(1) << P.tbits[i];
      const size_t before = items.size();
      for (uint64_t t = 0; t < P.ntiles; ++t) {
        if (k == 0) {
          if (t == 0) items.push_back(TItem{pi, 0, t});
          continue;
        }
        uint64_t base = 0;
        for (int r = 0, sh = 0; r < P.nruns; ++r) {
          base |= ((t >> sh) & ((uint64_t(1) << P.rlen[r]) - 1)) << P.rdst[r];
          sh += P.rlen[r];
        }
        if (base & ~support[c]) continue;
        items.push_back(TItem{pi, 0, t});
      }
      qpinfo[pi].nitems = int32_t(items.size() - before);
      support[c] |= tmask;
      L.tmax = std::max(L.tmax, int(kQTileBits));
    }
    L.nitems = items.size() - L.item_begin;
    large.push_back(L);
  }
  // persistent grids: as many blocks as the CUs hold at once for the LDS size
  auto grid_for = [&](const Level& L) {
    const size_t lds = (size_t(1) << L.tmax) * size_t(eb);
    const unsigned per_cu = std::max<unsigned>(1, unsigned((160u * 1024u) / std::max<size_t>(lds, 1)));
    const uint64_t want = std::min<uint64_t>(L.nitems, uint64_t(cus()) * std::min(per_cu, 8u));
    return unsigned(std::max<uint64_t>(want, 1));
  };
  small.grid = grid_for(small);
  for (auto& L : large) L.grid = grid_for(L);
  if (opts.phases) {
    // register-phase levels: as many blocks of 256 threads as the LDS holds
    // per CU (a 2^12-amplitude tile and the ticket: 4 of complex<float>, 2 of
    // complex<double>)
    const uint64_t per_cu =
        std::max<uint64_t>(1, (160u * 1024u) / ((uint64_t(1) << kQTileBits) * eb + 256u));
    for (auto& L : large) L.grid = unsigned(std::max<uint64_t>(1, std::min<uint64_t>(L.nitems, uint64_t(cus()) * per_cu)));
  }
  QSV_TRY(upload(d_circs, circs));
  QSV_TRY(upload(d_passes, passes));
  QSV_TRY(upload(d_ops, ops.empty() ? std::vector<TOp>(1) : ops));
  QSV_TRY(upload(d_items, items.empty() ? std::vector<TItem>(1) : items));
  QSV_TRY(upload(d_tickets, std::vector<unsigned long long>(1 + large.size(), 0)));
  QSV_TRY(upload(d_done, std::vector<unsigned int>(std::max(1, ncirc), 0)));
  QSV_TRY(upload(d_stamps, std::vector<unsigned long long>(ncirc + 1, 0)));
  if (opts.phases && !qpasses.empty()) {
    QSV_TRY(upload(d_qpasses, qpasses));
    QSV_TRY(upload(d_qphases, qphases));
    QSV_TRY(upload(d_qops, qops.empty() ? std::vector<QOp>(1) : qops));
    QSV_TRY(upload(d_qpinfo, qpinfo));
    for (int c : large_ids) large_phase_ids.push_back(c);
    QSV_TRY(upload(d_zero_ids, large_phase_ids));
  }
  return QSV_OK;
}

int Batch::enqueue(void* psi, hipStream_t st) {
  auto* tickets = static_cast<unsigned long long*>(d_tickets.p);
  auto* stamps = static_cast<unsigned long long*>(d_stamps.p);
  QSV_TRY(launch_tile_prologue(tickets, int(1 + large.size()), static_cast<unsigned int*>(d_done.p),
                               ncirc, stamps + ncirc, st));
  auto args = [&](const Level& L, int ticket) {
    TileLevelArgs a;
    a.psi = psi;
    a.circs = static_cast<const TCircuit*>(d_circs.p);
    a.passes = static_cast<const TPass*>(d_passes.p);
    a.ops = static_cast<const TOp*>(d_ops.p);
    a.items = static_cast<const TItem*>(d_items.p) + L.item_begin;
    a.nitems = L.nitems;
    a.ticket = tickets + ticket;
    a.done_tiles = static_cast<unsigned int*>(d_done.p);
    a.stamps = stamps;
    a.tmax = L.tmax;
    a.grid = L.grid;
    return a;
  };
  const bool dbl = dtype == QSV_C128;
  // small circuits on the side stream, concurrent with the large levels
  QSV_CHECK_HIP(hipEventRecord(fork, st));
  QSV_CHECK_HIP(hipStreamWaitEvent(side, fork, 0));
  if (small.nitems)
    QSV_TRY(dbl ? launch_tile_level<double>(args(small, 0), side) : launch_tile_level<float>(args(small, 0), side));
  auto qargs = [&](const Level& L, int ticket) {
    QTileLevelArgs a;
    a.psi = psi;
    a.circs = static_cast<const TCircuit*>(d_circs.p);
    a.passes = static_cast<const QPass*>(d_qpasses.p);
    a.pinfo = static_cast<const QPassInfo*>(d_qpinfo.p);
    a.phases = static_cast<const QPhase*>(d_qphases.p);
    a.ops = static_cast<const QOp*>(d_qops.p);
    a.items = static_cast<const TItem*>(d_items.p) + L.item_begin;
    a.nitems = L.nitems;
    a.ticket = tickets + ticket;
    a.done_tiles = static_cast<unsigned int*>(d_done.p);
    a.stamps = stamps;
    a.grid = L.grid;
    return a;
  };
  if (opts.phases && !large_phase_ids.empty()) {
    const int32_t* ids = static_cast<const int32_t*>(d_zero_ids.p);
    const auto* cs = static_cast<const TCircuit*>(d_circs.p);
    const int nz = int(large_phase_ids.size());
    QSV_TRY(dbl ? launch_zero_states<double>(psi, cs, ids, nz, 64, st)
                : launch_zero_states<float>(psi, cs, ids, nz, 64, st));
  }
  for (size_t k = 0; k < large.size(); ++k) {
    if (opts.phases)
      QSV_TRY(dbl ? launch_qtile_level<double>(qargs(large[k], int(1 + k)), st)
                  : launch_qtile_level<float>(qargs(large[k], int(1 + k)), st));
    else
      QSV_TRY(dbl ? launch_tile_level<double>(args(large[k], int(1 + k)), st)
                  : launch_tile_level<float>(args(large[k], int(1 + k)), st));
  }
  QSV_CHECK_HIP(hipEventRecord(join, side));
  QSV_CHECK_HIP(hipStreamWaitEvent(st, join, 0));
  return QSV_OK;
}

int Batch::run(void* psi, hipStream_t st, bool use_graph) {
  std::lock_guard<std::mutex> lk(mu);
  if (!side) {
    QSV_CHECK_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    QSV_CHECK_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    QSV_CHECK_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
  }
  if (!use_graph) return enqueue(psi, st);
  if (!exec || graph_psi != psi) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    // capture on a private stream so the caller's stream may be the null stream
    hipStream_t cap;
    QSV_CHECK_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    QSV_CHECK_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    const int st_enq = enqueue(psi, cap);
    hipGraph_t g = nullptr;
    const hipError_t e_end = hipStreamEndCapture(cap, &g);
    (void)hipStreamDestroy(cap);
    if (st_enq != QSV_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st_enq;
    }
    if (e_end != hipSuccess) return fail(QSV_EHIP, std::string("graph capture: ") + hipGetErrorString(e_end));
    graph = g;
    QSV_CHECK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    graph_psi = psi;
  }
  QSV_CHECK_HIP(hipGraphLaunch(exec, st));
  return QSV_OK;
}

}  // namespace qsv

namespace qsv {
namespace {

template <typename C>
__global__ __launch_bounds__(256) void k_max_abs_diff(const C* __restrict__ a, const C* __restrict__ b,
                                                      uint64_t count, double* partials) {
  __shared__ double sh[256];
  double m = 0;
  const uint64_t stride = uint64_t(gridDim.x) * 256;
  for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < count; i += stride) {
    const double dx = double(a[i].x) - double(b[i].x), dy = double(a[i].y) - double(b[i].y);
    m = fmax(m, sqrt(dx * dx + dy * dy));
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

}  // namespace
}  // namespace qsv

struct qsv_batch {
  qsv::Batch b;
};

using namespace qsv;

extern "C" {

int qsv_batch_create(int dtype, int ncirc, const int* nqubits, const qsv_gate* gates,
                     const size_t* gate_offsets, const qsv_batch_opts* opts, qsv_batch** out) {
  if (!out) return fail(QSV_EINVAL, "null output pointer");
  *out = nullptr;
  if (dtype != QSV_C64 && dtype != QSV_C128) return fail(QSV_EDTYPE, "dtype must be QSV_C64 or QSV_C128");
  if (ncirc < 1 || !nqubits || !gate_offsets) return fail(QSV_EINVAL, "empty or null batch");
  if (gate_offsets[ncirc] && !gates) return fail(QSV_EINVAL, "null gate array");
  for (int c = 0; c < ncirc; ++c)
    if (gate_offsets[c + 1] < gate_offsets[c]) return fail(QSV_EINVAL, "gate offsets must not decrease");
  qsv_batch_opts o{};
  if (opts) o = *opts;
  auto* b = new (std::nothrow) qsv_batch;
  if (!b) return fail(QSV_ENOMEM, "host allocation failed");
  const int st = b->b.build(dtype, ncirc, nqubits, gates, gate_offsets, o);
  if (st != QSV_OK) {
    delete b;
    return st;
  }
  *out = b;
  return QSV_OK;
}

int qsv_batch_info(const qsv_batch* b, qsv_batch_info_t* info, uint64_t* offsets) {
  if (!b) return fail(QSV_EINVAL, "null batch");
  const Batch& B = b->b;
  if (offsets) std::copy(B.offsets.begin(), B.offsets.end(), offsets);
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->ncirc = B.ncirc;
    info->total_amplitudes = B.offsets.back();
    info->npasses = int64_t(B.passes.size() + B.qpasses.size());
    info->nops = int64_t(B.ops.size() + B.qops.size());
    info->nlevels = int64_t(B.large.size()) + (B.small.nitems ? 1 : 0);
    info->opts = B.opts;
    info->nsmall = int64_t(B.small.nitems);  // one item per circuit run whole in LDS
    const double eb = B.dtype == QSV_C128 ? 16.0 : 8.0;
    double bytes = 0;
    for (const TPass& P : B.passes)
      bytes += (P.init_zero ? 1.0 : 2.0) * std::ldexp(eb, B.nq[size_t(P.circuit)]);
    // register-phase circuits: the zero fill writes each state once, then
    // a first pass writes its tile 0 and a later one reads and writes the
    // tiles that can hold nonzeros
    for (int c : B.large_phase_ids) bytes += std::ldexp(eb, B.nq[size_t(c)]);
    for (const QPassInfo& q : B.qpinfo) bytes += (q.first ? 1.0 : 2.0) * double(q.nitems) * std::ldexp(eb, kQTileBits);
    info->alg_bytes = bytes;
  }
  return QSV_OK;
}

int qsv_batch_run(qsv_batch* b, void* packed, int use_graph, void* stream) {
  if (!b) return fail(QSV_EINVAL, "null batch");
  if (!packed) return fail(QSV_EINVAL, "null state buffer");
  return b->b.run(packed, static_cast<hipStream_t>(stream), use_graph != 0);
}

int qsv_batch_latencies(qsv_batch* b, double* seconds, void* stream) {
  if (!b || !seconds) return fail(QSV_EINVAL, "null argument");
  Batch& B = b->b;
  auto st = static_cast<hipStream_t>(stream);
  std::vector<unsigned long long> s(B.ncirc + 1);
  QSV_CHECK_HIP(hipStreamSynchronize(st));
  QSV_CHECK_HIP(hipMemcpy(s.data(), B.d_stamps.p, s.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost));
  int dev = 0, khz = 0;
  QSV_CHECK_HIP(hipGetDevice(&dev));
  QSV_CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) return fail(QSV_EHIP, "wall clock rate unavailable");
  for (int c = 0; c < B.ncirc; ++c)
    seconds[c] = double(int64_t(s[c] - s[B.ncirc])) / (double(khz) * 1e3);
  return QSV_OK;
}

int qsv_batch_destroy(qsv_batch* b) {
  delete b;
  return QSV_OK;
}


int qsv_dev_max_abs_diff(const void* a, const void* b, uint64_t count, int dtype, double* out,
                         void* stream) {
  if (!a || !b || !out) return fail(QSV_EINVAL, "null argument");
  if (dtype != QSV_C64 && dtype != QSV_C128) return fail(QSV_EDTYPE, "dtype must be QSV_C64 or QSV_C128");
  auto st = static_cast<hipStream_t>(stream);
  constexpr int kBlocks = 1024;
  DevBuf part;
  QSV_TRY(upload(part, std::vector<double>(kBlocks, 0.0)));
  if (count) {
    if (dtype == QSV_C128)
      hipLaunchKernelGGL(k_max_abs_diff<double2>, dim3(kBlocks), dim3(256), 0, st,
                         static_cast<const double2*>(a), static_cast<const double2*>(b), count,
                         static_cast<double*>(part.p));
    else
      hipLaunchKernelGGL(k_max_abs_diff<float2>, dim3(kBlocks), dim3(256), 0, st,
                         static_cast<const float2*>(a), static_cast<const float2*>(b), count,
                         static_cast<double*>(part.p));
    QSV_CHECK_HIP(hipGetLastError());
  }
  std::vector<double> h(kBlocks);
  QSV_CHECK_HIP(hipMemcpyAsync(h.data(), part.p, kBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
  QSV_CHECK_HIP(hipStreamSynchronize(st));
  double m = 0;
  for (double v : h) m = std::max(m, v);
  *out = m;
  return QSV_OK;
}

int qsv_batch_autotune(int dtype, int ncirc, const int* nqubits, const qsv_gate* gates,
                       const size_t* gate_offsets, void* packed, void* scratch, int reps,
                       qsv_batch_opts* best, qsv_tune_result* results, int max_results,
                       int* nresults, void* stream) {
  if (!packed || !scratch || !best) return fail(QSV_EINVAL, "null argument");
  auto st = static_cast<hipStream_t>(stream);
  reps = std::max(1, reps);
  const int cap = dtype == QSV_C128 ? kTileMaxBits - 1 : kTileMaxBits;
  std::vector<qsv_batch_opts> cands;
  for (int t = 10; t <= std::min(13, cap); ++t)
    for (int k = 0; k <= 2; ++k) {
      qsv_batch_opts o{};
      o.tile_bits = t;
      o.small_bits = cap;
      o.fuse_k = k;
      cands.push_back(o);
    }
  // register-phase passes for the larger circuits, with the circuits up to
  // cap or up to 11 qubits whole in LDS (fusion width of those: 1 or 2)
  for (int sb : {cap, kQTileBits - 1})
    for (int k = 1; k <= 2; ++k) {
      qsv_batch_opts o{};
      o.tile_bits = kQTileBits;
      o.small_bits = sb;
      o.fuse_k = k;
      o.phases = 1;
      cands.push_back(o);
    }
  hipEvent_t e0, e1;
  QSV_CHECK_HIP(hipEventCreate(&e0));
  QSV_CHECK_HIP(hipEventCreate(&e1));
  double best_ms = 1e300;
  int nres = 0;
  uint64_t total = 0;
  int rc = QSV_OK;
  for (size_t i = 0; i < cands.size() && rc == QSV_OK; ++i) {
    qsv_batch* b = nullptr;
    rc = qsv_batch_create(dtype, ncirc, nqubits, gates, gate_offsets, &cands[i], &b);
    if (rc != QSV_OK) break;
    total = b->b.offsets.back();
    rc = b->b.run(packed, st, true);  // capture + warm-up
    if (rc == QSV_OK && hipEventRecord(e0, st) != hipSuccess) rc = fail(QSV_EHIP, "event record");
    for (int r = 0; r < reps && rc == QSV_OK; ++r) rc = b->b.run(packed, st, true);
    if (rc == QSV_OK && hipEventRecord(e1, st) != hipSuccess) rc = fail(QSV_EHIP, "event record");
    if (rc == QSV_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(QSV_EHIP, "event sync");
    float ms = 0;
    if (rc == QSV_OK) (void)hipEventElapsedTime(&ms, e0, e1);
    double diff = 0;
    const size_t bytes = total * (dtype == QSV_C128 ? 16 : 8);
    if (rc == QSV_OK) {
      if (i == 0) {
        if (hipMemcpyAsync(scratch, packed, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
          rc = fail(QSV_EHIP, "copy of reference states");
      } else {
        rc = qsv_dev_max_abs_diff(packed, scratch, total, dtype, &diff, stream);
      }
    }
    delete b;
    if (rc != QSV_OK) break;
    const double mean = double(ms) / reps;
    if (results && nres < max_results) results[nres] = qsv_tune_result{cands[i], mean, diff};
    ++nres;
    if (mean < best_ms) {
      best_ms = mean;
      *best = cands[i];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (nresults) *nresults = std::min(nres, max_results);
  return rc;
}

}  // extern "C"
