// Compiled circuits: lowering, optional fusion, device coefficient storage,
// and the launch sequence of one run.
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <memory>
#include <vector>

#include "fused.hpp"
#include "fusion.hpp"
#include "plain.hpp"
#include "program.hpp"

namespace qsv {

template <typename R>
size_t block_bytes(const Block& b) {
  using C = typename CT<R>::type;
  const size_t D = size_t(1) << b.q.size();
  const size_t pad = 256;
  auto up = [pad](size_t x) { return (x + pad - 1) / pad * pad; };
  if (b.diag) return up(D * sizeof(C));
  return up(4 * D * D * sizeof(R)) + up(D * D * sizeof(C));
}

template <typename R>
void block_image(const Block& b, unsigned char* dst, DeviceBlock& off) {
  using C = typename CT<R>::type;
  const size_t D = size_t(1) << b.q.size();
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  if (b.diag) {
    C* t = reinterpret_cast<C*>(dst);
    for (size_t i = 0; i < D; ++i) t[i] = mk<C>(R(b.m[i].real()), R(b.m[i].imag()));
    off.table = dst;
    return;
  }
  // U_real[2r+a][2c+b] for a,b in {0,1}: [[re, -im], [im, re]] of U[r][c]
  const size_t M = 2 * D;
  std::vector<R> ur(M * M);
  for (size_t r = 0; r < D; ++r)
    for (size_t c = 0; c < D; ++c) {
      const cd z = b.m[r * D + c];
      ur[(2 * r) * M + 2 * c] = R(z.real());
      ur[(2 * r) * M + 2 * c + 1] = R(-z.imag());
      ur[(2 * r + 1) * M + 2 * c] = R(z.imag());
      ur[(2 * r + 1) * M + 2 * c + 1] = R(z.real());
    }
  // lane order of the A operand: [(mt*KS + s)*64 + lane] = U_real[16mt + lane%16][4s + lane/16]
  R* ul = reinterpret_cast<R*>(dst);
  if (M >= 16) {
    const size_t MT = M / 16, KS = M / 4;
    for (size_t mt = 0; mt < MT; ++mt)
      for (size_t s = 0; s < KS; ++s)
        for (size_t lane = 0; lane < 64; ++lane)
          ul[(mt * KS + s) * 64 + lane] = ur[(16 * mt + lane % 16) * M + 4 * s + lane / 16];
  }
  C* mat = reinterpret_cast<C*>(dst + up(4 * D * D * sizeof(R)));
  for (size_t i = 0; i < D * D; ++i) mat[i] = mk<C>(R(b.m[i].real()), R(b.m[i].imag()));
  off.ulane = dst;
  off.matrix = mat;
}

template size_t block_bytes<double>(const Block&);
template size_t block_bytes<float>(const Block&);
template void block_image<double>(const Block&, unsigned char*, DeviceBlock&);
template void block_image<float>(const Block&, unsigned char*, DeviceBlock&);

int Program::build(int n_, int dtype_, const qsv_gate* gates, size_t ngates, int fuse_k,
                   int diag_k) {
  n = n_;
  dtype = dtype_;
  ngates_in = ngates;
  std::vector<KOp> lowered(ngates);
  for (size_t i = 0; i < ngates; ++i) QSV_TRY(lower_gate(gates[i], n, lowered[i]));
  // the compiled passes are the fastest form from 2^10 tiles up; they need
  // hiprtc, without it the interpreting kernel runs the same passes
  if (fuse_k == QSV_AUTO)
    fuse_k = n >= kQTileBits + 10 ? (qtile_jit_available() ? QSV_TILED_JIT : QSV_TILED) : kMaxFuse;
  const bool fixed_forms = fuse_k == QSV_TILED_JIT_FIXED;
  if (fixed_forms) fuse_k = QSV_TILED_JIT;
  if ((fuse_k == QSV_TILED || fuse_k == QSV_TILED_JIT) && n < kQTileBits)
    fuse_k = kMaxFuse;  // one tile would be the whole state
  if (fuse_k == QSV_TILED || fuse_k == QSV_TILED_JIT) {
    tiled = true;
    jit = fuse_k == QSV_TILED_JIT;
    // tile size of compiled passes: jit_tile_bits (2^13 for complex64 and
    // fixed-form complex128 passes, 2^12 for complex128 passes that branch
    // per rotation)
    const int tb = jit ? jit_tile_bits(dtype, n, fixed_forms) : kQTileBits;
    // A relabeling schedule may end out of place (no gate-free restore
    // pass) when a work buffer of the state's size can be held: allocated
    // here when the device has room for it and for the caller's state
    // besides (twice the state bytes and 2 GiB free), freed again when the
    // schedule does not use it.  QSV_JIT_OUT_OF_PLACE=0 turns it off.
    bool oop = false;
    if (jit) {
      const char* e = std::getenv("QSV_JIT_OUT_OF_PLACE");
      const size_t bytes = (size_t(1) << n) * (dtype == QSV_C128 ? 16 : 8);
      size_t free_b = 0, total_b = 0;
      if (!(e && e[0] == '0') && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
          free_b >= 2 * bytes + (size_t(2) << 30)) {
        if (hipMalloc(&work, bytes) == hipSuccess) {
          work_bytes = bytes;
          oop = true;
        } else {
          work = nullptr;
        }
      }
      (void)hipGetLastError();
    }
    QSV_TRY(compile_qtile(lowered, n, dtype == QSV_C128 ? 3 : 4, qt, tb, jit, oop));
    if (work && !qtile_uses_work(qt)) {
      (void)hipFree(work);
      work = nullptr;
      work_bytes = 0;
    }
    qtile_scale_rotations(qt);
    qt.fixed_forms = fixed_forms;
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    qt_pass_off = 0;
    qt_phase_off = up(qt.passes.size() * sizeof(QPass));
    qt_op_off = qt_phase_off + up(qt.phases.size() * sizeof(QPhase));
    image.assign(qt_op_off + up(qt.ops.size() * sizeof(QOp)), 0);
    std::memcpy(image.data() + qt_pass_off, qt.passes.data(), qt.passes.size() * sizeof(QPass));
    std::memcpy(image.data() + qt_phase_off, qt.phases.data(), qt.phases.size() * sizeof(QPhase));
    std::memcpy(image.data() + qt_op_off, qt.ops.data(), qt.ops.size() * sizeof(QOp));
    if (jit) QSV_TRY(qtile_jit_build(qt, dtype == QSV_C128, qj));
    return QSV_OK;
  }
  if (fuse_k >= 1) {
    FusionOptions fo;
    fo.max_k = fuse_k;
    fo.diag_k = diag_k > 0 ? diag_k : std::max(fuse_k, 8);
    fuse(lowered, n, fo, ops, blocks);
  } else {
    ops.clear();
    for (const KOp& o : lowered)
      if (o.kind != K_NOP) ops.push_back(o);
  }
  // device storage of the block coefficients
  const bool dbl = dtype == QSV_C128;
  std::vector<size_t> offs(blocks.size());
  size_t total = 0;
  for (size_t i = 0; i < blocks.size(); ++i) {
    offs[i] = total;
    total += dbl ? block_bytes<double>(blocks[i]) : block_bytes<float>(blocks[i]);
  }
  dblocks.assign(blocks.size(), DeviceBlock());
  image.assign(total, 0);
  for (size_t i = 0; i < blocks.size(); ++i) {
    DeviceBlock hb;
    if (dbl)
      block_image<double>(blocks[i], image.data() + offs[i], hb);
    else
      block_image<float>(blocks[i], image.data() + offs[i], hb);
    dblocks[i] = hb;  // host addresses until upload()
  }
  return QSV_OK;
}

int Program::upload() const {
  std::lock_guard<std::mutex> lk(mu);
  if (uploaded || image.empty()) {
    uploaded = true;
    return QSV_OK;
  }
  if (hipMalloc(&dbuf, image.size()) != hipSuccess) {
    (void)hipGetLastError();
    dbuf = nullptr;
    return fail(QSV_ENOMEM, "program: device allocation of coefficients failed");
  }
  QSV_CHECK_HIP(hipMemcpy(dbuf, image.data(), image.size(), hipMemcpyHostToDevice));
  auto rebase = [&](const void* p) -> const void* {
    if (!p) return nullptr;
    return static_cast<unsigned char*>(dbuf) +
           (static_cast<const unsigned char*>(p) - image.data());
  };
  for (auto& d : dblocks) {
    d.ulane = rebase(d.ulane);
    d.matrix = rebase(d.matrix);
    d.table = rebase(d.table);
  }
  uploaded = true;
  return QSV_OK;
}

Program::~Program() {
  if (dbuf) (void)hipFree(dbuf);
  if (work) (void)hipFree(work);
}

int Program::run(void* psi, hipStream_t st) const {
  QSV_TRY(upload());
  PlainTuning tu;
  const bool dbl = dtype == QSV_C128;
  if (tiled) {
    const unsigned char* base = static_cast<const unsigned char*>(dbuf);
    const QPass* passes = reinterpret_cast<const QPass*>(base + qt_pass_off);
    QTileLaunch a;
    a.psi = psi;
    a.phases = reinterpret_cast<const QPhase*>(base + qt_phase_off);
    a.ops = reinterpret_cast<const QOp*>(base + qt_op_off);
    if (jit) {
      for (size_t i = 0; i < qt.passes.size(); ++i)
        QSV_TRY(qtile_jit_launch(qj, i, psi, work, a.ops, st));
      return QSV_OK;
    }
    for (size_t i = 0; i < qt.passes.size(); ++i) {
      a.pass = passes + i;
      a.ntiles = qt.passes[i].ntiles;
      QSV_TRY(dbl ? launch_qtile<double>(a, st) : launch_qtile<float>(a, st));
    }
    return QSV_OK;
  }
  for (const KOp& op : ops) {
    if (op.kind == K_DENSEK || op.kind == K_DIAGK) {
      const DeviceBlock& db = dblocks[op.blk];
      QSV_TRY(dbl ? launch_fused<double>(op, db, psi, n, st) : launch_fused<float>(op, db, psi, n, st));
    } else {
      QSV_TRY(dbl ? launch_plain<double>(op, psi, n, st, tu) : launch_plain<float>(op, psi, n, st, tu));
    }
  }
  return QSV_OK;
}

void Program::info(qsv_program_info* out) const {
  std::memset(out, 0, sizeof(*out));
  out->ngates = int64_t(ngates_in);
  if (tiled) {
    const int eb = dtype == QSV_C128 ? 16 : 8;
    out->nlaunch = int64_t(qt.passes.size());
    out->ntiled_phases = int64_t(qt.phases.size());
    for (const QPass& P : qt.passes)
      out->ntiled_relabel += std::memcmp(P.tbits, P.tbits_out, sizeof(P.tbits)) != 0 ? 1 : 0;
    out->alg_bytes = double(qt.passes.size()) * 2.0 * std::ldexp(1.0, n) * eb;
    out->fp_lane_ops = qtile_fp_lane_ops(qt);
    out->work_bytes = double(work_bytes);
    return;
  }
  out->nlaunch = int64_t(ops.size());
  const int eb = dtype == QSV_C128 ? 16 : 8;
  for (const KOp& op : ops) {
    out->alg_bytes += algorithmic_bytes(op, n, eb);
    if (op.kind == K_DENSEK) {
      out->ndense[op.nq] += 1;
      out->flops += 8.0 * double(uint64_t(1) << op.nq) * std::ldexp(1.0, n);
    } else if (op.kind == K_DIAGK) {
      out->ndiag += 1;
    } else {
      out->nplain += 1;
    }
  }
}

}  // namespace qsv
