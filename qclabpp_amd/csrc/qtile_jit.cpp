// Source generation, hiprtc compilation and launch of specialised tiled
// passes.  The generated kernels follow the tiled pass kernel (qtile.hip):
// the same tile walk, prefetch, phase gather and write-back, with every
// descriptor field a literal.
#include "qtile_jit.hpp"

#include <dlfcn.h>
#include <hip/hiprtc.h>

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstddef>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <cstring>
#include <sstream>
#include <thread>

namespace qsv {

int device_cus();  // qtile.hip

namespace {

constexpr int NR = 1 << kQRegBits;
// grid of a pass launch: each block walks a contiguous range of at least
// kJitTilesPerBlock tiles, with at most kJitBlocksPerCU blocks per CU.  On
// the n=30 brickwork (2^18 tiles of 2^12, 2^17 of 2^13) 32 blocks per CU gave
// 111 ms (branching passes) and 100 ms (fixed forms), 128 gave 106 and 99
// ms, 2-8 (near-persistent grids) 104-128 ms; at n=28, 128 blocks per CU
// (two tiles each) took 32 ms against 26 ms for 32 (eight tiles each).
constexpr int kJitBlocksPerCU = 128;
constexpr uint64_t kJitTilesPerBlock = 8;

static_assert(sizeof(QOp) % sizeof(double) == 0 && offsetof(QOp, m) % sizeof(double) == 0 &&
                  offsetof(QOp, mf) % sizeof(float) == 0,
              "coefficients are addressed as arrays of the state's real type");

// Device prelude of every generated module: complex helpers and the pair
// updates, in the forms of the interpreted kernel's ops.
const char* kPrelude = R"QSV(
#define QC __attribute__((address_space(4)))
typedef REAL R;
typedef R RV __attribute__((ext_vector_type(2)));
struct __attribute__((aligned(2 * sizeof(R)))) C { R x, y; };
__device__ __forceinline__ C mk(R a, R b) { C c; c.x = a; c.y = b; return c; }
__device__ __forceinline__ C cmul(C a, C b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ C cmad(C a, C b, C c) {
  return mk(c.x + a.x * b.x - a.y * b.y, c.y + a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ C ldnt(const C* p) {
  const RV v = __builtin_nontemporal_load(reinterpret_cast<const RV*>(p));
  return mk(v.x, v.y);
}
__device__ __forceinline__ void stnt(C* p, C v) {
  const RV w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<RV*>(p));
}
__device__ __forceinline__ unsigned swz(unsigned p) { return p ^ (((p >> 4) ^ (p >> 8)) & 15u); }
__device__ __forceinline__ C cf2(const QC R* cf, int o) { return mk(cf[o], cf[o + 1]); }
// 2x2 [[m0, m1], [m2, m3]] on (a, b)
__device__ __forceinline__ void u_pair(C& a, C& b, C m0, C m1, C m2, C m3) {
  const C x = a, y = b;
  a = cmad(m1, y, cmul(m0, x));
  b = cmad(m3, y, cmul(m2, x));
}
__device__ __forceinline__ void u_real(C& a, C& b, R m0, R m1, R m2, R m3) {
  const C x = a, y = b;
  a = mk(m0 * x.x + m1 * y.x, m0 * x.y + m1 * y.y);
  b = mk(m2 * x.x + m3 * y.x, m2 * x.y + m3 * y.y);
}
// [[c, -s], [s, c]]
__device__ __forceinline__ void u_rot(C& a, C& b, R c, R s) {
  const C x = a, y = b;
  a = mk(c * x.x - s * y.x, c * x.y - s * y.y);
  b = mk(s * x.x + c * y.x, s * x.y + c * y.y);
}
// [[p, iq], [ir, s]]
__device__ __forceinline__ void u_xl(C& a, C& b, R p, R q, R r, R s) {
  const C x = a, y = b;
  a = mk(p * x.x - q * y.y, p * x.y + q * y.x);
  b = mk(s * y.x - r * x.y, s * y.y + r * x.x);
}
// scaled forms (the factor taken out is applied at the end of the pass):
// [[1, -k], [k, 1]]
__device__ __forceinline__ void u_ta(C& a, C& b, R k) {
  const C x = a, y = b;
  a = mk(x.x - k * y.x, x.y - k * y.y);
  b = mk(y.x + k * x.x, y.y + k * x.y);
}
// [[k, -1], [1, k]]
__device__ __forceinline__ void u_tb(C& a, C& b, R k) {
  const C x = a, y = b;
  a = mk(k * x.x - y.x, k * x.y - y.y);
  b = mk(k * y.x + x.x, k * y.y + x.y);
}
// [[k, 1], [-1, k]]
__device__ __forceinline__ void u_tbs(C& a, C& b, R k) {
  const C x = a, y = b;
  a = mk(k * x.x + y.x, k * x.y + y.y);
  b = mk(k * y.x - x.x, k * y.y - x.y);
}
// [[1, ik], [ik, 1]]
__device__ __forceinline__ void u_xa(C& a, C& b, R k) {
  const C x = a, y = b;
  a = mk(x.x - k * y.y, x.y + k * y.x);
  b = mk(y.x - k * x.y, y.y + k * x.x);
}
// [[k, i], [i, k]]
__device__ __forceinline__ void u_xb(C& a, C& b, R k) {
  const C x = a, y = b;
  a = mk(k * x.x - y.y, k * x.y + y.x);
  b = mk(k * y.x - x.y, k * y.y + x.x);
}
// 4x4 on (a0, a1, a2, a3), coefficients row-major from cf[o]
__device__ __forceinline__ void u_d2(C& a0, C& a1, C& a2, C& a3, const QC R* cf, int o) {
  const C x[4] = {a0, a1, a2, a3};
  C y[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    C acc = cmul(cf2(cf, o + 8 * r), x[0]);
#pragma unroll
    for (int c = 1; c < 4; ++c) acc = cmad(cf2(cf, o + 8 * r + 2 * c), x[c], acc);
    y[r] = acc;
  }
  a0 = y[0]; a1 = y[1]; a2 = y[2]; a3 = y[3];
}
)QSV";

// Op k applied to the register arrays `arrs`, one after the other (the
// data-dependent scaled forms inside one branch).
// fixed: the scaled forms' choice (which of cos and sin is larger) is taken
// from the op's coefficients at generation time, so the op has no branch
// (programs of the fixed-form kind only).
void emit_op(std::ostringstream& s, const QOp& o, size_t k, bool c128,
             const std::vector<std::string>& arrs = {"v"}, bool fixed = false) {
  // index of coefficient slot 0 of op k in the QOp array read as R[]
  const size_t stride = c128 ? sizeof(QOp) / sizeof(double) : sizeof(QOp) / sizeof(float);
  const size_t base = k * stride + (c128 ? offsetof(QOp, m) / sizeof(double) : offsetof(QOp, mf) / sizeof(float));
  auto at = [&](int i) { return std::to_string(base + i); };
  const int K = o.d & -o.d;
  auto pairs = [&](auto body) {
    for (int p = 0; p < NR; ++p)
      if (!(p & K)) body(p, p ^ o.d, bool(o.roles >> p & 1), bool(o.act >> p & 1));
  };
  std::string arr = arrs[0];
  auto v = [&arr](int p) { return arr + "[" + std::to_string(p) + "]"; };
  // body() once per array, with v() naming that array's registers
  auto each = [&](auto body) {
    for (const std::string& a : arrs) {
      arr = a;
      body();
    }
  };
  s << "    {\n";
  switch (o.kind) {
    case Q_DIAGR:
      each([&] {
        for (int p = 0; p < NR; ++p)
          if (o.b >> p & 1) s << "      " << v(p) << " = cmul(cf2(cf, " << at(2 * p) << "), " << v(p) << ");\n";
      });
      break;
    case Q_ROTP:
      s << "      const R c = cf[" << at(0) << "], sn = cf[" << at(4) << "];\n";
      each([&] {
        pairs([&](int p, int q, bool sw, bool) {
          s << "      u_rot(" << v(p) << ", " << v(q) << ", c, " << (sw ? "-sn" : "sn") << ");\n";
        });
      });
      break;
    case Q_REALP:
      s << "      const R m0 = cf[" << at(0) << "], m1 = cf[" << at(2) << "], m2 = cf[" << at(4)
        << "], m3 = cf[" << at(6) << "];\n";
      each([&] {
        pairs([&](int p, int q, bool sw, bool) {
          s << "      u_real(" << v(p) << ", " << v(q) << (sw ? ", m3, m2, m1, m0);\n" : ", m0, m1, m2, m3);\n");
        });
      });
      break;
    case Q_XLIKEP:
      s << "      const R xp = cf[" << at(0) << "], xq = cf[" << at(3) << "], xr = cf[" << at(5)
        << "], xs = cf[" << at(6) << "];\n";
      each([&] {
        pairs([&](int p, int q, bool sw, bool) {
          s << "      u_xl(" << v(p) << ", " << v(q) << (sw ? ", xs, xr, xq, xp);\n" : ", xp, xq, xr, xs);\n");
        });
      });
      break;
    case Q_ROTS:
    case Q_XLS: {
      // the form is data (it depends on the angle), so one compiled pass
      // serves every angle: a branch on a uniform value
      const bool rot = o.kind == Q_ROTS;
      auto first_form = [&] {
        each([&] {
          pairs([&](int p, int q, bool sw, bool) {
            if (rot) s << "        u_ta(" << v(p) << ", " << v(q) << (sw ? ", -k);\n" : ", k);\n");
            else s << "        u_xa(" << v(p) << ", " << v(q) << ", k);\n";
          });
        });
      };
      auto second_form = [&] {
        each([&] {
          pairs([&](int p, int q, bool sw, bool) {
            if (rot) s << "        " << (sw ? "u_tbs(" : "u_tb(") << v(p) << ", " << v(q) << ", k);\n";
            else s << "        u_xb(" << v(p) << ", " << v(q) << ", k);\n";
          });
        });
      };
      s << "      const R k = cf[" << at(0) << "];\n";
      if (fixed) {
        if ((c128 ? o.m[2] : double(o.mf[2])) != 0.0) first_form();
        else second_form();
        break;
      }
      s << "      if (cf[" << at(2) << "] != R(0)) {\n";
      first_form();
      s << "      } else {\n";
      second_form();
      s << "      }\n";
      break;
    }
    case Q_PAIR:
    case Q_CPAIR:
      s << "      const C m0 = cf2(cf, " << at(0) << "), m1 = cf2(cf, " << at(2) << "), m2 = cf2(cf, "
        << at(4) << "), m3 = cf2(cf, " << at(6) << ");\n";
      each([&] {
        pairs([&](int p, int q, bool sw, bool on) {
          if (o.kind == Q_CPAIR && !on) return;
          s << "      u_pair(" << v(p) << ", " << v(q) << (sw ? ", m3, m2, m1, m0);\n" : ", m0, m1, m2, m3);\n");
        });
      });
      break;
    case Q_DENSE2: {
      const int A = 1 << o.a, B = 1 << o.b;
      each([&] {
        for (int i = 0; i < NR; ++i)
          if (!(i & A) && !(i & B))
            s << "      u_d2(" << v(i) << ", " << v(i | B) << ", " << v(i | A) << ", " << v(i | A | B)
              << ", cf, " << at(0) << ");\n";
      });
      break;
    }
    default:
      break;
  }
  s << "    }\n";
}

// 64-bit literal
std::string u64(uint64_t x) { return std::to_string(x) + "ull"; }

// Global offset (in amplitudes) of tile position p under the address bits
// tb of the tile-local bits (the pass's tbits when reading, tbits_out when
// writing).
uint64_t gofs(const uint8_t* tb, int T, uint32_t p) {
  uint64_t o = 0;
  for (int b = 0; b < T; ++b)
    if (p >> b & 1u) o |= uint64_t(1) << tb[b];
  return o;
}

constexpr int kLaneBitsJit = 6;  // a wavefront's 64 lanes

// Tile-local bit of thread index bit j in a phase (ord[j]): the bits outside
// the phase's registers, in an order chosen per phase.
//  - `first` (the lines a direct first phase loads) leads, so 8 or 16
//    consecutive lanes read one 128-byte line;
//  - `lanes` (the lines a direct last phase stores) must sit among the 6
//    lane bits, so each line is written by one instruction;
//  - lane bits 0-2 take bits 0-2 of the slot swizzle's low nibble (local
//    bit mod 4, for local bits below 12) and lane bit 3 takes bit 3.  Measured
//    on gfx950: ds_read_b128 is conflict-free once lane bits 0-3 cover the
//    four nibble bits in any order, ds_write_b128 only when lane bits 0-2
//    cover nibble bits 0-2 (7.5 conflict cycles per write otherwise).  Lanes
//    in ascending bit order collided whenever the registers held low local
//    bits, as relabeled passes often do.
std::vector<int> lane_order(int T, uint32_t regmask, uint32_t first, uint32_t lanes) {
  std::vector<int> ord;
  auto free_bit = [&](int b) {
    return !(regmask >> b & 1u) && std::find(ord.begin(), ord.end(), b) == ord.end();
  };
  auto nib = [](int b) { return b < 12 ? b % 4 : -1; };
  auto nib_used = [&](int r) {
    for (size_t i = 0; i < ord.size() && i < 4; ++i)
      if (nib(ord[i]) == r) return true;
    return false;
  };
  for (int b = 0; b < T; ++b)
    if ((first >> b & 1u) && free_bit(b)) ord.push_back(b);
  while (ord.size() < 4) {
    // wanted nibble bits: an unused one of 0-2 for lane bits 0-2, 3 for lane bit 3
    auto wanted = [&](int b) {
      const int r = nib(b);
      if (r < 0 || nib_used(r)) return false;
      return ord.size() < 3 ? r < 3 : r == 3;
    };
    int pick = -1;
    for (int pass = 0; pass < 4 && pick < 0; ++pass)
      for (int b = 0; b < T && pick < 0; ++b) {
        if (!free_bit(b)) continue;
        const bool fresh = nib(b) >= 0 && !nib_used(nib(b));
        if ((pass == 0 && wanted(b) && (lanes >> b & 1u)) || (pass == 1 && wanted(b)) || (pass == 2 && fresh) ||
            pass == 3)
          pick = b;
      }
    if (pick < 0) break;
    ord.push_back(pick);
  }
  for (int b = 0; b < T; ++b)
    if ((lanes >> b & 1u) && free_bit(b)) ord.push_back(b);
  for (int b = 0; b < T; ++b)
    if (free_bit(b)) ord.push_back(b);
  return ord;
}

// Expression for the global offset of this thread's share of a phase.
std::string thread_gofs(const uint8_t* tb, const std::vector<int>& ord) {
  std::ostringstream e;
  e << "(0ull";
  for (size_t j = 0; j < ord.size(); ++j)
    e << " | ((unsigned long long)((tid >> " << j << ") & 1u) << " << int(tb[ord[j]]) << ")";
  e << ")";
  return e.str();
}

// Tile-local bits whose address bit lies inside a 128-byte line.
uint32_t line_mask(const uint8_t* tb, int T, bool c128) {
  uint32_t m = 0;
  for (int b = 0; b < T; ++b)
    if (tb[b] < (c128 ? 3 : 4)) m |= 1u << b;
  return m;
}

// Parameters of a pass kernel: the state in place, or for a pass at an
// out-of-place end the buffer it reads and the one it writes.
std::string pass_params(const QPass& P) {
  if (P.wsrc || P.wdst) return "(const C* __restrict__ pin, C* __restrict__ pout, const QC R* cf0, unsigned long long ntiles)";
  return "(C* __restrict__ psi, const QC R* cf0, unsigned long long ntiles)";
}
const char* src_name(const QPass& P) { return P.wsrc || P.wdst ? "pin" : "psi"; }
const char* dst_name(const QPass& P) { return P.wsrc || P.wdst ? "pout" : "psi"; }
bool same_out_runs(const QPass& P) {
  if (P.nruns_out != P.nruns) return false;
  for (int r = 0; r < P.nruns; ++r)
    if (P.rdst_out[r] != P.rdst[r] || P.rlen_out[r] != P.rlen[r]) return false;
  return true;
}
// tile_base(t): the tile's address bits outside the tile, on the read side;
// tile_base_out(t) the same on the written side when the pass moves them
void emit_tile_bases(std::ostringstream& s, const QPass& P) {
  auto one = [&](const char* name, int nruns, const uint8_t* dst, const uint8_t* len) {
    s << "  auto " << name << " = [&](unsigned long long t) {\n"
      << "    unsigned long long b = 0;\n";
    for (int r = 0, sh = 0; r < nruns; ++r) {
      s << "    b |= ((t >> " << sh << ") & " << u64((uint64_t(1) << len[r]) - 1) << ") << " << int(dst[r]) << ";\n";
      sh += len[r];
    }
    s << "    return b;\n"
      << "  };\n";
  };
  one("tile_base", P.nruns, P.rdst, P.rlen);
  if (!same_out_runs(P)) one("tile_base_out", P.nruns_out, P.rdst_out, P.rlen_out);
}
const char* out_base(const QPass& P) { return same_out_runs(P) ? "base" : "tile_base_out(t)"; }

// Kernel names: qtp_<pass>, or qtf_<pass> for fixed-form passes (so that
// profiles keep the two forms apart).
const char* pass_prefix(const QProgram& qp) { return qp.fixed_forms ? "qtf_" : "qtp_"; }

// A pass as one kernel.  Each tile goes HBM -> phases -> HBM.  A phase
// gathers its 16 register amplitudes from the LDS tile, applies its ops and
// writes them back; the first phase instead loads its registers straight
// from HBM (prefetched one tile ahead) and the last phase stores straight to
// HBM, whenever the lanes of a wave then still cover whole 128-byte lines
// (the phase's registers leave the line bits free).  A one-phase pass with
// both ends direct never touches LDS.
void emit_pass(std::ostringstream& s, const QProgram& qp, size_t pi, bool c128) {
  const QPass& P = qp.passes[pi];
  const int T = qp.tile_bits;
  const unsigned threads = 1u << (T - kQRegBits);
  const uint32_t tile_order = ((1u << kQRegBits) - 1) << (T - kQRegBits);  // registers j of position tid + threads*j
  const int K = P.phase_end - P.phase_begin;
  const QPhase& F = qp.phases[P.phase_begin];
  const QPhase& L = qp.phases[P.phase_end - 1];
  // tile-local bits of the 128-byte lines read and written (the same ones
  // unless the pass relabels qubits)
  const uint32_t line_in = line_mask(P.tbits, T, c128), line_out = line_mask(P.tbits_out, T, c128);
  const bool din = !(F.regmask & line_in);
  // A pass that relabels stores each amplitude to another address of its
  // tile than the one it came from, so no store may start before every
  // load of the tile has landed.  Passes of two or more phases guarantee
  // that with the barrier after their first scatter into LDS (which needs
  // the loaded values); a one-phase pass with both ends direct has no
  // barrier, so it stores through LDS instead.
  const bool relabels = std::memcmp(P.tbits, P.tbits_out, sizeof(P.tbits)) != 0;
  bool dout = !(L.regmask & line_out) && !(K == 1 && din && relabels);
  // per-phase thread orders (the first and last phases' also give the
  // direct load and store offsets); loads and stores through LDS use
  // positions tid + threads*j, whose lanes are local bits 0..5
  std::vector<std::vector<int>> ords(K);
  for (int k = 0; k < K; ++k) {
    const QPhase& Q = qp.phases[P.phase_begin + k];
    ords[k] = lane_order(T, Q.regmask, k == 0 && din ? line_in : 0u, k == K - 1 && dout ? line_out : 0u);
  }
  // a direct store needs every written line bit among the 6 lane bits
  // (always the case for the schedules built here; else store through LDS)
  if (dout)
    for (int j = kLaneBitsJit; j < int(ords[K - 1].size()); ++j)
      if (line_out >> ords[K - 1][j] & 1u) dout = false;
  if (!dout) ords[K - 1] = lane_order(T, L.regmask, K == 1 && din ? line_in : 0u, 0u);
  // waves per SIMD the register budget is sized for, from the blocks per CU
  // that LDS allows (160 KiB): complex128 tiles of 2^12 (64 KiB, 2 blocks of
  // 4 waves), complex64 tiles of 2^12 (32 KiB, 3 blocks of 4 waves) or 2^13
  // (64 KiB, 2 blocks of 8 waves)
  const int wpe = c128 ? (T == 12 ? 2 : 1) : (T == 12 ? 3 : 4);
  s << "extern \"C\" __global__ __launch_bounds__(" << threads << ") __attribute__((amdgpu_waves_per_eu(" << wpe << ")))\n"
    << "void " << pass_prefix(qp) << pi << pass_params(P) << " {\n";
  if (!(din && dout && K == 1)) s << "  __shared__ C tile[" << (1u << T) << "];\n";
  s << "  const unsigned tid = threadIdx.x;\n";
  emit_tile_bases(s, P);
  const std::string src = src_name(P), dst = dst_name(P), obase = out_base(P);
  // offsets of the 16 amplitudes a thread loads (and stores) at the ends:
  // the first (last) phase's registers, or tile positions tid + threads*j
  std::vector<uint64_t> in_ofs(NR), out_ofs(NR);
  for (int i = 0; i < NR; ++i) {
    in_ofs[i] = din ? gofs(P.tbits, T, qtile_swizzle(F.regpos[i])) : gofs(P.tbits, T, threads * uint32_t(i));
    out_ofs[i] = dout ? gofs(P.tbits_out, T, qtile_swizzle(L.regpos_out[i]))
                      : gofs(P.tbits_out, T, threads * uint32_t(i));
  }
  const std::vector<int> ord_tile = lane_order(T, tile_order, 0u, 0u);
  s << "  const unsigned long long gin = " << thread_gofs(P.tbits, din ? ords[0] : ord_tile) << ";\n"
    << "  const unsigned long long gout = " << thread_gofs(P.tbits_out, dout ? ords[K - 1] : ord_tile)
    << ";\n";
  auto arr = [&](const char* name, const std::vector<uint64_t>& v) {
    s << "  const unsigned long long " << name << "[16] = {";
    for (int i = 0; i < NR; ++i) s << (i ? ", " : "") << u64(v[i]);
    s << "};\n";
  };
  arr("oin", in_ofs);
  arr("oout", out_ofs);
  s << "  const unsigned tslot = swz(tid);\n"
    // each block walks a contiguous range of tiles: measured up to 20%
    // more HBM bandwidth than interleaved tiles on scattered tile layouts
    << "  const unsigned long long per = (ntiles + gridDim.x - 1) / gridDim.x;\n"
    << "  unsigned long long t = blockIdx.x * per;\n"
    << "  const unsigned long long tend = t + per < ntiles ? t + per : ntiles;\n"
    << "  C pre[16];\n"
    << "  if (t < tend) {\n"
    << "    const unsigned long long b0 = tile_base(t) + gin;\n"
    << "#pragma unroll\n"
    << "    for (int j = 0; j < 16; ++j) pre[j] = ldnt(&" << src << "[b0 + oin[j]]);\n"
    << "  }\n"
    << "  for (; t < tend; ++t) {\n"
    << "    const unsigned long long base = tile_base(t);\n"
    << "    C v[16];\n";
  if (din) {
    s << "#pragma unroll\n"
      << "    for (int j = 0; j < 16; ++j) v[j] = pre[j];\n";
  } else {
    s << "#pragma unroll\n"
      << "    for (int j = 0; j < 16; ++j) tile[tslot ^ swz(" << threads << "u * j)] = pre[j];\n"
      << "    __syncthreads();\n";
  }
  s << "    {\n"
    << "      const unsigned long long tn = t + 1;\n"
    << "      if (tn < tend) {\n"
    << "        const unsigned long long bn = tile_base(tn) + gin;\n"
    << "#pragma unroll\n"
    << "        for (int j = 0; j < 16; ++j) pre[j] = ldnt(&" << src << "[bn + oin[j]]);\n"
    << "      }\n"
    << "    }\n";
  for (int k = 0; k < K; ++k) {
    const int ph = P.phase_begin + k;
    const QPhase& Q = qp.phases[ph];
    const bool gather = !(k == 0 && din);
    const bool scatter = !(k == K - 1 && dout);
    // the coefficient pointer and the thread index pass through empty asm
    // statements at each phase, so coefficient loads and LDS addresses are
    // formed in the phase that uses them instead of being hoisted out of
    // the tile loop into registers held (and spilled) across the pass
    s << "    {  // phase " << ph << "\n"
      << "      const QC R* cf = cf0;\n"
      << "      asm volatile(\"\" : \"+s\"(cf));\n";
    if (gather || scatter) {
      s << "      unsigned y = tid;\n"
        << "      asm volatile(\"\" : \"+v\"(y));\n"
        << "      const unsigned x = 0u";
      // thread index bit j -> tile-local bit ords[k][j]
      for (size_t j = 0; j < ords[k].size(); ++j)
        s << " | (((y >> " << j << ") & 1u) << " << ords[k][j] << ")";
      s << ";\n";
      s << "      const unsigned tp = swz(x);\n";
    }
    if (gather) {
      for (int i = 0; i < NR; ++i) s << "      v[" << i << "] = tile[tp ^ " << Q.regpos[i] << "u];\n";
      // the LDS tile is written next by the following tile's first phase
      if (!scatter) s << "      __syncthreads();\n";
    }
    for (int o = Q.op_begin; o < Q.op_end; ++o) emit_op(s, qp.ops[o], size_t(o), c128, {"v"}, qp.fixed_forms);
    if (scatter) {
      for (int i = 0; i < NR; ++i) s << "      tile[tp ^ " << Q.regpos_out[i] << "u] = v[" << i << "];\n";
      s << "      __syncthreads();\n";
    } else {
      s << "      const unsigned long long bo = " << obase << " + gout;\n"
        << "#pragma unroll\n"
        << "      for (int j = 0; j < 16; ++j) stnt(&" << dst << "[bo + oout[j]], v[j]);\n";
    }
    s << "    }\n";
  }
  if (!dout) {
    s << "    {\n"
      << "      const unsigned long long bo = " << obase << " + gout;\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) stnt(&" << dst << "[bo + oout[j]], tile[tslot ^ swz(" << threads << "u * j)]);\n"
      << "      __syncthreads();\n"
      << "    }\n";
  }
  s << "  }\n"
    << "}\n";
}

// ---- passes through a half-tile LDS buffer ----------------------------------
//
// A phase change moves the tile through LDS in two rounds split on one
// tile-local bit b: round h moves the amplitudes whose position has bit b
// equal to h.  The buffer then holds half a tile (2^(T-1) amplitudes plus
// padding: 33 KiB for a complex128 tile of 2^12), so LDS no longer limits a
// CU to two blocks and the register budget sets occupancy: three waves per
// SIMD, no register prefetch (the other blocks hide the load latency).
//
// Slots: the position with bit b dropped and the top tile bit moved into its
// place ("compressed", T-1 bits), padded as q + (q >> 5) + (q >> 10).  Both
// maps are additive over disjoint bits, so a register's slot is the thread's
// base slot plus a literal (one address register per side, immediate
// offsets).  Bank rules measured on gfx950 with a probe kernel: ds_read_b128
// serves 32 lanes at a time and is conflict-free when their 16-byte slots
// differ mod 32; ds_write_b128 serves 8 lanes, slots distinct mod 8.  With
// this padding a compressed bit m moves the slot by 2^(m mod 5) mod 32 (its
// "class"), so reads are conflict-free when lane bits 0-4 hold the five
// classes, and writes when lane bits 0-2 hold classes 0-2.

// One side of a phase change: thread bit j -> local bit ord[j]; register i
// at local position offset off[i] (register bits only).
struct XSide {
  std::vector<int> ord;
  uint32_t off[NR];
  uint32_t regmask;
};

// Local bit m after compression by the split bit b (m != b); b < 0: none.
int xbit(int m, int b, int T) { return b >= 0 && m == T - 1 && b != T - 1 ? b : m; }

uint32_t xcompress(uint32_t p, int b, int T) {
  uint32_t q = 0;
  for (int m = 0; m < T; ++m)
    if (m != b && (p >> m & 1u)) q |= 1u << xbit(m, b, T);
  return q;
}

uint32_t xpad(uint32_t q) { return q + (q >> 5) + (q >> 10); }

int xclass(int m, int b, int T) { return xbit(m, b, T) % 5; }

// Conflict-free lanes of an order under split bit b: reads count the lane
// bits 0-4 of distinct classes (5 = conflict-free), writes whether lane
// bits 0-2 hold classes 0-2.
int read_lanes_ok(const std::vector<int>& ord, int b, int T) {
  unsigned seen = 0;
  int n = 0;
  for (size_t j = 0; j < 5 && j < ord.size(); ++j) {
    const int c = xclass(ord[j], b, T);
    if (!(seen >> c & 1u)) ++n;
    seen |= 1u << c;
  }
  return n;
}
bool write_lanes_ok(const std::vector<int>& ord, int b, int T, int wl) {
  unsigned seen = 0;
  for (size_t j = 0; j < size_t(wl) && j < ord.size(); ++j) {
    if (ord[j] == b) continue;  // that lane bit sits out each round in turn
    const int c = xclass(ord[j], b, T);
    if (c >= wl || (seen >> c & 1u)) return false;
    seen |= 1u << c;
  }
  return true;
}

// Thread order of a phase of the split kernels: as lane_order, with the
// classes of the padded compressed slots (split bit b of the phase change
// that fills the registers; b < 0 for a phase loaded from HBM): lane bits
// 0-2 take classes 0-2 and lane bits 3-4 classes 3-4 (after the lines a
// direct first phase loads, local bits 0-2: classes 0-2), preferring bits
// other than the top tile bit, whose class moves with the split bit; the
// lines a direct last phase stores stay among the 6 lane bits.
std::vector<int> lane_order_split(int T, uint32_t regmask, uint32_t first, uint32_t lanes, int b, int wl) {
  std::vector<int> ord;
  auto free_bit = [&](int m) {
    return !(regmask >> m & 1u) && std::find(ord.begin(), ord.end(), m) == ord.end();
  };
  auto used = [&](int c) {
    for (size_t i = 0; i < ord.size() && i < 5; ++i)
      if (xclass(ord[i], b, T) == c) return true;
    return false;
  };
  for (int m = 0; m < T; ++m)
    if ((first >> m & 1u) && free_bit(m)) ord.push_back(m);
  while (ord.size() < 5) {
    auto wanted = [&](int m) {
      const int c = xclass(m, b, T);
      return !used(c) && (int(ord.size()) < wl ? c < wl : c >= wl);
    };
    int pick = -1;
    for (int pass = 0; pass < 6 && pick < 0; ++pass)
      for (int m = 0; m < T && pick < 0; ++m) {
        if (!free_bit(m)) continue;
        const bool top = m == T - 1;
        if ((pass == 0 && wanted(m) && !top && (lanes >> m & 1u)) || (pass == 1 && wanted(m) && !top) ||
            (pass == 2 && wanted(m)) || (pass == 3 && !used(xclass(m, b, T)) && !top) ||
            (pass == 4 && !used(xclass(m, b, T))) || pass == 5)
          pick = m;
      }
    if (pick < 0) break;
    ord.push_back(pick);
  }
  for (int m = 0; m < T; ++m)
    if ((lanes >> m & 1u) && free_bit(m)) ord.push_back(m);
  for (int m = 0; m < T; ++m)
    if (free_bit(m)) ord.push_back(m);
  return ord;
}

// The exchange itself: registers v (side s) -> registers v (side d).
void emit_exchange(std::ostringstream& s, const XSide& src, const XSide& dst, int b, int T) {
  auto base = [&](const XSide& x, const char* name) {
    s << "      const unsigned " << name << "q = 0u";
    for (size_t j = 0; j < x.ord.size(); ++j)
      if (x.ord[j] != b) s << " | (((y >> " << j << ") & 1u) << " << xbit(x.ord[j], b, T) << ")";
    s << ";\n      const unsigned " << name << " = " << name << "q + (" << name << "q >> 5) + (" << name
      << "q >> 10);\n";
  };
  auto thread_bit = [&](const XSide& x) {
    for (size_t j = 0; j < x.ord.size(); ++j)
      if (x.ord[j] == b) return int(j);
    return -1;
  };
  s << "    {  // phase change through the half buffer, split on local bit " << b << "\n"
    << "      unsigned y = tid;\n"
    << "      asm volatile(\"\" : \"+v\"(y));\n";
  base(src, "as");
  base(dst, "ad");
  const int js = (src.regmask >> b & 1u) ? -1 : thread_bit(src);
  for (int h = 0; h < 2; ++h) {
    s << "      __syncthreads();\n";
    if (js >= 0) s << "      if (((y >> " << js << ") & 1u) == " << h << "u) {\n";
    for (int i = 0; i < NR; ++i)
      if (js >= 0 || int(src.off[i] >> b & 1u) == h)
        s << "      tile[as + " << xpad(xcompress(src.off[i], b, T)) << "u] = v[" << i << "];\n";
    if (js >= 0) s << "      }\n";
    s << "      __syncthreads();\n";
    for (int i = 0; i < NR; ++i)
      if (int(dst.off[i] >> b & 1u) == h)
        s << "      const C u" << i << " = tile[ad + " << xpad(xcompress(dst.off[i], b, T)) << "u];\n";
  }
  for (int i = 0; i < NR; ++i) s << "      v[" << i << "] = u" << i << ";\n";
  s << "    }\n";
}

// A pass as one kernel with half-buffer phase changes.  Each tile: load the
// registers (the first phase's, or tile positions tid + threads*j), phase
// changes and ops, store the registers (the last phase's, or positions
// tid + threads*j).  Returns false where a phase's register offsets leave
// its register bits (never for the schedules built here); the caller then
// emits the full-tile kernel.
bool emit_pass_split(std::ostringstream& s, const QProgram& qp, size_t pi, bool c128) {
  const QPass& P = qp.passes[pi];
  const int T = qp.tile_bits;
  const unsigned threads = 1u << (T - kQRegBits);
  const uint32_t tile_order = ((1u << kQRegBits) - 1) << (T - kQRegBits);
  const int K = P.phase_end - P.phase_begin;
  const QPhase& F = qp.phases[P.phase_begin];
  const QPhase& L = qp.phases[P.phase_end - 1];
  const uint32_t line_in = line_mask(P.tbits, T, c128), line_out = line_mask(P.tbits_out, T, c128);
  const bool din = !(F.regmask & line_in);
  const bool relabels = std::memcmp(P.tbits, P.tbits_out, sizeof(P.tbits)) != 0;
  // a relabeling one-phase pass with both ends direct would have no barrier
  // between the loads and the stores of its block; it stores through LDS
  bool dout = !(L.regmask & line_out) && !(K == 1 && din && relabels);
  // lane bits whose classes the writes constrain: 3 for ds_write_b128 (8
  // lanes per cycle), 4 for ds_write_b64 (16 lanes of 8 bytes)
  const int wl = c128 ? 3 : 4;
  // a direct store needs every written line bit among the 6 lane bits
  // (always the case for the schedules built here; else store through LDS)
  XSide tile_side;  // positions tid + threads*j
  tile_side.ord = lane_order_split(T, tile_order, 0u, 0u, -1, wl);
  tile_side.regmask = tile_order;
  for (int i = 0; i < NR; ++i) tile_side.off[i] = threads * uint32_t(i);
  // each phase at its gather and at its write-back
  std::vector<XSide> gat(K), sct(K);
  for (int k = 0; k < K; ++k) {
    const QPhase& Q = qp.phases[P.phase_begin + k];
    gat[k].regmask = sct[k].regmask = Q.regmask;
    for (int i = 0; i < NR; ++i) {
      gat[k].off[i] = qtile_swizzle(Q.regpos[i]);
      sct[k].off[i] = qtile_swizzle(Q.regpos_out[i]);
      if ((gat[k].off[i] | sct[k].off[i]) & ~Q.regmask) return false;
    }
  }
  // split bits (bin[k]: the phase change into phase k, bin[K]: into the
  // store layout) chosen with the thread order of the phase they fill: a
  // register bit of the receiving phase (every thread reads 8 amplitudes per
  // round, with no lane conditions: the compiler would otherwise merge and
  // spill the undefined halves of the register array), of the sending phase
  // too where possible (every thread also writes 8); then the reads and the
  // writes free of bank conflicts; then the highest bit.
  std::vector<int> bin(K + 1, -1);
  auto choose = [&](const XSide& src, uint32_t dmask, uint32_t lanes, std::vector<int>& ord_out) {
    int best = -1;
    long best_score = -1;
    for (int b = T - 1; b >= 0; --b) {
      if (!(dmask >> b & 1u)) continue;
      std::vector<int> ord = dmask == tile_order ? tile_side.ord : lane_order_split(T, dmask, 0u, lanes, b, wl);
      const long score = ((src.regmask >> b & 1u) ? 100 : 0) + 10 * read_lanes_ok(ord, b, T) +
                         (write_lanes_ok(src.ord, b, T, wl) ? 1 : 0);
      if (score > best_score) {
        best_score = score;
        best = b;
        ord_out = ord;
      }
    }
    return best;
  };
  auto build = [&]() {
    for (int k = 0; k < K; ++k) {
      const QPhase& Q = qp.phases[P.phase_begin + k];
      const uint32_t lanes = k == K - 1 && dout ? line_out : 0u;
      if (k == 0 && din) {
        gat[0].ord = lane_order_split(T, Q.regmask, line_in, lanes, -1, wl);
      } else {
        const XSide& src = k == 0 ? tile_side : sct[k - 1];
        bin[k] = choose(src, Q.regmask, lanes, gat[k].ord);
      }
      sct[k].ord = gat[k].ord;
    }
    if (!dout) {
      std::vector<int> ord;
      bin[K] = choose(sct[K - 1], tile_order, 0u, ord);
    }
  };
  build();
  if (dout) {
    for (int j = kLaneBitsJit; j < int(sct[K - 1].ord.size()); ++j)
      if (line_out >> sct[K - 1].ord[j] & 1u) dout = false;
    if (!dout) build();
  }
  const XSide& first = din ? gat[0] : tile_side;
  const XSide& last = dout ? sct[K - 1] : tile_side;
  std::vector<uint64_t> in_ofs(NR), out_ofs(NR);
  for (int i = 0; i < NR; ++i) {
    in_ofs[i] = gofs(P.tbits, T, first.off[i]);
    out_ofs[i] = gofs(P.tbits_out, T, last.off[i]);
  }
  const bool lds = !din || !dout || K > 1;
  // waves per SIMD the register budget is sized for: four blocks of 2^12 or
  // two of 2^13 amplitudes fit a CU's LDS (16 waves); fixed-form passes fit
  // the 128 registers of four waves per SIMD, passes that branch per
  // rotation need about 166 and run three
  const int wpe = c128 ? (qp.fixed_forms ? 4 : 3) : (qp.fixed_forms ? 6 : 4);
  s << "extern \"C\" __global__ __launch_bounds__(" << threads << ") __attribute__((amdgpu_waves_per_eu(" << wpe
    << ")))\n"
    << "void " << pass_prefix(qp) << pi << pass_params(P) << " {\n";
  if (lds) s << "  __shared__ C tile[" << xpad((1u << (T - 1)) - 1) + 1 << "];\n";
  s << "  const unsigned tid = threadIdx.x;\n";
  emit_tile_bases(s, P);
  const std::string src = src_name(P), dst = dst_name(P), obase = out_base(P);
  s << "  const unsigned long long gin = " << thread_gofs(P.tbits, first.ord) << ";\n"
    << "  const unsigned long long gout = " << thread_gofs(P.tbits_out, last.ord) << ";\n";
  auto arr = [&](const char* name, const std::vector<uint64_t>& v) {
    s << "  const unsigned long long " << name << "[16] = {";
    for (int i = 0; i < NR; ++i) s << (i ? ", " : "") << u64(v[i]);
    s << "};\n";
  };
  arr("oin", in_ofs);
  arr("oout", out_ofs);
  // each block walks a contiguous range of tiles
  s << "  const unsigned long long per = (ntiles + gridDim.x - 1) / gridDim.x;\n"
    << "  unsigned long long t = blockIdx.x * per;\n"
    << "  const unsigned long long tend = t + per < ntiles ? t + per : ntiles;\n"
    << "  for (; t < tend; ++t) {\n"
    << "    const unsigned long long base = tile_base(t);\n"
    << "    C v[16];\n"
    // the thread's offsets pass through empty asm statements inside the
    // loop, so the 16 load and 16 store addresses are formed per tile
    // instead of being hoisted out of the loop (and spilled)
    << "    {\n"
    << "      unsigned long long g = gin;\n"
    << "      asm volatile(\"\" : \"+v\"(g));\n"
    << "      const unsigned long long b0 = base + g;\n"
    << "#pragma unroll\n"
    << "      for (int j = 0; j < 16; ++j) v[j] = ldnt(&" << src << "[b0 + oin[j]]);\n"
    << "    }\n";
  if (!din) emit_exchange(s, tile_side, gat[0], bin[0], T);
  for (int k = 0; k < K; ++k) {
    const int ph = P.phase_begin + k;
    const QPhase& Q = qp.phases[ph];
    // the coefficient pointer passes through an empty asm statement at each
    // phase, so coefficient loads are formed in the phase that uses them
    // instead of being hoisted out of the tile loop
    s << "    {  // phase " << ph << "\n"
      << "      const QC R* cf = cf0;\n"
      << "      asm volatile(\"\" : \"+s\"(cf));\n";
    for (int o = Q.op_begin; o < Q.op_end; ++o) emit_op(s, qp.ops[o], size_t(o), c128, {"v"}, qp.fixed_forms);
    s << "    }\n";
    if (k + 1 < K) emit_exchange(s, sct[k], gat[k + 1], bin[k + 1], T);
  }
  if (!dout) emit_exchange(s, sct[K - 1], tile_side, bin[K], T);
  s << "    {\n"
    << "      unsigned long long g = gout;\n"
    << "      asm volatile(\"\" : \"+v\"(g));\n"
    << "      const unsigned long long bo = " << obase << " + g;\n"
    << "#pragma unroll\n"
    << "      for (int j = 0; j < 16; ++j) stnt(&" << dst << "[bo + oout[j]], v[j]);\n"
    << "    }\n"
    << "  }\n"
    << "}\n";
  return true;
}

// Compiled passes use the half buffer unless QSV_JIT_SPLIT=0 (full-tile
// kernels: for comparison).  Complex64 (tiles of 2^13, 33 KiB per block):
// the n=28 brickwork 20.6 -> 16.2 ms, n=33 666 -> 580 ms.
bool use_split() {
  const char* e = std::getenv("QSV_JIT_SPLIT");
  return !(e && e[0] == '0');
}

// ---- adjoint reverse sweep: passes over two states -------------------------
//
// The same tile walk and phases as emit_pass, over phi (psi, registers v)
// and lam (registers w) at once: every op is applied to both, and a Q_GEN
// op adds Re(k_p conj(w[p ^ d]) v[p]) over the registers into its row.
// Both states share one LDS tile: a phase change moves phi through it, then
// lam, so the block keeps the forward passes' LDS footprint (two blocks of
// 64 KiB per CU for complex128).  Row sums: each thread's FP64 sum is
// reduced over the 16 lanes of a DPP row (4 steps), and the row's first
// lane adds it to its own LDS slot (4 slots per wave), or, for passes with
// more than 128 rows, over the wave (readlane) into one slot per wave.  At
// the end each block adds its slots in a fixed order and writes one FP64
// partial per row and block, so the gradient does not depend on timing.

const char* kAdjPrelude = R"QSV(
__device__ __forceinline__ long long qta_bits(double x) { return __builtin_bit_cast(long long, x); }
__device__ __forceinline__ double qta_dbl(long long b) { return __builtin_bit_cast(double, b); }
template <int CTRL>
__device__ __forceinline__ double qta_dpp(double x) {
  const long long b = qta_bits(x);
  // every lane reads a valid lane of its own row (quad_perm, row_ror), so
  // the old value is never used and needs no initialising move
  const int lo = __builtin_amdgcn_mov_dpp(int(b), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), CTRL, 0xf, 0xf, true);
  return qta_dbl((long long)(unsigned)lo | ((long long)hi << 32));
}
// sum over the 16 lanes of each DPP row: quad_perm [1,0,3,2], [2,3,0,1],
// then row_ror 4 and 8; every lane of the row ends with the row's sum
__device__ __forceinline__ double qta_rowsum(double t) {
  t += qta_dpp<0xB1>(t);
  t += qta_dpp<0x4E>(t);
  t += qta_dpp<0x124>(t);
  t += qta_dpp<0x128>(t);
  return t;
}
__device__ __forceinline__ double qta_lane(double t, int lane) {
  const long long b = qta_bits(t);
  const int lo = __builtin_amdgcn_readlane(int(b), lane);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
  return qta_dbl((long long)(unsigned)lo | ((long long)hi << 32));
}
// 4 slots per wave: slot tid >> 4 (the DPP row)
__device__ __forceinline__ void qta_acc4(double t, double* row, unsigned tid) {
  t = qta_rowsum(t);
  if ((tid & 15u) == 0u) row[tid >> 4] += t;
}
// 1 slot per wave: the four row sums added in lane order
__device__ __forceinline__ void qta_acc1(double t, double* row, unsigned tid) {
  t = qta_rowsum(t);
  const double s = ((qta_lane(t, 0) + qta_lane(t, 16)) + qta_lane(t, 32)) + qta_lane(t, 48);
  if ((tid & 63u) == 0u) row[tid >> 6] += s;
}
)QSV";

// Rows a pass's Q_GEN ops write (pass-local indices 0..rows-1).
int adj_pass_rows(const QProgram& qp, size_t pi) {
  const QPass& P = qp.passes[pi];
  int rows = 0;
  for (int o = qp.phases[P.phase_begin].op_begin; o < qp.phases[P.phase_end - 1].op_end; ++o)
    if (qp.ops[o].kind == Q_GEN) rows = std::max(rows, qp.ops[o].a + 1);
  return rows;
}

// The weights k_p of one generator share their modulus (the generator's
// coefficients are 0 or of modulus 1, times one weight), which is taken
// out of the sum; each term then costs two FMAs with signs folded in:
//   Re(k conj(l) f) = kr (l.x f.x + l.y f.y) - ki (l.x f.y - l.y f.x).
double gen_modulus(const QOp& o) {
  for (int p = 0; p < NR; ++p)
    if (o.act >> p & 1u) return std::hypot(double(o.mf[2 * p]), double(o.mf[2 * p + 1]));
  return 1.0;
}

std::string lit(double x) {
  std::ostringstream e;
  e.precision(17);
  e << std::showpoint << x;
  return e.str();
}

void emit_gen_sum(std::ostringstream& s, const QOp& o, int spw) {
  const double m = gen_modulus(o);
  if (m != 1.0) s << "      t *= " << lit(m) << ";\n";
  s << "      qta_acc" << spw << "(t, acc + " << o.a << " * NSLOT, tid);\n";
}

void emit_gen(std::ostringstream& s, const QOp& o, bool c128, int spw) {
  const char* cv = c128 ? "" : "(double)";
  const double m = gen_modulus(o);
  s << "    {  // generator row " << o.a << "\n"
    << "      double t = 0.0;\n";
  for (int p = 0; p < NR; ++p) {
    if (!(o.act >> p & 1u)) continue;
    const int q = p ^ o.d;
    const double kr = o.mf[2 * p] / m, ki = o.mf[2 * p + 1] / m;  // 0 or +-1
    const std::string lx = std::string(cv) + "w[" + std::to_string(q) + "].x",
                      ly = std::string(cv) + "w[" + std::to_string(q) + "].y",
                      fx = std::string(cv) + "v[" + std::to_string(p) + "].x",
                      fy = std::string(cv) + "v[" + std::to_string(p) + "].y";
    auto sg = [](double k) { return k > 0 ? "" : "-"; };
    if (kr != 0)
      s << "      t = fma(" << sg(kr) << lx << ", " << fx << ", t);\n"
        << "      t = fma(" << sg(kr) << ly << ", " << fy << ", t);\n";
    if (ki != 0)
      s << "      t = fma(" << sg(-ki) << lx << ", " << fy << ", t);\n"
        << "      t = fma(" << sg(ki) << ly << ", " << fx << ", t);\n";
  }
  emit_gen_sum(s, o, spw);
  s << "    }\n";
}

// A run of generators with d = 0 (Z strings and projectors: RZ, RZZ, CRZ,
// P, CP) with no op between them sees the same registers, and each one's
// weights are imaginary: k_p = i kappa_p, Re(k conj(l) f) = -kappa_p u_p with
// u_p = Im(conj(l) f) = l.x f.y - l.y f.x.  The u_p are formed once for the
// run and each generator adds them with its signs.
bool diag_gen(const QOp& o) {
  if (o.kind != Q_GEN || o.d != 0) return false;
  for (int p = 0; p < NR; ++p)
    if ((o.act >> p & 1u) && o.mf[2 * p] != 0) return false;
  return true;
}

void emit_diag_gen_run(std::ostringstream& s, const QProgram& qp, int o0, int o1, bool c128, int spw) {
  const char* cv = c128 ? "" : "(double)";
  uint32_t used = 0;
  for (int o = o0; o < o1; ++o) used |= qp.ops[o].act;
  s << "    {  // generators " << qp.ops[o0].a << ".." << qp.ops[o1 - 1].a << ": one diagonal product\n"
    << "      double u[16];\n";
  for (int p = 0; p < NR; ++p)
    if (used >> p & 1u) {
      const std::string i = std::to_string(p);
      s << "      u[" << i << "] = fma(" << cv << "w[" << i << "].x, " << cv << "v[" << i << "].y, -(" << cv << "w["
        << i << "].y * " << cv << "v[" << i << "].x));\n";
    }
  for (int o = o0; o < o1; ++o) {
    const QOp& g = qp.ops[o];
    const double m = gen_modulus(g);
    s << "      {  // generator row " << g.a << "\n"
      << "        double t = 0.0";
    for (int p = 0; p < NR; ++p)
      if (g.act >> p & 1u) s << (g.mf[2 * p + 1] / m > 0 ? " - " : " + ") << "u[" << p << "]";
    s << ";\n";
    std::ostringstream tail;
    emit_gen_sum(tail, g, spw);
    s << "  " << tail.str();
    s << "      }\n";
  }
  s << "    }\n";
}

void emit_adj_pass(std::ostringstream& s, const QProgram& qp, size_t pi, bool c128) {
  const QPass& P = qp.passes[pi];
  const int T = qp.tile_bits;
  const unsigned threads = 1u << (T - kQRegBits);
  const int K = P.phase_end - P.phase_begin;
  const QPhase& F = qp.phases[P.phase_begin];
  const QPhase& L = qp.phases[P.phase_end - 1];
  const int rows = adj_pass_rows(qp, pi);
  const int spw = rows <= 128 ? 4 : 1;  // LDS slots per wave and row
  const unsigned nslot = threads / 64 * spw;
  const uint32_t line_in = line_mask(P.tbits, T, c128), line_out = line_mask(P.tbits_out, T, c128);
  const bool din = !(F.regmask & line_in);
  const bool relabels = std::memcmp(P.tbits, P.tbits_out, sizeof(P.tbits)) != 0;
  bool dout = !(L.regmask & line_out) && !(K == 1 && din && relabels);
  std::vector<std::vector<int>> ords(K);
  for (int k = 0; k < K; ++k) {
    const QPhase& Q = qp.phases[P.phase_begin + k];
    ords[k] = lane_order(T, Q.regmask, k == 0 && din ? line_in : 0u, k == K - 1 && dout ? line_out : 0u);
  }
  if (dout)
    for (int j = kLaneBitsJit; j < int(ords[K - 1].size()); ++j)
      if (line_out >> ords[K - 1][j] & 1u) dout = false;
  if (!dout) ords[K - 1] = lane_order(T, L.regmask, K == 1 && din ? line_in : 0u, 0u);
  const uint32_t tile_order = ((1u << kQRegBits) - 1) << (T - kQRegBits);
  // registers: two states of 16 amplitudes; complex128 keeps two waves per SIMD
  const int wpe = c128 ? 2 : 4;
  s << "#define NSLOT " << nslot << "u\n"
    << "extern \"C\" __global__ __launch_bounds__(" << threads << ") __attribute__((amdgpu_waves_per_eu(" << wpe
    << ")))\n"
    << "void qta_" << pi << "(C* __restrict__ psi, C* __restrict__ lam, const QC R* cf0, unsigned long long ntiles, "
    << "double* __restrict__ part) {\n";
  const bool lds = !(din && dout && K == 1);
  if (lds) s << "  __shared__ C tile[" << (1u << T) << "];\n";
  if (rows) s << "  __shared__ double acc[" << rows * nslot << "];\n";
  s << "  const unsigned tid = threadIdx.x;\n";
  if (rows)
    s << "  for (unsigned i = tid; i < " << rows * nslot << "u; i += " << threads << "u) acc[i] = 0.0;\n"
      << "  __syncthreads();\n";
  s << "  auto tile_base = [&](unsigned long long t) {\n"
    << "    unsigned long long b = 0;\n";
  for (int r = 0, sh = 0; r < P.nruns; ++r) {
    s << "    b |= ((t >> " << sh << ") & " << u64((uint64_t(1) << P.rlen[r]) - 1) << ") << " << int(P.rdst[r])
      << ";\n";
    sh += P.rlen[r];
  }
  s << "    return b;\n"
    << "  };\n";
  std::vector<uint64_t> in_ofs(NR), out_ofs(NR);
  for (int i = 0; i < NR; ++i) {
    in_ofs[i] = din ? gofs(P.tbits, T, qtile_swizzle(F.regpos[i])) : gofs(P.tbits, T, threads * uint32_t(i));
    out_ofs[i] = dout ? gofs(P.tbits_out, T, qtile_swizzle(L.regpos_out[i]))
                      : gofs(P.tbits_out, T, threads * uint32_t(i));
  }
  const std::vector<int> ord_tile = lane_order(T, tile_order, 0u, 0u);
  s << "  const unsigned long long gin = " << thread_gofs(P.tbits, din ? ords[0] : ord_tile) << ";\n"
    << "  const unsigned long long gout = " << thread_gofs(P.tbits_out, dout ? ords[K - 1] : ord_tile)
    << ";\n";
  auto arr = [&](const char* name, const std::vector<uint64_t>& v) {
    s << "  const unsigned long long " << name << "[16] = {";
    for (int i = 0; i < NR; ++i) s << (i ? ", " : "") << u64(v[i]);
    s << "};\n";
  };
  arr("oin", in_ofs);
  arr("oout", out_ofs);
  s << "  const unsigned tslot = swz(tid);\n"
    << "  const unsigned long long per = (ntiles + gridDim.x - 1) / gridDim.x;\n"
    << "  unsigned long long t = blockIdx.x * per;\n"
    << "  const unsigned long long tend = t + per < ntiles ? t + per : ntiles;\n"
    << "  for (; t < tend; ++t) {\n"
    << "    const unsigned long long base = tile_base(t);\n"
    << "    C v[16], w[16];\n";
  // the thread's tile-local position expression of phase k
  auto position = [&](int k) {
    std::ostringstream e;
    e << "      unsigned y = tid;\n"
      << "      asm volatile(\"\" : \"+v\"(y));\n"
      << "      const unsigned x = 0u";
    for (size_t j = 0; j < ords[k].size(); ++j) e << " | (((y >> " << j << ") & 1u) << " << ords[k][j] << ")";
    e << ";\n"
      << "      const unsigned tp = swz(x);\n";
    return e.str();
  };
  auto gather = [&](const char* a, int k) {
    const QPhase& Q = qp.phases[P.phase_begin + k];
    for (int i = 0; i < NR; ++i) s << "      " << a << "[" << i << "] = tile[tp ^ " << Q.regpos[i] << "u];\n";
  };
  auto scatter = [&](const char* a, int k) {
    const QPhase& Q = qp.phases[P.phase_begin + k];
    for (int i = 0; i < NR; ++i) s << "      tile[tp ^ " << Q.regpos_out[i] << "u] = " << a << "[" << i << "];\n";
  };
  // tile start
  if (din) {
    s << "    {\n"
      << "      const unsigned long long b0 = base + gin;\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) v[j] = ldnt(&psi[b0 + oin[j]]);\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) w[j] = ldnt(&lam[b0 + oin[j]]);\n"
      << "    }\n";
  } else {
    // both states in tile order, then each through LDS into phase 0's registers
    s << "    {\n"
      << "      const unsigned long long b0 = base + gin;\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) v[j] = ldnt(&psi[b0 + oin[j]]);\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) w[j] = ldnt(&lam[b0 + oin[j]]);\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) tile[tslot ^ swz(" << threads << "u * j)] = v[j];\n"
      << "      __syncthreads();\n"
      << position(0);
    gather("v", 0);
    s << "      __syncthreads();\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) tile[tslot ^ swz(" << threads << "u * j)] = w[j];\n"
      << "      __syncthreads();\n";
    gather("w", 0);
    s << "    }\n";
  }
  for (int k = 0; k < K; ++k) {
    const int ph = P.phase_begin + k;
    const QPhase& Q = qp.phases[ph];
    s << "    {  // phase " << ph << "\n"
      << "      const QC R* cf = cf0;\n"
      << "      asm volatile(\"\" : \"+s\"(cf));\n";
    for (int o = Q.op_begin; o < Q.op_end; ++o) {
      if (diag_gen(qp.ops[o])) {
        int e = o + 1;
        while (e < Q.op_end && diag_gen(qp.ops[e])) ++e;
        if (e - o >= 2) {
          emit_diag_gen_run(s, qp, o, e, c128, spw);
          o = e - 1;
          continue;
        }
      }
      if (qp.ops[o].kind == Q_GEN) {
        emit_gen(s, qp.ops[o], c128, spw);
      } else {
        // one state after the other: a shared branch for the scaled forms
        // holds both states' temporaries at once and spills
        emit_op(s, qp.ops[o], size_t(o), c128, {"v"});
        emit_op(s, qp.ops[o], size_t(o), c128, {"w"});
      }
    }
    s << "    }\n";
    if (k + 1 < K) {
      // phase change: phi, then lam, through the shared tile
      s << "    {\n" << position(k);
      scatter("v", k);
      s << "    }\n"
        << "    __syncthreads();\n"
        << "    {\n" << position(k + 1);
      gather("v", k + 1);
      s << "    }\n"
        << "    __syncthreads();\n"
        << "    {\n" << position(k);
      scatter("w", k);
      s << "    }\n"
        << "    __syncthreads();\n"
        << "    {\n" << position(k + 1);
      gather("w", k + 1);
      s << "    }\n";
      // the last phase stores straight to HBM: the next tile's first LDS
      // writes must wait until every thread has gathered
      if (k + 1 == K - 1 && dout) s << "    __syncthreads();\n";
    }
  }
  if (dout) {
    if (K == 1 && !din) s << "    __syncthreads();\n";
    s << "    {\n"
      << "      const unsigned long long bo = base + gout;\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) stnt(&psi[bo + oout[j]], v[j]);\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) stnt(&lam[bo + oout[j]], w[j]);\n"
      << "    }\n";
  } else {
    s << "    {\n" << position(K - 1);
    scatter("v", K - 1);
    s << "    }\n"
      << "    __syncthreads();\n"
      << "    {\n"
      << "      const unsigned long long bo = base + gout;\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) stnt(&psi[bo + oout[j]], tile[tslot ^ swz(" << threads << "u * j)]);\n"
      << "    }\n"
      << "    __syncthreads();\n"
      << "    {\n" << position(K - 1);
    scatter("w", K - 1);
    s << "    }\n"
      << "    __syncthreads();\n"
      << "    {\n"
      << "      const unsigned long long bo = base + gout;\n"
      << "#pragma unroll\n"
      << "      for (int j = 0; j < 16; ++j) stnt(&lam[bo + oout[j]], tile[tslot ^ swz(" << threads << "u * j)]);\n"
      << "    }\n"
      << "    __syncthreads();\n";
  }
  s << "  }\n";
  if (rows) {
    s << "  __syncthreads();\n"
      << "  for (unsigned r = tid; r < " << rows << "u; r += " << threads << "u) {\n"
      << "    double sum = 0.0;\n"
      << "    for (unsigned k = 0; k < NSLOT; ++k) sum += acc[r * NSLOT + k];\n"
      << "    const unsigned long long at = 2ull * ((unsigned long long)r * " << kAdjPartStride << "ull + blockIdx.x);\n"
      << "    part[at] = sum;\n"
      << "    part[at + 1] = 0.0;\n"
      << "  }\n";
  }
  s << "}\n"
    << "#undef NSLOT\n";
}

// ---- hiprtc, loaded on first use ----------------------------------------
struct Rtc {
  bool ok = false;
  std::string why;
  hiprtcResult (*create)(hiprtcProgram*, const char*, const char*, int, const char**, const char**) = nullptr;
  hiprtcResult (*compile)(hiprtcProgram, int, const char**) = nullptr;
  hiprtcResult (*log_size)(hiprtcProgram, size_t*) = nullptr;
  hiprtcResult (*log)(hiprtcProgram, char*) = nullptr;
  hiprtcResult (*code_size)(hiprtcProgram, size_t*) = nullptr;
  hiprtcResult (*code)(hiprtcProgram, char*) = nullptr;
  hiprtcResult (*destroy)(hiprtcProgram*) = nullptr;
};

const Rtc& rtc() {
  static Rtc r = [] {
    Rtc x;
    void* h = dlopen("libhiprtc.so.7", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("/opt/rocm/lib/libhiprtc.so.7", RTLD_NOW | RTLD_LOCAL);
    if (!h) {
      const char* e = dlerror();
      x.why = std::string("cannot load libhiprtc.so.7: ") + (e ? e : "");
      return x;
    }
    auto sym = [&](auto& f, const char* name) {
      f = reinterpret_cast<std::remove_reference_t<decltype(f)>>(dlsym(h, name));
      return f != nullptr;
    };
    if (!sym(x.create, "hiprtcCreateProgram") || !sym(x.compile, "hiprtcCompileProgram") ||
        !sym(x.log_size, "hiprtcGetProgramLogSize") || !sym(x.log, "hiprtcGetProgramLog") ||
        !sym(x.code_size, "hiprtcGetCodeSize") || !sym(x.code, "hiprtcGetCode") ||
        !sym(x.destroy, "hiprtcDestroyProgram")) {
      x.why = "libhiprtc.so.7 lacks an expected entry point";
      return x;
    }
    x.ok = true;
    return x;
  }();
  return r;
}

}  // namespace

struct QJitModule {
  hipModule_t mod = nullptr;
  ~QJitModule() {
    if (mod) (void)hipModuleUnload(mod);
  }
};

namespace {

struct CacheKey {
  int device;
  std::string src;
  bool operator<(const CacheKey& o) const { return device != o.device ? device < o.device : src < o.src; }
};
std::mutex g_cache_mu;
std::map<CacheKey, std::shared_ptr<QJitModule>> g_cache;
constexpr size_t kCacheMax = 256;  // modules kept for reuse; programs hold their own references
std::atomic<int64_t> g_mod_hits{0}, g_disk_hits{0}, g_compiles{0};

// ---- code objects on disk ------------------------------------------------
// A compiled group of passes is also kept as a file named by two 64-bit
// hashes of its source (and the source length), so a new process with a
// circuit of a known structure loads it instead of compiling.  The directory
// is $QSV_JIT_CACHE_DIR, else $XDG_CACHE_HOME/qclabpp_amd/jit, else
// $HOME/.cache/qclabpp_amd/jit; QSV_JIT_CACHE_DIR set to an empty string
// turns it off.  Files are written to a temporary name and renamed, so a
// reader never sees a partial one; a file that fails to load is compiled
// again and replaced.

std::string disk_cache_dir() {
  if (const char* e = std::getenv("QSV_JIT_CACHE_DIR")) return e;
  if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) return std::string(x) + "/qclabpp_amd/jit";
  if (const char* h = std::getenv("HOME"); h && *h) return std::string(h) + "/.cache/qclabpp_amd/jit";
  return "";
}

std::string disk_cache_file(const std::string& dir, const std::string& src) {
  uint64_t a = 1469598103934665603ull, b = 0x9e3779b97f4a7c15ull;
  for (unsigned char c : src) {
    a = (a ^ c) * 1099511628211ull;
    b = (b + c) * 0xff51afd7ed558ccdull;
    b ^= b >> 29;
  }
  char name[80];
  std::snprintf(name, sizeof name, "/qtp_%016llx%016llx_%zu.co", (unsigned long long)a,
                (unsigned long long)b, src.size());
  return dir + name;
}

bool disk_cache_read(const std::string& path, std::vector<char>& code) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !code.empty();
}

void disk_cache_write(const std::string& dir, const std::string& path, const std::vector<char>& code) {
  // mkdir -p
  for (size_t i = 1; i <= dir.size(); ++i)
    if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
  const std::string tmp = path + ".tmp" + std::to_string(getpid()) + "_" +
                          std::to_string(std::hash<std::thread::id>()(std::this_thread::get_id()));
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    f.write(code.data(), std::streamsize(code.size()));
    if (!f) {
      f.close();
      std::remove(tmp.c_str());
      return;
    }
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

}  // namespace

bool qtile_jit_available() { return rtc().ok; }

void qtile_jit_cache_stats(int64_t* mem_hits, int64_t* disk_hits, int64_t* compiles) {
  if (mem_hits) *mem_hits = g_mod_hits.load();
  if (disk_hits) *disk_hits = g_disk_hits.load();
  if (compiles) *compiles = g_compiles.load();
}

std::string qtile_jit_source(const QProgram& qp, bool c128, size_t p0, size_t p1) {
  std::ostringstream s;
  s << "#define REAL " << (c128 ? "double" : "float") << "\n" << kPrelude;
  const bool split = use_split();
  for (size_t i = p0; i < p1 && i < qp.passes.size(); ++i) {
    std::ostringstream k;
    if (split && emit_pass_split(k, qp, i, c128)) {
      s << k.str();
      continue;
    }
    emit_pass(s, qp, i, c128);
  }
  return s.str();
}

int qtile_jit_compile(const std::string& src, std::vector<char>& code) {
  const Rtc& r = rtc();
  if (!r.ok) return fail(QSV_EUNSUP, "specialised tiled passes need hiprtc: " + r.why);
  // QSV_JIT_DUMP_DIR: keep every generated source there, for inspection
  if (const char* d = std::getenv("QSV_JIT_DUMP_DIR"); d && *d) {
    std::ofstream f(disk_cache_file(d, src).substr(0, disk_cache_file(d, src).size() - 3) + ".hip");
    f << src;
  }
  hiprtcProgram prog = nullptr;
  if (r.create(&prog, src.c_str(), "qtile_pass.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return fail(QSV_EHIP, "hiprtcCreateProgram failed");
  // no packed FP32 instructions, as in the rest of the device code
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = r.compile(prog, int(sizeof(opts) / sizeof(opts[0])), opts);
  if (cr != HIPRTC_SUCCESS) {
    size_t n = 0;
    std::string log;
    if (r.log_size(prog, &n) == HIPRTC_SUCCESS && n) {
      log.resize(n);
      r.log(prog, &log[0]);
    }
    r.destroy(&prog);
    return fail(QSV_EHIP, "hiprtc compile of a tiled pass failed: " + log.substr(0, 2000));
  }
  size_t n = 0;
  r.code_size(prog, &n);
  code.resize(n);
  r.code(prog, code.data());
  r.destroy(&prog);
  return n ? QSV_OK : fail(QSV_EHIP, "hiprtc returned an empty code object");
}

namespace {

// Generate (source of passes [p0, p1)), compile or take from the caches,
// and load the kernels named prefix + pass index.
template <typename Gen>
int build_modules(const QProgram& qp, const char* prefix, Gen gen, QJit& out) {
  out = QJit();
  const size_t np = qp.passes.size();
  if (!np) return QSV_OK;
  int dev = 0;
  QSV_CHECK_HIP(hipGetDevice(&dev));
  // split the passes into groups of similar op counts, one module each,
  // compiled on parallel threads
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t ng = std::min<size_t>(np, hw);
  size_t total = 0;
  for (const QPass& P : qp.passes)
    total += size_t(qp.phases[P.phase_end - 1].op_end - qp.phases[P.phase_begin].op_begin) + 4;
  std::vector<size_t> cut{0};
  size_t acc = 0;
  for (size_t i = 0; i < np; ++i) {
    const QPass& P = qp.passes[i];
    acc += size_t(qp.phases[P.phase_end - 1].op_end - qp.phases[P.phase_begin].op_begin) + 4;
    if (acc * ng >= total * cut.size() && cut.size() < ng && i + 1 < np) cut.push_back(i + 1);
  }
  cut.push_back(np);
  const size_t groups = cut.size() - 1;
  std::vector<std::string> src(groups);
  std::vector<std::shared_ptr<QJitModule>> mods(groups);
  std::vector<std::vector<char>> code(groups);
  std::vector<int> st(groups, QSV_OK);
  std::vector<std::string> err(groups);
  std::vector<size_t> todo;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (size_t g = 0; g < groups; ++g) {
      src[g] = gen(cut[g], cut[g + 1]);
      auto it = g_cache.find(CacheKey{dev, src[g]});
      if (it != g_cache.end()) {
        mods[g] = it->second;
        ++g_mod_hits;
      } else {
        todo.push_back(g);
      }
    }
  }
  const std::string dir = disk_cache_dir();
  std::vector<std::string> path(groups);
  std::vector<char> from_disk(groups, 0);
  std::vector<std::thread> th;
  for (size_t g : todo)
    th.emplace_back([&, g] {
      if (!dir.empty()) {
        path[g] = disk_cache_file(dir, src[g]);
        if (disk_cache_read(path[g], code[g])) {
          from_disk[g] = 1;
          return;
        }
      }
      st[g] = qtile_jit_compile(src[g], code[g]);
      if (st[g] != QSV_OK) err[g] = get_error();
    });
  for (std::thread& t : th) t.join();
  for (size_t g : todo)
    if (st[g] != QSV_OK) return fail(st[g], err[g]);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (size_t g : todo) {
      auto m = std::make_shared<QJitModule>();
      if (from_disk[g] && hipModuleLoadData(&m->mod, code[g].data()) != hipSuccess) {
        // unreadable cached object: compile it again and replace the file
        (void)hipGetLastError();
        m->mod = nullptr;
        from_disk[g] = 0;
        QSV_TRY(qtile_jit_compile(src[g], code[g]));
      }
      if (!from_disk[g]) {
        QSV_CHECK_HIP(hipModuleLoadData(&m->mod, code[g].data()));
        ++g_compiles;
        if (!dir.empty()) disk_cache_write(dir, path[g], code[g]);
      } else {
        ++g_disk_hits;
      }
      mods[g] = m;
      if (g_cache.size() >= kCacheMax) g_cache.clear();
      g_cache[CacheKey{dev, src[g]}] = m;
    }
  }
  out.mods = mods;
  out.fn.resize(np);
  out.ntiles.resize(np);
  out.wsrc.resize(np);
  out.wdst.resize(np);
  out.threads = 1u << (qp.tile_bits - kQRegBits);
  for (size_t g = 0; g < groups; ++g)
    for (size_t i = cut[g]; i < cut[g + 1]; ++i) {
      const std::string name = prefix + std::to_string(i);
      QSV_CHECK_HIP(hipModuleGetFunction(&out.fn[i], mods[g]->mod, name.c_str()));
      out.ntiles[i] = qp.passes[i].ntiles;
      out.wsrc[i] = qp.passes[i].wsrc;
      out.wdst[i] = qp.passes[i].wdst;
    }
  return QSV_OK;
}

}  // namespace

int qtile_jit_build(const QProgram& qp, bool c128, QJit& out) {
  return build_modules(qp, pass_prefix(qp), [&](size_t p0, size_t p1) { return qtile_jit_source(qp, c128, p0, p1); },
                       out);
}

std::string qtile_jit_adjoint_source(const QProgram& qp, bool c128, size_t p0, size_t p1) {
  std::ostringstream s;
  s << "#define REAL " << (c128 ? "double" : "float") << "\n" << kPrelude << kAdjPrelude;
  for (size_t i = p0; i < p1 && i < qp.passes.size(); ++i) emit_adj_pass(s, qp, i, c128);
  return s.str();
}

int qtile_jit_adjoint_check(const QProgram& qp) {
  if (qtile_uses_work(qp)) return fail(QSV_EINVAL, "adjoint passes run in place");
  for (size_t i = 0; i < qp.passes.size(); ++i) {
    const int rows = adj_pass_rows(qp, i);
    const unsigned threads = 1u << (qp.tile_bits - kQRegBits);
    const size_t bytes = size_t(rows) * (rows <= 128 ? threads / 16 : threads / 64) * sizeof(double);
    if (bytes > kAdjAccBytesMax)
      return fail(QSV_EUNSUP, "adjoint pass with " + std::to_string(rows) + " generator rows exceeds LDS");
  }
  return QSV_OK;
}

int qtile_jit_build_adjoint(const QProgram& qp, bool c128, QJit& out) {
  QSV_TRY(qtile_jit_adjoint_check(qp));
  return build_modules(qp, "qta_",
                       [&](size_t p0, size_t p1) { return qtile_jit_adjoint_source(qp, c128, p0, p1); }, out);
}

int qtile_jit_launch_adjoint(const QJit& j, size_t i, void* phi, void* lam, const QOp* ops, double* part,
                             hipStream_t st) {
  if (i >= j.fn.size()) return fail(QSV_EINVAL, "adjoint pass index out of range");
  uint64_t ntiles = j.ntiles[i];
  const void* cf = ops;
  void* args[] = {&phi, &lam, &cf, &ntiles, &part};
  QSV_CHECK_HIP(hipModuleLaunchKernel(j.fn[i], kAdjPartStride, 1, 1, j.threads, 1, 1, 0, st, args, nullptr));
  return QSV_OK;
}

int qtile_jit_launch(const QJit& j, size_t i, void* psi, void* work, const QOp* ops, hipStream_t st) {
  if (i >= j.fn.size()) return fail(QSV_EINVAL, "tiled pass index out of range");
  uint64_t ntiles = j.ntiles[i];
  const uint64_t want = std::min(uint64_t(device_cus()) * kJitBlocksPerCU,
                                 std::max<uint64_t>(1, ntiles / kJitTilesPerBlock));
  const unsigned grid = unsigned(ntiles < want ? ntiles : want);
  const void* cf = ops;
  if (j.wsrc[i] || j.wdst[i]) {
    if (!work) return fail(QSV_EINVAL, "tiled pass needs the work buffer");
    const void* pin = j.wsrc[i] ? work : psi;
    void* pout = j.wdst[i] ? work : psi;
    void* args[] = {&pin, &pout, &cf, &ntiles};
    QSV_CHECK_HIP(hipModuleLaunchKernel(j.fn[i], grid, 1, 1, j.threads, 1, 1, 0, st, args, nullptr));
    return QSV_OK;
  }
  void* args[] = {&psi, &cf, &ntiles};
  QSV_CHECK_HIP(hipModuleLaunchKernel(j.fn[i], grid, 1, 1, j.threads, 1, 1, 0, st, args, nullptr));
  return QSV_OK;
}

}  // namespace qsv
