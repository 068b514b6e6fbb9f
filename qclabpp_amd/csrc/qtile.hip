// Tiled pass kernel: one launch streams the state through LDS tiles of
// 2^kQTileBits amplitudes and applies a pass's gates phase by phase, each
// phase in registers (16 amplitudes per thread over the phase's 4 bits).
#include "qtile.hpp"
#include "tile_launch.hpp"

namespace qsv {

namespace {

constexpr int NR = 1 << kQRegBits;

// Pass, phase and op descriptors are read through the constant address
// space: the loads are wave-uniform and the descriptors never change during
// a launch, so they become scalar loads into SGPRs.
#define QCONST __attribute__((address_space(4)))
typedef const QCONST QOp* QOpPtr;
typedef const QCONST QPhase* QPhasePtr;
typedef const QCONST QPass* QPassPtr;
constexpr int TE = 1 << kQTileBits;
static_assert(TE == kQThreads * NR, "one tile = one register set per thread");


__device__ __forceinline__ uint32_t pdep12(uint32_t x, uint32_t mask) {
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < kQTileBits; ++b)
    if (mask >> b & 1u) {
      r |= (x & 1u) << b;
      x >>= 1;
    }
  return r;
}

// Coefficient i of an op, from the table of the state's precision.
template <typename C>
__device__ __forceinline__ C qc(const QCONST QOp& o, int i) {
  if constexpr (sizeof(C) == sizeof(double2))
    return mk<C>(o.m[2 * i], o.m[2 * i + 1]);
  else
    return mk<C>(o.mf[2 * i], o.mf[2 * i + 1]);
}
template <typename C>
__device__ __forceinline__ decltype(C::x) qre(const QCONST QOp& o, int i) {
  if constexpr (sizeof(C) == sizeof(double2)) return o.m[i];
  else return o.mf[i];
}

// Results are chosen with value selects, never with conditional stores, so
// v stays in registers.  Selects on wave-uniform values stay scalar.
template <typename C>
__device__ __forceinline__ C qsel(bool c, C a, C b) {
  C r;
  r.x = c ? a.x : b.x;
  r.y = c ? a.y : b.y;
  return r;
}

// Pair ops over the register pairs (p, p ^ D), p the base (bit lowbit(D)
// clear).  `roles` bit p: base p holds the |1> side, so the 2x2 runs with
// rows and columns swapped; the choice is a select on uniform coefficients
// (scalar instructions), compiled out (SEL false) for ops with no roles.
template <int D, bool SEL, typename C>
__device__ __forceinline__ void q_pair(C (&v)[NR], const QCONST QOp& o) {
  constexpr int K = D & -D;
  const C m0 = qc<C>(o, 0), m1 = qc<C>(o, 1), m2 = qc<C>(o, 2), m3 = qc<C>(o, 3);
  const uint32_t roles = o.roles;
#pragma unroll
  for (int p = 0; p < NR; ++p)
    if (!(p & K)) {
      const bool sw = SEL && (roles >> p & 1);
      const C a = qsel(sw, m3, m0), b = qsel(sw, m2, m1), c = qsel(sw, m1, m2), d = qsel(sw, m0, m3);
      const C x = v[p], y = v[p ^ D];
      v[p] = cmad(b, y, cmul(a, x));
      v[p ^ D] = cmad(d, y, cmul(c, x));
    }
}

// Q_PAIR on the pairs whose base has its `act` bit set
template <int D, typename C>
__device__ __forceinline__ void q_cpair(C (&v)[NR], const QCONST QOp& o) {
  constexpr int K = D & -D;
  const C m0 = qc<C>(o, 0), m1 = qc<C>(o, 1), m2 = qc<C>(o, 2), m3 = qc<C>(o, 3);
  const uint32_t roles = o.roles, act = o.act;
#pragma unroll
  for (int p = 0; p < NR; ++p)
    if (!(p & K)) {
      const bool sw = roles >> p & 1, on = act >> p & 1;
      const C a = qsel(sw, m3, m0), b = qsel(sw, m2, m1), c = qsel(sw, m1, m2), d = qsel(sw, m0, m3);
      const C x = v[p], y = v[p ^ D];
      v[p] = qsel(on, cmad(b, y, cmul(a, x)), x);
      v[p ^ D] = qsel(on, cmad(d, y, cmul(c, x)), y);
    }
}

// real 2x2 (4 FP ops per amplitude)
template <int D, bool SEL, typename C>
__device__ __forceinline__ void q_realp(C (&v)[NR], const QCONST QOp& o) {
  using R = decltype(C::x);
  constexpr int K = D & -D;
  const R m0 = qre<C>(o, 0), m1 = qre<C>(o, 2), m2 = qre<C>(o, 4), m3 = qre<C>(o, 6);
  const uint32_t roles = o.roles;
#pragma unroll
  for (int p = 0; p < NR; ++p)
    if (!(p & K)) {
      const bool sw = SEL && (roles >> p & 1);
      const R a = sw ? m3 : m0, b = sw ? m2 : m1, c = sw ? m1 : m2, d = sw ? m0 : m3;
      const C x = v[p], y = v[p ^ D];
      v[p] = mk<C>(a * x.x + b * y.x, a * x.y + b * y.y);
      v[p ^ D] = mk<C>(c * x.x + d * y.x, c * x.y + d * y.y);
    }
}

// real rotation [[c, -s], [s, c]]: swapping rows and columns negates s
template <int D, bool SEL, typename C>
__device__ __forceinline__ void q_rotp(C (&v)[NR], const QCONST QOp& o) {
  using R = decltype(C::x);
  constexpr int K = D & -D;
  const R c = qre<C>(o, 0), s = qre<C>(o, 4);
  const uint32_t roles = o.roles;
#pragma unroll
  for (int p = 0; p < NR; ++p)
    if (!(p & K)) {
      const R t = SEL && (roles >> p & 1) ? -s : s;
      const C x = v[p], y = v[p ^ D];
      v[p] = mk<C>(c * x.x - t * y.x, c * x.y - t * y.y);
      v[p ^ D] = mk<C>(t * x.x + c * y.x, t * x.y + c * y.y);
    }
}

// [[p, iq], [ir, s]] with p, q, r, s real (4 FP ops per amplitude);
// swapping rows and columns exchanges p with s and q with r
template <int D, bool SEL, typename C>
__device__ __forceinline__ void q_xlikep(C (&v)[NR], const QCONST QOp& o) {
  using R = decltype(C::x);
  constexpr int K = D & -D;
  const R P = qre<C>(o, 0), Q = qre<C>(o, 3), Rr = qre<C>(o, 5), S = qre<C>(o, 6);
  const uint32_t roles = o.roles;
#pragma unroll
  for (int p = 0; p < NR; ++p)
    if (!(p & K)) {
      const bool sw = SEL && (roles >> p & 1);
      const R pp = sw ? S : P, qq = sw ? Rr : Q, rr = sw ? Q : Rr, ss = sw ? P : S;
      const C x = v[p], y = v[p ^ D];
      v[p] = mk<C>(pp * x.x - qq * y.y, pp * x.y + qq * y.x);
      v[p ^ D] = mk<C>(ss * y.x - rr * x.y, ss * y.y + rr * x.x);
    }
}

// scaled rotation (Q_ROTS): [[1, -k], [k, 1]] or [[k, -1], [1, k]] as re(m1)
// is 1 or 0; swapping rows and columns negates the off-diagonal entries
template <int D, bool SEL, typename C>
__device__ __forceinline__ void q_rots(C (&v)[NR], const QCONST QOp& o) {
  using R = decltype(C::x);
  constexpr int K = D & -D;
  const R k = qre<C>(o, 0);
  const uint32_t roles = o.roles;
  if (qre<C>(o, 2) != R(0)) {
#pragma unroll
    for (int p = 0; p < NR; ++p)
      if (!(p & K)) {
        const R t = SEL && (roles >> p & 1) ? -k : k;
        const C x = v[p], y = v[p ^ D];
        v[p] = mk<C>(x.x - t * y.x, x.y - t * y.y);
        v[p ^ D] = mk<C>(y.x + t * x.x, y.y + t * x.y);
      }
  } else {
#pragma unroll
    for (int p = 0; p < NR; ++p)
      if (!(p & K)) {
        const R g = SEL && (roles >> p & 1) ? R(-1) : R(1);
        const C x = v[p], y = v[p ^ D];
        v[p] = mk<C>(k * x.x - g * y.x, k * x.y - g * y.y);
        v[p ^ D] = mk<C>(k * y.x + g * x.x, k * y.y + g * x.y);
      }
  }
}

// scaled RX-like (Q_XLS): [[1, ik], [ik, 1]] or [[k, i], [i, k]] as re(m1)
// is 1 or 0 (symmetric, so roles do not matter)
template <int D, bool SEL, typename C>
__device__ __forceinline__ void q_xls(C (&v)[NR], const QCONST QOp& o) {
  using R = decltype(C::x);
  constexpr int K = D & -D;
  const R k = qre<C>(o, 0);
  if (qre<C>(o, 2) != R(0)) {
#pragma unroll
    for (int p = 0; p < NR; ++p)
      if (!(p & K)) {
        const C x = v[p], y = v[p ^ D];
        v[p] = mk<C>(x.x - k * y.y, x.y + k * y.x);
        v[p ^ D] = mk<C>(y.x - k * x.y, y.y + k * x.x);
      }
  } else {
#pragma unroll
    for (int p = 0; p < NR; ++p)
      if (!(p & K)) {
        const C x = v[p], y = v[p ^ D];
        v[p] = mk<C>(k * x.x - y.y, k * x.y + y.x);
        v[p ^ D] = mk<C>(k * y.x - x.y, k * y.y + x.x);
      }
  }
}

// 4x4 on register bits (A, B), row index 2*bit(A) + bit(B)
template <int A, int B, typename C>
__device__ __forceinline__ void q_dense2(C (&v)[NR], const QCONST QOp& o) {
#pragma unroll
  for (int i = 0; i < NR; ++i)
    if (!(i & (1 << A)) && !(i & (1 << B))) {
      const int ix[4] = {i, i | (1 << B), i | (1 << A), i | (1 << A) | (1 << B)};
      C x[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) x[p] = v[ix[p]];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        C acc = cmul(qc<C>(o, 4 * r), x[0]);
#pragma unroll
        for (int c = 1; c < 4; ++c) acc = cmad(qc<C>(o, 4 * r + c), x[c], acc);
        v[ix[r]] = acc;
      }
    }
}

#define QPAIRS(X) \
  X(0, 1) X(0, 2) X(0, 3) X(1, 0) X(1, 2) X(1, 3) X(2, 0) X(2, 1) X(2, 3) X(3, 0) X(3, 1) X(3, 2)
#define QDISTS(X) \
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15)

// FN<D, SEL> for the op's distance, with selects only if some base has a role
#define QBYDIST_SEL(FN, SEL)                                                    \
  switch (o.d) {                                                                \
    case 1: FN<1, SEL>(v, o); break;    case 2: FN<2, SEL>(v, o); break;        \
    case 3: FN<3, SEL>(v, o); break;    case 4: FN<4, SEL>(v, o); break;        \
    case 5: FN<5, SEL>(v, o); break;    case 6: FN<6, SEL>(v, o); break;        \
    case 7: FN<7, SEL>(v, o); break;    case 8: FN<8, SEL>(v, o); break;        \
    case 9: FN<9, SEL>(v, o); break;    case 10: FN<10, SEL>(v, o); break;      \
    case 11: FN<11, SEL>(v, o); break;  case 12: FN<12, SEL>(v, o); break;      \
    case 13: FN<13, SEL>(v, o); break;  case 14: FN<14, SEL>(v, o); break;      \
    case 15: FN<15, SEL>(v, o); break;  default: break;                         \
  }
#define QBYDIST(FN)          \
  if (o.roles) {             \
    QBYDIST_SEL(FN, true)    \
  } else {                   \
    QBYDIST_SEL(FN, false)   \
  }

template <typename C>
__device__ __forceinline__ void apply_op(const QCONST QOp& o, C (&v)[NR]) {
  switch (o.kind) {
    case Q_DIAGR: {
      const int mask = o.b;  // factors other than 1
#pragma unroll
      for (int i = 0; i < NR; ++i)
        if (mask >> i & 1) v[i] = cmul(qc<C>(o, i), v[i]);
      break;
    }
    case Q_ROTP:
      QBYDIST(q_rotp);
      break;
    case Q_XLIKEP:
      QBYDIST(q_xlikep);
      break;
    case Q_ROTS:
      QBYDIST(q_rots);
      break;
    case Q_XLS:
      QBYDIST_SEL(q_xls, false);
      break;
    case Q_REALP:
      QBYDIST(q_realp);
      break;
    case Q_PAIR:
      QBYDIST(q_pair);
      break;
    case Q_CPAIR:
      switch (o.d) {
#define X(D) case D: q_cpair<D>(v, o); break;
        QDISTS(X)
#undef X
        default: break;
      }
      break;
    case Q_DENSE2: {
      switch (o.a * 4 + o.b) {
#define X(A, B) case A * 4 + B: q_dense2<A, B>(v, o); break;
        QPAIRS(X)
#undef X
        default: break;
      }
      break;
    }
    default:
      break;
  }
}

// Two waves per SIMD (at most 256 registers): a third of the time goes to
// LDS and HBM waits, and a single wave per SIMD cannot cover them.
template <typename R>
__global__ __launch_bounds__(kQThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_qtile(typename CT<R>::type* __restrict__ psi,
                                                     QPassPtr pass, QPhasePtr phases, QOpPtr ops,
                                                     uint64_t ntiles) {
  using C = typename CT<R>::type;
  __shared__ C tile[TE];
  const int tid = threadIdx.x;
  // global offset of tile position p = tid + 256*j: the low 8 position bits
  // come from the thread, the high 4 from j
  uint64_t lo = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (tid >> i & 1) lo |= uint64_t(1) << pass->tbits[i];
  uint64_t hb[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) hb[i] = uint64_t(1) << pass->tbits[8 + i];
  auto gofs = [&](int j) {
    uint64_t o = lo;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (j >> i & 1) o |= hb[i];
    return o;
  };
  const int phb = pass->phase_begin, phe = pass->phase_end, nruns = pass->nruns;
  const uint32_t tslot = qtile_swizzle(uint32_t(tid));  // slot part of this thread's tile positions

  auto tile_base = [&](uint64_t t) {
    uint64_t base = 0;
    int sh = 0;
    for (int r = 0; r < nruns; ++r) {
      const int len = pass->rlen[r];
      base |= ((t >> sh) & ((uint64_t(1) << len) - 1)) << pass->rdst[r];
      sh += len;
    }
    return base;
  };
  // Software pipeline: the next tile's amplitudes are loaded into registers
  // while the current tile's phases run, so HBM and VALU work overlap within
  // the block (blocks running the same code in step would not overlap them).
  // Each block walks a contiguous range of tiles (measured up to 20% more
  // HBM bandwidth than interleaved tiles on scattered tile layouts).
  const uint64_t per = (ntiles + gridDim.x - 1) / gridDim.x;
  uint64_t t = uint64_t(blockIdx.x) * per;
  const uint64_t tend = t + per < ntiles ? t + per : ntiles;
  C pre[NR];
  if (t < tend) {
    const uint64_t b0 = tile_base(t);
#pragma unroll
    for (int j = 0; j < NR; ++j) pre[j] = ld_stream(&psi[b0 + gofs(j)]);
  }
  for (; t < tend; ++t) {
    const uint64_t base = tile_base(t);
#pragma unroll
    for (int j = 0; j < NR; ++j) tile[tslot ^ qtile_swizzle(uint32_t(kQThreads * j))] = pre[j];
    __syncthreads();
    const uint64_t tn = t + 1;
    if (tn < tend) {
      const uint64_t bn = tile_base(tn);
#pragma unroll
      for (int j = 0; j < NR; ++j) pre[j] = ld_stream(&psi[bn + gofs(j)]);
    }
    for (int ph = phb; ph < phe; ++ph) {
      const QCONST QPhase& P = phases[ph];
      const uint32_t tp = qtile_swizzle(pdep12(uint32_t(tid), ~P.regmask & uint32_t(TE - 1)));
      C v[NR];
#pragma unroll
      for (int i = 0; i < NR; ++i) v[i] = tile[tp ^ P.regpos[i]];
      const int ob = P.op_begin, oe = P.op_end;
      for (int o = ob; o < oe; ++o) apply_op(ops[o], v);
#pragma unroll
      for (int i = 0; i < NR; ++i) tile[tp ^ P.regpos_out[i]] = v[i];
      __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) st_stream(&psi[base + gofs(j)], tile[tslot ^ qtile_swizzle(uint32_t(kQThreads * j))]);
    __syncthreads();
  }
}

// One dependency level of the batch engine's register-phase circuits: the
// passes of k_qtile over many programs.  Persistent blocks take (pass, tile)
// work items from a ticket counter; an item loads its tile (a circuit's first
// pass sets |0...0> instead: the amplitude 1 sits at position 0 of tile 0),
// runs the pass's phases as k_qtile does, and stores the tile.  The block
// that stores the last tile of a circuit's last pass stamps its completion.
// Items of one level never depend on each other.  Each block takes its next
// ticket when it starts an item and loads that item's amplitudes into
// registers while the current item's phases run.
template <typename R>
__global__ __launch_bounds__(kQThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_qtile_level(
    typename CT<R>::type* __restrict__ psi, const TCircuit* __restrict__ circs, QPassPtr passes,
    const QPassInfo* __restrict__ pinfo, QPhasePtr phases, QOpPtr ops, const TItem* __restrict__ items,
    uint64_t nitems, unsigned long long* ticket, unsigned int* done_tiles, unsigned long long* stamps) {
  using C = typename CT<R>::type;
  __shared__ C tile[TE];
  __shared__ unsigned long long next_ticket;
  const int tid = threadIdx.x;
  const uint32_t tslot = qtile_swizzle(uint32_t(tid));
  // the ticket read from LDS is the same in every lane: made scalar, the
  // item, pass and phase descriptors are read with scalar loads
  auto uniform = [](unsigned long long x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
    return (uint64_t(hi) << 32) | lo;
  };
  // global offset of this thread's tile position tid + 256*j for an item:
  // the tile's base and the thread's low 8 bits (lo), then j's 4 bits
  struct Where {
    uint64_t lo;
    uint64_t hb[4];
    __device__ uint64_t at(int j) const {
      uint64_t x = lo;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (j >> i & 1) x |= hb[i];
      return x;
    }
  };
  auto where = [&](const TItem& item) {
    const QCONST QPass& pass = passes[item.pass];
    Where w;
    uint64_t o = 0;
    int sh = 0;
    for (int r = 0; r < pass.nruns; ++r) {
      const int len = pass.rlen[r];
      o |= ((item.tile >> sh) & ((uint64_t(1) << len) - 1)) << pass.rdst[r];
      sh += len;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (tid >> i & 1) o |= uint64_t(1) << pass.tbits[i];
    w.lo = o;
#pragma unroll
    for (int i = 0; i < 4; ++i) w.hb[i] = uint64_t(1) << pass.tbits[8 + i];
    return w;
  };
  auto load = [&](uint64_t it, C (&pre)[NR]) {
    if (it >= nitems) return;
    const TItem item = items[it];
    const QPassInfo pi = pinfo[item.pass];
    if (pi.first) return;
    const C* state = psi + circs[pi.circuit].offset;
    const Where w = where(item);
#pragma unroll
    for (int j = 0; j < NR; ++j) pre[j] = ld_stream(&state[w.at(j)]);
  };
  if (tid == 0) next_ticket = atomicAdd(ticket, 1ull);
  __syncthreads();
  uint64_t it = uniform(next_ticket);
  C pre[NR];
  load(it, pre);
  while (it < nitems) {
    __syncthreads();  // every lane has read next_ticket
    if (tid == 0) next_ticket = atomicAdd(ticket, 1ull);
    const TItem item = items[it];
    const QCONST QPass& pass = passes[item.pass];
    const QPassInfo pi = pinfo[item.pass];
#pragma unroll
    for (int j = 0; j < NR; ++j)
      tile[tslot ^ qtile_swizzle(uint32_t(kQThreads * j))] =
          pi.first ? mk<C>(item.tile == 0 && tid == 0 && j == 0 ? 1 : 0, 0) : pre[j];
    __syncthreads();
    const uint64_t itn = uniform(next_ticket);
    load(itn, pre);
    for (int ph = pass.phase_begin; ph < pass.phase_end; ++ph) {
      const QCONST QPhase& P = phases[ph];
      const uint32_t tp = qtile_swizzle(pdep12(uint32_t(tid), ~P.regmask & uint32_t(TE - 1)));
      C v[NR];
#pragma unroll
      for (int i = 0; i < NR; ++i) v[i] = tile[tp ^ P.regpos[i]];
      const int ob = P.op_begin, oe = P.op_end;
      for (int o = ob; o < oe; ++o) apply_op(ops[o], v);
#pragma unroll
      for (int i = 0; i < NR; ++i) tile[tp ^ P.regpos_out[i]] = v[i];
      __syncthreads();
    }
    C* state = psi + circs[pi.circuit].offset;
    const Where w = where(item);
#pragma unroll
    for (int j = 0; j < NR; ++j) st_stream(&state[w.at(j)], tile[tslot ^ qtile_swizzle(uint32_t(kQThreads * j))]);
    if (pi.last) {
      __threadfence();
      __syncthreads();
      if (tid == 0) {
        const unsigned int prev = atomicAdd(&done_tiles[pi.circuit], 1u);
        if (int64_t(prev) + 1 == int64_t(pi.nitems)) stamps[pi.circuit] = wall_clock64();
      }
    }
    __syncthreads();  // the tile is free again before the next item
    it = itn;
  }
}

// Zeros over the states of the listed circuits (blockIdx.y picks one), 16
// bytes per store: the batch's register-phase circuits before their first
// passes, which then only compute the tile that holds |0...0>.
template <typename R>
__global__ __launch_bounds__(256) void k_zero_states(typename CT<R>::type* __restrict__ psi,
                                                     const TCircuit* __restrict__ circs,
                                                     const int32_t* __restrict__ ids) {
  using C = typename CT<R>::type;
  const TCircuit c = circs[ids[blockIdx.y]];
  C* state = psi + c.offset;
  constexpr int per = 16 / int(sizeof(C));  // amplitudes per 16-byte store
  const uint64_t nvec = (uint64_t(1) << c.n) / per;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nvec;
       i += uint64_t(gridDim.x) * blockDim.x) {
    if constexpr (per == 1) {
      st_stream(&state[i], mk<C>(0, 0));
    } else {
      st_stream(&state[2 * i], mk<C>(0, 0));
      st_stream(&state[2 * i + 1], mk<C>(0, 0));
    }
  }
}

}  // namespace

// Blocks per CU in the grid: 2 are resident at a time (two waves per SIMD,
// 256 threads per block); launching 16 times more lets blocks start and end
// out of step, so one block's HBM phase overlaps another's VALU phase
// (measured: 3% faster than a grid of exactly the resident blocks).
constexpr int kQBlocksPerCU = 32;

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0, c = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && c > 0)
      cus = c;
    else
      cus = 256;
  }
  return cus;
}

template <typename R>
int launch_qtile(const QTileLaunch& a, hipStream_t st) {
  using C = typename CT<R>::type;
  const uint64_t want = uint64_t(device_cus()) * kQBlocksPerCU;
  const uint64_t grid = a.ntiles < want ? a.ntiles : want;
  hipLaunchKernelGGL((k_qtile<R>), dim3(unsigned(grid)), dim3(kQThreads), 0, st,
                     static_cast<C*>(a.psi), (QPassPtr)a.pass, (QPhasePtr)a.phases, (QOpPtr)a.ops,
                     a.ntiles);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template int launch_qtile<double>(const QTileLaunch&, hipStream_t);
template int launch_qtile<float>(const QTileLaunch&, hipStream_t);

template <typename R>
int launch_qtile_level(const QTileLevelArgs& a, hipStream_t st) {
  using C = typename CT<R>::type;
  hipLaunchKernelGGL((k_qtile_level<R>), dim3(a.grid), dim3(kQThreads), 0, st, static_cast<C*>(a.psi), a.circs,
                     (QPassPtr)a.passes, a.pinfo, (QPhasePtr)a.phases, (QOpPtr)a.ops, a.items, a.nitems, a.ticket,
                     a.done_tiles, a.stamps);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template int launch_qtile_level<double>(const QTileLevelArgs&, hipStream_t);
template int launch_qtile_level<float>(const QTileLevelArgs&, hipStream_t);

template <typename R>
int launch_zero_states(void* psi, const TCircuit* circs, const int32_t* ids, int nids, unsigned blocks_per,
                       hipStream_t st) {
  using C = typename CT<R>::type;
  if (nids <= 0) return QSV_OK;
  hipLaunchKernelGGL((k_zero_states<R>), dim3(blocks_per, unsigned(nids)), dim3(256), 0, st, static_cast<C*>(psi),
                     circs, ids);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template int launch_zero_states<double>(void*, const TCircuit*, const int32_t*, int, unsigned, hipStream_t);
template int launch_zero_states<float>(void*, const TCircuit*, const int32_t*, int, unsigned, hipStream_t);

}  // namespace qsv
