// Tiled pass kernel: one launch streams the state through LDS tiles of
// 2^kQTileBits amplitudes and applies a pass's gates phase by phase, each
// phase in registers (16 amplitudes per thread over the phase's 4 bits).
#include "qtile.hpp"

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

template <typename C>
__device__ __forceinline__ C qc(const QCONST QOp& o, int i) {
  using R = decltype(C::x);
  return mk<C>(R(o.m[2 * i]), R(o.m[2 * i + 1]));
}

template <int J, typename C>
__device__ __forceinline__ void q_dense1(C (&v)[NR], C m0, C m1, C m2, C m3) {
#pragma unroll
  for (int i = 0; i < NR; ++i)
    if (!(i & (1 << J))) {
      const C x = v[i], y = v[i | (1 << J)];
      v[i] = cmad(m1, y, cmul(m0, x));
      v[i | (1 << J)] = cmad(m3, y, cmul(m2, x));
    }
}

// The control state is a run-time value: results are chosen with value
// selects, never with conditional stores, so v stays in registers.
template <typename C>
__device__ __forceinline__ C qsel(bool c, C a, C b) {
  C r;
  r.x = c ? a.x : b.x;
  r.y = c ? a.y : b.y;
  return r;
}

template <int TB, int CB, typename C>
__device__ __forceinline__ void q_cdense1(C (&v)[NR], C m0, C m1, C m2, C m3, int cs) {
#pragma unroll
  for (int i = 0; i < NR; ++i)
    if (!(i & (1 << TB))) {
      const bool on = ((i >> CB) & 1) == cs;
      const C x = v[i], y = v[i | (1 << TB)];
      v[i] = qsel(on, cmad(m1, y, cmul(m0, x)), x);
      v[i | (1 << TB)] = qsel(on, cmad(m3, y, cmul(m2, x)), y);
    }
}

template <int TB, int CB, int CS, typename C>
__device__ __forceinline__ void q_cx(C (&v)[NR]) {
#pragma unroll
  for (int i = 0; i < NR; ++i)
    if (!(i & (1 << TB)) && ((i >> CB) & 1) == CS) {
      const C x = v[i];
      v[i] = v[i | (1 << TB)];
      v[i | (1 << TB)] = x;
    }
}

// real 2x2 (RotationY): 4 FP ops per amplitude
template <int J, typename C>
__device__ __forceinline__ void q_real1(C (&v)[NR], decltype(C::x) a, decltype(C::x) b,
                                        decltype(C::x) c, decltype(C::x) d) {
#pragma unroll
  for (int i = 0; i < NR; ++i)
    if (!(i & (1 << J))) {
      const C x = v[i], y = v[i | (1 << J)];
      v[i] = mk<C>(a * x.x + b * y.x, a * x.y + b * y.y);
      v[i | (1 << J)] = mk<C>(c * x.x + d * y.x, c * x.y + d * y.y);
    }
}

// [[p, iq], [ir, s]] with p, q, r, s real (RotationX): 4 FP ops per amplitude
template <int J, typename C>
__device__ __forceinline__ void q_xlike1(C (&v)[NR], decltype(C::x) p, decltype(C::x) q,
                                         decltype(C::x) r, decltype(C::x) s) {
#pragma unroll
  for (int i = 0; i < NR; ++i)
    if (!(i & (1 << J))) {
      const C x = v[i], y = v[i | (1 << J)];
      v[i] = mk<C>(p * x.x - q * y.y, p * x.y + q * y.x);
      v[i | (1 << J)] = mk<C>(s * y.x - r * x.y, s * y.y + r * x.x);
    }
}

// exchange |A=1,B=0> with |A=0,B=1>, both times lam
template <int A, int B, typename C>
__device__ __forceinline__ void q_swap(C (&v)[NR], C lam, bool scale) {
#pragma unroll
  for (int i = 0; i < NR; ++i)
    if ((i & (1 << A)) && !(i & (1 << B))) {
      const int j = i ^ (1 << A) ^ (1 << B);
      const C x = v[i], y = v[j];
      v[i] = qsel(scale, cmul(lam, y), y);
      v[j] = qsel(scale, cmul(lam, x), x);
    }
}

// 4x4 on (A, B), row index 2*bit(A) + bit(B)
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

template <typename C>
__device__ __forceinline__ void apply_op(const QCONST QOp& o, C (&v)[NR]) {
  switch (o.kind) {
    case Q_DENSE1: {
      const C m0 = qc<C>(o, 0), m1 = qc<C>(o, 1), m2 = qc<C>(o, 2), m3 = qc<C>(o, 3);
      switch (o.a) {
        case 0: q_dense1<0>(v, m0, m1, m2, m3); break;
        case 1: q_dense1<1>(v, m0, m1, m2, m3); break;
        case 2: q_dense1<2>(v, m0, m1, m2, m3); break;
        default: q_dense1<3>(v, m0, m1, m2, m3); break;
      }
      break;
    }
    case Q_DIAGR: {
      const int mask = o.b;  // factors other than 1
#pragma unroll
      for (int i = 0; i < NR; ++i)
        if (mask >> i & 1) v[i] = cmul(qc<C>(o, i), v[i]);
      break;
    }
    case Q_REAL1: {
      using R = decltype(C::x);
      const R a = R(o.m[0]), b = R(o.m[2]), c = R(o.m[4]), d = R(o.m[6]);
      switch (o.a) {
        case 0: q_real1<0>(v, a, b, c, d); break;
        case 1: q_real1<1>(v, a, b, c, d); break;
        case 2: q_real1<2>(v, a, b, c, d); break;
        default: q_real1<3>(v, a, b, c, d); break;
      }
      break;
    }
    case Q_XLIKE1: {
      using R = decltype(C::x);
      const R p = R(o.m[0]), q = R(o.m[3]), r = R(o.m[5]), ss = R(o.m[6]);
      switch (o.a) {
        case 0: q_xlike1<0>(v, p, q, r, ss); break;
        case 1: q_xlike1<1>(v, p, q, r, ss); break;
        case 2: q_xlike1<2>(v, p, q, r, ss); break;
        default: q_xlike1<3>(v, p, q, r, ss); break;
      }
      break;
    }
    case Q_CDENSE1: {
      const C m0 = qc<C>(o, 0), m1 = qc<C>(o, 1), m2 = qc<C>(o, 2), m3 = qc<C>(o, 3);
      switch (o.a * 4 + o.b) {
#define X(A, B) case A * 4 + B: q_cdense1<A, B>(v, m0, m1, m2, m3, o.cs); break;
        QPAIRS(X)
#undef X
        default: break;
      }
      break;
    }
    case Q_CX: {
      switch ((o.a * 4 + o.b) * 2 + o.cs) {
#define X(A, B) case (A * 4 + B) * 2: q_cx<A, B, 0>(v); break; \
                case (A * 4 + B) * 2 + 1: q_cx<A, B, 1>(v); break;
        QPAIRS(X)
#undef X
        default: break;
      }
      break;
    }
    case Q_SWAP: {
      const C lam = qc<C>(o, 0);
      const bool scale = !(o.m[0] == 1.0 && o.m[1] == 0.0);
      switch (o.a * 4 + o.b) {
#define X(A, B) case A * 4 + B: q_swap<A, B>(v, lam, scale); break;
        QPAIRS(X)
#undef X
        default: break;
      }
      break;
    }
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
  uint64_t t = blockIdx.x;
  C pre[NR];
  if (t < ntiles) {
    const uint64_t b0 = tile_base(t);
#pragma unroll
    for (int j = 0; j < NR; ++j) pre[j] = ld_stream(&psi[b0 + gofs(j)]);
  }
  for (; t < ntiles; t += gridDim.x) {
    const uint64_t base = tile_base(t);
#pragma unroll
    for (int j = 0; j < NR; ++j) tile[tslot ^ qtile_swizzle(uint32_t(kQThreads * j))] = pre[j];
    __syncthreads();
    const uint64_t tn = t + gridDim.x;
    if (tn < ntiles) {
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
      for (int i = 0; i < NR; ++i) tile[tp ^ P.regpos[i]] = v[i];
      __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) st_stream(&psi[base + gofs(j)], tile[tslot ^ qtile_swizzle(uint32_t(kQThreads * j))]);
    __syncthreads();
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

}  // namespace qsv
