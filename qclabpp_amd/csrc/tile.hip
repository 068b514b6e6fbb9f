// LDS tile-pass kernels of the batch engine.
//
// One launch processes one dependency level: a flat list of work items
// (pass, tile).  Work items are handed out through a ticket counter
// (persistent blocks: the grid is sized to what the GPU holds at once, each
// block loops until the tickets run out, so every wave reaches the exit).
// An item loads its tile of 2^t amplitudes into LDS -- or starts from |0..0>
// on a circuit's first pass -- applies the pass's ops with a block barrier
// between ops, and stores the tile back.  Items of one level never depend on
// each other; levels are ordered by the stream (or the captured graph).
//
// When the last tile of a circuit's last pass is stored, the block stamps the
// circuit's completion time (wall_clock64) for latency statistics.
#include "tile.hpp"
#include "tile_launch.hpp"

namespace qsv {

namespace {

__device__ __forceinline__ uint64_t ins0(uint64_t k, int b) {
  const uint64_t lo = k & ((uint64_t(1) << b) - 1);
  return ((k ^ lo) << 1) | lo;
}

template <typename C>
__device__ __forceinline__ C coef(const TOp& o, int i) {
  return mk<C>(static_cast<decltype(C::x)>(o.m[2 * i]), static_cast<decltype(C::x)>(o.m[2 * i + 1]));
}

// bit of an element for operand spec s (local bit, or ~global bit)
__device__ __forceinline__ int obit(int s, uint32_t e, uint64_t base) {
  return s >= 0 ? int((e >> s) & 1u) : int((base >> (~s)) & 1u);
}

template <typename C>
__device__ void apply_top(const TOp& o, C* L, uint32_t size, uint64_t base) {
  const int tid = threadIdx.x;
  switch (o.kind) {
    case T_DENSE1: {
      const C m0 = coef<C>(o, 0), m1 = coef<C>(o, 1), m2 = coef<C>(o, 2), m3 = coef<C>(o, 3);
      const uint32_t s = 1u << o.a;
      for (uint32_t p = tid; p < size / 2; p += kTileThreads) {
        const uint32_t i0 = uint32_t(ins0(p, o.a)), i1 = i0 | s;
        const C x = L[i0], y = L[i1];
        L[i0] = cmad(m1, y, cmul(m0, x));
        L[i1] = cmad(m3, y, cmul(m2, x));
      }
      break;
    }
    case T_DIAG1: {
      const C d0 = coef<C>(o, 0), d1 = coef<C>(o, 1);
      for (uint32_t e = tid; e < size; e += kTileThreads) L[e] = cmul(obit(o.a, e, base) ? d1 : d0, L[e]);
      break;
    }
    case T_DIAG2: {
      const C d[4] = {coef<C>(o, 0), coef<C>(o, 1), coef<C>(o, 2), coef<C>(o, 3)};
      for (uint32_t e = tid; e < size; e += kTileThreads)
        L[e] = cmul(d[2 * obit(o.a, e, base) + obit(o.b, e, base)], L[e]);
      break;
    }
    case T_CDENSE1: {
      const C m0 = coef<C>(o, 0), m1 = coef<C>(o, 1), m2 = coef<C>(o, 2), m3 = coef<C>(o, 3);
      const uint32_t s = 1u << o.a;
      for (uint32_t p = tid; p < size / 2; p += kTileThreads) {
        const uint32_t i0 = uint32_t(ins0(p, o.a)), i1 = i0 | s;
        if (obit(o.b, i0, base) != o.cs) continue;
        const C x = L[i0], y = L[i1];
        L[i0] = cmad(m1, y, cmul(m0, x));
        L[i1] = cmad(m3, y, cmul(m2, x));
      }
      break;
    }
    case T_SWAP: {
      const C lam = coef<C>(o, 0);
      const int lo = o.a < o.b ? o.a : o.b, hi = o.a < o.b ? o.b : o.a;
      for (uint32_t p = tid; p < size / 4; p += kTileThreads) {
        const uint32_t i = uint32_t(ins0(ins0(p, lo), hi));
        const uint32_t ia = i | (1u << o.a), ib = i | (1u << o.b);
        const C x = L[ia], y = L[ib];
        L[ia] = cmul(lam, y);
        L[ib] = cmul(lam, x);
      }
      break;
    }
    case T_DENSE2: {
      const int lo = o.a < o.b ? o.a : o.b, hi = o.a < o.b ? o.b : o.a;
      for (uint32_t p = tid; p < size / 4; p += kTileThreads) {
        const uint32_t i = uint32_t(ins0(ins0(p, lo), hi));
        const uint32_t ix[4] = {i, i | (1u << o.b), i | (1u << o.a), i | (1u << o.a) | (1u << o.b)};
        C x[4];
        for (int q = 0; q < 4; ++q) x[q] = L[ix[q]];
        for (int r = 0; r < 4; ++r) {
          C acc = cmul(coef<C>(o, 4 * r), x[0]);
          for (int c = 1; c < 4; ++c) acc = cmad(coef<C>(o, 4 * r + c), x[c], acc);
          L[ix[r]] = acc;
        }
      }
      break;
    }
    default:
      break;
  }
}

__device__ __forceinline__ uint64_t deposit(uint64_t w, const TPass& P) {
  uint64_t a = 0;
  int sh = 0;
  for (int r = 0; r < P.nruns; ++r) {
    a |= ((w >> sh) & ((uint64_t(1) << P.rlen[r]) - 1)) << P.rdst[r];
    sh += P.rlen[r];
  }
  return a;
}

__device__ __forceinline__ uint64_t pdep_tile(uint32_t e, const TPass& P) {
  uint64_t off = 0;
  for (int i = 0; i < P.t; ++i)
    if (e >> i & 1u) off |= uint64_t(1) << P.tbits[i];
  return off;
}

template <typename C>
__global__ __launch_bounds__(kTileThreads) void k_tile_level(
    C* __restrict__ psi, const TCircuit* __restrict__ circs, const TPass* __restrict__ passes,
    const TOp* __restrict__ ops, const TItem* __restrict__ items, uint64_t nitems,
    unsigned long long* ticket, unsigned int* done_tiles, unsigned long long* stamps) {
  extern __shared__ unsigned char smem[];
  C* L = reinterpret_cast<C*>(smem);
  __shared__ unsigned long long cur;
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(ticket, 1ull);
    __syncthreads();
    const unsigned long long it = cur;
    __syncthreads();
    if (it >= nitems) break;
    const TItem item = items[it];
    const TPass P = passes[item.pass];
    C* state = psi + circs[P.circuit].offset;
    const uint64_t base = deposit(item.tile, P);
    const uint32_t size = 1u << P.t;
    for (uint32_t e = threadIdx.x; e < size; e += kTileThreads) {
      if (P.init_zero)
        L[e] = mk<C>((base == 0 && e == 0) ? 1 : 0, 0);
      else
        L[e] = state[base + pdep_tile(e, P)];
    }
    __syncthreads();
    for (int o = P.op_begin; o < P.op_end; ++o) {
      apply_top<C>(ops[o], L, size, base);
      __syncthreads();
    }
    for (uint32_t e = threadIdx.x; e < size; e += kTileThreads) state[base + pdep_tile(e, P)] = L[e];
    if (P.last) {
      __threadfence();
      __syncthreads();
      if (threadIdx.x == 0) {
        const unsigned int prev = atomicAdd(&done_tiles[P.circuit], 1u);
        if (uint64_t(prev) + 1 == P.ntiles) stamps[P.circuit] = wall_clock64();
      }
    }
    __syncthreads();
  }
}

// Reset the ticket and completion counters of one run and stamp its start;
// a kernel rather than a memset so the run can be captured in a graph.
__global__ void k_tile_prologue(unsigned long long* tickets, int ntickets, unsigned int* done,
                                int ncirc, unsigned long long* start_stamp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ntickets) tickets[i] = 0;
  if (i < ncirc) done[i] = 0;
  if (i == 0) *start_stamp = wall_clock64();
}

}  // namespace

template <typename R>
int launch_tile_level(const TileLevelArgs& a, hipStream_t st) {
  using C = typename CT<R>::type;
  const size_t lds = (size_t(1) << a.tmax) * sizeof(C);
  QSV_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_level<C>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL((k_tile_level<C>), dim3(a.grid), dim3(kTileThreads), lds, st,
                     static_cast<C*>(a.psi), a.circs, a.passes, a.ops, a.items, a.nitems, a.ticket,
                     a.done_tiles, a.stamps);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

int launch_tile_prologue(unsigned long long* tickets, int ntickets, unsigned int* done, int ncirc,
                         unsigned long long* start_stamp, hipStream_t st) {
  const int m = std::max(ntickets, ncirc);
  const int threads = 256;
  hipLaunchKernelGGL(k_tile_prologue, dim3((m + threads - 1) / threads > 0 ? (m + threads - 1) / threads : 1),
                     dim3(threads), 0, st, tickets, ntickets, done, ncirc, start_stamp);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template int launch_tile_level<double>(const TileLevelArgs&, hipStream_t);
template int launch_tile_level<float>(const TileLevelArgs&, hipStream_t);

}  // namespace qsv
