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
      const C d0 = coef<C>(o, 0), d1 = coef<C>(o, 1), d2 = coef<C>(o, 2), d3 = coef<C>(o, 3);
      for (uint32_t e = tid; e < size; e += kTileThreads) {
        const int ba = obit(o.a, e, base), bb = obit(o.b, e, base);
        const C d = ba ? (bb ? d3 : d2) : (bb ? d1 : d0);
        L[e] = cmul(d, L[e]);
      }
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
        const uint32_t i1 = i | (1u << o.b), i2 = i | (1u << o.a), i3 = i2 | (1u << o.b);
        const C x0 = L[i], x1 = L[i1], x2 = L[i2], x3 = L[i3];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          C acc = cmul(coef<C>(o, 4 * r), x0);
          acc = cmad(coef<C>(o, 4 * r + 1), x1, acc);
          acc = cmad(coef<C>(o, 4 * r + 2), x2, acc);
          acc = cmad(coef<C>(o, 4 * r + 3), x3, acc);
          L[r == 0 ? i : r == 1 ? i1 : r == 2 ? i2 : i3] = acc;
        }
      }
      break;
    }
    default:
      break;
  }
}

template <typename C>
__global__ __launch_bounds__(kTileThreads) void k_tile_level(
    C* __restrict__ psi, const TCircuit* __restrict__ circs, const TPass* __restrict__ passes,
    const TOp* __restrict__ ops, const TItem* __restrict__ items, uint64_t nitems,
    unsigned long long* ticket, unsigned int* done_tiles, unsigned long long* stamps) {
  extern __shared__ unsigned char smem[];
  C* L = reinterpret_cast<C*>(smem);
  __shared__ unsigned long long cur;
  __shared__ TPass sp;  // the current item's pass (dynamically indexed in LDS, not in registers)
  for (;;) {
    if (threadIdx.x == 0) cur = atomicAdd(ticket, 1ull);
    __syncthreads();
    const unsigned long long it = cur;
    if (it >= nitems) break;
    const TItem item = items[it];
    if (threadIdx.x == 0) sp = passes[item.pass];
    __syncthreads();
    const int t = sp.t;
    C* state = psi + circs[sp.circuit].offset;
    uint64_t base = 0;
    {
      int sh = 0;
      for (int r = 0; r < sp.nruns; ++r) {
        base |= ((item.tile >> sh) & ((uint64_t(1) << sp.rlen[r]) - 1)) << sp.rdst[r];
        sh += sp.rlen[r];
      }
    }
    const uint32_t size = 1u << t;
    // element e = threadIdx.x + 256*k: offset = pdep(low 8 bits) | pdep(k << 8)
    uint64_t lane_off = 0;
    for (int i = 0; i < t && i < 8; ++i)
      if (threadIdx.x >> i & 1u) lane_off |= uint64_t(1) << sp.tbits[i];
    const int init = sp.init_zero;
    for (uint32_t e = threadIdx.x, k = 0; e < size; e += kTileThreads, ++k) {
      uint64_t off = lane_off;
      for (int i = 8; i < t; ++i)
        if (k >> (i - 8) & 1u) off |= uint64_t(1) << sp.tbits[i];
      L[e] = init ? mk<C>((base == 0 && e == 0) ? 1 : 0, 0) : state[base + off];
    }
    __syncthreads();
    const int ob = sp.op_begin, oe = sp.op_end;
    for (int o = ob; o < oe; ++o) {
      apply_top<C>(ops[o], L, size, base);
      __syncthreads();
    }
    for (uint32_t e = threadIdx.x, k = 0; e < size; e += kTileThreads, ++k) {
      uint64_t off = lane_off;
      for (int i = 8; i < t; ++i)
        if (k >> (i - 8) & 1u) off |= uint64_t(1) << sp.tbits[i];
      state[base + off] = L[e];
    }
    if (sp.last) {
      __threadfence();
      __syncthreads();
      if (threadIdx.x == 0) {
        const unsigned int prev = atomicAdd(&done_tiles[sp.circuit], 1u);
        if (uint64_t(prev) + 1 == sp.ntiles) stamps[sp.circuit] = wall_clock64();
      }
    }
    __syncthreads();  // all threads are done with `sp` and `cur` before they are rewritten
  }
}

// The generator's inner product, accumulated from the values before an op
// by the thread that updates them (see TGen::fused).
struct GenAcc {
  bool on;        // the tile lies inside the generator's projector
  int zs;         // parity of the Z mask outside the tile
  uint32_t floc, zloc, ploc, pvloc;
  double sr, si;
  template <typename C>
  __device__ __forceinline__ void add(uint32_t e, C f, C lp) {  // lp = lam at e ^ floc
    if (!on || (e & ploc) != pvloc) return;
    double tr = double(lp.x) * double(f.x) + double(lp.y) * double(f.y);
    double ti = double(lp.x) * double(f.y) - double(lp.y) * double(f.x);
    if ((__popc(e & zloc) & 1) ^ zs) {
      tr = -tr;
      ti = -ti;
    }
    sr += tr;
    si += ti;
  }
};

// apply_top on both tiles in one loop, adding the generator's inner product
// over the elements each thread updates when ga.on.
template <typename C>
__device__ void apply2_top(const TOp& o, C* F, C* L, uint32_t size, uint64_t base, GenAcc& ga) {
  const int tid = threadIdx.x;
  switch (o.kind) {
    case T_DENSE1:
    case T_CDENSE1: {
      const C m0 = coef<C>(o, 0), m1 = coef<C>(o, 1), m2 = coef<C>(o, 2), m3 = coef<C>(o, 3);
      const uint32_t s = 1u << o.a;
      const bool ctl = o.kind == T_CDENSE1;
      const bool fl = ga.floc != 0;  // the generator flips the target (floc == s)
      for (uint32_t p = tid; p < size / 2; p += kTileThreads) {
        const uint32_t i0 = uint32_t(ins0(p, o.a)), i1 = i0 | s;
        if (ctl && obit(o.b, i0, base) != o.cs) continue;
        const C f0 = F[i0], f1 = F[i1], l0 = L[i0], l1 = L[i1];
        ga.add(i0, f0, fl ? l1 : l0);
        ga.add(i1, f1, fl ? l0 : l1);
        F[i0] = cmad(m1, f1, cmul(m0, f0));
        F[i1] = cmad(m3, f1, cmul(m2, f0));
        L[i0] = cmad(m1, l1, cmul(m0, l0));
        L[i1] = cmad(m3, l1, cmul(m2, l0));
      }
      break;
    }
    case T_DIAG1:
    case T_DIAG2: {
      const C d0 = coef<C>(o, 0), d1 = coef<C>(o, 1);
      const C d2 = o.kind == T_DIAG2 ? coef<C>(o, 2) : d0, d3 = o.kind == T_DIAG2 ? coef<C>(o, 3) : d1;
      for (uint32_t e = tid; e < size; e += kTileThreads) {
        const int ba = obit(o.a, e, base), bb = o.kind == T_DIAG2 ? obit(o.b, e, base) : 0;
        const C d = o.kind == T_DIAG2 ? (ba ? (bb ? d3 : d2) : (bb ? d1 : d0)) : (ba ? d1 : d0);
        const C f = F[e], l = L[e];
        ga.add(e, f, l);
        F[e] = cmul(d, f);
        L[e] = cmul(d, l);
      }
      break;
    }
    case T_DENSE2: {
      const int lo = o.a < o.b ? o.a : o.b, hi = o.a < o.b ? o.b : o.a;
      // quad member r: bit 1 of r is bit a, bit 0 is bit b; the partner of r
      // under the generator's flip is r ^ fr
      const int fr = int((ga.floc >> o.a) & 1u) << 1 | int((ga.floc >> o.b) & 1u);
      for (uint32_t p = tid; p < size / 4; p += kTileThreads) {
        const uint32_t i = uint32_t(ins0(ins0(p, lo), hi));
        const uint32_t idx[4] = {i, i | (1u << o.b), i | (1u << o.a), i | (1u << o.a) | (1u << o.b)};
        C x[4], y[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          x[r] = F[idx[r]];
          y[r] = L[idx[r]];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          C lp = y[0];
#pragma unroll
          for (int q = 1; q < 4; ++q)
            if ((r ^ fr) == q) lp = y[q];
          ga.add(idx[r], x[r], lp);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          C af = cmul(coef<C>(o, 4 * r), x[0]), al = cmul(coef<C>(o, 4 * r), y[0]);
#pragma unroll
          for (int c = 1; c < 4; ++c) {
            af = cmad(coef<C>(o, 4 * r + c), x[c], af);
            al = cmad(coef<C>(o, 4 * r + c), y[c], al);
          }
          F[idx[r]] = af;
          L[idx[r]] = al;
        }
      }
      break;
    }
    default:  // T_SWAP and the rest: never fused with a generator
      apply_top<C>(o, F, size, base);
      apply_top<C>(o, L, size, base);
      break;
  }
}

// Deterministic sum over a wavefront (fixed butterfly order); every lane
// ends with the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One pass of the adjoint reverse sweep over tiles of both states: for each
// op, the inner product <lam|G|phi> of its generator over the tile (before
// the op), then the op on both tiles.  Per-wave accumulators in LDS, summed
// per block in wave order at the end: partials[row * gridDim.x + block].
template <typename C>
__global__ __launch_bounds__(kTileThreads) void k_adjoint_tiles(C* __restrict__ phi, C* __restrict__ lam,
                                                                const TPass P, const TOp* __restrict__ ops,
                                                                const TGen* __restrict__ gens, int nrows,
                                                                double2* __restrict__ partials) {
  extern __shared__ unsigned char smem[];
  const int t = P.t;
  const uint32_t size = 1u << t;
  C* Lf = reinterpret_cast<C*>(smem);
  C* Ll = Lf + size;
  double2* acc = reinterpret_cast<double2*>(Ll + size);  // [nrows][4 waves]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < nrows * 4; i += kTileThreads) acc[i] = make_double2(0, 0);
  // element e = threadIdx.x + 256*k: offset = pdep(low 8 bits) | pdep(k << 8)
  uint64_t lane_off = 0;
  for (int i = 0; i < t && i < 8; ++i)
    if (threadIdx.x >> i & 1u) lane_off |= uint64_t(1) << P.tbits[i];
  __syncthreads();
  const int nops = P.op_end - P.op_begin;
  for (uint64_t tile = blockIdx.x; tile < P.ntiles; tile += gridDim.x) {
    uint64_t base = 0;
    {
      int sh = 0;
      for (int r = 0; r < P.nruns; ++r) {
        base |= ((tile >> sh) & ((uint64_t(1) << P.rlen[r]) - 1)) << P.rdst[r];
        sh += P.rlen[r];
      }
    }
    for (uint32_t e = threadIdx.x, k = 0; e < size; e += kTileThreads, ++k) {
      uint64_t off = lane_off;
      for (int i = 8; i < t; ++i)
        if (k >> (i - 8) & 1u) off |= uint64_t(1) << P.tbits[i];
      Lf[e] = phi[base + off];
      Ll[e] = lam[base + off];
    }
    __syncthreads();
    for (int o = 0; o < nops; ++o) {
      const TGen g = gens[o];
      if (g.row >= 0 && g.fused) {  // block-uniform: inner product inside the update
        GenAcc ga{(base & g.pout) == g.pvout, int(__popcll(base & g.zout) & 1), g.floc, g.zloc, g.ploc, g.pvloc, 0.0, 0.0};
        apply2_top<C>(ops[o], Lf, Ll, size, base, ga);
        const double sr = wave_sum(ga.sr), si = wave_sum(ga.si);
        if (lane == 0) {
          acc[g.row * 4 + wave].x += sr;
          acc[g.row * 4 + wave].y += si;
        }
        if (g.sync) __syncthreads();  // block-uniform
        continue;
      }
      if (g.row >= 0) {  // block-uniform
        double sr = 0, si = 0;
        if ((base & g.pout) == g.pvout) {
          const int zs = __popcll(base & g.zout) & 1;
          for (uint32_t e = threadIdx.x; e < size; e += kTileThreads) {
            if ((e & g.ploc) != g.pvloc) continue;
            const C f = Lf[e], l = Ll[e ^ g.floc];
            double tr = double(l.x) * double(f.x) + double(l.y) * double(f.y);
            double ti = double(l.x) * double(f.y) - double(l.y) * double(f.x);
            if ((__popc(e & g.zloc) & 1) ^ zs) {
              tr = -tr;
              ti = -ti;
            }
            sr += tr;
            si += ti;
          }
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (lane == 0) {
          acc[g.row * 4 + wave].x += sr;
          acc[g.row * 4 + wave].y += si;
        }
        __syncthreads();  // the op below rewrites what other threads just read
      }
      GenAcc none{false, 0, 0u, 0u, 0u, 0u, 0.0, 0.0};
      apply2_top<C>(ops[o], Lf, Ll, size, base, none);
      if (g.sync) __syncthreads();  // block-uniform
    }
    for (uint32_t e = threadIdx.x, k = 0; e < size; e += kTileThreads, ++k) {
      uint64_t off = lane_off;
      for (int i = 8; i < t; ++i)
        if (k >> (i - 8) & 1u) off |= uint64_t(1) << P.tbits[i];
      phi[base + off] = Lf[e];
      lam[base + off] = Ll[e];
    }
    __syncthreads();
  }
  for (int r = threadIdx.x; r < nrows; r += kTileThreads) {
    double2 s = acc[r * 4];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      s.x += acc[r * 4 + w].x;
      s.y += acc[r * 4 + w].y;
    }
    partials[size_t(r) * gridDim.x + blockIdx.x] = s;
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

template <typename R>
int launch_adjoint_tiles(void* phi, void* lam, const TPass& pass, const TOp* ops, const TGen* gens,
                         int nrows, void* partials, int blocks, hipStream_t st) {
  using C = typename CT<R>::type;
  if (nrows < 0 || nrows > kAdjRowsMax) return fail(QSV_EINVAL, "adjoint tiles: too many rows in one pass");
  if (pass.t < 1 || pass.t > kTileMaxBits) return fail(QSV_EINVAL, "adjoint tiles: bad tile size");
  const size_t lds = 2 * (size_t(1) << pass.t) * sizeof(C) + size_t(kAdjRowsMax) * 4 * sizeof(double2);
  QSV_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_adjoint_tiles<C>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL((k_adjoint_tiles<C>), dim3(blocks), dim3(kTileThreads), lds, st, static_cast<C*>(phi),
                     static_cast<C*>(lam), pass, ops, gens, nrows, static_cast<double2*>(partials));
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template int launch_adjoint_tiles<double>(void*, void*, const TPass&, const TOp*, const TGen*, int, void*, int,
                                          hipStream_t);
template int launch_adjoint_tiles<float>(void*, void*, const TPass&, const TOp*, const TGen*, int, void*, int,
                                         hipStream_t);
template int launch_tile_level<double>(const TileLevelArgs&, hipStream_t);
template int launch_tile_level<float>(const TileLevelArgs&, hipStream_t);

}  // namespace qsv
