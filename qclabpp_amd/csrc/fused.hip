// Kernels of fused blocks.
//
// Dense block (k = 3..5 qubits): out_g = U in_g for every group g of 2^k
// amplitudes that differ only in the block's bits.  As a real GEMM with
// interleaved complex values: U_real (2^(k+1) x 2^(k+1)) times the groups'
// (re, im) columns, on v_mfma_f64_16x16x4_f64 (complex<double>) or
// v_mfma_f32_16x16x4_f32 (complex<float>).  A wavefront owns 16 groups per
// step (the MFMA's N = 16): it stages their 16 * 2^k amplitudes through LDS
// with coalesced loads (the tile's bits are the block's bits plus the four
// lowest other bits, so every load is a run of >= 16 consecutive
// amplitudes), reads the B operand from LDS, keeps U_real in registers for
// the whole kernel, accumulates 2^(k+1)/16 output tiles, writes them back to
// LDS and stores the tile with the same coalesced pattern.
//
// MFMA operand layouts (16x16x4, 64 lanes): A[i][k] in lane i + 16k,
// B[k][j] in lane j + 16k; D[i][j] in lane j + 16*(i/4), element i%4 for
// FP32, and in lane j + 16*(i%4), element i/4 for FP64 (MF<R>::row).
//
// Diagonal block (k <= 10 qubits): each amplitude is multiplied by the table
// entry its block bits select; the table sits in LDS.
#include "fused.hpp"

namespace qsv {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename R> struct MF;
// row(lane, r): the output row (0..15) that element r of `lane` holds; the
// column is lane % 16 for both.  Measured on gfx950: the FP64 variant
// interleaves rows (lane/16 + 4r), the FP32 variant groups them (4*(lane/16) + r).
template <> struct MF<double> {
  using acc = v4d;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct MF<float> {
  using acc = v4f;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

__device__ __forceinline__ uint32_t pext32(uint32_t x, uint32_t mask) {
  uint32_t r = 0;
  int k = 0;
  for (int b = 0; b < 32 && mask; ++b) {
    const uint32_t low = mask & (0u - mask);
    if (x & low) r |= 1u << k;
    ++k;
    mask ^= low;
  }
  return r;
}

__device__ __forceinline__ uint64_t tile_offset(uint32_t t, const TileMap& tm) {
  uint64_t off = 0;
#pragma unroll 1
  for (int i = 0; i < tm.ntb; ++i)
    if (t >> i & 1u) off |= uint64_t(1) << tm.tbits[i];
  return off;
}

__device__ __forceinline__ uint64_t deposit_runs(uint64_t w, const TileMap& tm) {
  uint64_t a = 0;
  int sh = 0;
#pragma unroll 1
  for (int r = 0; r < tm.nruns; ++r) {
    const uint64_t len = tm.rlen[r];
    a |= ((w >> sh) & ((uint64_t(1) << len) - 1)) << tm.rdst[r];
    sh += int(len);
  }
  return a;
}

// Orders one wavefront's LDS accesses: every lane's earlier LDS writes are
// visible to every lane of the same wave afterwards.  Each wave of the block
// kernel owns its own tile region, so waves never wait for each other and
// their load, MFMA and store phases interleave freely.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// NB MFMA column tiles (16 groups each) per wave tile.
template <typename R, int K, int NB>
__global__ __launch_bounds__(256) void k_fused_mfma(typename CT<R>::type* __restrict__ psi,
                                                    const R* __restrict__ ulane, TileMap tm,
                                                    uint64_t ntiles) {
  using C = typename CT<R>::type;
  using Acc = typename MF<R>::acc;
  constexpr int A = 1 << K;     // amplitudes per group
  constexpr int GW = 16 * NB;   // groups per wave tile
  constexpr int TE = A * GW;    // amplitudes per wave tile
  constexpr int M = 2 * A;      // real rows of U_real
  constexpr int MT = M / 16;    // output row tiles
  constexpr int KS = M / 4;     // K steps
  constexpr int EPL = TE / 64;  // tile amplitudes per lane
  // LDS tile of one wave: amplitude i of group g sits in slot GW*i + (g ^ (i & (GW-1))).
  // Whichever address bits the block and group bits are, the 16 lanes of an
  // MFMA operand row (16 groups, one i) and the 16 consecutive tile positions
  // of a global access map to 16 different bank groups.
  __shared__ C tile[4][TE];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  R areg[MT][KS];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int s = 0; s < KS; ++s) areg[mt][s] = ulane[(mt * KS + s) * 64 + lane];
  const int g = lane & 15;
  uint64_t eoff[EPL];
  int slot[EPL];
  const uint64_t loff = tile_offset(uint32_t(lane), tm);
#pragma unroll
  for (int e = 0; e < EPL; ++e) {
    const uint32_t t = uint32_t(lane + 64 * e);  // natural tile position
    eoff[e] = loff | tile_offset(uint32_t(e) << 6, tm);
    const int i = int(pext32(t, tm.local_block_mask)), gg = int(pext32(t, tm.local_group_mask));
    slot[e] = GW * i + (gg ^ (i & (GW - 1)));
  }

  for (uint64_t w0 = uint64_t(blockIdx.x) * 4; w0 < ntiles; w0 += uint64_t(gridDim.x) * 4) {
    const uint64_t w = w0 + wave;
    const bool act = w < ntiles;
    const uint64_t base = act ? deposit_runs(w, tm) : 0;
    if (act) {
      C v[EPL];
#pragma unroll
      for (int e = 0; e < EPL; ++e) v[e] = ld_stream(&psi[base + eoff[e]]);
#pragma unroll
      for (int e = 0; e < EPL; ++e) tile[wave][slot[e]] = v[e];
    }
    wave_sync();
    Acc acc[NB][MT];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int gc = 16 * c + g;
      R b[KS];
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int i = 2 * s + (lane >> 5);
        const C v = tile[wave][GW * i + (gc ^ (i & (GW - 1)))];
        b[s] = (lane & 16) ? v.y : v.x;
      }
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) acc[c][mt] = Acc{0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[c][mt] = MF<R>::mma(areg[mt][s], b[s], acc[c][mt]);
    }
    wave_sync();
    R* tre = reinterpret_cast<R*>(&tile[wave][0]);
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * mt + MF<R>::row(lane, r);  // real row: 2*amp + (re|im)
          const int i = row >> 1, gc = 16 * c + g;
          tre[2 * (GW * i + (gc ^ (i & (GW - 1)))) + (row & 1)] = acc[c][mt][r];
        }
    wave_sync();
    if (act) {
#pragma unroll
      for (int e = 0; e < EPL; ++e) st_stream(&psi[base + eoff[e]], tile[wave][slot[e]]);
    }
  }
}

// Small states (fewer than k+4 qubits): one thread per group, U row-major.
template <typename R>
__global__ __launch_bounds__(64) void k_densek_simple(typename CT<R>::type* __restrict__ psi,
                                                      const typename CT<R>::type* __restrict__ U,
                                                      int k, uint64_t groups, uint32_t bits_lo,
                                                      BlockBits bb) {
  using C = typename CT<R>::type;
  const uint64_t g = uint64_t(blockIdx.x) * 64 + threadIdx.x;
  if (g >= groups) return;
  // insert zeros at the block's bit positions (ascending) to get the base
  uint64_t base = g;
  for (int p = 0; p < k; ++p) {
    const uint64_t lo = base & ((uint64_t(1) << bb.b[p]) - 1);
    base = ((base ^ lo) << 1) | lo;
  }
  const int D = 1 << k;
  C x[32];
  for (int r = 0; r < D; ++r) {
    uint64_t a = base;
    for (int p = 0; p < k; ++p)
      if (r >> p & 1) a |= uint64_t(1) << bb.b[p];
    x[r] = ld_stream(&psi[a]);
  }
  for (int r = 0; r < D; ++r) {
    C acc = mk<C>(0, 0);
    for (int c = 0; c < D; ++c) acc = cmad(U[r * D + c], x[c], acc);
    uint64_t a = base;
    for (int p = 0; p < k; ++p)
      if (r >> p & 1) a |= uint64_t(1) << bb.b[p];
    st_stream(&psi[a], acc);
  }
  (void)bits_lo;
}

template <typename R, int NP>
__global__ __launch_bounds__(256) void k_diagk(typename CT<R>::type* __restrict__ psi, uint64_t nel,
                                               const typename CT<R>::type* __restrict__ table,
                                               int k, BlockBits bb, uint64_t blk0) {
  using C = typename CT<R>::type;
  __shared__ C tab[1 << kMaxDiag];
  const int D = 1 << k;
  for (int i = threadIdx.x; i < D; i += 256) tab[i] = table[i];
  __syncthreads();
  const uint64_t e0 = (blk0 + blockIdx.x) * (256 * NP) + threadIdx.x;
  C x[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * 256;
    if (e < nel) x[j] = ld_stream(&psi[e]);
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * 256;
    if (e < nel) {
      int idx = 0;
      for (int p = 0; p < k; ++p) idx |= int((e >> bb.b[p]) & 1) << p;
      st_stream(&psi[e], cmul(tab[idx], x[j]));
    }
  }
}

// Column tiles per wave tile: two for k <= 4 (twice the bytes per wave
// iteration), one for k = 5 (its A operand already fills the registers).
template <int K> constexpr int kColTiles = K <= 4 ? 2 : 1;

template <int K, typename R>
void launch_mfma(void* psi, const void* ulane, const TileMap& tm, uint64_t ntiles, hipStream_t st) {
  using C = typename CT<R>::type;
  const uint64_t want = (ntiles + 3) / 4;
  const unsigned grid = unsigned(want < 4096 ? want : 4096);
  hipLaunchKernelGGL((k_fused_mfma<R, K, kColTiles<K>>), dim3(grid), dim3(256), 0, st,
                     static_cast<C*>(psi), static_cast<const R*>(ulane), tm, ntiles);
}

}  // namespace

TileMap make_tile_map(const std::vector<int>& qubits, int n, int group_bits) {
  TileMap tm{};
  const int k = int(qubits.size());
  uint64_t bmask = 0;
  for (int q : qubits) bmask |= uint64_t(1) << (n - 1 - q);
  // group bits: the lowest bits outside the block
  uint64_t gmask = 0;
  int got = 0;
  for (int b = 0; b < n && got < group_bits; ++b)
    if (!(bmask >> b & 1)) {
      gmask |= uint64_t(1) << b;
      ++got;
    }
  const uint64_t tmask = bmask | gmask;
  tm.ntb = 0;
  for (int b = 0; b < n; ++b)
    if (tmask >> b & 1) {
      if (bmask >> b & 1) tm.local_block_mask |= 1u << tm.ntb;
      else tm.local_group_mask |= 1u << tm.ntb;
      tm.tbits[tm.ntb++] = uint8_t(b);
    }
  // runs of bits outside the tile, ascending
  tm.nruns = 0;
  for (int b = 0; b < n;) {
    if (tmask >> b & 1) {
      ++b;
      continue;
    }
    int e = b;
    while (e < n && !(tmask >> e & 1)) ++e;
    tm.rdst[tm.nruns] = uint8_t(b);
    tm.rlen[tm.nruns] = uint8_t(e - b);
    ++tm.nruns;
    b = e;
  }
  (void)k;
  return tm;
}

template <typename R>
int launch_fused(const KOp& op, const DeviceBlock& db, void* psi, int n, hipStream_t st) {
  using C = typename CT<R>::type;
  const int k = op.nq;
  BlockBits bb{};
  for (int p = 0; p < k; ++p) bb.b[p] = uint8_t(n - 1 - op.q[k - 1 - p]);
  const uint64_t N = uint64_t(1) << n;
  if (op.kind == K_DIAGK) {
    constexpr int NP = 4;
    const uint64_t blocks = (N + 256 * NP - 1) / (256 * NP);
    QSV_LAUNCH_BLOCKS((k_diagk<R, NP>), blocks, 256, 0, st, static_cast<C*>(psi), N,
                      static_cast<const C*>(db.table), k, bb);
  } else if (op.kind == K_DENSEK) {
    if (k >= 3 && n >= k + 6) {
      const int gbits = 4 + (k <= 4 ? 1 : 0);  // log2(16 * column tiles)
      const TileMap tm = make_tile_map(std::vector<int>(op.q, op.q + k), n, gbits);
      const uint64_t ntiles = N >> (k + gbits);
      switch (k) {
        case 3: launch_mfma<3, R>(psi, db.ulane, tm, ntiles, st); break;
        case 4: launch_mfma<4, R>(psi, db.ulane, tm, ntiles, st); break;
        default: launch_mfma<5, R>(psi, db.ulane, tm, ntiles, st); break;
      }
    } else {
      const uint64_t groups = N >> k;
      hipLaunchKernelGGL((k_densek_simple<R>), dim3(unsigned((groups + 63) / 64)), dim3(64), 0, st,
                         static_cast<C*>(psi), static_cast<const C*>(db.matrix), k, groups, 0u,
                         bb);
    }
  } else {
    return fail(QSV_EINVAL, "launch_fused: not a fused op");
  }
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template int launch_fused<double>(const KOp&, const DeviceBlock&, void*, int, hipStream_t);
template int launch_fused<float>(const KOp&, const DeviceBlock&, void*, int, hipStream_t);

}  // namespace qsv
