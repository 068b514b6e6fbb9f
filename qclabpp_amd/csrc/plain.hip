// Plain per-gate kernels: one HBM sweep per gate, one kernel per sweep class
// of src/qgates/apply.hpp:258-371.  Index math follows the reference's masks
// (apply.hpp:226-256): a compact loop index k gets zero bits inserted at the
// gate's bit positions; bit(q) = n-1-q (qubit 0 is the most significant).
//
// Layout: the state is one contiguous array of 2^n complex values
// (double2 for complex<double>, float2 for complex<float>) in HBM.
// Work split: TPB threads per block, NP items per thread, items of one
// thread TPB apart so every load instruction of a wavefront reads
// consecutive amplitudes (a contiguous 1 KiB run for complex<double>).
// Gates whose pair stride is below the wavefront width use the "lanes"
// variant: each lane loads one amplitude and exchanges it with its partner
// lane through a cross-lane shuffle, so loads stay fully contiguous.
#include "plain.hpp"

namespace qsv {

namespace {

__device__ __forceinline__ uint64_t ins0(uint64_t k, int b) {
  const uint64_t lo = k & ((uint64_t(1) << b) - 1);
  return ((k ^ lo) << 1) | lo;
}

template <typename C>
__device__ __forceinline__ C shfl_xor_c(C v, int mask) {
  C r;
  r.x = __shfl_xor(v.x, mask);
  r.y = __shfl_xor(v.y, mask);
  return r;
}

template <typename C>
struct M2 {
  C a00, a01, a10, a11;
};
template <typename C>
struct M4 {
  C a[16];
};

// ---- apply2: 2x2 on pairs (a, a|s) -----------------------------------------
// Pair sweeps with the partner 2^b >= 64 amplitudes away: lanes 32-63 of a
// wave hold the amplitudes 2^b above those of lanes 0-31, so each load
// instruction reads 32 |0>-side and 32 |1>-side amplitudes, and
// v_permlane32_swap gives every lane both amplitudes of its pair.  NP pairs
// per lane pair: 1 or 2, whichever measured faster for the target bit on the
// MI355X (2^30 complex128 amplitudes, every b from 6 to 29, two runs that
// agreed within 1%); the choice follows the byte-address bit of the partner
// offset, so complex64 uses the same table one amplitude bit higher.  Over
// all b: 6.19 TB/s on average with the table against 5.98 for two loads per
// lane, one per side.
__device__ __forceinline__ void swap_halves(uint32_t v, uint32_t& lo, uint32_t& hi) {
  // lo: every lane gets v of lane & 31; hi: v of lane | 32
  const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  lo = r[0];
  hi = r[1];
}
template <typename C>
__device__ __forceinline__ void pair_across_halves(const C v, C& x, C& y) {
  using R = decltype(C::x);
  // a copy of each component: __builtin_bit_cast of a vector element
  // expression reads element 0
  const R re = v.x, im = v.y;
  if constexpr (sizeof(R) == 8) {
    const uint64_t br = __builtin_bit_cast(uint64_t, re), bi = __builtin_bit_cast(uint64_t, im);
    uint32_t l0, h0, l1, h1, l2, h2, l3, h3;
    swap_halves(uint32_t(br), l0, h0);
    swap_halves(uint32_t(br >> 32), l1, h1);
    swap_halves(uint32_t(bi), l2, h2);
    swap_halves(uint32_t(bi >> 32), l3, h3);
    x = mk<C>(__builtin_bit_cast(double, (uint64_t(l1) << 32) | l0), __builtin_bit_cast(double, (uint64_t(l3) << 32) | l2));
    y = mk<C>(__builtin_bit_cast(double, (uint64_t(h1) << 32) | h0), __builtin_bit_cast(double, (uint64_t(h3) << 32) | h2));
  } else {
    uint32_t l0, h0, l1, h1;
    swap_halves(__builtin_bit_cast(uint32_t, re), l0, h0);
    swap_halves(__builtin_bit_cast(uint32_t, im), l1, h1);
    x = mk<C>(__builtin_bit_cast(float, l0), __builtin_bit_cast(float, l1));
    y = mk<C>(__builtin_bit_cast(float, h0), __builtin_bit_cast(float, h1));
  }
}

inline int pair_np(int b, size_t amp_bytes) {
  // byte-address bits of the partner offset where one pair per lane pair won
  constexpr uint64_t one = (1ull << 10) | (1ull << 11) | (1ull << 13) | (1ull << 21) | (1ull << 28) |
                           (1ull << 29) | (1ull << 32) | (1ull << 33);
  const int byte_bit = b + (amp_bytes == 16 ? 4 : 3);
  return byte_bit < 64 && (one >> byte_bit & 1) ? 1 : 2;
}

template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_dense1(C* __restrict__ psi, uint64_t cnt, int b,
                                                M2<C> m, uint64_t blk0) {
  const int lane = threadIdx.x & 63, i = lane & 31;
  const bool upper = lane >= 32;
  const uint64_t wave = (blk0 + blockIdx.x) * (TPB / 64) + (threadIdx.x >> 6);
  const uint64_t kb = wave * (32 * NP) + i;
  const uint64_t side = upper ? uint64_t(1) << b : 0;
  C v[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = kb + uint64_t(j) * 32;
    v[j] = mk<C>(0, 0);
    if (k < cnt) v[j] = ld_stream(&psi[ins0(k, b) | side]);
  }
  // the lane's row of the 2x2: [a00 a01] for the |0> side, [a10 a11] for |1>
  const C r0 = upper ? m.a10 : m.a00, r1 = upper ? m.a11 : m.a01;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = kb + uint64_t(j) * 32;
    C x, y;
    pair_across_halves(v[j], x, y);
    if (k < cnt) st_stream(&psi[ins0(k, b) | side], cmad(r1, y, cmul(r0, x)));
  }
}

// Same operation for b < 6: partner amplitude lives in lane (lane ^ 2^b).
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_dense1_lanes(C* __restrict__ psi, uint64_t nel, int b,
                                                      M2<C> m, uint64_t blk0) {
  const int s = 1 << b;
  const uint64_t e0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  C x[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * TPB;
    x[j] = e < nel ? ld_stream(&psi[e]) : mk<C>(0, 0);
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * TPB;
    const C y = shfl_xor_c(x[j], s);
    const bool hi = (e >> b) & 1;
    const C cs = hi ? m.a11 : m.a00;  // coefficient on own amplitude
    const C co = hi ? m.a10 : m.a01;  // coefficient on partner amplitude
    if (e < nel) st_stream(&psi[e], cmad(co, y, cmul(cs, x[j])));
  }
}

// ---- RotationZ-class diagonal on every amplitude ---------------------------
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_diag1(C* __restrict__ psi, uint64_t nel, int b, C d0,
                                               C d1, uint64_t blk0) {
  const uint64_t e0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  C x[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * TPB;
    if (e < nel) x[j] = ld_stream(&psi[e]);
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * TPB;
    if (e < nel) st_stream(&psi[e], cmul(((e >> b) & 1) ? d1 : d0, x[j]));
  }
}

// ---- apply2b: scale the |1> half ------------------------------------------------
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_phase1(C* __restrict__ psi, uint64_t cnt, int b, C lam, uint64_t blk0) {
  const uint64_t s = uint64_t(1) << b;
  const uint64_t k0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  C x[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    if (k < cnt) x[j] = ld_stream(&psi[ins0(k, b) | s]);
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    if (k < cnt) st_stream(&psi[ins0(k, b) | s], cmul(lam, x[j]));
  }
}

// ---- apply4 (controlled): 2x2 on the control==cs half ------------------------
// The pair layout of k_dense1 over the control half (0.74 of HBM peak against
// 0.69 with two loads per lane, over the neighbour-qubit CX gates of an n=30
// complex128 brickwork, the control half read and written in place).
// blo < bhi are the two bit positions; a1 carries the control bit, s the
// target.  Each wave takes 32*NP pairs; lanes whose pair is out of range
// still take part in the swaps (both lanes of a pair are in or out together).
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_cdense1(C* __restrict__ psi, uint64_t cnt, int blo,
                                                 int bhi, uint64_t a1, uint64_t s, M2<C> m, uint64_t blk0) {
  const int lane = threadIdx.x & 63, i = lane & 31;
  const bool upper = lane >= 32;
  const uint64_t wave = (blk0 + blockIdx.x) * (TPB / 64) + (threadIdx.x >> 6);
  const uint64_t kb = wave * (32 * NP) + i;
  const uint64_t side = upper ? s : 0;
  C v[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = kb + uint64_t(j) * 32;
    v[j] = mk<C>(0, 0);
    if (k < cnt) v[j] = ld_stream(&psi[(ins0(ins0(k, blo), bhi) | a1) | side]);
  }
  // the lane's row of the 2x2: [a00 a01] for the |0> side, [a10 a11] for |1>
  const C r0 = upper ? m.a10 : m.a00, r1 = upper ? m.a11 : m.a01;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = kb + uint64_t(j) * 32;
    C x, y;
    pair_across_halves(v[j], x, y);
    if (k < cnt) st_stream(&psi[(ins0(ins0(k, blo), bhi) | a1) | side], cmad(r1, y, cmul(r0, x)));
  }
}

// Controlled 2x2 with target bit < 6: shuffle across lanes; store where the
// control bit matches, or every element (unchanged where it does not) when
// `all` is set -- a control bit inside a 128-byte line, where storing only
// the matching amplitudes would write every line in part.
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_cdense1_lanes(C* __restrict__ psi, uint64_t nel, int bt,
                                                       int bc, int cs, int all, M2<C> m,
                                                       uint64_t blk0) {
  const int s = 1 << bt;
  const uint64_t e0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  C x[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * TPB;
    x[j] = e < nel ? ld_stream(&psi[e]) : mk<C>(0, 0);
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * TPB;
    const C y = shfl_xor_c(x[j], s);
    const bool hi = (e >> bt) & 1;
    const C cself = hi ? m.a11 : m.a00;
    const C cother = hi ? m.a10 : m.a01;
    const bool on = int((e >> bc) & 1) == cs;
    if (e < nel && (on || all)) st_stream(&psi[e], on ? cmad(cother, y, cmul(cself, x[j])) : x[j]);
  }
}

// ---- apply4b: scale one quarter ------------------------------------------------
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_cphase(C* __restrict__ psi, uint64_t cnt, int blo,
                                                int bhi, uint64_t b1, C lam, uint64_t blk0) {
  const uint64_t k0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  C x[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    if (k < cnt) x[j] = ld_stream(&psi[ins0(ins0(k, blo), bhi) | b1]);
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    if (k < cnt) st_stream(&psi[ins0(ins0(k, blo), bhi) | b1], cmul(lam, x[j]));
  }
}

// ---- apply4bc: exchange |01>,|10> and scale both by lam ---------------------
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_swap(C* __restrict__ psi, uint64_t cnt, int blo, int bhi,
                                              C lam, int scale, uint64_t blk0) {
  const uint64_t k0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  const uint64_t sl = uint64_t(1) << blo, sh = uint64_t(1) << bhi;
  C x[NP], y[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    if (k < cnt) {
      const uint64_t i = ins0(ins0(k, blo), bhi);
      x[j] = ld_stream(&psi[i | sh]);
      y[j] = ld_stream(&psi[i | sl]);
    }
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    if (k < cnt) {
      const uint64_t i = ins0(ins0(k, blo), bhi);
      st_stream(&psi[i | sh], scale ? cmul(lam, y[j]) : y[j]);
      st_stream(&psi[i | sl], scale ? cmul(lam, x[j]) : x[j]);
    }
  }
}

// ---- SWAP / iSWAP whose lower bit is below the lane count ------------------
// Each thread owns the pair (e, e|sh) for one e with bit bhi clear, loaded and
// stored whole, so every 128-byte line is read and written in full; the
// |01>,|10> exchange crosses to lane ^ 2^blo.  Thread k < cnt = N/2.
template <typename C>
__global__ __launch_bounds__(TPB) void k_swap_lanes(C* __restrict__ psi, uint64_t cnt, int blo,
                                                    int bhi, C lam, int scale, uint64_t blk0) {
  const uint64_t k = (blk0 + blockIdx.x) * TPB + threadIdx.x;
  const uint64_t sh = uint64_t(1) << bhi;
  const uint64_t e = ins0(k < cnt ? k : 0, bhi);
  C a = ld_stream(&psi[e]), b = ld_stream(&psi[e | sh]);
  const bool lo = (e >> blo) & 1;
  // lo = 0 holds (|00>, |10>) and hands |10> over; lo = 1 holds (|01>, |11>) and hands |01>
  const C v = shfl_xor_c(lo ? a : b, 1 << blo);
  const C w = scale ? cmul(lam, v) : v;
  if (lo) a = w;
  else b = w;
  if (k < cnt) {
    st_stream(&psi[e], a);
    st_stream(&psi[e | sh], b);
  }
}

// ---- controlled 2x2, target bit below the lane count, control bit above it -----
// Sweeps only the half with control bit == cs (thread k < N/2 owns element
// ins0(k, bc) | cs<<bc); the target partner is lane ^ 2^bt.
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_cdense1_lanes_half(C* __restrict__ psi, uint64_t cnt,
                                                            int bt, int bc, int cs, M2<C> m,
                                                            uint64_t blk0) {
  const uint64_t k0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  C x[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    x[j] = ld_stream(&psi[ins0(k < cnt ? k : 0, bc) | (uint64_t(cs) << bc)]);
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    const uint64_t e = ins0(k < cnt ? k : 0, bc) | (uint64_t(cs) << bc);
    const C y = shfl_xor_c(x[j], 1 << bt);
    const bool hi = (e >> bt) & 1;
    const C cself = hi ? m.a11 : m.a00;
    const C cother = hi ? m.a10 : m.a01;
    if (k < cnt) st_stream(&psi[e], cmad(cother, y, cmul(cself, x[j])));
  }
}

// ---- phase on a quarter whose lower bit lies inside a 128-byte line ----------
// Sweeps the half with bit bh == vh in whole lines and scales the amplitudes
// whose bit bl == vl (thread k < N/2 owns element ins0(k, bh) | vh<<bh).
template <typename C>
__global__ __launch_bounds__(TPB) void k_cphase_half(C* __restrict__ psi, uint64_t cnt, int bh,
                                                     int vh, int bl, int vl, C lam, uint64_t blk0) {
  const uint64_t k = (blk0 + blockIdx.x) * TPB + threadIdx.x;
  if (k >= cnt) return;
  const uint64_t e = ins0(k, bh) | (uint64_t(vh) << bh);
  const C x = ld_stream(&psi[e]);
  st_stream(&psi[e], int((e >> bl) & 1) == vl ? cmul(lam, x) : x);
}

// ---- controlled 2x2 whose control bit is inside a 128-byte line --------------
// Sweeps every target pair (like k_dense1) and applies the matrix where the
// control bit equals cs, identity elsewhere, so whole lines are written.
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_cdense1_full(C* __restrict__ psi, uint64_t cnt, int bt,
                                                      int bc, int cs, M2<C> m, uint64_t blk0) {
  const uint64_t s = uint64_t(1) << bt;
  const uint64_t k0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  C x[NP], y[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    if (k < cnt) {
      const uint64_t a = ins0(k, bt);
      x[j] = ld_stream(&psi[a]);
      y[j] = ld_stream(&psi[a | s]);
    }
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    if (k < cnt) {
      const uint64_t a = ins0(k, bt);
      const bool on = int((a >> bc) & 1) == cs;
      st_stream(&psi[a], on ? cmad(m.a01, y[j], cmul(m.a00, x[j])) : x[j]);
      st_stream(&psi[a | s], on ? cmad(m.a11, y[j], cmul(m.a10, x[j])) : y[j]);
    }
  }
}

// ---- apply4 (dense): 4x4 on quads ----------------------------------------------
// b0 = bit of the lower-numbered qubit (more significant), b1 of the other;
// quad position p = 2*bit(b0) + bit(b1).
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_dense2(C* __restrict__ psi, uint64_t cnt, int b0, int b1,
                                                M4<C> m, uint64_t blk0) {
  const uint64_t k0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  const uint64_t s0 = uint64_t(1) << b0, s1 = uint64_t(1) << b1;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t k = k0 + uint64_t(j) * TPB;
    if (k < cnt) {
      const uint64_t i = ins0(ins0(k, b1), b0);
      const uint64_t ix[4] = {i, i | s1, i | s0, i | s0 | s1};
      C x[4], y[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) x[p] = ld_stream(&psi[ix[p]]);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        C acc = cmul(m.a[4 * r], x[0]);
#pragma unroll
        for (int c = 1; c < 4; ++c) acc = cmad(m.a[4 * r + c], x[c], acc);
        y[r] = acc;
      }
#pragma unroll
      for (int p = 0; p < 4; ++p) st_stream(&psi[ix[p]], y[p]);
    }
  }
}

// ---- diagonal 4x4 on every amplitude ------------------------------------------
template <typename C, int NP>
__global__ __launch_bounds__(TPB) void k_diag2(C* __restrict__ psi, uint64_t nel, int b0, int b1,
                                               C d0, C d1, C d2, C d3, uint64_t blk0) {
  const uint64_t e0 = (blk0 + blockIdx.x) * (TPB * NP) + threadIdx.x;
  C x[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * TPB;
    if (e < nel) x[j] = ld_stream(&psi[e]);
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint64_t e = e0 + uint64_t(j) * TPB;
    if (e < nel) {
      const int p = int(((e >> b0) & 1) << 1) | int((e >> b1) & 1);
      const C d = p == 0 ? d0 : p == 1 ? d1 : p == 2 ? d2 : d3;
      st_stream(&psi[e], cmul(d, x[j]));
    }
  }
}

// ---- psi := |idx> --------------------------------------------------------------
template <typename C>
__global__ __launch_bounds__(TPB) void k_set_basis(C* __restrict__ psi, uint64_t nel, uint64_t idx, uint64_t blk0) {
  const uint64_t e = (blk0 + blockIdx.x) * TPB + threadIdx.x;
  if (e < nel) st_stream(&psi[e], mk<C>(e == idx ? 1 : 0, 0));
}

template <typename C>
C to_c(const cd& z) {
  return mk<C>(static_cast<decltype(C::x)>(z.real()), static_cast<decltype(C::x)>(z.imag()));
}

inline uint64_t blocks_for(uint64_t items, int per_block) {
  return (items + per_block - 1) / per_block;
}

}  // namespace

template <typename R>
int launch_plain(const KOp& op, void* p, int n, hipStream_t st, const PlainTuning& tu) {
  using C = typename CT<R>::type;
  C* psi = static_cast<C*>(p);
  const uint64_t N = uint64_t(1) << n;
  // items per thread, measured on gfx950 at n=30: one for the full and half
  // sweeps (more waves in flight), two for the controlled quarter-pair sweep
  constexpr int NP = 1;
  const int per = TPB * NP;
  // address bits inside one 128-byte line: a sweep that touches only some
  // amplitudes of a line is replaced by one that reads and writes whole lines
  const int line_bits = sizeof(C) == 16 ? 3 : 4;
  auto bit = [n](int q) { return n - 1 - q; };
  switch (op.kind) {
    case K_NOP:
      return QSV_OK;
    case K_DENSE1: {
      const M2<C> m{to_c<C>(op.m[0]), to_c<C>(op.m[1]), to_c<C>(op.m[2]), to_c<C>(op.m[3])};
      const int b = bit(op.q[0]);
      if (b < tu.lane_bits && n >= 6) {
        QSV_LAUNCH_BLOCKS((k_dense1_lanes<C, NP>), blocks_for(N, per), TPB, 0, st,
                           psi, N, b, m);
      } else {
        if (pair_np(b, sizeof(C)) == 2)
          QSV_LAUNCH_BLOCKS((k_dense1<C, 2>), blocks_for(N / 2, TPB), TPB, 0, st, psi, N / 2, b, m);
        else
          QSV_LAUNCH_BLOCKS((k_dense1<C, 1>), blocks_for(N / 2, TPB / 2), TPB, 0, st, psi, N / 2, b, m);
      }
      break;
    }
    case K_DIAG1:
      QSV_LAUNCH_BLOCKS((k_diag1<C, NP>), blocks_for(N, per), TPB, 0, st, psi, N,
                         bit(op.q[0]), to_c<C>(op.m[0]), to_c<C>(op.m[1]));
      break;
    case K_PHASE1:
      if (bit(op.q[0]) < line_bits) {  // the |1> half interleaves within lines: sweep all
        QSV_LAUNCH_BLOCKS((k_diag1<C, NP>), blocks_for(N, per), TPB, 0, st, psi, N, bit(op.q[0]),
                          mk<C>(1, 0), to_c<C>(op.m[0]));
      } else {
        QSV_LAUNCH_BLOCKS((k_phase1<C, NP>), blocks_for(N / 2, per), TPB, 0, st, psi,
                          N / 2, bit(op.q[0]), to_c<C>(op.m[0]));
      }
      break;
    case K_CDENSE1: {
      const M2<C> m{to_c<C>(op.m[0]), to_c<C>(op.m[1]), to_c<C>(op.m[2]), to_c<C>(op.m[3])};
      const int bc = bit(op.q[0]), bt = bit(op.q[1]);
      // The half sweep moves whole lines when the control bit lies outside
      // them (n=30 complex128, brickwork CX: 2.76 ms, 0.78 of HBM peak for
      // its S bytes).  With the control bit inside a line it would move
      // 32-byte pieces of every line (6.97 ms); sweeping and storing all of
      // the state instead takes 5.07 ms for the 2S counted for such a gate.
      if (bt < tu.lane_bits && bc > bt && bc >= line_bits && n >= 7) {
        QSV_LAUNCH_BLOCKS((k_cdense1_lanes_half<C, 1>), blocks_for(N / 2, TPB), TPB, 0, st, psi,
                          N / 2, bt, bc, op.cs, m);
      } else if (bt < tu.lane_bits && n >= 6) {
        QSV_LAUNCH_BLOCKS((k_cdense1_lanes<C, NP>), blocks_for(N, per), TPB, 0, st,
                          psi, N, bt, bc, op.cs, int(bc < line_bits), m);
      } else if (bc < line_bits) {
        QSV_LAUNCH_BLOCKS((k_cdense1_full<C, NP>), blocks_for(N / 2, per), TPB, 0, st, psi, N / 2,
                          bt, bc, op.cs, m);
      } else {
        // NP pairs per lane pair: 32*NP pairs per wave, TPB/64 waves per block
        const uint64_t a1 = uint64_t(op.cs) << bc;
        const int blo = std::min(bc, bt), bhi = std::max(bc, bt);
        if (pair_np(bt, sizeof(C)) == 2)
          QSV_LAUNCH_BLOCKS((k_cdense1<C, 2>), blocks_for(N / 4, TPB), TPB, 0, st, psi, N / 4, blo,
                            bhi, a1, uint64_t(1) << bt, m);
        else
          QSV_LAUNCH_BLOCKS((k_cdense1<C, 1>), blocks_for(N / 4, TPB / 2), TPB, 0, st, psi, N / 4,
                            blo, bhi, a1, uint64_t(1) << bt, m);
      }
      break;
    }
    case K_CPHASE: {
      const int bc = bit(op.q[0]), bt = bit(op.q[1]);
      if (std::min(bc, bt) < line_bits && std::max(bc, bt) >= line_bits) {
        // the quarter's lines are the half with the upper bit set as the pattern asks
        const bool ch = bc > bt;
        QSV_LAUNCH_BLOCKS((k_cphase_half<C>), blocks_for(N / 2, TPB), TPB, 0, st, psi, N / 2,
                          ch ? bc : bt, ch ? op.cs : 1, ch ? bt : bc, ch ? 1 : op.cs,
                          to_c<C>(op.m[0]));
      } else if (std::min(bc, bt) < line_bits) {  // both bits inside a line: sweep all
        const C one = mk<C>(1, 0), lam = to_c<C>(op.m[0]);
        // k_diag2 index p = 2*bit(bc) + bit(bt): the phase sits at (cs, 1)
        QSV_LAUNCH_BLOCKS((k_diag2<C, NP>), blocks_for(N, per), TPB, 0, st, psi, N, bc, bt, one,
                          op.cs == 0 ? lam : one, one, op.cs == 1 ? lam : one);
      } else {
        const uint64_t b1 = (uint64_t(op.cs) << bc) | (uint64_t(1) << bt);
        QSV_LAUNCH_BLOCKS((k_cphase<C, NP>), blocks_for(N / 4, per), TPB, 0, st, psi,
                          N / 4, std::min(bc, bt), std::max(bc, bt), b1, to_c<C>(op.m[0]));
      }
      break;
    }
    case K_SWAP: {
      const int b0 = bit(op.q[0]), b1 = bit(op.q[1]);
      const int blo = std::min(b0, b1), bhi = std::max(b0, b1);
      const int scale = op.m[0] != cd(1.0, 0.0);
      if (blo < line_bits && n >= 7) {  // the pairs interleave within lines: whole-line pairs
        QSV_LAUNCH_BLOCKS((k_swap_lanes<C>), blocks_for(N / 2, TPB), TPB, 0, st, psi, N / 2, blo,
                          bhi, to_c<C>(op.m[0]), scale);
      } else {
        QSV_LAUNCH_BLOCKS((k_swap<C, NP>), blocks_for(N / 4, per), TPB, 0, st, psi,
                          N / 4, blo, bhi, to_c<C>(op.m[0]), scale);
      }
      break;
    }
    case K_DENSE2: {
      M4<C> m;
      for (int i = 0; i < 16; ++i) m.a[i] = to_c<C>(op.m[i]);
      QSV_LAUNCH_BLOCKS((k_dense2<C, 1>), blocks_for(N / 4, TPB), TPB, 0, st, psi,
                         N / 4, bit(op.q[0]), bit(op.q[1]), m);
      break;
    }
    case K_DIAG2:
      QSV_LAUNCH_BLOCKS((k_diag2<C, NP>), blocks_for(N, per), TPB, 0, st, psi, N,
                         bit(op.q[0]), bit(op.q[1]), to_c<C>(op.m[0]), to_c<C>(op.m[1]),
                         to_c<C>(op.m[2]), to_c<C>(op.m[3]));
      break;
    default:
      return fail(QSV_EUNSUP, "kernel kind " + std::to_string(op.kind) + " has no plain kernel");
  }
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template <typename R>
int launch_set_basis(void* p, int n, uint64_t idx, hipStream_t st) {
  using C = typename CT<R>::type;
  const uint64_t N = uint64_t(1) << n;
  QSV_LAUNCH_BLOCKS((k_set_basis<C>), blocks_for(N, TPB), TPB, 0, st,
                     static_cast<C*>(p), N, idx);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template int launch_plain<double>(const KOp&, void*, int, hipStream_t, const PlainTuning&);
template int launch_plain<float>(const KOp&, void*, int, hipStream_t, const PlainTuning&);
template int launch_set_basis<double>(void*, int, uint64_t, hipStream_t);
template int launch_set_basis<float>(void*, int, uint64_t, hipStream_t);

}  // namespace qsv
