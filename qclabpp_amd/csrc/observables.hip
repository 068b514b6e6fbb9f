// Pauli-operator kernels with FP64 accumulation.
//
// A Pauli string maps a basis state |y> to w(y)|y ^ flip> (X and Y flip a
// bit; Z and Y contribute (-1)^bit; each Y adds a factor i).  Expectation
// values and adjoint-gradient terms are inner products <lam|G|phi>; H|psi>
// gathers in[x ^ flip] once per distinct flip mask.  Every reduction runs on a
// fixed grid (kRedBlocks x kRedThreads) with per-thread FP64 sums, a fixed
// shared-memory tree per block, and a separate fixed-order pass over the
// block partials, so results are bit-reproducible run to run.
#include <algorithm>
#include <map>

#include "observables.hpp"

namespace qsv {

int pauli_from_term(const qsv_pauli_term& t, int n, PauliOp& o) {
  o = PauliOp();
  if (t.nfactors < 0 || t.nfactors > QSV_MAX_PAULI_FACTORS)
    return fail(QSV_EINVAL, "pauli term: nfactors out of range");
  int ny = 0;
  uint64_t used = 0;
  for (int f = 0; f < t.nfactors; ++f) {
    const int q = t.qubits[f];
    const char p = t.paulis[f];
    if (p == 'I' || p == 'i') continue;
    if (q < 0 || q >= n) return fail(QSV_EQUBIT, "pauli term: qubit out of range");
    const uint64_t b = uint64_t(1) << (n - 1 - q);
    if (used & b) return fail(QSV_EQUBIT, "pauli term: repeated qubit");
    used |= b;
    switch (p) {
      case 'X': case 'x': o.flip |= b; break;
      case 'Y': case 'y': o.flip |= b; o.zmask |= b; ++ny; break;
      case 'Z': case 'z': o.zmask |= b; break;
      default: return fail(QSV_EINVAL, "pauli term: factor must be I, X, Y or Z");
    }
  }
  static const double re[4] = {1, 0, -1, 0}, im[4] = {0, 1, 0, -1};
  o.phase_re = re[ny & 3];
  o.phase_im = im[ny & 3];
  o.coeff = t.coeff;
  return QSV_OK;
}

namespace {

__device__ __forceinline__ double2 cmul_conj(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

template <typename C>
__device__ __forceinline__ double2 to_d(C v) {
  return make_double2(double(v.x), double(v.y));
}

__device__ __forceinline__ double2 block_sum(double2 v, double2* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[threadIdx.x].x += sh[threadIdx.x + s].x;
      sh[threadIdx.x].y += sh[threadIdx.x + s].y;
    }
    __syncthreads();
  }
  return sh[0];
}

// partial[block] = sum over y handled by the block of
//   (-1)^popc(y & zmask) conj(lam[y ^ flip]) phi[y],   (y & pmask) == pval
template <typename C>
__global__ __launch_bounds__(kRedThreads) void k_pauli_inner(const C* __restrict__ lam,
                                                             const C* __restrict__ phi,
                                                             uint64_t N, uint64_t flip,
                                                             uint64_t zmask, uint64_t pmask,
                                                             uint64_t pval, double2* partials) {
  __shared__ double2 sh[kRedThreads];
  double2 acc = make_double2(0, 0);
  const uint64_t stride = uint64_t(gridDim.x) * kRedThreads;
  for (uint64_t y = uint64_t(blockIdx.x) * kRedThreads + threadIdx.x; y < N; y += stride) {
    if ((y & pmask) != pval) continue;
    double2 t = cmul_conj(to_d(lam[y ^ flip]), to_d(phi[y]));
    if (__popcll(y & zmask) & 1) t = make_double2(-t.x, -t.y);
    acc.x += t.x;
    acc.y += t.y;
  }
  const double2 s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}


// rows x kRedBlocks partials -> rows values, fixed order.
__global__ __launch_bounds__(kRedThreads) void k_reduce_rows(const double2* __restrict__ partials,
                                                             double2* __restrict__ out) {
  __shared__ double2 sh[kRedThreads];
  const double2* row = partials + size_t(blockIdx.x) * kRedBlocks;
  double2 acc = make_double2(0, 0);
  for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) {
    acc.x += row[i].x;
    acc.y += row[i].y;
  }
  const double2 s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

constexpr int kDiagMax = 32;

// Up to 32 diagonal Pauli terms in one pass; partials[t * kRedBlocks + block].
// The grid's stride is 2^sb (sb = 18), so the amplitudes x = x0 + k 2^sb of a
// thread share their low sb bits: a term's sign is (-1)^popc(x0 & z), fixed
// for the thread, times (-1)^popc(k & (z >> sb)).  Terms with no bit at or
// above sb share one running sum; the others accumulate with the second
// factor only, and the first is applied at the end.  Negation is exact, so
// the sums are bit-identical to applying the full sign per amplitude.
template <typename C>
__global__ __launch_bounds__(kRedThreads) void k_diag_expect(const C* __restrict__ psi, uint64_t N,
                                                             const uint64_t* __restrict__ zm,
                                                             int nt, double2* partials) {
  __shared__ double2 sh[kRedThreads];
  __shared__ uint64_t z[kDiagMax];
  if (threadIdx.x < nt) z[threadIdx.x] = zm[threadIdx.x];
  __syncthreads();
  const uint64_t stride = uint64_t(gridDim.x) * kRedThreads;  // a power of two
  const int sb = __ffsll(static_cast<long long>(stride)) - 1;
  const uint64_t x0 = uint64_t(blockIdx.x) * kRedThreads + threadIdx.x;
  double all = 0;
  double acc[kDiagMax];
#pragma unroll
  for (int t = 0; t < kDiagMax; ++t) acc[t] = 0;
  uint64_t k = 0;
  for (uint64_t x = x0; x < N; x += stride, ++k) {
    const double2 v = to_d(psi[x]);
    const double p = v.x * v.x + v.y * v.y;
    all += p;
#pragma unroll
    for (int t = 0; t < kDiagMax; ++t) {
      const uint64_t zh = t < nt ? z[t] >> sb : 0;
      if (zh) acc[t] += (__popcll(k & zh) & 1) ? -p : p;
    }
  }
  for (int t = 0; t < nt; ++t) {
    const double s = (z[t] >> sb) ? acc[t] : all;
    const double2 r = block_sum(make_double2((__popcll(x0 & z[t]) & 1) ? -s : s, 0), sh);
    if (threadIdx.x == 0) partials[size_t(t) * kRedBlocks + blockIdx.x] = r;
    __syncthreads();
  }
}

// out[x] = sum_groups in[x ^ flip_g] * sum_{t in g} c_t (-1)^popc((x^flip_g) & z_t) [proj]
template <typename C>
__global__ __launch_bounds__(256) void k_pauli_apply(const C* __restrict__ in, C* __restrict__ out,
                                                     uint64_t N, const DevPauli* __restrict__ terms,
                                                     const int* __restrict__ gstart, int ngroups,
                                                     uint64_t blk0) {
  extern __shared__ unsigned char smem[];
  DevPauli* st = reinterpret_cast<DevPauli*>(smem);
  int* gs = reinterpret_cast<int*>(st + gstart[ngroups]);
  const int nterms = gstart[ngroups];
  for (int i = threadIdx.x; i < nterms; i += blockDim.x) st[i] = terms[i];
  for (int i = threadIdx.x; i <= ngroups; i += blockDim.x) gs[i] = gstart[i];
  __syncthreads();
  const uint64_t x = (blk0 + blockIdx.x) * 256 + threadIdx.x;
  if (x >= N) return;
  double2 acc = make_double2(0, 0);
  for (int g = 0; g < ngroups; ++g) {
    const uint64_t y = x ^ st[gs[g]].flip;
    const double2 v = to_d(in[y]);
    double cr = 0, ci = 0;
    for (int t = gs[g]; t < gs[g + 1]; ++t) {
      if ((y & st[t].pmask) != st[t].pval) continue;
      const double sg = (__popcll(y & st[t].zmask) & 1) ? -1.0 : 1.0;
      cr += sg * st[t].c_re;
      ci += sg * st[t].c_im;
    }
    acc.x += cr * v.x - ci * v.y;
    acc.y += cr * v.y + ci * v.x;
  }
  out[x] = mk<C>(static_cast<decltype(C::x)>(acc.x), static_cast<decltype(C::x)>(acc.y));
}

// Pauli sums whose terms flip at most one bit (diagonal terms, single X or
// Y factors times diagonal ones: the transverse-field Ising Hamiltonian),
// in a few passes over tiles of 2^12 amplitudes instead of one gather of
// the state per flip mask.  Pass 0 takes the tiles of address bits 0..11;
// pass j >= 1 the line bits 0..L-1 (L = 3 for complex128, 4 for complex64)
// and the bits lo..lo+11-L.  A pass handles the flips of the bits [flo, fhi)
// within its tile (pass 0 also the diagonal terms): each block loads its
// tile of `in` into LDS, where the partner in[x ^ 2^b] of every such flip
// bit b is, and writes out[x] (pass 0) or adds to it (later passes).
constexpr int kFlipTileBits = 12;
constexpr int kFlipThreads = 512;  // 8 amplitudes per thread, 16 waves per CU (two 64 KiB tiles)
constexpr int kFlipGroupsMax = 32;  // groups per launch
typedef const __attribute__((address_space(4))) DevPauli* DevPauliConst;
// The groups one launch handles: group index and the tile-local bit of its
// flip (-1: the diagonal group).
struct FlipGroups {
  int32_t n;
  int16_t g[kFlipGroupsMax];
  int8_t lb[kFlipGroupsMax];
};
template <typename C>
__global__ __launch_bounds__(kFlipThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_pauli_flip_tiles(
    const C* __restrict__ in, C* __restrict__ out, DevPauliConst terms, const int* __restrict__ gstart,
    FlipGroups fg, int line, int lo, int first, uint64_t blk0) {
  constexpr int TN = 1 << kFlipTileBits;
  constexpr int PER = TN / kFlipThreads;
  __shared__ C tile[TN];
  // the tile's base address: tile index bits below lo go to address bits
  // line..lo-1, the rest above the tile's bits
  const int hi = lo + kFlipTileBits - line;  // first address bit above the tile
  const uint64_t t = blk0 + blockIdx.x;
  const uint64_t low_n = uint64_t(lo - line);
  const uint64_t base = ((t & ((uint64_t(1) << low_n) - 1)) << line) | ((t >> low_n) << hi);
  const uint32_t lmask = (1u << line) - 1;
  auto addr = [&](uint32_t p) { return base | (p & lmask) | (uint64_t(p >> line) << lo); };
  // the tile, and the partial sums of earlier launches
  double2 acc[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t p = threadIdx.x + uint32_t(kFlipThreads) * j;
    tile[p] = in[addr(p)];
    acc[j] = first ? make_double2(0, 0) : to_d(out[addr(p)]);
  }
  __syncthreads();
  // groups outside, the thread's amplitudes inside: each term's scalar
  // loads serve all of them
  for (int i = 0; i < fg.n; ++i) {
    const int g = fg.g[i], lb = fg.lb[i];
    const int k0 = gstart[g], k1 = gstart[g + 1];
    const uint64_t f = terms[k0].flip;
    double cr[PER], ci[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) cr[j] = ci[j] = 0;
    for (int k = k0; k < k1; ++k) {
      const uint64_t zm = terms[k].zmask, pm = terms[k].pmask, pv = terms[k].pval;
      const double tr = terms[k].c_re, tim = terms[k].c_im;
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        const uint64_t y = addr(threadIdx.x + uint32_t(kFlipThreads) * j) ^ f;
        const double on = (y & pm) == pv ? ((__popcll(y & zm) & 1) ? -1.0 : 1.0) : 0.0;
        cr[j] += on * tr;
        ci[j] += on * tim;
      }
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t p = threadIdx.x + uint32_t(kFlipThreads) * j;
      const double2 v = to_d(tile[lb < 0 ? p : p ^ (1u << lb)]);
      acc[j].x += cr[j] * v.x - ci[j] * v.y;
      acc[j].y += cr[j] * v.y + ci[j] * v.x;
    }
  }
#pragma unroll
  for (int j = 0; j < PER; ++j)
    out[addr(threadIdx.x + uint32_t(kFlipThreads) * j)] =
        mk<C>(static_cast<decltype(C::x)>(acc[j].x), static_cast<decltype(C::x)>(acc[j].y));
}

}  // namespace

bool pauli_flips_single(const std::vector<DevPauli>& terms, const std::vector<int>& group_start) {
  for (size_t g = 0; g + 1 < group_start.size(); ++g)
    if (__builtin_popcountll(terms[size_t(group_start[g])].flip) > 1) return false;
  return true;
}

template <typename R>
int launch_pauli_apply_tiles(const void* in, void* out, int n, const std::vector<DevPauli>& host_terms,
                             const std::vector<int>& host_group_start, const void* dev_terms,
                             const void* dev_group_start, hipStream_t st) {
  using C = typename CT<R>::type;
  if (n < kFlipTileBits) return fail(QSV_EINVAL, "tiled Pauli sum needs at least 12 qubits");
  if (host_terms.size() > size_t(kMaxPauliTerms)) return fail(QSV_EINVAL, "Pauli sum has too many terms");
  if (!pauli_flips_single(host_terms, host_group_start))
    return fail(QSV_EINVAL, "tiled Pauli sum: a term flips more than one qubit");
  const int line = sizeof(C) == 16 ? 3 : 4;
  const int ngroups = int(host_group_start.size()) - 1;
  const uint64_t ntiles = uint64_t(1) << (n - kFlipTileBits);
  const auto* ti = static_cast<const C*>(in);
  auto* to = static_cast<C*>(out);
  const auto dt = (DevPauliConst)dev_terms;
  const auto* dg = static_cast<const int*>(dev_group_start);
  bool first = true;
  // one tile range: launches of at most kFlipGroupsMax of its groups each,
  // the first of the sum writing `out`, the others adding to it
  auto run = [&](int lo, int flo, int fhi, bool diag) -> int {
    FlipGroups fg{};
    auto flush = [&]() -> int {
      if (fg.n == 0) return QSV_OK;
      QSV_LAUNCH_BLOCKS((k_pauli_flip_tiles<C>), ntiles, kFlipThreads, 0, st, ti, to, dt, dg, fg, line, lo,
                        first ? 1 : 0);
      QSV_CHECK_HIP(hipGetLastError());
      first = false;
      fg.n = 0;
      return QSV_OK;
    };
    for (int g = 0; g < ngroups; ++g) {
      const uint64_t f = host_terms[size_t(host_group_start[g])].flip;
      int lb;
      if (f == 0) {
        if (!diag) continue;
        lb = -1;
      } else {
        const int b = __builtin_ctzll(f);
        if (b < flo || b >= fhi) continue;
        lb = b < line ? b : b - lo + line;
      }
      fg.g[fg.n] = int16_t(g);
      fg.lb[fg.n] = int8_t(lb);
      if (++fg.n == kFlipGroupsMax) QSV_TRY(flush());
    }
    return flush();
  };
  // the tiles of address bits 0..11: their flips and the diagonal terms
  QSV_TRY(run(line, 0, kFlipTileBits, true));
  // then the flips of the next 12 - line bits each, in tiles of the line bits
  // and a 12 - line bit range holding them (slid down onto the top bits for
  // the last ones)
  const int step = kFlipTileBits - line;
  for (int f0 = kFlipTileBits; f0 < n; f0 += step)
    QSV_TRY(run(std::min(f0, n - step), f0, std::min(n, f0 + step), false));
  if (first) {  // no term at all reached a launch: out = 0
    FlipGroups none{};
    QSV_LAUNCH_BLOCKS((k_pauli_flip_tiles<C>), ntiles, kFlipThreads, 0, st, ti, to, dt, dg, none, line, line, 1);
    QSV_CHECK_HIP(hipGetLastError());
  }
  return QSV_OK;
}

void group_terms(const std::vector<PauliOp>& terms, std::vector<DevPauli>& out,
                 std::vector<int>& group_start) {
  std::map<uint64_t, std::vector<const PauliOp*>> by_flip;
  for (const PauliOp& t : terms) by_flip[t.flip].push_back(&t);
  out.clear();
  group_start.clear();
  for (auto& kv : by_flip) {
    group_start.push_back(int(out.size()));
    for (const PauliOp* t : kv.second)
      out.push_back(DevPauli{t->flip, t->zmask, t->pmask, t->pval, t->coeff * t->phase_re,
                             t->coeff * t->phase_im});
  }
  group_start.push_back(int(out.size()));
}

template <typename R>
int launch_pauli_inner(const void* lam, const void* phi, int n, const PauliOp& g, void* partials,
                       hipStream_t st) {
  using C = typename CT<R>::type;
  hipLaunchKernelGGL((k_pauli_inner<C>), dim3(kRedBlocks), dim3(kRedThreads), 0, st,
                     static_cast<const C*>(lam), static_cast<const C*>(phi), uint64_t(1) << n,
                     g.flip, g.zmask, g.pmask, g.pval, static_cast<double2*>(partials));
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

int launch_reduce_rows(const void* partials, int rows, void* out, hipStream_t st) {
  if (rows <= 0) return QSV_OK;
  hipLaunchKernelGGL(k_reduce_rows, dim3(rows), dim3(kRedThreads), 0, st,
                     static_cast<const double2*>(partials), static_cast<double2*>(out));
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template <typename R>
int launch_diag_expect(const void* psi, int n, const std::vector<PauliOp>& terms, void* partials,
                       void* tmp_terms, hipStream_t st) {
  using C = typename CT<R>::type;
  const int nt = int(terms.size());
  if (nt > kDiagMax) return fail(QSV_EINVAL, "too many diagonal terms for one pass");
  std::vector<uint64_t> z(nt);
  for (int i = 0; i < nt; ++i) z[i] = terms[i].zmask;
  QSV_CHECK_HIP(hipMemcpyAsync(tmp_terms, z.data(), nt * sizeof(uint64_t), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL((k_diag_expect<C>), dim3(kRedBlocks), dim3(kRedThreads), 0, st,
                     static_cast<const C*>(psi), uint64_t(1) << n,
                     static_cast<const uint64_t*>(tmp_terms), nt, static_cast<double2*>(partials));
  QSV_CHECK_HIP(hipGetLastError());
  // keep `z` alive until the copy has been consumed
  QSV_CHECK_HIP(hipStreamSynchronize(st));
  return QSV_OK;
}

template <typename R>
int launch_pauli_apply(const void* in, void* out, int n, const void* dev_terms,
                       const void* dev_group_start, int ngroups, int nterms, hipStream_t st) {
  using C = typename CT<R>::type;
  const uint64_t N = uint64_t(1) << n;
  if (nterms > kMaxPauliTerms) return fail(QSV_EINVAL, "Pauli sum has too many terms");
  // terms + group table in LDS
  const size_t lds = size_t(nterms) * sizeof(DevPauli) + size_t(ngroups + 1) * sizeof(int);
  QSV_LAUNCH_BLOCKS((k_pauli_apply<C>), (N + 255) / 256, 256, lds, st, static_cast<const C*>(in),
                    static_cast<C*>(out), N, static_cast<const DevPauli*>(dev_terms),
                    static_cast<const int*>(dev_group_start), ngroups);
  QSV_CHECK_HIP(hipGetLastError());
  return QSV_OK;
}

template int launch_pauli_inner<double>(const void*, const void*, int, const PauliOp&, void*, hipStream_t);
template int launch_pauli_inner<float>(const void*, const void*, int, const PauliOp&, void*, hipStream_t);
template int launch_diag_expect<double>(const void*, int, const std::vector<PauliOp>&, void*, void*, hipStream_t);
template int launch_diag_expect<float>(const void*, int, const std::vector<PauliOp>&, void*, void*, hipStream_t);
template int launch_pauli_apply<double>(const void*, void*, int, const void*, const void*, int, int, hipStream_t);
template int launch_pauli_apply_tiles<double>(const void*, void*, int, const std::vector<DevPauli>&,
                                              const std::vector<int>&, const void*, const void*, hipStream_t);
template int launch_pauli_apply_tiles<float>(const void*, void*, int, const std::vector<DevPauli>&,
                                             const std::vector<int>&, const void*, const void*, hipStream_t);
template int launch_pauli_apply<float>(const void*, void*, int, const void*, const void*, int, int, hipStream_t);

}  // namespace qsv
