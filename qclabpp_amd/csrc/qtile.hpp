// Tiled execution of one large circuit: each launch ("pass") streams the
// state through LDS once, in tiles of 2^T amplitudes (T address bits, the
// lowest line bits always among them), and applies every gate of the pass
// to the tile.  Inside a pass the gates are grouped into phases: a phase
// moves each thread's 2^RB amplitudes (RB tile bits, the phase's qubits)
// from LDS into registers, applies its gates there one by one on the VALU,
// and writes them back.  A phase costs one LDS round trip; a pass costs one
// HBM round trip.
//
// No counterpart in the reference, which sweeps the state once per gate
// (include/qclab/QCircuit.hpp:130-136, src/qgates/apply.hpp:258-371); the
// result is the same product of the same lowered gates.
#pragma once

#include <vector>

#include "common.hpp"

namespace qsv {

constexpr int kQTileBits = 12;     // tile: 2^12 amplitudes (64 KiB of complex<double>)
constexpr int kQRegBits = 4;       // registers: 16 amplitudes per thread
constexpr int kQThreads = 256;     // (1 << kQTileBits) / (1 << kQRegBits)
constexpr int kQTileBitsMax = 13;  // compiled passes may use 2^13-amplitude tiles

// Ops act on a thread's 16 registers.  Inside a phase the host tracks which
// register holds which logical amplitude (an affine relabeling over GF(2)):
// X, CX, SWAP and other permutation gates only update that relabeling, plus
// a diagonal table for their phases, so they cost no register moves; the
// phase's write-back puts every register at its logical tile position.
//
// Pair ops update the register pairs (p, p ^ d) with p the member whose
// bit lowbit(d) is clear ("base").  Bit p of `roles` set: base p holds the
// |1> side of its pair, so the 2x2 is applied with rows and columns swapped.
enum QKind : int32_t {
  Q_PAIR = 1,     // complex 2x2, m[0..3] row-major
  Q_DIAGR = 2,    // v[p] *= m[p] for the registers p in mask b, m[0..15]
  Q_CPAIR = 3,    // Q_PAIR on the pairs whose base p has bit p of `act` set
  Q_DENSE2 = 6,   // 4x4 on register bits (a, b), row = 2*bit(a) + bit(b), m[0..15]
  Q_REALP = 7,    // real 2x2: re(m[0..3])
  Q_ROTP = 8,     // real rotation [[c, -s], [s, c]]: c = re(m0), s = re(m2)
  Q_XLIKEP = 9,   // [[p, iq], [ir, s]], p q r s real: re(m0) im(m1) im(m2) re(m3)
  // Scaled forms (see qtile_scale_rotations): the op without a real factor
  // lam that the pass's last op applies for it.
  // k = re(m0); re(m1) = 1 selects the first form, 0 the second.
  Q_ROTS = 10,    // [[1, -k], [k, 1]] or [[k, -1], [1, k]]
  Q_XLS = 11,     // [[1, ik], [ik, 1]] or [[k, i], [i, k]]
  // Adjoint reverse sweep only (two states, phi in v and lam in w): result
  // row `a` (pass-local) gets sum over registers p in `act` of
  // Re(k_p * conj(w[p ^ d]) * v[p]), k_p = (mf[2p], mf[2p+1]) -- small
  // integer constants of the circuit's structure -- and m[0] holds the
  // factor |lam|^2 the scaled rotations took out of both states before it.
  Q_GEN = 12,
};

struct QOp {
  int32_t kind;
  int32_t a, b;
  int32_t d;             // pair distance (a register index mask)
  uint32_t roles, act;   // per base register, see above
  int32_t pad[2];
  double m[32];          // complex coefficients, (re, im) interleaved
  float mf[32];          // the same in single precision
};

struct QPhase {
  uint32_t regmask;                     // tile-local bits held in registers
  uint16_t regpos[1 << kQRegBits];      // LDS slot part of register p at gather (swizzled)
  uint16_t regpos_out[1 << kQRegBits];  // LDS slot part of register p at write-back
  int32_t op_begin, op_end;
};

// LDS slot of tile position p: the low 4 bits are XOR-ed with the two
// nibbles above them, so 16 lanes that differ in any of the three nibbles'
// bits mostly land in different bank groups.  XOR-linear: the slot of
// (a | b), a and b without common bits, is qtile_swizzle(a) ^ qtile_swizzle(b).
inline __host__ __device__ uint32_t qtile_swizzle(uint32_t p) {
  return p ^ (((p >> 4) ^ (p >> 8)) & 15u);
}

// A pass reads its tile with tile-local bit i at address bit tbits[i] and
// writes it back with bit i at tbits_out[i].  Both are the same set of
// address bits (the pass runs in place); tbits_out permutes them when the
// pass moves other qubits onto the line bits for the passes after it (a
// relabeling of qubits to address bits, undone by the end of the program).
// Local bits 0..L-1 are the line bits read (address bits 0..L-1), and the
// line bits written are among local bits 0..2L-1, so the lanes of a wave
// cover whole 128-byte lines on both sides.  Without relabeling the local
// bits are the address bits in ascending order and tbits_out == tbits.
//
// Out-of-place end (compiled passes only): a relabeling schedule whose last
// gate pass would leave qubits away from their home addresses may instead
// have its second-to-last pass write a work buffer of the state's size
// (wdst) and its last pass read that buffer (wsrc) and write the state with
// every qubit home: the last pass's tbits_out and its out-runs (where the
// tile index bits go) then permute all address bits, not only its tile's,
// and no gate-free restore pass is needed.
struct QPass {
  int32_t phase_begin, phase_end;
  int32_t nruns;
  uint8_t tbits[16];      // address bit of tile-local bit i when read
  uint8_t tbits_out[16];  // address bit of tile-local bit i when written
  uint8_t rdst[24];       // runs of address bits outside the tile (read side)
  uint8_t rlen[24];
  uint8_t wsrc, wdst;     // reads / writes the work buffer instead of the state
  uint8_t nruns_out, pad_;
  uint8_t rdst_out[24];   // the same tile index runs on the written side
  uint8_t rlen_out[24];
  uint64_t ntiles;
};

struct QProgram {
  int n = 0;
  std::vector<QPass> passes;
  std::vector<QPhase> phases;
  std::vector<QOp> ops;
  size_t gates_absorbed = 0;
  int tile_bits = kQTileBits;
  // compiled passes only: each scaled rotation's form is fixed by its
  // coefficients when the pass is generated (no branch on the form)
  bool fixed_forms = false;
};

// Compile lowered, unfused gates (n >= tile_bits) into passes and phases.
// line_bits: address bits inside one 128-byte line, always kept in the tile.
// tile_bits: address bits per tile (kQTileBits for the interpreting kernel,
// up to kQTileBitsMax for compiled passes).  relabel: passes may move qubits
// onto the line bits (tbits_out != tbits; compiled passes only), so a tile
// need not spend its line bits on qubits the next gates do not touch.
// out_of_place: a relabeling schedule may end out of place (QPass wsrc /
// wdst), which needs a work buffer of the state's size at run time.
int compile_qtile(const std::vector<KOp>& in, int n, int line_bits, QProgram& out,
                  int tile_bits = kQTileBits, bool relabel = false, bool out_of_place = false);

// Whether a program's passes use the work buffer (an out-of-place end).
bool qtile_uses_work(const QProgram& qp);

// Rewrites the real rotations of every pass into scaled forms that cost one
// FMA per real output component instead of two (a rotation [[c, -s], [s, c]]
// is max(|c|, |s|) times a matrix with a unit entry pair), factors the most
// common entry out of each diagonal op so its unit entries drop out, and
// folds the product of the factors into a diagonal op at the end of the
// pass.  Each factor is at least 1/sqrt(2) in modulus; after 120 of them the
// next rotation is kept as a complex 2x2 times their product, which resets
// it, so the stored amplitudes stay within 2^60 of their true scale in
// passes of any length.  The reset points depend on the structure only.
void qtile_scale_rotations(QProgram& qp);

// Hits and misses of the schedule cache: compile_qtile remembers, per
// circuit structure (gate kinds, qubits and control states, not angles),
// which tie rule won, and on a hit runs that rule alone.
void qtile_schedule_cache_stats(int64_t* hits, int64_t* misses);

// Floating-point lane operations (FMA and multiply) of one run's gate
// arithmetic as the compiled passes generate it: per op and thread, 4 per
// register pair of a scaled rotation, 8 of a real 2x2, 16 of a complex one,
// 4 per register of a diagonal and 64 per register quad of a 4x4; times
// the 2^n / 16 threads of a pass.
double qtile_fp_lane_ops(const QProgram& qp);

struct QTileLaunch {
  void* psi;
  const QPass* pass;      // device
  const QPhase* phases;   // device
  const QOp* ops;         // device
  uint64_t ntiles;
};

template <typename R>
int launch_qtile(const QTileLaunch& a, hipStream_t st);

}  // namespace qsv
