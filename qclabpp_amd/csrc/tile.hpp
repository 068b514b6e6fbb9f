// LDS tile passes: a pass loads a tile of 2^t amplitudes (the values of t
// chosen bits, the others fixed) into LDS, applies a list of gates that act
// only on the tile's bits (diagonal gates may act on any bit), and stores it
// back.  Used by the batch engine: a circuit whose state fits one tile runs
// entirely in LDS; a larger one runs as a sequence of passes, each absorbing
// as many gates as the dependencies allow.
#pragma once

#include <vector>

#include "common.hpp"

namespace qsv {

constexpr int kTileMaxBits = 14;  // 2^14 complex<float> = 128 KiB of LDS
constexpr int kTileThreads = 256;

// One operation inside a tile.  Bit operands: >= 0 is a tile-local bit,
// < 0 encodes an address bit outside the tile as ~bit (diagonal ops only).
enum TKind : int32_t {
  T_DENSE1 = 1,  // 2x2 on local bit a                      m[0..3]
  T_DIAG1 = 2,   // multiply by m[bit(a)]                    m[0..1]
  T_DIAG2 = 3,   // multiply by m[2*bit(a) + bit(b)]         m[0..3]
  T_CDENSE1 = 4, // 2x2 on local bit a where bit(b) == cs    m[0..3]
  T_SWAP = 5,    // exchange |a=1,b=0> and |a=0,b=1>, times m[0]
  T_DENSE2 = 6,  // 4x4 on local bits (a, b), row = 2*bit(a)+bit(b)  m[0..15]
};

struct TOp {
  int32_t kind;
  int32_t a, b, cs;
  double m[32];  // complex coefficients, interleaved
};

struct TPass {
  int32_t circuit;
  int32_t t;            // tile bits
  int32_t op_begin, op_end;
  int32_t init_zero;    // first pass of a circuit: the tile starts as |0...0>
  int32_t last;         // last pass of its circuit
  int32_t nruns;
  uint8_t tbits[kTileMaxBits];
  uint8_t rdst[kTileMaxBits + 2];
  uint8_t rlen[kTileMaxBits + 2];
  uint64_t ntiles;
};

struct TCircuit {
  uint64_t offset;  // amplitude offset in the packed buffer
  int32_t n;
  int32_t npasses;
};

struct TItem {
  int32_t pass;
  int32_t pad;
  uint64_t tile;
};

// Host compile of one circuit into passes over tiles of `t` bits.
// `merge1q` multiplies runs of 1-qubit gates on the same qubit into one.
// `origin` (optional, with merge1q false): index in `ops` of each TOp.
int compile_tiled(const std::vector<KOp>& ops, int n, int t, bool merge1q, int circuit_id,
                  std::vector<TPass>& passes, std::vector<TOp>& tops, std::vector<int>* origin = nullptr);

// The adjoint reverse sweep in LDS tile passes: both states' tiles are in
// LDS, and each op is preceded, when it carries a parameter, by its
// generator's inner product <lam|G|phi> over the tile.  G's masks are split
// into tile-local bits (`*loc`) and address bits outside the tile (`*out`,
// fixed per tile).  `row` is the op's result row within its pass, -1 if none.
// `fused`: the thread that updates an element also holds its partner under
// the generator's flip (inside the op's pair or quad), and the elements the
// op skips lie outside the generator's projector, so the inner product is
// accumulated inside the op's own loop from the values before the update;
// otherwise it takes a sweep of its own before the op.
// `sync`: a barrier must follow the op, because the next op (or the store
// of the tile) has threads read elements that other threads wrote.  An op's
// elements belong to threads by its loop: diagonal ops, and pair or quad ops
// whose bits are all at local bit 8 or above, give thread tid the elements
// e with e % 256 == tid, as the tile's load and store do.
struct TGen {
  int32_t row;
  uint32_t floc, zloc, ploc, pvloc;
  int32_t fused;
  int32_t sync;
  int32_t pad;
  uint64_t zout, pout, pvout;
};
constexpr int kAdjRowsMax = 128;  // result rows per pass (LDS accumulators)

// One pass: grid of `blocks` blocks (each takes tiles b, b + blocks, ...),
// partial sums of row r of this pass at partials[r * blocks + block].
template <typename R>
int launch_adjoint_tiles(void* phi, void* lam, const TPass& pass, const TOp* ops, const TGen* gens,
                         int nrows, void* partials, int blocks, hipStream_t st);

}  // namespace qsv
