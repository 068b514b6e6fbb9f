// Gate fusion and compiled programs.
#pragma once

#include <vector>

#include "common.hpp"

namespace qsv {

constexpr int kMaxFuse = 5;   // dense blocks: at most 2^5 x 2^5
constexpr int kMaxDiag = 10;  // merged diagonals: at most 2^10 phases

// A fused block: its qubits (ascending, MSB numbering) and its matrix in the
// block's own basis (row index bit p <-> the p-th lowest address bit of the
// block, i.e. qubit q[k-1-p]); row-major 2^k x 2^k, or 2^k phases when
// diagonal.
struct Block {
  std::vector<int> q;
  bool diag = false;
  std::vector<cd> m;
};

struct FusionOptions {
  int max_k = 5;     // largest dense block (1..5); 0 disables fusion
  int diag_k = 8;    // largest merged diagonal
};

// Greedy fusion: gates are merged into open blocks while the union of their
// qubits fits; blocks are emitted when a gate that does not fit touches them.
// Single-gate blocks are emitted as the original op.  Output ops of kind
// K_DENSEK / K_DIAGK index into `blocks`.
void fuse(const std::vector<KOp>& in, int n, const FusionOptions& opt, std::vector<KOp>& out,
          std::vector<Block>& blocks);

// Host reference of one op on a small dense vector (used to build block
// matrices); `v` has 2^k entries, qubits are already local (MSB numbering).
void apply_op_host(const KOp& op, int k, cd* v, const std::vector<Block>* blocks = nullptr);

bool is_diagonal(const KOp& op);

}  // namespace qsv
