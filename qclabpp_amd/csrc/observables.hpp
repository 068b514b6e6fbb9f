// Pauli operators, expectation values and the H|psi> product, with FP64
// accumulation for every element type.
#pragma once

#include <vector>

#include "common.hpp"

namespace qsv {

// G|y> = w(y) |y ^ flip>, w(y) = phase * (-1)^popcount(y & zmask) when
// (y & pmask) == pval, else 0.  Covers Pauli strings (phase = i^#Y, zmask =
// Z and Y bits, flip = X and Y bits), and projector-restricted Paulis
// (controlled-rotation and phase-gate generators).
struct PauliOp {
  uint64_t flip = 0;
  uint64_t zmask = 0;
  uint64_t pmask = 0;
  uint64_t pval = 0;
  double phase_re = 1, phase_im = 0;
  double coeff = 1;  // real weight (Hamiltonian coefficient)
};

// Build from a qsv_pauli_term (validates qubits against n).
int pauli_from_term(const qsv_pauli_term& t, int n, PauliOp& out);

// Deterministic FP64 reductions: every launch uses kRedBlocks blocks.
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

// z = <lam| G |phi> (complex, FP64); lam may equal phi.  Writes kRedBlocks
// partial sums to `partials` (double2 x kRedBlocks) without synchronising.
template <typename R>
int launch_pauli_inner(const void* lam, const void* phi, int n, const PauliOp& g, void* partials,
                       hipStream_t st);

// Reduce `rows` rows of kRedBlocks partial sums each into rows complex values.
int launch_reduce_rows(const void* partials, int rows, void* out, hipStream_t st);

// Diagonal terms (flip == 0, no projector) in one pass: sum_x |psi_x|^2 *
// sign_t(x) for up to 32 terms; partials layout [term][kRedBlocks] (real).
template <typename R>
int launch_diag_expect(const void* psi, int n, const std::vector<PauliOp>& terms, void* partials,
                       void* tmp_terms, hipStream_t st);

// out = sum_t coeff_t P_t in (terms grouped by flip mask on the host).
constexpr int kMaxPauliTerms = 1024;
template <typename R>
int launch_pauli_apply(const void* in, void* out, int n, const void* dev_terms,
                       const void* dev_group_start, int ngroups, int nterms, hipStream_t st);

// Device image of a Pauli sum for launch_pauli_apply.
struct DevPauli {
  uint64_t flip, zmask, pmask, pval;
  double c_re, c_im;  // coeff * phase
};

// The same for sums whose terms each flip at most one bit (n >= 12): a few
// passes over 2^12-amplitude tiles, each holding the partners of the flips
// it handles, instead of one gather of the state per flip mask.
template <typename R>
int launch_pauli_apply_tiles(const void* in, void* out, int n, const std::vector<DevPauli>& host_terms,
                             const std::vector<int>& host_group_start, const void* dev_terms,
                             const void* dev_group_start, hipStream_t st);
// Whether every group (terms sharing a flip mask) flips at most one bit.
bool pauli_flips_single(const std::vector<DevPauli>& terms, const std::vector<int>& group_start);
void group_terms(const std::vector<PauliOp>& terms, std::vector<DevPauli>& out,
                 std::vector<int>& group_start);

}  // namespace qsv
