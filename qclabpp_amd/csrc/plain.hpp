// Launchers of the plain per-gate kernels.
#pragma once

#include "common.hpp"

namespace qsv {

constexpr int TPB = 256;  // threads per block of the sweep kernels

struct PlainTuning {
  // Dense 1-qubit and controlled gates whose (target) bit is below this use
  // the cross-lane shuffle kernels; at most 6 (the wavefront is 64 lanes).
  int lane_bits = 6;
};

template <typename R>
int launch_plain(const KOp& op, void* psi, int n, hipStream_t st, const PlainTuning& tu);

template <typename R>
int launch_set_basis(void* psi, int n, uint64_t idx, hipStream_t st);

}  // namespace qsv
