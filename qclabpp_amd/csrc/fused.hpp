// Launchers of the fused-block kernels.
#pragma once

#include <vector>

#include "common.hpp"
#include "fusion.hpp"

namespace qsv {

// Address bits of a block, ascending: row bit p of the block matrix is
// address bit b[p].
struct BlockBits {
  uint8_t b[kMaxDiag];
};

// Bit layout of one MFMA wave tile: the block's bits plus the four lowest
// bits outside it (the 16 groups), and the runs of remaining bits that
// enumerate the tiles.
struct TileMap {
  uint8_t tbits[12];
  int ntb;
  uint32_t local_block_mask;
  uint32_t local_group_mask;
  int nruns;
  uint8_t rdst[16];
  uint8_t rlen[16];
};

TileMap make_tile_map(const std::vector<int>& qubits, int n, int group_bits);

// Device copies of one block's coefficients.
struct DeviceBlock {
  const void* ulane = nullptr;   // dense: U_real in MFMA A-operand lane order
  const void* matrix = nullptr;  // dense: U row-major complex
  const void* table = nullptr;   // diagonal: 2^k phases
};

template <typename R>
int launch_fused(const KOp& op, const DeviceBlock& db, void* psi, int n, hipStream_t st);

// Bytes of device storage a block needs for element type R, and the host
// image of that storage.
template <typename R>
size_t block_bytes(const Block& b);
template <typename R>
void block_image(const Block& b, unsigned char* dst, DeviceBlock& offsets);

}  // namespace qsv
