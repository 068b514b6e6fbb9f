// Launchers of the tile-pass kernels.
#pragma once

#include "tile.hpp"

namespace qsv {

struct TileLevelArgs {
  void* psi;                       // packed states
  const TCircuit* circs;
  const TPass* passes;
  const TOp* ops;
  const TItem* items;              // this level's work items
  uint64_t nitems;
  unsigned long long* ticket;      // this level's ticket counter
  unsigned int* done_tiles;        // per circuit
  unsigned long long* stamps;      // per circuit completion stamps
  int tmax;                        // largest tile bits of the level (LDS size)
  unsigned grid;                   // persistent blocks
};

template <typename R>
int launch_tile_level(const TileLevelArgs& a, hipStream_t st);

int launch_tile_prologue(unsigned long long* tickets, int ntickets, unsigned int* done, int ncirc,
                         unsigned long long* start_stamp, hipStream_t st);

}  // namespace qsv
