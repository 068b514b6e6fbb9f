// Launchers of the tile-pass kernels.
#pragma once

#include "qtile.hpp"
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

// Per pass of the batch engine's register-phase circuits (qtile.hpp passes).
struct QPassInfo {
  int32_t circuit;  // index into the batch's circuits
  int32_t first;    // the circuit's first pass: its tiles start as |0...0>
  int32_t last;     // the circuit's last pass: completion stamping
  int32_t nitems;   // work items of the pass (tiles that can hold nonzeros)
};

// One dependency level of register-phase passes (k_qtile_level, qtile.hip).
struct QTileLevelArgs {
  void* psi;                       // packed states
  const TCircuit* circs;
  const QPass* passes;             // every register-phase pass of the batch
  const QPassInfo* pinfo;
  const QPhase* phases;
  const QOp* ops;
  const TItem* items;              // this level's work items (pass, tile)
  uint64_t nitems;
  unsigned long long* ticket;
  unsigned int* done_tiles;
  unsigned long long* stamps;
  unsigned grid;                   // persistent blocks
};

template <typename R>
int launch_qtile_level(const QTileLevelArgs& a, hipStream_t st);

// Zeros over the states of circuits ids[0..nids) (device array), blocks_per
// blocks of 256 threads per circuit.
template <typename R>
int launch_zero_states(void* psi, const TCircuit* circs, const int32_t* ids, int nids, unsigned blocks_per,
                       hipStream_t st);

}  // namespace qsv
