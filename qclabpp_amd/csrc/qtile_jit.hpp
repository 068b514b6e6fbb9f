// Specialised tiled passes: the passes, phases and ops of a compiled
// QProgram are emitted as straight-line HIP source, one kernel per pass,
// and compiled at program creation with hiprtc for the device.  The op
// kinds, register pairs, roles and LDS slots become constants, so a pass
// runs without the op interpreter's dispatch, selects and register copies.
// Coefficients stay in device memory (the QOp array of the program), so
// programs with the same structure share one compiled module.
//
// No counterpart in the reference (src/qgates/apply.hpp:258-371 sweeps the
// state once per gate); the result is the product of the same lowered gates.
#pragma once

#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "qtile.hpp"

namespace qsv {

struct QJitModule;  // a loaded code object, shared by programs of equal structure

struct QJit {
  std::vector<std::shared_ptr<QJitModule>> mods;
  std::vector<hipFunction_t> fn;  // per pass
  std::vector<uint64_t> ntiles;   // per pass
  std::vector<uint8_t> wsrc, wdst;  // per pass: reads / writes the work buffer
  unsigned threads = kQThreads;   // block size: 2^(tile bits - register bits)
};

// Tile bits of compiled passes (their phase changes go through a half-tile
// LDS buffer): 2^13-amplitude tiles for complex64 (33 KiB per block) and for
// fixed-form complex128 passes (66 KiB: two blocks of eight waves per CU;
// the n=30 brickwork takes 12 passes instead of 13, 98 ms against 103).
// Complex128 passes that branch per rotation need about 166 registers, so a
// 2^13 tile would leave one block per CU: they keep 2^12.
// QSV_JIT_TILE_BITS=12 or 13 overrides the choice.
inline int jit_tile_bits(int dtype, int n, bool fixed_forms = false) {
  int t = (dtype == QSV_C64 || fixed_forms) && n >= kQTileBitsMax ? kQTileBitsMax : kQTileBits;
  if (const char* e = std::getenv("QSV_JIT_TILE_BITS")) {
    const int v = std::atoi(e);
    if (v >= kQTileBits && v <= kQTileBitsMax && n >= v) t = v;
  }
  return t;
}

// Source of the kernels for passes [p0, p1) of qp (test hook and build step).
std::string qtile_jit_source(const QProgram& qp, bool c128, size_t p0, size_t p1);

// Generate, compile (or take from the process-wide cache) and load the
// pass kernels for the current device.
int qtile_jit_build(const QProgram& qp, bool c128, QJit& out);

// Launch pass i.  ops: the program's device QOp array; work: a buffer of the
// state's size, used by the passes of an out-of-place end (else unused).
int qtile_jit_launch(const QJit& j, size_t i, void* psi, void* work, const QOp* ops, hipStream_t st);

// Compile one source to a gfx950 code object (no device needed).
int qtile_jit_compile(const std::string& src, std::vector<char>& code);

// True when hiprtc loads, so compiled passes can be built.
bool qtile_jit_available();

// Adjoint reverse sweep: the passes of a program whose ops include Q_GEN,
// each compiled as a kernel over two states (phi and lam).  A launch runs
// kAdjPartStride blocks; block b writes the FP64 partial sum of pass-local
// row r at part[2 * (r * kAdjPartStride + b)] (imaginary slot zero), so the
// caller points `part` at its first row of a rows x kAdjPartStride table.
constexpr int kAdjPartStride = 1024;
constexpr size_t kAdjAccBytesMax = 64 * 1024;  // LDS row accumulators of one block
std::string qtile_jit_adjoint_source(const QProgram& qp, bool c128, size_t p0, size_t p1);
int qtile_jit_adjoint_check(const QProgram& qp);
int qtile_jit_build_adjoint(const QProgram& qp, bool c128, QJit& out);
int qtile_jit_launch_adjoint(const QJit& j, size_t i, void* phi, void* lam, const QOp* ops, double* part,
                             hipStream_t st);

// Module groups found in the process-wide cache, loaded from the on-disk
// cache, and compiled, since the library was loaded.
void qtile_jit_cache_stats(int64_t* mem_hits, int64_t* disk_hits, int64_t* compiles);

}  // namespace qsv
