// A compiled circuit: its launch sequence and device-resident coefficients.
#pragma once

#include <mutex>
#include <vector>

#include "common.hpp"
#include "fused.hpp"
#include "fusion.hpp"
#include "qtile.hpp"
#include "qtile_jit.hpp"

struct qsv_program_info;

namespace qsv {

struct Program {
  int n = 0;
  int dtype = QSV_C128;
  size_t ngates_in = 0;
  std::vector<KOp> ops;
  std::vector<Block> blocks;
  // Block coefficients: host image built at compile time, copied to the
  // device by upload() (first run).
  std::vector<unsigned char> image;
  mutable std::vector<DeviceBlock> dblocks;
  // tiled execution (fuse_k == QSV_TILED): passes, phases and ops; their
  // device copies sit in dbuf at the offsets below
  bool tiled = false;
  bool jit = false;  // QSV_TILED_JIT: the passes run as compiled kernels
  QProgram qt;
  QJit qj;
  size_t qt_pass_off = 0, qt_phase_off = 0, qt_op_off = 0;
  mutable void* dbuf = nullptr;
  // work buffer of the state's size for a schedule with an out-of-place end
  // (the last two passes: state -> work -> state); runs of one such program
  // on different streams must not overlap
  void* work = nullptr;
  size_t work_bytes = 0;
  mutable bool uploaded = false;
  mutable std::mutex mu;

  int upload() const;
  int build(int n, int dtype, const qsv_gate* gates, size_t ngates, int fuse_k, int diag_k);
  int run(void* psi, hipStream_t st) const;
  void info(qsv_program_info* out) const;
  ~Program();
};

}  // namespace qsv
