// Host compile of a circuit into LDS tile passes.
#include "tile.hpp"

#include <algorithm>
#include <set>

namespace qsv {

namespace {

bool diag_kind(int k) {
  return k == K_DIAG1 || k == K_PHASE1 || k == K_CPHASE || k == K_DIAG2 || k == K_NOP;
}

// qubits whose bits must be inside the tile for `op`
std::vector<int> local_qubits(const KOp& op) {
  switch (op.kind) {
    case K_DENSE1: return {op.q[0]};
    case K_CDENSE1: return {op.q[1]};  // target; the control may be outside
    case K_SWAP: case K_DENSE2: return {op.q[0], op.q[1]};
    default: return {};  // diagonal
  }
}

std::vector<int> all_qubits(const KOp& op) { return std::vector<int>(op.q, op.q + op.nq); }

void set_c(TOp& t, int i, cd z) {
  t.m[2 * i] = z.real();
  t.m[2 * i + 1] = z.imag();
}

}  // namespace

int compile_tiled(const std::vector<KOp>& in, int n, int t, bool merge1q, int circuit_id,
                  std::vector<TPass>& passes, std::vector<TOp>& tops, std::vector<int>* origin) {
  if (origin && merge1q) return fail(QSV_EINVAL, "tile compile: origins need unmerged ops");
  if (t < 1) return fail(QSV_EINVAL, "tile bits must be positive");
  t = std::min(t, n);
  if (t > kTileMaxBits) return fail(QSV_EINVAL, "tile too large");
  // 1. optionally merge runs of 1-qubit ops on the same qubit (ops on other
  //    qubits in between commute with them)
  std::vector<KOp> ops;
  std::vector<int> src;            // index in `in` of each op
  std::vector<int> last1q(n, -1);  // index in `ops` of a mergeable 1-qubit op per qubit
  for (size_t ii = 0; ii < in.size(); ++ii) {
    const KOp& op = in[ii];
    if (op.kind == K_NOP) continue;
    const bool one = op.nq == 1 && (op.kind == K_DENSE1 || op.kind == K_DIAG1 || op.kind == K_PHASE1);
    if (merge1q && one && last1q[op.q[0]] >= 0) {
      KOp& prev = ops[last1q[op.q[0]]];
      // matrices of both as 2x2 (row-major), product op * prev
      auto as2 = [](const KOp& o, cd* m) {
        if (o.kind == K_DENSE1) { for (int i = 0; i < 4; ++i) m[i] = o.m[i]; }
        else if (o.kind == K_DIAG1) { m[0] = o.m[0]; m[1] = 0.0; m[2] = 0.0; m[3] = o.m[1]; }
        else { m[0] = 1.0; m[1] = 0.0; m[2] = 0.0; m[3] = o.m[0]; }
      };
      cd a[4], b[4], c[4];
      as2(op, a);
      as2(prev, b);
      c[0] = a[0] * b[0] + a[1] * b[2];
      c[1] = a[0] * b[1] + a[1] * b[3];
      c[2] = a[2] * b[0] + a[3] * b[2];
      c[3] = a[2] * b[1] + a[3] * b[3];
      prev.kind = K_DENSE1;
      for (int i = 0; i < 4; ++i) prev.m[i] = c[i];
      continue;
    }
    for (int q : all_qubits(op)) last1q[q] = -1;
    ops.push_back(op);
    src.push_back(int(ii));
    if (one) last1q[op.q[0]] = int(ops.size()) - 1;
  }
  // 2. passes
  const int first_pass = int(passes.size());
  std::vector<KOp> rem = ops;
  std::vector<int> rem_src = src;
  const int c_low = std::min(5, t);
  bool first = true;
  if (rem.empty()) {  // the identity circuit still needs its |0..0> written
    rem.push_back(KOp());
    rem_src.push_back(-1);
  }
  while (!rem.empty()) {
    std::vector<int> tq;  // tile qubits
    auto add = [&](int q) {
      if (int(tq.size()) < t && std::find(tq.begin(), tq.end(), q) == tq.end()) tq.push_back(q);
    };
    if (n <= t) {
      for (int q = 0; q < n; ++q) add(q);
    } else {
      // the first pending op always fits (progress), then the lowest address
      // bits (contiguous runs), then the qubits the next ops need
      for (const KOp& op : rem)
        if (op.kind != K_NOP) {
          for (int q : local_qubits(op)) add(q);
          break;
        }
      for (int j = 0; j < c_low; ++j) add(n - 1 - j);
      for (const KOp& op : rem) {
        if (int(tq.size()) >= t) break;
        for (int q : local_qubits(op)) add(q);
      }
      for (int q = 0; int(tq.size()) < t && q < n; ++q) add(q);
    }
    std::set<int> tset(tq.begin(), tq.end());
    std::vector<int> tb;  // tile bits ascending
    for (int q : tq) tb.push_back(n - 1 - q);
    std::sort(tb.begin(), tb.end());
    auto local = [&](int q) -> int {
      const int b = n - 1 - q;
      const auto it = std::lower_bound(tb.begin(), tb.end(), b);
      if (it != tb.end() && *it == b) return int(it - tb.begin());
      return ~b;
    };
    TPass P{};
    P.circuit = circuit_id;
    P.t = t;
    P.op_begin = int(tops.size());
    P.init_zero = first ? 1 : 0;
    for (int i = 0; i < t; ++i) P.tbits[i] = uint8_t(tb[i]);
    P.nruns = 0;
    {
      uint64_t tm = 0;
      for (int b : tb) tm |= uint64_t(1) << b;
      for (int b = 0; b < n;) {
        if (tm >> b & 1) { ++b; continue; }
        int e = b;
        while (e < n && !(tm >> e & 1)) ++e;
        P.rdst[P.nruns] = uint8_t(b);
        P.rlen[P.nruns] = uint8_t(e - b);
        ++P.nruns;
        b = e;
      }
    }
    P.ntiles = uint64_t(1) << (n - t);
    std::set<int> blocked;
    std::vector<KOp> deferred;
    std::vector<int> deferred_src;
    for (size_t ri = 0; ri < rem.size(); ++ri) {
      const KOp& op = rem[ri];
      bool ok = op.kind != K_NOP;
      for (int q : local_qubits(op)) ok = ok && tset.count(q);
      for (int q : all_qubits(op)) ok = ok && !blocked.count(q);
      if (op.kind == K_NOP) continue;
      if (!ok) {
        for (int q : all_qubits(op)) blocked.insert(q);
        deferred.push_back(op);
        deferred_src.push_back(rem_src[ri]);
        continue;
      }
      if (origin) origin->push_back(rem_src[ri]);
      TOp to{};
      switch (op.kind) {
        case K_DENSE1:
          to.kind = T_DENSE1; to.a = local(op.q[0]);
          for (int i = 0; i < 4; ++i) set_c(to, i, op.m[i]);
          break;
        case K_DIAG1:
          to.kind = T_DIAG1; to.a = local(op.q[0]);
          set_c(to, 0, op.m[0]); set_c(to, 1, op.m[1]);
          break;
        case K_PHASE1:
          to.kind = T_DIAG1; to.a = local(op.q[0]);
          set_c(to, 0, 1.0); set_c(to, 1, op.m[0]);
          break;
        case K_CDENSE1:
          to.kind = T_CDENSE1; to.a = local(op.q[1]); to.b = local(op.q[0]); to.cs = op.cs;
          for (int i = 0; i < 4; ++i) set_c(to, i, op.m[i]);
          break;
        case K_CPHASE:
          to.kind = T_DIAG2; to.a = local(op.q[0]); to.b = local(op.q[1]);
          for (int i = 0; i < 4; ++i) set_c(to, i, 1.0);
          set_c(to, 2 * op.cs + 1, op.m[0]);
          break;
        case K_SWAP:
          to.kind = T_SWAP; to.a = local(op.q[0]); to.b = local(op.q[1]);
          set_c(to, 0, op.m[0]);
          break;
        case K_DENSE2:
          to.kind = T_DENSE2; to.a = local(op.q[0]); to.b = local(op.q[1]);
          for (int i = 0; i < 16; ++i) set_c(to, i, op.m[i]);
          break;
        case K_DIAG2:
          to.kind = T_DIAG2; to.a = local(op.q[0]); to.b = local(op.q[1]);
          for (int i = 0; i < 4; ++i) set_c(to, i, op.m[i]);
          break;
        default:
          return fail(QSV_EUNSUP, "tile passes take unfused ops of at most 2 qubits");
      }
      tops.push_back(to);
    }
    P.op_end = int(tops.size());
    passes.push_back(P);
    first = false;
    if (deferred.size() == rem.size() && P.op_end == P.op_begin && !rem.empty() &&
        rem[0].kind != K_NOP)
      return fail(QSV_EINVAL, "tile compile made no progress");
    rem.swap(deferred);
    rem_src.swap(deferred_src);
  }
  passes.back().last = 1;
  (void)first_pass;
  return QSV_OK;
}

}  // namespace qsv
