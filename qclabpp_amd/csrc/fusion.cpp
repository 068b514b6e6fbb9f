// Gate fusion: collapse runs of gates into dense 2^k x 2^k blocks (k <= 5)
// and runs of diagonal gates into phase tables.  No counterpart in the
// reference, which applies one sweep per gate (include/qclab/QCircuit.hpp:130-136).
// Block matrices are products of the same per-gate operations the plain
// kernels apply, in the MSB (kron) convention of QCircuit::matrix()
// (include/qclab/QCircuit.hpp:121-127).
#include "fusion.hpp"

#include <algorithm>
#include <numeric>

namespace qsv {

bool is_diagonal(const KOp& op) {
  switch (op.kind) {
    case K_DIAG1: case K_PHASE1: case K_CPHASE: case K_DIAG2: case K_DIAGK: case K_NOP:
      return true;
    default:
      return false;
  }
}

static std::vector<int> op_qubits(const KOp& op) {
  std::vector<int> q(op.q, op.q + op.nq);
  std::sort(q.begin(), q.end());
  return q;
}

void apply_op_host(const KOp& op, int k, cd* v, const std::vector<Block>* blocks) {
  const uint64_t N = uint64_t(1) << k;
  auto bit = [k](int q) { return uint64_t(1) << (k - 1 - q); };
  switch (op.kind) {
    case K_NOP:
      return;
    case K_DENSE1: {
      const uint64_t s = bit(op.q[0]);
      for (uint64_t a = 0; a < N; ++a)
        if (!(a & s)) {
          const cd x = v[a], y = v[a | s];
          v[a] = op.m[0] * x + op.m[1] * y;
          v[a | s] = op.m[2] * x + op.m[3] * y;
        }
      return;
    }
    case K_DIAG1: {
      const uint64_t s = bit(op.q[0]);
      for (uint64_t a = 0; a < N; ++a) v[a] *= (a & s) ? op.m[1] : op.m[0];
      return;
    }
    case K_PHASE1: {
      const uint64_t s = bit(op.q[0]);
      for (uint64_t a = 0; a < N; ++a)
        if (a & s) v[a] *= op.m[0];
      return;
    }
    case K_CDENSE1: {
      const uint64_t c = bit(op.q[0]), s = bit(op.q[1]);
      for (uint64_t a = 0; a < N; ++a)
        if (!(a & s) && (((a & c) != 0) == (op.cs == 1))) {
          const cd x = v[a], y = v[a | s];
          v[a] = op.m[0] * x + op.m[1] * y;
          v[a | s] = op.m[2] * x + op.m[3] * y;
        }
      return;
    }
    case K_CPHASE: {
      const uint64_t c = bit(op.q[0]), s = bit(op.q[1]);
      for (uint64_t a = 0; a < N; ++a)
        if ((a & s) && (((a & c) != 0) == (op.cs == 1))) v[a] *= op.m[0];
      return;
    }
    case K_SWAP: {
      const uint64_t s0 = bit(op.q[0]), s1 = bit(op.q[1]);
      for (uint64_t a = 0; a < N; ++a)
        if ((a & s0) && !(a & s1)) {
          const uint64_t b = (a ^ s0) | s1;
          const cd x = v[a], y = v[b];
          v[a] = op.m[0] * y;
          v[b] = op.m[0] * x;
        }
      return;
    }
    case K_DENSE2: {
      const uint64_t s0 = bit(op.q[0]), s1 = bit(op.q[1]);
      for (uint64_t a = 0; a < N; ++a)
        if (!(a & s0) && !(a & s1)) {
          const uint64_t ix[4] = {a, a | s1, a | s0, a | s0 | s1};
          cd x[4], y[4];
          for (int p = 0; p < 4; ++p) x[p] = v[ix[p]];
          for (int r = 0; r < 4; ++r) {
            y[r] = 0;
            for (int c = 0; c < 4; ++c) y[r] += op.m[4 * r + c] * x[c];
          }
          for (int p = 0; p < 4; ++p) v[ix[p]] = y[p];
        }
      return;
    }
    case K_DIAG2: {
      const uint64_t s0 = bit(op.q[0]), s1 = bit(op.q[1]);
      for (uint64_t a = 0; a < N; ++a) v[a] *= op.m[((a & s0) ? 2 : 0) + ((a & s1) ? 1 : 0)];
      return;
    }
    case K_DENSEK: case K_DIAGK: {
      const Block& B = (*blocks)[op.blk];
      const int kb = int(B.q.size());
      const uint64_t D = uint64_t(1) << kb;
      std::vector<uint64_t> sb(kb);  // address bit of block row bit p
      for (int p = 0; p < kb; ++p) sb[p] = bit(B.q[kb - 1 - p]);
      uint64_t bm = 0;
      for (auto s : sb) bm |= s;
      std::vector<cd> x(D), y(D);
      for (uint64_t g = 0; g < N; ++g) {
        if (g & bm) continue;
        for (uint64_t r = 0; r < D; ++r) {
          uint64_t a = g;
          for (int p = 0; p < kb; ++p)
            if (r >> p & 1) a |= sb[p];
          x[r] = v[a];
        }
        for (uint64_t r = 0; r < D; ++r) {
          if (B.diag) {
            y[r] = B.m[r] * x[r];
          } else {
            cd acc = 0;
            for (uint64_t c = 0; c < D; ++c) acc += B.m[r * D + c] * x[c];
            y[r] = acc;
          }
        }
        for (uint64_t r = 0; r < D; ++r) {
          uint64_t a = g;
          for (int p = 0; p < kb; ++p)
            if (r >> p & 1) a |= sb[p];
          v[a] = y[r];
        }
      }
      return;
    }
    default:
      return;
  }
}

namespace {

struct Open {
  std::vector<int> q;    // sorted
  std::vector<int> ops;  // indices into the input, ascending
  bool diag = true;
};

void emit(const Open& b, const std::vector<KOp>& in, std::vector<KOp>& out,
          std::vector<Block>& blocks) {
  if (b.ops.size() == 1) {
    out.push_back(in[b.ops[0]]);
    return;
  }
  const int k = int(b.q.size());
  const uint64_t D = uint64_t(1) << k;
  // remap every op to local qubit numbers (position in the sorted block)
  std::vector<KOp> local;
  local.reserve(b.ops.size());
  for (int i : b.ops) {
    KOp o = in[i];
    for (int j = 0; j < o.nq; ++j)
      o.q[j] = int(std::lower_bound(b.q.begin(), b.q.end(), o.q[j]) - b.q.begin());
    local.push_back(o);
  }
  Block B;
  B.q = b.q;
  B.diag = b.diag;
  if (b.diag) {
    std::vector<cd> v(D, cd(1.0, 0.0));
    for (const KOp& o : local) apply_op_host(o, k, v.data());
    // row index r (block basis) = local MSB index with bit p <-> local qubit k-1-p
    B.m = v;
  } else {
    B.m.assign(D * D, 0.0);
    std::vector<cd> col(D);
    for (uint64_t c = 0; c < D; ++c) {
      std::fill(col.begin(), col.end(), cd(0.0, 0.0));
      col[c] = 1.0;
      for (const KOp& o : local) apply_op_host(o, k, col.data());
      for (uint64_t r = 0; r < D; ++r) B.m[r * D + c] = col[r];
    }
  }
  KOp f;
  f.nq = k;
  for (int j = 0; j < k; ++j) f.q[j] = b.q[j];
  if (k == 1 && !b.diag) {
    f.kind = K_DENSE1;
    for (int i = 0; i < 4; ++i) f.m[i] = B.m[i];
  } else if (k == 1) {
    f.kind = K_DIAG1;
    f.m[0] = B.m[0];
    f.m[1] = B.m[1];
  } else if (k == 2 && !b.diag) {
    f.kind = K_DENSE2;  // row = 2*bit(q[0]) + bit(q[1]): the same convention
    for (int i = 0; i < 16; ++i) f.m[i] = B.m[i];
  } else if (k == 2) {
    f.kind = K_DIAG2;
    for (int i = 0; i < 4; ++i) f.m[i] = B.m[i];
  } else {
    f.kind = b.diag ? K_DIAGK : K_DENSEK;
    f.blk = int(blocks.size());
    blocks.push_back(std::move(B));
  }
  out.push_back(f);
}

}  // namespace

void fuse(const std::vector<KOp>& in, int n, const FusionOptions& opt, std::vector<KOp>& out,
          std::vector<Block>& blocks) {
  (void)n;
  out.clear();
  blocks.clear();
  const int max_k = std::max(1, std::min(opt.max_k, kMaxFuse));
  const int diag_k = std::max(max_k, std::min(opt.diag_k, kMaxDiag));
  std::vector<Open> open;
  for (int i = 0; i < int(in.size()); ++i) {
    const KOp& op = in[i];
    if (op.kind == K_NOP) continue;
    const std::vector<int> Q = op_qubits(op);
    const bool d = is_diagonal(op);
    std::vector<int> hit;  // indices into `open`
    for (int b = 0; b < int(open.size()); ++b)
      for (int q : Q)
        if (std::binary_search(open[b].q.begin(), open[b].q.end(), q)) {
          hit.push_back(b);
          break;
        }
    std::vector<int> closed;
    while (true) {
      std::vector<int> U = Q;
      bool all_diag = d;
      for (int b : hit) {
        U.insert(U.end(), open[b].q.begin(), open[b].q.end());
        all_diag = all_diag && open[b].diag;
      }
      std::sort(U.begin(), U.end());
      U.erase(std::unique(U.begin(), U.end()), U.end());
      const int lim = all_diag ? diag_k : max_k;
      if (int(U.size()) <= lim) {
        Open nb;
        nb.q = U;
        nb.diag = all_diag;
        for (int b : hit) nb.ops.insert(nb.ops.end(), open[b].ops.begin(), open[b].ops.end());
        nb.ops.push_back(i);
        std::sort(nb.ops.begin(), nb.ops.end());
        // remove merged blocks (descending index) and append the new one
        std::vector<int> rm = hit;
        std::sort(rm.rbegin(), rm.rend());
        for (int b : rm) open.erase(open.begin() + b);
        open.push_back(std::move(nb));
        break;
      }
      if (hit.empty()) {  // the gate alone exceeds the limit (max_k < 2)
        out.push_back(op);
        break;
      }
      // close the largest intersecting block and retry with the rest
      auto it = std::max_element(hit.begin(), hit.end(), [&](int a, int b) {
        return open[a].q.size() < open[b].q.size();
      });
      const int b = *it;
      emit(open[b], in, out, blocks);
      open.erase(open.begin() + b);
      hit.erase(it);
      for (int& h : hit)
        if (h > b) --h;
    }
  }
  for (const Open& b : open) emit(b, in, out, blocks);
}

}  // namespace qsv
