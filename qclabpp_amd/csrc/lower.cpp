// Host-side lowering of QCLAB++ gate records to kernel operations.
//
// Each gate's matrix is written the way its class defines matrix() in the
// reference (cited per case), the Op is folded in as M, M^T or M^H
// (include/qclab/dense/transpose.hpp:12-35), and the result is classified
// into the same sweep class the reference binds the gate to
// (src/qgates/*.cpp; dispatch of src/qgates/apply.hpp:258-371).
#include <cmath>
#include <mutex>
#include <string>

#include "common.hpp"

namespace qsv {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string& msg) { g_err = msg; }
const char* get_error() { return g_err.c_str(); }

namespace {

const cd I1(0.0, 1.0);

// (cos, sin) of a gate: rotations use theta/2 (QRotation.hpp:44-46),
// Phase/CPhase use theta (QAngle.hpp:30-33).  Host libm, like the reference.
void angle_cs(const qsv_gate& g, bool half, double& c, double& s) {
  if (g.has_cs) {
    c = g.cs[0];
    s = g.cs[1];
  } else {
    const double a = half ? g.theta / 2 : g.theta;
    c = std::cos(a);
    s = std::sin(a);
  }
}

// Apply Op to a dim x dim row-major matrix in place.
void fold_op(int op, cd* m, int dim) {
  if (op == QSV_OP_N) return;
  if (op == QSV_OP_C)
    for (int i = 0; i < dim * dim; ++i) m[i] = std::conj(m[i]);
  for (int r = 0; r < dim; ++r)
    for (int c = r + 1; c < dim; ++c) std::swap(m[r * dim + c], m[c * dim + r]);
}

// Column-major interleaved (re,im) -> row-major complex.
void load_colmajor(const double* src, int dim, cd* m) {
  for (int r = 0; r < dim; ++r)
    for (int c = 0; c < dim; ++c) {
      const int k = r + c * dim;
      m[r * dim + c] = cd(src[2 * k], src[2 * k + 1]);
    }
}

void set2(cd* m, cd a00, cd a01, cd a10, cd a11) {
  m[0] = a00; m[1] = a01; m[2] = a10; m[3] = a11;
}

bool is_rotation(int kind) {
  switch (kind) {
    case QSV_GATE_RX: case QSV_GATE_RY: case QSV_GATE_RZ:
    case QSV_GATE_CRX: case QSV_GATE_CRY: case QSV_GATE_CRZ:
    case QSV_GATE_RXX: case QSV_GATE_RYY: case QSV_GATE_RZZ:
    case QSV_GATE_P: case QSV_GATE_CP:
      return true;
    default:
      return false;
  }
}

}  // namespace

int lower_gate(const qsv_gate& g, int n, KOp& o) {
  o = KOp();
  if (g.op != QSV_OP_N && g.op != QSV_OP_T && g.op != QSV_OP_C)
    return fail(QSV_EINVAL, "gate op must be 'N', 'T' or 'C'");
  if (g.kind < 0 || g.kind >= QSV_GATE_COUNT)
    return fail(QSV_EINVAL, "unknown gate kind " + std::to_string(g.kind));
  const bool two = g.kind >= QSV_GATE_CX && g.kind != QSV_GATE_U1;
  const int q0 = g.q[0], q1 = g.q[1];
  if (q0 < 0 || q0 >= n)
    return fail(QSV_EQUBIT, "qubit " + std::to_string(q0) + " out of range for n=" + std::to_string(n));
  if (two) {
    if (q1 < 0 || q1 >= n)
      return fail(QSV_EQUBIT, "qubit " + std::to_string(q1) + " out of range for n=" + std::to_string(n));
    if (q0 == q1) return fail(QSV_EQUBIT, "2-qubit gate on a repeated qubit");
  }
  const bool controlled = g.kind >= QSV_GATE_CX && g.kind <= QSV_GATE_CU;
  if (controlled && g.ctrl_state != 0 && g.ctrl_state != 1)
    return fail(QSV_EINVAL, "control state must be 0 or 1");
  if (is_rotation(g.kind)) {
    o.param = g.param;
    o.gen = g.kind;
  }

  double c = 1, s = 0;
  cd m[16];
  const int op = g.op;
  switch (g.kind) {
    case QSV_GATE_I:  // Identity.cpp:7-11: no-op
      o.kind = K_NOP;
      return QSV_OK;
    case QSV_GATE_H: {  // Hadamard.hpp:46-51
      const double r = 1.0 / std::sqrt(2.0);
      set2(m, r, r, r, -r);
      break;
    }
    case QSV_GATE_X:  // PauliX.hpp:46-49
      set2(m, 0.0, 1.0, 1.0, 0.0);
      break;
    case QSV_GATE_Y:  // PauliY.hpp:46-49
      set2(m, 0.0, -I1, I1, 0.0);
      break;
    case QSV_GATE_RX:  // RotationX.hpp:104-108
      angle_cs(g, true, c, s);
      set2(m, c, cd(0, -s), cd(0, -s), c);
      break;
    case QSV_GATE_RY:  // RotationY.hpp:104-107
      angle_cs(g, true, c, s);
      set2(m, c, -s, s, c);
      break;
    case QSV_GATE_U1:
      load_colmajor(g.mat, 2, m);
      break;
    default:
      break;
  }

  switch (g.kind) {
    case QSV_GATE_H: case QSV_GATE_X: case QSV_GATE_Y:
    case QSV_GATE_RX: case QSV_GATE_RY: case QSV_GATE_U1:
      fold_op(op, m, 2);
      o.kind = K_DENSE1;
      o.nq = 1;
      o.q[0] = q0;
      for (int i = 0; i < 4; ++i) o.m[i] = m[i];
      return QSV_OK;
    case QSV_GATE_RZ: {  // RotationZ.hpp:104-108, full sweep (RotationZ.cpp:7-14)
      angle_cs(g, true, c, s);
      cd d[4] = {cd(c, -s), 0.0, 0.0, cd(c, s)};
      fold_op(op, d, 2);
      o.kind = K_DIAG1;
      o.nq = 1;
      o.q[0] = q0;
      o.m[0] = d[0];
      o.m[1] = d[3];
      return QSV_OK;
    }
    case QSV_GATE_Z: case QSV_GATE_P: case QSV_GATE_T: case QSV_GATE_S: {
      // |1>-half sweep (PauliZ.cpp, Phase.cpp, Phase45.cpp, Phase90.cpp)
      cd lam;
      if (g.kind == QSV_GATE_Z) {
        lam = -1.0;
      } else if (g.kind == QSV_GATE_P) {  // Phase.hpp:119-122
        angle_cs(g, false, c, s);
        lam = cd(c, s);
      } else if (g.kind == QSV_GATE_T) {  // Phase45.hpp:46-51
        const double r = 1.0 / std::sqrt(2.0);
        lam = cd(r, r);
      } else {  // Phase90.hpp:46-49
        lam = I1;
      }
      if (op == QSV_OP_C) lam = std::conj(lam);
      o.kind = K_PHASE1;
      o.nq = 1;
      o.q[0] = q0;
      o.m[0] = lam;
      return QSV_OK;
    }
    default:
      break;
  }

  if (controlled) {
    o.nq = 2;
    o.q[0] = q0;  // control
    o.q[1] = q1;  // target
    o.cs = g.ctrl_state;
    if (g.kind == QSV_GATE_CZ || g.kind == QSV_GATE_CP) {  // CZ.cpp, CPhase.cpp
      cd lam = -1.0;
      if (g.kind == QSV_GATE_CP) {
        angle_cs(g, false, c, s);
        lam = cd(c, s);
      }
      if (op == QSV_OP_C) lam = std::conj(lam);
      o.kind = K_CPHASE;
      o.m[0] = lam;
      return QSV_OK;
    }
    switch (g.kind) {  // the controlled 1-qubit gate's matrix
      case QSV_GATE_CX: set2(m, 0.0, 1.0, 1.0, 0.0); break;
      case QSV_GATE_CY: set2(m, 0.0, -I1, I1, 0.0); break;
      case QSV_GATE_CRX:
        angle_cs(g, true, c, s);
        set2(m, c, cd(0, -s), cd(0, -s), c);
        break;
      case QSV_GATE_CRY:
        angle_cs(g, true, c, s);
        set2(m, c, -s, s, c);
        break;
      case QSV_GATE_CRZ:
        angle_cs(g, true, c, s);
        set2(m, cd(c, -s), 0.0, 0.0, cd(c, s));
        break;
      case QSV_GATE_CU: load_colmajor(g.mat, 2, m); break;
      default: break;
    }
    fold_op(op, m, 2);
    o.kind = K_CDENSE1;
    for (int i = 0; i < 4; ++i) o.m[i] = m[i];
    return QSV_OK;
  }

  if (g.kind == QSV_GATE_SWAP || g.kind == QSV_GATE_ISWAP) {  // apply4bc class
    cd lam = 1.0;
    if (g.kind == QSV_GATE_ISWAP) lam = (op == QSV_OP_C) ? -I1 : I1;  // iSWAP.hpp:68-74
    o.kind = K_SWAP;
    o.nq = 2;
    o.q[0] = std::min(q0, q1);
    o.q[1] = std::max(q0, q1);
    o.m[0] = lam;
    return QSV_OK;
  }

  // 2-qubit dense / diagonal (QGate2.cpp:7-15; RotationXX/YY/ZZ.hpp:136-143)
  o.nq = 2;
  o.q[0] = std::min(q0, q1);
  o.q[1] = std::max(q0, q1);
  for (int i = 0; i < 16; ++i) m[i] = 0.0;
  switch (g.kind) {
    case QSV_GATE_RXX: {
      angle_cs(g, true, c, s);
      const cd d = c, x = cd(0, -s);
      m[0] = d; m[3] = x; m[5] = d; m[6] = x; m[9] = x; m[10] = d; m[12] = x; m[15] = d;
      break;
    }
    case QSV_GATE_RYY: {
      angle_cs(g, true, c, s);
      const cd d = c, y = cd(0, s);
      m[0] = d; m[3] = y; m[5] = d; m[6] = -y; m[9] = -y; m[10] = d; m[12] = y; m[15] = d;
      break;
    }
    case QSV_GATE_RZZ: {
      angle_cs(g, true, c, s);
      cd d[16] = {};
      d[0] = cd(c, -s); d[5] = cd(c, s); d[10] = cd(c, s); d[15] = cd(c, -s);
      fold_op(op, d, 4);
      o.kind = K_DIAG2;
      o.m[0] = d[0]; o.m[1] = d[5]; o.m[2] = d[10]; o.m[3] = d[15];
      return QSV_OK;
    }
    case QSV_GATE_U2:
      // QGate2::apply passes qubits() unsorted to masks(), which assumes
      // q0 < q1 (src/qgates/apply.hpp:241-256); only that order is defined.
      if (q0 > q1) return fail(QSV_EQUBIT, "QGate2 qubits must be ascending");
      load_colmajor(g.mat, 4, m);
      break;
    default:
      return fail(QSV_EINVAL, "unhandled gate kind " + std::to_string(g.kind));
  }
  fold_op(op, m, 4);
  // lambda_QGate2 receives (a, a|bit(q0), a|bit(q1), a|both)
  // (src/qgates/apply.hpp:289-308), so matrix position p addresses the
  // amplitude with bit(q0) = p&1 and bit(q1) = p>>1.  The engine's 2-qubit
  // kernels index by 2*bit(q0)+bit(q1); swap positions 1 and 2 to match.
  {
    const int p[4] = {0, 2, 1, 3};
    o.kind = K_DENSE2;
    for (int r = 0; r < 4; ++r)
      for (int cc = 0; cc < 4; ++cc) o.m[p[r] * 4 + p[cc]] = m[r * 4 + cc];
  }
  return QSV_OK;
}

// Minimal HBM traffic of one launch: every 128-byte line that holds a touched
// amplitude is read and written once.  A sweep that touches the amplitudes
// with some bits fixed touches 2^-f of the lines, f = the fixed bits at or
// above the line's own address bits (3 for complex128, 4 for complex64).
double algorithmic_bytes(const KOp& op, int n, int eb) {
  const double S = std::ldexp(1.0, n) * eb;
  const int line_bits = eb == 16 ? 3 : 4;
  auto above = [&](int q) { return (n - 1 - q) >= line_bits ? 1 : 0; };
  switch (op.kind) {
    case K_DENSE1: case K_DIAG1: case K_DENSE2: case K_DIAG2:
    case K_DENSEK: case K_DIAGK:
      return 2 * S;
    case K_PHASE1:                       // the |1> half of the target
      return 2 * S / (1 << above(op.q[0]));
    case K_CDENSE1:                      // the control == cs half
      return 2 * S / (1 << above(op.q[0]));
    case K_CPHASE:                       // control == cs and target == 1
      return 2 * S / (1 << (above(op.q[0]) + above(op.q[1])));
    case K_SWAP:                         // |01> and |10>: half the lines only if both bits are above
      return above(op.q[0]) && above(op.q[1]) ? S : 2 * S;
    default:
      return 0;
  }
}

}  // namespace qsv
