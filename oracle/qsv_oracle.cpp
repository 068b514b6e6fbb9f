// ORACLE -- test infrastructure only.
//
// CPU restatement of QCLAB++'s state-vector hot path, used by tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker.
// Nothing in the product (qclabpp_amd/, include/) may call it.
//
// It restates, function by function:
//   masks()                       src/qgates/apply.hpp:226-256
//   apply2 / apply2b              src/qgates/apply.hpp:258-287
//   apply4 (dense, controlled)    src/qgates/apply.hpp:289-332
//   apply4b / apply4bc            src/qgates/apply.hpp:334-371
//   lambda_* (incl. Op branches)  src/qgates/apply.hpp:10-223
//   operateInPlace                include/qclab/dense/transpose.hpp:12-35
//   gate -> (lambda, sweep)       src/qgates/{Hadamard,PauliX,...,iSWAP}.cpp
//   QCircuit::apply order         include/qclab/QCircuit.hpp:130-143
// with std::complex arithmetic and host libm cos/sin, as the reference.
//
// Pinned by: the reference's own known-answer vectors (tests/golden/kat_*.json,
// extracted from test/qgates/*.cpp), the reference itself compiled from its
// sources into oracle/_ref (tests/test_oracle_vs_ref.py), and closed forms
// (DFT for the QFT, brute-force kron unitaries, norm invariants).
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

#include "qsv.h"

namespace {

template <typename T>
using R_of = typename T::value_type;

// ---- operateInPlace on a small column-major matrix (transpose.hpp:12-35) ----
template <typename T>
void operate_in_place(int op, T* A, int size) {
  if (op == QSV_OP_N) return;
  if (op == QSV_OP_C)
    for (int k = 0; k < size * size; ++k) A[k] = std::conj(A[k]);
  for (int j = 0; j < size - 1; ++j)
    for (int i = j + 1; i < size; ++i) std::swap(A[i + j * size], A[j + i * size]);
}

// ---- masks (apply.hpp:226-256) -----------------------------------------------
void masks1(int n, int q, uint64_t& mL, uint64_t& mR) {
  const int nL = q, nR = n - q - 1;
  mL = ((1ULL << nL) - 1) << nR;
  mR = (1ULL << nR) - 1;
}
void masks2(int n, int q0, int q1, uint64_t& mL, uint64_t& mC, uint64_t& mR) {
  const int nL = q0, nC = q1 - q0 - 1, nR = n - q1 - 1;
  mL = ((1ULL << nL) - 1) << (nR + nC);
  mC = ((1ULL << nC) - 1) << nR;
  mR = (1ULL << nR) - 1;
}

// ---- sweeps (apply.hpp:258-371) ------------------------------------------------
template <typename F>
void apply2(int n, int q, const F& f) {
  uint64_t mL, mR;
  masks1(n, q, mL, mR);
  const uint64_t cnt = 1ULL << (n - 1), b1 = 1ULL << (n - q - 1);
#pragma omp parallel for schedule(static)
  for (int64_t kk = 0; kk < int64_t(cnt); ++kk) {
    const uint64_t k = uint64_t(kk);
    const uint64_t a = ((k & mL) << 1) | (k & mR);
    f(a, a | b1);
  }
}

template <typename F>
void apply2b(int n, int q, const F& f) {
  uint64_t mL, mR;
  masks1(n, q, mL, mR);
  const uint64_t cnt = 1ULL << (n - 1), b1 = 1ULL << (n - q - 1);
#pragma omp parallel for schedule(static)
  for (int64_t kk = 0; kk < int64_t(cnt); ++kk) {
    const uint64_t k = uint64_t(kk);
    f(((k & mL) << 1) | b1 | (k & mR));
  }
}

template <typename F>
void apply4_dense(int n, int q0, int q1, const F& f) {
  uint64_t mL, mC, mR;
  masks2(n, q0, q1, mL, mC, mR);
  const uint64_t cnt = 1ULL << (n - 2);
  const uint64_t b1 = 1ULL << (n - q0 - 1), c1 = 1ULL << (n - q1 - 1), d1 = b1 | c1;
#pragma omp parallel for schedule(static)
  for (int64_t kk = 0; kk < int64_t(cnt); ++kk) {
    const uint64_t k = uint64_t(kk);
    const uint64_t a = (k & mR) + ((k & mC) << 1) + ((k & mL) << 2);
    f(a, a | b1, a | c1, a | d1);
  }
}

template <typename F>
void apply4_ctrl(int n, int q0, int q1, int control, int target, int cstate, const F& f) {
  uint64_t mL, mC, mR;
  masks2(n, q0, q1, mL, mC, mR);
  const uint64_t cnt = 1ULL << (n - 2);
  uint64_t a1 = 0, b1 = 1ULL << (n - target - 1);
  if (cstate == 1) {
    a1 += 1ULL << (n - control - 1);
    b1 += 1ULL << (n - control - 1);
  }
#pragma omp parallel for schedule(static)
  for (int64_t kk = 0; kk < int64_t(cnt); ++kk) {
    const uint64_t k = uint64_t(kk);
    const uint64_t i = ((k & mL) << 2) | ((k & mC) << 1) | (k & mR);
    f(i | a1, i | b1);
  }
}

template <typename F>
void apply4b(int n, int q0, int q1, int control, int target, int cstate, const F& f) {
  uint64_t mL, mC, mR;
  masks2(n, q0, q1, mL, mC, mR);
  const uint64_t cnt = 1ULL << (n - 2);
  uint64_t b1 = 1ULL << (n - target - 1);
  if (cstate == 1) b1 += 1ULL << (n - control - 1);
#pragma omp parallel for schedule(static)
  for (int64_t kk = 0; kk < int64_t(cnt); ++kk) {
    const uint64_t k = uint64_t(kk);
    f(((k & mL) << 2) | ((k & mC) << 1) | (k & mR) | b1);
  }
}

template <typename F>
void apply4bc(int n, int q0, int q1, const F& f) {
  uint64_t mL, mC, mR;
  masks2(n, q0, q1, mL, mC, mR);
  const uint64_t cnt = 1ULL << (n - 2);
  const uint64_t b1 = 1ULL << (n - q0 - 1), c1 = 1ULL << (n - q1 - 1);
#pragma omp parallel for schedule(static)
  for (int64_t kk = 0; kk < int64_t(cnt); ++kk) {
    const uint64_t k = uint64_t(kk);
    const uint64_t a = ((k & mL) << 2) | ((k & mC) << 1) | (k & mR);
    f(a | b1, a | c1);
  }
}

// ---- 2x2 pair update shared by the dense 1-qubit lambdas ---------------------
template <typename T>
struct Pair2 {  // v[a] = m11 v[a] + m12 v[b]; v[b] = m21 v[a] + m22 v[b]
  T* v;
  T m11, m12, m21, m22;
  void operator()(uint64_t a, uint64_t b) const {
    const T x1 = v[a], x2 = v[b];
    v[a] = m11 * x1 + m12 * x2;
    v[b] = m21 * x1 + m22 * x2;
  }
};

// (cos, sin) as stored by QRotation (theta/2) or QAngle (theta).
template <typename R>
void cs_of(const qsv_gate& g, bool half, R& c, R& s) {
  if (g.has_cs) {
    c = R(g.cs[0]);
    s = R(g.cs[1]);
  } else {
    const R th = R(g.theta);
    const R a = half ? th / 2 : th;
    c = std::cos(a);
    s = std::sin(a);
  }
}

template <typename T>
void mat_from(const qsv_gate& g, int size, T* A) {  // column-major
  for (int k = 0; k < size * size; ++k) A[k] = T(R_of<T>(g.mat[2 * k]), R_of<T>(g.mat[2 * k + 1]));
}

// The controlled 1-qubit gate's matrix (column-major), for the generic
// QControlledGate2 path (src/qgates/QControlledGate2.cpp:7-20).
template <typename T>
void controlled_matrix(const qsv_gate& g, T* A) {
  using R = R_of<T>;
  R c, s;
  switch (g.kind) {
    case QSV_GATE_CRX:  // RotationX.hpp:104-108
      cs_of(g, true, c, s);
      A[0] = c; A[1] = T(0, -s); A[2] = T(0, -s); A[3] = c;
      break;
    case QSV_GATE_CRY:  // RotationY.hpp:104-107, column-major
      cs_of(g, true, c, s);
      A[0] = c; A[1] = s; A[2] = -s; A[3] = c;
      break;
    default:
      mat_from(g, 2, A);
  }
}

template <typename T>
int apply_gate(int n, const qsv_gate& g, T* v) {
  using R = R_of<T>;
  const int op = g.op;
  const int q = g.q[0];
  const bool two = g.kind >= QSV_GATE_CX;
  if (q < 0 || q >= n) return QSV_EQUBIT;
  if (two && (g.q[1] < 0 || g.q[1] >= n || g.q[1] == q)) return QSV_EQUBIT;
  R c, s;
  switch (g.kind) {
    case QSV_GATE_I:  // Identity.cpp:7-11
      return QSV_OK;
    case QSV_GATE_H: {  // lambda_Hadamard apply.hpp:29-42 + apply2
      const R sqrt2 = R(1) / std::sqrt(R(2));
      apply2(n, q, [=](uint64_t a, uint64_t b) {
        const T x1 = v[a], x2 = v[b];
        v[a] = sqrt2 * x1 + sqrt2 * x2;
        v[b] = sqrt2 * x1 - sqrt2 * x2;
      });
      return QSV_OK;
    }
    case QSV_GATE_X:  // lambda_PauliX apply.hpp:45-52
      apply2(n, q, [=](uint64_t a, uint64_t b) { std::swap(v[a], v[b]); });
      return QSV_OK;
    case QSV_GATE_Y: {  // lambda_PauliY apply.hpp:55-68
      T l1(0, -1), l2(0, 1);
      if (op == QSV_OP_T) std::swap(l1, l2);
      apply2(n, q, [=](uint64_t a, uint64_t b) {
        std::swap(v[a], v[b]);
        v[a] *= l1;
        v[b] *= l2;
      });
      return QSV_OK;
    }
    case QSV_GATE_Z:  // lambda_PauliZ apply.hpp:71-78 + apply2b
      apply2b(n, q, [=](uint64_t b) { v[b] = -v[b]; });
      return QSV_OK;
    case QSV_GATE_RX: {  // lambda_RotationX apply.hpp:81-95
      cs_of(g, true, c, s);
      T sx(0, -s);
      if (op == QSV_OP_C) sx = std::conj(sx);
      apply2(n, q, [=](uint64_t a, uint64_t b) {
        const T x1 = v[a], x2 = v[b];
        v[a] = c * x1 + sx * x2;
        v[b] = sx * x1 + c * x2;
      });
      return QSV_OK;
    }
    case QSV_GATE_RY: {  // lambda_RotationY apply.hpp:98-113
      cs_of(g, true, c, s);
      R s1 = -s, s2 = s;
      if (op != QSV_OP_N) std::swap(s1, s2);
      apply2(n, q, [=](uint64_t a, uint64_t b) {
        const T x1 = v[a], x2 = v[b];
        v[a] = c * x1 + s1 * x2;
        v[b] = s2 * x1 + c * x2;
      });
      return QSV_OK;
    }
    case QSV_GATE_RZ: {  // lambda_RotationZ apply.hpp:116-128, full sweep
      cs_of(g, true, c, s);
      T l1(c, -s), l2(c, s);
      if (op == QSV_OP_C) std::swap(l1, l2);
      apply2(n, q, [=](uint64_t a, uint64_t b) {
        v[a] *= l1;
        v[b] *= l2;
      });
      return QSV_OK;
    }
    case QSV_GATE_P: case QSV_GATE_T: case QSV_GATE_S: {  // apply.hpp:131-168
      T lam;
      if (g.kind == QSV_GATE_P) {
        cs_of(g, false, c, s);
        lam = T(c, s);
      } else if (g.kind == QSV_GATE_T) {
        const R sqrt2 = R(1) / std::sqrt(R(2));
        lam = T(sqrt2, sqrt2);
      } else {
        lam = T(0, 1);
      }
      if (op == QSV_OP_C) lam = std::conj(lam);
      apply2b(n, q, [=](uint64_t b) { v[b] *= lam; });
      return QSV_OK;
    }
    case QSV_GATE_U1: {  // lambda_QGate1 apply.hpp:10-26
      T A[4];
      mat_from(g, 2, A);
      operate_in_place(op, A, 2);
      apply2(n, q, Pair2<T>{v, A[0], A[2], A[1], A[3]});
      return QSV_OK;
    }
    default:
      break;
  }

  const int q1 = g.q[1];
  const int lo = std::min(q, q1), hi = std::max(q, q1);
  switch (g.kind) {
    case QSV_GATE_CX:  // CX.cpp:7-18: lambda_PauliX + apply4 controlled
      apply4_ctrl(n, lo, hi, q, q1, g.ctrl_state,
                  [=](uint64_t a, uint64_t b) { std::swap(v[a], v[b]); });
      return QSV_OK;
    case QSV_GATE_CY: {  // CY.cpp:7-18
      T l1(0, -1), l2(0, 1);
      if (op == QSV_OP_T) std::swap(l1, l2);
      apply4_ctrl(n, lo, hi, q, q1, g.ctrl_state, [=](uint64_t a, uint64_t b) {
        std::swap(v[a], v[b]);
        v[a] *= l1;
        v[b] *= l2;
      });
      return QSV_OK;
    }
    case QSV_GATE_CRZ: {  // CRotationZ.cpp:7-19: lambda_RotationZ
      cs_of(g, true, c, s);
      T l1(c, -s), l2(c, s);
      if (op == QSV_OP_C) std::swap(l1, l2);
      apply4_ctrl(n, lo, hi, q, q1, g.ctrl_state, [=](uint64_t a, uint64_t b) {
        v[a] *= l1;
        v[b] *= l2;
      });
      return QSV_OK;
    }
    case QSV_GATE_CRX: case QSV_GATE_CRY: case QSV_GATE_CU: {  // QControlledGate2.cpp:7-20
      T A[4];
      controlled_matrix(g, A);
      operate_in_place(op, A, 2);
      apply4_ctrl(n, lo, hi, q, q1, g.ctrl_state, Pair2<T>{v, A[0], A[2], A[1], A[3]});
      return QSV_OK;
    }
    case QSV_GATE_CZ:  // CZ.cpp:7-18: lambda_PauliZ + apply4b
      apply4b(n, lo, hi, q, q1, g.ctrl_state, [=](uint64_t b) { v[b] = -v[b]; });
      return QSV_OK;
    case QSV_GATE_CP: {  // CPhase.cpp:7-19: lambda_Phase + apply4b
      cs_of(g, false, c, s);
      T lam(c, s);
      if (op == QSV_OP_C) lam = std::conj(lam);
      apply4b(n, lo, hi, q, q1, g.ctrl_state, [=](uint64_t b) { v[b] *= lam; });
      return QSV_OK;
    }
    case QSV_GATE_SWAP:  // SWAP.cpp:8-16: lambda_SWAP + apply4bc
      apply4bc(n, lo, hi, [=](uint64_t b, uint64_t cc) { std::swap(v[b], v[cc]); });
      return QSV_OK;
    case QSV_GATE_ISWAP: {  // iSWAP.cpp:10-18: lambda_iSWAP
      T lam(0, 1);
      if (op == QSV_OP_C) lam = T(0, -1);
      apply4bc(n, lo, hi, [=](uint64_t b, uint64_t cc) {
        std::swap(v[b], v[cc]);
        v[b] *= lam;
        v[cc] *= lam;
      });
      return QSV_OK;
    }
    default:
      break;
  }

  // QGate2::apply (QGate2.cpp:7-15): lambda_QGate2 + apply4 dense, qubits()
  // ascending for the rotation gates (QRotationGate2.hpp setQubits).
  T A[16];
  for (auto& x : A) x = T(0);
  switch (g.kind) {
    case QSV_GATE_RXX: {  // RotationXX.hpp:136-143 (row-order arguments)
      cs_of(g, true, c, s);
      const T d = c, o(0, -s);
      A[0] = d; A[5] = d; A[10] = d; A[15] = d;
      A[3 * 4 + 0] = o; A[2 * 4 + 1] = o; A[1 * 4 + 2] = o; A[0 * 4 + 3] = o;
      break;
    }
    case QSV_GATE_RYY: {  // RotationYY.hpp:136-143
      cs_of(g, true, c, s);
      const T d = c, o(0, s);
      A[0] = d; A[5] = d; A[10] = d; A[15] = d;
      // rows: (d,0,0,o) (0,d,-o,0) (0,-o,d,0) (o,0,0,d); store column-major
      A[0 + 3 * 4] = o; A[1 + 2 * 4] = -o; A[2 + 1 * 4] = -o; A[3 + 0 * 4] = o;
      break;
    }
    case QSV_GATE_RZZ: {  // RotationZZ.hpp:136-143
      cs_of(g, true, c, s);
      const T a(c, -s), b(c, s);
      A[0] = a; A[5] = b; A[10] = b; A[15] = a;
      break;
    }
    case QSV_GATE_U2:
      if (q > q1) return QSV_EQUBIT;
      mat_from(g, 4, A);
      break;
    default:
      return QSV_EINVAL;
  }
  operate_in_place(op, A, 4);
  const T m11 = A[0], m12 = A[4], m13 = A[8], m14 = A[12];
  const T m21 = A[1], m22 = A[5], m23 = A[9], m24 = A[13];
  const T m31 = A[2], m32 = A[6], m33 = A[10], m34 = A[14];
  const T m41 = A[3], m42 = A[7], m43 = A[11], m44 = A[15];
  const int g0 = (g.kind == QSV_GATE_U2) ? q : lo, g1 = (g.kind == QSV_GATE_U2) ? q1 : hi;
  apply4_dense(n, g0, g1, [=](uint64_t a, uint64_t b, uint64_t cc, uint64_t d) {
    const T x1 = v[a], x2 = v[b], x3 = v[cc], x4 = v[d];
    v[a] = m11 * x1 + m12 * x2 + m13 * x3 + m14 * x4;
    v[b] = m21 * x1 + m22 * x2 + m23 * x3 + m24 * x4;
    v[cc] = m31 * x1 + m32 * x2 + m33 * x3 + m34 * x4;
    v[d] = m41 * x1 + m42 * x2 + m43 * x3 + m44 * x4;
  });
  return QSV_OK;
}

template <typename T>
int apply_gates(int n, const qsv_gate* g, size_t ng, T* v) {
  for (size_t i = 0; i < ng; ++i) {
    const int st = apply_gate<T>(n, g[i], v);
    if (st != QSV_OK) return st;
  }
  return QSV_OK;
}

}  // namespace

extern "C" {

int oracle_apply_gate(int n, int dtype, const qsv_gate* g, void* state) {
  if (!g || !state || n < 1 || n > 40) return QSV_EINVAL;
  if (dtype == QSV_C128) return apply_gate(n, *g, static_cast<std::complex<double>*>(state));
  if (dtype == QSV_C64) return apply_gate(n, *g, static_cast<std::complex<float>*>(state));
  return QSV_EDTYPE;
}

// QCircuit::simulate / apply(NoTrans): gates in order (QCircuit.hpp:130-136).
int oracle_apply_gates(int n, int dtype, const qsv_gate* g, size_t ng, void* state) {
  if ((!g && ng) || !state || n < 1 || n > 40) return QSV_EINVAL;
  if (dtype == QSV_C128) return apply_gates(n, g, ng, static_cast<std::complex<double>*>(state));
  if (dtype == QSV_C64) return apply_gates(n, g, ng, static_cast<std::complex<float>*>(state));
  return QSV_EDTYPE;
}

// QCircuit::apply with a [Conj]Trans op: gates in reverse, each with `op`
// (QCircuit.hpp:137-141).  The records' own op fields are overridden.
int oracle_apply_gates_op(int n, int dtype, const qsv_gate* g, size_t ng, int op, void* state) {
  if ((!g && ng) || !state || n < 1 || n > 40) return QSV_EINVAL;
  std::vector<qsv_gate> seq(g, g + ng);
  if (op != QSV_OP_N) {
    std::vector<qsv_gate> rev(seq.rbegin(), seq.rend());
    seq.swap(rev);
  }
  for (auto& x : seq) x.op = op;
  return oracle_apply_gates(n, dtype, seq.data(), seq.size(), state);
}

// ---- observables (no reference counterpart; built from its gate semantics) --
// <psi|P|psi> for each Pauli string: apply PauliX/Y/Z (lambda_PauliX/Y/Z via
// apply2/apply2b, restated above) to a copy and take the inner product with
// psi, accumulated in double.
int oracle_expectation(int n, int dtype, const void* state, const qsv_pauli_term* terms,
                       size_t nterms, double* values) {
  if (!state || (!terms && nterms) || !values) return QSV_EINVAL;
  const size_t N = size_t(1) << n;
  auto run = [&](auto* psi) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(psi)>>;
    std::vector<T> tmp(N);
    for (size_t t = 0; t < nterms; ++t) {
      std::memcpy(tmp.data(), psi, N * sizeof(T));
      for (int f = 0; f < terms[t].nfactors; ++f) {
        qsv_gate g{};
        g.op = QSV_OP_N;
        g.q[0] = terms[t].qubits[f];
        const char p = terms[t].paulis[f];
        if (p == 'I') continue;
        g.kind = p == 'X' ? QSV_GATE_X : p == 'Y' ? QSV_GATE_Y : p == 'Z' ? QSV_GATE_Z : -1;
        if (g.kind < 0) return QSV_EINVAL;
        const int st = apply_gate(n, g, tmp.data());
        if (st) return st;
      }
      double re = 0;
#pragma omp parallel for reduction(+ : re) schedule(static)
      for (int64_t x = 0; x < int64_t(N); ++x) {
        const std::complex<double> a(psi[x]), b(tmp[x]);
        re += (std::conj(a) * b).real();
      }
      values[t] = re;
    }
    return QSV_OK;
  };
  if (dtype == QSV_C128) return run(static_cast<const std::complex<double>*>(state));
  if (dtype == QSV_C64) return run(static_cast<const std::complex<float>*>(state));
  return QSV_EDTYPE;
}

// E(theta) = <psi|H|psi>, psi = U(theta)|0...0>, parameters overriding the
// gates' angles (gate.param indexes theta).
static int energy_at(int n, const qsv_gate* g, size_t ng, const double* theta,
                     const qsv_pauli_term* terms, size_t nterms, double* e) {
  std::vector<qsv_gate> gs(g, g + ng);
  for (auto& x : gs)
    if (x.param >= 0) {
      x.theta = theta[x.param];
      x.has_cs = 0;
    }
  std::vector<std::complex<double>> psi(size_t(1) << n);
  psi[0] = 1;
  int st = apply_gates(n, gs.data(), gs.size(), psi.data());
  if (st) return st;
  std::vector<double> v(nterms);
  st = oracle_expectation(n, QSV_C128, psi.data(), terms, nterms, v.data());
  if (st) return st;
  double s = 0;
  for (size_t t = 0; t < nterms; ++t) s += terms[t].coeff * v[t];
  *e = s;
  return QSV_OK;
}

int oracle_energy(int n, const qsv_gate* g, size_t ng, const double* theta,
                  const qsv_pauli_term* terms, size_t nterms, double* e) {
  if (!e || (!g && ng)) return QSV_EINVAL;
  return energy_at(n, g, ng, theta, terms, nterms, e);
}

// Parameter-shift gradient (exact for gates exp(-i theta P/2), P^2 = I, and
// for Phase): dE/dtheta_p = [E(theta + pi/2 e_p) - E(theta - pi/2 e_p)] / 2,
// for the parameters listed in `which`.
int oracle_param_shift(int n, const qsv_gate* g, size_t ng, const double* theta, size_t nparams,
                       const qsv_pauli_term* terms, size_t nterms, const int* which,
                       size_t nwhich, double* grad) {
  if (!theta || !grad || (!g && ng) || (!which && nwhich)) return QSV_EINVAL;
  std::vector<double> th(theta, theta + nparams);
  const double h = 3.14159265358979323846 / 2;
  for (size_t i = 0; i < nwhich; ++i) {
    const int p = which[i];
    if (p < 0 || size_t(p) >= nparams) return QSV_EINVAL;
    double ep = 0, em = 0;
    th[p] = theta[p] + h;
    int st = energy_at(n, g, ng, th.data(), terms, nterms, &ep);
    if (st) return st;
    th[p] = theta[p] - h;
    st = energy_at(n, g, ng, th.data(), terms, nterms, &em);
    if (st) return st;
    th[p] = theta[p];
    grad[i] = (ep - em) / 2;
  }
  return QSV_OK;
}

}  // extern "C"
