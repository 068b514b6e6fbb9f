// ORACLE -- test infrastructure only.
//
// C entry points around the real QCLAB++ library, compiled from the
// reference's own sources (oracle/Makefile -> oracle/_ref/libqclab_ref.so).
// It turns qsv_gate records into the reference's gate objects and runs the
// reference's own QObject::apply / QCircuit::simulate / QCircuit::matrix.
// Used to pin the restatement in qsv_oracle.cpp and as the CPU baseline.
#include <chrono>
#include <complex>
#include <cstring>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include <cstdio>
#include <sstream>
#include <string>

#include "qclab/QCircuit.hpp"
#include "qclab/io/QASMFile.hpp"
#include "qclab/qgates/CNOT.hpp"
#include "qclab/qgates/CPhase.hpp"
#include "qclab/qgates/CRotationX.hpp"
#include "qclab/qgates/CRotationY.hpp"
#include "qclab/qgates/CRotationZ.hpp"
#include "qclab/qgates/CX.hpp"
#include "qclab/qgates/CY.hpp"
#include "qclab/qgates/CZ.hpp"
#include "qclab/qgates/Hadamard.hpp"
#include "qclab/qgates/Identity.hpp"
#include "qclab/qgates/PauliX.hpp"
#include "qclab/qgates/PauliY.hpp"
#include "qclab/qgates/PauliZ.hpp"
#include "qclab/qgates/Phase.hpp"
#include "qclab/qgates/Phase45.hpp"
#include "qclab/qgates/Phase90.hpp"
#include "qclab/qgates/RotationX.hpp"
#include "qclab/qgates/RotationXX.hpp"
#include "qclab/qgates/RotationY.hpp"
#include "qclab/qgates/RotationYY.hpp"
#include "qclab/qgates/RotationZ.hpp"
#include "qclab/qgates/RotationZZ.hpp"
#include "qclab/qgates/SWAP.hpp"
#include "qclab/qgates/iSWAP.hpp"
#include "qsv.h"
#include "qasm_records.hpp"

namespace {

using qclab::Op;

Op to_op(int op) { return op == 'T' ? Op::Trans : op == 'C' ? Op::ConjTrans : Op::NoTrans; }

// Concrete dense gates over the reference's abstract QGate1 / QGate2 /
// QControlledGate2, so generic matrices go through the reference's generic
// paths (QGate1.cpp:7-13, QGate2.cpp:7-15, QControlledGate2.cpp:7-20).
template <typename T>
struct Dense1 : qclab::qgates::QGate1<T> {
  qclab::dense::SquareMatrix<T> m;
  Dense1(int q, const double* mat) : qclab::qgates::QGate1<T>(q), m(2) {
    for (int k = 0; k < 4; ++k) m.ptr()[k] = T(mat[2 * k], mat[2 * k + 1]);
  }
  bool fixed() const override { return true; }
  qclab::dense::SquareMatrix<T> matrix() const override { return m; }
  int toQASM(std::ostream&, const int = 0) const override { return -1; }
  bool equals(const qclab::QObject<T>&) const override { return false; }
};

template <typename T>
struct Dense2 : qclab::qgates::QGate2<T> {
  qclab::dense::SquareMatrix<T> m;
  int q0, q1;
  Dense2(int a, int b, const double* mat) : m(4), q0(a), q1(b) {
    for (int k = 0; k < 16; ++k) m.ptr()[k] = T(mat[2 * k], mat[2 * k + 1]);
  }
  bool fixed() const override { return true; }
  bool controlled() const override { return false; }
  int qubit() const override { return q0; }
  std::vector<int> qubits() const override { return {q0, q1}; }
  void setQubits(const int* q) override { q0 = q[0]; q1 = q[1]; }
  qclab::dense::SquareMatrix<T> matrix() const override { return m; }
  int toQASM(std::ostream&, const int = 0) const override { return -1; }
  bool equals(const qclab::QObject<T>&) const override { return false; }
};

template <typename T>
struct DenseC : qclab::qgates::QControlledGate2<T> {
  std::unique_ptr<Dense1<T>> g;
  DenseC(int c, int t, int cs, const double* mat)
      : qclab::qgates::QControlledGate2<T>(c, cs), g(std::make_unique<Dense1<T>>(t, mat)) {}
  bool fixed() const override { return true; }
  int qubit() const override { return std::min(this->control_, g->qubit()); }
  int target() const override { return g->qubit(); }
  const qclab::qgates::QGate1<T>* gate() const override { return g.get(); }
  void setTarget(const int t) override { g->setQubit(t); }
  int toQASM(std::ostream&, const int = 0) const override { return -1; }
  bool equals(const qclab::QObject<T>&) const override { return false; }
};

template <typename T>
std::unique_ptr<qclab::QObject<T>> make_gate(const qsv_gate& g) {
  using namespace qclab::qgates;
  using R = typename T::value_type;
  const int a = g.q[0], b = g.q[1], cs = g.ctrl_state;
  const R th = R(g.theta), c = R(g.cs[0]), s = R(g.cs[1]);
  const bool ex = g.has_cs != 0;
  switch (g.kind) {
    case QSV_GATE_I: return std::make_unique<Identity<T>>(a);
    case QSV_GATE_H: return std::make_unique<Hadamard<T>>(a);
    case QSV_GATE_X: return std::make_unique<PauliX<T>>(a);
    case QSV_GATE_Y: return std::make_unique<PauliY<T>>(a);
    case QSV_GATE_Z: return std::make_unique<PauliZ<T>>(a);
    case QSV_GATE_RX:
      return ex ? std::make_unique<RotationX<T>>(a, c, s) : std::make_unique<RotationX<T>>(a, th);
    case QSV_GATE_RY:
      return ex ? std::make_unique<RotationY<T>>(a, c, s) : std::make_unique<RotationY<T>>(a, th);
    case QSV_GATE_RZ:
      return ex ? std::make_unique<RotationZ<T>>(a, c, s) : std::make_unique<RotationZ<T>>(a, th);
    case QSV_GATE_P:
      return ex ? std::make_unique<Phase<T>>(a, c, s) : std::make_unique<Phase<T>>(a, th);
    case QSV_GATE_T: return std::make_unique<Phase45<T>>(a);
    case QSV_GATE_S: return std::make_unique<Phase90<T>>(a);
    case QSV_GATE_U1: return std::make_unique<Dense1<T>>(a, g.mat);
    case QSV_GATE_CX: return std::make_unique<CX<T>>(a, b, cs);
    case QSV_GATE_CY: return std::make_unique<CY<T>>(a, b, cs);
    case QSV_GATE_CZ: return std::make_unique<CZ<T>>(a, b, cs);
    case QSV_GATE_CP:
      return ex ? std::make_unique<CPhase<T>>(a, b, c, s, cs)
                : std::make_unique<CPhase<T>>(a, b, th, cs);
    case QSV_GATE_CRX:
      return ex ? std::make_unique<CRotationX<T>>(a, b, c, s, cs)
                : std::make_unique<CRotationX<T>>(a, b, th, cs);
    case QSV_GATE_CRY:
      return ex ? std::make_unique<CRotationY<T>>(a, b, c, s, cs)
                : std::make_unique<CRotationY<T>>(a, b, th, cs);
    case QSV_GATE_CRZ:
      return ex ? std::make_unique<CRotationZ<T>>(a, b, c, s, cs)
                : std::make_unique<CRotationZ<T>>(a, b, th, cs);
    case QSV_GATE_CU: return std::make_unique<DenseC<T>>(a, b, cs, g.mat);
    case QSV_GATE_SWAP: return std::make_unique<SWAP<T>>(a, b);
    case QSV_GATE_ISWAP: return std::make_unique<iSWAP<T>>(a, b);
    case QSV_GATE_RXX:
      return ex ? std::make_unique<RotationXX<T>>(a, b, c, s)
                : std::make_unique<RotationXX<T>>(a, b, th);
    case QSV_GATE_RYY:
      return ex ? std::make_unique<RotationYY<T>>(a, b, c, s)
                : std::make_unique<RotationYY<T>>(a, b, th);
    case QSV_GATE_RZZ:
      return ex ? std::make_unique<RotationZZ<T>>(a, b, c, s)
                : std::make_unique<RotationZZ<T>>(a, b, th);
    case QSV_GATE_U2: return std::make_unique<Dense2<T>>(a, b, g.mat);
    default: return nullptr;
  }
}

// The reference asserts (and, built with NDEBUG, silently misbehaves) on
// invalid qubits; reject those here instead.
int valid(int n, const qsv_gate& g) {
  if (g.kind < 0 || g.kind >= QSV_GATE_COUNT) return QSV_EINVAL;
  if (g.q[0] < 0 || g.q[0] >= n) return QSV_EQUBIT;
  if (g.kind >= QSV_GATE_CX && (g.q[1] < 0 || g.q[1] >= n || g.q[1] == g.q[0])) return QSV_EQUBIT;
  if (g.kind == QSV_GATE_U2 && g.q[0] > g.q[1]) return QSV_EQUBIT;
  return QSV_OK;
}

template <typename T>
int build_circuit(int n, const qsv_gate* gs, size_t ng, qclab::QCircuit<T>& circ) {
  for (size_t i = 0; i < ng; ++i) {
    const int st = valid(n, gs[i]);
    if (st) return st;
    auto p = make_gate<T>(gs[i]);
    if (!p) return QSV_EINVAL;
    circ.push_back(std::move(p));
  }
  return QSV_OK;
}

template <typename T>
int apply_one(int n, const qsv_gate& g, void* state) {
  const int st = valid(n, g);
  if (st) return st;
  auto gate = make_gate<T>(g);
  if (!gate) return QSV_EINVAL;
  T* s = static_cast<T*>(state);
  std::vector<T> v(s, s + (size_t(1) << n));
  gate->apply(to_op(g.op), n, v);
  std::memcpy(s, v.data(), v.size() * sizeof(T));
  return QSV_OK;
}

template <typename T>
int circuit_apply(int n, const qsv_gate* gs, size_t ng, int op, void* state, double* secs) {
  qclab::QCircuit<T> circ(n);
  const int st = build_circuit(n, gs, ng, circ);
  if (st) return st;
  T* s = static_cast<T*>(state);
  std::vector<T> v(s, s + (size_t(1) << n));
  const auto t0 = std::chrono::steady_clock::now();
  if (op == 'N')
    circ.simulate(v);
  else
    circ.apply(to_op(op), n, v);
  const auto t1 = std::chrono::steady_clock::now();
  if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();
  std::memcpy(s, v.data(), v.size() * sizeof(T));
  return QSV_OK;
}

template <typename T>
int circuit_matrix(int n, const qsv_gate* gs, size_t ng, void* out) {
  qclab::QCircuit<T> circ(n);
  const int st = build_circuit(n, gs, ng, circ);
  if (st) return st;
  const auto m = circ.matrix();
  std::memcpy(out, m.ptr(), sizeof(T) * m.size() * m.size());
  return QSV_OK;
}

}  // namespace

extern "C" {

// gate.apply(op, n, vector)  (QObject.hpp:52-53)
int ref_apply_gate(int n, int dtype, const qsv_gate* g, void* state) {
  if (!g || !state) return QSV_EINVAL;
  if (dtype == QSV_C128) return apply_one<std::complex<double>>(n, *g, state);
  if (dtype == QSV_C64) return apply_one<std::complex<float>>(n, *g, state);
  return QSV_EDTYPE;
}

// QCircuit(n) + push_back(gates) + simulate(vector) for op 'N', or
// apply(op, n, vector) otherwise (QCircuit.hpp:130-143,164-166).  `secs`
// receives the wall time of that call alone (steady_clock), as
// test/timings/run.hpp:54-56 times simulate().
int ref_circuit_apply(int n, int dtype, const qsv_gate* g, size_t ng, int op, void* state,
                      double* secs) {
  if ((!g && ng) || !state) return QSV_EINVAL;
  if (dtype == QSV_C128) return circuit_apply<std::complex<double>>(n, g, ng, op, state, secs);
  if (dtype == QSV_C64) return circuit_apply<std::complex<float>>(n, g, ng, op, state, secs);
  return QSV_EDTYPE;
}

// QCircuit::matrix() (QCircuit.hpp:121-127), column-major 2^n x 2^n.
int ref_circuit_matrix(int n, int dtype, const qsv_gate* g, size_t ng, void* out) {
  if ((!g && ng) || !out || n > 12) return QSV_EINVAL;
  if (dtype == QSV_C128) return circuit_matrix<std::complex<double>>(n, g, ng, out);
  if (dtype == QSV_C64) return circuit_matrix<std::complex<float>>(n, g, ng, out);
  return QSV_EDTYPE;
}

// QCircuit::simulate on |0...0> built inside the library (no copies), as
// test/timings/run.hpp:49-56 does; `secs` gets the simulate() wall time.
int ref_time_simulate_zero(int n, int dtype, const qsv_gate* g, size_t ng, double* secs) {
  if ((!g && ng) || !secs || n < 1 || n > 34) return QSV_EINVAL;
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    qclab::QCircuit<T> circ(n);
    const int st = build_circuit(n, g, ng, circ);
    if (st) return st;
    std::vector<T> v(size_t(1) << n, T(0));
    v[0] = 1;
    const auto t0 = std::chrono::steady_clock::now();
    circ.simulate(v);
    *secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return QSV_OK;
  };
  if (dtype == QSV_C128) return run(std::complex<double>());
  if (dtype == QSV_C64) return run(std::complex<float>());
  return QSV_EDTYPE;
}

// io::QASMFile(filename).gates<QObject<complex<double>>>()
// (include/qclab/io/QASMFile.hpp:42-152, src/io/QASMFile.cpp:16-53): the
// gates the reference reads from a file, as records (qasm_records.hpp).
// *count gets the number of gates (records beyond `cap` are dropped),
// *nqubits the register size; "Failed parsing" goes to stderr as in the
// reference.  Inputs on which the reference never returns (a last command
// without ';', an empty command between two ';') must not be given.
int ref_qasm_gates(const char* path, qsv_gate* out, size_t cap, size_t* count, int* nqubits) {
  if (!path || !count || !nqubits) return QSV_EINVAL;
  using T = std::complex<double>;
  qclab::io::QASMFile f(path);
  const auto gates = f.gates<qclab::QObject<T>>();
  *nqubits = f.nbQubits();
  *count = gates.size();
  for (size_t i = 0; i < gates.size() && i < cap; ++i)
    if (!qsv_oracle::qasm_record(*gates[i], out[i])) return QSV_EUNSUP;
  return QSV_OK;
}

// QCircuit::toQASM (include/qclab/QCircuit.hpp:211-218 with the gates'
// toQASM and include/qclab/qasm.hpp) of the records' circuit, complex<double>
// (angles written "%.15f").  *len gets the text length; the text (NUL
// terminated) is copied when it fits in `cap`.  Returns the writer's status
// (nonzero: a gate without a QASM form).
int ref_circuit_qasm(int n, const qsv_gate* g, size_t ng, char* out, size_t cap, size_t* len) {
  if ((!g && ng) || !len) return QSV_EINVAL;
  using T = std::complex<double>;
  qclab::QCircuit<T> circ(n);
  const int st = build_circuit(n, g, ng, circ);
  if (st) return st;
  std::ostringstream text;
  const int w = circ.toQASM(text);
  const std::string s = text.str();
  *len = s.size();
  if (out && cap > s.size()) std::memcpy(out, s.c_str(), s.size() + 1);
  return w == 0 ? QSV_OK : QSV_EUNSUP;
}

int ref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
