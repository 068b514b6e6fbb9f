// Tests of the QCLAB++-compatible C++ classes (include/qclab) running on the
// MI355X engine.  Each gate class is applied with Op N/T/C at several qubit
// placements and offsets, and the result is compared with the dense product
// op(G) * v built from the class's own matrix() on the host.  Circuits add
// nesting, pointer gates, [Conj]Trans reversal, matrix() columns, device
// vectors and QASM round trips (mirroring test/QCircuit.cpp and
// test/qgates/*.cpp of the reference).
//
//   dropin_test          every check (needs a GPU)
//   dropin_test --host   only the checks that do not touch the GPU
//   dropin_test --qasm-dump FILE ref|ext
//                        the gates io::QASMFile reads from FILE in the given
//                        dialect, one record per line (no GPU)
//   dropin_test --qasm-rewrite FILE
//                        FILE read in the extended dialect and written back
//                        with QCircuit::toQASM (no GPU)
//   dropin_test --time-brickwork N DEPTH SEED
//                        QCircuit<complex<double>>::simulate of the seeded
//                        brickwork on a std::vector from |0...0>, twice,
//                        with the library's default form; prints JSON
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <sstream>
#include <string>

#include "qclab/QCircuit.hpp"
#include "qclab/io/QASMFile.hpp"
#include "qclab/qgates/CPhase.hpp"
#include "qclab/qgates/CRotationX.hpp"
#include "qclab/qgates/CRotationY.hpp"
#include "qclab/qgates/CRotationZ.hpp"
#include "qclab/qgates/Identity.hpp"
#include "qclab/qgates/Phase45.hpp"
#include "qclab/qgates/Phase90.hpp"
#include "qclab/qgates/RotationXX.hpp"
#include "qclab/qgates/RotationYY.hpp"
#include "qclab/qgates/RotationZZ.hpp"
#include "qclab/qgates/iSWAP.hpp"
#include "qasm_records.hpp"
#include "qclab/qgates/CNOT.hpp"
#include "qclab/qgates/PointerGate1.hpp"
#include "qclab/qgates/PointerGate2.hpp"
#include "qclab/qgates/QGate2.hpp"

using namespace qclab;
using namespace qclab::qgates;

static int g_checks = 0, g_fail = 0;
#define CHECK(cond, ...)                                              \
  do {                                                                \
    ++g_checks;                                                       \
    if (!(cond)) {                                                    \
      ++g_fail;                                                       \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                \
      std::printf(__VA_ARGS__);                                       \
      std::printf("\n");                                              \
    }                                                                 \
  } while (0)

template <typename T>
double tol() { return std::is_same_v<T, std::complex<double>> ? 1e-12 : 2e-5; }

template <typename T>
std::vector<T> random_state(int n, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd;
  std::vector<T> v(size_t(1) << n);
  double s = 0;
  for (auto& x : v) { x = T(nd(rng), nd(rng)); s += std::norm(std::complex<double>(x)); }
  for (auto& x : v) x /= real_t<T>(std::sqrt(s));
  return v;
}

template <typename T>
double max_diff(const std::vector<T>& a, const std::vector<T>& b) {
  double m = 0;
  for (size_t i = 0; i < a.size(); ++i) m = std::max(m, double(std::abs(a[i] - b[i])));
  return m;
}

// op(G) * v on the host from the object's own matrix
template <typename T>
std::vector<T> host_apply(const QObject<T>& g, Op op, int n, const std::vector<T>& v, int offset) {
  dense::SquareMatrix<T> M = g.matrix_in(n, offset);
  dense::operateInPlace(op, M);
  std::vector<T> out(v.size(), T(0));
  for (int64_t c = 0; c < M.cols(); ++c)
    for (int64_t r = 0; r < M.rows(); ++r) out[r] += M(r, c) * v[c];
  return out;
}

template <typename T>
std::vector<std::unique_ptr<QObject<T>>> every_gate(int a, int b) {
  using R = real_t<T>;
  std::vector<std::unique_ptr<QObject<T>>> g;
  const R th = R(0.7311), c = std::cos(R(0.4)), s = std::sin(R(0.4));
  g.push_back(std::make_unique<Identity<T>>(a));
  g.push_back(std::make_unique<Hadamard<T>>(a));
  g.push_back(std::make_unique<PauliX<T>>(a));
  g.push_back(std::make_unique<PauliY<T>>(a));
  g.push_back(std::make_unique<PauliZ<T>>(a));
  g.push_back(std::make_unique<Phase45<T>>(a));
  g.push_back(std::make_unique<Phase90<T>>(a));
  g.push_back(std::make_unique<RotationX<T>>(a, th));
  g.push_back(std::make_unique<RotationY<T>>(a, c, s));
  g.push_back(std::make_unique<RotationZ<T>>(a, -th));
  g.push_back(std::make_unique<Phase<T>>(a, th));
  g.push_back(std::make_unique<Phase<T>>(a, c, s));
  for (int cs = 0; cs < 2; ++cs) {
    g.push_back(std::make_unique<CX<T>>(a, b, cs));
    g.push_back(std::make_unique<CNOT<T>>(b, a, cs));
    g.push_back(std::make_unique<CY<T>>(a, b, cs));
    g.push_back(std::make_unique<CZ<T>>(b, a, cs));
    g.push_back(std::make_unique<CPhase<T>>(a, b, th, cs));
    g.push_back(std::make_unique<CRotationX<T>>(b, a, th, cs));
    g.push_back(std::make_unique<CRotationY<T>>(a, b, c, s, cs));
    g.push_back(std::make_unique<CRotationZ<T>>(b, a, -th, cs));
  }
  g.push_back(std::make_unique<SWAP<T>>(a, b));
  g.push_back(std::make_unique<iSWAP<T>>(b, a));
  g.push_back(std::make_unique<RotationXX<T>>(a, b, th));
  g.push_back(std::make_unique<RotationYY<T>>(b, a, th));
  g.push_back(std::make_unique<RotationZZ<T>>(a, b, c, s));
  return g;
}

// a user gate defined only by its matrix (dense 2x2 path)
template <typename T>
class MyGate1 : public QGate1<T> {
 public:
  MyGate1(int q) : QGate1<T>(q) {}
  bool fixed() const override { return true; }
  dense::SquareMatrix<T> matrix() const override {
    return dense::SquareMatrix<T>(T(0.6, 0.0), T(0.0, 0.8), T(0.0, 0.8), T(0.6, 0.0));
  }
  int toQASM(std::ostream&, const int = 0) const override { return -1; }

 protected:
  bool equals(const QObject<T>& o) const override { return dynamic_cast<const MyGate1<T>*>(&o); }
};

// a user 2-qubit gate with a matrix that is symmetric under exchanging its
// qubits (where QGate2's vector convention and the kron order agree)
template <typename T>
class MyGate2 : public QGate2<T> {
 public:
  MyGate2(int a, int b) : q_{std::min(a, b), std::max(a, b)} {}
  bool fixed() const override { return true; }
  bool controlled() const override { return false; }
  int qubit() const override { return q_[0]; }
  std::vector<int> qubits() const override { return {q_[0], q_[1]}; }
  void setQubits(const int* q) override { q_[0] = q[0]; q_[1] = q[1]; }
  dense::SquareMatrix<T> matrix() const override {
    const T a(0.5, 0.5), b(0.5, -0.5), o(0);
    return dense::SquareMatrix<T>(a, o, o, b, o, a, b, o, o, b, a, o, b, o, o, a);
  }
  int toQASM(std::ostream&, const int = 0) const override { return -1; }

 protected:
  bool equals(const QObject<T>& o) const override { return dynamic_cast<const MyGate2<T>*>(&o); }
  int q_[2];
};

// a user 2-qubit gate whose matrix changes under exchanging its qubits:
// matrix() applies it in kron order, simulate() in QGate2's vector basis
template <typename T>
class MyGate2A : public QGate2<T> {
 public:
  MyGate2A(int a, int b) : q_{std::min(a, b), std::max(a, b)} {}
  bool fixed() const override { return true; }
  bool controlled() const override { return false; }
  int qubit() const override { return q_[0]; }
  std::vector<int> qubits() const override { return {q_[0], q_[1]}; }
  void setQubits(const int* q) override { q_[0] = q[0]; q_[1] = q[1]; }
  dense::SquareMatrix<T> matrix() const override {
    dense::SquareMatrix<T> m(4, T(0));
    m(0, 0) = T(1);
    m(1, 3) = T(1);
    m(2, 1) = T(0, 1);
    m(3, 2) = T(0.6, 0.8);
    return m;
  }
  int toQASM(std::ostream&, const int = 0) const override { return -1; }

 protected:
  bool equals(const QObject<T>& o) const override { return dynamic_cast<const MyGate2A<T>*>(&o); }
  int q_[2];
};

template <typename T>
double mat_diff(const dense::SquareMatrix<T>& a, const dense::SquareMatrix<T>& b) {
  double m = 0;
  for (int64_t c = 0; c < a.cols(); ++c)
    for (int64_t r = 0; r < a.rows(); ++r) m = std::max(m, double(std::abs(a(r, c) - b(r, c))));
  return m;
}

template <typename T>
void gate_checks(const char* tn) {
  const int n = 6;
  const Op ops[3] = {Op::NoTrans, Op::Trans, Op::ConjTrans};
  const int placements[][3] = {{0, 1, 0}, {2, 5, 0}, {4, 1, 0}, {0, 3, 2}, {3, 0, 1}};
  unsigned seed = 1;
  for (const auto& pl : placements) {
    auto gates = every_gate<T>(pl[0], pl[1]);
    gates.push_back(std::make_unique<MyGate1<T>>(pl[0]));
    gates.push_back(std::make_unique<MyGate2<T>>(pl[0], pl[1]));
    for (const auto& g : gates)
      for (Op op : ops) {
        auto v = random_state<T>(n, seed++);
        const auto want = host_apply(*g, op, n, v, pl[2]);
        g->apply(op, n, v, pl[2]);
        const double d = max_diff(v, want);
        CHECK(d < tol<T>(), "%s gate qubits %d,%d offset %d op %c: max diff %.3e", tn, pl[0], pl[1],
              pl[2], char(op), d);
      }
  }
}

template <typename T>
QCircuit<T> sample_circuit(int n, unsigned seed) {
  using R = real_t<T>;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> ang(0, 2 * M_PI);
  QCircuit<T> c(n);
  for (int layer = 0; layer < 4; ++layer) {
    for (int q = 0; q < n; ++q) {
      const int k = int(rng() % 3);
      if (k == 0) c.push_back(std::make_unique<RotationX<T>>(q, R(ang(rng))));
      if (k == 1) c.push_back(std::make_unique<RotationY<T>>(q, R(ang(rng))));
      if (k == 2) c.push_back(std::make_unique<Phase<T>>(q, R(ang(rng))));
    }
    for (int q = layer % 2; q + 1 < n; q += 2) c.push_back(std::make_unique<CX<T>>(q, q + 1));
    c.push_back(std::make_unique<RotationZZ<T>>(0, n - 1, R(ang(rng))));
    c.push_back(std::make_unique<CPhase<T>>(n - 1, 1, R(ang(rng)), 0));
    c.push_back(std::make_unique<iSWAP<T>>(1, 2));
  }
  return c;
}

template <typename T>
void circuit_checks(const char* tn) {
  const int n = 7;
  // whole circuit against its own matrix; ConjTrans undoes it
  QCircuit<T> c = sample_circuit<T>(n, 11);
  const auto v0 = random_state<T>(n, 99);
  auto v = v0;
  c.simulate(v);
  const auto want = host_apply<T>(c, Op::NoTrans, n, v0, 0);
  CHECK(max_diff(v, want) < 10 * tol<T>(), "%s circuit simulate vs matrix: %.3e", tn, max_diff(v, want));
  c.apply(Op::ConjTrans, n, v);
  CHECK(max_diff(v, v0) < 10 * tol<T>(), "%s ConjTrans round trip: %.3e", tn, max_diff(v, v0));
  auto vt = v0;
  c.apply(Op::Trans, n, vt);
  const auto want_t = host_apply<T>(c, Op::Trans, n, v0, 0);
  CHECK(max_diff(vt, want_t) < 10 * tol<T>(), "%s Trans: %.3e", tn, max_diff(vt, want_t));

  // nested circuit with an offset inside a larger register, plus pointer gates
  QCircuit<T> outer(n + 2);
  auto inner = std::make_unique<QCircuit<T>>(3, 2);
  inner->push_back(std::make_unique<Hadamard<T>>(0));
  inner->push_back(std::make_unique<CX<T>>(0, 2));
  inner->push_back(std::make_unique<RotationY<T>>(1, real_t<T>(0.3)));
  outer.push_back(std::move(inner));
  Hadamard<T> h(1);
  RotationXX<T> rxx(0, 2, real_t<T>(1.1));
  outer.push_back(std::make_unique<PointerGate1<T>>(&h, 3));
  outer.push_back(std::make_unique<PointerGate2<T>>(&rxx, 4));
  outer.push_back(std::make_unique<CZ<T>>(8, 0));
  const auto w0 = random_state<T>(n + 2, 5);
  for (int off = 0; off < 2; ++off) {
    auto w = random_state<T>(n + 3, 6);
    const auto w_in = w;
    outer.apply(Op::NoTrans, n + 3, w, off);
    const auto want_w = host_apply<T>(outer, Op::NoTrans, n + 3, w_in, off);
    CHECK(max_diff(w, want_w) < 10 * tol<T>(), "%s nested/pointer offset %d: %.3e", tn, off,
          max_diff(w, want_w));
  }
  // matrix columns: simulate(e_j) is column j of matrix()
  QCircuit<T> small = sample_circuit<T>(4, 3);
  const auto M = small.matrix();
  double worst = 0;
  for (int j = 0; j < 16; ++j) {
    std::vector<T> e(16, T(0));
    e[j] = T(1);
    small.simulate(e);
    for (int r = 0; r < 16; ++r) worst = std::max(worst, double(std::abs(e[r] - M(r, j))));
  }
  CHECK(worst < 10 * tol<T>(), "%s simulate(e_j) vs matrix(): %.3e", tn, worst);

  // matrix() on the GPU (the identity as a 2k-qubit state) against the
  // host construction, in every compiled form
  {
    QCircuit<T> m7 = sample_circuit<T>(n, 13);
    m7.push_back(std::make_unique<MyGate1<T>>(3));
    m7.push_back(std::make_unique<MyGate2A<T>>(2, 3));
    m7.push_back(std::make_unique<MyGate2A<T>>(4, 6));
    m7.push_back(std::make_unique<CRotationY<T>>(6, 0, real_t<T>(0.7), 0));
    const auto Mh = m7.matrix_host();
    for (const char* form : {"0", "5", "tiled", "auto"}) {
      setenv("QCLAB_FUSE_K", form, 1);
      const double d = mat_diff(m7.matrix_gpu(), Mh);
      CHECK(d < 10 * tol<T>(), "%s matrix_gpu vs matrix_host (form %s): %.3e", tn, form, d);
    }
    unsetenv("QCLAB_FUSE_K");
    // nested circuit with an offset and pointer gates; matrix() ignores the
    // top circuit's own offset
    const double dn = mat_diff(outer.matrix_gpu(), outer.matrix_host());
    CHECK(dn < 10 * tol<T>(), "%s nested matrix_gpu vs matrix_host: %.3e", tn, dn);
    QCircuit<T> shifted(4, 3);
    shifted.push_back(std::make_unique<Hadamard<T>>(0));
    shifted.push_back(std::make_unique<CX<T>>(0, 3));
    shifted.push_back(std::make_unique<MyGate2A<T>>(1, 2));
    const double ds = mat_diff(shifted.matrix_gpu(), shifted.matrix_host());
    CHECK(ds < 10 * tol<T>(), "%s offset circuit matrix_gpu vs matrix_host: %.3e", tn, ds);
    // nrmF: both unitaries and the reduction on the GPU, against the host
    QCircuit<T> m7b = sample_circuit<T>(n, 13);
    m7b.push_back(std::make_unique<MyGate1<T>>(3));
    m7b.push_back(std::make_unique<MyGate2A<T>>(2, 3));
    m7b.push_back(std::make_unique<MyGate2A<T>>(4, 6));
    m7b.push_back(std::make_unique<CRotationY<T>>(6, 0, real_t<T>(0.75), 0));
    const double fg = nrmF(m7, m7b), fz = nrmF(m7, m7);
    setenv("QCLAB_MATRIX_GPU_MIN", "99", 1);
    const double fh = nrmF(m7, m7b);
    CHECK(m7 == m7 && m7 != m7b, "%s circuit equality by host matrix()", tn);
    unsetenv("QCLAB_MATRIX_GPU_MIN");
    CHECK(m7 == m7 && m7 != m7b, "%s circuit equality by GPU matrix()", tn);
    const double rel = std::is_same_v<T, std::complex<double>> ? 1e-10 : 1e-3;
    CHECK(fz == 0.0 && fg > 0 && std::abs(fg - fh) < rel * fh,
          "%s nrmF GPU %.17g host %.17g self %.3g", tn, fg, fh, fz);
  }

  // device-resident vector
  auto host = v0;
  c.simulate(host);
  T* d = nullptr;
  const size_t bytes = v0.size() * sizeof(T);
  CHECK(hipMalloc(&d, bytes) == hipSuccess, "hipMalloc");
  CHECK(hipMemcpy(d, v0.data(), bytes, hipMemcpyHostToDevice) == hipSuccess, "upload");
  c.simulate_device(d);
  std::vector<T> back(v0.size());
  CHECK(hipMemcpy(back.data(), d, bytes, hipMemcpyDeviceToHost) == hipSuccess, "download");
  (void)hipFree(d);
  CHECK(max_diff(back, host) < tol<T>(), "%s simulate_device vs simulate: %.3e", tn, max_diff(back, host));

  // the compiled forms agree: plain sweeps, fused blocks and tiled passes
  {
    const int nb = 14;
    QCircuit<T> big = sample_circuit<T>(nb, 21);
    const auto s0 = random_state<T>(nb, 31);
    std::vector<std::vector<T>> res;
    for (const char* form : {"0", "5", "tiled"}) {
      setenv("QCLAB_FUSE_K", form, 1);
      auto x = s0;
      big.simulate(x);
      res.push_back(x);
    }
    unsetenv("QCLAB_FUSE_K");
    CHECK(max_diff(res[0], res[1]) < 10 * tol<T>(), "%s plain vs fused: %.3e", tn, max_diff(res[0], res[1]));
    CHECK(max_diff(res[0], res[2]) < 10 * tol<T>(), "%s plain vs tiled: %.3e", tn, max_diff(res[0], res[2]));
  }

  // the default form from 22 qubits up is the compiled tiled passes; a
  // second simulate of the same structure with other angles reuses the
  // schedule and the compiled kernels (no hiprtc compile) and matches the
  // fused form
  {
    const int nb = 22;
    auto brick = [&](double phase) {
      QCircuit<T> c(nb);
      for (int layer = 0; layer < 6; ++layer) {
        for (int qb = 0; qb < nb; ++qb) {
          const real_t<T> th = real_t<T>(0.3 + 0.11 * qb + 0.7 * layer + phase);
          if ((qb + layer) % 3 == 0) c.push_back(std::make_unique<RotationX<T>>(qb, th));
          else if ((qb + layer) % 3 == 1) c.push_back(std::make_unique<RotationY<T>>(qb, th));
          else c.push_back(std::make_unique<RotationZ<T>>(qb, th));
        }
        for (int qb = layer % 2; qb + 1 < nb; qb += 2) c.push_back(std::make_unique<CX<T>>(qb, qb + 1));
      }
      return c;
    };
    unsetenv("QCLAB_FUSE_K");
    qsv_cache_stats s0{}, s1{}, s2{};
    CHECK(qsv_cache_stats_get(&s0) == QSV_OK, "cache stats");
    auto a = random_state<T>(nb, 41), b = a;
    QCircuit<T> c1 = brick(0.0);
    c1.simulate(a);
    CHECK(qsv_cache_stats_get(&s1) == QSV_OK, "cache stats");
    CHECK(s1.schedule_hits + s1.schedule_misses > s0.schedule_hits + s0.schedule_misses,
          "%s default form at n=22 builds tiled passes", tn);
    CHECK(s1.module_compiles + s1.module_hits + s1.module_disk_hits >
              s0.module_compiles + s0.module_hits + s0.module_disk_hits,
          "%s default form at n=22 runs compiled passes", tn);
    QCircuit<T> c2 = brick(0.37);
    auto a2 = a;
    c2.simulate(a2);
    CHECK(qsv_cache_stats_get(&s2) == QSV_OK, "cache stats");
    CHECK(s2.schedule_hits == s1.schedule_hits + 1 && s2.schedule_misses == s1.schedule_misses,
          "%s second simulate of one structure reuses the schedule (hits %lld -> %lld)", tn,
          (long long)s1.schedule_hits, (long long)s2.schedule_hits);
    CHECK(s2.module_compiles == s1.module_compiles && s2.module_hits > s1.module_hits,
          "%s second simulate of one structure reuses the compiled passes", tn);
    setenv("QCLAB_FUSE_K", "5", 1);
    c1.simulate(b);
    unsetenv("QCLAB_FUSE_K");
    double scale = 0;
    for (const T& x : b) scale = std::max(scale, double(std::abs(x)));
    CHECK(max_diff(a, b) < tol<T>() * scale, "%s compiled passes vs fused at n=22: %.3e", tn, max_diff(a, b));
  }

  // QASM round trip through the extended reader (no Phase: it is written as rz)
  QCircuit<T> q(5);
  q.push_back(std::make_unique<Hadamard<T>>(0));
  q.push_back(std::make_unique<RotationX<T>>(1, real_t<T>(0.25)));
  q.push_back(std::make_unique<CPhase<T>>(0, 3, real_t<T>(-1.5)));
  q.push_back(std::make_unique<CRotationY<T>>(4, 2, real_t<T>(2.0), 0));
  q.push_back(std::make_unique<RotationZZ<T>>(1, 4, real_t<T>(0.5)));
  q.push_back(std::make_unique<iSWAP<T>>(2, 3));
  q.push_back(std::make_unique<Phase45<T>>(4));
  q.push_back(std::make_unique<Phase90<T>>(3));
  q.push_back(std::make_unique<SWAP<T>>(0, 4));
  std::ostringstream text;
  text << "OPENQASM 2.0;\ninclude \"qelib1.inc\";\nqreg q[5];\n";
  CHECK(q.toQASM(text) == 0, "toQASM");
  auto f = io::QASMFile::from_string(text.str(), io::QASMFile::Dialect::Extended);
  CHECK(f.error().empty() && f.nbQubits() == 5, "extended parse: %s", f.error().c_str());
  QCircuit<T> q2(5);
  for (auto& g : f.gates<QObject<T>>()) q2.push_back(std::move(g));
  // the gate controlled on |0> is written as x, cry, x
  CHECK(q2.nbGates() == q.nbGates() + 2, "%s QASM round trip gate count %zu", tn, q2.nbGates());
  auto a = random_state<T>(5, 8), b = a;
  q.simulate(a);
  q2.simulate(b);
  // angles are written "%.15f" for double and "%.7f" for float
  const double qtol = std::is_same_v<T, std::complex<double>> ? 1e-13 : 1e-5;
  CHECK(max_diff(a, b) < qtol, "%s QASM round trip state: %.3e", tn, max_diff(a, b));
}

// checks with no GPU work: headers, QASM readers, matrices
template <typename T>
void host_checks(const char* tn) {
  // the reference dialect: numeric angles read, `pi` stops the list
  const std::string text =
      "OPENQASM 2.0;\ninclude \"qelib1.inc\";\nqreg q[4];\ncreg c[4];\n"
      "h q[0];\nrx(0.5) q[1];\ncx q[1], q[3];\n// a comment\nswap q[0],q[2]; p(1.25) q[2];\n"
      "s q[1];\nrz(pi/2) q[0];\nx q[3];\n";
  std::streambuf* old = std::cerr.rdbuf();
  std::ostringstream err;
  std::cerr.rdbuf(err.rdbuf());
  auto ref = io::QASMFile::from_string(text);
  auto gates = ref.gates<QObject<T>>();
  std::cerr.rdbuf(old);
  CHECK(ref.nbQubits() == 4, "%s reference dialect qubits", tn);
  CHECK(gates.size() == 5, "%s reference dialect keeps the 5 gates before rz(pi/2), got %zu", tn, gates.size());
  CHECK(err.str().find("Failed parsing QASM file!") != std::string::npos, "%s truncation message", tn);
  if (gates.size() == 5) {
    CHECK(*gates[1] == RotationX<T>(1, real_t<T>(0.5)), "%s rx parsed", tn);
    CHECK(dynamic_cast<const Phase<T>*>(gates[4].get()) != nullptr, "%s p parsed", tn);
  }
  // the extended dialect reads all of it, `pi` included
  std::cerr.rdbuf(err.rdbuf());
  auto ext = io::QASMFile::from_string(text, io::QASMFile::Dialect::Extended);
  std::cerr.rdbuf(old);
  auto eg = ext.gates<QObject<T>>();
  CHECK(eg.size() == 8 && ext.error().empty(), "%s extended dialect: %zu gates, '%s'", tn, eg.size(),
        ext.error().c_str());
  if (eg.size() == 8) {
    const auto* rz = dynamic_cast<const RotationZ<T>*>(eg[6].get());
    CHECK(rz && std::abs(double(rz->theta()) - M_PI / 2) < 1e-6, "%s rz(pi/2)", tn);
  }
  auto bc = io::QASMFile::from_string("qreg a[2]; qreg b[3]; h b; cx a[1], b[0]; rzz(-pi/4*2) a[0],b[2];",
                                      io::QASMFile::Dialect::Extended);
  CHECK(bc.nbQubits() == 5 && bc.gates<QObject<T>>().size() == 5, "%s broadcast/multi-register", tn);

  // toQASM formats (include/qclab/qasm.hpp of the reference)
  std::ostringstream s1;
  CX<T>(2, 0, 0).toQASM(s1, 1);
  CHECK(s1.str() == "x q[3];\ncx q[3], q[1];\nx q[3];\n", "%s CX controlState 0 QASM: '%s'", tn, s1.str().c_str());
  std::ostringstream s2;
  Phase<T>(1, real_t<T>(0.5)).toQASM(s2);
  CHECK(s2.str().rfind("rz(0.5", 0) == 0, "%s Phase written as rz: '%s'", tn, s2.str().c_str());

  // matrices and equality semantics
  CHECK(CX<T>(0, 1).matrix() == dense::SquareMatrix<T>(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0),
        "%s CX(0,1) matrix", tn);
  CHECK(CX<T>(1, 0).matrix() == dense::SquareMatrix<T>(1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0),
        "%s CX(1,0) matrix", tn);
  CHECK(CX<T>(0, 1) == CX<T>(2, 5) && CX<T>(0, 1) != CX<T>(1, 0), "%s CX equality", tn);
  RotationX<T> r1(0, real_t<T>(0.3)), r2(0, real_t<T>(0.4));
  CHECK(std::abs(double((r1 * r2).theta()) - 0.7) < 1e-6, "%s rotation product", tn);
  QCircuit<T> c1(2), c2(2);
  c1.push_back(std::make_unique<CZ<T>>(0, 1));
  c2.push_back(std::make_unique<CZ<T>>(1, 0));
  CHECK(c1 == c2 && nrmF(c1, c2) == 0.0, "%s circuit equality by matrix", tn);
  CHECK(!c1.canInsert(Hadamard<T>(2)) && c1.canInsert(Hadamard<T>(1)), "%s canInsert", tn);
}

// --qasm-dump: "nq N" then "kind q0 q1 ctrl_state theta" per gate
int qasm_dump(const char* path, const char* dialect) {
  using T = std::complex<double>;
  const auto d = std::strcmp(dialect, "ext") == 0 ? io::QASMFile::Dialect::Extended
                                                  : io::QASMFile::Dialect::Reference;
  io::QASMFile f(path, d);
  const auto gates = f.gates<QObject<T>>();
  std::printf("nq %d\n", f.nbQubits());
  for (const auto& g : gates) {
    qsv_gate r;
    if (!qsv_oracle::qasm_record(*g, r)) {
      std::printf("unknown gate\n");
      return 1;
    }
    std::printf("%d %d %d %d %.17g\n", r.kind, r.q[0], r.q[1], r.ctrl_state, r.theta);
  }
  return 0;
}

// --qasm-rewrite: read (extended dialect), write back with QCircuit::toQASM
int qasm_rewrite(const char* path) {
  using T = std::complex<double>;
  io::QASMFile f(path, io::QASMFile::Dialect::Extended);
  if (!f.error().empty()) {
    std::fprintf(stderr, "%s\n", f.error().c_str());
    return 1;
  }
  QCircuit<T> c(f.nbQubits());
  for (auto& g : f.gates<QObject<T>>()) c.push_back(std::move(g));
  std::ostringstream text;
  if (c.toQASM(text) != 0) return 1;
  std::fputs(text.str().c_str(), stdout);
  return 0;
}

// The brickwork of the benchmark built through the drop-in classes: per
// layer a rotation on every qubit (axis then angle from SplitMix64, as
// qclabpp_amd/workloads.py draws them), then CNOT(i, i+1) for i of the
// layer's parity.
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

int time_brickwork(int n, int depth, uint64_t seed) {
  using T = std::complex<double>;
  SplitMix64 rng{seed};
  QCircuit<T> c(n);
  for (int layer = 0; layer < depth; ++layer) {
    for (int q = 0; q < n; ++q) {
      const int axis = std::min(int(rng.uniform() * 3), 2);
      const double th = rng.uniform() * 2 * M_PI;
      if (axis == 0) c.push_back(std::make_unique<RotationX<T>>(q, th));
      else if (axis == 1) c.push_back(std::make_unique<RotationY<T>>(q, th));
      else c.push_back(std::make_unique<RotationZ<T>>(q, th));
    }
    for (int i = layer % 2; i + 1 < n; i += 2) c.push_back(std::make_unique<CNOT<T>>(i, i + 1));
  }
  std::vector<T> v(size_t(1) << n);
  double secs[2];
  for (int rep = 0; rep < 2; ++rep) {
    std::fill(v.begin(), v.end(), T(0));
    v[0] = 1;
    const auto t0 = std::chrono::steady_clock::now();
    c.simulate(v);
    secs[rep] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  double norm = 0;
  for (const T& x : v) norm += std::norm(x);
  qsv_cache_stats st{};
  (void)qsv_cache_stats_get(&st);
  std::printf("{\"n\": %d, \"gates\": %zu, \"first_call_s\": %.6f, \"s\": %.6f, \"norm\": %.15f, "
              "\"schedule_hits\": %lld, \"module_compiles\": %lld, \"module_hits\": %lld}\n",
              n, c.nbGates(), secs[0], secs[1], norm, (long long)st.schedule_hits,
              (long long)st.module_compiles, (long long)st.module_hits);
  return std::abs(norm - 1.0) < 1e-9 ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc > 4 && std::strcmp(argv[1], "--time-brickwork") == 0)
    return time_brickwork(std::atoi(argv[2]), std::atoi(argv[3]), std::strtoull(argv[4], nullptr, 10));
  if (argc > 3 && std::strcmp(argv[1], "--qasm-dump") == 0) return qasm_dump(argv[2], argv[3]);
  if (argc > 2 && std::strcmp(argv[1], "--qasm-rewrite") == 0) return qasm_rewrite(argv[2]);
  const bool host_only = argc > 1 && std::strcmp(argv[1], "--host") == 0;
  host_checks<std::complex<double>>("c128");
  host_checks<std::complex<float>>("c64");
  if (!host_only) {
    if (qsv_device_count() < 1) {
      std::printf("FAIL: no GPU visible (run with --host for the host checks)\n");
      return 2;
    }
    gate_checks<std::complex<double>>("c128");
    gate_checks<std::complex<float>>("c64");
    circuit_checks<std::complex<double>>("c128");
    circuit_checks<std::complex<float>>("c64");
    bool threw = false;
    try {
      std::vector<std::complex<double>> bad(8);
      Hadamard<std::complex<double>>(5).apply(Op::NoTrans, 3, bad);
    } catch (const backend::error& e) {
      threw = e.code() == QSV_EQUBIT;
    }
    CHECK(threw, "an out-of-range qubit raises backend::error(QSV_EQUBIT)");
  }
  std::printf("%s: %d checks, %d failed\n", host_only ? "host" : "all", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
