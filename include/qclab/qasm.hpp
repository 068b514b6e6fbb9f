// OpenQASM text of single gates (reference include/qclab/qasm.hpp):
// "name q[i];\n", "name q[i], q[j];\n", angles printed with %.15f (double)
// or %.7f (float).
#pragma once

#include <cstdio>
#include <sstream>
#include <string>

namespace qclab {

inline std::string qasm(const float angle) {
  char b[64];
  const int n = std::snprintf(b, sizeof(b), "%.7f", angle);
  return std::string(b, n);
}
inline std::string qasm(const double angle) {
  char b[64];
  const int n = std::snprintf(b, sizeof(b), "%.15f", angle);
  return std::string(b, n);
}
inline std::string qasm1(const char* type, const int q) {
  std::ostringstream s;
  s << type << " q[" << q << "];\n";
  return s.str();
}
inline std::string qasm2(const char* type, const int q0, const int q1) {
  std::ostringstream s;
  s << type << " q[" << q0 << "], q[" << q1 << "];\n";
  return s.str();
}
template <typename T>
std::string qasm1(const char* type, const int q, const T angle) {
  std::ostringstream s;
  s << type << "(" << qasm(angle) << ") q[" << q << "];\n";
  return s.str();
}
template <typename T>
std::string qasm2(const char* type, const int q0, const int q1, const T angle) {
  std::ostringstream s;
  s << type << "(" << qasm(angle) << ") q[" << q0 << "], q[" << q1 << "];\n";
  return s.str();
}

inline std::string qasmI(const int q) { return qasm1("i", q); }
inline std::string qasmH(const int q) { return qasm1("h", q); }
inline std::string qasmX(const int q) { return qasm1("x", q); }
inline std::string qasmY(const int q) { return qasm1("y", q); }
inline std::string qasmZ(const int q) { return qasm1("z", q); }
inline std::string qasmS(const int q) { return qasm1("s", q); }
inline std::string qasmT(const int q) { return qasm1("t", q); }
template <typename T> std::string qasmRx(const int q, const T a) { return qasm1("rx", q, a); }
template <typename T> std::string qasmRy(const int q, const T a) { return qasm1("ry", q, a); }
template <typename T> std::string qasmRz(const int q, const T a) { return qasm1("rz", q, a); }
template <typename T> std::string qasmRxx(const int a, const int b, const T t) { return qasm2("rxx", a, b, t); }
template <typename T> std::string qasmRyy(const int a, const int b, const T t) { return qasm2("ryy", a, b, t); }
template <typename T> std::string qasmRzz(const int a, const int b, const T t) { return qasm2("rzz", a, b, t); }
inline std::string qasmCNOT(const int c, const int t) { return qasm2("cx", c, t); }
inline std::string qasmCX(const int c, const int t) { return qasm2("cx", c, t); }
inline std::string qasmCY(const int c, const int t) { return qasm2("cy", c, t); }
inline std::string qasmCZ(const int c, const int t) { return qasm2("cz", c, t); }
template <typename T> std::string qasmCRx(const int c, const int t, const T a) { return qasm2("crx", c, t, a); }
template <typename T> std::string qasmCRy(const int c, const int t, const T a) { return qasm2("cry", c, t, a); }
template <typename T> std::string qasmCRz(const int c, const int t, const T a) { return qasm2("crz", c, t, a); }
template <typename T> std::string qasmCPhase(const int c, const int t, const T a) { return qasm2("cp", c, t, a); }
inline std::string qasmSWAP(const int a, const int b) { return qasm2("swap", a, b); }
inline std::string qasmiSWAP(const int a, const int b) { return qasm2("iswap", a, b); }

}  // namespace qclab
