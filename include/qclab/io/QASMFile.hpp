// OpenQASM 2.0 reader (reference include/qclab/io/QASMFile.hpp:42-152,
// src/io/QASMFile.cpp:16-53).
//
// Dialect::Reference (the default) reads exactly what the reference reads:
// whitespace is removed, commands split at ';', everything before the first
// `qreg` is ignored, and the commands h x y z rx ry rz p cx cy cz swap
// (numeric angles only) become gates; other commands are skipped.  A command
// that fails to parse (e.g. an angle written with `pi`) stops the gate list
// there, prints "Failed parsing QASM file!" and returns the gates read so far.
// Where the reference's string helpers never return (a last command without
// ';', an empty command between two ';'), this reader ends the command.
//
// Dialect::Extended reads the circuits external tools write: angle
// expressions (pi, + - * / ^, unary minus, parentheses, sin cos tan exp ln
// sqrt), several qregs (concatenated in declaration order), register
// broadcast (`h q;`), and the gates id h x y z s sdg t tdg rx ry rz p u1 cx
// cy cz cp cu1 crx cry crz swap iswap rxx ryy rzz -- every gate the QASM
// writer emits.  creg / barrier / measure / reset are skipped and counted;
// an unknown gate or a `gate` definition stops the list as above.
#pragma once

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "qclab/io/util.hpp"
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

namespace qclab::io {

class QASMFile {
 public:
  enum class Dialect { Reference, Extended };

  /// One gate statement of the Extended dialect, resolved to global qubits.
  struct Statement {
    std::string name;
    std::vector<double> params;
    std::vector<int> qubits;
  };

  QASMFile(const std::string filename, const Dialect dialect = Dialect::Reference)
      : dialect_(dialect), nbQubits_(0) {
    std::ifstream stream(filename);
    if (!stream.good()) throw std::runtime_error("QASMFile: cannot open " + filename);
    std::stringstream text;
    text << stream.rdbuf();
    if (dialect == Dialect::Reference) parse_reference(text.str());
    else parse_extended(text.str());
  }

  /// From QASM text instead of a file.
  static QASMFile from_string(const std::string& text, const Dialect dialect = Dialect::Reference) {
    QASMFile f(dialect);
    if (dialect == Dialect::Reference) f.parse_reference(text);
    else f.parse_extended(text);
    return f;
  }

  int nbQubits() const { return nbQubits_; }
  Dialect dialect() const { return dialect_; }
  /// Extended dialect: statements skipped as non-unitary (creg, barrier, ...).
  int skipped() const { return skipped_; }
  /// Extended dialect: the parse error that ended the gate list ("" if none).
  const std::string& error() const { return error_; }

  /// The gates of the file, as objects of the gate hierarchy.
  template <typename G>
  std::vector<std::unique_ptr<G>> gates() const {
    return dialect_ == Dialect::Reference ? reference_gates<G>() : extended_gates<G>();
  }

 private:
  explicit QASMFile(const Dialect d) : dialect_(d), nbQubits_(0) {}

  // ---- reference dialect --------------------------------------------------
  void parse_reference(const std::string& text) {
    bool hasQreg = false;
    std::istringstream lines(text);
    std::string str;
    while (std::getline(lines, str)) {
      str.erase(std::remove_if(str.begin(), str.end(), [](unsigned char c) { return std::isspace(c); }),
                str.end());
      if (str.empty() || str.compare(0, 2, "//") == 0) continue;
      for (auto& command : split<std::string>(str, ";")) {
        if (!hasQreg) {
          if (command.compare(0, 4, "qreg") == 0) {
            std::string rest = command.substr(4);
            qregName_ = left_of(rest, "[");
            nbQubits_ = read_value<int>(left_of(rest, "]"));
            hasQreg = true;
          }
        } else {
          commands_.push_back(command);
        }
      }
    }
  }

  int parseQubit(std::string& str) const { return read_value<int>(left_of(str, "]")); }

  template <typename GATE, typename G>
  bool parse1const(const std::string& name, std::string command, std::unique_ptr<G>& gate) const {
    const auto n1 = name.size(), n2 = qregName_.size();
    if (command.compare(0, n1 + n2, name + qregName_) != 0) return false;
    command.erase(0, n1 + n2 + 1);
    gate = std::make_unique<GATE>(parseQubit(command));
    return true;
  }

  template <typename GATE, typename G>
  bool parse1a(const std::string& name, std::string command, std::unique_ptr<G>& gate) const {
    using R = real_t<typename GATE::value_type>;
    const auto n1 = name.size(), n2 = qregName_.size();
    if (command.compare(0, n1 + 1, name + "(") != 0) return false;
    command.erase(0, n1 + 1);
    const R angle = read_value<R>(left_of(command, ")"));
    command.erase(0, n2 + 1);
    const int qubit = parseQubit(command);
    gate = std::make_unique<GATE>(qubit, angle);
    return true;
  }

  template <typename GATE, typename G>
  bool parse2const(const std::string& name, std::string command, std::unique_ptr<G>& gate) const {
    const auto n1 = name.size(), n2 = qregName_.size();
    if (command.compare(0, n1 + n2, name + qregName_) != 0) return false;
    command.erase(0, n1 + n2 + 1);
    const int q0 = parseQubit(command);
    command.erase(0, n2 + 2);
    const int q1 = parseQubit(command);
    gate = std::make_unique<GATE>(q0, q1);
    return true;
  }

  template <typename G>
  std::vector<std::unique_ptr<G>> reference_gates() const {
    using T = typename G::value_type;
    using namespace qclab::qgates;
    std::vector<std::unique_ptr<G>> out;
    std::unique_ptr<G> g;
    try {
      for (const auto& c : commands_) {
        if (parse1const<Hadamard<T>>("h", c, g) || parse1const<PauliX<T>>("x", c, g) ||
            parse1const<PauliY<T>>("y", c, g) || parse1const<PauliZ<T>>("z", c, g) ||
            parse1a<RotationX<T>>("rx", c, g) || parse1a<RotationY<T>>("ry", c, g) ||
            parse1a<RotationZ<T>>("rz", c, g) || parse1a<Phase<T>>("p", c, g) ||
            parse2const<CX<T>>("cx", c, g) || parse2const<CY<T>>("cy", c, g) ||
            parse2const<CZ<T>>("cz", c, g) || parse2const<SWAP<T>>("swap", c, g))
          out.push_back(std::move(g));
      }
    } catch (const std::exception&) {
      std::cerr << "Failed parsing QASM file!" << std::endl;
    }
    return out;
  }

  // ---- extended dialect ---------------------------------------------------
  struct Lexer {
    const std::string& s;
    size_t i = 0;
    explicit Lexer(const std::string& str) : s(str) {}
    void ws() { while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i; }
    bool eat(char c) { ws(); if (i < s.size() && s[i] == c) { ++i; return true; } return false; }
    void expect(char c) { if (!eat(c)) fail(std::string("expected '") + c + "'"); }
    bool done() { ws(); return i >= s.size(); }
    std::string ident() {
      ws();
      const size_t b = i;
      while (i < s.size() && (std::isalnum(static_cast<unsigned char>(s[i])) || s[i] == '_')) ++i;
      if (b == i) fail("expected a name");
      return s.substr(b, i - b);
    }
    [[noreturn]] void fail(const std::string& what) const {
      throw std::runtime_error(what + " in '" + s + "'");
    }
    // expr := term (('+'|'-') term)* ; term := unary (('*'|'/') unary)* ;
    // unary := '-' unary | power ; power := atom ('^' unary)?
    double expr() {
      double v = term();
      for (;;) {
        if (eat('+')) v += term();
        else if (eat('-')) v -= term();
        else return v;
      }
    }
    double term() {
      double v = unary();
      for (;;) {
        if (eat('*')) v *= unary();
        else if (eat('/')) v /= unary();
        else return v;
      }
    }
    double unary() {
      if (eat('-')) return -unary();
      if (eat('+')) return unary();
      const double a = atom();
      if (eat('^')) return std::pow(a, unary());
      return a;
    }
    double atom() {
      ws();
      if (eat('(')) { const double v = expr(); expect(')'); return v; }
      if (i < s.size() && (std::isdigit(static_cast<unsigned char>(s[i])) || s[i] == '.')) {
        size_t used = 0;
        const double v = std::stod(s.substr(i), &used);
        i += used;
        return v;
      }
      const std::string name = ident();
      if (name == "pi") return 3.14159265358979323846;
      double (*fn)(double) = nullptr;
      if (name == "sin") fn = [](double x) { return std::sin(x); };
      else if (name == "cos") fn = [](double x) { return std::cos(x); };
      else if (name == "tan") fn = [](double x) { return std::tan(x); };
      else if (name == "exp") fn = [](double x) { return std::exp(x); };
      else if (name == "ln") fn = [](double x) { return std::log(x); };
      else if (name == "sqrt") fn = [](double x) { return std::sqrt(x); };
      else fail("unknown symbol '" + name + "'");
      expect('(');
      const double v = expr();
      expect(')');
      return fn(v);
    }
  };

  // one register argument: a list of global qubits (one, or the register)
  std::vector<int> argument(Lexer& lx) const {
    const std::string reg = lx.ident();
    const auto it = std::find_if(qregs_.begin(), qregs_.end(), [&](const auto& r) { return r.first == reg; });
    if (it == qregs_.end()) lx.fail("unknown register '" + reg + "'");
    const int base = it->second.first, size = it->second.second;
    if (lx.eat('[')) {
      lx.ws();
      size_t used = 0;
      const int idx = std::stoi(lx.s.substr(lx.i), &used);
      lx.i += used;
      lx.expect(']');
      if (idx < 0 || idx >= size) lx.fail("qubit index out of range");
      return {base + idx};
    }
    std::vector<int> all(size);
    for (int k = 0; k < size; ++k) all[k] = base + k;
    return all;
  }

  static std::string strip_comments(const std::string& text) {
    std::string out;
    std::istringstream lines(text);
    std::string line;
    while (std::getline(lines, line)) {
      const auto c = line.find("//");
      out += (c == std::string::npos ? line : line.substr(0, c));
      out += '\n';
    }
    return out;
  }

  static int arity(const std::string& n) {
    static const char* one[] = {"id", "h", "x", "y", "z", "s", "sdg", "t", "tdg", "rx", "ry", "rz", "p", "u1"};
    static const char* two[] = {"cx", "CX", "cnot", "cy", "cz", "cp", "cu1", "cphase", "crx", "cry", "crz",
                                "swap", "iswap", "rxx", "ryy", "rzz"};
    for (const char* o : one) if (n == o) return 1;
    for (const char* t : two) if (n == t) return 2;
    return 0;
  }
  static int nparams(const std::string& n) {
    static const char* withp[] = {"rx", "ry", "rz", "p", "u1", "cp", "cu1", "cphase", "crx", "cry", "crz",
                                  "rxx", "ryy", "rzz"};
    for (const char* w : withp) if (n == w) return 1;
    return 0;
  }

  void parse_extended(const std::string& text) {
    const std::string body = strip_comments(text);
    size_t start = 0;
    try {
      while (start < body.size()) {
        size_t end = body.find(';', start);
        std::string stmt = body.substr(start, end == std::string::npos ? std::string::npos : end - start);
        start = end == std::string::npos ? body.size() : end + 1;
        Lexer lx(stmt);
        if (lx.done()) continue;
        const std::string head = lx.ident();
        if (head == "OPENQASM" || head == "include") continue;
        if (head == "qreg") {
          const std::string name = lx.ident();
          lx.expect('[');
          lx.ws();
          size_t used = 0;
          const int size = std::stoi(lx.s.substr(lx.i), &used);
          lx.i += used;
          lx.expect(']');
          qregs_.push_back({name, {nbQubits_, size}});
          nbQubits_ += size;
          continue;
        }
        if (head == "creg" || head == "barrier" || head == "measure" || head == "reset") {
          ++skipped_;
          continue;
        }
        if (head == "gate" || head == "opaque" || head == "if")
          lx.fail("'" + head + "' statements are not supported");
        const int ar = arity(head);
        if (ar == 0) lx.fail("unknown gate '" + head + "'");
        std::vector<double> params;
        if (lx.eat('(')) {
          if (!lx.eat(')')) {
            do params.push_back(lx.expr()); while (lx.eat(','));
            lx.expect(')');
          }
        }
        if (int(params.size()) != nparams(head)) lx.fail("wrong number of parameters");
        std::vector<std::vector<int>> args;
        do args.push_back(argument(lx)); while (lx.eat(','));
        if (!lx.done()) lx.fail("unexpected text");
        if (int(args.size()) != ar) lx.fail("wrong number of qubits");
        // broadcast: registers of equal size apply element-wise
        size_t width = 1;
        for (const auto& a : args) width = std::max(width, a.size());
        for (const auto& a : args)
          if (a.size() != 1 && a.size() != width) lx.fail("register sizes differ");
        for (size_t k = 0; k < width; ++k) {
          Statement st{head, params, {}};
          for (const auto& a : args) st.qubits.push_back(a.size() == 1 ? a[0] : a[k]);
          if (st.qubits.size() == 2 && st.qubits[0] == st.qubits[1]) lx.fail("repeated qubit");
          statements_.push_back(std::move(st));
        }
      }
    } catch (const std::exception& e) {
      error_ = e.what();
      std::cerr << "Failed parsing QASM file!" << std::endl;
    }
  }

  template <typename G>
  std::vector<std::unique_ptr<G>> extended_gates() const {
    using T = typename G::value_type;
    using R = real_t<T>;
    using namespace qclab::qgates;
    const R h = R(1) / std::sqrt(R(2));
    std::vector<std::unique_ptr<G>> out;
    for (const auto& st : statements_) {
      const std::string& n = st.name;
      const int a = st.qubits[0], b = st.qubits.size() > 1 ? st.qubits[1] : 0;
      const R th = st.params.empty() ? R(0) : R(st.params[0]);
      if (n == "id") out.push_back(std::make_unique<Identity<T>>(a));
      else if (n == "h") out.push_back(std::make_unique<Hadamard<T>>(a));
      else if (n == "x") out.push_back(std::make_unique<PauliX<T>>(a));
      else if (n == "y") out.push_back(std::make_unique<PauliY<T>>(a));
      else if (n == "z") out.push_back(std::make_unique<PauliZ<T>>(a));
      else if (n == "s") out.push_back(std::make_unique<Phase90<T>>(a));
      else if (n == "t") out.push_back(std::make_unique<Phase45<T>>(a));
      else if (n == "sdg") out.push_back(std::make_unique<Phase<T>>(a, R(0), R(-1)));
      else if (n == "tdg") out.push_back(std::make_unique<Phase<T>>(a, h, -h));
      else if (n == "rx") out.push_back(std::make_unique<RotationX<T>>(a, th));
      else if (n == "ry") out.push_back(std::make_unique<RotationY<T>>(a, th));
      else if (n == "rz") out.push_back(std::make_unique<RotationZ<T>>(a, th));
      else if (n == "p" || n == "u1") out.push_back(std::make_unique<Phase<T>>(a, th));
      else if (n == "cx" || n == "CX" || n == "cnot") out.push_back(std::make_unique<CX<T>>(a, b));
      else if (n == "cy") out.push_back(std::make_unique<CY<T>>(a, b));
      else if (n == "cz") out.push_back(std::make_unique<CZ<T>>(a, b));
      else if (n == "cp" || n == "cu1" || n == "cphase") out.push_back(std::make_unique<CPhase<T>>(a, b, th));
      else if (n == "crx") out.push_back(std::make_unique<CRotationX<T>>(a, b, th));
      else if (n == "cry") out.push_back(std::make_unique<CRotationY<T>>(a, b, th));
      else if (n == "crz") out.push_back(std::make_unique<CRotationZ<T>>(a, b, th));
      else if (n == "swap") out.push_back(std::make_unique<SWAP<T>>(a, b));
      else if (n == "iswap") out.push_back(std::make_unique<iSWAP<T>>(a, b));
      else if (n == "rxx") out.push_back(std::make_unique<RotationXX<T>>(a, b, th));
      else if (n == "ryy") out.push_back(std::make_unique<RotationYY<T>>(a, b, th));
      else if (n == "rzz") out.push_back(std::make_unique<RotationZZ<T>>(a, b, th));
    }
    return out;
  }

  Dialect dialect_;
  int nbQubits_;
  int skipped_ = 0;
  std::string error_;
  std::string qregName_;
  std::vector<std::string> commands_;
  std::vector<std::pair<std::string, std::pair<int, int>>> qregs_;
  std::vector<Statement> statements_;
};

}  // namespace qclab::io
