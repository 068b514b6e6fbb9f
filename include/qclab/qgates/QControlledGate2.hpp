// Base class of controlled 1-qubit gates
// (reference include/qclab/qgates/QControlledGate2.hpp).
#pragma once

#include <algorithm>
#include <cassert>

#include "qclab/qgates/QGate1.hpp"
#include "qclab/qgates/QGate2.hpp"

namespace qclab {
namespace qgates {

template <typename T>
class QControlledGate2 : public QGate2<T> {
 public:
  using gate_type = QGate1<T>;

  QControlledGate2(const int control, const int controlState = 1)
      : control_(control), controlState_(controlState) {
    assert(control >= 0);
    assert(controlState == 0 || controlState == 1);
  }

  bool controlled() const override { return true; }
  std::vector<int> qubits() const override {
    return {std::min(control(), target()), std::max(control(), target())};
  }
  void setQubits(const int* qubits) override {
    assert(qubits[0] >= 0 && qubits[1] >= 0 && qubits[0] != qubits[1]);
    control_ = qubits[0];
    setTarget(qubits[1]);
  }

  /// Controlled matrix in the order of qubits() (ascending).
  dense::SquareMatrix<T> matrix() const override {
    const dense::SquareMatrix<T> I2 = dense::eye<T>(2);
    const dense::SquareMatrix<T> P0(T(1), T(0), T(0), T(0)), P1(T(0), T(0), T(0), T(1));
    const dense::SquareMatrix<T> G = gate()->matrix();
    const dense::SquareMatrix<T>& on = controlState_ == 1 ? P1 : P0;
    const dense::SquareMatrix<T>& off = controlState_ == 1 ? P0 : P1;
    if (control() < target()) return dense::kron(off, I2) + dense::kron(on, G);
    return dense::kron(I2, off) + dense::kron(G, on);
  }

  int control() const { return control_; }
  virtual int target() const = 0;
  int controlState() const { return controlState_; }
  virtual const gate_type* gate() const = 0;
  void setControl(const int control) {
    assert(control >= 0 && control != target());
    control_ = control;
  }
  virtual void setTarget(const int target) = 0;
  void setControlState(const int controlState) {
    assert(controlState == 0 || controlState == 1);
    controlState_ = controlState;
  }

  /// Generic controlled gates run as a controlled dense 2x2 (QSV_GATE_CU).
  void records(std::vector<qsv_gate>& out, Op op, int offset) const override {
    qsv_gate g = record(QSV_GATE_CU, op, offset);
    backend::set_mat(g, gate()->matrix());
    out.push_back(g);
  }

 protected:
  qsv_gate record(int kind, Op op, int offset) const {
    qsv_gate g = backend::make(kind, control() + offset, target() + offset, char(op));
    g.ctrl_state = controlState_;
    return g;
  }
  // QASM of a gate controlled on |0>: X on the control before and after
  template <typename F>
  int qasm_controlled(std::ostream& stream, const int offset, F&& body) const {
    if (controlState_ == 0) stream << qasmX(control() + offset);
    stream << body(control() + offset, target() + offset);
    if (controlState_ == 0) stream << qasmX(control() + offset);
    return 0;
  }
  bool same_layout(const QControlledGate2<T>& o) const {
    return o.controlState() == controlState() &&
           ((o.control() < o.target()) == (control() < target()));
  }

  int control_;
  int controlState_;
};

}  // namespace qgates
}  // namespace qclab
