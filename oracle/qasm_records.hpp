// ORACLE -- test infrastructure only.
//
// One QCLAB++ gate object as a qsv_gate record: kind, qubits (control first
// for controlled gates), control state and angle, read through the gate
// classes' public accessors (qubit(), qubits(), control(), target(),
// controlState(), theta()).  The same function is compiled against the
// reference's headers (ref_driver.cpp, for io::QASMFile of
// include/qclab/io/QASMFile.hpp:42-152) and against this repository's
// drop-in headers (tests/cpp/dropin_test.cpp), so both QASM readers are
// compared gate by gate with exact angles.
#pragma once

#include "qsv.h"

namespace qsv_oracle {

template <typename T>
bool qasm_record(const qclab::QObject<T>& g, qsv_gate& r) {
  namespace Q = qclab::qgates;
  r = qsv_gate{};
  r.param = -1;
  r.ctrl_state = 1;
  r.op = QSV_OP_N;
  auto one = [&](int kind, int q) {
    r.kind = kind;
    r.q[0] = q;
    return true;
  };
  auto ctl = [&](int kind, const Q::QControlledGate2<T>& c) {
    r.kind = kind;
    r.q[0] = c.control();
    r.q[1] = c.target();
    r.ctrl_state = c.controlState();
    return true;
  };
  auto two = [&](int kind, const std::vector<int>& q) {
    r.kind = kind;
    r.q[0] = q[0];
    r.q[1] = q[1];
    return true;
  };
  if (auto* p = dynamic_cast<const Q::Identity<T>*>(&g)) return one(QSV_GATE_I, p->qubit());
  if (auto* p = dynamic_cast<const Q::Hadamard<T>*>(&g)) return one(QSV_GATE_H, p->qubit());
  if (auto* p = dynamic_cast<const Q::PauliX<T>*>(&g)) return one(QSV_GATE_X, p->qubit());
  if (auto* p = dynamic_cast<const Q::PauliY<T>*>(&g)) return one(QSV_GATE_Y, p->qubit());
  if (auto* p = dynamic_cast<const Q::PauliZ<T>*>(&g)) return one(QSV_GATE_Z, p->qubit());
  if (auto* p = dynamic_cast<const Q::Phase45<T>*>(&g)) return one(QSV_GATE_T, p->qubit());
  if (auto* p = dynamic_cast<const Q::Phase90<T>*>(&g)) return one(QSV_GATE_S, p->qubit());
  if (auto* p = dynamic_cast<const Q::RotationX<T>*>(&g)) { r.theta = double(p->theta()); return one(QSV_GATE_RX, p->qubit()); }
  if (auto* p = dynamic_cast<const Q::RotationY<T>*>(&g)) { r.theta = double(p->theta()); return one(QSV_GATE_RY, p->qubit()); }
  if (auto* p = dynamic_cast<const Q::RotationZ<T>*>(&g)) { r.theta = double(p->theta()); return one(QSV_GATE_RZ, p->qubit()); }
  if (auto* p = dynamic_cast<const Q::Phase<T>*>(&g)) { r.theta = double(p->theta()); return one(QSV_GATE_P, p->qubit()); }
  if (auto* p = dynamic_cast<const Q::CX<T>*>(&g)) return ctl(QSV_GATE_CX, *p);
  if (auto* p = dynamic_cast<const Q::CY<T>*>(&g)) return ctl(QSV_GATE_CY, *p);
  if (auto* p = dynamic_cast<const Q::CZ<T>*>(&g)) return ctl(QSV_GATE_CZ, *p);
  if (auto* p = dynamic_cast<const Q::CPhase<T>*>(&g)) { r.theta = double(p->theta()); return ctl(QSV_GATE_CP, *p); }
  if (auto* p = dynamic_cast<const Q::CRotationX<T>*>(&g)) { r.theta = double(p->theta()); return ctl(QSV_GATE_CRX, *p); }
  if (auto* p = dynamic_cast<const Q::CRotationY<T>*>(&g)) { r.theta = double(p->theta()); return ctl(QSV_GATE_CRY, *p); }
  if (auto* p = dynamic_cast<const Q::CRotationZ<T>*>(&g)) { r.theta = double(p->theta()); return ctl(QSV_GATE_CRZ, *p); }
  if (auto* p = dynamic_cast<const Q::SWAP<T>*>(&g)) return two(QSV_GATE_SWAP, p->qubits());
  if (auto* p = dynamic_cast<const Q::iSWAP<T>*>(&g)) return two(QSV_GATE_ISWAP, p->qubits());
  if (auto* p = dynamic_cast<const Q::RotationXX<T>*>(&g)) { r.theta = double(p->theta()); return two(QSV_GATE_RXX, p->qubits()); }
  if (auto* p = dynamic_cast<const Q::RotationYY<T>*>(&g)) { r.theta = double(p->theta()); return two(QSV_GATE_RYY, p->qubits()); }
  if (auto* p = dynamic_cast<const Q::RotationZZ<T>*>(&g)) { r.theta = double(p->theta()); return two(QSV_GATE_RZZ, p->qubits()); }
  return false;
}

}  // namespace qsv_oracle
