"""Gate records named after QCLAB++'s gate classes.

Each constructor takes the arguments of the reference class's constructor
(include/qclab/qgates/*.hpp): qubits first, then the angle `theta` or an
explicit (cos, sin) pair, then the control state for controlled gates.
Rotations store (cos(theta/2), sin(theta/2)) and Phase/CPhase store
(cos(theta), sin(theta)), computed by the engine with the host libm as the
reference does (include/qclab/QRotation.hpp:44-46, QAngle.hpp:30-33).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace
from typing import Iterable, Sequence

import numpy as np

from ._lib import QsvGate

class K:
    """Gate kinds (enum qsv_gate_kind of include/qsv.h)."""
    I, H, X, Y, Z, RX, RY, RZ, P, T, S, U1 = range(12)
    CX, CY, CZ, CP, CRX, CRY, CRZ, CU, SWAP, ISWAP, RXX, RYY, RZZ, U2 = range(12, 26)


I, H, X, Y, Z, RX, RY, RZ, P, T, S, U1 = (K.I, K.H, K.X, K.Y, K.Z, K.RX, K.RY, K.RZ, K.P, K.T,
                                          K.S, K.U1)
_CX, _CY, _CZ, CP, CRX, CRY, CRZ, CU, _SWAP, ISWAP, RXX, RYY, RZZ, U2 = range(12, 26)

NAMES = {
    I: "Identity", H: "Hadamard", X: "PauliX", Y: "PauliY", Z: "PauliZ", RX: "RotationX",
    RY: "RotationY", RZ: "RotationZ", P: "Phase", T: "Phase45", S: "Phase90", U1: "QGate1",
    _CX: "CX", _CY: "CY", _CZ: "CZ", CP: "CPhase", CRX: "CRotationX", CRY: "CRotationY",
    CRZ: "CRotationZ", CU: "QControlledGate2", _SWAP: "SWAP", ISWAP: "iSWAP",
    RXX: "RotationXX", RYY: "RotationYY", RZZ: "RotationZZ", U2: "QGate2",
}
ROTATIONS = {RX, RY, RZ, CRX, CRY, CRZ, RXX, RYY, RZZ}  # exp(-i theta/2 P)
PHASES = {P, CP}
CONTROLLED = {_CX, _CY, _CZ, CP, CRX, CRY, CRZ, CU}
ONE_QUBIT = {I, H, X, Y, Z, RX, RY, RZ, P, T, S, U1}


@dataclass
class Gate:
    kind: int
    qubits: tuple
    theta: float = 0.0
    cs: tuple | None = None          # explicit (cos, sin)
    ctrl_state: int = 1
    param: int = -1                  # index into a parameter vector
    mat: np.ndarray | None = None    # U1/CU: 2x2, U2: 4x4 (row/col as math)
    op: str = "N"                    # 'N', 'T' or 'C': the gate applies op(matrix)

    @property
    def name(self) -> str:
        return NAMES[self.kind]

    @property
    def nb_qubits(self) -> int:
        return 1 if self.kind in ONE_QUBIT else 2

    def record(self, op: str | None = None, offset: int = 0) -> QsvGate:
        """The engine record; `op` overrides the gate's own op."""
        op = self.op if op is None else op
        g = QsvGate()
        g.kind = self.kind
        g.q[0] = self.qubits[0] + offset
        g.q[1] = (self.qubits[1] + offset) if len(self.qubits) > 1 else 0
        g.ctrl_state = self.ctrl_state
        g.op = ord(op)
        g.param = self.param
        g.theta = float(self.theta)
        if self.cs is not None:
            g.has_cs = 1
            g.cs[0], g.cs[1] = float(self.cs[0]), float(self.cs[1])
        if self.mat is not None:
            m = np.asarray(self.mat, dtype=np.complex128)
            flat = m.reshape(m.shape).flatten(order="F")  # column-major
            for k, z in enumerate(flat):
                g.mat[2 * k] = z.real
                g.mat[2 * k + 1] = z.imag
        return g

    def with_theta(self, theta: float) -> "Gate":
        return replace(self, theta=float(theta), cs=None)


def inverse(gates: Iterable[Gate]) -> list:
    """The inverse circuit: the gates in reverse order, each applied as its
    conjugate transpose (QCircuit::apply with Op::ConjTrans,
    include/qclab/QCircuit.hpp:137-142).  Gates must carry op 'N'."""
    out = []
    for g in reversed(list(gates)):
        if g.op != "N":
            raise ValueError("inverse() takes gates applied with op 'N'")
        out.append(replace(g, op="C"))
    return out


def records(gates: Iterable[Gate], op: str | None = None, offset: int = 0):
    gl = list(gates)
    arr = (QsvGate * max(1, len(gl)))()
    if not gl:
        return arr, 0
    # filled column by column through a numpy view of the records (the same
    # fields Gate.record sets, without a ctypes object per gate)
    v = np.ctypeslib.as_array(arr)
    v["kind"] = [g.kind for g in gl]
    v["q"][:, 0] = [g.qubits[0] + offset for g in gl]
    v["q"][:, 1] = [(g.qubits[1] + offset) if len(g.qubits) > 1 else 0 for g in gl]
    v["ctrl_state"] = [g.ctrl_state for g in gl]
    v["op"] = [ord(g.op if op is None else op) for g in gl]
    v["param"] = [g.param for g in gl]
    v["theta"] = [float(g.theta) for g in gl]
    for i, g in enumerate(gl):
        if g.cs is not None or g.mat is not None:
            arr[i] = g.record(op, offset)
    return arr, len(gl)


def _rot(kind, qubits, theta, cos, sin, **kw):
    cs = None if cos is None else (cos, sin)
    return Gate(kind, tuple(qubits), theta=float(theta or 0.0), cs=cs, **kw)


# ---- 1-qubit gates ---------------------------------------------------------
def Identity(q=0): return Gate(I, (q,))
def Hadamard(q=0): return Gate(H, (q,))
def PauliX(q=0): return Gate(X, (q,))
def PauliY(q=0): return Gate(Y, (q,))
def PauliZ(q=0): return Gate(Z, (q,))
def Phase45(q=0): return Gate(T, (q,))
def Phase90(q=0): return Gate(S, (q,))
def RotationX(q=0, theta=0.0, cos=None, sin=None, param=-1): return _rot(RX, (q,), theta, cos, sin, param=param)
def RotationY(q=0, theta=0.0, cos=None, sin=None, param=-1): return _rot(RY, (q,), theta, cos, sin, param=param)
def RotationZ(q=0, theta=0.0, cos=None, sin=None, param=-1): return _rot(RZ, (q,), theta, cos, sin, param=param)
def Phase(q=0, theta=0.0, cos=None, sin=None, param=-1): return _rot(P, (q,), theta, cos, sin, param=param)
def QGate1(q, mat): return Gate(U1, (q,), mat=np.asarray(mat, dtype=np.complex128))


# ---- 2-qubit gates ---------------------------------------------------------
def CX(control=0, target=1, controlState=1): return Gate(_CX, (control, target), ctrl_state=controlState)
CNOT = CX
def CY(control=0, target=1, controlState=1): return Gate(_CY, (control, target), ctrl_state=controlState)
def CZ(control=0, target=1, controlState=1): return Gate(_CZ, (control, target), ctrl_state=controlState)
def CPhase(control=0, target=1, theta=0.0, cos=None, sin=None, controlState=1, param=-1):
    return _rot(CP, (control, target), theta, cos, sin, ctrl_state=controlState, param=param)
def CRotationX(control=0, target=1, theta=0.0, cos=None, sin=None, controlState=1, param=-1):
    return _rot(CRX, (control, target), theta, cos, sin, ctrl_state=controlState, param=param)
def CRotationY(control=0, target=1, theta=0.0, cos=None, sin=None, controlState=1, param=-1):
    return _rot(CRY, (control, target), theta, cos, sin, ctrl_state=controlState, param=param)
def CRotationZ(control=0, target=1, theta=0.0, cos=None, sin=None, controlState=1, param=-1):
    return _rot(CRZ, (control, target), theta, cos, sin, ctrl_state=controlState, param=param)
def QControlledGate2(control, target, mat, controlState=1):
    return Gate(CU, (control, target), ctrl_state=controlState, mat=np.asarray(mat, dtype=np.complex128))
def SWAP(q0=0, q1=1): return Gate(_SWAP, (q0, q1))
def iSWAP(q0=0, q1=1): return Gate(ISWAP, (q0, q1))
def RotationXX(q0=0, q1=1, theta=0.0, cos=None, sin=None, param=-1): return _rot(RXX, (q0, q1), theta, cos, sin, param=param)
def RotationYY(q0=0, q1=1, theta=0.0, cos=None, sin=None, param=-1): return _rot(RYY, (q0, q1), theta, cos, sin, param=param)
def RotationZZ(q0=0, q1=1, theta=0.0, cos=None, sin=None, param=-1): return _rot(RZZ, (q0, q1), theta, cos, sin, param=param)
def QGate2(q0, q1, mat): return Gate(U2, (q0, q1), mat=np.asarray(mat, dtype=np.complex128))
