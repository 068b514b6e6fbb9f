"""Pauli-sum observables, expectation values and the adjoint gradient.

QCLAB++ has no observables or gradients; these build on its gate
conventions (PauliX/Y/Z matrices, MSB qubit order, rotations
exp(-i theta P / 2)).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import QSV_AUTO, QsvPauliTerm, check, lib
from .engine import _check_state, stream_handle
from .gates import records


@dataclass
class PauliSum:
    """sum_t coeff_t P_t, each P_t a list of (qubit, 'X'|'Y'|'Z')."""

    terms: list = field(default_factory=list)

    def add(self, coeff: float, factors) -> "PauliSum":
        self.terms.append((float(coeff), list(factors)))
        return self

    def records(self):
        arr = (QsvPauliTerm * max(1, len(self.terms)))()
        for i, (c, fs) in enumerate(self.terms):
            if len(fs) > 16:
                raise ValueError("at most 16 factors per Pauli term")
            arr[i].coeff = c
            arr[i].nfactors = len(fs)
            for j, (q, p) in enumerate(fs):
                arr[i].qubits[j] = q
                arr[i].paulis[j] = ord(p)
        return arr, len(self.terms)

    def matrix(self, n: int) -> np.ndarray:
        """Dense 2^n x 2^n matrix (MSB order; for small n)."""
        P = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
             "Z": np.diag([1, -1])}
        H = np.zeros((1 << n, 1 << n), dtype=np.complex128)
        for c, fs in self.terms:
            ops = ["I"] * n
            for q, p in fs:
                ops[q] = p
            m = np.array([[1.0]])
            for o in ops:
                m = np.kron(m, P[o])
            H += c * m
        return H


def tfim(n: int, J: float = 1.0, h: float = 0.7) -> PauliSum:
    """Transverse-field Ising chain, open boundary: J sum Z_i Z_{i+1} + h sum X_i."""
    H = PauliSum()
    for i in range(n - 1):
        H.add(J, [(i, "Z"), (i + 1, "Z")])
    for i in range(n):
        H.add(h, [(i, "X")])
    return H


def zz_chain(n: int) -> PauliSum:
    """<Z_i Z_{i+1}> for i = 0..n-2, unit coefficients."""
    H = PauliSum()
    for i in range(n - 1):
        H.add(1.0, [(i, "Z"), (i + 1, "Z")])
    return H


def expectation(psi, H: PauliSum, stream=None):
    """(per-term <P_t>, sum_t coeff_t <P_t>) with FP64 accumulation."""
    n, dt = _check_state(psi)
    arr, cnt = H.records()
    vals = (C.c_double * max(1, cnt))()
    tot = C.c_double()
    check(lib().qsv_dev_expectation(C.c_void_p(psi.data_ptr()), n, dt, arr, cnt, vals,
                                    C.byref(tot), stream_handle(stream)))
    return np.array(vals[:cnt]), tot.value


def norm2(psi, stream=None) -> float:
    re, _ = inner(psi, psi, stream)
    return re


def inner(a, b, stream=None):
    n, dt = _check_state(a)
    re, im = C.c_double(), C.c_double()
    check(lib().qsv_dev_inner(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), n, dt,
                              C.byref(re), C.byref(im), stream_handle(stream)))
    return re.value, im.value


def apply_pauli_sum(psi_in, psi_out, H: PauliSum, stream=None) -> None:
    n, dt = _check_state(psi_in)
    arr, cnt = H.records()
    check(lib().qsv_dev_apply_pauli_sum(C.c_void_p(psi_in.data_ptr()), C.c_void_p(psi_out.data_ptr()),
                                        n, dt, arr, cnt, stream_handle(stream)))


def adjoint_gradient(psi, lam, gates, H: PauliSum, theta=None, nparams=None, fuse_k: int = QSV_AUTO,
                     stream=None):
    """(E, dE/dtheta) by the adjoint method; psi holds psi0 on entry and is
    rewound to it; lam is scratch.  fuse_k selects the forward pass's form
    (default QSV_AUTO, the engine's fastest)."""
    n, dt = _check_state(psi)
    arr, cnt = records(gates)
    if nparams is None:
        nparams = 1 + max([g.param for g in gates if g.param >= 0], default=-1)
    th = None
    if theta is not None:
        theta = np.ascontiguousarray(np.asarray(theta, dtype=np.float64))
        if theta.size != nparams:
            raise ValueError("theta must hold one angle per parameter")
        th = theta.ctypes.data_as(C.POINTER(C.c_double))
    harr, hcnt = H.records()
    g = (C.c_double * max(1, nparams))()
    e = C.c_double()
    check(lib().qsv_dev_adjoint_gradient(C.c_void_p(psi.data_ptr()), C.c_void_p(lam.data_ptr()), n, dt,
                                         arr, cnt, th, nparams, harr, hcnt, fuse_k, C.byref(e), g,
                                         stream_handle(stream)))
    return e.value, np.array(g[:nparams])


def adjoint_profile() -> dict:
    """GPU milliseconds of the last adjoint_gradient call by phase
    (qsv_adjoint_profile_get): forward, hamiltonian (lam = H psi and E),
    reverse (the reverse sweep's launches), with the launch count and
    whether the compiled two-state passes ran."""
    from ._lib import QsvAdjointProfile
    p = QsvAdjointProfile()
    check(lib().qsv_adjoint_profile_get(C.byref(p)))
    return {"forward_ms": p.forward_ms, "hamiltonian_ms": p.hamiltonian_ms, "reverse_ms": p.reverse_ms,
            "reverse_launches": p.reverse_launches, "compiled": bool(p.compiled)}
