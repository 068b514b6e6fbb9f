"""Gate instances covering every QCLAB++ gate class, Op and qubit layout.

Shared by the oracle-vs-reference tests and the GPU parity tests.
"""
from __future__ import annotations

import math

import numpy as np

from qclabpp_amd import gates as G


def random_state(n: int, seed: int, dtype=np.complex128) -> np.ndarray:
    rs = np.random.default_rng(seed)
    v = rs.standard_normal(1 << n) + 1j * rs.standard_normal(1 << n)
    v /= np.linalg.norm(v)
    return v.astype(dtype)


def random_unitary(dim: int, seed: int) -> np.ndarray:
    rs = np.random.default_rng(seed)
    z = rs.standard_normal((dim, dim)) + 1j * rs.standard_normal((dim, dim))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def one_qubit_gates(q: int, seed: int = 0) -> list:
    th = 0.3 + 0.7 * q + 0.1 * seed
    return [
        G.Identity(q), G.Hadamard(q), G.PauliX(q), G.PauliY(q), G.PauliZ(q),
        G.RotationX(q, th), G.RotationY(q, -th), G.RotationZ(q, th * 1.3),
        G.RotationX(q, cos=0.6, sin=0.8), G.Phase(q, th), G.Phase45(q), G.Phase90(q),
        G.QGate1(q, random_unitary(2, 100 + q + seed)),
        G.QGate1(q, np.array([[1, 2j], [3, -4 + 1j]])),  # non-unitary, asymmetric
    ]


def two_qubit_gates(a: int, b: int, seed: int = 0) -> list:
    th = 0.45 + 0.3 * a - 0.2 * b + 0.05 * seed
    lo, hi = min(a, b), max(a, b)
    out = []
    for cs in (1, 0):
        out += [
            G.CX(a, b, cs), G.CY(a, b, cs), G.CZ(a, b, cs), G.CPhase(a, b, th, controlState=cs),
            G.CRotationX(a, b, th, controlState=cs), G.CRotationY(a, b, -th, controlState=cs),
            G.CRotationZ(a, b, th, controlState=cs),
            G.QControlledGate2(a, b, random_unitary(2, 7 + a + 3 * b), controlState=cs),
        ]
    out += [
        G.SWAP(a, b), G.iSWAP(a, b), G.RotationXX(a, b, th), G.RotationYY(a, b, -th),
        G.RotationZZ(a, b, th), G.RotationXX(a, b, cos=0.0, sin=1.0),
        G.QGate2(lo, hi, random_unitary(4, 11 + a + 5 * b)),
        G.QGate2(lo, hi, np.arange(16).reshape(4, 4) + 1j * np.eye(4)),  # asymmetric
    ]
    return out


def all_cases(n: int, seed: int = 0):
    """(gate, op) pairs over every qubit (1q) and ordered qubit pair (2q)."""
    cases = []
    for q in range(n):
        for g in one_qubit_gates(q, seed):
            for op in "NTC":
                cases.append((g, op))
    for a in range(n):
        for b in range(n):
            if a == b:
                continue
            for g in two_qubit_gates(a, b, seed):
                for op in "NTC":
                    cases.append((g, op))
    return cases
