"""The fusion pass, checked on the CPU (no GPU needed: block coefficients are
built on the host and only uploaded at the first run).

The product of every launch's operator (qsv_program_op), embedded in the
full 2^n space, must equal the reference's own QCircuit::matrix() of the
whole circuit (include/qclab/QCircuit.hpp:121-127, via oracle/_ref) -- and,
where _ref is absent, the brute-force kron product of the oracle's per-gate
action on basis vectors.
"""
import numpy as np
import pytest

import oracle
from qclabpp_amd import engine as E
from qclabpp_amd import gates as G
from qclabpp_amd import workloads as W
from tests.gatecases import random_unitary


def embed(n, qubits, diag, m):
    """Full 2^n operator of a k-qubit op (row bit p <-> qubits[k-1-p])."""
    k = len(qubits)
    N = 1 << n
    bits = [n - 1 - qubits[k - 1 - p] for p in range(k)]
    idx = np.arange(N)
    loc = np.zeros(N, dtype=np.int64)
    for p, b in enumerate(bits):
        loc |= ((idx >> b) & 1) << p
    rest = idx.copy()
    for b in bits:
        rest &= ~(1 << b)
    if diag:
        return np.diag(m[loc])
    full = np.zeros((N, N), dtype=np.complex128)
    same = rest[:, None] == rest[None, :]
    full[same] = m[loc[:, None], loc[None, :]][same]
    return full


def program_unitary(n, prog):
    U = np.eye(1 << n, dtype=np.complex128)
    for qubits, diag, m in prog.launches():
        U = embed(n, qubits, diag, m) @ U
    return U


def oracle_unitary(n, circ):
    N = 1 << n
    cols = []
    for c in range(N):
        e = np.zeros(N, np.complex128)
        e[c] = 1
        cols.append(oracle.apply(n, circ, e))
    return np.stack(cols, axis=1)


def circuits(n):
    rs = np.random.default_rng(5)
    mixed = []
    for i in range(40):
        a, b = rs.choice(n, 2, replace=False)
        th = rs.uniform(-3, 3)
        mixed += [G.RotationY(a, th), G.CX(a, b), G.RotationZ(b, -th), G.CPhase(b, a, th),
                  G.SWAP(a, b), G.RotationZZ(min(a, b), max(a, b), th), G.Hadamard(b),
                  G.CRotationX(a, b, th, controlState=0), G.iSWAP(a, b), G.PauliY(a),
                  G.Phase45(b), G.QGate2(min(a, b), max(a, b), random_unitary(4, i))]
    return {"brickwork": W.brickwork(n, 8, seed=3), "qft": W.qft(n), "tfim": W.tfim_ansatz(n, 2),
            "mixed": mixed}


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("name", ["brickwork", "qft", "tfim", "mixed"])
def test_fused_unitary_matches(name, k):
    n = 6
    circ = circuits(n)[name]
    prog = E.Program(n, circ, fuse_k=k)
    got = program_unitary(n, prog)
    if oracle.ref_available() and name != "mixed":
        want = oracle.ref_matrix(n, circ)  # QCircuit::matrix() of the reference
    else:
        want = oracle_unitary(n, circ)  # QGate2 with an asymmetric matrix: vector semantics
    assert np.abs(got - want).max() < 1e-12


def test_fusion_counts_brickwork30():
    """Fewer launches and algorithmic bytes than the plain path."""
    circ = W.brickwork(30, 20)
    plain = E.Program(30, circ, fuse_k=0).info()
    f5 = E.Program(30, circ, fuse_k=5).info()
    assert plain["nlaunch"] == 890
    assert f5["nlaunch"] < plain["nlaunch"] / 8
    assert f5["alg_bytes"] < plain["alg_bytes"] / 8
