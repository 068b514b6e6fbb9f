"""Pin the oracle's QCircuit::matrix() restatement.

Closed form: the QFT's unitary is the DFT matrix with the e^{+2 pi i rc/N}
sign.  Where oracle/_ref is built: the reference's own matrix() on the same
circuits, with every gate class and Op, QGate2 on adjacent qubits (the
reference's QGate2::apply(Side, ...) asserts qubits[0] + 1 == qubits[1],
src/qgates/QGate2.cpp:41).
"""
import dataclasses

import numpy as np
import pytest

import oracle
from qclabpp_amd import gates as G
from qclabpp_amd import workloads as W
from tests.gatecases import one_qubit_gates, random_unitary, two_qubit_gates


def dft(k):
    N = 1 << k
    r = np.arange(N)
    return np.exp(2j * np.pi * np.outer(r, r) / N) / np.sqrt(N)


@pytest.mark.parametrize("k", [1, 3, 6])
def test_qft_matrix_is_dft(k):
    assert np.abs(oracle.matrix(k, W.qft(k)) - dft(k)).max() < 1e-13


def test_matrix_is_unitary_and_columns_are_simulate():
    k = 5
    circ = W.brickwork(k, 4, seed=5) + [G.QGate2(1, 3, random_unitary(4, 3))]
    M = oracle.matrix(k, circ)
    assert np.abs(M @ M.conj().T - np.eye(1 << k)).max() < 1e-13
    # columns of gates without QGate2 are simulate() of basis states
    plain = W.brickwork(k, 4, seed=5)
    Mp = oracle.matrix(k, plain)
    for c in (0, 7, 31):
        e = np.zeros(1 << k, np.complex128)
        e[c] = 1
        assert np.abs(Mp[:, c] - oracle.apply(k, plain, e)).max() == 0


def test_qgate2_kron_order():
    """matrix() of a lone QGate2 on qubits (0, 1) of 2 is its matrix()."""
    A = np.arange(16).reshape(4, 4) + 1j * np.eye(4)
    M = oracle.matrix(2, [G.QGate2(0, 1, A)])
    assert np.abs(M - A).max() == 0
    # and the vector path uses the transposed basis (apply.hpp:289-308)
    p = [0, 2, 1, 3]
    e = np.zeros(4, np.complex128)
    e[1] = 1
    assert np.abs(oracle.apply(2, [G.QGate2(0, 1, A)], e) - A[np.ix_(p, p)][:, 1]).max() == 0


def mixed_circuit(k, seed):
    out = []
    rs = np.random.default_rng(seed)
    for q in range(k):
        out += one_qubit_gates(q, seed)[1:]
    for a in range(k - 1):
        for g in two_qubit_gates(a, a + 1, seed):
            out.append(g)
    for g in two_qubit_gates(k - 1, 0, seed):
        if g.kind not in (G.U2, G.RXX, G.RYY, G.RZZ):  # non-adjacent: QGate2 family asserts
            out.append(g)
    rs.shuffle(out)
    return out


def test_record_ops_transpose_and_adjoint():
    """A record's own Op: matrix([g with T]) = matrix([g])^T, C gives ^H."""
    k = 2
    for g in mixed_circuit(k, 4):
        M = oracle.matrix(k, [g])
        assert np.abs(oracle.matrix(k, [dataclasses.replace(g, op="T")]) - M.T).max() < 1e-14, g.name
        assert np.abs(oracle.matrix(k, [dataclasses.replace(g, op="C")]) - M.conj().T).max() < 1e-14, g.name


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("k", [2, 4])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_matrix_vs_reference(k, dtype):
    """Gate objects carry no Op in the reference (matrix() applies each with
    Op::NoTrans), so every record here has op 'N'."""
    circ = mixed_circuit(k, 3 + k)
    got = oracle.matrix(k, circ, dtype)
    want = oracle.ref_matrix(k, circ, dtype)
    scale = max(1.0, float(np.abs(want).max()))
    tol = (1e-12 if dtype == np.complex128 else 2e-5) * scale
    assert np.abs(got - want).max() <= tol
