"""Pin the CPU restatement against QCLAB++ itself, compiled from its sources.

Skipped where oracle/_ref was not built (the reference tree is absent).
"""
import numpy as np
import pytest

import oracle
from qclabpp_amd import workloads as W
from tests.gatecases import all_cases, random_state

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


def ulps(dtype, k=8):
    return k * np.finfo(np.float64 if dtype == np.complex128 else np.float32).eps


@pytest.mark.parametrize("n", [2, 3, 5])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_every_gate_op_layout(n, dtype):
    psi = random_state(n, 1234 + n, dtype)
    for gate, op in all_cases(n):
        got = oracle.apply_gate(n, gate, psi, op)
        want = oracle.ref_apply_gate(n, gate, psi, op)
        # Same formulas, same compiler; only FMA contraction of the 4-term
        # sums of the dense 2-qubit lambda may differ by an ulp or two.
        assert np.abs(got - want).max() <= ulps(dtype), (gate.name, gate.qubits, op, n)


@pytest.mark.parametrize("op", ["N", "T", "C"])
def test_circuit_apply_order(op):
    n = 6
    circ = W.brickwork(n, 5, seed=77) + W.qft(n)[:12]
    psi = random_state(n, 5, np.complex128)
    got = oracle.apply(n, circ, psi, op)
    want, _ = oracle.ref_apply(n, circ, psi, op)
    assert np.abs(got - want).max() <= ulps(np.complex128, 64)


def test_tfim_ansatz_and_qft():
    n = 8
    for circ in (W.tfim_ansatz(n, 2), W.qft(n), W.trotter(n)):
        psi = random_state(n, 9, np.complex128)
        got = oracle.apply(n, circ, psi)
        want, _ = oracle.ref_apply(n, circ, psi)
        assert np.abs(got - want).max() <= ulps(np.complex128, 64)
