"""Pin the oracle on the reference's own known-answer vectors.

753 host vector-apply checks of test/qgates/*.cpp (700 asserted with ==,
the rest with 10-100 eps helpers) and the circuit checks of
test/QCircuit.cpp:163-236 (EXPECT_NEAR at 10 eps; the expected values there
are printed to 15 digits, so 1e-14 is used).
"""
import numpy as np

import oracle
from tests.katcases import load

CASES, CIRCUITS = load()


def test_case_count():
    assert len(CASES) == 753


def test_gate_kats():
    bad = []
    for gate, op, n, vin, vexp, exact, where in CASES:
        got = oracle.apply_gate(n, gate, vin, op)
        if exact:
            ok = np.array_equal(got, vexp)
        else:
            ok = np.abs(got - vexp).max() <= 100 * np.finfo(np.float64).eps
        if not ok:
            bad.append((where, gate.name, gate.qubits, op, n))
    assert not bad, bad[:10]


def test_gate_kats_complex64():
    """The tests are also instantiated for complex<float>."""
    bad = []
    for gate, op, n, vin, vexp, exact, where in CASES:
        got = oracle.apply_gate(n, gate, vin.astype(np.complex64), op)
        if np.abs(got - vexp).max() > 100 * np.finfo(np.float32).eps * max(1, np.abs(vexp).max()):
            bad.append((where, gate.name, op))
    assert not bad, bad[:10]


def test_circuit_kats():
    for gates, op, n, vin, vexp, where in CIRCUITS:
        got = oracle.apply(n, gates, vin, op)
        assert np.abs(got - vexp).max() < 1e-14, where
