"""The reference's own known-answer vectors through the HIP engine.

Same 753 cases as tests/test_oracle_kat.py; the engine computes with its own
complex FMA order, so equality is to 1e-14 absolute (values up to ~360).
"""
import numpy as np
import pytest

from qclabpp_amd import engine as E
from tests.katcases import load

pytestmark = pytest.mark.gpu

CASES, CIRCUITS = load()


def test_gate_kats_gpu():
    import torch
    print(f"{len(CASES)} known-answer cases", flush=True)
    bad = []
    for gate, op, n, vin, vexp, exact, where in CASES:
        for dt, tol in ((torch.complex128, 1e-14), (torch.complex64, 2e-5)):
            d = torch.from_numpy(vin).to(dt).cuda()
            E.apply_gate(d, gate, op)
            got = d.cpu().numpy()
            scale = max(1.0, np.abs(vexp).max())
            if np.abs(got - vexp).max() > tol * scale:
                bad.append((where, gate.name, gate.qubits, op, str(dt)))
    assert not bad, bad[:10]


def test_circuit_kats_gpu():
    import torch
    for gates, op, n, vin, vexp, where in CIRCUITS:
        if op != "N":
            continue  # reverse-order circuit apply is exercised via the adjoint path
        d = torch.from_numpy(vin).cuda()
        E.apply_gates(d, gates)
        got = d.cpu().numpy()
        assert np.abs(got - vexp).max() < 1e-14, where
