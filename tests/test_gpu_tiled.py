"""GPU parity of tiled programs (fuse_k = QSV_TILED): passes over LDS tiles
of 2^12 amplitudes, gates applied in registers phase by phase.  Checked
against the oracle on mixed circuits with every gate class, both control
states, Op N/C and every qubit, and at full size through the inverse."""
import numpy as np
import pytest

import oracle
from qclabpp_amd import engine as E
from qclabpp_amd import gates as G
from qclabpp_amd import workloads as W
from qclabpp_amd.rng import normal_state
from tests.gatecases import one_qubit_gates, random_state, two_qubit_gates

pytestmark = pytest.mark.gpu


def run(prog, psi0):
    import torch
    d = torch.from_numpy(psi0.copy()).cuda()
    prog.run(d)
    torch.cuda.synchronize()
    return d.cpu().numpy()


def every_kind(n, seed):
    """All gate classes on spread-out qubit pairs and every single qubit,
    with ConjTrans copies, shuffled deterministically."""
    rs = np.random.default_rng(seed)
    circ = []
    for q in range(n):
        circ += one_qubit_gates(q, seed)
    pairs = [(0, n - 1), (n - 1, 0), (1, 2), (n - 2, n - 1), (3, n // 2), (n // 2, 4), (5, 6)]
    for a, b in pairs:
        circ += two_qubit_gates(a, b, seed)
    # unitary gates only: repeated non-unitary matrices would blow the norm up
    circ = [g for g in circ
            if g.mat is None or np.allclose(g.mat @ g.mat.conj().T, np.eye(len(g.mat)))]
    circ += [g.__class__(**{**g.__dict__, "op": "C"}) for g in circ[::3]]
    order = rs.permutation(len(circ))
    return [circ[i] for i in order]


@pytest.mark.parametrize("n", [12, 13, 15, 18])
def test_tiled_every_kind_c128(n):
    print(f"tiled every kind n={n}", flush=True)
    circ = every_kind(n, n)
    psi = random_state(n, 5, np.complex128)
    want = oracle.apply(n, circ, psi)
    prog = E.Program(n, circ, fuse_k=E.TILED)
    got = run(prog, psi)
    assert np.abs(got - want).max() < 1e-12


@pytest.mark.parametrize("n", [13, 17])
def test_tiled_every_kind_c64(n):
    circ = every_kind(n, 2 * n)
    psi = random_state(n, 6, np.complex128)
    want = oracle.apply(n, circ, psi)
    got = run(E.Program(n, circ, fuse_k=E.TILED, dtype="c64"), psi.astype(np.complex64))
    assert np.abs(got - want).max() < 5e-5


@pytest.mark.parametrize("n", [16, 20])
def test_tiled_brickwork_and_ansatz(n):
    circ = W.brickwork(n, 20, seed=7) + W.tfim_ansatz(n, 2)
    psi = random_state(n, 9, np.complex128)
    want = oracle.apply(n, circ, psi)
    prog = E.Program(n, circ, fuse_k=E.TILED)
    assert prog.info()["nlaunch"] >= 2
    got = run(prog, psi)
    assert np.abs(got - want).max() < 1e-12


def test_tiled_qft20_is_dft():
    n = 20
    x = normal_state(W.SEED_QFT, n)
    want = np.fft.ifft(x) * np.sqrt(1 << n)
    got = run(E.Program(n, W.qft(n), fuse_k=E.TILED), x)
    assert np.abs(got - want).max() < 1e-12
