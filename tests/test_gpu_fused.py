"""GPU parity of compiled programs (plain and fused, MFMA dense blocks and
merged diagonals) against the oracle and closed forms."""
import numpy as np
import pytest

import oracle
from qclabpp_amd import engine as E
from qclabpp_amd import gates as G
from qclabpp_amd import workloads as W
from qclabpp_amd.rng import normal_state
from tests.gatecases import random_state, random_unitary
from tests.parity import C64, C128, assert_close

pytestmark = pytest.mark.gpu


def run(prog, psi0):
    import torch
    d = torch.from_numpy(psi0.copy()).cuda()
    prog.run(d)
    torch.cuda.synchronize()
    return d.cpu().numpy()


def mixed(n, count, seed):
    rs = np.random.default_rng(seed)
    out = []
    for i in range(count):
        a, b = rs.choice(n, 2, replace=False)
        th = rs.uniform(-3, 3)
        out += [G.RotationY(a, th), G.CX(a, b), G.RotationZ(b, -th), G.CPhase(b, a, th),
                G.SWAP(a, b), G.RotationZZ(min(a, b), max(a, b), th), G.Hadamard(b),
                G.CRotationX(a, b, th, controlState=0), G.iSWAP(a, b), G.PauliY(a),
                G.Phase45(b), G.QGate2(min(a, b), max(a, b), random_unitary(4, i)),
                G.RotationX(rs.integers(n), th), G.CZ(a, b)]
    return out


@pytest.mark.parametrize("k", [0, 3, 4, 5])
@pytest.mark.parametrize("n", [7, 12, 18])
def test_programs_vs_oracle_c128(n, k):
    print(f"program n={n} k={k}", flush=True)
    circ = W.brickwork(n, 12, seed=11) + mixed(n, 10, seed=n) + W.tfim_ansatz(n, 1)
    psi = random_state(n, 3, np.complex128)
    want = oracle.apply(n, circ, psi)
    got = run(E.Program(n, circ, fuse_k=k), psi)
    assert_close(got, want, C128, f"n={n} k={k}")


@pytest.mark.parametrize("k", [0, 4, 5])
def test_programs_c64(k):
    n = 16
    circ = W.brickwork(n, 10, seed=12)
    psi = random_state(n, 4, np.complex128)
    want = oracle.apply(n, circ, psi)
    got = run(E.Program(n, circ, fuse_k=k, dtype="c64"), psi.astype(np.complex64))
    assert_close(got, want, C64, f"c64 k={k}")


def test_every_block_qubit_layout():
    """Dense blocks on every window of 5 qubits, including the lowest bits."""
    n = 14
    circ = []
    for lo in range(n - 4):
        qs = list(range(lo, lo + 5))
        for q in qs:
            circ.append(G.RotationX(q, 0.1 * q + lo))
        for a, b in zip(qs, qs[1:]):
            circ.append(G.CX(a, b))
    circ += [G.RotationY(q, 0.3) for q in (0, 3, 7, 11, 13)] + [G.CX(0, 13), G.CX(3, 7)]
    psi = random_state(n, 8, np.complex128)
    want = oracle.apply(n, circ, psi)
    for k in (3, 4, 5):
        got = run(E.Program(n, circ, fuse_k=k), psi)
        assert_close(got, want, C128, f"k={k}")


def test_fused_qft20_is_dft():
    n = 20
    x = normal_state(W.SEED_QFT, n)
    want = np.fft.ifft(x) * np.sqrt(1 << n)
    for k in (4, 5):
        got = run(E.Program(n, W.qft(n), fuse_k=k), x)
        assert_close(got, want, C128, f"QFT k={k}")
