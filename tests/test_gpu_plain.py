"""GPU parity of the plain per-gate kernels against the oracle.

Every gate class x Op x qubit layout on seeded random states, through the
C ABI; tolerances are written per dtype below.
"""
import sys

import numpy as np
import pytest

import oracle
from qclabpp_amd import engine as E
from qclabpp_amd import workloads as W
from qclabpp_amd import gates as G
from tests.gatecases import all_cases, one_qubit_gates, random_state, two_qubit_gates
from tests.parity import C128, assert_close, rel_err

pytestmark = pytest.mark.gpu

# Max |amplitude difference| allowed per gate, relative to the largest
# amplitude of the expected state (one gate: a few roundings).
TOL = {np.complex128: 1e-13, np.complex64: 2e-6}


def torch_dtype(dt):
    import torch
    return torch.complex128 if dt == np.complex128 else torch.complex64


def to_dev(v):
    import torch
    return torch.from_numpy(v.copy()).cuda()


def from_dev(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [2, 3, 5])
def test_every_gate_small(n, dtype):
    print(f"plain gates n={n} {dtype.__name__}", flush=True)
    psi = random_state(n, 10 + n, dtype)
    for gate, op in all_cases(n):
        want = oracle.apply_gate(n, gate, psi, op)
        d = to_dev(psi)
        E.apply_gate(d, gate, op)
        got = from_dev(d)
        err = rel_err(got, want)
        assert err <= TOL[dtype], (gate.name, gate.qubits, op, err)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_every_qubit_large(dtype):
    """n=14: all 14 qubits as target (both lane and pair kernels), a spread
    of control/target pairs, multi-block grids.  Pairs with a bit inside one
    128-byte line (bits 0-2 for complex128, 0-3 for complex64) take the
    whole-line variants of the half- and quarter-touch sweeps."""
    n = 14
    print(f"plain gates n={n} {dtype.__name__}", flush=True)
    psi = random_state(n, 99, dtype)
    pairs = [(0, 13), (13, 0), (12, 13), (13, 12), (3, 9), (9, 3), (0, 1), (7, 8), (2, 12),
             (10, 2), (2, 10), (11, 4), (4, 11)]
    cases = [(g, op) for q in range(n) for g in one_qubit_gates(q) for op in "NC"]
    cases += [(g, op) for a, b in pairs for g in two_qubit_gates(a, b) for op in "NC"]
    for gate, op in cases:
        want = oracle.apply_gate(n, gate, psi, op)
        d = to_dev(psi)
        E.apply_gate(d, gate, op)
        got = from_dev(d)
        err = rel_err(got, want)
        assert err <= TOL[dtype], (gate.name, gate.qubits, op, err)


def test_brickwork_16_vs_oracle():
    n = 16
    circ = W.brickwork(n, 20, seed=4242)
    psi0 = np.zeros(1 << n, np.complex128)
    psi0[0] = 1
    want = oracle.apply(n, circ, psi0)
    d = E.new_state(n)
    E.apply_gates(d, circ)
    got = from_dev(d)
    assert_close(got, want, C128, "brickwork n=16")


def test_qft_20_is_dft():
    """QFT with the + sign maps x to sqrt(N) * ifft(x) (closed form)."""
    from qclabpp_amd.rng import normal_state
    n = 20
    x = normal_state(W.SEED_QFT, n)
    d = to_dev(x)
    E.apply_gates(d, W.qft(n, +1))
    got = from_dev(d)
    want = np.fft.ifft(x) * np.sqrt(1 << n)
    assert_close(got, want, C128, "QFT n=20")


def test_host_vector_entry():
    n = 10
    circ = W.qft(n) + W.brickwork(n, 4, seed=3)
    psi = random_state(n, 1, np.complex128)
    got = E.simulate_host(psi, circ)
    want = oracle.apply(n, circ, psi)
    assert_close(got, want, C128, "host entry")


@pytest.mark.parametrize("n,dt", [(24, np.complex128), (25, np.complex64), (13, np.complex64)])
def test_host_entry_staged_transfers(n, dt):
    """The std::vector entry moves the state through pinned double buffers of
    64 MiB: several chunks (n=24 complex128: 4; n=25 complex64: 4) and one
    partial chunk (n=13).  Same result as the device-resident program."""
    import ctypes as C

    import torch

    from qclabpp_amd import _lib
    from qclabpp_amd import gates as G
    circ = W.brickwork(n, 3, seed=5)
    x = random_state(n, 7, np.complex128).astype(dt)
    h = x.copy()
    arr, cnt = G.records(circ)
    code = _lib.QSV_C128 if dt == np.complex128 else _lib.QSV_C64
    _lib.check(_lib.lib().qsv_host_simulate(h.ctypes.data_as(C.c_void_p), n, code, arr, cnt, 5))
    d = torch.from_numpy(x.copy()).cuda()
    E.Program(n, circ, fuse_k=5, dtype="c128" if dt == np.complex128 else "c64").run(d)
    torch.cuda.synchronize()
    assert np.array_equal(h, d.cpu().numpy())
