"""Full-size states checked through size-independent properties: the
30-qubit complex128 brickwork and its inverse return |0...0> in every
compiled form, as does the 33-qubit complex64 one through compiled passes; a 33-qubit
complex64 state (64 GiB, sweeps split over several launches) follows closed
forms after Hadamard layers, diagonal gates, merged diagonals and Pauli
products, with FP64 norms and expectation values."""
import cmath
import math

import numpy as np
import pytest

from qclabpp_amd import engine as E
from qclabpp_amd import gates as G
from qclabpp_amd import observables as O
from qclabpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _free():
    import torch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("fuse_k,depth", [(5, 20), (4, 20), (0, 3), (E.TILED, 20), (E.TILED_JIT, 20),
                                           (E.TILED_JIT_FIXED, 20)])
def test_brickwork30_c128_inverse_returns_to_zero(fuse_k, depth):
    import torch
    n = 30
    circ = W.brickwork(n, depth)
    fwd = E.Program(n, circ, fuse_k=fuse_k)
    back = E.Program(n, G.inverse(circ), fuse_k=fuse_k)
    if fuse_k in (E.TILED_JIT, E.TILED_JIT_FIXED):  # compiled passes that relabel qubits, at full size
        assert fwd.info()["tiled_relabel"] > 0 and back.info()["tiled_relabel"] > 0
    psi = E.new_state(n, torch.complex128)
    fwd.run(psi)
    torch.cuda.synchronize()
    assert abs(psi[0].item()) < 0.5           # the state moved away from |0...0>
    assert abs(O.norm2(psi) - 1.0) < 1e-12
    back.run(psi)
    torch.cuda.synchronize()
    assert abs(psi[0].item() - 1.0) < 1e-11
    assert abs(O.norm2(psi) - 1.0) < 1e-12
    del psi
    _free()


def test_brickwork33_c64_compiled_inverse_returns_to_zero():
    """64 GiB complex64 state through relabeling compiled passes (2^13 tiles)
    and back; FP64 norms."""
    import torch
    n = 33
    circ = W.brickwork(n, 20, seed=W.SEED_PREC)
    fwd = E.Program(n, circ, fuse_k=E.TILED_JIT, dtype="c64")
    back = E.Program(n, G.inverse(circ), fuse_k=E.TILED_JIT, dtype="c64")
    assert fwd.info()["tiled_relabel"] > 0
    psi = E.new_state(n, torch.complex64)
    fwd.run(psi)
    torch.cuda.synchronize()
    assert abs(psi[0].item()) < 0.5
    assert abs(O.norm2(psi) - 1.0) < 1e-5
    back.run(psi)
    torch.cuda.synchronize()
    assert abs(psi[0].item() - 1.0) < 1e-4
    assert abs(O.norm2(psi) - 1.0) < 1e-5
    del psi
    _free()


def test_state33_c64_closed_forms():
    import torch
    n = 33
    N = 1 << n
    psi = E.new_state(n, torch.complex64)      # 64 GiB: set_basis spans 8 launches
    assert psi[0].item() == 1 and psi[N - 1].item() == 0 and psi[N // 2 + 7].item() == 0
    assert O.norm2(psi) == 1.0
    # H on every qubit (dense sweeps, lane-exchange form on the low bits)
    E.Program(n, [G.Hadamard(q) for q in range(n)], fuse_k=0, dtype="c64").run(psi)
    amp = 2.0 ** (-n / 2)
    th, ph = 0.8, -1.3
    # RZ on qubit 0 (bit 32: a diagonal sweep over two launches), Phase on qubit 32
    E.Program(n, [G.RotationZ(0, th), G.Phase(n - 1, ph)], fuse_k=0, dtype="c64").run(psi)
    torch.cuda.synchronize()
    for x in (0, 1, N // 2, N // 2 + 1, N - 2, N - 1, 123456789):
        hi, lo = (x >> (n - 1)) & 1, x & 1
        want = amp * cmath.exp(1j * (th / 2 if hi else -th / 2)) * (cmath.exp(1j * ph) if lo else 1)
        assert abs(psi[x].item() - want) < 1e-6 * amp, x
    assert abs(O.norm2(psi) - 1.0) < 1e-5
    # FP64 expectation values: <X_0> = cos(th), <X_32> = cos(ph), <Z_0 Z_1> = 0
    H = O.PauliSum().add(1.0, [(0, "X")]).add(1.0, [(n - 1, "X")]).add(1.0, [(0, "Z"), (1, "Z")])
    vals, _ = O.expectation(psi, H)
    assert abs(vals[0] - math.cos(th)) < 1e-5
    assert abs(vals[1] - math.cos(ph)) < 1e-5
    assert abs(vals[2]) < 1e-6
    # X_0 applied into a second state (the Pauli-product sweep spans 8 launches)
    out = torch.empty_like(psi)
    O.apply_pauli_sum(psi, out, O.PauliSum().add(1.0, [(0, "X")]))
    re, im = O.inner(psi, out)
    assert abs(re - math.cos(th)) < 1e-5 and abs(im) < 1e-6
    del out
    _free()
    # merged diagonal: a chain of CZ, RZZ, RZ, CPhase, RZZ over qubits {0,32,5,3,1}
    # becomes one phase-table sweep (two launches at this size)
    diag = [G.CZ(0, n - 1), G.RotationZZ(5, n - 1, 0.9), G.RotationZ(5, 0.3), G.CPhase(5, 3, 0.7),
            G.RotationZZ(1, 3, -0.4)]
    prog = E.Program(n, diag, fuse_k=5, dtype="c64")
    assert prog.info()["ndiag"] == 1 and prog.info()["nlaunch"] == 1
    xs = (0, N - 1, 0b1011 << 27, (1 << 32) | 1, (1 << 27) | (1 << 29), 987654321)
    before = {x: psi[x].item() for x in xs}
    prog.run(psi)
    torch.cuda.synchronize()
    for x, v in before.items():
        bit = lambda q: (x >> (n - 1 - q)) & 1  # noqa: E731
        f = (-1) ** (bit(0) & bit(n - 1))
        f *= cmath.exp(1j * (-0.45 if bit(5) == bit(n - 1) else 0.45))
        f *= cmath.exp(1j * (0.15 if bit(5) else -0.15))
        f *= cmath.exp(1j * 0.7) if bit(5) and bit(3) else 1
        f *= cmath.exp(1j * (0.2 if bit(1) == bit(3) else -0.2))
        assert abs(psi[x].item() - v * f) < 1e-6 * amp, x
    del psi
    _free()
