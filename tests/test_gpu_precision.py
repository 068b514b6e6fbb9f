"""complex64 path: fused simulation error against complex128 and the oracle,
and FP64 reductions of complex64 states (norm, <Z_i Z_{i+1}>) against numpy
float64 sums of the same amplitudes."""
import numpy as np
import pytest

import oracle
from qclabpp_amd import engine as E
from qclabpp_amd import observables as O
from qclabpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _zz_numpy(n, psi):
    """<Z_i Z_{i+1}> of a host state in float64 (qubit q <-> bit n-1-q)."""
    p = np.abs(psi.astype(np.complex128)) ** 2
    idx = np.arange(1 << n, dtype=np.uint64)
    out = []
    for i in range(n - 1):
        b0 = (idx >> np.uint64(n - 1 - i)) & np.uint64(1)
        b1 = (idx >> np.uint64(n - 2 - i)) & np.uint64(1)
        sign = 1.0 - 2.0 * (b0 ^ b1).astype(np.float64)
        out.append(float(np.dot(sign, p)))
    return np.array(out)


@pytest.mark.parametrize("n", [14, 20])
def test_c64_against_c128_and_oracle(n):
    import torch
    circ = W.brickwork(n, 20, seed=W.SEED_PREC)
    states = {}
    for dt, tdt in (("c128", torch.complex128), ("c64", torch.complex64)):
        psi = E.new_state(n, tdt)
        E.Program(n, circ, fuse_k=5, dtype=dt).run(psi)
        torch.cuda.synchronize()
        states[dt] = psi
    ref0 = np.zeros(1 << n, np.complex128)
    ref0[0] = 1
    want = oracle.apply(n, circ, ref0)
    h128 = states["c128"].cpu().numpy()
    h64 = states["c64"].cpu().numpy()
    assert np.abs(h128 - want).max() < 1e-12
    # complex64 after 20 layers of fused blocks: a few hundred ulps of 2^-24
    d64 = np.abs(h64.astype(np.complex128) - want).max()
    assert d64 < 2e-5, d64
    # FP64 reductions on the complex64 state match float64 sums of the same amplitudes
    nrm_gpu = O.norm2(states["c64"])
    nrm_np = float(np.sum(np.abs(h64.astype(np.complex128)) ** 2))
    assert abs(nrm_gpu - nrm_np) < 1e-12 * max(1.0, nrm_np)
    assert abs(nrm_gpu - 1.0) < 1e-5
    vals, tot = O.expectation(states["c64"], O.zz_chain(n))
    zz_np = _zz_numpy(n, h64)
    assert np.abs(vals - zz_np).max() < 1e-12
    assert abs(tot - zz_np.sum()) < 1e-11
    vals128, _ = O.expectation(states["c128"], O.zz_chain(n))
    assert np.abs(vals - vals128).max() < 1e-5


def test_c64_norm_is_fp64_accumulated():
    """A uniform complex64 state over 2^24 amplitudes: every |a|^2 is 2^-24,
    so an FP32 running sum would stall far below 1; the FP64 reduction is
    exact."""
    import torch
    n = 24
    psi = torch.full((1 << n,), complex(2.0 ** (-n / 2), 0.0), dtype=torch.complex64, device="cuda")
    assert O.norm2(psi) == 1.0
    vals, _ = O.expectation(psi, O.zz_chain(n))
    assert np.abs(vals).max() == 0.0
