"""The batched-ansatz and adjoint-gradient workloads at their full sizes.

* 512 hardware-efficient-ansatz circuits (n ~ U{8..22}, 4..16 layers of RY
  and RZ on every qubit then a CZ ladder, seed 20261204), complex64, in one
  packed buffer with the tile width, small-circuit cut-off and fusion width
  the autotuner picks on the MI355X (2^12 tiles, circuits up to 14 qubits
  whole in 128 KiB of LDS, fuse_k 2).  A seeded subset holding every qubit
  count 8..22 is checked against the oracle's per-circuit simulate(), every
  circuit's FP64 norm against 1, and graph replays against the eager run.
* The RY/RZ/RZZ ansatz with 710 parameters at n=24 (seed 20261205) and the
  same structure at n=30: the adjoint gradient against central differences
  on the GPU and against the oracle's CPU parameter-shift gradient.
"""
import numpy as np
import pytest

import oracle
from qclabpp_amd import engine as E
from qclabpp_amd import observables as O
from qclabpp_amd import workloads as W
from tests.parity import C64, rel_err

pytestmark = pytest.mark.gpu

TUNED_C64 = {"tile_bits": 12, "small_bits": 14, "fuse_k": 2}


def _zero(n):
    v = np.zeros(1 << n, np.complex128)
    v[0] = 1
    return v


def test_hea_batch_512_c64():
    import torch

    from qclabpp_amd import batch as B
    spec = W.hea_batch(512)
    assert min(spec.nq) == 8 and max(spec.nq) == 22
    b = B.Batch(spec.circuits, spec.nq, dtype="c64", **TUNED_C64)
    print(f"batch 512: passes={b.info['npasses']} levels={b.info['nlevels']} "
          f"small={b.info['nsmall']} amplitudes={b.info['total_amplitudes']}", flush=True)
    assert b.info["nsmall"] > 0  # circuits that run whole in LDS
    eager = b.new_buffer()
    b.run(eager, use_graph=False)
    torch.cuda.synchronize()
    # every circuit's FP64 norm
    norms = np.array([O.norm2(b.state(eager, c)) for c in range(len(spec.nq))])
    assert np.abs(norms - 1).max() < 1e-5, np.abs(norms - 1).max()
    # graph replays equal the eager run, with eager work between them
    packed = b.new_buffer()
    for rep in range(3):
        packed.zero_()
        b.run(packed, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(packed, eager), rep
    # the oracle on the smallest circuit of every qubit count 8..22 (the
    # first of each in circuit order), n=13-14 included (whole in LDS)
    first = {}
    for c, n in enumerate(spec.nq):
        first.setdefault(n, c)
    assert sorted(first) == list(range(8, 23))
    host = eager.cpu().numpy()
    worst = 0.0
    for n, c in sorted(first.items()):
        want = oracle.apply(n, spec.circuits[c], _zero(n))
        got = host[int(b.offsets[c]):int(b.offsets[c + 1])]
        err = rel_err(got, want)
        print(f"  circuit {c}: n={n} layers={spec.layers[c]} rel err {err:.2e}", flush=True)
        assert err <= C64, (c, n, err)
        worst = max(worst, err)
    lat = b.latencies()
    assert np.all(lat > 0) and np.all(np.isfinite(lat))


def _params_by_kind(circ, counts, seed):
    """Seeded parameter indices: counts[k] of each gate kind k (RY, RZ, RZZ)."""
    rs = np.random.default_rng(seed)
    kinds = {}
    for g in circ:
        if g.param >= 0:
            kinds.setdefault(g.name, []).append(g.param)
    assert sorted(kinds) == ["RotationY", "RotationZ", "RotationZZ"], sorted(kinds)
    out = []
    for k in sorted(kinds):
        out += [int(p) for p in rs.choice(kinds[k], counts[k], replace=False)]
    return sorted(out)


def test_adjoint_tfim_24_710_params():
    """n=24, 10 layers, 710 parameters (complex128): the adjoint gradient
    against central differences (h=1e-6, 18 parameters over RY, RZ and
    RZZ, 1e-8) and against the oracle's CPU parameter shift (4 parameters:
    an RY, an RZ and two RZZ; exact up to rounding, 1e-12)."""
    import torch

    from qclabpp_amd.autograd import finite_difference
    n, L = 24, 10
    circ = W.tfim_ansatz(n, L)
    theta = np.array(W.params_of(circ))
    assert theta.size == 710
    H = O.tfim(n)
    psi = E.new_state(n)
    lam = torch.empty_like(psi)
    e, g = O.adjoint_gradient(psi, lam, circ, H)
    back = psi.cpu().numpy()
    assert abs(back[0] - 1) < 1e-11 and np.abs(back[1:]).max() < 1e-11  # psi rewound
    del psi, lam
    fd_which = _params_by_kind(circ, {"RotationY": 6, "RotationZ": 6, "RotationZZ": 6}, seed=1)
    fd = finite_difference(circ, H, theta, n, h=1e-6, which=fd_which)
    err_fd = float(np.abs(g[fd_which] - fd).max())
    ps_which = _params_by_kind(circ, {"RotationY": 1, "RotationZ": 1, "RotationZZ": 2}, seed=2)
    print(f"adjoint n=24: E={e:.12f} |adj-FD|={err_fd:.2e} ps params {ps_which}", flush=True)
    assert err_fd < 1e-8
    ps = oracle.param_shift(n, circ, H, theta, ps_which)
    err_ps = float(np.abs(g[ps_which] - ps).max())
    print(f"adjoint n=24: |adj-PS|={err_ps:.2e}", flush=True)
    assert err_ps < 1e-12
    assert abs(e - oracle.energy(n, circ, H, theta)) < 1e-11


def test_adjoint_tfim_30():
    """The same ansatz at n=30 (890 parameters, two 16 GiB states): two
    parameters, the first RY and the last RZZ, against central differences
    on the GPU, and psi rewound to |0...0> by the reverse sweep."""
    import torch

    from qclabpp_amd.autograd import finite_difference
    n, L = 30, 10
    circ = W.tfim_ansatz(n, L)
    theta = np.array(W.params_of(circ))
    H = O.tfim(n)
    psi = E.new_state(n)
    lam = torch.empty_like(psi)
    e, g = O.adjoint_gradient(psi, lam, circ, H)
    amp0 = complex(psi[0].item())
    rest = float(psi[1:].abs().max().item())
    del psi, lam
    torch.cuda.empty_cache()
    which = [0, theta.size - 1]
    fd = finite_difference(circ, H, theta, n, h=1e-6, which=which)
    err = float(np.abs(g[which] - fd).max())
    print(f"adjoint n=30: E={e:.12f} |adj-FD|={err:.2e} amp0={amp0} rest={rest:.1e}", flush=True)
    assert err < 1e-8
    assert abs(amp0 - 1) < 1e-10 and rest < 1e-10
    torch.cuda.empty_cache()
