"""Batched circuits (LDS tile passes, graph replay, autotuner) against the
oracle, circuit by circuit."""
import numpy as np
import pytest

import oracle
from qclabpp_amd import batch as B
from qclabpp_amd import gates as G
from qclabpp_amd import workloads as W
from qclabpp_amd.rng import SplitMix64
from tests.gatecases import random_unitary
from tests.parity import C64, C128, rel_err

pytestmark = pytest.mark.gpu


def mixed_circuit(n, seed):
    rs = np.random.default_rng(seed)
    out = [G.Hadamard(q) for q in range(n)]
    for i in range(3 * n):
        a, b = (int(x) for x in rs.choice(n, 2, replace=False)) if n > 1 else (0, 0)
        th = float(rs.uniform(-3, 3))
        out += [G.RotationY(a, th), G.RotationZ(b, -th), G.Phase45(a)]
        if n > 1:
            out += [G.CX(a, b), G.CZ(a, b), G.CPhase(b, a, th), G.SWAP(a, b), G.iSWAP(a, b),
                    G.CRotationX(a, b, th, controlState=0), G.RotationZZ(min(a, b), max(a, b), th),
                    G.RotationXX(min(a, b), max(a, b), -th),
                    G.QGate2(min(a, b), max(a, b), random_unitary(4, seed * 100 + i))]
    return out


def zero(n, dtype):
    v = np.zeros(1 << n, dtype)
    v[0] = 1
    return v


def check_all(batch, packed, circuits, nq, tol, which=None):
    """Every circuit (or those listed in `which`) against the oracle,
    relative to the largest expected amplitude."""
    import torch
    torch.cuda.synchronize()
    host = packed.cpu().numpy()
    worst = 0.0
    for c in (range(len(circuits)) if which is None else which):
        n = nq[c]
        want = oracle.apply(n, circuits[c], zero(n, np.complex128))
        got = host[int(batch.offsets[c]):int(batch.offsets[c + 1])]
        err = rel_err(got, want)
        assert err <= tol, (c, n, err)
        worst = max(worst, err)
    return worst


@pytest.mark.parametrize("dtype,tol", [("c128", C128), ("c64", C64)])
@pytest.mark.parametrize("fuse_k", [0, 1, 2])
@pytest.mark.parametrize("tile_bits,small_bits", [(6, 6), (8, 10), (10, 13)])
def test_mixed_batch_vs_oracle(dtype, tol, fuse_k, tile_bits, small_bits):
    print(f"batch {dtype} fuse_k={fuse_k} tile={tile_bits} small={small_bits}", flush=True)
    nq = [1, 2, 3, 5, 7, 9, 11, 12, 14, 15, 8, 6, 13, 4, 10]
    circuits = [mixed_circuit(n, 17 + i) for i, n in enumerate(nq)]
    b = B.Batch(circuits, nq, dtype=dtype, tile_bits=tile_bits, small_bits=small_bits, fuse_k=fuse_k)
    packed = b.new_buffer()
    b.run(packed, use_graph=False)
    check_all(b, packed, circuits, nq, tol)


@pytest.mark.parametrize("dtype,tol", [("c128", C128), ("c64", C64)])
@pytest.mark.parametrize("fuse_k", [1, 2])
@pytest.mark.parametrize("small_bits", [6, 11, 13])
def test_register_phase_batch_vs_oracle(dtype, tol, fuse_k, small_bits):
    """phases=1: circuits above max(small_bits, 11) qubits run register-phase
    passes of 2^12-amplitude tiles (first pass from |0...0> without a read),
    the rest whole in LDS; the circuits of 12 to 15 qubits take one to
    several passes each."""
    print(f"batch phases {dtype} fuse_k={fuse_k} small={small_bits}", flush=True)
    nq = [1, 2, 3, 5, 7, 9, 11, 12, 14, 15, 8, 6, 13, 4, 10, 16]
    circuits = [mixed_circuit(n, 31 + i) for i, n in enumerate(nq)]
    b = B.Batch(circuits, nq, dtype=dtype, small_bits=small_bits, fuse_k=fuse_k, phases=1)
    assert b.info["phases"] == 1
    assert b.info["nsmall"] == sum(1 for n in nq if n <= max(small_bits, 11))
    packed = b.new_buffer()
    packed.fill_(float("nan"))               # every amplitude is written
    b.run(packed, use_graph=False)
    check_all(b, packed, circuits, nq, tol)


@pytest.mark.parametrize("phases", [0, 1])
def test_graph_replays_match_eager(phases):
    import torch
    spec = W.hea_batch(48, seed=7, nmin=4, nmax=17, lmin=2, lmax=6)
    b = B.Batch(spec.circuits, spec.nq, dtype="c64", tile_bits=10, small_bits=12, fuse_k=1,
                phases=phases)
    eager = b.new_buffer()
    b.run(eager, use_graph=False)
    torch.cuda.synchronize()
    ref = eager.clone()
    packed = b.new_buffer()
    for rep in range(4):
        packed.zero_()                       # eager work between the replays
        b.run(packed, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(packed, ref), rep
    check_all(b, packed, spec.circuits, spec.nq, C64)
    lat = b.latencies()
    assert np.all(lat > 0) and np.all(np.isfinite(lat))


def test_autotuner_candidates_agree():
    import torch
    spec = W.hea_batch(24, seed=3, nmin=6, nmax=16, lmin=2, lmax=5)
    nq = spec.nq
    total = sum(1 << n for n in nq)
    packed = torch.empty(total, dtype=torch.complex128, device="cuda")
    scratch = torch.empty_like(packed)
    best, results = B.autotune(spec.circuits, nq, packed, scratch, dtype="c128", reps=1)
    # tile bits 10..13 (complex128 caps at 13) x fuse_k 0..2, then register-phase
    # passes: small_bits 13 and 11 x fuse_k 1..2
    assert len(results) == 16
    assert sum(r["phases"] for r in results) == 4
    assert max(r["max_abs_diff"] for r in results) < 1e-12
    b = B.Batch(spec.circuits, nq, dtype="c128", **best)
    out = b.new_buffer()
    b.run(out)
    check_all(b, out, spec.circuits, nq, C128)
