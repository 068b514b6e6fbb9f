"""GPU observables and the adjoint gradient against the oracle, closed forms,
parameter shift and central finite differences."""
import math

import numpy as np
import pytest

import oracle
from qclabpp_amd import engine as E
from qclabpp_amd import gates as G
from qclabpp_amd import observables as O
from qclabpp_amd import workloads as W
from tests.gatecases import random_state

pytestmark = pytest.mark.gpu


def dev(v):
    import torch
    return torch.from_numpy(v.copy()).cuda()


def mixed_H(n):
    H = O.tfim(n)
    H.add(0.3, [(1, "Y"), (3, "X"), (n - 1, "Z")]).add(-1.1, [(0, "Y")])
    H.add(0.5, [(2, "X"), (n - 2, "Y")]).add(0.25, [(0, "Z"), (2, "Z"), (4, "Z")])
    return H


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 2e-5)])
@pytest.mark.parametrize("n", [6, 14, 20])
def test_expectation_vs_oracle(n, dtype, tol):
    """Per-term <psi|P|psi> against the oracle; at n=20 the diagonal terms on
    qubits 0 and 1 (address bits 18 and 19) vary along each thread's
    amplitudes, the others do not."""
    print(f"expectation n={n} {dtype.__name__}", flush=True)
    psi = random_state(n, 7, np.complex128)
    H = mixed_H(n)
    want = oracle.expectation(n, psi, H)
    vals, tot = O.expectation(dev(psi.astype(dtype)), H)
    assert np.abs(vals - want).max() < tol
    assert abs(tot - sum(c * v for (c, _), v in zip(H.terms, want))) < 10 * tol


def test_apply_pauli_sum_vs_matrix():
    import torch
    n = 6
    psi = random_state(n, 3)
    H = mixed_H(n)
    out = torch.empty(1 << n, dtype=torch.complex128, device="cuda")
    O.apply_pauli_sum(dev(psi), out, H)
    assert np.abs(out.cpu().numpy() - H.matrix(n) @ psi).max() < 1e-14


def single_flip_H(n):
    """Terms that each flip at most one qubit (the tiled H|psi> path): the
    TFIM chain plus Y factors, Z strings and X on the last qubit."""
    H = O.tfim(n)
    H.add(0.3, [(1, "Y"), (n - 1, "Z")]).add(-1.1, [(0, "Y")]).add(0.4, [(n - 1, "X"), (0, "Z")])
    H.add(0.25, [(0, "Z"), (2, "Z"), (4, "Z")]).add(0.6, [(n // 2, "Y"), (n // 2 + 1, "Z")])
    return H


def oracle_apply_sum(n, H, psi):
    """sum_t c_t P_t psi with each Pauli factor applied as the oracle's gate."""
    ctor = {"X": G.PauliX, "Y": G.PauliY, "Z": G.PauliZ}
    out = np.zeros_like(psi)
    for c, fs in H.terms:
        out += c * oracle.apply(n, [ctor[p](q) for q, p in fs], psi)
    return out


@pytest.mark.parametrize("tiles", ["1", "0"])
@pytest.mark.parametrize("dt,tol", [("c128", 1e-13), ("c64", 2e-5)])
@pytest.mark.parametrize("n,kind", [(12, "single"), (14, "single"), (22, "single"), (14, "mixed")])
def test_apply_pauli_sum_vs_oracle(n, kind, dt, tol, tiles, monkeypatch):
    """H|psi> at 12 to 22 qubits against the oracle's gates.  With
    QSV_PAULI_TILES=1, sums whose terms flip at most one qubit take the tiled
    passes (one to three of them, every flip bit in exactly one), which
    otherwise serve states above 1 GiB; other sums, and every sum with =0,
    take the per-flip-mask gather."""
    import torch
    monkeypatch.setenv("QSV_PAULI_TILES", tiles)
    print(f"apply pauli sum n={n} {kind} {dt} tiles={tiles}", flush=True)
    H = single_flip_H(n) if kind == "single" else mixed_H(n)
    psi = random_state(n, 11 + n, np.complex128)
    want = oracle_apply_sum(n, H, psi)
    cdt = np.complex128 if dt == "c128" else np.complex64
    out = torch.full((1 << n,), float("nan"), dtype=torch.complex128 if dt == "c128" else torch.complex64,
                     device="cuda")
    O.apply_pauli_sum(dev(psi.astype(cdt)), out, H)
    got = out.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert np.abs(got - want).max() / np.abs(want).max() < tol


def test_inner_and_norm():
    n = 12
    a, b = random_state(n, 1), random_state(n, 2)
    re, im = O.inner(dev(a), dev(b))
    z = np.vdot(a, b)
    assert abs(re - z.real) < 1e-14 and abs(im - z.imag) < 1e-14
    assert abs(O.norm2(dev(a)) - 1) < 1e-14


@pytest.mark.parametrize("th", [0.3, -2.0])
def test_adjoint_single_ry_closed_form(th):
    import torch
    H = O.PauliSum().add(1.0, [(0, "Z")])
    psi = E.new_state(1)
    lam = torch.empty_like(psi)
    e, g = O.adjoint_gradient(psi, lam, [G.RotationY(0, th, param=0)], H)
    assert abs(e - math.cos(th)) < 1e-15 and abs(g[0] + math.sin(th)) < 1e-15


@pytest.mark.parametrize("fuse_k", [0, 5])
def test_adjoint_vs_param_shift_tfim(fuse_k):
    import torch
    n = 8
    circ = W.tfim_ansatz(n, 2, seed=99)
    theta = np.array(W.params_of(circ))
    H = O.tfim(n)
    psi = E.new_state(n)
    lam = torch.empty_like(psi)
    e, g = O.adjoint_gradient(psi, lam, circ, H, fuse_k=fuse_k)
    assert abs(e - oracle.energy(n, circ, H, theta)) < 1e-12
    ps = oracle.param_shift(n, circ, H, theta, list(range(theta.size)))
    assert np.abs(g - ps).max() < 1e-12
    # the reverse sweep rewinds psi to |0...0>
    back = psi.cpu().numpy()
    assert abs(back[0] - 1) < 1e-12 and np.abs(back[1:]).max() < 1e-12


@pytest.fixture
def adjoint_form(monkeypatch):
    """Select the reverse sweep: compiled two-state passes ("jit") or the
    interpreted LDS passes ("interp")."""
    def use(form):
        monkeypatch.setenv("QSV_ADJOINT", form)
    return use


@pytest.mark.parametrize("n,form", [(6, "interp"), (15, "interp"), (12, "jit"), (15, "jit")])
def test_adjoint_all_differentiable_kinds_vs_fd(n, form, adjoint_form):
    """Every differentiable kind against central differences, through both
    reverse sweeps.  At n=15 the interpreted sweep's tiles (2^11 amplitudes)
    and the compiled sweep's (2^12) leave qubits outside the tile, so
    generators and controls on address bits outside it are covered; at
    n=12 the compiled sweep's one tile is the whole state."""
    import torch
    from qclabpp_amd.autograd import finite_difference
    adjoint_form(form)
    print(f"adjoint kinds n={n} form={form}", flush=True)
    rs = np.random.default_rng(4)
    circ = [G.Hadamard(q) for q in range(n)]
    p = 0
    for ctor in (G.RotationX, G.RotationY, G.RotationZ, G.Phase):
        for q in range(n):
            circ.append(ctor(q, rs.uniform(-3, 3), param=p))
            p += 1
    for ctor in (G.RotationXX, G.RotationYY, G.RotationZZ):
        for q in range(n - 1):
            circ.append(ctor(q, q + 1, rs.uniform(-3, 3), param=p))
            p += 1
        circ.append(ctor(0, n - 1, rs.uniform(-3, 3), param=p))
        p += 1
    for ctor in (G.CRotationX, G.CRotationY, G.CRotationZ, G.CPhase):
        for cs in (1, 0):
            for c, t in ((2, 4), (5, 1), (0, n - 1), (n - 1, 3)):
                circ.append(ctor(c, t, rs.uniform(-3, 3), controlState=cs, param=p))
                p += 1
    circ.append(G.RotationY(0, 0.7, param=0))  # a shared parameter
    theta = np.array([next(g.theta for g in circ if g.param == i) for i in range(p)])
    H = mixed_H(n)
    psi = E.new_state(n)
    lam = torch.empty_like(psi)
    e, g = O.adjoint_gradient(psi, lam, circ, H, theta=theta, fuse_k=0)
    fd = finite_difference(circ, H, theta, n, h=1e-6, fuse_k=0)
    assert np.abs(g - fd).max() < 1e-8


@pytest.mark.parametrize("form", ["interp", "jit"])
def test_adjoint_pass_with_many_parametrised_gates_vs_fd(form, adjoint_form):
    """384 parametrised gates on 11 qubits of a 13-qubit state fit one tile
    pass of the reverse sweep: the interpreted sweep runs it in several
    launches of at most 128 gradient rows each, the compiled one in a single
    launch whose rows reduce over whole waves (one LDS slot per wave)."""
    import torch
    from qclabpp_amd.autograd import finite_difference
    adjoint_form(form)
    n = 13
    rs = np.random.default_rng(12)
    circ = [G.Hadamard(q) for q in range(n)]
    p = 0
    for _ in range(12):
        for q in range(2, n):
            circ.append(G.RotationY(q, rs.uniform(-3, 3), param=p))
            circ.append(G.RotationZ(q, rs.uniform(-3, 3), param=p + 1))
            p += 2
        for q in range(2, n - 1):
            circ.append(G.RotationZZ(q, q + 1, rs.uniform(-3, 3), param=p))
            p += 1
    assert p == 384
    theta = np.array([next(g.theta for g in circ if g.param == i) for i in range(p)])
    H = mixed_H(n)
    psi = E.new_state(n)
    lam = torch.empty_like(psi)
    e, g = O.adjoint_gradient(psi, lam, circ, H, theta=theta, fuse_k=0)
    fd = finite_difference(circ, H, theta, n, h=1e-6, fuse_k=0)
    assert abs(e - oracle.energy(n, circ, H, theta)) < 1e-11
    assert np.abs(g - fd).max() < 1e-8


def test_adjoint_rejects_non_differentiable_gate_before_touching_psi():
    import torch
    from qclabpp_amd._lib import QsvError
    n = 6
    h = G.Hadamard(2)
    h.param = 0
    circ = [G.RotationY(0, 0.4, param=1), h]
    psi0 = random_state(n, 21)
    psi = dev(psi0)
    lam = torch.empty_like(psi)
    with pytest.raises(QsvError, match="not differentiable"):
        O.adjoint_gradient(psi, lam, circ, O.tfim(n), fuse_k=0)
    assert np.array_equal(psi.cpu().numpy(), psi0)


def test_autograd_function():
    import torch
    from qclabpp_amd.autograd import energy
    n = 6
    circ = W.tfim_ansatz(n, 2, seed=5)
    theta = torch.tensor(W.params_of(circ), dtype=torch.float64, requires_grad=True)
    H = O.tfim(n)
    e = energy(theta, circ, H, n)
    e.backward()
    ps = oracle.param_shift(n, circ, H, theta.detach().numpy(), list(range(theta.numel())))
    assert abs(e.item() - oracle.energy(n, circ, H, theta.detach().numpy())) < 1e-12
    assert np.abs(theta.grad.numpy() - ps).max() < 1e-12


@pytest.mark.parametrize("n", [10, 13])
def test_adjoint_complex64_vs_complex128(n):
    """The reverse sweep's tile passes in FP32 arithmetic with FP64 inner
    products, against the complex128 gradient (n=13: full-size tiles of both
    precisions, 2^11 and 2^12 amplitudes)."""
    import torch
    print(f"adjoint c64 vs c128 n={n}", flush=True)
    circ = W.tfim_ansatz(n, 3, seed=9)
    H = O.tfim(n)
    out = {}
    for dt in (torch.complex128, torch.complex64):
        psi = E.new_state(n, dt)
        lam = torch.empty_like(psi)
        out[dt] = O.adjoint_gradient(psi, lam, circ, H, fuse_k=0)
    (e128, g128), (e64, g64) = out[torch.complex128], out[torch.complex64]
    assert abs(e64 - e128) < 1e-5
    assert np.abs(np.asarray(g64) - np.asarray(g128)).max() < 1e-4


@pytest.mark.parametrize("dt,tol", [("c128", 1e-12), ("c64", 2e-4)])
@pytest.mark.parametrize("n", [12, 14, 17])
def test_adjoint_compiled_matches_interpreted(n, dt, tol, adjoint_form):
    """The compiled reverse sweep (two states per pass, generator rows
    reduced per DPP row into LDS slots) against the interpreted one on every
    differentiable kind with fixed gates between them (permutations fold
    into the register map, so generators see relabeled registers); psi is
    rewound by both."""
    import torch
    from tests.test_jit_compile import differentiable_circuit
    print(f"adjoint compiled vs interpreted n={n} {dt}", flush=True)
    circ, p = differentiable_circuit(n, n)
    # the TFIM ansatz after it, its parameters shared with the first ones
    circ += W.tfim_ansatz(n, 2, seed=n)
    H = mixed_H(n)
    cdt = torch.complex128 if dt == "c128" else torch.complex64
    res = {}
    for form in ("interp", "jit"):
        adjoint_form(form)
        psi = E.new_state(n, cdt)
        lam = torch.empty_like(psi)
        e, g = O.adjoint_gradient(psi, lam, circ, H)
        prof = O.adjoint_profile()
        assert prof["compiled"] == (form == "jit") and prof["reverse_launches"] >= 1
        assert prof["forward_ms"] > 0 and prof["reverse_ms"] > 0
        back = psi.cpu().numpy()
        assert abs(back[0] - 1) < (1e-11 if dt == "c128" else 1e-4)
        res[form] = (e, np.asarray(g))
    (ei, gi), (ej, gj) = res["interp"], res["jit"]
    scale = max(1.0, float(np.abs(gi).max()))
    assert abs(ei - ej) < tol * max(1.0, abs(ei))
    assert np.abs(gi - gj).max() < tol * scale, np.abs(gi - gj).max()


@pytest.mark.parametrize("n", [12, 14])
def test_adjoint_compiled_vs_param_shift(n, adjoint_form):
    """The compiled reverse sweep against the oracle's CPU parameter shift
    on the TFIM ansatz (every parameter)."""
    import torch
    adjoint_form("jit")
    circ = W.tfim_ansatz(n, 2, seed=99)
    theta = np.array(W.params_of(circ))
    H = O.tfim(n)
    psi = E.new_state(n)
    lam = torch.empty_like(psi)
    e, g = O.adjoint_gradient(psi, lam, circ, H)
    assert abs(e - oracle.energy(n, circ, H, theta)) < 1e-12
    ps = oracle.param_shift(n, circ, H, theta, list(range(theta.size)))
    assert np.abs(g - ps).max() < 1e-12
