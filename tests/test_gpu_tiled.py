"""GPU parity of tiled programs: passes over LDS tiles of 2^12 amplitudes,
gates applied in registers phase by phase, run by the interpreting pass
kernel (fuse_k = QSV_TILED) and as passes compiled per circuit structure
(QSV_TILED_JIT).  Checked against the oracle on mixed circuits with every
gate class, both control states, Op N/C and every qubit."""
import numpy as np
import pytest

import oracle
from qclabpp_amd import engine as E
from qclabpp_amd import gates as G
from qclabpp_amd import workloads as W
from qclabpp_amd.rng import normal_state
from tests.gatecases import one_qubit_gates, random_state, two_qubit_gates
from tests.parity import C64, C128, assert_close, tol

pytestmark = pytest.mark.gpu

FORMS = [pytest.param(E.TILED, id="interp"), pytest.param(E.TILED_JIT, id="jit"),
         pytest.param(E.TILED_JIT_FIXED, id="jitfixed")]
JIT_FORMS = [pytest.param(E.TILED_JIT, id="jit"), pytest.param(E.TILED_JIT_FIXED, id="jitfixed")]


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


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("n", [12, 13, 15, 18])
def test_tiled_every_kind_c128(n, form):
    print(f"tiled every kind n={n} form={form}", flush=True)
    circ = every_kind(n, n)
    psi = random_state(n, 5, np.complex128)
    want = oracle.apply(n, circ, psi)
    prog = E.Program(n, circ, fuse_k=form)
    got = run(prog, psi)
    assert_close(got, want, C128, f"n={n}")


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("n", [13, 17])
def test_tiled_every_kind_c64(n, form):
    circ = every_kind(n, 2 * n)
    psi = random_state(n, 6, np.complex128)
    want = oracle.apply(n, circ, psi)
    got = run(E.Program(n, circ, fuse_k=form, dtype="c64"), psi.astype(np.complex64))
    assert_close(got, want, C64, f"c64 n={n}")


@pytest.mark.parametrize("form,count", [pytest.param(E.TILED, 5000, id="interp"),
                                        pytest.param(E.TILED_JIT, 600, id="jit"),
                                        pytest.param(E.TILED_JIT_FIXED, 600, id="jitfixed")])
@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_tiled_deep_pass_keeps_scale(dt, form, count):
    """Thousands of rotations on a 12-qubit state are one pass (its tile
    is the whole state), interpreted and compiled.  The rotations run in
    scaled forms whose factors multiply up over the pass; after 120 scaled
    rotations the next runs as a complex 2x2 that resets the product, so the
    stored amplitudes stay within 2^60 of their true scale and neither
    precision overflows.  The compiled pass is kept to 600 rotations
    (4 resets) so that its straight-line code compiles in seconds."""
    print(f"tiled deep pass {dt} form={form} rotations={count}", flush=True)
    n = 12
    rs = np.random.default_rng(31)
    circ = []
    for i in range(count):
        q = int(rs.integers(n))
        circ.append((G.RotationX, G.RotationY)[i % 2](q, float(rs.uniform(0, 2 * np.pi))))
        if i % 50 == 49:
            circ.append(G.CX(q, (q + 1) % n))
    psi = random_state(n, 8, np.complex128)
    want = oracle.apply(n, circ, psi)
    cdt = np.complex128 if dt == "c128" else np.complex64
    prog = E.Program(n, circ, fuse_k=form, dtype=dt)
    assert prog.info()["nlaunch"] == 1
    got = run(prog, psi.astype(cdt))
    assert np.all(np.isfinite(got))
    # thousands of gates in one pass: 1e-11 relative for complex128 (the
    # rounding of 5000 gates adds up), FP32 as everywhere else
    assert_close(got, want, 1e-11 if dt == "c128" else C64, f"deep {dt}")


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("n", [16, 20])
def test_tiled_brickwork_and_ansatz(n, form):
    circ = W.brickwork(n, 20, seed=7) + W.tfim_ansatz(n, 2)
    psi = random_state(n, 9, np.complex128)
    want = oracle.apply(n, circ, psi)
    prog = E.Program(n, circ, fuse_k=form)
    assert prog.info()["nlaunch"] >= 2
    got = run(prog, psi)
    assert_close(got, want, C128, f"n={n}")


@pytest.mark.parametrize("form", FORMS)
def test_tiled_qft20_is_dft(form):
    n = 20
    x = normal_state(W.SEED_QFT, n)
    want = np.fft.ifft(x) * np.sqrt(1 << n)
    got = run(E.Program(n, W.qft(n), fuse_k=form), x)
    assert_close(got, want, C128, "QFT")


def test_jit_structure_shared_across_angles():
    """Programs whose gates differ only in angles share compiled passes
    (the second creation finds them cached) and each gives its own result."""
    import time
    n = 16
    a = W.brickwork(n, 8, seed=11)
    # same structure, new angles (none near 0 or pi)
    b = [g.with_theta(0.5 + 0.9 * abs(np.sin(3.0 * g.theta))) if g.theta else g for g in a]
    psi = random_state(n, 4, np.complex128)
    t0 = time.perf_counter()
    pa = E.Program(n, a, fuse_k=E.TILED_JIT)
    t1 = time.perf_counter()
    pb = E.Program(n, b, fuse_k=E.TILED_JIT)
    t2 = time.perf_counter()
    print(f"jit create: first {t1 - t0:.2f} s, same structure {t2 - t1:.3f} s", flush=True)
    assert_close(run(pa, psi), oracle.apply(n, a, psi), C128, "angles a")
    assert_close(run(pb, psi), oracle.apply(n, b, psi), C128, "angles b")
    assert t2 - t1 < 0.5 * (t1 - t0)


@pytest.mark.parametrize("form", FORMS)
def test_tiled_brickwork_c64(form):
    """complex64 passes (compiled ones use 2^13-amplitude tiles) against the
    oracle, through the same circuit in complex128 as well."""
    n = 20
    circ = W.brickwork(n, 20, seed=8) + W.tfim_ansatz(n, 1)
    psi = random_state(n, 3, np.complex128)
    want = oracle.apply(n, circ, psi)
    got = run(E.Program(n, circ, fuse_k=form, dtype="c64"), psi.astype(np.complex64))
    assert_close(got, want, C64, "c64 brickwork")


def _structured(n):
    """Circuits whose passes take each load/store route of the compiled
    kernels: gates only on the line qubits (first and last phases gather
    and scatter through LDS), one gate far from them (a one-phase pass with
    no LDS at all), and mixes whose first or last phase holds line bits."""
    hi = [0, 1, 2, 3]
    line = [n - 1, n - 2, n - 3]
    rot = lambda q, t: [G.RotationX(q, t), G.RotationY(q, 0.7 * t), G.RotationZ(q, 0.3 * t)]
    c_line = sum((rot(q, 0.4 + q) for q in line), []) + [G.CX(n - 1, n - 2), G.CZ(n - 3, n - 1)]
    c_one = [G.RotationY(0, 0.9)]
    c_hi_then_line = sum((rot(q, 0.2 * q + 0.1) for q in hi), []) + [G.CX(0, 3)] + c_line
    c_line_then_hi = c_line + sum((rot(q, 0.5 * q + 0.3) for q in hi), []) + [G.SWAP(1, 2)]
    return {"line_only": c_line, "one_gate": c_one, "high_then_line": c_hi_then_line,
            "line_then_high": c_line_then_hi}


@pytest.mark.parametrize("form", JIT_FORMS)
@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("name", ["line_only", "one_gate", "high_then_line", "line_then_high"])
def test_jit_load_store_routes(name, dt, form):
    n = 15
    circ = _structured(n)[name]
    psi = random_state(n, 12, np.complex128)
    want = oracle.apply(n, circ, psi)
    x = psi if dt == "c128" else psi.astype(np.complex64)
    got = run(E.Program(n, circ, fuse_k=form, dtype=dt), x)
    assert_close(got, want, tol(dt), f"{name} {dt}")


@pytest.mark.parametrize("oop", [pytest.param(True, id="oop"), pytest.param(False, id="restore")])
@pytest.mark.parametrize("form", JIT_FORMS)
@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("n", [18, 20, 22])
def test_jit_relabeling_passes(n, dt, form, oop, monkeypatch):
    """Compiled passes that move qubits onto the line bits (and back by the
    end of the program): brickwork plus every gate kind, against the
    oracle, with at least one relabeling pass in the schedule.  The schedule
    ends out of place when it can (the last pass reads a work buffer that the
    pass before it wrote, and writes the state with every qubit home), or,
    with QSV_JIT_OUT_OF_PLACE=0, with a gate-free restore pass."""
    if not oop:
        monkeypatch.setenv("QSV_JIT_OUT_OF_PLACE", "0")
    circ = W.brickwork(n, 20, seed=30 + n) + every_kind(n, n) + W.brickwork(n, 6, seed=n)
    psi = random_state(n, 40 + n, np.complex128)
    want = oracle.apply(n, circ, psi)
    prog = E.Program(n, circ, fuse_k=form, dtype=dt)
    info = prog.info()
    print(f"relabel n={n} {dt}: passes={info['nlaunch']} relabeling={info['tiled_relabel']} "
          f"work={info['work_bytes']:.0f}", flush=True)
    assert info["tiled_relabel"] > 0
    state_bytes = (1 << n) * (16 if dt == "c128" else 8)
    assert info["work_bytes"] in (0, state_bytes)
    if not oop:
        assert info["work_bytes"] == 0
    x = psi if dt == "c128" else psi.astype(np.complex64)
    got = run(prog, x)
    assert_close(got, want, tol(dt), f"relabel n={n} {dt}")


def test_jit_out_of_place_end_is_used():
    """Some relabeling schedules end out of place: the n=20 complex128
    brickwork-and-every-kind circuit in fixed forms does, and its result
    matches the oracle on two runs in a row."""
    n = 20
    circ = W.brickwork(n, 20, seed=30 + n) + every_kind(n, n) + W.brickwork(n, 6, seed=n)
    prog = E.Program(n, circ, fuse_k=E.TILED_JIT_FIXED)
    info = prog.info()
    print(f"out-of-place end: passes={info['nlaunch']} work={info['work_bytes']:.0f}", flush=True)
    assert info["work_bytes"] == (1 << n) * 16
    psi = random_state(n, 77, np.complex128)
    assert_close(run(prog, psi), oracle.apply(n, circ, psi), C128, "out-of-place end")
    # the same program twice on one stream: the work buffer is reused
    assert_close(run(prog, psi), oracle.apply(n, circ, psi), C128, "second run")
