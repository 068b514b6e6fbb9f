"""The specialised tiled passes (QSV_TILED_JIT) generate and compile for
gfx950 without a GPU: every pass of several circuits goes through source
generation and hiprtc here, so a generator error fails the CPU suite."""
from __future__ import annotations

import ctypes as C
import math

import pytest

from qclabpp_amd import _lib
from qclabpp_amd import gates as G
from qclabpp_amd import workloads as W
from tests.test_gpu_tiled import every_kind


def _compile(n, circ, dtype):
    if not _lib.LIB_PATH.exists():
        pytest.skip("engine library not built")
    arr, cnt = G.records(circ)
    size = C.c_size_t()
    npass = C.c_int64()
    st = _lib.lib().qsv_tiled_jit_compile_check(n, dtype, arr, cnt, C.byref(size), C.byref(npass),
                                                None)
    if st == _lib.QSV_EUNSUP and b"hiprtc" in _lib.lib().qsv_last_error():
        pytest.skip("hiprtc not loadable here")
    _lib.check(st)
    assert npass.value >= 1 and size.value > 0
    return npass.value, size.value


def _source_hash(n, circ, dtype):
    """(pass count, hash of the generated source of every pass), without
    compiling: the compiled-module cache is keyed by that source."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("engine library not built")
    arr, cnt = G.records(circ)
    npass = C.c_int64()
    h = C.c_uint64()
    _lib.check(_lib.lib().qsv_tiled_jit_compile_check(n, dtype, arr, cnt, None, C.byref(npass),
                                                      C.byref(h)))
    return npass.value, h.value


def _check(n, circ, dtype):
    return _compile(n, circ, dtype)[0]


def _form_check(n, circ, dtype, form, compile_=True):
    """(pass count, source hash) of QSV_TILED_JIT or QSV_TILED_JIT_FIXED,
    compiling every pass when compile_ is set."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("engine library not built")
    arr, cnt = G.records(circ)
    size = C.c_size_t()
    npass = C.c_int64()
    h = C.c_uint64()
    st = _lib.lib().qsv_tiled_jit_compile_check_form(n, dtype, arr, cnt, form,
                                                     C.byref(size) if compile_ else None,
                                                     C.byref(npass), C.byref(h))
    if st == _lib.QSV_EUNSUP and b"hiprtc" in _lib.lib().qsv_last_error():
        pytest.skip("hiprtc not loadable here")
    _lib.check(st)
    return npass.value, h.value


@pytest.mark.parametrize("dtype", [_lib.QSV_C128, _lib.QSV_C64])
def test_generic_angles_share_compiled_code(dtype):
    """Angles are data: two circuits of one structure whose generic angles
    differ, including which of cos and sin is larger (the two scaled
    rotation forms), compile to passes of the same number and code size."""
    import random

    def circ(seed):
        r = random.Random(seed)
        out = []
        for layer in range(6):
            for q in range(14):
                ctor = (G.RotationX, G.RotationY, G.RotationZ)[(q + layer) % 3]
                out.append(ctor(q, r.uniform(0.1, 3.0)))
            for i in range(layer % 2, 13, 2):
                out.append(G.CX(i, i + 1))
        return out

    assert _compile(14, circ(1), dtype) == _compile(14, circ(2), dtype)
    assert _source_hash(14, circ(1), dtype) == _source_hash(14, circ(2), dtype)


@pytest.mark.parametrize("dtype", [_lib.QSV_C128, _lib.QSV_C64])
def test_deep_pass_source_is_angle_independent(dtype):
    """A deep circuit on 12 qubits (360 rotations) is one pass whose scaled
    rotations reset their factored-out product twice.  The reset points depend on the
    structure only, so every angle set -- small and large angles, cos or sin
    the larger, angles that take the product to its bound fastest -- gives
    the same generated source and shares one compiled module."""
    import random

    n = 12

    def circ(angle):
        out = []
        for layer in range(30):
            for q in range(n):
                out.append((G.RotationX, G.RotationY)[(q + layer) % 2](q, angle(layer, q)))
            for i in range(layer % 2, n - 1, 2):
                out.append(G.CX(i, i + 1))
        return out

    r = random.Random(5)
    sets = [lambda l, q: r.uniform(0, 2 * math.pi),
            lambda l, q: 0.01 + 0.001 * q,                  # |cos| ~ 1: factors near 1
            lambda l, q: math.pi / 2 + 0.001 * (l + q),     # |cos| ~ |sin|: factors near 2^-1/2
            lambda l, q: 3.0 - 0.01 * l]
    hashes = {_source_hash(n, circ(a), dtype) for a in sets}
    assert len(hashes) == 1, hashes
    assert next(iter(hashes))[0] == 1  # the whole circuit is one pass


@pytest.mark.parametrize("dtype", [_lib.QSV_C128, _lib.QSV_C64])
def test_every_gate_kind_compiles(dtype):
    # every gate class, both control states, Op N/C: exercises every op form
    _check(13, every_kind(13, 13), dtype)


def test_brickwork_and_dense_passes_compile():
    n = 14
    circ = W.brickwork(n, 6, seed=4) + [G.RotationXX(0, n - 1, 0.3), G.iSWAP(2, 9), G.CX(5, 6, 0)]
    assert _check(n, circ, _lib.QSV_C128) >= 1


def test_bad_arguments_are_rejected():
    if not _lib.LIB_PATH.exists():
        pytest.skip("engine library not built")
    lib = _lib.lib()
    assert lib.qsv_tiled_jit_compile_check(8, _lib.QSV_C128, None, 0, None, None, None) == _lib.QSV_EINVAL
    assert lib.qsv_tiled_jit_compile_check(14, 7, None, 0, None, None, None) == _lib.QSV_EDTYPE


@pytest.mark.parametrize("dtype", [_lib.QSV_C128, _lib.QSV_C64])
def test_relabeling_schedules_compile(dtype):
    """Schedules whose passes move qubits onto the line bits: the scheduler
    checks that every pass writes exactly its own tile's addresses and that
    the lines read and written stay within a wave's lanes, then every pass
    compiles."""
    n = 20
    assert _check(n, W.brickwork(n, 20, seed=50) + every_kind(n, 20), dtype) >= 2


def _adjoint_compile(n, circ, dtype):
    if not _lib.LIB_PATH.exists():
        pytest.skip("engine library not built")
    arr, cnt = G.records(circ)
    size = C.c_size_t()
    npass = C.c_int64()
    rows = C.c_int64()
    st = _lib.lib().qsv_adjoint_compile_check(n, dtype, arr, cnt, C.byref(size), C.byref(npass),
                                              C.byref(rows))
    if st == _lib.QSV_EUNSUP and b"hiprtc" in _lib.lib().qsv_last_error():
        pytest.skip("hiprtc not loadable here")
    _lib.check(st)
    return npass.value, rows.value, size.value


def differentiable_circuit(n, seed):
    """Every differentiable kind (both control states, generators on
    scattered qubits), shared parameters, and fixed gates between them."""
    import numpy as np
    rs = np.random.default_rng(seed)
    circ = [G.Hadamard(q) for q in range(n)]
    p = 0
    for ctor in (G.RotationX, G.RotationY, G.RotationZ, G.Phase):
        for q in range(n):
            circ.append(ctor(q, rs.uniform(-3, 3), param=p))
            p += 1
    circ += [G.CX(0, n - 1), G.SWAP(1, 4), G.iSWAP(2, n - 2), G.PauliY(3)]
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
    return circ, p


@pytest.mark.parametrize("dtype", [_lib.QSV_C128, _lib.QSV_C64])
def test_adjoint_passes_compile(dtype):
    """The compiled reverse sweep (two states, generator rows) generates and
    compiles for every differentiable kind and for the TFIM ansatz; every
    parametrised gate gets one row."""
    circ, p = differentiable_circuit(14, 3)
    npass, rows, size = _adjoint_compile(14, circ, dtype)
    assert npass >= 1 and size > 0
    assert rows == sum(1 for g in circ if g.param >= 0)
    tf = W.tfim_ansatz(16, 3)
    npass, rows, size = _adjoint_compile(16, tf, dtype)
    assert rows == len(tf) and npass >= 2


@pytest.mark.parametrize("dtype", [_lib.QSV_C128, _lib.QSV_C64])
def test_fixed_form_passes_compile(dtype):
    """QSV_TILED_JIT_FIXED emits each scaled rotation in the form its angle
    picks, without the branch: every pass of a brickwork and of a circuit
    with every gate kind compiles."""
    from qclabpp_amd import engine as E
    _form_check(14, W.brickwork(14, 8, seed=3), dtype, E.TILED_JIT_FIXED)
    _form_check(13, every_kind(13, 13), dtype, E.TILED_JIT_FIXED)


def test_fixed_form_source_follows_the_forms():
    """Fixed-form passes depend on which of |cos| and |sin| is the larger
    per rotation, not on the angles otherwise: angles that keep every form
    share the source, a flipped form changes it, and the angle-generic
    source is the same for all three angle sets."""
    from qclabpp_amd import engine as E
    n = 12

    def circ(angles):
        out = []
        for layer in range(4):
            for q in range(n):
                out.append(G.RotationY(q, angles[layer * n + q]))
            for i in range(layer % 2, n - 1, 2):
                out.append(G.CX(i, i + 1))
        return out

    base = [0.3 + 0.01 * i for i in range(4 * n)]           # |cos| > |sin| everywhere
    same = [0.2 + 0.005 * i for i in range(4 * n)]          # the same forms
    flip = list(base)
    flip[5] = 2.0                                            # |sin| > |cos| for one rotation
    f = [_form_check(n, circ(a), _lib.QSV_C128, E.TILED_JIT_FIXED, compile_=False)[1] for a in (base, same, flip)]
    assert f[0] == f[1] and f[0] != f[2]
    g = {_form_check(n, circ(a), _lib.QSV_C128, E.TILED_JIT, compile_=False)[1] for a in (base, same, flip)}
    assert len(g) == 1
    assert _lib.lib().qsv_tiled_jit_compile_check_form(n, _lib.QSV_C128, None, 0, 0, None, None,
                                                       None) == _lib.QSV_EINVAL
