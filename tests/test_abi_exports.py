"""The C-ABI library loads without a GPU and exports what include/qsv.h declares."""
from __future__ import annotations

import ctypes
import subprocess

import pytest

from qclabpp_amd import _lib


def _loaded():
    if not _lib.LIB_PATH.exists():
        pytest.skip("engine library not built")
    return _lib.lib()


def test_every_declared_function_is_exported():
    lib = _loaded()
    declared = _lib.header_functions()
    assert len(declared) >= 25
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, f"declared in include/qsv.h but not exported: {missing}"


def test_python_signatures_cover_the_header():
    """Every declared function has a ctypes signature, and no stale ones."""
    assert sorted(_lib.SIGNATURES) == _lib.header_functions()


def test_exports_are_plain_c_symbols():
    """No C++ mangling on the boundary (nm lists the names verbatim)."""
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True)
    if out.returncode != 0:
        pytest.skip("nm unavailable")
    names = {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}
    for name in _lib.header_functions():
        assert name in names, name


def test_calls_that_need_no_gpu():
    lib = _loaded()
    assert b"gfx950" in lib.qsv_version()
    assert lib.qsv_device_count() >= 0
    # argument checks fail before any device work
    st = lib.qsv_host_simulate(None, 3, 2, None, 0, 0)
    assert st == _lib.QSV_EINVAL
    assert b"null" in lib.qsv_last_error()
    buf = (ctypes.c_double * 16)()
    assert lib.qsv_host_simulate(buf, 3, 9, None, 0, 0) == 5  # QSV_EDTYPE


def test_matrix_argument_checks_need_no_gpu():
    from qclabpp_amd import gates as G
    lib = _loaded()
    buf = (ctypes.c_double * 8)()
    arr, cnt = G.records([G.Hadamard(2)])
    assert lib.qsv_host_circuit_matrix(buf, 2, _lib.QSV_C128, arr, cnt, 0) == _lib.QSV_EQUBIT
    assert lib.qsv_host_circuit_matrix(buf, 21, _lib.QSV_C128, arr, cnt, 0) == _lib.QSV_EINVAL
    assert lib.qsv_host_circuit_matrix(None, 2, _lib.QSV_C128, arr, cnt, 0) == _lib.QSV_EINVAL
    out = ctypes.c_double()
    assert lib.qsv_circuit_nrmf(2, 9, arr, cnt, arr, cnt, 0, ctypes.byref(out)) == _lib.QSV_EDTYPE


def test_records_match_per_gate_records():
    """The column-wise record builder writes the same bytes as Gate.record,
    for every op override and offset, including explicit (cos, sin) and
    matrix gates."""
    import numpy as np

    from qclabpp_amd import gates as G
    from qclabpp_amd import workloads as W

    rng = np.random.default_rng(3)
    u1 = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0]
    mixed = [G.Hadamard(1), G.RotationY(2, 0.3, param=4), G.RotationZ(0, cos=0.6, sin=0.8),
             G.QGate1(3, u1), G.CX(0, 2, controlState=0), G.CPhase(1, 3, 0.7, param=0), G.CZ(1, 2)]
    for circ in (mixed, W.tfim_ansatz(12, 3), W.brickwork(10, 4)):
        for op, off in ((None, 0), ("C", 3), ("T", 1)):
            arr, cnt = G.records(circ, op, off)
            assert cnt == len(circ)
            for i, g in enumerate(circ):
                assert bytes(arr[i]) == bytes(g.record(op, off))
    arr, cnt = G.records([])
    assert cnt == 0
