"""qsv_dev_expectation_noalloc: the expectation values with the reduction
scratch in the library's own fixed device array instead of an allocation.
Checked through the C ABI against qsv_dev_expectation (same kernels, same
fixed reduction grid: bit-identical) and the oracle, on two calls in a row,
the second smaller; sizes beyond its maximum are rejected."""
import ctypes as C

import numpy as np
import pytest

import oracle
from qclabpp_amd import _lib
from qclabpp_amd import observables as O
from tests.gatecases import random_state

pytestmark = pytest.mark.gpu


def _call(name, psi, n, dtype, H):
    arr, cnt = H.records()
    vals = (C.c_double * max(1, cnt))()
    tot = C.c_double()
    st = getattr(_lib.lib(), name)(C.c_void_p(psi.data_ptr()), n, dtype, arr, cnt, vals,
                                   C.byref(tot), None)
    return st, np.array(vals[:cnt]), tot.value


def _hamiltonian(n, extra):
    H = O.tfim(n)
    H.add(0.3, [(1, "Y"), (3, "X"), (n - 1, "Z")]).add(-1.1, [(0, "Y")])
    H.add(0.25, [(0, "Z"), (2, "Z"), (4, "Z")])
    for i in range(extra):  # more diagonal and off-diagonal terms
        H.add(0.01 * (i + 1), [(i % n, "Z"), ((3 * i + 1) % n, "Z")] if i % 2 else [(i % n, "X")])
    return H


@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_noalloc_expectation_matches_allocating_entry(dt):
    import torch
    code = _lib.QSV_C128 if dt == "c128" else _lib.QSV_C64
    npdt = np.complex128 if dt == "c128" else np.complex64
    # two calls in a row, the second with a smaller state and fewer terms
    for n, extra in ((20, 150), (11, 3)):
        H = _hamiltonian(n, extra)
        psi0 = random_state(n, 7 + n, np.complex128).astype(npdt)
        psi = torch.from_numpy(psi0.copy()).cuda()
        st_a, v_a, t_a = _call("qsv_dev_expectation", psi, n, code, H)
        st_b, v_b, t_b = _call("qsv_dev_expectation_noalloc", psi, n, code, H)
        print(f"noalloc {dt} n={n} terms={len(H.terms)}: total {t_b:.15f}", flush=True)
        assert st_a == 0 and st_b == 0, _lib.lib().qsv_last_error()
        assert np.array_equal(v_a, v_b) and t_a == t_b
        want = oracle.expectation(n, psi0.astype(np.complex128), H)
        tol = 1e-12 if dt == "c128" else 1e-5
        assert np.abs(v_b - want).max() < tol
        assert np.array_equal(psi.cpu().numpy(), psi0)  # the state is only read


def test_noalloc_expectation_rejects_too_many_terms():
    import torch
    n = 8
    H = _hamiltonian(n, 300)
    assert len(H.terms) > 256
    psi = torch.from_numpy(random_state(n, 3, np.complex128)).cuda()
    st, _, _ = _call("qsv_dev_expectation_noalloc", psi, n, _lib.QSV_C128, H)
    assert st == _lib.QSV_EINVAL
    assert b"at most 256 terms" in _lib.lib().qsv_last_error()
