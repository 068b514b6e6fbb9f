"""GPU unitaries (QCircuit::matrix) and nrmF against the oracle and closed
forms: every gate class with per-record Ops, every compiled form, the QFT
against the DFT matrix, unitarity, and nrmF's scaled Frobenius norm."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

import oracle
from qclabpp_amd import _lib
from qclabpp_amd import engine as E
from qclabpp_amd import gates as G
from qclabpp_amd import workloads as W
from tests.test_oracle_matrix import dft, mixed_circuit

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 2e-5}


def torch_dtype(dt):
    import torch
    return torch.complex128 if dt == "c128" else torch.complex64


def with_ops(circ, seed):
    ops = np.random.default_rng(seed).choice(list("NTC"), len(circ))
    return [dataclasses.replace(g, op=str(o)) for g, o in zip(circ, ops)]


@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("k,form", [(3, 0), (3, 5), (5, 0), (5, 3), (5, E.AUTO), (7, 5), (7, E.TILED),
                                    (7, E.AUTO)])
def test_matrix_vs_oracle(k, form, dt):
    circ = with_ops(mixed_circuit(k, 10 + k), k)
    circ += [G.QGate2(0, k - 1, np.arange(16).reshape(4, 4) + 1j * np.eye(4))]  # non-adjacent
    print(f"matrix k={k} form={form} {dt} gates={len(circ)}", flush=True)
    got = E.circuit_matrix(k, circ, torch_dtype(dt), fuse_k=form).cpu().numpy()
    want = oracle.matrix(k, circ)
    scale = max(1.0, float(np.abs(want).max()))
    assert np.abs(got - want).max() <= TOL[dt] * scale * 10


@pytest.mark.parametrize("k,form", [(10, E.AUTO), (12, E.AUTO), (12, E.TILED_JIT)])
def test_qft_matrix_is_dft(k, form):
    print(f"qft matrix k={k} form={form}", flush=True)
    got = E.circuit_matrix(k, W.qft(k), fuse_k=form).cpu().numpy()
    assert np.abs(got - dft(k)).max() < 1e-12


def test_brickwork_matrix_is_unitary():
    import torch
    k = 11
    M = E.circuit_matrix(k, W.brickwork(k, 10, seed=8))
    err = (M @ M.conj().t() - torch.eye(1 << k, dtype=M.dtype, device=M.device)).abs().max().item()
    assert err < 1e-12


def test_host_entry_matches_device():
    import torch
    k = 8
    circ = with_ops(mixed_circuit(k, 2), 2)
    N = 1 << k
    host = np.empty(N * N, np.complex128)
    arr, cnt = G.records(circ)
    _lib.check(_lib.lib().qsv_host_circuit_matrix(host.ctypes.data_as(C.c_void_p), k, _lib.QSV_C128,
                                                   arr, cnt, E.AUTO))
    dev = E.circuit_matrix(k, circ).cpu().numpy()
    assert np.array_equal(host.reshape(N, N).T, dev)
    torch.cuda.synchronize()


@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_nrmf_semantics(dt):
    import torch
    rs = np.random.default_rng(4)
    a = (rs.standard_normal(1 << 14) + 1j * rs.standard_normal(1 << 14)) * 1e-3
    b = a.copy()
    b[::7] += 1e-4 * (rs.standard_normal(b[::7].size) + 1j)
    npdt = np.complex128 if dt == "c128" else np.complex64
    ta = torch.from_numpy(a.astype(npdt)).cuda()
    tb = torch.from_numpy(b.astype(npdt)).cuda()
    want = np.linalg.norm(a.astype(npdt).astype(np.complex128) - b.astype(npdt).astype(np.complex128))
    assert abs(E.nrmf(ta, tb) - want) <= 1e-12 * want
    assert E.nrmf(ta, ta) == 0.0


@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_circuit_nrmf(dt):
    k = 6
    c1 = W.brickwork(k, 6, seed=1)
    c2 = list(c1)
    c2[3] = c2[3].with_theta(c2[3].theta + 1e-3)
    want = np.linalg.norm(oracle.matrix(k, c1) - oracle.matrix(k, c2))
    got = E.circuit_nrmf(k, c1, c2, dtype=dt)
    assert abs(got - want) <= (1e-9 if dt == "c128" else 1e-3) * want
    assert E.circuit_nrmf(k, c1, c1, dtype=dt) == 0.0

