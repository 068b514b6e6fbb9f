"""Device-resident state-vector API over the C ABI.

A state is a contiguous torch tensor of 2^n complex128 or complex64 values
on the GPU (the `simulate_device(T*)` analogue of
include/qclab/QCircuit.hpp:170-172).  torch provides memory and streams
only; every gate runs in the HIP engine.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import QSV_C64, QSV_C128, check, lib
from .gates import records


def _torch():
    import torch
    return torch


def dtype_code(x) -> int:
    torch = _torch()
    if x.dtype == torch.complex128:
        return QSV_C128
    if x.dtype == torch.complex64:
        return QSV_C64
    raise TypeError(f"state must be complex128 or complex64, got {x.dtype}")


def nqubits(x) -> int:
    n = int(x.numel()).bit_length() - 1
    if x.numel() != 1 << n or n < 1:
        raise ValueError(f"state length {x.numel()} is not a power of two >= 2")
    return n


def _check_state(psi):
    if not psi.is_cuda:
        raise ValueError("state must live on the GPU")
    if not psi.is_contiguous():
        raise ValueError("state must be contiguous")
    return nqubits(psi), dtype_code(psi)


def stream_handle(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def new_state(n: int, dtype=None, device="cuda", basis: int = 0):
    """|basis> on the GPU."""
    torch = _torch()
    dtype = dtype or torch.complex128
    psi = torch.empty(1 << n, dtype=dtype, device=device)
    set_basis(psi, basis)
    return psi


def set_basis(psi, index: int = 0, stream=None) -> None:
    n, dt = _check_state(psi)
    check(lib().qsv_dev_set_basis(C.c_void_p(psi.data_ptr()), n, dt, index,
                                  stream_handle(stream)))


def apply_gate(psi, gate, op: str | None = None, stream=None) -> None:
    n, dt = _check_state(psi)
    rec = gate.record(op)
    check(lib().qsv_dev_apply_gate(C.c_void_p(psi.data_ptr()), n, dt, C.byref(rec),
                                   stream_handle(stream)))


def apply_gates(psi, gates, stream=None) -> None:
    """Plain path: one kernel per gate, in order."""
    n, dt = _check_state(psi)
    arr, cnt = records(gates)
    check(lib().qsv_dev_apply_gates(C.c_void_p(psi.data_ptr()), n, dt, arr, cnt,
                                    stream_handle(stream)))


def simulate_host(vec: np.ndarray, gates) -> np.ndarray:
    """std::vector path: upload, run every gate on the GPU, download."""
    out = np.ascontiguousarray(vec).copy()
    if out.dtype == np.complex128:
        dt = QSV_C128
    elif out.dtype == np.complex64:
        dt = QSV_C64
    else:
        raise TypeError("vector must be complex128 or complex64")
    n = int(out.size).bit_length() - 1
    arr, cnt = records(gates)
    check(lib().qsv_host_apply_gates(out.ctypes.data_as(C.c_void_p), n, dt, arr, cnt))
    return out


TILED = _lib.QSV_TILED
AUTO = _lib.QSV_AUTO
TILED_JIT = _lib.QSV_TILED_JIT
TILED_JIT_FIXED = _lib.QSV_TILED_JIT_FIXED


def circuit_matrix(k: int, gates, dtype=None, fuse_k: int = AUTO, stream=None):
    """QCircuit::matrix() on the GPU (reference include/qclab/QCircuit.hpp:
    121-127): the 2^k x 2^k unitary of a k-qubit gate list as a device tensor
    M with M[r, c] = <r|U|c> (MSB kron convention).  The engine fills a
    column-major buffer; the returned tensor is its transposed view."""
    torch = _torch()
    dtype = dtype or torch.complex128
    dt = dtype_code(torch.empty(0, dtype=dtype))
    N = 1 << k
    buf = torch.empty(N * N, dtype=dtype, device="cuda")
    arr, cnt = records(gates)
    check(lib().qsv_dev_circuit_matrix(C.c_void_p(buf.data_ptr()), k, dt, arr, cnt, fuse_k,
                                       stream_handle(stream)))
    return buf.view(N, N).t()


def nrmf(a, b, stream=None) -> float:
    """nrmF's scaled Frobenius norm of a - b (reference QCircuit.hpp:406-430)
    for two device tensors of equal size, dtype and layout."""
    if a.shape != b.shape or a.dtype != b.dtype:
        raise ValueError("nrmf: tensors differ in shape or dtype")
    if not (a.is_cuda and b.is_cuda):
        raise ValueError("nrmf: tensors must live on the GPU")
    # the norm pairs elements by storage offset: both must be dense in the
    # same layout (contiguous, or both transposed views of contiguous buffers)
    dense = a.numel() == 0 or (a.stride() == b.stride() and
                               (a.is_contiguous() or (a.dim() == 2 and a.t().is_contiguous())))
    if not dense:
        raise ValueError("nrmf: tensors must be dense with equal strides")
    dt = dtype_code(a)
    out = C.c_double()
    check(lib().qsv_dev_nrmf(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), a.numel(), dt,
                             C.byref(out), stream_handle(stream)))
    return out.value


def circuit_nrmf(k: int, gates1, gates2, dtype: str = "c128", fuse_k: int = AUTO) -> float:
    """nrmF(circuit1, circuit2) with both unitaries built and reduced on the GPU."""
    dt = QSV_C128 if dtype in ("c128", "complex128") else QSV_C64
    a1, c1 = records(gates1)
    a2, c2 = records(gates2)
    out = C.c_double()
    check(lib().qsv_circuit_nrmf(k, dt, a1, c1, a2, c2, fuse_k, C.byref(out)))
    return out.value


class Program:
    """A compiled circuit (qsv_program): plain (fuse_k=0), fused (1..5) or
    tiled passes (fuse_k=TILED)."""

    def __init__(self, n: int, gates, fuse_k: int = 5, dtype: str = "c128", diag_k: int = 0):
        self.n = n
        self.dtype = QSV_C128 if dtype in ("c128", "complex128") else QSV_C64
        self.fuse_k = fuse_k
        arr, cnt = records(gates)
        h = C.c_void_p()
        check(lib().qsv_program_create(n, self.dtype, arr, cnt, fuse_k, diag_k, C.byref(h)))
        self._h = h
        # kept here: module globals may already be gone when a Program is
        # collected at interpreter exit
        self._destroy = lib().qsv_program_destroy

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._destroy(h)
            self._h = None

    def run(self, psi, stream=None) -> None:
        n, dt = _check_state(psi)
        if n != self.n or dt != self.dtype:
            raise ValueError("state does not match the program's qubit count / dtype")
        check(lib().qsv_program_run(self._h, C.c_void_p(psi.data_ptr()), stream_handle(stream)))

    def info(self) -> dict:
        inf = _lib.QsvProgramInfo()
        check(lib().qsv_program_info_get(self._h, C.byref(inf)))
        return {"ngates": inf.ngates, "nlaunch": inf.nlaunch, "nplain": inf.nplain,
                "ndiag": inf.ndiag, "ndense": list(inf.ndense), "alg_bytes": inf.alg_bytes,
                "flops": inf.flops, "tiled_phases": inf.ntiled_phases,
                "tiled_relabel": inf.ntiled_relabel, "fp_lane_ops": inf.fp_lane_ops,
                "work_bytes": inf.work_bytes}

    def launches(self):
        """Every launch of one run, in order: (qubits, diag, coefficients) with
        coefficients the row-major matrix or the diagonal (row bit p <->
        qubits[k-1-p])."""
        out = []
        for i in range(self.info()["nlaunch"]):
            k = C.c_int()
            qs = (C.c_int * 16)()
            dg = C.c_int()
            co = (C.c_double * (2 * 4 ** 10))()
            check(lib().qsv_program_op(self._h, i, C.byref(k), qs, C.byref(dg), co))
            kk = k.value
            D = 1 << kk
            vals = np.array(co[: 2 * (D if dg.value else D * D)]).view(np.complex128)
            out.append((list(qs[:kk]), bool(dg.value), vals if dg.value else vals.reshape(D, D)))
        return out
