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


def apply_gate(psi, gate, op: str = "N", stream=None) -> None:
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
