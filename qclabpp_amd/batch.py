"""Batched simulation of many independent circuits (packed states).

Each circuit starts from |0...0>; circuit c's final state occupies
amplitudes [offsets[c], offsets[c+1]) of one device buffer.  See
qsv_batch_* in include/qsv.h for the execution model (LDS tile passes, one
hipGraph per run, autotuned tile width and fusion width).
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import QSV_C64, QSV_C128, QsvBatchOpts, check, lib
from .engine import stream_handle
from .gates import records


def _pack(circuits):
    flat = [g for c in circuits for g in c]
    arr, cnt = records(flat)
    offs = np.zeros(len(circuits) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(c) for c in circuits])
    return arr, offs


def _opts(tile_bits=0, small_bits=0, fuse_k=1, phases=0):
    o = QsvBatchOpts()
    o.tile_bits, o.small_bits, o.fuse_k, o.phases = tile_bits, small_bits, fuse_k, phases
    return o


class Batch:
    def __init__(self, circuits, nqubits, dtype: str = "c64", tile_bits: int = 0,
                 small_bits: int = 0, fuse_k: int = 1, phases: int = 0):
        self.dtype = QSV_C64 if dtype in ("c64", "complex64") else QSV_C128
        self.nq = np.ascontiguousarray(np.asarray(nqubits, dtype=np.int32))
        self._arr, self._goff = _pack(circuits)
        goff = self._goff.astype(np.uintp)
        o = _opts(tile_bits, small_bits, fuse_k, phases)
        h = C.c_void_p()
        check(lib().qsv_batch_create(self.dtype, len(circuits),
                                     self.nq.ctypes.data_as(C.POINTER(C.c_int)), self._arr,
                                     goff.ctypes.data_as(C.POINTER(C.c_size_t)), C.byref(o),
                                     C.byref(h)))
        self._h = h
        inf = _lib.QsvBatchInfo()
        offs = np.zeros(len(circuits) + 1, dtype=np.uint64)
        check(lib().qsv_batch_info(h, C.byref(inf), offs.ctypes.data_as(C.POINTER(C.c_uint64))))
        self.offsets = offs
        self.info = {"ncirc": inf.ncirc, "nsmall": inf.nsmall, "npasses": inf.npasses,
                     "nops": inf.nops, "nlevels": inf.nlevels,
                     "total_amplitudes": inf.total_amplitudes,
                     "tile_bits": inf.opts.tile_bits, "small_bits": inf.opts.small_bits,
                     "fuse_k": inf.opts.fuse_k, "phases": inf.opts.phases,
                     "alg_bytes": inf.alg_bytes}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().qsv_batch_destroy(h)
            self._h = None

    def new_buffer(self, device="cuda"):
        import torch
        dt = torch.complex64 if self.dtype == QSV_C64 else torch.complex128
        return torch.empty(int(self.offsets[-1]), dtype=dt, device=device)

    def run(self, packed, use_graph: bool = True, stream=None) -> None:
        if packed.numel() < int(self.offsets[-1]):
            raise ValueError("packed buffer too small")
        check(lib().qsv_batch_run(self._h, C.c_void_p(packed.data_ptr()), int(use_graph),
                                  stream_handle(stream)))

    def latencies(self, stream=None) -> np.ndarray:
        out = np.zeros(len(self.nq), dtype=np.float64)
        check(lib().qsv_batch_latencies(self._h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                        stream_handle(stream)))
        return out

    def state(self, packed, c: int):
        return packed[int(self.offsets[c]):int(self.offsets[c + 1])]


def autotune(circuits, nqubits, packed, scratch, dtype: str = "c64", reps: int = 3, stream=None):
    """Time every (tile_bits, small_bits, fuse_k, phases) candidate; returns
    (best opts dict, results)."""
    dt = QSV_C64 if dtype in ("c64", "complex64") else QSV_C128
    nq = np.ascontiguousarray(np.asarray(nqubits, dtype=np.int32))
    arr, goff = _pack(circuits)
    goff = goff.astype(np.uintp)
    best = QsvBatchOpts()
    res = (_lib.QsvTuneResult * 32)()
    nres = C.c_int()
    check(lib().qsv_batch_autotune(dt, len(circuits), nq.ctypes.data_as(C.POINTER(C.c_int)), arr,
                                   goff.ctypes.data_as(C.POINTER(C.c_size_t)),
                                   C.c_void_p(packed.data_ptr()), C.c_void_p(scratch.data_ptr()),
                                   reps, C.byref(best), res, 32, C.byref(nres), stream_handle(stream)))
    results = [{"tile_bits": r.opts.tile_bits, "small_bits": r.opts.small_bits,
                "fuse_k": r.opts.fuse_k, "phases": r.opts.phases, "ms": r.ms,
                "max_abs_diff": r.max_abs_diff} for r in res[:nres.value]]
    return {"tile_bits": best.tile_bits, "small_bits": best.small_bits, "fuse_k": best.fuse_k,
            "phases": best.phases}, results


def signature(circuits, nqubits, dtype: str) -> str:
    """Cache key of a batch: dtype, qubit counts and gate kinds/qubits."""
    h = hashlib.sha256(dtype.encode())
    h.update(np.asarray(nqubits, dtype=np.int32).tobytes())
    for c in circuits:
        h.update(np.array([(g.kind, *g.qubits, g.ctrl_state) + ((0,) * (2 - len(g.qubits)))
                           for g in c], dtype=np.int32).tobytes())
    return h.hexdigest()[:24]


def autotune_cached(circuits, nqubits, packed, scratch, dtype: str = "c64", reps: int = 3,
                    cache: str | None = None):
    """autotune() with a JSON cache keyed by signature()."""
    path = Path(cache or os.environ.get("QSV_TUNE_CACHE",
                                        Path.home() / ".cache" / "qclabpp_amd" / "batch_tune.json"))
    key = signature(circuits, nqubits, dtype)
    db = {}
    if path.exists():
        try:
            db = json.loads(path.read_text())
        except (OSError, ValueError):
            db = {}
    if key in db:
        return db[key]["best"], db[key]["results"], True
    best, results = autotune(circuits, nqubits, packed, scratch, dtype, reps)
    db[key] = {"best": best, "results": results}
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(db, indent=1))
    except OSError:
        pass
    return best, results, False
