"""ctypes binding of the C ABI in include/qsv.h (libqclabpp_amd.so).

The product path has no CPU fallback: if the HIP library is not built, or no
GPU is visible, calls raise instead of computing something else.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libqclabpp_amd.so"
HEADER = PKG.parent / "include" / "qsv.h"

QSV_OK, QSV_EINVAL, QSV_EQUBIT, QSV_ENOMEM, QSV_EHIP, QSV_EDTYPE, QSV_ENODEV, QSV_EUNSUP = range(8)
QSV_TILED = -1  # fuse_k value selecting tiled passes (include/qsv.h)
QSV_AUTO = -2   # fuse_k value letting the engine choose
QSV_TILED_JIT = -3  # tiled passes compiled per circuit structure
QSV_TILED_JIT_FIXED = -4  # the same with each rotation's scaled form fixed at creation
QSV_C64, QSV_C128 = 1, 2


class QsvGate(C.Structure):
    """Mirror of `qsv_gate` (include/qsv.h)."""

    _fields_ = [
        ("kind", C.c_int32),
        ("q", C.c_int32 * 2),
        ("ctrl_state", C.c_int32),
        ("op", C.c_int32),
        ("param", C.c_int32),
        ("has_cs", C.c_int32),
        ("reserved", C.c_int32),
        ("theta", C.c_double),
        ("cs", C.c_double * 2),
        ("mat", C.c_double * 32),
    ]


class QsvPauliTerm(C.Structure):
    """Mirror of `qsv_pauli_term` (include/qsv.h)."""

    _fields_ = [
        ("coeff", C.c_double),
        ("nfactors", C.c_int32),
        ("reserved", C.c_int32),
        ("qubits", C.c_int32 * 16),
        ("paulis", C.c_byte * 16),
    ]


class QsvBatchOpts(C.Structure):
    """Mirror of `qsv_batch_opts` (include/qsv.h)."""

    _fields_ = [("tile_bits", C.c_int32), ("small_bits", C.c_int32), ("fuse_k", C.c_int32),
                ("phases", C.c_int32)]


class QsvBatchInfo(C.Structure):
    """Mirror of `qsv_batch_info_t` (include/qsv.h)."""

    _fields_ = [("ncirc", C.c_int64), ("nsmall", C.c_int64), ("npasses", C.c_int64),
                ("nops", C.c_int64), ("nlevels", C.c_int64), ("total_amplitudes", C.c_uint64),
                ("opts", QsvBatchOpts), ("alg_bytes", C.c_double)]


class QsvTuneResult(C.Structure):
    """Mirror of `qsv_tune_result` (include/qsv.h)."""

    _fields_ = [("opts", QsvBatchOpts), ("ms", C.c_double), ("max_abs_diff", C.c_double)]


class QsvCacheStats(C.Structure):
    """Mirror of `qsv_cache_stats` (include/qsv.h)."""

    _fields_ = [("schedule_hits", C.c_int64), ("schedule_misses", C.c_int64),
                ("module_hits", C.c_int64), ("module_disk_hits", C.c_int64),
                ("module_compiles", C.c_int64)]


def cache_stats() -> dict:
    """The compile caches' counters (qsv_cache_stats_get)."""
    st = QsvCacheStats()
    check(lib().qsv_cache_stats_get(C.byref(st)))
    return {k: getattr(st, k) for k, _ in QsvCacheStats._fields_}


class QsvAdjointProfile(C.Structure):
    """Mirror of `qsv_adjoint_profile` (include/qsv.h)."""

    _fields_ = [("forward_ms", C.c_double), ("hamiltonian_ms", C.c_double), ("reverse_ms", C.c_double),
                ("reverse_launches", C.c_int64), ("compiled", C.c_int32), ("reserved", C.c_int32)]


class QsvProgramInfo(C.Structure):
    """Mirror of `qsv_program_info` (include/qsv.h)."""

    _fields_ = [
        ("ngates", C.c_int64),
        ("nlaunch", C.c_int64),
        ("nplain", C.c_int64),
        ("ndiag", C.c_int64),
        ("ndense", C.c_int64 * 6),
        ("alg_bytes", C.c_double),
        ("flops", C.c_double),
        ("ntiled_phases", C.c_int64),
        ("ntiled_relabel", C.c_int64),
        ("fp_lane_ops", C.c_double),
        ("work_bytes", C.c_double),
    ]


class QsvError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"qsv error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Load the engine; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP engine first "
                "(python -m qclabpp_amd.build)")
        # torch ships its own libamdhip64 (soname libamdhip64.so.7, loaded
        # under the file name libamdhip64.so).  Loaded first, it satisfies the
        # engine's libamdhip64.so.7 dependency, so the process has one HIP
        # runtime; loading the engine first would map /opt/rocm's runtime as
        # well, and the second runtime then finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def check(status: int) -> None:
    if status != QSV_OK:
        msg = lib().qsv_last_error().decode(errors="replace")
        raise QsvError(status, msg)


_P = C.c_void_p
_SZ = C.c_size_t
_I = C.c_int
_U64 = C.c_uint64
_D = C.c_double
_GP = C.POINTER(QsvGate)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); every function declared in include/qsv.h.
SIGNATURES = {
    "qsv_last_error": (C.c_char_p, []),
    "qsv_version": (C.c_char_p, []),
    "qsv_device_count": (_I, []),
    "qsv_dev_set_basis": (_I, [_P, _I, _I, _U64, _P]),
    "qsv_dev_apply_gate": (_I, [_P, _I, _I, _GP, _P]),
    "qsv_dev_apply_gates": (_I, [_P, _I, _I, _GP, _SZ, _P]),
    "qsv_host_apply_gates": (_I, [_P, _I, _I, _GP, _SZ]),
    "qsv_program_create": (_I, [_I, _I, _GP, _SZ, _I, _I, C.POINTER(_P)]),
    "qsv_program_run": (_I, [_P, _P, _P]),
    "qsv_tiled_jit_compile_check": (_I, [_I, _I, _GP, _SZ, C.POINTER(_SZ), C.POINTER(C.c_int64),
                                         C.POINTER(C.c_uint64)]),
    "qsv_tiled_jit_compile_check_form": (_I, [_I, _I, _GP, _SZ, _I, C.POINTER(_SZ), C.POINTER(C.c_int64),
                                              C.POINTER(C.c_uint64)]),
    "qsv_program_info_get": (_I, [_P, C.POINTER(QsvProgramInfo)]),
    "qsv_program_op": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), _DP]),
    "qsv_program_destroy": (_I, [_P]),
    "qsv_release_cache": (_I, []),
    "qsv_cache_stats_get": (_I, [C.POINTER(QsvCacheStats)]),
    "qsv_host_simulate": (_I, [_P, _I, _I, _GP, _SZ, _I]),
    "qsv_dev_simulate": (_I, [_P, _I, _I, _GP, _SZ, _I, _P]),
    "qsv_dev_expectation": (_I, [_P, _I, _I, C.POINTER(QsvPauliTerm), _SZ, _DP, _DP, _P]),
    "qsv_dev_expectation_noalloc": (_I, [_P, _I, _I, C.POINTER(QsvPauliTerm), _SZ, _DP, _DP, _P]),
    "qsv_dev_inner": (_I, [_P, _P, _I, _I, _DP, _DP, _P]),
    "qsv_dev_apply_pauli_sum": (_I, [_P, _P, _I, _I, C.POINTER(QsvPauliTerm), _SZ, _P]),
    "qsv_batch_create": (_I, [_I, _I, C.POINTER(_I), _GP, C.POINTER(_SZ), C.POINTER(QsvBatchOpts),
                              C.POINTER(_P)]),
    "qsv_batch_info": (_I, [_P, C.POINTER(QsvBatchInfo), C.POINTER(_U64)]),
    "qsv_batch_run": (_I, [_P, _P, _I, _P]),
    "qsv_batch_latencies": (_I, [_P, _DP, _P]),
    "qsv_batch_destroy": (_I, [_P]),
    "qsv_batch_autotune": (_I, [_I, _I, C.POINTER(_I), _GP, C.POINTER(_SZ), _P, _P, _I,
                                C.POINTER(QsvBatchOpts), C.POINTER(QsvTuneResult), _I,
                                C.POINTER(_I), _P]),
    "qsv_dev_max_abs_diff": (_I, [_P, _P, _U64, _I, _DP, _P]),
    "qsv_dev_circuit_matrix": (_I, [_P, _I, _I, _GP, _SZ, _I, _P]),
    "qsv_host_circuit_matrix": (_I, [_P, _I, _I, _GP, _SZ, _I]),
    "qsv_dev_nrmf": (_I, [_P, _P, _U64, _I, _DP, _P]),
    "qsv_circuit_nrmf": (_I, [_I, _I, _GP, _SZ, _GP, _SZ, _I, _DP]),
    "qsv_adjoint_profile_get": (_I, [C.POINTER(QsvAdjointProfile)]),
    "qsv_adjoint_compile_check": (_I, [_I, _I, _GP, _SZ, C.POINTER(_SZ), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64)]),
    "qsv_dev_adjoint_gradient": (_I, [_P, _P, _I, _I, _GP, _SZ, _DP, _SZ, C.POINTER(QsvPauliTerm),
                                      _SZ, _I, _DP, _DP, _P]),
}


def _declare(l: C.CDLL) -> None:
    for name, (res, args) in SIGNATURES.items():
        f = getattr(l, name)
        f.restype = res
        f.argtypes = args


def header_functions() -> list[str]:
    """Names of the functions include/qsv.h declares."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qsv_[a-z0-9_]+)\s*\(", text)))
