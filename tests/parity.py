"""Parity bars shared by the GPU tests.

Errors are measured relative to the largest amplitude of the expected
state, so the bar means the same at every qubit count: on a unit-norm
state of 2^22 amplitudes the entries are about 1e-3, and an absolute bound
of 1e-12 would only ask for about 1e-9 relative.
"""
from __future__ import annotations

import numpy as np

# complex128 against the oracle: 1e-12 relative, the bar of the path.
C128 = 1e-12
# complex64 (FP32 gate arithmetic) against the complex128 oracle.
C64 = 5e-5


def rel_err(got, want) -> float:
    got = np.asarray(got)
    want = np.asarray(want)
    scale = float(np.abs(want).max()) if want.size else 0.0
    diff = float(np.abs(got.astype(np.complex128) - want).max()) if want.size else 0.0
    return diff / scale if scale > 0 else diff


def assert_close(got, want, rel: float, what: str = "") -> float:
    e = rel_err(got, want)
    assert e <= rel, f"{what}: max|got - want| / max|want| = {e:.3e} > {rel:.0e}"
    return e


def tol(dt: str) -> float:
    return C128 if dt == "c128" else C64
