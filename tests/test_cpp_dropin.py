"""The QCLAB++-compatible C++ classes (include/qclab) over the engine.

The program tests/cpp/dropin_test.cpp checks every gate class against the
dense product of its own matrix() for Op N/T/C, circuits (nesting, offsets,
pointer gates, [Conj]Trans, matrix() columns, device vectors) and the QASM
readers; the reference's own tests of these classes are test/qgates/*.cpp and
test/QCircuit.cpp.  It is built by __graft_entry__.build().
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "tests" / "cpp" / "_bin" / "dropin_test"
LIB = ROOT / "qclabpp_amd" / "libqclabpp_amd.so"


def _binary() -> Path:
    if not BIN.exists():
        raise RuntimeError(f"{BIN} is missing: run __graft_entry__.build() first")
    return BIN


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_headers_compile_with_gpp():
    """The headers compile warning-free with g++ (the reference's compiler)."""
    src = ROOT / "tests" / "cpp" / "headers_gpp.cpp"
    res = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                          f"-I{ROOT / 'include'}", str(src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_host_checks():
    """QASM readers, matrices, equality and QASM output: no GPU needed."""
    if not LIB.exists():
        pytest.skip("engine library not built")
    res = subprocess.run([str(_binary()), "--host"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failed" in res.stdout


@pytest.mark.gpu
def test_dropin_on_gpu(tmp_path):
    """Every gate class, circuits and device vectors on the GPU, and the
    default compiled form at 22 qubits with its caches (compiled code
    objects go to a fresh directory)."""
    import os
    print(f"running {BIN}", flush=True)
    env = dict(os.environ, QSV_JIT_CACHE_DIR=str(tmp_path))
    res = subprocess.run([str(_binary())], capture_output=True, text=True, timeout=300, env=env)
    print(res.stdout[-4000:])
    assert res.returncode == 0, res.stdout[-4000:] + res.stderr[-2000:]
    assert "0 failed" in res.stdout


@pytest.mark.gpu
def test_cpp_brickwork_timing_mode(tmp_path):
    """The C++ timing mode the benchmark uses: a seeded brickwork built from
    the drop-in gate classes, simulated twice on a std::vector; the second
    call reuses the schedule and compiled passes of the first."""
    import json
    import os
    env = dict(os.environ, QSV_JIT_CACHE_DIR=str(tmp_path))
    res = subprocess.run([str(_binary()), "--time-brickwork", "22", "4", "20261202"], capture_output=True,
                         text=True, timeout=300, env=env)
    assert res.returncode == 0, res.stdout + res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["gates"] == 4 * 22 + 2 * 11 + 2 * 10 and abs(out["norm"] - 1) < 1e-9
    assert out["schedule_hits"] >= 1 and out["module_hits"] >= 1
