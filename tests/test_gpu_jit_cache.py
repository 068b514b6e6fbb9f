"""Compiled tiled passes are cached on disk across processes: a second
process that builds a program of a known circuit structure loads the code
objects instead of compiling them with hiprtc, and runs them correctly."""
import json
import os
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

# simulates one circuit through the host-vector entry with compiled passes
# (no torch), and prints the cache counters and the final state
SCRIPT = r"""
import ctypes as C, json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from qclabpp_amd import _lib, gates as G, workloads as W
n = 14
scale = float(sys.argv[2])
circ = [g.with_theta(g.theta * scale) if g.theta else g for g in W.brickwork(n, 8, seed=3)]
arr, cnt = G.records(circ)
v = np.zeros(1 << n, np.complex128)
v[0] = 1
_lib.check(_lib.lib().qsv_host_simulate(v.ctypes.data_as(C.c_void_p), n, _lib.QSV_C128, arr, cnt,
                                        _lib.QSV_TILED_JIT))
np.save(sys.argv[3], v)
print(json.dumps(_lib.cache_stats()))
"""


def _circuit(scale):
    from qclabpp_amd import workloads as W
    return [g.with_theta(g.theta * scale) if g.theta else g for g in W.brickwork(14, 8, seed=3)]


def _run(cache_dir, scale, out):
    env = dict(os.environ, QSV_JIT_CACHE_DIR=str(cache_dir))
    res = subprocess.run([sys.executable, "-c", SCRIPT, str(ROOT), str(scale), str(out)],
                         capture_output=True, text=True, timeout=120, env=env)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    return json.loads(res.stdout.strip().splitlines()[-1])


def test_code_objects_cached_on_disk(tmp_path):
    import numpy as np

    import oracle
    cache = tmp_path / "jit"
    first = _run(cache, 1.0, tmp_path / "a.npy")
    print("first process:", first, flush=True)
    assert first["module_compiles"] > 0 and first["module_disk_hits"] == 0
    files = sorted(p.name for p in cache.iterdir())
    assert files and all(f.startswith("qtp_") and f.endswith(".co") for f in files), files
    # same structure, other angles: loaded from disk, nothing compiled
    second = _run(cache, 0.7, tmp_path / "b.npy")
    print("second process:", second, flush=True)
    assert second["module_compiles"] == 0
    assert second["module_disk_hits"] == first["module_compiles"]
    n = 14
    psi0 = np.zeros(1 << n, np.complex128)
    psi0[0] = 1
    for scale, name in ((1.0, "a.npy"), (0.7, "b.npy")):
        want = oracle.apply(n, _circuit(scale), psi0)
        got = np.load(tmp_path / name)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), scale
