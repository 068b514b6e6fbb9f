"""Build the HIP engine (libqclabpp_amd.so) for gfx950, in-tree.

Usage: python -m qclabpp_amd.build   (or build() from __graft_entry__)

Every call recompiles every source: the result is the "does it build" check
as well as the library that tests, smoke() and bench.py load.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = ROOT / "_build" / "obj"
LIB = PKG / "libqclabpp_amd.so"

ARCH = os.environ.get("QSV_OFFLOAD_ARCH", "gfx950")

# No packed-FP32 instructions anywhere in the device code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP engine cannot be built")


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    out = OBJ / (src.name + ".o")
    cmd = [_hipcc(), *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return out


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    workers = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


CPP_TESTS = ROOT / "tests" / "cpp"
CPP_TEST_BIN = CPP_TESTS / "_bin" / "dropin_test"


def build_cpp_tests(verbose: bool = True) -> Path:
    """Compile the C++ drop-in test program against the built library.

    The binary finds libqclabpp_amd.so through a relative rpath, so the tree
    can be copied elsewhere and still run it.
    """
    CPP_TEST_BIN.parent.mkdir(parents=True, exist_ok=True)
    cmd = [_hipcc(), "-std=c++17", "-O1", f"-I{ROOT / 'include'}", f"-I{ROOT / 'oracle'}",
           str(CPP_TESTS / "dropin_test.cpp"), "-o", str(CPP_TEST_BIN),
           f"-L{PKG}", "-lqclabpp_amd", "-Wl,-rpath,$ORIGIN/../../../qclabpp_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"C++ test build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {CPP_TEST_BIN}", file=sys.stderr)
    return CPP_TEST_BIN


if __name__ == "__main__":
    build()
    build_cpp_tests()
