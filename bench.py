"""Benchmark: 30-qubit random brickwork (depth 20, complex128) on one MI355X.

A step is one full simulation of the 890-gate circuit on the resident
2^30-amplitude state (16 GiB).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the engine is single-GPU; sharding is a design note only).", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from qclabpp_amd import engine as E
    from qclabpp_amd import workloads as W
    from qclabpp_amd._lib import lib

    n, depth = args.n, args.depth
    circ = W.brickwork(n, depth)
    nfull = sum(1 for g in circ if g.nb_qubits == 1)
    nhalf = len(circ) - nfull
    S = (1 << n) * 16
    alg_bytes = nfull * 2 * S + nhalf * S

    psi = E.new_state(n, torch.complex128)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        E.apply_gates(psi, circ)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        E.apply_gates(psi, circ)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    gates_per_s = len(circ) / (ms / 1e3)

    # Dominant kernel: dense 1-qubit sweep (RX/RY); time launches on a
    # high-stride qubit with HIP events on the launch stream.
    from qclabpp_amd import gates as G
    probe = [G.RotationX(3, 0.1)] * 10
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    E.apply_gates(psi, probe[:2])
    e0.record(stream)
    E.apply_gates(psi, probe)
    e1.record(stream)
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / len(probe)
    k_bytes = 2 * S
    achieved = k_bytes / (k_ms / 1e3) / 1e9

    out = {
        "metric": "30q brickwork gate-passes/s (complex128)",
        "value": gates_per_s,
        "unit": "gate-passes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded brickwork, |0...0>)",
        "config": {"workload": f"brickwork n={n} depth={depth} plain per-gate", "gates": len(circ)},
        "hbm_GBps_plain": alg_bytes / (ms / 1e3) / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": None},
        "cpu_baseline": None,
        "wall_s": wall,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
