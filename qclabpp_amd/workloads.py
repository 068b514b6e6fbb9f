"""Seeded circuit families of the benchmark and parity suites.

All randomness comes from qclabpp_amd.rng.SplitMix64, drawn in the order
stated per family, so every host builds bit-identical circuits.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from . import gates as G
from .rng import SplitMix64

SEED_QFT, SEED_BRICK, SEED_PREC, SEED_BATCH, SEED_ADJ = 20261201, 20261202, 20261203, 20261204, 20261205


def qft(n: int, sign: int = +1) -> list:
    """QFT as in test/timings/run_qft.cpp:10-33, with phase angle
    sign*2*pi/2^k: Hadamard(j), then CPhase(control=j+k-1, target=j) for
    k = 2..n-j, then SWAP(j, n-1-j) for j < n/2.  sign=+1 maps x to
    sqrt(N)*ifft(x); sign=-1 (the reference harness's choice) to fft(x)/sqrt(N)."""
    out = []
    for j in range(n):
        out.append(G.Hadamard(j))
        for k in range(2, n - j + 1):
            out.append(G.CPhase(j + k - 1, j, sign * 2 * math.pi / (1 << k)))
    for j in range(n // 2):
        out.append(G.SWAP(j, n - j - 1))
    return out


def brickwork(n: int, depth: int, seed: int = SEED_BRICK) -> list:
    """Random brickwork: per layer, a rotation on every qubit (axis uniform
    over X/Y/Z, then angle uniform in [0, 2pi), drawn in (layer, qubit)
    order), then CNOT(i, i+1) for i of the layer's parity."""
    rng = SplitMix64(seed)
    out = []
    ctor = (G.RotationX, G.RotationY, G.RotationZ)
    for layer in range(depth):
        for q in range(n):
            axis = min(int(rng.uniform() * 3), 2)
            theta = rng.uniform() * 2 * math.pi
            out.append(ctor[axis](q, theta))
        for i in range(layer % 2, n - 1, 2):
            out.append(G.CX(i, i + 1))
    return out


def trotter(n: int) -> list:
    """The reference's Trotter timing circuit (test/timings/run_trotter.cpp:8-35)."""
    out = []
    pi = math.pi
    for lyr in range(n):
        os_ = 1 if lyr % 2 else 0
        os_c = (1 - os_) if n % 2 else os_
        for i in range(os_, n - os_c):
            out.append(G.RotationZ(i, pi))
        for i in range(os_, n - os_c - 1, 2):
            out += [G.RotationX(i, pi / 2), G.RotationX(i + 1, pi / 2), G.CX(i, i + 1),
                    G.RotationX(i, pi), G.RotationZ(i + 1, pi), G.CX(i, i + 1),
                    G.RotationX(i, -pi / 2), G.RotationX(i + 1, -pi / 2)]
    return out


def hea(n: int, layers: int, rng: SplitMix64, first_param: int = 0) -> list:
    """Hardware-efficient ansatz: per layer RY then RZ on every qubit
    (angles uniform in [-pi, pi)), then CZ(i, i+1) for i = 0..n-2."""
    out = []
    p = first_param
    for _ in range(layers):
        for q in range(n):
            out.append(G.RotationY(q, rng.uniform_range(-math.pi, math.pi), param=p))
            out.append(G.RotationZ(q, rng.uniform_range(-math.pi, math.pi), param=p + 1))
            p += 2
        for i in range(n - 1):
            out.append(G.CZ(i, i + 1))
    return out


@dataclass
class BatchSpec:
    nq: list
    layers: list
    circuits: list


def hea_batch(count: int = 512, seed: int = SEED_BATCH, nmin: int = 8, nmax: int = 22,
              lmin: int = 4, lmax: int = 16) -> BatchSpec:
    """`count` HEA circuits; per circuit n ~ U{nmin..nmax} then L ~ U{lmin..lmax},
    then its angles, all from one stream."""
    rng = SplitMix64(seed)
    nq, ls, circs = [], [], []
    for _ in range(count):
        n = rng.integer(nmin, nmax)
        L = rng.integer(lmin, lmax)
        nq.append(n)
        ls.append(L)
        circs.append(hea(n, L, rng))
    return BatchSpec(nq, ls, circs)


def tfim_ansatz(n: int, layers: int, seed: int = SEED_ADJ) -> list:
    """Per layer: RY on every qubit, RZ on every qubit, RotationZZ(i, i+1)
    on the n-1 nearest-neighbour pairs; one parameter per gate, angles
    uniform in [-pi, pi) in circuit order.  n=24, 10 layers -> 710 params."""
    rng = SplitMix64(seed)
    out = []
    p = 0
    for _ in range(layers):
        for q in range(n):
            out.append(G.RotationY(q, rng.uniform_range(-math.pi, math.pi), param=p))
            p += 1
        for q in range(n):
            out.append(G.RotationZ(q, rng.uniform_range(-math.pi, math.pi), param=p))
            p += 1
        for i in range(n - 1):
            out.append(G.RotationZZ(i, i + 1, rng.uniform_range(-math.pi, math.pi), param=p))
            p += 1
    return out


def params_of(circuit: list) -> list:
    """theta of every parametrised gate, indexed by its `param`."""
    ps = {g.param: g.theta for g in circuit if g.param >= 0}
    return [ps[i] for i in range(len(ps))]
