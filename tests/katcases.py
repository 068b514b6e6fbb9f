"""Known-answer cases of QCLAB++'s own tests (tests/golden/kat_vectors.json)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from qclabpp_amd import gates as G

KAT = Path(__file__).resolve().parent / "golden" / "kat_vectors.json"

_CTOR = {
    "Identity": G.Identity, "Hadamard": G.Hadamard, "PauliX": G.PauliX, "PauliY": G.PauliY,
    "PauliZ": G.PauliZ, "Phase45": G.Phase45, "Phase90": G.Phase90,
    "RotationX": G.RotationX, "RotationY": G.RotationY, "RotationZ": G.RotationZ, "Phase": G.Phase,
    "CX": G.CX, "CNOT": G.CX, "CY": G.CY, "CZ": G.CZ,
    "CPhase": G.CPhase, "CRotationX": G.CRotationX, "CRotationY": G.CRotationY,
    "CRotationZ": G.CRotationZ, "SWAP": G.SWAP, "iSWAP": G.iSWAP,
    "RotationXX": G.RotationXX, "RotationYY": G.RotationYY, "RotationZZ": G.RotationZZ,
}


def make_gate(d: dict):
    cls, q = d["cls"], d["q"]
    ctor = _CTOR[cls]
    kw = {}
    if "cs" in d:
        if "theta" in d:
            kw["theta"] = d["theta"]   # let the engine derive (cos, sin) with libm
        else:
            kw["cos"], kw["sin"] = d["cs"]
    if "cstate" in d:
        kw["controlState"] = d["cstate"]
    return ctor(*q, **kw)


def load():
    data = json.loads(KAT.read_text())
    out = []
    for c in data["gate_cases"]:
        vin = np.array([complex(a, b) for a, b in c["input"]], dtype=np.complex128)
        vexp = np.array([complex(a, b) for a, b in c["expected"]], dtype=np.complex128)
        out.append((make_gate(c["gate"]), c["op"], c["n"], vin, vexp, c["exact"], c["where"]))
    circ = []
    for c in data["circuit_cases"]:
        gates = [make_gate(g) for g in c["gates"]]
        vin = np.array([complex(a, b) for a, b in c["input"]], dtype=np.complex128)
        vexp = np.array([complex(a, b) for a, b in c["expected"]], dtype=np.complex128)
        circ.append((gates, c["op"], c["n"], vin, vexp, c["where"]))
    return out, circ
