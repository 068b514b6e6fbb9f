"""The QASM reader and writer against the reference compiled from its own
sources (oracle/_ref).

* Reference dialect: every file of tests/golden/qasm/ is read by the
  reference's io::QASMFile (include/qclab/io/QASMFile.hpp:42-152,
  src/io/QASMFile.cpp:16-53) and by this repository's reader
  (include/qclab/io/QASMFile.hpp, Dialect::Reference); both gate lists are
  turned into records by the same code (oracle/qasm_records.hpp) and must be
  equal, angles bit for bit.  The corpus covers comments, creg/barrier/
  measure and unknown gates (skipped), several registers (only the first
  counts), a register name longer than one letter, `pi` (the reference stops
  reading at the first angle it cannot convert), a bad qubit index, missing
  ';' between commands on one line, CRLF line ends and tabs, and number
  forms std::stod reads by prefix (2*0.5 -> 2, 1.5pi -> 1.5).
  Not in the corpus: a last command without ';' and an empty command between
  two ';' (the reference's string helpers never return on them; the
  repository's reader ends the command there), and a trailing comment on a
  gate line (for the same reason).
* Writer: circuits of every gate the reference writes go through the
  reference's QCircuit::toQASM (qasm.hpp, "%.15f" angles), are read back by
  the extended dialect and written again by this repository's toQASM: the
  two texts must be equal, and the state of the circuit read back must
  match the original's to 1e-13 relative (the angles carry 15 decimals).
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from qclabpp_amd import gates as G

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "tests" / "cpp" / "_bin" / "dropin_test"
CORPUS = sorted((ROOT / "tests" / "golden" / "qasm").glob("*.qasm"))


def _need():
    if not oracle.ref_available():
        pytest.skip("reference library oracle/_ref not built (needs the reference tree)")
    if not BIN.exists():
        pytest.skip("tests/cpp/_bin/dropin_test not built")


def _ours(path, dialect):
    res = subprocess.run([str(BIN), "--qasm-dump", str(path), dialect], capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = res.stdout.strip().splitlines()
    assert lines[0].startswith("nq ")
    gates = []
    for ln in lines[1:]:
        k, a, b, cs, th = ln.split()
        gates.append((int(k), int(a), int(b), int(cs), float(th)))
    return int(lines[0].split()[1]), gates


def test_corpus_present():
    assert len(CORPUS) >= 10


@pytest.mark.parametrize("path", CORPUS, ids=[p.stem for p in CORPUS])
def test_reference_dialect_matches_reference_reader(path):
    _need()
    nq_ref, ref = oracle.ref_qasm_gates(path)
    nq, ours = _ours(path, "ref")
    assert nq == nq_ref
    assert ours == ref, f"\nref:  {ref}\nours: {ours}"


def test_corpus_exercises_truncation_and_prefix_reads():
    """Pins on the corpus itself, so a file edit that drops a case fails."""
    _need()
    got = {p.stem: oracle.ref_qasm_gates(p)[1] for p in CORPUS}
    assert len(got["pi_truncates"]) == 3 and len(got["pi_midline"]) == 1
    assert len(got["bad_qubit"]) == 1
    assert [g[4] for g in got["number_forms"][4:6]] == [2.0, 1.5]
    assert len(got["skips_unknown"]) == 2
    assert got["two_registers"][0][:2] == (G.H, 0) and len(got["two_registers"]) == 3


def _writable(n, seed):
    """Every gate kind the reference writes, on scattered qubits; both
    control states; no Phase (written as rz, another gate up to a global
    phase) and no generic matrices (no QASM form)."""
    rs = np.random.default_rng(seed)
    th = lambda: float(rs.uniform(-np.pi, np.pi))
    out = []
    for q in range(n):
        out += [G.Hadamard(q), G.RotationX(q, th()), G.RotationY(q, th()), G.RotationZ(q, th())]
    for i in range(2 * n):
        a, b = (int(x) for x in rs.choice(n, 2, replace=False))
        cs = int(rs.integers(2))
        out += [G.CX(a, b, cs), G.CY(a, b, cs), G.CZ(a, b, cs), G.CPhase(a, b, th(), controlState=cs),
                G.CRotationX(a, b, th(), controlState=cs), G.CRotationY(a, b, th(), controlState=cs),
                G.CRotationZ(a, b, th(), controlState=cs), G.SWAP(a, b), G.iSWAP(a, b),
                G.RotationXX(min(a, b), max(a, b), th()), G.RotationYY(min(a, b), max(a, b), th()),
                G.RotationZZ(min(a, b), max(a, b), th()), G.PauliX(a), G.PauliY(b), G.PauliZ(a),
                G.Phase45(b), G.Phase90(a), G.Identity(b)]
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_writer_round_trip_against_reference_writer(seed, tmp_path):
    _need()
    n = 6
    circ = _writable(n, seed)
    text = oracle.ref_circuit_qasm(n, circ)
    f = tmp_path / "ref_written.qasm"
    f.write_text(f"OPENQASM 2.0;\ninclude \"qelib1.inc\";\nqreg q[{n}];\n" + text)
    res = subprocess.run([str(BIN), "--qasm-rewrite", str(f)], capture_output=True, text=True,
                         timeout=60)
    assert res.returncode == 0, res.stderr
    assert res.stdout == text  # this repository's writer reproduces the reference's text
    nq, back = _ours(f, "ext")
    assert nq == n
    # controlled gates on |0> are written as x, c-gate, x; the identity
    # writes nothing (Identity.hpp:68-73 of the reference)
    assert len(back) == len(circ) + 2 * sum(1 for g in circ if g.ctrl_state == 0 and len(g.qubits) == 2
                                            and g.kind in G.CONTROLLED) \
        - sum(1 for g in circ if g.kind == G.I)
    gates = [G.Gate(k, (a,) if k in G.ONE_QUBIT else (a, b), theta=t, ctrl_state=cs)
             for k, a, b, cs, t in back]
    rs = np.random.default_rng(seed)
    psi = rs.standard_normal(1 << n) + 1j * rs.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    want = oracle.apply(n, circ, psi)
    got = oracle.apply(n, gates, psi)
    rel = np.abs(got - want).max() / np.abs(want).max()
    assert rel < 1e-13, rel
