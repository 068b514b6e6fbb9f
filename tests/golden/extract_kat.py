"""Write tests/golden/kat_vectors.json: the reference's known-answer vectors.

Reads QCLAB++'s own gate tests (test/qgates/*.cpp under the reference tree)
as text and records, for every host-side `gate.apply(op, n, vec)` followed by
a check, the gate (class, qubits, angle, control state), the Op, n, the
input vector and the expected vector the test asserts.  The C++ statements of
those test bodies are interpreted by a small evaluator (declarations of
vectors, scalars, gates and their setters); offload (#ifdef
QCLAB_OMP_OFFLOADING) blocks are skipped since they repeat the host checks.

The fixture is data: inputs and expected outputs.  It also records the
circuit-level checks of test/QCircuit.cpp:163-236.

Usage: python tests/golden/extract_kat.py [/root/reference]
"""
from __future__ import annotations

import cmath
import json
import math
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent / "kat_vectors.json"

ONE_FIXED = {"Identity", "Hadamard", "PauliX", "PauliY", "PauliZ", "Phase45", "Phase90"}
ONE_ROT = {"RotationX", "RotationY", "RotationZ", "Phase"}
CTRL_FIXED = {"CX", "CNOT", "CY", "CZ"}
CTRL_ROT = {"CPhase", "CRotationX", "CRotationY", "CRotationZ"}
TWO_FIXED = {"SWAP", "iSWAP"}
TWO_ROT = {"RotationXX", "RotationYY", "RotationZZ"}
HALF = {"RotationX", "RotationY", "RotationZ", "CRotationX", "CRotationY", "CRotationZ",
        "RotationXX", "RotationYY", "RotationZZ"}  # QRotation: theta/2


class Angle:
    """QAngle / QRotation value: holds (cos, sin) of the stored angle."""

    def __init__(self, c, s):
        self.c, self.s = c, s


def strip_offload(text: str) -> str:
    out, skip = [], 0
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("#ifdef QCLAB_OMP_OFFLOADING"):
            skip += 1
            out.append("")
            continue
        if skip and s.startswith("#if"):
            skip += 1
            out.append("")
            continue
        if skip and s.startswith("#endif"):
            skip -= 1
            out.append("")
            continue
        out.append("" if skip else re.sub(r"//.*", "", line))  # keep line numbers
    return "\n".join(out)


def py_expr(e: str) -> str:
    # library functions get a prefix so test variables named cos/sin don't shadow them
    e = re.sub(r"std::(\w+)\s*\(", r"_f_\1(", e)
    e = e.replace("std::", "").replace("qclab::", "")
    e = re.sub(r"\bT\s*\(", "_T(", e)
    e = re.sub(r"\bR\s*\(", "float(", e)
    e = e.replace("{", "[").replace("}", "]")
    return e


class Interp:
    def __init__(self):
        self.ns = {"_T": lambda a=0.0, b=0.0: complex(a, b), "_f_cos": math.cos, "_f_sin": math.sin,
                   "_f_sqrt": math.sqrt, "_f_atan": math.atan, "_f_exp": cmath.exp,
                   "_f_conj": lambda z: complex(z).conjugate(), "_f_abs": abs, "M_PI": math.pi}
        self.vecs = {}
        self.gates = {}
        self.pending = {}
        self.cases = []
        self.skipped = []

    def ev(self, e):
        return eval(py_expr(e), {"__builtins__": {"float": float, "complex": complex, "int": int}}, self.ns)

    # ---- gates -------------------------------------------------------------
    def declare_gate(self, cls, name, args):
        a = [x.strip() for x in split_args(args)] if args.strip() else []
        g = {"cls": cls}
        if cls in ONE_FIXED:
            g["q"] = [int(self.ev(a[0])) if a else 0]
        elif cls in ONE_ROT:
            if not a:
                g.update(q=[0], cs=[1.0, 0.0])
            elif len(a) == 1:
                g.update(q=[0], **self.angle(cls, a[0]))
            elif len(a) == 2 and not re.fullmatch(r"\d+", a[0]):
                g.update(q=[0], cs=[float(self.ev(a[0])), float(self.ev(a[1]))])
            elif len(a) == 2:
                g.update(q=[int(self.ev(a[0]))], **self.angle(cls, a[1]))
            else:
                g.update(q=[int(self.ev(a[0]))], cs=[float(self.ev(a[1])), float(self.ev(a[2]))])
        elif cls in CTRL_FIXED:
            g.update(q=[int(self.ev(a[0])), int(self.ev(a[1]))] if a else [0, 1],
                     cstate=int(self.ev(a[2])) if len(a) > 2 else 1)
        elif cls in CTRL_ROT:
            if not a:
                g.update(q=[0, 1], cs=[1.0, 0.0], cstate=1)
            else:
                q = [int(self.ev(a[0])), int(self.ev(a[1]))]
                rest = a[2:]
                if len(rest) >= 2 and "cos" in rest[0]:
                    g.update(q=q, cs=[float(self.ev(rest[0])), float(self.ev(rest[1]))],
                             cstate=int(self.ev(rest[2])) if len(rest) > 2 else 1)
                else:
                    g.update(q=q, **self.angle(cls, rest[0]),
                             cstate=int(self.ev(rest[1])) if len(rest) > 1 else 1)
        elif cls in TWO_FIXED:
            q = [int(self.ev(a[0])), int(self.ev(a[1]))] if a else [0, 1]
            g["q"] = sorted(q)
        elif cls in TWO_ROT:
            if not a:
                g.update(q=[0, 1], cs=[1.0, 0.0])
            elif len(a) == 3:
                g.update(q=sorted([int(self.ev(a[0])), int(self.ev(a[1]))]), **self.angle(cls, a[2]))
            elif len(a) == 4:
                g.update(q=sorted([int(self.ev(a[0])), int(self.ev(a[1]))]),
                         cs=[float(self.ev(a[2])), float(self.ev(a[3]))])
            else:
                raise ValueError(f"{cls} args {a}")
        else:
            raise ValueError(f"unsupported class {cls}")
        self.gates[name] = g

    def angle(self, cls, expr):
        v = self.ev(expr)
        if isinstance(v, Angle):
            return {"cs": [v.c, v.s]}
        th = float(v)
        if cls in HALF:
            return {"cs": [math.cos(th / 2), math.sin(th / 2)], "theta": th}
        return {"cs": [math.cos(th), math.sin(th)], "theta": th}

    def setter(self, name, meth, arg):
        g = self.gates[name]
        cls = g["cls"]
        if meth == "setQubit":
            g["q"] = [int(self.ev(arg))]
        elif meth == "setQubits":
            arr = self.ns[re.match(r"&\s*(\w+)", arg.strip()).group(1)]
            if not isinstance(arr, list):
                arr = [arr]
            if cls in CTRL_FIXED | CTRL_ROT:
                g["q"] = [int(arr[0]), int(arr[1])]
            elif cls in ONE_FIXED | ONE_ROT:
                g["q"] = [int(arr[0])]
            else:
                g["q"] = sorted([int(arr[0]), int(arr[1])])
        elif meth == "setControl":
            g["q"][0] = int(self.ev(arg))
        elif meth == "setTarget":
            g["q"][1] = int(self.ev(arg))
        elif meth == "setControlState":
            g["cstate"] = int(self.ev(arg))
        elif meth == "update":
            g.update(self.angle(cls, arg))
        # makeFixed / makeVariable / setOffset do not change apply().

    # ---- statements ----------------------------------------------------------
    def run(self, stmt: str, where: str):
        s = " ".join(stmt.split())
        while s.startswith("}") or s.startswith("{"):
            s = s[1:].strip()
        if not s or s.startswith("#") or s.startswith("using ") or s.startswith("template"):
            return
        m = re.fullmatch(r"qclab::qgates::(\w+)\s*<\s*T\s*>\s+(\w+)\s*(?:\((.*)\))?", s)
        if m:
            try:
                self.declare_gate(m.group(1), m.group(2), m.group(3) or "")
            except Exception as e:  # noqa: BLE001 - record and continue
                self.gates.pop(m.group(2), None)
                self.skipped.append((where, s, str(e)))
            return
        m = re.fullmatch(r"qclab::(QAngle|QRotation)\s*<\s*R\s*>\s+(\w+)\s*\((.*)\)", s)
        if m:
            a = [x.strip() for x in split_args(m.group(3))]
            if len(a) == 1:
                th = float(self.ev(a[0]))
                h = th / 2 if m.group(1) == "QRotation" else th
                self.ns[m.group(2)] = Angle(math.cos(h), math.sin(h))
            else:
                self.ns[m.group(2)] = Angle(float(self.ev(a[0])), float(self.ev(a[1])))
            return
        m = re.fullmatch(r"(?:const\s+)?auto\s+(\w+)\s*=\s*(\w+)\.qubits\(\s*\)", s)
        if m and m.group(2) in self.gates:
            self.ns[m.group(1)] = sorted(self.gates[m.group(2)]["q"])
            return
        m = re.fullmatch(r"(\w+)\.(setQubits?|setControl|setTarget|setControlState|update|makeFixed|makeVariable|setOffset)\s*\((.*)\)", s)
        if m and m.group(1) in self.gates:
            self.setter(m.group(1), m.group(2), m.group(3))
            return
        m = re.fullmatch(r"(\w+)\.apply\(\s*qclab::Op::(\w+)\s*,\s*(\w+)\s*,\s*(\w+)\s*\)", s)
        if m:
            gname, op, n, vec = m.groups()
            if gname not in self.gates:
                self.skipped.append((where, s, "unknown gate"))
                return
            g = dict(self.gates[gname])
            g["q"] = list(g["q"])
            self.pending[vec] = {"gate": g, "op": {"NoTrans": "N", "Trans": "T", "ConjTrans": "C"}[op],
                                 "n": int(self.ev(n)), "input": list(self.vecs[vec])}
            return
        m = re.fullmatch(r"EXPECT_TRUE\(\s*(\w+)\s*==\s*(\w+)\s*\)", s)
        m2 = re.fullmatch(r"(?:EXPECT_TRUE\(\s*)?(?:check_\w+|test_\w+_check)\(\s*(\w+)\s*,\s*(\w+)\s*\)\s*\)?", s)
        if m or m2:
            exact = m is not None
            vec, chk = (m or m2).groups()
            if vec in self.pending and chk in self.vecs:
                c = dict(self.pending[vec])
                c.update(expected=list(self.vecs[chk]), exact=exact, where=where)
                self.cases.append(c)
            return
        m = re.fullmatch(r"(?:const\s+)?std::vector\s*<\s*int\s*>\s+(\w+)\s*=\s*(\{.*\})", s)
        if m:
            self.ns[m.group(1)] = [int(x) for x in self.ev(m.group(2))]
            return
        m = re.fullmatch(r"FORINT (\w+) (\d+) (\d+) @@ (\w+)\[\1\]\s*=\s*(.+)", s)
        if m:
            var, a, b, vec, rhs = m.groups()
            for k in range(int(a), int(b)):
                self.ns[var] = k
                self.vecs[vec][k] = complex(self.ev(rhs))
            self.ns.pop(var, None)
            return
        m = re.fullmatch(r"for\s*\(\s*auto\s*&\s*(\w+)\s*:\s*(\w+)\s*\)\s*(.+)", s)
        if m:
            var, idx, body = m.groups()
            for k in self.ns[idx]:
                for part in body.split("@@"):
                    pm = re.fullmatch(r"(\w+)\[(.+?)\]\s*\*=\s*(.+)", part.strip())
                    if not pm:
                        raise ValueError(f"loop body {part}")
                    vec, ie, val = pm.groups()
                    i = int(eval(ie, {"__builtins__": {}}, {var: k}))
                    self.vecs[vec][i] = self.vecs[vec][i] * complex(self.ev(val))
            return
        m = re.fullmatch(r"(?:const\s+)?V\s+(\w+)\s*\(\s*(\d+)\s*\)", s)
        if m:
            self.vecs[m.group(1)] = [0j] * int(m.group(2))
            return
        m = re.fullmatch(r"iota\(\s*(\w+)\.begin\(\)\s*,\s*\w+\.end\(\)\s*,\s*(\d+)\s*\)", s.replace("std::", ""))
        if m:
            v = self.vecs[m.group(1)]
            self.vecs[m.group(1)] = [complex(int(m.group(2)) + i) for i in range(len(v))]
            return
        m = re.fullmatch(r"(?:const\s+)?(?:V|auto)\s+(\w+)\s*=\s*(.+)", s) or \
            re.fullmatch(r"(\w+)\s*=\s*(.+)", s)
        if m:
            name, rhs = m.groups()
            rhs = rhs.strip()
            if rhs in self.vecs:
                self.vecs[name] = list(self.vecs[rhs])
                self.pending.pop(name, None)
                return
            vm = re.fullmatch(r"V\s*\(\s*(\d+)\s*\)", rhs)
            if vm:
                self.vecs[name] = [0j] * int(vm.group(1))
                return
            if rhs.startswith("{"):
                self.vecs[name] = [complex(x) for x in self.ev(rhs)]
                return
            im = re.fullmatch(r"(\w+)\[(\d+)\]", name)
            if im:
                self.ns[im.group(1)][int(im.group(2))] = int(self.ev(rhs))
                return
        m = re.fullmatch(r"(\w+)\[(\d+)\]\s*=\s*([^=].*)", s)
        if m and isinstance(self.ns.get(m.group(1)), list):
            self.ns[m.group(1)][int(m.group(2))] = int(self.ev(m.group(3)))
            return
        m = re.fullmatch(r"int\s+(\w+)\s*\[\s*\d*\s*\]\s*=\s*(\{.*\})", s)
        if m:
            self.ns[m.group(1)] = [int(x) for x in self.ev(m.group(2))]
            return
        m = re.fullmatch(r"(?:const\s+)?(?:R|T|double|float|int|auto)\s+(\w+)\s*=\s*(.+)", s)
        if m:
            try:
                self.ns[m.group(1)] = self.ev(m.group(2))
            except Exception:  # noqa: BLE001 - not a value this extractor needs
                pass
            return


def split_args(s: str) -> list:
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return out


def join_for_bodies(text: str) -> str:
    """`for ( auto& i: idx ) { a ; b ; }` -> one statement with '@@' between parts."""
    def rep(m):
        body = "@@".join(x.strip() for x in m.group(3).split(";") if x.strip())
        return f"for ( auto& {m.group(1)}: {m.group(2)} ) {body} ;"
    text = re.sub(r"for\s*\(\s*auto\s*&\s*(\w+)\s*:\s*(\w+)\s*\)\s*\{([^}]*)\}", rep, text)
    # `for ( int i = a; i < b; ++i ) stmt ;` -> one statement
    return re.sub(r"for\s*\(\s*int\s+(\w+)\s*=\s*(\d+)\s*;\s*\1\s*<\s*(\d+)\s*;\s*\+\+\1\s*\)\s*([^;]*);",
                  lambda m: f"FORINT {m.group(1)} {m.group(2)} {m.group(3)} @@ {m.group(4).strip()} ;", text)


def extract_file(path: Path):
    text = join_for_bodies(strip_offload(path.read_text()))
    it = Interp()
    # statements end at ';'; line of the statement start is kept for citations
    pos = 0
    for stmt in text.split(";"):
        line = text.count("\n", 0, pos) + 1
        pos += len(stmt) + 1
        try:
            it.run(stmt, f"test/qgates/{path.name}:{line}")
        except Exception as e:  # noqa: BLE001 - statements outside the apply checks
            it.skipped.append((f"test/qgates/{path.name}:{line}", " ".join(stmt.split())[:80], repr(e)))
    return it


def circuit_cases():
    """test/QCircuit.cpp:163-236: circuit (H, Y, Z on qubit 0) applied with
    NoTrans / ConjTrans on 1- and 2-qubit vectors, and simulate()."""
    def cd(vals):
        return [[complex(v).real, complex(v).imag] for v in vals]
    gates = [{"cls": "Hadamard", "q": [0]}, {"cls": "PauliY", "q": [0]}, {"cls": "PauliZ", "q": [0]}]
    return [
        {"gates": gates, "op": "N", "n": 1, "input": cd([3, 5]),
         "expected": cd([1.414213562373095j, -5.656854249492380j]), "where": "test/QCircuit.cpp:175-182"},
        {"gates": gates, "op": "C", "n": 1, "input": cd([3, 5]),
         "expected": cd([5.656854249492380j, 1.414213562373095j]), "where": "test/QCircuit.cpp:184-190"},
        {"gates": gates, "op": "N", "n": 2, "input": cd([3, 5, 2, 7]),
         "expected": cd([-0.707106781186547j, 1.414213562373094j, -3.535533905932737j,
                         -8.485281374238570j]), "where": "test/QCircuit.cpp:192-205"},
        {"gates": gates, "op": "C", "n": 2, "input": cd([3, 5, 2, 7]),
         "expected": cd([3.535533905932737j, 8.485281374238570j, -0.707106781186547j,
                         1.414213562373094j]), "where": "test/QCircuit.cpp:207-219"},
        {"gates": gates, "op": "N", "n": 1, "input": cd([3, 5]),
         "expected": cd([1.414213562373095j, -5.656854249492380j]), "where": "test/QCircuit.cpp:222-236 (simulate)"},
    ]


def main():
    files = sorted((REF / "test" / "qgates").glob("*.cpp"))
    allcases, skipped, counts = [], [], {}
    for f in files:
        it = extract_file(f)
        counts[f.name] = len(it.cases)
        for c in it.cases:
            c["input"] = [[z.real, z.imag] for z in c["input"]]
            c["expected"] = [[z.real, z.imag] for z in c["expected"]]
        allcases += it.cases
        skipped += it.skipped
    data = {"source": "QCLAB++ test/qgates/*.cpp host vector-apply checks; test/QCircuit.cpp:163-236",
            "counts": counts, "gate_cases": allcases, "circuit_cases": circuit_cases()}
    OUT.write_text(json.dumps(data, separators=(",", ":")))
    print(f"{len(allcases)} gate cases from {len(files)} files -> {OUT}")
    for k, v in counts.items():
        print(f"  {k}: {v}")
    for s in skipped:
        print("skipped:", s)


if __name__ == "__main__":
    main()
