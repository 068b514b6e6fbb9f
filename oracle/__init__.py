"""ORACLE -- test infrastructure only.

ctypes access to the CPU restatement (liboracle.so) and to the reference
compiled from its own sources (_ref/libqclab_ref.so).  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this package,
and only to check or time against it.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import subprocess
from pathlib import Path

import numpy as np

from qclabpp_amd._lib import QsvGate
from qclabpp_amd.gates import U2, records

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libqclab_ref.so"
REFERENCE_TREE = Path("/root/reference")

_o = None
_r = None


def build(with_ref: bool | None = None) -> None:
    """Compile the restatement and, where the reference tree exists, _ref."""
    targets = ["oracle"]
    if with_ref or (with_ref is None and (REFERENCE_TREE / "src" / "qgates").is_dir()):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", str(HERE), "-j8", *targets], check=True)


def oracle_lib() -> C.CDLL:
    global _o
    if _o is None:
        if not ORACLE_SO.exists():
            build(with_ref=False)
        _o = C.CDLL(str(ORACLE_SO))
        gp = C.POINTER(QsvGate)
        _o.oracle_apply_gate.argtypes = [C.c_int, C.c_int, gp, C.c_void_p]
        _o.oracle_apply_gates.argtypes = [C.c_int, C.c_int, gp, C.c_size_t, C.c_void_p]
        _o.oracle_apply_gates_op.argtypes = [C.c_int, C.c_int, gp, C.c_size_t, C.c_int, C.c_void_p]
        _declare_extra(_o)
    return _o


def _declare_extra(o) -> None:
    from qclabpp_amd._lib import QsvPauliTerm
    gp = C.POINTER(QsvGate)
    dp = C.POINTER(C.c_double)
    tp = C.POINTER(QsvPauliTerm)
    o.oracle_expectation.argtypes = [C.c_int, C.c_int, C.c_void_p, tp, C.c_size_t, dp]
    o.oracle_energy.argtypes = [C.c_int, gp, C.c_size_t, dp, tp, C.c_size_t, dp]
    o.oracle_param_shift.argtypes = [C.c_int, gp, C.c_size_t, dp, C.c_size_t, tp, C.c_size_t,
                                     C.POINTER(C.c_int), C.c_size_t, dp]


def ref_available() -> bool:
    return REF_SO.exists()


def ref_lib() -> C.CDLL:
    global _r
    if _r is None:
        _r = C.CDLL(str(REF_SO))
        gp = C.POINTER(QsvGate)
        _r.ref_apply_gate.argtypes = [C.c_int, C.c_int, gp, C.c_void_p]
        _r.ref_circuit_apply.argtypes = [C.c_int, C.c_int, gp, C.c_size_t, C.c_int, C.c_void_p,
                                         C.POINTER(C.c_double)]
        _r.ref_circuit_matrix.argtypes = [C.c_int, C.c_int, gp, C.c_size_t, C.c_void_p]
        _r.ref_time_simulate_zero.argtypes = [C.c_int, C.c_int, gp, C.c_size_t,
                                              C.POINTER(C.c_double)]
        _r.ref_qasm_gates.argtypes = [C.c_char_p, gp, C.c_size_t, C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_int)]
        _r.ref_circuit_qasm.argtypes = [C.c_int, gp, C.c_size_t, C.c_char_p, C.c_size_t,
                                        C.POINTER(C.c_size_t)]
    return _r


def ref_qasm_gates(path) -> tuple[int, list]:
    """The reference's io::QASMFile on a file: (nbQubits, gates), each gate
    a tuple (kind, q0, q1, ctrl_state, theta) as qasm_records.hpp makes it."""
    from qclabpp_amd._lib import QsvGate
    cap = 4096
    arr = (QsvGate * cap)()
    cnt = C.c_size_t()
    nq = C.c_int()
    st = ref_lib().ref_qasm_gates(str(path).encode(), arr, cap, C.byref(cnt), C.byref(nq))
    if st:
        raise RuntimeError(f"reference QASM reader status {st}")
    out = [(g.kind, g.q[0], g.q[1], g.ctrl_state, g.theta) for g in arr[:min(cnt.value, cap)]]
    return nq.value, out


def ref_circuit_qasm(n: int, gates) -> str:
    """The reference's QCircuit::toQASM text of a gate list (complex<double>)."""
    arr, cnt = records(gates)
    ln = C.c_size_t()
    st = ref_lib().ref_circuit_qasm(n, arr, cnt, None, 0, C.byref(ln))
    if st:
        raise RuntimeError(f"reference QASM writer status {st}")
    buf = C.create_string_buffer(ln.value + 1)
    st = ref_lib().ref_circuit_qasm(n, arr, cnt, buf, ln.value + 1, C.byref(ln))
    if st:
        raise RuntimeError(f"reference QASM writer status {st}")
    return buf.value.decode()


def ref_time_simulate_zero(n: int, gates, dtype=np.complex128) -> float:
    """Wall seconds of the reference's QCircuit::simulate on |0...0>."""
    arr, cnt = records(gates)
    secs = C.c_double(0)
    st = ref_lib().ref_time_simulate_zero(n, _dtype_code(np.zeros(1, dtype)), arr, cnt,
                                          C.byref(secs))
    if st:
        raise RuntimeError(f"reference driver status {st}")
    return secs.value


def ref_threads() -> int:
    return int(ref_lib().ref_max_threads())


def _dtype_code(arr: np.ndarray) -> int:
    if arr.dtype == np.complex128:
        return 2
    if arr.dtype == np.complex64:
        return 1
    raise TypeError(f"unsupported dtype {arr.dtype}")


def _ptr(arr: np.ndarray):
    assert arr.flags.c_contiguous
    return arr.ctypes.data_as(C.c_void_p)


def apply(n: int, gates, psi: np.ndarray, op: str | None = None) -> np.ndarray:
    """Oracle run of `gates` on a copy of psi.  op=None applies each gate with
    its own op, in order; op='N'/'T'/'C' is QCircuit::apply(op): every gate
    with that op, in reverse order for 'T'/'C' (QCircuit.hpp:130-143)."""
    out = np.ascontiguousarray(psi).copy()
    arr, cnt = records(gates)
    if op is None:
        st = oracle_lib().oracle_apply_gates(n, _dtype_code(out), arr, cnt, _ptr(out))
    else:
        st = oracle_lib().oracle_apply_gates_op(n, _dtype_code(out), arr, cnt, ord(op), _ptr(out))
    if st:
        raise RuntimeError(f"oracle status {st}")
    return out


def matrix(k: int, gates, dtype=np.complex128) -> np.ndarray:
    """QCircuit::matrix() restated (include/qclab/QCircuit.hpp:121-127): the
    2^k x 2^k unitary, column c = the gates applied to |c> by the restated
    sweeps.  A QGate2 record acts there in the transposed vector basis of
    src/qgates/apply.hpp:289-308, while matrix() applies QGate2::matrix() in
    kron order (src/qgates/QGate2.cpp:32-109); so each QGate2 record is
    given P M P (basis states 01 and 10 exchanged), which the vector sweep
    turns back into M in kron order.  Every other gate acts identically in
    both.  Pure-Python column loop: for k <= 8."""
    N = 1 << k
    recs = []
    for g in gates:
        if g.kind == U2:
            p = [0, 2, 1, 3]
            g = dataclasses.replace(g, mat=np.asarray(g.mat, dtype=np.complex128)[np.ix_(p, p)])
        recs.append(g)
    out = np.zeros((N, N), dtype=dtype)
    for c in range(N):
        e = np.zeros(N, dtype=dtype)
        e[c] = 1
        out[:, c] = apply(k, recs, e)
    return out


def apply_gate(n: int, gate, psi: np.ndarray, op: str | None = None) -> np.ndarray:
    out = np.ascontiguousarray(psi).copy()
    rec = gate.record(op)
    st = oracle_lib().oracle_apply_gate(n, _dtype_code(out), C.byref(rec), _ptr(out))
    if st:
        raise RuntimeError(f"oracle status {st}")
    return out


def ref_apply_gate(n: int, gate, psi: np.ndarray, op: str | None = None) -> np.ndarray:
    out = np.ascontiguousarray(psi).copy()
    rec = gate.record(op)
    st = ref_lib().ref_apply_gate(n, _dtype_code(out), C.byref(rec), _ptr(out))
    if st:
        raise RuntimeError(f"reference driver status {st}")
    return out


def ref_apply(n: int, gates, psi: np.ndarray, op: str = "N"):
    """Reference QCircuit(n) + simulate/apply(op); returns (state, seconds)."""
    out = np.ascontiguousarray(psi).copy()
    arr, cnt = records(gates)
    secs = C.c_double(0)
    st = ref_lib().ref_circuit_apply(n, _dtype_code(out), arr, cnt, ord(op), _ptr(out),
                                     C.byref(secs))
    if st:
        raise RuntimeError(f"reference driver status {st}")
    return out, secs.value


def ref_matrix(n: int, gates, dtype=np.complex128) -> np.ndarray:
    """Reference QCircuit::matrix(): returns the 2^n x 2^n unitary."""
    N = 1 << n
    out = np.zeros(N * N, dtype=dtype)
    arr, cnt = records(gates)
    st = ref_lib().ref_circuit_matrix(n, _dtype_code(out), arr, cnt, _ptr(out))
    if st:
        raise RuntimeError(f"reference driver status {st}")
    return out.reshape(N, N, order="F")


def expectation(n: int, psi: np.ndarray, H) -> np.ndarray:
    """Per-term <psi|P_t|psi> through the restated PauliX/Y/Z gates."""
    psi = np.ascontiguousarray(psi)
    arr, cnt = H.records()
    vals = (C.c_double * max(1, cnt))()
    st = oracle_lib().oracle_expectation(n, _dtype_code(psi), _ptr(psi), arr, cnt, vals)
    if st:
        raise RuntimeError(f"oracle status {st}")
    return np.array(vals[:cnt])


def _theta_arg(theta):
    th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64))
    return th, th.ctypes.data_as(C.POINTER(C.c_double))


def energy(n: int, gates, H, theta) -> float:
    arr, cnt = records(gates)
    harr, hcnt = H.records()
    th, thp = _theta_arg(theta)
    e = C.c_double()
    st = oracle_lib().oracle_energy(n, arr, cnt, thp, harr, hcnt, C.byref(e))
    if st:
        raise RuntimeError(f"oracle status {st}")
    return e.value


def param_shift(n: int, gates, H, theta, which) -> np.ndarray:
    """CPU parameter-shift dE/dtheta_p for p in `which`."""
    arr, cnt = records(gates)
    harr, hcnt = H.records()
    th, thp = _theta_arg(theta)
    w = (C.c_int * max(1, len(which)))(*which)
    g = (C.c_double * max(1, len(which)))()
    st = oracle_lib().oracle_param_shift(n, arr, cnt, thp, th.size, harr, hcnt, w, len(which), g)
    if st:
        raise RuntimeError(f"oracle status {st}")
    return np.array(g[:len(which)])
