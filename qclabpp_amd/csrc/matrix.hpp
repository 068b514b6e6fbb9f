// Unitary of a circuit on the GPU (QCircuit::matrix(), reference
// include/qclab/QCircuit.hpp:121-127) and the Frobenius distance of two
// unitaries (nrmF, QCircuit.hpp:406-430).
//
// A 2^k x 2^k column-major matrix is a state of 2k qubits: entry (r, c) sits
// at address c*2^k + r.  In MSB qubit order the column index then holds
// qubits 0..k-1 and the row index qubits k..2k-1, and column c of U is U|c>.
// So U is the circuit, its qubits shifted by k, applied to the identity, and
// every kernel and compiled form of simulate() runs on it unchanged.
#pragma once

#include <vector>

#include "common.hpp"

namespace qsv {

// Gate records of a k-qubit circuit moved to the row qubits k..2k-1 of the
// 2k-qubit matrix state.  QGate2 records (QSV_GATE_U2) are re-indexed so
// that they act in the kron order of QGate2::matrix(), as the reference's
// QCircuit::matrix() applies them (QGate2::apply(Side, ...),
// src/qgates/QGate2.cpp:22-60), rather than in the vector basis of
// simulate() (src/qgates/apply.hpp:289-308).
int matrix_records(int k, const qsv_gate* gates, size_t ngates, std::vector<qsv_gate>& out);

// Writes the 2^k x 2^k identity (4^k amplitudes) into m.
int launch_set_identity(void* m, int k, int dtype, hipStream_t st);

// Scaled Frobenius norm of a - b over `count` amplitudes: scale = max |a_i - b_i|
// (1 if all are equal), result scale * sqrt(sum |a_i - b_i|^2 / scale^2), all in
// FP64 with per-block partial sums added in block order (deterministic).
int dev_nrmf(const void* a, const void* b, uint64_t count, int dtype, double* out, hipStream_t st);

}  // namespace qsv
