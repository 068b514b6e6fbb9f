/*
 * qsv.h -- C ABI of the MI355X state-vector engine (libqclabpp_amd.so).
 *
 * This is the drop-in boundary for QCLAB++'s simulation hot path:
 *   QCircuit<T>::simulate / apply(Op, nbQubits, std::vector<T>&, offset)
 *     (include/qclab/QCircuit.hpp:130-143,164-166)
 *   -> QObject<T>::apply (include/qclab/QObject.hpp:52-53)
 *   -> per-gate sweeps apply2/apply2b/apply4/apply4b/apply4bc
 *     (src/qgates/apply.hpp:258-371), and their OpenMP-offload twins
 *     apply_device* / simulate_device (src/qgates/apply.hpp:373-532,
 *     include/qclab/QCircuit.hpp:145-173).
 *
 * Conventions (identical to the reference):
 *   - qubit 0 is the most significant bit: qubit q has amplitude stride
 *     2^(n-1-q)                                    (src/qgates/apply.hpp:264)
 *   - rotations carry (cos(theta/2), sin(theta/2)); Phase/CPhase carry
 *     (cos(theta), sin(theta))      (include/qclab/QRotation.hpp:44-46,
 *                                    include/qclab/qgates/Phase.hpp:51-55)
 *   - Op 'N','T','C' applies M, M^T, M^H of the gate's matrix
 *                                    (include/qclab/dense/transpose.hpp:12-35)
 *   - generic matrices are column-major complex (re,im interleaved),
 *     element (i,j) at [2*(i + j*size)]    (include/qclab/dense/SquareMatrix.hpp:125-132)
 *   - a QGate2 matrix (QSV_GATE_U2) is applied as QGate2::apply does: for
 *     qubits q0 < q1, matrix position p addresses the amplitude with
 *     bit(q0) = p & 1 and bit(q1) = p >> 1 (src/qgates/apply.hpp:289-308).
 *     For a matrix that is not symmetric under exchanging the two qubits
 *     this differs from the kron order of QGate2::matrix() and
 *     QCircuit::matrix(); the built-in 2-qubit gates are all symmetric.
 *
 * Errors: every function returns an int status (QSV_OK == 0); no exception
 * crosses this boundary.  qsv_last_error() returns a thread-local message
 * for the last failing call on the calling thread.
 *
 * Streams: entry points that take a `stream` argument enqueue their work on
 * that HIP stream (NULL = the default stream) and return once it is
 * enqueued, except where a host result (a number, a host buffer) is
 * returned: those synchronise the stream first.
 */
#ifndef QSV_H
#define QSV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define QSV_OK        0
#define QSV_EINVAL    1  /* bad argument: null pointer, size, op code      */
#define QSV_EQUBIT    2  /* qubit index out of range or repeated           */
#define QSV_ENOMEM    3  /* device or host allocation failed               */
#define QSV_EHIP      4  /* a HIP runtime call failed                      */
#define QSV_EDTYPE    5  /* unsupported element type                       */
#define QSV_ENODEV    6  /* no usable GPU                                  */
#define QSV_EUNSUP    7  /* valid request this build does not implement    */

/* ---- element types ----------------------------------------------------- */
#define QSV_C64   1   /* std::complex<float>  : 8 bytes per amplitude  */
#define QSV_C128  2   /* std::complex<double> : 16 bytes per amplitude */

/* ---- Op (include/qclab/util.hpp:15-20) ---------------------------------- */
#define QSV_OP_N  'N'
#define QSV_OP_T  'T'
#define QSV_OP_C  'C'

/* ---- gate kinds: one per QCLAB++ gate class ------------------------------ */
enum qsv_gate_kind {
  QSV_GATE_I     = 0,   /* Identity        include/qclab/qgates/Identity.hpp     */
  QSV_GATE_H     = 1,   /* Hadamard        Hadamard.hpp:46-51                    */
  QSV_GATE_X     = 2,   /* PauliX          PauliX.hpp:46-49                      */
  QSV_GATE_Y     = 3,   /* PauliY          PauliY.hpp:46-49                      */
  QSV_GATE_Z     = 4,   /* PauliZ          PauliZ.hpp:46-49                      */
  QSV_GATE_RX    = 5,   /* RotationX       RotationX.hpp:104-108                 */
  QSV_GATE_RY    = 6,   /* RotationY       RotationY.hpp:104-107                 */
  QSV_GATE_RZ    = 7,   /* RotationZ       RotationZ.hpp:104-108                 */
  QSV_GATE_P     = 8,   /* Phase           Phase.hpp:119-122                     */
  QSV_GATE_T     = 9,   /* Phase45         Phase45.hpp:46-51                     */
  QSV_GATE_S     = 10,  /* Phase90         Phase90.hpp:46-49                     */
  QSV_GATE_U1    = 11,  /* QGate1 / PointerGate1: dense 2x2 from mat[]           */
  QSV_GATE_CX    = 12,  /* CX / CNOT       CX.hpp, CNOT.hpp:11-12                */
  QSV_GATE_CY    = 13,  /* CY                                                    */
  QSV_GATE_CZ    = 14,  /* CZ                                                    */
  QSV_GATE_CP    = 15,  /* CPhase          CPhase.hpp:196-203                    */
  QSV_GATE_CRX   = 16,  /* CRotationX                                            */
  QSV_GATE_CRY   = 17,  /* CRotationY                                            */
  QSV_GATE_CRZ   = 18,  /* CRotationZ                                            */
  QSV_GATE_CU    = 19,  /* QControlledGate2: controlled 2x2 from mat[]           */
  QSV_GATE_SWAP  = 20,  /* SWAP            SWAP.hpp:68-73                        */
  QSV_GATE_ISWAP = 21,  /* iSWAP           iSWAP.hpp:68-74                       */
  QSV_GATE_RXX   = 22,  /* RotationXX      RotationXX.hpp:136-143                */
  QSV_GATE_RYY   = 23,  /* RotationYY      RotationYY.hpp:136-143                */
  QSV_GATE_RZZ   = 24,  /* RotationZZ      RotationZZ.hpp:136-143                */
  QSV_GATE_U2    = 25,  /* QGate2: dense 4x4 from mat[]                          */
  QSV_GATE_COUNT = 26
};

/*
 * One gate application.  The same record drives the GPU engine, the CPU
 * oracle and the reference driver, so a circuit is a plain array of these.
 *
 *   q[0]      : 1-qubit gates: the qubit.  Controlled gates: the control.
 *               2-qubit gates: the first qubit.
 *   q[1]      : controlled gates: the target.  2-qubit gates: second qubit.
 *   ctrl_state: control state 0 or 1 (QControlledGate2.hpp:29-36)
 *   op        : QSV_OP_N / QSV_OP_T / QSV_OP_C
 *   param     : index of this gate's angle in a parameter vector, or -1
 *   has_cs    : 1 -> use (cs[0], cs[1]) = (cos, sin) verbatim, as the
 *               (cos,sin) constructors do; 0 -> derive them from theta with
 *               the host libm, as QRotation(theta)/QAngle(theta) do
 *   theta     : the rotation / phase parameter
 *   mat       : U1/CU: 2x2, U2: 4x4, column-major complex (re,im)
 */
typedef struct qsv_gate {
  int32_t kind;
  int32_t q[2];
  int32_t ctrl_state;
  int32_t op;
  int32_t param;
  int32_t has_cs;
  int32_t reserved;
  double  theta;
  double  cs[2];
  double  mat[32];
} qsv_gate;

/*
 * One term coeff * P of a Pauli-sum observable, P a tensor product of
 * nfactors single-qubit Paulis ('X', 'Y', 'Z'; 'I' factors are ignored) on
 * distinct qubits.  QCLAB++ has no observables; Paulis follow its PauliX /
 * PauliY / PauliZ matrices (PauliX.hpp:46-49, PauliY.hpp:46-49,
 * PauliZ.hpp:46-49) and its MSB qubit order.
 */
#define QSV_MAX_PAULI_FACTORS 16
typedef struct qsv_pauli_term {
  double  coeff;
  int32_t nfactors;
  int32_t reserved;
  int32_t qubits[QSV_MAX_PAULI_FACTORS];
  char    paulis[QSV_MAX_PAULI_FACTORS];
} qsv_pauli_term;

/* ---- library ------------------------------------------------------------ */
const char* qsv_last_error(void);
/* Version string; also names the GPU architecture the code objects target. */
const char* qsv_version(void);
/* Number of visible GPUs (0 when none); never fails. */
int qsv_device_count(void);

/* ---- device-resident state (the simulate_device(T*) analogue) ------------
 * `psi` is a device pointer to 2^n amplitudes of `dtype`.  These calls never
 * allocate; they enqueue kernels on `stream`.                               */

/* psi := |index> */
int qsv_dev_set_basis(void* psi, int n, int dtype, uint64_t index, void* stream);

/* Apply one gate with the plain per-gate kernel of its sweep class
 * (one launch; the QObject::apply_device analogue).                         */
int qsv_dev_apply_gate(void* psi, int n, int dtype, const qsv_gate* gate,
                       void* stream);

/* Apply `ngates` gates in order (op 'N' per gate record), one plain launch
 * per gate (QCircuit::simulate_device, include/qclab/QCircuit.hpp:170-172). */
int qsv_dev_apply_gates(void* psi, int n, int dtype, const qsv_gate* gates,
                        size_t ngates, void* stream);

/* ---- compiled circuits ----------------------------------------------------
 * A program is a gate list lowered once to kernel operations, optionally
 * fused: with fuse_k in 1..5 runs of gates collapse into dense 2^k x 2^k
 * blocks (applied with MFMA for k >= 3) and runs of diagonal gates into
 * phase tables over up to diag_k qubits (diag_k <= 10; 0 picks 8).
 * fuse_k == 0 keeps one plain launch per gate.
 * fuse_k == QSV_TILED runs the circuit as tiled passes: each launch streams
 * the state once through LDS tiles of 2^12 amplitudes and applies every gate
 * of the pass there, one by one in registers (states of fewer than 12
 * qubits use fuse_k = 5 instead).  Coefficients live on the device for the
 * program's lifetime; running it allocates nothing.                       */
#define QSV_TILED (-1)
/* fuse_k == QSV_AUTO picks QSV_TILED_JIT from 22 qubits up (at least 1024
 * tiles to spread over the GPU), QSV_TILED there where hiprtc cannot be
 * loaded, and fuse_k = 5 below 22 qubits.                                */
#define QSV_AUTO  (-2)
/* fuse_k == QSV_TILED_JIT runs the same tiled passes as QSV_TILED, each pass
 * compiled at program creation into a kernel of its own (hiprtc, gfx950)
 * with the gate structure as constants: faster passes for a compile of a
 * few seconds, done once per circuit structure.  Programs whose gates
 * differ only in angles share the schedule and the compiled passes (both
 * cached per process; the code objects also on disk, see
 * qsv_cache_stats).  Creation fails with QSV_EUNSUP where hiprtc cannot be
 * loaded.                                                                 */
#define QSV_TILED_JIT (-3)
/* fuse_k == QSV_TILED_JIT_FIXED compiles the same passes as QSV_TILED_JIT
 * with each rotation's scaled form (which of |cos| and |sin| is factored
 * out of its 2x2) taken from its angle at creation, so the passes carry no
 * branch per rotation: faster passes for circuits simulated repeatedly with
 * the same angles.  Programs share compiled passes only where their
 * rotations also share these forms (a new angle set usually compiles anew);
 * parameter sweeps use QSV_TILED_JIT.  No counterpart in the reference.    */
#define QSV_TILED_JIT_FIXED (-4)
typedef struct qsv_program qsv_program;

typedef struct qsv_program_info {
  int64_t ngates;      /* gates given                                   */
  int64_t nlaunch;     /* kernel launches per run                       */
  int64_t nplain;      /* launches that are single (unfused) gates      */
  int64_t ndiag;       /* merged diagonal passes (k >= 3)               */
  int64_t ndense[6];   /* dense fused blocks by size k (k >= 3)         */
  double  alg_bytes;   /* algorithmic HBM bytes of one run              */
  double  flops;       /* useful flops of the dense fused blocks        */
  int64_t ntiled_phases; /* tiled programs: register phases of one run  */
  int64_t ntiled_relabel; /* compiled tiled passes that write their tile
                             with qubits moved to other address bits     */
  double  fp_lane_ops; /* tiled programs: floating-point lane operations
                          (FMA and multiply; FP64 for complex128, FP32 for
                          complex64) of one run's gate arithmetic, as the
                          compiled passes generate it                    */
  double  work_bytes;  /* compiled tiled programs whose relabeling ends out
                          of place: device bytes of the program's work
                          buffer (one state); 0 otherwise.  Runs of such a
                          program on different streams must not overlap */
} qsv_program_info;

int qsv_program_create(int n, int dtype, const qsv_gate* gates, size_t ngates, int fuse_k,
                       int diag_k, qsv_program** out);
int qsv_program_run(const qsv_program* prog, void* psi, void* stream);
/* Build check of QSV_TILED_JIT without a GPU: schedule the circuit into
 * tiled passes, generate every pass kernel and, when code_bytes is given,
 * compile it for gfx950 with hiprtc, without loading it.  *code_bytes
 * (optional) receives the total code-object size, *npasses (optional) the
 * pass count and *source_hash (optional) a 64-bit FNV-1a hash of the
 * generated source of every pass, which is what the compiled-module cache
 * is keyed by.  No counterpart in the reference; for build systems and CI
 * hosts without a device.                                                  */
int qsv_tiled_jit_compile_check(int n, int dtype, const qsv_gate* gates, size_t ngates,
                                size_t* code_bytes, int64_t* npasses, uint64_t* source_hash);
/* The same for fuse_k = QSV_TILED_JIT or QSV_TILED_JIT_FIXED (the check
 * above is this one with QSV_TILED_JIT).  No counterpart in the reference. */
int qsv_tiled_jit_compile_check_form(int n, int dtype, const qsv_gate* gates, size_t ngates,
                                     int fuse_k, size_t* code_bytes, int64_t* npasses,
                                     uint64_t* source_hash);
int qsv_program_info_get(const qsv_program* prog, qsv_program_info* out);
/* Launch i of one run (0-based, launch order): the qubits it acts on
 * (ascending, k of them), and its operator as interleaved complex doubles --
 * the 2^k x 2^k matrix row-major, or only its diagonal when *diag is set.
 * Row bit p of the matrix is qubit qubits[k-1-p] (MSB convention of
 * QCircuit::matrix(), include/qclab/QCircuit.hpp:121-127).  `coeffs` needs
 * 2*4^k doubles.  Returns QSV_EINVAL past the last launch.                 */
int qsv_program_op(const qsv_program* prog, int i, int* k, int* qubits, int* diag,
                   double* coeffs);
int qsv_program_destroy(qsv_program* prog);

/* ---- observables and gradients (no counterpart in the reference) ---------
 * All reductions accumulate in FP64 (also for QSV_C64 states) on a fixed
 * grid, so results are reproducible bit for bit.  These calls synchronise
 * `stream` before returning their host results.                           */

/* values[t] = <psi|P_t|psi> (coefficient not applied; may be NULL);
 * *total = sum_t coeff_t values[t] (may be NULL).                          */
int qsv_dev_expectation(const void* psi, int n, int dtype, const qsv_pauli_term* terms,
                        size_t nterms, double* values, double* total, void* stream);
/* qsv_dev_expectation without any device allocation: the reduction
 * partials live in a fixed array in the library's own device memory (one
 * per GPU, 4 MiB, held until the process ends) instead of the library's
 * growable scratch.  No counterpart in the reference; for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  At most 256 terms (QSV_EINVAL beyond).  Calls
 * on one GPU share that array: calls that overlap on one GPU (other host
 * threads, other streams) race.                                            */
int qsv_dev_expectation_noalloc(const void* psi, int n, int dtype, const qsv_pauli_term* terms,
                                size_t nterms, double* values, double* total, void* stream);
/* <a|b> = sum conj(a_x) b_x. */
int qsv_dev_inner(const void* a, const void* b, int n, int dtype, double* re, double* im,
                  void* stream);
/* out = sum_t coeff_t P_t in  (out must not alias in; at most 1024 terms). */
int qsv_dev_apply_pauli_sum(const void* in, void* out, int n, int dtype,
                            const qsv_pauli_term* terms, size_t nterms, void* stream);
/* Adjoint gradient of E = <psi|H|psi>, psi = U(theta) psi0.
 *   psi    : in: psi0; out: psi0 recomputed by the reverse sweep
 *   lam    : scratch state of the same size
 *   theta  : nparams angles overriding each parametrised gate's theta
 *            (gate.param indexes theta), or NULL to use the records' own
 *   fuse_k : form of the forward pass (a qsv_program_create fuse_k;
 *            QSV_AUTO picks the fastest)
 *   energy : E;  grad[p] = dE/dtheta_p summed over the gates sharing p
 * Differentiable kinds: RX RY RZ RXX RYY RZZ CRX CRY CRZ P CP, Op 'N'.
 * The reverse sweep runs as passes compiled for the circuit's structure
 * over both states from 12 qubits up where hiprtc loads (the schedule and
 * kernels cached like QSV_TILED_JIT programs), else as interpreted LDS
 * tile passes; QSV_ADJOINT=interp|jit in the environment forces one.
 * The compiled sweep's schedule and kernels are prepared on one extra host
 * thread, started and joined within the call, while the forward pass and
 * H|phi> run on `stream`; the call returns after the stream is idle.      */
int qsv_dev_adjoint_gradient(void* psi, void* lam, int n, int dtype, const qsv_gate* gates,
                             size_t ngates, const double* theta, size_t nparams,
                             const qsv_pauli_term* terms, size_t nterms, int fuse_k,
                             double* energy, double* grad, void* stream);
/* GPU time of the last qsv_dev_adjoint_gradient call, by phase (HIP events
 * on its stream): the forward pass; lam = H psi and the energy; the reverse
 * sweep's launches alone (from the first to the last reverse kernel, no host
 * work in between).  `compiled` is 1 when the reverse sweep ran as compiled
 * two-state passes, 0 for the interpreted LDS passes.  Each reverse launch
 * reads and writes both states once (4 * 2^n * sizeof(complex)).  No
 * counterpart in the reference.                                            */
typedef struct qsv_adjoint_profile {
  double forward_ms, hamiltonian_ms, reverse_ms;
  int64_t reverse_launches;
  int32_t compiled, reserved;
} qsv_adjoint_profile;
int qsv_adjoint_profile_get(qsv_adjoint_profile* out);
/* Build check of the compiled reverse sweep without a GPU: the inverse
 * gates, last first, scheduled into two-state tiled passes with each
 * parametrised gate's generator measured before it; every pass generated
 * and compiled for gfx950 with hiprtc, not loaded.  *code_bytes,
 * *npasses and *nrows (each optional) receive the code size, the pass count
 * and the generator rows.  No counterpart in the reference; for build
 * systems and CI hosts without a device.                                  */
int qsv_adjoint_compile_check(int n, int dtype, const qsv_gate* gates, size_t ngates,
                              size_t* code_bytes, int64_t* npasses, int64_t* nrows);

/* ---- batched circuits (no counterpart in the reference) --------------------
 * Many independent circuits of uneven qubit counts, each started from
 * |0...0>, simulated together.  Their final states are packed in one device
 * buffer: circuit c occupies amplitudes [offsets[c], offsets[c+1]),
 * offsets[c+1] = offsets[c] + 2^n_c.  Gates of circuit c are
 * gates[gate_offsets[c] .. gate_offsets[c+1]).                            */
typedef struct qsv_batch qsv_batch;

typedef struct qsv_batch_opts {
  int32_t tile_bits;   /* LDS tile width for circuits above small_bits (0 = default) */
  int32_t small_bits;  /* circuits with n <= small_bits run whole in LDS (0 = largest) */
  int32_t fuse_k;      /* 0: gates as given; 1: merge runs of 1-qubit gates;
                          2: also fuse into blocks of at most 2 qubits           */
  int32_t phases;      /* 0: circuits above small_bits run tile passes that apply
                          one gate per LDS round trip; 1: they run register-
                          phase passes of 2^12-amplitude tiles (16 amplitudes
                          per thread, gates on the register bits, the unfused
                          gates of the circuit); circuits up to
                          max(small_bits, 11) qubits stay whole in LDS        */
} qsv_batch_opts;

typedef struct qsv_batch_info_t {
  int64_t  ncirc;
  int64_t  nsmall;     /* circuits run whole in LDS                         */
  int64_t  npasses;    /* tile passes over all circuits                     */
  int64_t  nops;       /* tile operations over all passes                   */
  int64_t  nlevels;    /* dependent launches of one run                     */
  uint64_t total_amplitudes;
  qsv_batch_opts opts; /* the options in effect                             */
  double   alg_bytes;  /* algorithmic HBM bytes of one run: every pass reads
                          and writes its circuit's state once (a circuit's
                          first pass only writes it)                       */
} qsv_batch_info_t;

typedef struct qsv_tune_result {
  qsv_batch_opts opts;
  double ms;            /* mean time of one run                             */
  double max_abs_diff;  /* vs the first candidate's states                  */
} qsv_tune_result;

int qsv_batch_create(int dtype, int ncirc, const int* nqubits, const qsv_gate* gates,
                     const size_t* gate_offsets, const qsv_batch_opts* opts, qsv_batch** out);
/* offsets may be NULL; otherwise receives ncirc+1 amplitude offsets.      */
int qsv_batch_info(const qsv_batch* batch, qsv_batch_info_t* info, uint64_t* offsets);
/* Simulate every circuit into `packed`.  use_graph != 0 replays a hipGraph
 * captured at the first call for this buffer (capture allocates nothing in
 * the graph: counters are reset by a kernel node).                        */
int qsv_batch_run(qsv_batch* batch, void* packed, int use_graph, void* stream);
/* Per-circuit completion time of the last run, seconds from its start.    */
int qsv_batch_latencies(qsv_batch* batch, double* seconds, void* stream);
int qsv_batch_destroy(qsv_batch* batch);
/* Time every candidate (tile_bits in 10..13 that fit, fuse_k in 0..2, and
 * register-phase passes with small_bits at its largest and at 11) on
 * `packed` (reps runs each, graph replay), check that all candidates give
 * the same states (max |difference| reported per candidate), and return the
 * fastest in *best.  `scratch` holds a copy of the first candidate's states
 * (same size as packed).  results may be NULL.                            */
int qsv_batch_autotune(int dtype, int ncirc, const int* nqubits, const qsv_gate* gates,
                       const size_t* gate_offsets, void* packed, void* scratch, int reps,
                       qsv_batch_opts* best, qsv_tune_result* results, int max_results,
                       int* nresults, void* stream);
/* max_i |a_i - b_i| over `count` amplitudes (FP64 result). */
int qsv_dev_max_abs_diff(const void* a, const void* b, uint64_t count, int dtype, double* out,
                         void* stream);

/* ---- host-vector entry points (the std::vector<T>& API) -------------------
 * `host` holds 2^n amplitudes; it is uploaded once, every gate runs on the
 * GPU, and the result is downloaded once.  Both transfers go through two
 * pinned 64 MiB staging buffers (host copy of one chunk overlapping the DMA
 * of the other; 55 GB/s for the two transfers of a 16 GiB state on the
 * MI355X host).  The device buffer and the staging buffers are cached by
 * the library between calls (qsv_release_cache frees them).  Calls
 * serialise on that cache.                                                 */
int qsv_host_apply_gates(void* host, int n, int dtype, const qsv_gate* gates,
                         size_t ngates);
/* Same, running the gate list as one compiled program with fusion width
 * fuse_k (0 = plain launches, 1..5 = fused, QSV_TILED = tiled passes,
 * QSV_AUTO = the engine's choice).                                        */
int qsv_host_simulate(void* host, int n, int dtype, const qsv_gate* gates, size_t ngates,
                      int fuse_k);
/* Device-resident variant: compile, run on `psi`, synchronise `stream`.   */
int qsv_dev_simulate(void* psi, int n, int dtype, const qsv_gate* gates, size_t ngates,
                     int fuse_k, void* stream);
int qsv_release_cache(void);

/* Counters of the compile caches since the library was loaded, so callers
 * (and tests) can see that a repeated simulate of one circuit structure
 * skipped the schedule search and the compile.  No counterpart in the
 * reference.
 *   schedule_hits/misses: tiled schedules whose tie rule was known for the
 *     circuit structure (one rule run) / searched over every rule;
 *   module_hits: compiled pass groups found in the process-wide cache;
 *   module_disk_hits: loaded from the on-disk cache
 *     ($QSV_JIT_CACHE_DIR, else ~/.cache/qclabpp_amd/jit; "" disables it);
 *   module_compiles: compiled with hiprtc.                                 */
typedef struct qsv_cache_stats {
  int64_t schedule_hits, schedule_misses;
  int64_t module_hits, module_disk_hits, module_compiles;
} qsv_cache_stats;
int qsv_cache_stats_get(qsv_cache_stats* out);

/* ---- unitaries (QCircuit::matrix, nrmF) -----------------------------------
 * Replace QCircuit<T>::matrix() (include/qclab/QCircuit.hpp:121-127, the
 * gates' apply(Side::Right, ...) on an identity, e.g.
 * src/qgates/QGate1.cpp:27-82) and nrmF (QCircuit.hpp:406-430).
 * `mat` receives the 2^k x 2^k unitary of the k-qubit gate list, column-
 * major (4^k amplitudes), in the MSB kron convention of matrix().  The
 * matrix is simulated as a 2k-qubit state (entry (r, c) at c*2^k + r: the
 * row index is qubits k..2k-1), so every compiled form applies (fuse_k as
 * for qsv_host_simulate).  Gate qubits must lie in [0, k), k in [1, 20].
 * QGate2 records (QSV_GATE_U2) act in the kron order of QGate2::matrix(),
 * as in the reference's matrix() (src/qgates/QGate2.cpp:32-109), not in
 * the transposed vector basis of its simulate() (src/qgates/apply.hpp:
 * 289-308); every other gate acts identically in both.                    */
int qsv_dev_circuit_matrix(void* mat, int k, int dtype, const qsv_gate* gates, size_t ngates,
                           int fuse_k, void* stream);
/* Same into host memory (through the cached device buffer, as
 * qsv_host_simulate).                                                     */
int qsv_host_circuit_matrix(void* mat, int k, int dtype, const qsv_gate* gates, size_t ngates,
                            int fuse_k);
/* Frobenius norm of a - b over `count` amplitudes with nrmF's max-abs
 * scaling: scale = max |a_i - b_i| (1 when all are equal), result
 * scale * sqrt(sum |a_i - b_i|^2 / scale^2), accumulated in FP64 in a fixed
 * order (QCircuit.hpp:406-430).                                            */
int qsv_dev_nrmf(const void* a, const void* b, uint64_t count, int dtype, double* out,
                 void* stream);
/* nrmF(circuit1, circuit2) for two k-qubit gate lists: both unitaries are
 * built on the GPU (two 4^k-amplitude device buffers, freed on return) and
 * reduced there; only the FP64 result comes back.                         */
int qsv_circuit_nrmf(int k, int dtype, const qsv_gate* gates1, size_t ngates1,
                     const qsv_gate* gates2, size_t ngates2, int fuse_k, double* out);

#ifdef __cplusplus
}
#endif

#endif /* QSV_H */
