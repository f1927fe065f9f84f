/*
 * qsim.h -- C ABI of the MI355X (gfx950) gate-application library libqsim_hip.so.
 *
 * This library replaces the device-runtime boundary of nasir26/fpga-mps-simulator:
 * the pyxrt buffer objects and the two HLS kernels they launch.
 *
 *   statevector : quantum_fpga_kernel_complex128(double* state_real, double* state_imag,
 *                   int* gate_sequence, int num_gates, int num_qubits)
 *                   (opt_kernel.cpp:200-206), launched by QuantumFPGACircuit._launch_kernel
 *                   (fpga_simulator_v07.py:486-488)
 *   MPS         : fpga_mps_simulator(fp64* bank0..bank7, int* gate_seq, int num_gates,
 *                   int nq, int chi_max) (fpga_mps_simulator.cpp:392-404), launched from
 *                   _run_circuit_fpga (fpga_mps_simulator.py:512, 594), followed there by a
 *                   host scipy SVD (fpga_mps_simulator.py:597-634)
 *
 * Conventions
 *  - Every function returns a status: QS_OK (0) or a negative QS_E_* code.  The message of
 *    the last failure on the calling thread is qs_last_error().  Nothing is clamped
 *    silently: an argument beyond a documented limit is an error (the HLS kernels clamp
 *    chi to 64 and gates to 512 without telling the host).
 *  - Device pointers are plain device addresses (e.g. torch.Tensor.data_ptr() on ROCm).
 *  - `stream` is a hipStream_t passed as void*; NULL is the null stream.  Calls are
 *    asynchronous on that stream unless documented as synchronous.
 *  - Calls may run concurrently from several threads but must not be given overlapping
 *    buffers at the same time.
 *  - Statevectors are complex128, interleaved (re, im), index bit q = qubit q
 *    (little-endian, as the reference's collected statevector).
 *  - MPS site tensors are complex128 (chi_l, 2, chi_r) in C order, as mps_core.MPSState.
 */
#ifndef QSIM_H
#define QSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QS_ABI_VERSION 2

#define QS_OK         0
#define QS_E_ARG     -1   /* bad argument (null pointer, qubit out of range, ...) */
#define QS_E_CHI     -2   /* bond dimension beyond the supported maximum */
#define QS_E_HIP     -3   /* HIP runtime error */
#define QS_E_NOCONV  -4   /* Jacobi SVD did not converge within the sweep limit */
#define QS_E_NODEV   -5   /* no usable gfx950 device */
#define QS_E_RANGE   -6   /* size beyond a documented limit (qubits, gates, workspace) */

/* ---------------------------------------------------------------- lifecycle */

/* ABI version of the loaded library (QS_ABI_VERSION). */
int qs_abi_version(void);

/* Select and initialise `device` (hipSetDevice + capability check).  Replaces
 * xrt.device(i)/load_xclbin/xrt.kernel (fpga_simulator_v07.py:399-404,
 * fpga_mps_simulator.py:675-681).  Fails with QS_E_NODEV when the device is not gfx950. */
int qs_init(int device);

/* Thread-local message of the last failure ("" if none). */
const char* qs_last_error(void);

/* Maximum qubits of the statevector path (state must fit one allocation). */
int qs_sv_max_qubits(void);

/* ------------------------------------------------------------- statevector */

/* Gate kinds of qs_gate.  Matrices are complex, row-major, (re, im) pairs. */
#define QS_GATE_U1    0   /* 2x2 on q[0]                                              */
#define QS_GATE_U2    1   /* 4x4 on (q[0], q[1]); row index = 2*bit(q[0]) + bit(q[1])   */
#define QS_GATE_DIAG1 2   /* diagonal (d0, d1) on q[0]                                 */
#define QS_GATE_DIAG2 3   /* diagonal (d00, d01, d10, d11) on (q[0], q[1]), same order */

/* One gate.  `ctrl_mask` lists control qubits: the gate acts only on the amplitudes
 * whose index has all those bits set (controls must not overlap the targets). */
typedef struct qs_gate {
    int32_t  kind;
    int32_t  q[2];
    int32_t  reserved;
    uint64_t ctrl_mask;
    double   m[32];
} qs_gate;

/* |0...0> into a 2^nq complex128 state.  Replaces _distribute_statevector
 * (fpga_simulator_v07.py:430-436). */
int qs_sv_set_zero(void* state, int nq, void* stream);

/* Apply `ng` gates in order.  Replaces the per-chunk upload + launch of
 * quantum_fpga_kernel_complex128 (opt_kernel.cpp:200-353,
 * fpga_simulator_v07.py:475-500).  The host plans fused passes: each pass reads
 * and writes the state once and applies every gate whose targets fall in its
 * tile of qubits; diagonal gates and controls may lie anywhere.  No gate-count limit. */
int qs_sv_apply(void* state, int nq, const qs_gate* gates, int ng, void* stream);

/* Apply gates given in the reference's 8-word instruction format
 * [op, q0, q1, q2, f32 p0, f32 p1, f32 p2, 0] with the SV opcode table
 * (id0 h1 x2 y3 z4 s5 t6 sdg7 tdg8 sx9 sxdg10 rx11 ry12 rz13 p14 u2 15 u3 16
 * cx17 cz18; fpga_simulator_v07.py:69-78, 456-473; opt_kernel.cpp:39-58, 240-353).
 * Parameters are decoded from float32 exactly as the HLS kernel does.  Unlike the
 * kernel, num_gates is not clamped to 512: any count is applied. */
int qs_sv_apply_words(void* state, int nq, const int32_t* gate_sequence, int num_gates,
                      void* stream);

/* Plan once, launch many (no counterpart in the reference, whose driver re-packs and
 * re-uploads the gate words for every run, fpga_simulator_v07.py:456-500).
 * qs_sv_plan_create builds the fused-pass plan of `gates` for an nq-qubit state (the
 * plan qs_sv_apply would build) and uploads it, synchronously, into device memory of the
 * current device owned by the plan.  qs_sv_plan_launch applies the planned gates to
 * `state` on `stream`: it plans nothing, copies nothing from the host and allocates
 * nothing, so it can be recorded into a HIP graph; its result equals qs_sv_apply's bit for
 * bit.  The plan's block-factor table is written once by qs_sv_plan_create and only read by
 * launches, so one plan may be launched on several streams at once (on different
 * states).  qs_sv_plan_info reports the qubit count, the
 * passes and the device bytes the plan holds.  qs_sv_plan_destroy frees it (hipFree: it
 * waits for the device, so launches in flight finish first); NULL is accepted. */
typedef struct qs_sv_plan qs_sv_plan;
int qs_sv_plan_create(int nq, const qs_gate* gates, int ng, qs_sv_plan** plan_out);
int qs_sv_plan_launch(const qs_sv_plan* plan, void* state, void* stream);
int qs_sv_plan_info(const qs_sv_plan* plan, int* nq, int* npasses, size_t* device_bytes);
int qs_sv_plan_destroy(qs_sv_plan* plan);

/* Per-plan pass kernels (no counterpart in the reference).  qs_sv_plan_specialize
 * compiles, with hipRTC for gfx950, one kernel per distinct pass of the plan: the pass's
 * ops as straight-line code whose op kinds, register bits, controls and aux words are
 * compile-time constants (the op-stream interpreter branches on them per op).  The op
 * data stays in the plan's tables, so later launches of the plan run these kernels and
 * give the same state bit for bit.  Kernels live in a process-wide cache keyed by device
 * and kernel text: passes with the same op sequence, and plans of the same circuit
 * structure with other parameters, reuse them.  *kernels_compiled (may be NULL) receives
 * the number compiled by this call.  On failure (QS_E_HIP, the compiler log in
 * qs_last_error) the plan keeps launching the interpreter.  The device must be the
 * plan's.  Not thread-safe with launches of the same plan.
 * qs_sv_jit_check builds the plan of `gates` and compiles its pass kernels without
 * loading them (host only, no device needed): *nkernels distinct kernels, *code_bytes
 * bytes of code objects (either may be NULL). */
int qs_sv_plan_specialize(qs_sv_plan* plan, int* kernels_compiled);
int qs_sv_jit_check(int nq, const qs_gate* gates, int ng, int* nkernels, size_t* code_bytes);
/* Source text of pass `pass`'s kernel (host only; buffer contract as qs_sv_plan_json). */
int qs_sv_jit_source(int nq, const qs_gate* gates, int ng, int pass, char* buf, size_t cap,
                     size_t* needed);

/* Number of fused passes qs_sv_apply would launch for this gate list (host only). */
int qs_sv_plan_passes(int nq, const qs_gate* gates, int ng, int* npasses);

/* Diagnostics (host only, no device needed): the fused-pass plan as JSON
 * {"passes":[{"K","R","load_regs":[...],"tq":[...],"runs":[[src,dst,mask],...],
 *              "stages":[{"regs":[...],"gates":[...]}]}]}: the tile qubits, the map from
 * tile index bits to the other qubits as runs, and per stage the register bits and ops.
 * Writes at most `cap` bytes into `buf` (NUL-terminated when it fits) and the size
 * required into *needed.  No counterpart in the reference; it lets the planner be
 * tested without a GPU. */
int qs_sv_plan_json(int nq, const qs_gate* gates, int ng, char* buf, size_t cap,
                    size_t* needed);

/* Instrumentation, no counterpart in the reference.  While enabled (per host
 * thread), every fused-pass kernel that qs_sv_apply / qs_sv_apply_words launch is
 * bracketed by HIP events recorded on the caller's stream.  qs_sv_pass_times waits
 * for the recorded events, writes up to `cap` kernel durations in milliseconds (launch
 * order) into ms_out and the number recorded into *count, then clears the record. */
int qs_sv_time_passes(int enable);
int qs_sv_pass_times(double* ms_out, int cap, int* count);

/* sum |a_i|^2 into *out_host.  Synchronous. */
int qs_sv_norm_sq(const void* state, int nq, double* out_host, void* stream);

/* state *= (re + i im). */
int qs_sv_scale(void* state, int nq, double re, double im, void* stream);

/* Reset qubit q: zero the |1> half, renormalise, |0..0> if nothing is left
 * (fpga_simulator_v07.py:502-513, on the device).  Synchronous (reads the norm). */
int qs_sv_reset_qubit(void* state, int nq, int q, void* stream);

/* probs[i] = |a_i|^2 (float64, device buffer of 2^nq). */
int qs_sv_probabilities(const void* state, int nq, double* probs, void* stream);

/* Measurement shots, replacing the host-side
 *   probs = |sv|^2 / sum |sv|^2;  indices = np.random.choice(2^nq, shots, p=probs)
 * of run_circuit (fpga_simulator_v07.py:592-609).  numpy's choice draws one uniform per
 * shot (random_sample) and returns, for each, the first index whose cumulative
 * probability exceeds it; here the caller passes those uniforms (device buffer of
 * `shots` float64 in [0, 1)) and receives the indices (device buffer of `shots`
 * int64).  Indices differ from numpy's only when a uniform lies within rounding
 * distance of a cumulative-probability boundary.  `workspace` is a device buffer of at
 * least qs_sv_sample_workspace_bytes(nq) bytes.  Fails with QS_E_ARG when sum |a|^2 is
 * not positive (numpy raises "probabilities do not sum to 1").  Synchronises `stream`
 * once (normalisation check); the indices are written on `stream`. */
size_t qs_sv_sample_workspace_bytes(int nq);
int qs_sv_sample(const void* state, int nq, const double* uniforms, int shots,
                 int64_t* indices, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------- statevector: gradients */
/* No counterpart in the reference (it has no differentiation); these serve the adjoint
 * gradient of <psi|H|psi> for circuits of rotations exp(-i theta P / 2). */

/* out[2k], out[2k+1] = Re, Im of <lam| P_k |psi>, P_k = X (0), Y (1) or Z (2) on qubits[k].
 * Synchronous. */
int qs_sv_pauli_inner(const void* lam, const void* psi, int nq, const int32_t* kinds,
                      const int32_t* qubits, int count, double* out_host, void* stream);

/* out_host[0], out_host[1] = Re, Im of <a|b>.  Synchronous. */
int qs_sv_inner(const void* a, const void* b, int nq, double* out_host, void* stream);

/* out = H psi, H = -J sum_i Z_i Z_{i+1} - h sum_i X_i (open chain); out != psi. */
int qs_sv_apply_tfim(const void* psi, void* out, int nq, double J, double h, void* stream);

/* --------------------------------------------------------------------- MPS */

#define QS_PREC_FP64       0   /* contraction FP64, SVD FP64                        */
#define QS_PREC_FP32_MFMA  1   /* contraction complex FP32 on MFMA, SVD FP64        */

/* Largest chi_max the two-site update supports (128).  Up to 64 the SVD keeps theta in
 * registers; the 65..128 class keeps the columns in a working copy in the workspace.  Both
 * classes take the requested contraction precision (the 65..128 class contracts in
 * 128 x 128 blocks of theta, four workgroups per update). */
int qs_mps_max_chi(void);

/* One two-site update of a chain: sites i and i+1 of the chain whose slots start at
 * `sites` (slot k at sites + k*site_stride complex elements, each slot holding a
 * (chi_l, 2, chi_r) tensor in C order) and whose bond dimensions are `bond`
 * (device int32[n+1], bond[k] = left dimension of site k, bond[0] = bond[n] = 1).
 * The update contracts the two sites with the gate, SVDs the (2chi_l x 2chi_r) matrix,
 * keeps clip(#{S > cutoff*S[0]}, 1, chi_max) values, renormalises them, writes
 * U*sqrt(S) into site i and sqrt(S)*Vh into site i+1, stores the kept rank in
 * bond[i+1] and the norm_log increment in *dlog (mps_core.py:277-300, 326-350). */
typedef struct qs_mps_task {
    double*  sites;        /* device, complex128 slots                         */
    int32_t* bond;         /* device, int32[n+1]                               */
    double*  dlog;         /* device, receives the norm_log increment           */
    int64_t  site_stride;  /* complex elements per slot, >= chi_max*2*chi_max  */
    int32_t  site;         /* left site i, 0 <= i < n_sites - 1                */
    int32_t  chi_max;
    double   cutoff;
    int32_t  n_sites;      /* n, the chain's site count (ABI version 2)        */
    int32_t  reserved;     /* 0                                                */
    double   gate[32];     /* 4x4, row index = 2*sigma_i + sigma_{i+1}         */
} qs_mps_task;

/* Workspace bytes qs_mps_update_batch needs for `ntasks` tasks of at most chi_max
 * (per task 4 chi_max^2 complex elements, 8 chi_max^2 when chi_max > 64). */
size_t qs_mps_workspace_bytes(int ntasks, int chi_max);

/* Run `ntasks` independent two-site updates in one launch sequence (no two tasks may
 * touch the same site).  Replaces the per-gate kernel launch + theta read-back + host
 * SVD + site write-back of _run_circuit_fpga (fpga_mps_simulator.py:566-634) and the
 * device contraction contract_and_apply_2q (fpga_mps_simulator.cpp:287-371).
 * `workspace` is a device buffer of at least qs_mps_workspace_bytes(ntasks, max chi_max).
 * When `status_dev` is non-null it receives per-task statuses (device int32[ntasks]):
 * the number of Jacobi sweeps used (>= 1) on success, QS_E_NOCONV or QS_E_CHI. */
int qs_mps_update_batch(const qs_mps_task* tasks, int ntasks, int precision,
                        void* workspace, size_t ws_bytes, int32_t* status_dev,
                        void* stream);

/* Allocation-free form of qs_mps_update_batch.  It has no counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The workspace is a static
 * device array of QS_MPS_STATIC_SCRATCH_BYTES (32 MiB) that the library holds on every
 * device it runs on until the process ends; the call allocates nothing (no hipMalloc,
 * hipFree or allocating library routine) and refuses, with QS_E_RANGE, a batch whose
 * qs_mps_workspace_bytes(ntasks, max chi_max) exceeds that size (e.g. more than 127 tasks
 * at chi_max 64).  It runs on the current device; tasks, statuses and results are as for
 * qs_mps_update_batch, bit for bit.  Every call on one device shares that device's
 * array: calls that overlap in time on one device (two host threads, or two streams
 * whose work overlaps) race and must be serialised by the caller. */
#define QS_MPS_STATIC_SCRATCH_BYTES (32u << 20)
int qs_mps_update_batch_static(const qs_mps_task* tasks, int ntasks, int precision,
                               int32_t* status_dev, void* stream);

/* Graph-capturable form of qs_mps_update_batch, in two steps (no counterpart in the
 * reference, whose host loop launches one update at a time and waits for it,
 * fpga_mps_simulator.py:566-634).
 * qs_mps_batch_plan validates `tasks` on the host exactly as qs_mps_update_batch does
 * and writes the launch table (tasks grouped by SVD size class) into `table`, host
 * memory of at least qs_mps_batch_table_bytes(ntasks) bytes; counts[5] receives the
 * number of tasks of class 8, 16, 32, 64 and 128, *chi_max_out the largest chi_max.
 * The caller copies the table to device memory.
 * qs_mps_batch_launch runs the batch from that device copy: it reads no host memory
 * beyond its arguments, copies nothing and allocates nothing, so it can be recorded
 * into a HIP graph and replayed while the table, chains, dlog slots, workspace and
 * status buffer stay allocated.  `chi_max` sizes the per-task workspace slots
 * (qs_mps_workspace_bytes(ntasks, chi_max)); a task whose chi_max needs a larger slot
 * gets status QS_E_CHI and writes nothing.  Statuses as for qs_mps_update_batch. */
size_t qs_mps_batch_table_bytes(int ntasks);
int qs_mps_batch_plan(const qs_mps_task* tasks, int ntasks, void* table, size_t table_bytes,
                      int32_t counts[5], int* chi_max_out);
int qs_mps_batch_launch(const void* table_dev, int ntasks, const int32_t counts[5],
                        int chi_max, int precision, void* workspace, size_t ws_bytes,
                        int32_t* status_dev, void* stream);

/* Reverse-mode step of the two-site update (no counterpart in the reference, which has no
 * differentiation): for each task, the cotangents of the sites before one update, given
 * the sites before (a: (cl, 2, cm), b: (cm, 2, cr)) and after (a2: (cl, 2, k), b2: (k, 2,
 * cr)) it and the cotangents ga2, gb2 of the sites after it (device complex128, C order;
 * cotangent convention dL = Re sum conj(g) dz).  Writes ga (cl, 2, cm) and gb (cm, 2, cr).
 * The SVD backward uses F_ij = (s_j^2 - s_i^2) / ((s_j^2 - s_i^2)^2 + eps s_max^4) so that
 * near-degenerate singular values stay finite; S is read back as Re diag(U^H theta V) with U,
 * V from a2, b2 (whose column / row norms are sqrt(S_t)).  One workgroup per task; tasks
 * must not write overlapping outputs.  `work`: qs_mps_backward_work_bytes(ntasks, chi_max)
 * device bytes, chi_max bounding every dimension; status_dev (device int32[ntasks], may be
 * null) receives QS_OK or QS_E_ARG per task. */
typedef struct qs_mps_bwd_task {
    const double* a;
    const double* b;
    const double* a2;
    const double* b2;
    const double* ga2;
    const double* gb2;
    double*       ga;
    double*       gb;
    int32_t       cl, cm, cr, k;
    double        eps;
    double        gate[32];   /* the update's 4x4 gate, row index 2 sigma_i + sigma_{i+1} */
} qs_mps_bwd_task;
size_t qs_mps_backward_work_bytes(int ntasks, int chi_max);
int qs_mps_backward_batch(const qs_mps_bwd_task* tasks, int ntasks, int chi_max, void* work,
                          size_t work_bytes, int32_t* status_dev, void* stream);

/* Kernel timing of the MPS batches (instrumentation, no counterpart in the reference).
 * While enabled, every contraction and SVD launch of qs_mps_update_batch /
 * qs_mps_batch_launch is bracketed by HIP events on its stream.  qs_mps_kernel_times
 * waits for them and returns, in launch order, up to `cap` entries: kind (0 = complex
 * FP32 contraction on MFMA, 1 = FP64 contraction, 2 = SVD), the launch's task count and
 * its duration in ms; *count receives the number recorded.  Reading clears the record. */
int qs_mps_time_kernels(int enable);
int qs_mps_kernel_times(int32_t* kind, int32_t* tasks, double* ms, int cap, int* count);

/* Dense statevector of a device chain (mps_core.py:433-461): out[idx] with bit q of idx
 * the value of qubit q (site 0 the least significant), times exp(norm_log).  `out` and
 * `work` are device buffers of at least 2^n complex128 (buffer_bytes each); every bond
 * must satisfy bond[i+1] <= 2^(n-i-1), as SVD truncation leaves it.  One launch per site;
 * synchronises `stream` once to read the bond dimensions. */
int qs_mps_statevector(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                       int chi_max, double norm_log, double* out, double* work,
                       size_t buffer_bytes, void* stream);

/* Shot sampling of a chain on the device (mps_core.py:354-420, called by the driver at
 * fpga_mps_simulator.py:639-643, where it runs on the host).
 * qs_mps_right_envs writes the right environments R_0..R_n of the chain (sites, bond,
 * site_stride as for qs_mps_task; chi_max <= 128) into `envs`, device memory of at least
 * qs_mps_env_bytes(n, chi_max) bytes: R_n = [[1]], R_i = sum_s A_i[:, s, :] R_{i+1}
 * A_i[:, s, :]^H divided by its trace (when above 1e-300), each (bond[i] x bond[i])
 * complex128 with row stride chi_max, at envs + i * chi_max^2 complex elements.
 * qs_mps_sample then draws `shots` bitstrings by sequential conditional sampling: for
 * shot t and site i it reads the uniform uniforms[t * n + i] (device, in [0, 1)) and
 * takes outcome 1 exactly when u >= p_0 / (p_0 + p_1), as numpy's
 * Generator.choice(2, p=[p_0, p_1]) does with the same uniform; bits[t * n + i]
 * (device uint8) receives the outcome.  Drawing the uniforms in that order from the
 * reference's generator reproduces its samples. */
size_t qs_mps_env_bytes(int n, int chi_max);
int qs_mps_right_envs(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                      int chi_max, double* envs, size_t env_bytes, void* stream);
int qs_mps_sample(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                  int chi_max, const double* envs, const double* uniforms, int shots,
                  uint8_t* bits, void* stream);

/* Transfer-matrix read-outs of device chains (no counterpart in the reference, whose
 * observables need the dense statevector, mps_core.py:433-461, or host copies).  Chains are
 * given as for qs_mps_right_envs (sites, bond, site_stride, n, chi_max <= 128).
 * qs_mps_left_envs writes L_0 = [[1]] and L_{i+1}[a', b'] = sum_s sum_{a,b}
 * conj(A_i[a, s, a']) L_i[a, b] A_i[b, s, b'], each divided by its trace, (bond[i] x bond[i])
 * at envs + i * chi_max^2 complex elements (row stride chi_max; qs_mps_env_bytes bytes).
 * qs_mps_expect writes, for span 1, out[i] = <psi|O_i|psi> / <psi|psi> for i < n with O =
 * op_a, and for span 2 out[i] = <psi|O_i O'_{i+1}|psi> / <psi|psi> for i < n - 1 with
 * O' = op_b (2x2 complex row-major, host memory; out is device float64), from the left
 * environments and the right environments of qs_mps_right_envs; `work` is device memory of
 * qs_mps_expect_work_bytes(chi_max).  qs_mps_overlap computes <bra|ket> of two chains of
 * n sites (norm_log factors excluded) into out_host = {log |<bra|ket>|, arg <bra|ket>};
 * `work` is device memory of qs_mps_overlap_work_bytes(n, chi_max); synchronous.
 * Every transfer is renormalised on the way (trace or entry sum), so chains of thousands
 * of sites neither overflow nor underflow. */
int qs_mps_left_envs(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                     int chi_max, double* envs, size_t env_bytes, double* log_traces,
                     void* stream);
size_t qs_mps_expect_work_bytes(int chi_max);
int qs_mps_expect(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                  int chi_max, const double* lenvs, const double* renvs, const double op_a[8],
                  const double op_b[8], int span, double* out, void* work, size_t work_bytes,
                  void* stream);
size_t qs_mps_overlap_work_bytes(int n, int chi_max);
int qs_mps_overlap(const double* bra_sites, const int32_t* bra_bond, int64_t bra_stride,
                   const double* ket_sites, const int32_t* ket_bond, int64_t ket_stride, int n,
                   int chi_max, void* work, size_t work_bytes, double out_host[2],
                   void* stream);

/* Entanglement entropy on the device, by the reference's definition (mps_core.py:465-483:
 * sites 0..b contracted into one matrix M, its singular values S > 1e-15, p = S^2 / sum
 * S^2, -sum p log p), from the left environments and log traces of qs_mps_left_envs:
 * S^2 are the eigenvalues of M^H M, which is L_{b+1} times the product of the traces.
 * One workgroup per requested bond (bonds: host int32[nbonds], each in [0, n-2]) runs a
 * cyclic Jacobi eigenvalue iteration on a copy of L_{b+1} in `work` (device memory of
 * qs_mps_entropy_work_bytes(nbonds, chi_max)); entropy: device float64[nbonds].  The
 * eigenvalues of a Gram matrix carry an absolute error of about 1e-16 of the largest, so
 * entropies match the reference's SVD of M to about 1e-13. */
size_t qs_mps_entropy_work_bytes(int nbonds, int chi_max);
int qs_mps_entanglement_entropy(const int32_t* bond, int n, int chi_max, const double* lenvs,
                                const double* log_traces, const int32_t* bonds, int nbonds,
                                double* entropy, void* work, size_t work_bytes, void* stream);

/* SVD tile shape, the autotuner's knob (no counterpart in the reference, whose SVD is
 * LAPACK gesdd on the host).  The register SVD of size class 8, 16, 32 or 64 runs one
 * team of `lanes` lanes per column pair (lanes * class threads per task, 2 class / lanes
 * rows per lane).  qs_mps_svd_lane_choices lists the widths built for a class (up to
 * `cap` into out[], the number in *count); qs_mps_svd_lanes returns the width in use
 * (default 16), qs_mps_set_svd_lanes selects one (QS_E_ARG if not built).  Every width
 * computes the same decomposition; the summation order of the column inner products,
 * and so the last bits of the result, depends on it.  The setting is process-wide. */
int qs_mps_svd_lane_choices(int chi_class, int* out, int cap, int* count);
int qs_mps_svd_lanes(int chi_class);
int qs_mps_set_svd_lanes(int chi_class, int lanes);

/* Many one-site gates in one launch sequence (the driver's batched flush of up to 508
 * one-qubit instructions per kernel call, fpga_mps_simulator.py:498-526, 562): gate j is the
 * 2x2 complex row-major matrix mats[8j .. 8j+7] (host memory) on site site_list[j] (host)
 * of a chain of n sites.  Sites must be distinct (compose the gates of one site first);
 * gates travel in the launch arguments, 32 per launch. */
int qs_mps_apply_1q_batch(double* sites, const int32_t* bond, int64_t site_stride, int n,
                          const int32_t* site_list, const double* mats, int count,
                          void* stream);

/* Reset of one site on the device (fpga_mps_simulator.py:530-552): the sigma = 1 half is
 * zeroed and the site divided by its norm, or set to the |0> tensor (t[0,0,0] = 1) when
 * that norm is at most 1e-15. */
int qs_mps_reset_site(double* sites, const int32_t* bond, int64_t site_stride, int n, int site,
                      void* stream);

/* A'[a, s', b] = sum_s m[s', s] A[a, s, b] on one site, dims read from bond[site],
 * bond[site+1] on the device (fpga_mps_simulator.cpp:232-276, mps_core.py:267-273).
 * m: 2x2 complex row-major, host memory (copied into the launch). */
int qs_mps_apply_1q(double* sites, const int32_t* bond, int64_t site_stride, int site,
                    const double m[8], void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QSIM_H */
