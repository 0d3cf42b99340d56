/*
 * libqulacs_mi355 -- C ABI of the MI355X (gfx950) state-vector and
 * density-matrix kernels.
 *
 * The state-vector entry points keep the names, argument order and meaning of
 * qulacs's GPU seam, the `*_host` family that cppsim calls whenever
 * `state->get_device_name() == "gpu"` (reference: src/gpusim/update_ops_cuda.h,
 * src/gpusim/stat_ops.h, src/gpusim/memory_ops.h, src/gpusim/util_func.h).  Each
 * declaration below cites the reference declaration it replaces.
 *
 * Types (reference src/gpusim/util_type.h:9-17):
 *   ITYPE      = unsigned long long (dimension / basis index)
 *   UINT       = unsigned int
 *   qm_complex = two doubles (re, im); layout-compatible with std::complex<double>
 *                and returned in the same registers on x86-64 SysV.
 * `state`  : device pointer to dim complex128 values (qubit q = index bit q).
 * `stream` : pointer to a hipStream_t (as returned by allocate_cuda_stream_host),
 *            or NULL for the null stream.
 * `device_number` : HIP device the state lives on; selected if not current.
 * Matrices are host pointers, row-major; for a k-qubit gate bit i of the
 * row/column index belongs to target_qubit_index_list[i] (reference
 * src/csim/utility.cpp:77-93).  They may be reused as soon as the call returns.
 *
 * Behaviour that differs from the reference, on purpose:
 *   * errors do not exit the process (reference: src/gpusim/util.cuh:240-254);
 *     an invalid argument or a HIP failure is recorded and the call returns
 *     without touching the state.  Read it with qm_get_last_error().
 *   * update calls synchronise their stream before returning, as the reference
 *     does (src/gpusim/update_ops_single.cu:130), unless qm_set_synchronous(0)
 *     makes them stream-ordered only.  Reductions always return a host value
 *     and therefore always wait for their stream.
 *
 * The density-matrix entry points (dm_*_host) have no GPU counterpart in the
 * reference, whose density matrix only runs on the CPU (src/csim/update_ops_dm.cpp);
 * they mirror those csim functions' arguments with the device/stream tail added.
 */
#ifndef QULACS_MI355_H_
#define QULACS_MI355_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned long long ITYPE;
typedef unsigned int UINT;
typedef struct qm_complex {
    double re;
    double im;
} qm_complex;

#define QM_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ errors */
#define QM_OK 0
#define QM_ERR_INVALID_ARGUMENT 1
#define QM_ERR_HIP 2
#define QM_ERR_OUT_OF_MEMORY 3
#define QM_ERR_UNSUPPORTED 4

/* Last error recorded on the calling thread (0 if none); copies its message
 * into buf (may be NULL).  No reference counterpart: the reference exits. */
QM_EXPORT int qm_get_last_error(char* buf, size_t buf_len);
QM_EXPORT void qm_clear_last_error(void);
/* 1 (default): update calls wait for their stream, as the reference does.
 * 0: update calls only enqueue work.  No reference counterpart. */
QM_EXPORT void qm_set_synchronous(int on);
QM_EXPORT int qm_get_synchronous(void);
/* Waits for the given stream (NULL = null stream) on device_number. */
QM_EXPORT void qm_synchronize(void* stream, UINT device_number);
/* Library version string. */
QM_EXPORT const char* qm_version(void);

/* ------------------------------------------------------------------ devices / memory
 * reference src/gpusim/util_func.h:9-25, src/gpusim/memory_ops.h:7-20 */
QM_EXPORT int get_num_device(void);                                      /* util_func.h:9  */
QM_EXPORT void set_device(unsigned int device_num);                      /* util_func.h:10 */
QM_EXPORT int get_current_device(void);                                  /* util_func.h:11 */
QM_EXPORT void* allocate_cuda_stream_host(unsigned int max_cuda_stream,
    unsigned int device_number);                                         /* memory_ops.h:7  */
QM_EXPORT void release_cuda_stream_host(void* cuda_stream,
    unsigned int max_cuda_stream, unsigned int device_number);           /* memory_ops.h:9  */
QM_EXPORT void* allocate_quantum_state_host(ITYPE dim,
    unsigned int device_number);                                         /* memory_ops.h:11 */
QM_EXPORT void initialize_quantum_state_host(void* state, ITYPE dim,
    void* stream, unsigned int device_number);                           /* memory_ops.h:13 */
QM_EXPORT void release_quantum_state_host(void* state,
    unsigned int device_number);                                         /* memory_ops.h:15 */
/* Haar-random state from a counter-based generator of this library; the
 * bitstream differs from the reference's hipRAND XORWOW (memory_ops.cu:86-161). */
QM_EXPORT void initialize_Haar_random_state_host(void* state, ITYPE dim,
    void* stream, unsigned int device_number);                           /* memory_ops.h:17 */
QM_EXPORT void initialize_Haar_random_state_with_seed_host(void* state,
    ITYPE dim, UINT seed, void* stream, unsigned int device_number);     /* memory_ops.h:19 */
QM_EXPORT void set_computational_basis_host(ITYPE comp_basis, void* state,
    ITYPE dim, void* stream, unsigned int device_number);                /* util_func.h:12 */
QM_EXPORT void copy_quantum_state_from_device_to_device(void* state_gpu_copy,
    const void* state_gpu, ITYPE dim, void* stream,
    unsigned int device_number);                                         /* util_func.h:14 */
QM_EXPORT void copy_quantum_state_from_host_to_device(void* state_gpu_copy,
    const void* state, ITYPE dim, void* stream,
    unsigned int device_number);                                         /* util_func.h:16 */
QM_EXPORT void copy_quantum_state_from_cppstate_host(void* state_gpu_copy,
    const qm_complex* cppstate, ITYPE dim, void* stream,
    UINT device_number);                                                 /* util_func.h:18 */
QM_EXPORT void copy_quantum_state_from_device_to_host(void* state_cpu_copy,
    const void* state_gpu_original, ITYPE dim, void* stream,
    unsigned int device_number);                                         /* util_func.h:20 */
QM_EXPORT void get_quantum_state_host(void* state_gpu, void* psi_cpu_copy,
    ITYPE dim, void* stream, unsigned int device_number);                /* util_func.h:23 */
QM_EXPORT void print_quantum_state_host(void* state, ITYPE dim,
    unsigned int device_number);                                         /* util_func.h:25 */

/* ------------------------------------------------------------------ named gates
 * reference src/gpusim/update_ops_cuda.h:10-55 */
QM_EXPORT void H_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);   /* :10 */
QM_EXPORT void X_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);   /* :12 */
QM_EXPORT void Y_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);   /* :14 */
QM_EXPORT void Z_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);   /* :16 */
QM_EXPORT void CZ_gate_host(unsigned int control_qubit_index, unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);     /* :18 */
QM_EXPORT void CNOT_gate_host(unsigned int control_qubit_index, unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);   /* :21 */
QM_EXPORT void SWAP_gate_host(unsigned int target_qubit_index0, unsigned int target_qubit_index1, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :24 */
QM_EXPORT void RX_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :27 */
QM_EXPORT void RY_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :29 */
QM_EXPORT void RZ_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :31 */
QM_EXPORT void S_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);          /* :33 */
QM_EXPORT void Sdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);       /* :35 */
QM_EXPORT void T_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);          /* :37 */
QM_EXPORT void Tdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);       /* :39 */
QM_EXPORT void sqrtX_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);      /* :41 */
QM_EXPORT void sqrtXdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);   /* :43 */
QM_EXPORT void sqrtY_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);      /* :45 */
QM_EXPORT void sqrtYdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);   /* :47 */
QM_EXPORT void P0_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);         /* :49 */
QM_EXPORT void P1_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);         /* :51 */
QM_EXPORT void normalize_host(double squared_norm, void* state, ITYPE dim, void* stream, unsigned int device_number);          /* :53 */

/* ------------------------------------------------------------------ single-qubit matrices
 * reference src/gpusim/update_ops_cuda.h:57-80 */
QM_EXPORT void single_qubit_Pauli_gate_host(UINT target_qubit_index, UINT Pauli_operator_type, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :57 */
QM_EXPORT void single_qubit_Pauli_rotation_gate_host(unsigned int target_qubit_index, unsigned int op_idx, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :60 */
QM_EXPORT void single_qubit_dense_matrix_gate_host(unsigned int target_qubit_index, const qm_complex matrix[4], void* state, ITYPE dim, void* stream, unsigned int device_number); /* :63 */
QM_EXPORT void single_qubit_diagonal_matrix_gate_host(unsigned int target_qubit_index, const qm_complex diagonal_matrix[2], void* state, ITYPE dim, void* stream, unsigned int device_number); /* :66 */
QM_EXPORT void single_qubit_control_single_qubit_dense_matrix_gate_host(unsigned int control_qubit_index, unsigned int control_value, unsigned int target_qubit_index, const qm_complex matrix[4], void* state, ITYPE dim, void* stream, unsigned int device_number); /* :69 */
QM_EXPORT void single_qubit_phase_gate_host(unsigned int target_qubit_index, qm_complex phase, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :73 */
QM_EXPORT void multi_qubit_control_single_qubit_dense_matrix_gate_host(const UINT* control_qubit_index_list, const UINT* control_value_list, UINT control_qubit_index_count, UINT target_qubit_index, const qm_complex matrix[4], void* state, ITYPE dim, void* stream, unsigned int device_number); /* :76 */

/* ------------------------------------------------------------------ multi-qubit matrices
 * reference src/gpusim/update_ops_cuda.h:83-140 */
QM_EXPORT void penta_qubit_dense_matrix_gate_host(const unsigned int target_qubit_index[5], const qm_complex matrix[1024], void* state, ITYPE dim, void* stream, UINT device_number); /* :83 */
QM_EXPORT void quad_qubit_dense_matrix_gate_host(const unsigned int target_qubit_index[4], const qm_complex matrix[256], void* state, ITYPE dim, void* stream, unsigned int device_number); /* :86 */
QM_EXPORT void triple_qubit_dense_matrix_gate_host(unsigned int target0_qubit_index, unsigned int target1_qubit_index, unsigned int target2_qubit_index, const qm_complex matrix[64], void* state, ITYPE dim, void* stream, unsigned int device_number); /* :89 */
QM_EXPORT void double_qubit_dense_matrix_gate_host(unsigned int target0_qubit_index, unsigned int target1_qubit_index, const qm_complex matrix[16], void* state, ITYPE dim, void* stream, unsigned int device_number); /* :93 */
QM_EXPORT void multi_qubit_Pauli_gate_Z_mask_host(ITYPE phase_flip_mask, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :97 */
QM_EXPORT void multi_qubit_Pauli_gate_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :99 */
QM_EXPORT void multi_qubit_Pauli_rotation_gate_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :103 */
QM_EXPORT void multi_qubit_Pauli_rotation_gate_Z_mask_host(ITYPE phase_flip_mask, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :107 */
QM_EXPORT void multi_qubit_Pauli_gate_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :110 */
QM_EXPORT void multi_qubit_Pauli_gate_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :114 */
QM_EXPORT void multi_qubit_Pauli_rotation_gate_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :117 */
QM_EXPORT void multi_qubit_Pauli_rotation_gate_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :121 */
QM_EXPORT void multi_qubit_dense_matrix_gate_host(const UINT* target_qubit_index_list, UINT target_qubit_index_count, const qm_complex* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :124 */
QM_EXPORT void single_qubit_control_multi_qubit_dense_matrix_gate_host(UINT control_qubit_index, UINT control_value, const UINT* target_qubit_index_list, UINT target_qubit_index_count, const qm_complex* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :128 */
QM_EXPORT void multi_qubit_control_multi_qubit_dense_matrix_gate_host(const UINT* control_qubit_index_list, const UINT* control_value_list, UINT control_qubit_index_count, const UINT* target_qubit_index_list, UINT target_qubit_index_count, const qm_complex* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :133 */
/* Whole-state diagonal: state[i] *= diagonal_matrix[i] for all i < dim. */
QM_EXPORT void multi_qubit_diagonal_matrix_gate_host(const qm_complex* diagonal_matrix, void* state, ITYPE dim, void* stream, UINT device_number); /* :138 */

/* Diagonal gate on a target list, optionally controlled: the GPU form of
 * csim multi_qubit_control_multi_qubit_diagonal_matrix_gate
 * (src/csim/update_ops_matrix_diagonal_multi.cpp:60-112), which the reference
 * GPU path lacks (src/cppsim/gate_matrix_diagonal.cpp:101-104). */
QM_EXPORT void qm_multi_qubit_control_multi_qubit_diagonal_matrix_gate_host(const UINT* control_qubit_index_list, const UINT* control_value_list, UINT control_qubit_index_count, const UINT* target_qubit_index_list, UINT target_qubit_index_count, const qm_complex* diagonal_element, void* state, ITYPE dim, void* stream, unsigned int device_number);

/* ------------------------------------------------------------------ reductions / state arithmetic
 * reference src/gpusim/stat_ops.h:7-63 */
QM_EXPORT double state_norm_squared_cublas_host(void* state, ITYPE dim); /* :7 (no BLAS here: same as state_norm_squared_host on the null stream) */
QM_EXPORT double state_norm_squared_host(void* state, ITYPE dim, void* stream, unsigned int device_number); /* :8 */
QM_EXPORT double measurement_distribution_entropy_host(void* state, ITYPE dim, void* stream, unsigned int device_number); /* :10 */
QM_EXPORT void state_add_host(void* state_added, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :12 */
QM_EXPORT void state_multiply_host(qm_complex coef, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :14 */
QM_EXPORT qm_complex inner_product_host(const void* bra_state, const void* ket_state, ITYPE dim, void* stream, unsigned int device_number); /* :16 */
QM_EXPORT double expectation_value_single_qubit_Pauli_operator_host(unsigned int operator_index, unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :18 */
QM_EXPORT double M0_prob_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :26 */
QM_EXPORT double M1_prob_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :28 */
QM_EXPORT double marginal_prob_host(UINT* sorted_target_qubit_index_list, UINT* measured_value_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :30 */
QM_EXPORT double expectation_value_multi_qubit_Pauli_operator_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :33 */
QM_EXPORT double expectation_value_multi_qubit_Pauli_operator_Z_mask_host(ITYPE phase_flip_mask, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :37 */
QM_EXPORT double expectation_value_multi_qubit_Pauli_operator_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :40 */
QM_EXPORT double expectation_value_multi_qubit_Pauli_operator_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, void* state, ITYPE dim, void* stream, unsigned int device_number); /* :44 */
QM_EXPORT qm_complex transition_amplitude_multi_qubit_Pauli_operator_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number); /* :47 */
QM_EXPORT qm_complex transition_amplitude_multi_qubit_Pauli_operator_Z_mask_host(ITYPE phase_flip_mask, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number); /* :51 */
QM_EXPORT qm_complex transition_amplitude_multi_qubit_Pauli_operator_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number); /* :54 */
QM_EXPORT qm_complex transition_amplitude_multi_qubit_Pauli_operator_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number); /* :59 */

/* state += coef * state_added (csim state_add_with_coef,
 * src/csim/update_ops_named_state.cpp:62-76; cppsim reaches it through
 * QuantumStateBase::add_state_with_coef, src/cppsim/state.hpp). */
QM_EXPORT void qm_state_add_with_coef_host(qm_complex coef, const void* state_added, void* state, ITYPE dim, void* stream, unsigned int device_number);

/* Whole-observable expectation in one call: returns sum_k coef_k <P_k> with
 * P_k given by its masks (bit_flip, phase_flip, Y count; the pivot is derived).
 * Replaces the per-term loop of GeneralQuantumOperator::get_expectation_value
 * (src/cppsim/general_quantum_operator.cpp:105-141), which on the GPU issues one
 * allocation, launch and synchronisation per term. */
QM_EXPORT qm_complex qm_observable_expectation_host(const ITYPE* bit_flip_masks, const ITYPE* phase_flip_masks, const UINT* global_phase_90rot_counts, const qm_complex* coefs, UINT term_count, void* state, ITYPE dim, void* stream, unsigned int device_number);

/* Allocation-free variant of qm_observable_expectation_host, with no
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, same kernels and the same result bit for bit;
 * it calls no hipMalloc / hipFree.  Its term table, partial sums and results
 * live in one library-owned device array of about 4 MiB per device, held until
 * the process ends.  At most 128 terms per call (more is
 * QM_ERR_INVALID_ARGUMENT).  Calls that overlap on one device share that
 * array and race: serialise them. */
QM_EXPORT qm_complex qm_observable_expectation_noalloc_host(const ITYPE* bit_flip_masks, const ITYPE* phase_flip_masks, const UINT* global_phase_90rot_counts, const qm_complex* coefs, UINT term_count, void* state, ITYPE dim, void* stream, unsigned int device_number);

/* Sampling with caller-supplied uniforms: out[c] = lower_bound(cdf, u[c]) - 1
 * over the sequential un-normalised CDF cdf[0] = 0, cdf[i+1] = cdf[i] + |psi_i|^2,
 * exactly as QuantumStateCpu::sampling (src/cppsim/state.hpp:909-930), with
 * |psi|^2 = fma(re, re, im*im).  The CDF is built on the device with the same
 * sequence of IEEE roundings, so the indices equal the sequential CPU loop's.
 * uniforms and out are host arrays of `count` entries. */
QM_EXPORT void qm_sampling_host(const double* uniforms, UINT count, ITYPE* out, void* state, ITYPE dim, void* stream, unsigned int device_number);
/* The CDF qm_sampling_host searches: cdf_out[0..dim] (host, dim + 1 doubles),
 * equal bit for bit to the sequential loop's stacked probabilities. */
QM_EXPORT void qm_sampling_cdf_host(double* cdf_out, void* state, ITYPE dim, void* stream, unsigned int device_number);
/* Sampling with a boundary check (no reference counterpart; the reference's
 * contract is that of qm_sampling_host).  Two states that differ by rounding
 * (a fused circuit against the gate-by-gate loop, or either against the CPU)
 * give CDFs that differ by a few units in the last place, so a uniform that
 * lies that close to a CDF step can select a neighbouring index.  Each sample
 * whose uniform u is within delta = margin * cdf[dim] of the CDF values that
 * bracket it (cdf[j-1] < u <= cdf[j]) gets flags[c] = 1; every sample with
 * flags[c] = 0 is the same for any CDF whose entries lie within delta of this
 * one.  `flagged` (may be NULL) receives the number of flagged samples.
 * margin = 0 flags nothing; qm_sampling_margin gives the margin for a state
 * that `gate_count` gates produced. */
QM_EXPORT void qm_sampling_checked_host(const double* uniforms, UINT count, double margin, ITYPE* out, unsigned char* flags, UINT* flagged, void* state, ITYPE dim, void* stream, unsigned int device_number);
/* Relative CDF margin covering the rounding difference between two
 * evaluations of a circuit of `gate_count` gates on a normalised state. */
QM_EXPORT double qm_sampling_margin(UINT gate_count);

/* ------------------------------------------------------------------ density matrix
 * `state` holds dim*dim complex128 values, row-major rho[y*dim + x]; `dim` is
 * 2^n as in csim's dm_* functions (src/csim/update_ops_dm.hpp, stat_ops_dm.hpp). */
QM_EXPORT void dm_initialize_quantum_state_host(void* state, ITYPE dim, void* stream, unsigned int device_number); /* memory_ops_dm.cpp:22-37 */
QM_EXPORT void dm_initialize_with_pure_state_host(void* state, const void* pure_state, ITYPE dim, void* stream, unsigned int device_number); /* memory_ops_dm.cpp:39-51 */
QM_EXPORT void dm_single_qubit_dense_matrix_gate_host(UINT target_qubit_index, const qm_complex matrix[4], void* state, ITYPE dim, void* stream, unsigned int device_number); /* update_ops_dm.cpp:33-102 */
QM_EXPORT void dm_multi_qubit_dense_matrix_gate_host(const UINT* target_qubit_index_list, UINT target_qubit_index_count, const qm_complex* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number); /* update_ops_dm.cpp:342-498 */
QM_EXPORT void dm_multi_qubit_control_multi_qubit_dense_matrix_gate_host(const UINT* control_qubit_index_list, const UINT* control_value_list, UINT control_qubit_index_count, const UINT* target_qubit_index_list, UINT target_qubit_index_count, const qm_complex* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number); /* update_ops_dm.cpp:500-688 */
/* General CPTP map on k qubits given as its superoperator S = sum_i K_i (x) conj(K_i)
 * (4^k x 4^k, row index = (row bits of the k targets) * 2^k + (column bits)),
 * applied in one pass.  Covers QuantumGate_Probabilistic / CPTP on a density
 * matrix (src/cppsim/gate_general.hpp:265-285) without its ~10 passes. */
QM_EXPORT void dm_superoperator_host(const UINT* target_qubit_index_list, UINT target_qubit_index_count, const qm_complex* superop, void* state, ITYPE dim, void* stream, unsigned int device_number);
/* Single-qubit depolarizing channel with probability p (DepolarizingNoise,
 * src/cppsim/gate_factory.cpp:454-467) in one pass. */
QM_EXPORT void dm_depolarizing_host(UINT target_qubit_index, double p, void* state, ITYPE dim, void* stream, unsigned int device_number);
/* Single-qubit channels on distinct qubits (1..4 of them), each given by its
 * 4 x 4 superoperator (row index = row bit * 2 + column bit, as for
 * dm_superoperator_host), applied together in one pass.  Each superoperator
 * must map {00, 11} and {01, 10} onto themselves (Pauli channels, amplitude and
 * phase damping); anything else is QM_ERR_INVALID_ARGUMENT.  Replaces one
 * QuantumGate_Probabilistic / CPTP update per channel on a density matrix
 * (src/cppsim/gate_general.hpp:265-285), e.g. the noise after a multi-qubit gate. */
QM_EXPORT void qm_dm_single_qubit_channels_host(const UINT* qubits, UINT count, const qm_complex* superops, void* state, ITYPE dim, void* stream, unsigned int device_number);
QM_EXPORT double dm_state_norm_squared_host(void* state, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops_dm.cpp:12-26 */
QM_EXPORT double dm_measurement_distribution_entropy_host(void* state, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops_dm.cpp:28-47 */
QM_EXPORT double dm_M0_prob_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops_dm.cpp:49-63 */
QM_EXPORT double dm_marginal_prob_host(const UINT* sorted_target_qubit_index_list, const UINT* measured_value_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops_dm.cpp:89-114 */
QM_EXPORT double dm_expectation_value_multi_qubit_Pauli_operator_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops_dm.cpp:164-191 */
QM_EXPORT void dm_normalize_host(double squared_norm, void* state, ITYPE dim, void* stream, unsigned int device_number); /* update_ops_dm.cpp:14-31 */
/* DM sampling over |rho_ii| (src/cppsim/state_dm.hpp:383-408), same contract as qm_sampling_host. */
QM_EXPORT void dm_sampling_host(const double* uniforms, UINT count, ITYPE* out, void* state, ITYPE dim, void* stream, unsigned int device_number);
/* The CDF dm_sampling_host searches (dim + 1 doubles, host), and the sampling
 * with a boundary check (as qm_sampling_checked_host).  No reference counterpart. */
QM_EXPORT void dm_sampling_cdf_host(double* cdf_out, void* state, ITYPE dim, void* stream, unsigned int device_number);
QM_EXPORT void dm_sampling_checked_host(const double* uniforms, UINT count, double margin, ITYPE* out, unsigned char* flags, UINT* flagged, void* state, ITYPE dim, void* stream, unsigned int device_number);

/* ------------------------------------------------------------------ state utilities
 * No GPU counterpart in the reference (its QuantumStateGpu has none of these);
 * the csim names and arguments of src/csim/stat_ops.cpp:111-160 and
 * stat_ops_dm.cpp:193-298 with the device / stream pair appended.  dst is a
 * separate, already allocated device buffer of the output size. */
QM_EXPORT void state_tensor_product_host(const void* state_left, ITYPE dim_left, const void* state_right, ITYPE dim_right, void* state_dst, void* stream, unsigned int device_number); /* stat_ops.cpp:111-121 */
QM_EXPORT void state_permutate_qubit_host(const UINT* qubit_order, const void* state_src, void* state_dst, UINT qubit_count, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops.cpp:122-135 */
QM_EXPORT void state_drop_qubits_host(const UINT* target, const UINT* projection, UINT target_count, const void* state_src, void* state_dst, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops.cpp:136-160 */
QM_EXPORT void dm_state_tensor_product_host(const void* state_left, ITYPE dim_left, const void* state_right, ITYPE dim_right, void* state_dst, void* stream, unsigned int device_number); /* stat_ops_dm.cpp:193-212 */
QM_EXPORT void dm_state_permutate_qubit_host(const UINT* qubit_order, const void* state_src, void* state_dst, UINT qubit_count, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops_dm.cpp:213-232 */
QM_EXPORT void dm_state_partial_trace_from_density_matrix_host(const UINT* target, UINT target_count, const void* state_src, void* state_dst, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops_dm.cpp:233-265 */
QM_EXPORT void dm_state_partial_trace_from_state_vector_host(const UINT* target, UINT target_count, const void* state_src, void* state_dst, ITYPE dim, void* stream, unsigned int device_number); /* stat_ops_dm.cpp:266-298 */

/* ------------------------------------------------------------------ batched circuits
 * Many small circuits in one launch.  Replaces, for each circuit of a batch,
 * QuantumCircuit::update_quantum_state on |0...0> (src/cppsim/circuit.cpp:46-57)
 * followed by Observable::get_expectation_value (src/cppsim/observable.cpp:56-58,
 * general_quantum_operator.cpp:105-141): no reference counterpart runs more than
 * one circuit per call.  A batch is built on the host (add circuits, compile),
 * then run any number of times; each run is one kernel launch in which every
 * work-group takes circuits from a queue and keeps the state in LDS (up to 13
 * qubits; 14..20 qubits in 13-qubit tiles staged through device memory).
 *
 * Gate kinds: QM_BATCH_DENSE (1..4 qubits, row-major 2^k x 2^k matrix, bit i of
 * the matrix index <-> qubits[i]), QM_BATCH_DIAG (1..4 qubits, 2^k diagonal),
 * QM_BATCH_CZ (qubits[0], qubits[1]), QM_BATCH_CNOT (qubits[0] = control,
 * qubits[1] = target).  `matrix` is host memory read during add_circuit. */
#define QM_BATCH_DENSE 0u
#define QM_BATCH_DIAG 1u
#define QM_BATCH_CZ 2u
#define QM_BATCH_CNOT 3u
#define QM_BATCH_MAX_QUBITS 20u
typedef struct qm_batch_gate {
    UINT kind;
    UINT qubit_count;
    UINT qubits[4];
    const qm_complex* matrix;
} qm_batch_gate;
typedef struct qm_batch qm_batch;
QM_EXPORT qm_batch* qm_batch_create(void);
QM_EXPORT void qm_batch_destroy(qm_batch* batch);
/* Adds one circuit on `qubit_count` (1..20) qubits starting from |0...0> and its
 * observable sum_k coefs[k] P_k (P_k given by masks as in
 * qm_observable_expectation_host; real coefficients).  Returns its index, or -1. */
QM_EXPORT int qm_batch_add_circuit(qm_batch* batch, UINT qubit_count, const qm_batch_gate* gates, UINT gate_count, const ITYPE* bit_flip_masks, const ITYPE* phase_flip_masks, const UINT* global_phase_90rot_counts, const double* coefs, UINT term_count);
/* Schedules every circuit into register-blocked passes of `block_qubits` (2..4;
 * 0 = 4) and uploads the programs.  keep_states != 0 keeps every final state
 * in device memory (qm_batch_get_state).  Returns QM_OK or an error code. */
QM_EXPORT int qm_batch_compile(qm_batch* batch, UINT block_qubits, int keep_states, unsigned int device_number);
/* Runs every circuit; energies[i] (host array) receives circuit i's expectation
 * value.  Returns QM_OK or an error code. */
QM_EXPORT int qm_batch_run(qm_batch* batch, double* energies, void* stream, unsigned int device_number);
/* As qm_batch_run but leaves the energies in device memory (returned pointer,
 * valid until the next compile or destroy) without synchronising the stream. */
QM_EXPORT const double* qm_batch_run_async(qm_batch* batch, void* stream, unsigned int device_number);
/* Copies circuit i's final state (2^n complex128) to host memory after a run
 * compiled with keep_states. */
QM_EXPORT int qm_batch_get_state(qm_batch* batch, UINT index, qm_complex* out, void* stream, unsigned int device_number);
/* Qubits held in LDS per tile (8..13, default 13); takes effect at the next
 * compile.  A smaller window halves LDS per work-group (two work-groups per CU
 * at 12) and tiles larger circuits. */
QM_EXPORT int qm_batch_set_local_qubits(qm_batch* batch, UINT local_qubits);
/* Captures one run (queue reset + the launch) on `stream` (a created stream,
 * not NULL) into a HIP graph; qm_batch_run_graph replays it.  A compile drops
 * the graph.  The graph holds kernel nodes only. */
QM_EXPORT int qm_batch_capture(qm_batch* batch, void* stream, unsigned int device_number);
QM_EXPORT int qm_batch_run_graph(qm_batch* batch, double* energies, void* stream, unsigned int device_number);
/* Removes every circuit (device memory is kept for the next compile). */
QM_EXPORT void qm_batch_clear(qm_batch* batch);
/* Device address of circuit i's final state (compiled with keep_states). */
QM_EXPORT void* qm_batch_state_device_ptr(qm_batch* batch, UINT index);
/* Schedule statistics: number of circuits, stages, passes and sub-operations. */
QM_EXPORT void qm_batch_stats(const qm_batch* batch, ITYPE* circuits, ITYPE* stages, ITYPE* passes, ITYPE* subops);

/* ------------------------------------------------------------------ fused circuits
 * One circuit applied in place to an existing state with the batch scheduler:
 * gates are grouped into stages whose qubits fit a window of `local_qubits`
 * (8..13; qubits 0..3 always included), one launch per stage, each work-group
 * holding 2^local_qubits-amplitude tiles in LDS and applying every gate of the
 * stage to a tile before writing it back.  Replaces QuantumCircuit::update_quantum_state's
 * one launch per gate (src/cppsim/circuit.cpp:46-57) for circuits of 4..40
 * qubits made of the batch gate kinds.  create() does host work only. */
#define QM_CIRCUIT_MAX_QUBITS 40u
typedef struct qm_circuit qm_circuit;
/* block_qubits 2..4, local_qubits 8..13; 0 picks by qubit count (3 and 12 up to
 * 19 qubits, else 4 and 13; the width grows to the largest dense gate, at most
 * 4 qubits); NULL on error. */
QM_EXPORT qm_circuit* qm_circuit_create(UINT qubit_count, const qm_batch_gate* gates, UINT gate_count, UINT block_qubits, UINT local_qubits);
QM_EXPORT void qm_circuit_destroy(qm_circuit* circuit);
/* Applies the circuit to `state` (dim = 2^qubit_count) in stream order. */
QM_EXPORT int qm_circuit_apply(qm_circuit* circuit, void* state, ITYPE dim, void* stream, unsigned int device_number);
QM_EXPORT void qm_circuit_stats(const qm_circuit* circuit, ITYPE* stages, ITYPE* passes, ITYPE* subops);

/* ------------------------------------------------------------------ adjoint gradients
 * d<psi(theta)|H|psi(theta)>/d theta for a circuit of fixed gates (batch kinds
 * on 1..2 qubits) and parametric rotations exp(i coef theta[param] P / 2) with
 * P a Pauli string on 1..4 qubits, for `batch` parameter sets at once.
 * Restates ParametricQuantumCircuit::backprop (src/vqcsim/parametric_circuit.cpp:205-335):
 * gates are grouped into passes over blocks of 3 qubits held in registers; on
 * the way back each pass adds its rotations' gradient terms and applies the
 * inverses to both states. */
#define QM_ADJ_ROT 4u
typedef struct qm_adjoint_gate {
    UINT kind;                 /* QM_BATCH_DENSE/DIAG/CZ/CNOT (fixed) or QM_ADJ_ROT */
    UINT qubit_count;
    UINT qubits[4];
    const qm_complex* matrix;  /* fixed dense / diagonal gates (host) */
    UINT paulis[4];            /* QM_ADJ_ROT: 1 = X, 2 = Y, 3 = Z on qubits[i] */
    UINT param;                /* QM_ADJ_ROT: index into theta */
    double coef;               /* QM_ADJ_ROT: angle = coef * theta[param] */
} qm_adjoint_gate;
typedef struct qm_adjoint qm_adjoint;
QM_EXPORT qm_adjoint* qm_adjoint_create(UINT qubit_count, const qm_adjoint_gate* gates, UINT gate_count, UINT param_count);
QM_EXPORT void qm_adjoint_destroy(qm_adjoint* circuit);
/* Register-block passes one sweep issues (0: one kernel per gate, for states
 * below 3 qubits or rotations with more than 3 X/Y qubits) and gates. */
QM_EXPORT void qm_adjoint_stats(const qm_adjoint* circuit, ITYPE* passes, ITYPE* gates);
/* Device memory qm_adjoint_run needs for `batch` parameter sets. */
QM_EXPORT ITYPE qm_adjoint_workspace_bytes(const qm_adjoint* circuit, UINT batch);
/* theta: device, batch x param_count (row-major); energy: device, batch;
 * grad: device, batch x param_count; workspace: device, qm_adjoint_workspace_bytes.
 * The observable is sum_k coefs[k] P_k in mask form (host arrays, real coefficients). */
QM_EXPORT int qm_adjoint_run(qm_adjoint* circuit, UINT batch, const double* theta, const ITYPE* bit_flip_masks, const ITYPE* phase_flip_masks, const UINT* global_phase_90rot_counts, const double* coefs, UINT term_count, double* energy, double* grad, void* workspace, void* stream, unsigned int device_number);

/* ------------------------------------------------------------------ request server
 * Circuits of mixed size (1..40 qubits) submitted from any thread and run on
 * one device with continuous batching: each request is a circuit from
 * |0...0> (batch gate kinds), a real-coefficient Pauli observable and uniforms;
 * it returns <H> and the sampled indices (as qm_sampling_host).  Requests are
 * admitted while the sum of 16 * 2^n bytes stays within `hbm_budget_bytes`;
 * requests of up to `small_max_qubits` (<= 16) qubits are batched into one
 * launch, larger ones run on up to `large_slots` streams at once.  No reference
 * counterpart (the reference runs one circuit per call, src/cppsim/circuit.cpp:46-57). */
typedef struct qm_server qm_server;
QM_EXPORT qm_server* qm_server_create(ITYPE hbm_budget_bytes, UINT large_slots, UINT small_max_qubits, unsigned int device_number);
/* Copies the request; returns its id (>= 0) or -1 on an invalid request. */
QM_EXPORT long long qm_server_submit(qm_server* server, UINT qubit_count, const qm_batch_gate* gates, UINT gate_count, const ITYPE* bit_flip_masks, const ITYPE* phase_flip_masks, const UINT* global_phase_90rot_counts, const double* coefs, UINT term_count, const double* uniforms, UINT sample_count);
/* Waits for request `id` (once per id): energy, samples (sample_count entries)
 * and the time from submission to completion. */
QM_EXPORT int qm_server_wait(qm_server* server, long long id, double* energy, ITYPE* samples, double* latency_seconds);
/* As qm_server_wait, plus the boundary check of each sample (flags: sample_count
 * bytes, as qm_sampling_checked_host with qm_sampling_margin(gate_count)) and
 * the number of flagged samples.  A request's samples are computed on the
 * state its fused circuit produced, which differs from the gate-by-gate state
 * by rounding; the unflagged samples are the ones both states agree on. */
QM_EXPORT int qm_server_wait_checked(qm_server* server, long long id, double* energy, ITYPE* samples, unsigned char* flags, UINT* flagged, double* latency_seconds);
QM_EXPORT void qm_server_stats(qm_server* server, ITYPE* admitted, ITYPE* completed, ITYPE* failed, ITYPE* batches, ITYPE* max_batch, ITYPE* peak_bytes);
/* Finishes every submitted request, then frees the server. */
QM_EXPORT void qm_server_destroy(qm_server* server);

#ifdef __cplusplus
}
#endif

#endif /* QULACS_MI355_H_ */
