// C++-linkage entry points with the exact signatures of the reference's GPU
// seam (src/gpusim/update_ops_cuda.h, stat_ops.h, memory_ops.h, util_func.h),
// whose declarations have C++ linkage and std::complex<double> arguments.  A
// cppsim built with its GPU option links against these mangled names; each one
// forwards to the extern "C" entry point of the same name (bound here through
// an asm label, since both cannot be declared under one C++ name).
#include <complex>

typedef unsigned long long ITYPE;
typedef unsigned int UINT;
typedef std::complex<double> CPPCTYPE;
struct QmC {
    double re, im;
};
static_assert(sizeof(QmC) == sizeof(CPPCTYPE), "complex layout");

#define QM_CXX_EXPORT __attribute__((visibility("default")))

static inline const QmC* qc(const CPPCTYPE* p) { return reinterpret_cast<const QmC*>(p); }
static inline QmC qc(CPPCTYPE z) { return QmC{z.real(), z.imag()}; }
static inline CPPCTYPE cc(QmC z) { return CPPCTYPE(z.re, z.im); }

extern "C" void qmc_H_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("H_gate_host");
QM_CXX_EXPORT void H_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_H_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_X_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("X_gate_host");
QM_CXX_EXPORT void X_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_X_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_Y_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("Y_gate_host");
QM_CXX_EXPORT void Y_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_Y_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_Z_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("Z_gate_host");
QM_CXX_EXPORT void Z_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_Z_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_CZ_gate_host(unsigned int control_qubit_index, unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("CZ_gate_host");
QM_CXX_EXPORT void CZ_gate_host(unsigned int control_qubit_index, unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_CZ_gate_host(control_qubit_index, target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_CNOT_gate_host(unsigned int control_qubit_index, unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("CNOT_gate_host");
QM_CXX_EXPORT void CNOT_gate_host(unsigned int control_qubit_index, unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_CNOT_gate_host(control_qubit_index, target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_SWAP_gate_host(unsigned int target_qubit_index0, unsigned int target_qubit_index1, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("SWAP_gate_host");
QM_CXX_EXPORT void SWAP_gate_host(unsigned int target_qubit_index0, unsigned int target_qubit_index1, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_SWAP_gate_host(target_qubit_index0, target_qubit_index1, state, dim, stream, device_number); }

extern "C" void qmc_RX_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("RX_gate_host");
QM_CXX_EXPORT void RX_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_RX_gate_host(target_qubit_index, angle, state, dim, stream, device_number); }

extern "C" void qmc_RY_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("RY_gate_host");
QM_CXX_EXPORT void RY_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_RY_gate_host(target_qubit_index, angle, state, dim, stream, device_number); }

extern "C" void qmc_RZ_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("RZ_gate_host");
QM_CXX_EXPORT void RZ_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_RZ_gate_host(target_qubit_index, angle, state, dim, stream, device_number); }

extern "C" void qmc_S_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("S_gate_host");
QM_CXX_EXPORT void S_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_S_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_Sdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("Sdag_gate_host");
QM_CXX_EXPORT void Sdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_Sdag_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_T_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("T_gate_host");
QM_CXX_EXPORT void T_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_T_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_Tdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("Tdag_gate_host");
QM_CXX_EXPORT void Tdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_Tdag_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_sqrtX_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("sqrtX_gate_host");
QM_CXX_EXPORT void sqrtX_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_sqrtX_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_sqrtXdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("sqrtXdag_gate_host");
QM_CXX_EXPORT void sqrtXdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_sqrtXdag_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_sqrtY_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("sqrtY_gate_host");
QM_CXX_EXPORT void sqrtY_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_sqrtY_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_sqrtYdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("sqrtYdag_gate_host");
QM_CXX_EXPORT void sqrtYdag_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_sqrtYdag_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_P0_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("P0_gate_host");
QM_CXX_EXPORT void P0_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_P0_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_P1_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("P1_gate_host");
QM_CXX_EXPORT void P1_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_P1_gate_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_normalize_host(double squared_norm, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("normalize_host");
QM_CXX_EXPORT void normalize_host(double squared_norm, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_normalize_host(squared_norm, state, dim, stream, device_number); }

extern "C" void qmc_single_qubit_Pauli_gate_host(UINT target_qubit_index, UINT Pauli_operator_type, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("single_qubit_Pauli_gate_host");
QM_CXX_EXPORT void single_qubit_Pauli_gate_host(UINT target_qubit_index, UINT Pauli_operator_type, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_single_qubit_Pauli_gate_host(target_qubit_index, Pauli_operator_type, state, dim, stream, device_number); }

extern "C" void qmc_single_qubit_Pauli_rotation_gate_host(unsigned int target_qubit_index, unsigned int op_idx, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("single_qubit_Pauli_rotation_gate_host");
QM_CXX_EXPORT void single_qubit_Pauli_rotation_gate_host(unsigned int target_qubit_index, unsigned int op_idx, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_single_qubit_Pauli_rotation_gate_host(target_qubit_index, op_idx, angle, state, dim, stream, device_number); }

extern "C" void qmc_single_qubit_dense_matrix_gate_host(unsigned int target_qubit_index, const QmC* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("single_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void single_qubit_dense_matrix_gate_host(unsigned int target_qubit_index, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_single_qubit_dense_matrix_gate_host(target_qubit_index, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_single_qubit_diagonal_matrix_gate_host(unsigned int target_qubit_index, const QmC* diagonal_matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("single_qubit_diagonal_matrix_gate_host");
QM_CXX_EXPORT void single_qubit_diagonal_matrix_gate_host(unsigned int target_qubit_index, const CPPCTYPE* diagonal_matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_single_qubit_diagonal_matrix_gate_host(target_qubit_index, qc(diagonal_matrix), state, dim, stream, device_number); }

extern "C" void qmc_single_qubit_control_single_qubit_dense_matrix_gate_host(unsigned int control_qubit_index, unsigned int control_value, unsigned int target_qubit_index, const QmC* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("single_qubit_control_single_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void single_qubit_control_single_qubit_dense_matrix_gate_host(unsigned int control_qubit_index, unsigned int control_value, unsigned int target_qubit_index, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_single_qubit_control_single_qubit_dense_matrix_gate_host(control_qubit_index, control_value, target_qubit_index, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_single_qubit_phase_gate_host(unsigned int target_qubit_index, QmC phase, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("single_qubit_phase_gate_host");
QM_CXX_EXPORT void single_qubit_phase_gate_host(unsigned int target_qubit_index, CPPCTYPE phase, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_single_qubit_phase_gate_host(target_qubit_index, qc(phase), state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_control_single_qubit_dense_matrix_gate_host(const UINT* control_qubit_index_list, const UINT* control_value_list, UINT control_qubit_index_count, UINT target_qubit_index, const QmC* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_control_single_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void multi_qubit_control_single_qubit_dense_matrix_gate_host(const UINT* control_qubit_index_list, const UINT* control_value_list, UINT control_qubit_index_count, UINT target_qubit_index, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_control_single_qubit_dense_matrix_gate_host(control_qubit_index_list, control_value_list, control_qubit_index_count, target_qubit_index, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_penta_qubit_dense_matrix_gate_host(const unsigned int* target_qubit_index, const QmC* matrix, void* state, ITYPE dim, void* stream, UINT device_number) __asm__("penta_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void penta_qubit_dense_matrix_gate_host(const unsigned int* target_qubit_index, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, UINT device_number) { qmc_penta_qubit_dense_matrix_gate_host(target_qubit_index, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_quad_qubit_dense_matrix_gate_host(const unsigned int* target_qubit_index, const QmC* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("quad_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void quad_qubit_dense_matrix_gate_host(const unsigned int* target_qubit_index, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_quad_qubit_dense_matrix_gate_host(target_qubit_index, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_triple_qubit_dense_matrix_gate_host(unsigned int target0_qubit_index, unsigned int target1_qubit_index, unsigned int target2_qubit_index, const QmC* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("triple_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void triple_qubit_dense_matrix_gate_host(unsigned int target0_qubit_index, unsigned int target1_qubit_index, unsigned int target2_qubit_index, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_triple_qubit_dense_matrix_gate_host(target0_qubit_index, target1_qubit_index, target2_qubit_index, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_double_qubit_dense_matrix_gate_host(unsigned int target0_qubit_index, unsigned int target1_qubit_index, const QmC* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("double_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void double_qubit_dense_matrix_gate_host(unsigned int target0_qubit_index, unsigned int target1_qubit_index, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_double_qubit_dense_matrix_gate_host(target0_qubit_index, target1_qubit_index, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_Pauli_gate_Z_mask_host(ITYPE phase_flip_mask, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_Pauli_gate_Z_mask_host");
QM_CXX_EXPORT void multi_qubit_Pauli_gate_Z_mask_host(ITYPE phase_flip_mask, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_Pauli_gate_Z_mask_host(phase_flip_mask, state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_Pauli_gate_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_Pauli_gate_XZ_mask_host");
QM_CXX_EXPORT void multi_qubit_Pauli_gate_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_Pauli_gate_XZ_mask_host(bit_flip_mask, phase_flip_mask, global_phase_90rot_count, pivot_qubit_index, state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_Pauli_rotation_gate_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_Pauli_rotation_gate_XZ_mask_host");
QM_CXX_EXPORT void multi_qubit_Pauli_rotation_gate_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_Pauli_rotation_gate_XZ_mask_host(bit_flip_mask, phase_flip_mask, global_phase_90rot_count, pivot_qubit_index, angle, state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_Pauli_rotation_gate_Z_mask_host(ITYPE phase_flip_mask, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_Pauli_rotation_gate_Z_mask_host");
QM_CXX_EXPORT void multi_qubit_Pauli_rotation_gate_Z_mask_host(ITYPE phase_flip_mask, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_Pauli_rotation_gate_Z_mask_host(phase_flip_mask, angle, state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_Pauli_gate_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_Pauli_gate_partial_list_host");
QM_CXX_EXPORT void multi_qubit_Pauli_gate_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_Pauli_gate_partial_list_host(target_qubit_index_list, Pauli_operator_type_list, target_qubit_index_count, state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_Pauli_gate_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_Pauli_gate_whole_list_host");
QM_CXX_EXPORT void multi_qubit_Pauli_gate_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_Pauli_gate_whole_list_host(Pauli_operator_type_list, qubit_count, state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_Pauli_rotation_gate_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_Pauli_rotation_gate_partial_list_host");
QM_CXX_EXPORT void multi_qubit_Pauli_rotation_gate_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_Pauli_rotation_gate_partial_list_host(target_qubit_index_list, Pauli_operator_type_list, target_qubit_index_count, angle, state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_Pauli_rotation_gate_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_Pauli_rotation_gate_whole_list_host");
QM_CXX_EXPORT void multi_qubit_Pauli_rotation_gate_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, double angle, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_Pauli_rotation_gate_whole_list_host(Pauli_operator_type_list, qubit_count, angle, state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_dense_matrix_gate_host(const UINT* target_qubit_index_list, UINT target_qubit_index_count, const QmC* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void multi_qubit_dense_matrix_gate_host(const UINT* target_qubit_index_list, UINT target_qubit_index_count, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_dense_matrix_gate_host(target_qubit_index_list, target_qubit_index_count, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_single_qubit_control_multi_qubit_dense_matrix_gate_host(UINT control_qubit_index, UINT control_value, const UINT* target_qubit_index_list, UINT target_qubit_index_count, const QmC* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("single_qubit_control_multi_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void single_qubit_control_multi_qubit_dense_matrix_gate_host(UINT control_qubit_index, UINT control_value, const UINT* target_qubit_index_list, UINT target_qubit_index_count, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_single_qubit_control_multi_qubit_dense_matrix_gate_host(control_qubit_index, control_value, target_qubit_index_list, target_qubit_index_count, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_control_multi_qubit_dense_matrix_gate_host(const UINT* control_qubit_index_list, const UINT* control_value_list, UINT control_qubit_index_count, const UINT* target_qubit_index_list, UINT target_qubit_index_count, const QmC* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("multi_qubit_control_multi_qubit_dense_matrix_gate_host");
QM_CXX_EXPORT void multi_qubit_control_multi_qubit_dense_matrix_gate_host(const UINT* control_qubit_index_list, const UINT* control_value_list, UINT control_qubit_index_count, const UINT* target_qubit_index_list, UINT target_qubit_index_count, const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_multi_qubit_control_multi_qubit_dense_matrix_gate_host(control_qubit_index_list, control_value_list, control_qubit_index_count, target_qubit_index_list, target_qubit_index_count, qc(matrix), state, dim, stream, device_number); }

extern "C" void qmc_multi_qubit_diagonal_matrix_gate_host(const QmC* diagonal_matrix, void* state, ITYPE dim, void* stream, UINT device_number) __asm__("multi_qubit_diagonal_matrix_gate_host");
QM_CXX_EXPORT void multi_qubit_diagonal_matrix_gate_host(const CPPCTYPE* diagonal_matrix, void* state, ITYPE dim, void* stream, UINT device_number) { qmc_multi_qubit_diagonal_matrix_gate_host(qc(diagonal_matrix), state, dim, stream, device_number); }

extern "C" double qmc_state_norm_squared_cublas_host(void* state, ITYPE dim) __asm__("state_norm_squared_cublas_host");
QM_CXX_EXPORT double state_norm_squared_cublas_host(void* state, ITYPE dim) { return qmc_state_norm_squared_cublas_host(state, dim); }

extern "C" double qmc_state_norm_squared_host(void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("state_norm_squared_host");
QM_CXX_EXPORT double state_norm_squared_host(void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_state_norm_squared_host(state, dim, stream, device_number); }

extern "C" double qmc_measurement_distribution_entropy_host(void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("measurement_distribution_entropy_host");
QM_CXX_EXPORT double measurement_distribution_entropy_host(void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_measurement_distribution_entropy_host(state, dim, stream, device_number); }

extern "C" void qmc_state_add_host(void* state_added, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("state_add_host");
QM_CXX_EXPORT void state_add_host(void* state_added, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_state_add_host(state_added, state, dim, stream, device_number); }

extern "C" void qmc_state_multiply_host(QmC coef, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("state_multiply_host");
QM_CXX_EXPORT void state_multiply_host(CPPCTYPE coef, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_state_multiply_host(qc(coef), state, dim, stream, device_number); }

extern "C" QmC qmc_inner_product_host(const void* bra_state, const void* ket_state, ITYPE dim, void* stream, unsigned int device_number) __asm__("inner_product_host");
QM_CXX_EXPORT CPPCTYPE inner_product_host(const void* bra_state, const void* ket_state, ITYPE dim, void* stream, unsigned int device_number) { return cc(qmc_inner_product_host(bra_state, ket_state, dim, stream, device_number)); }

extern "C" double qmc_expectation_value_single_qubit_Pauli_operator_host(unsigned int operator_index, unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("expectation_value_single_qubit_Pauli_operator_host");
QM_CXX_EXPORT double expectation_value_single_qubit_Pauli_operator_host(unsigned int operator_index, unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_expectation_value_single_qubit_Pauli_operator_host(operator_index, target_qubit_index, state, dim, stream, device_number); }

extern "C" double qmc_M0_prob_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("M0_prob_host");
QM_CXX_EXPORT double M0_prob_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_M0_prob_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" double qmc_M1_prob_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("M1_prob_host");
QM_CXX_EXPORT double M1_prob_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_M1_prob_host(target_qubit_index, state, dim, stream, device_number); }

extern "C" double qmc_marginal_prob_host(UINT* sorted_target_qubit_index_list, UINT* measured_value_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("marginal_prob_host");
QM_CXX_EXPORT double marginal_prob_host(UINT* sorted_target_qubit_index_list, UINT* measured_value_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_marginal_prob_host(sorted_target_qubit_index_list, measured_value_list, target_qubit_index_count, state, dim, stream, device_number); }

extern "C" double qmc_expectation_value_multi_qubit_Pauli_operator_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("expectation_value_multi_qubit_Pauli_operator_XZ_mask_host");
QM_CXX_EXPORT double expectation_value_multi_qubit_Pauli_operator_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_expectation_value_multi_qubit_Pauli_operator_XZ_mask_host(bit_flip_mask, phase_flip_mask, global_phase_90rot_count, pivot_qubit_index, state, dim, stream, device_number); }

extern "C" double qmc_expectation_value_multi_qubit_Pauli_operator_Z_mask_host(ITYPE phase_flip_mask, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("expectation_value_multi_qubit_Pauli_operator_Z_mask_host");
QM_CXX_EXPORT double expectation_value_multi_qubit_Pauli_operator_Z_mask_host(ITYPE phase_flip_mask, void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_expectation_value_multi_qubit_Pauli_operator_Z_mask_host(phase_flip_mask, state, dim, stream, device_number); }

extern "C" double qmc_expectation_value_multi_qubit_Pauli_operator_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("expectation_value_multi_qubit_Pauli_operator_partial_list_host");
QM_CXX_EXPORT double expectation_value_multi_qubit_Pauli_operator_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_expectation_value_multi_qubit_Pauli_operator_partial_list_host(target_qubit_index_list, Pauli_operator_type_list, target_qubit_index_count, state, dim, stream, device_number); }

extern "C" double qmc_expectation_value_multi_qubit_Pauli_operator_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("expectation_value_multi_qubit_Pauli_operator_whole_list_host");
QM_CXX_EXPORT double expectation_value_multi_qubit_Pauli_operator_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, void* state, ITYPE dim, void* stream, unsigned int device_number) { return qmc_expectation_value_multi_qubit_Pauli_operator_whole_list_host(Pauli_operator_type_list, qubit_count, state, dim, stream, device_number); }

extern "C" QmC qmc_transition_amplitude_multi_qubit_Pauli_operator_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number) __asm__("transition_amplitude_multi_qubit_Pauli_operator_XZ_mask_host");
QM_CXX_EXPORT CPPCTYPE transition_amplitude_multi_qubit_Pauli_operator_XZ_mask_host(ITYPE bit_flip_mask, ITYPE phase_flip_mask, UINT global_phase_90rot_count, UINT pivot_qubit_index, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number) { return cc(qmc_transition_amplitude_multi_qubit_Pauli_operator_XZ_mask_host(bit_flip_mask, phase_flip_mask, global_phase_90rot_count, pivot_qubit_index, state_bra, state_ket, dim, stream, device_number)); }

extern "C" QmC qmc_transition_amplitude_multi_qubit_Pauli_operator_Z_mask_host(ITYPE phase_flip_mask, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number) __asm__("transition_amplitude_multi_qubit_Pauli_operator_Z_mask_host");
QM_CXX_EXPORT CPPCTYPE transition_amplitude_multi_qubit_Pauli_operator_Z_mask_host(ITYPE phase_flip_mask, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number) { return cc(qmc_transition_amplitude_multi_qubit_Pauli_operator_Z_mask_host(phase_flip_mask, state_bra, state_ket, dim, stream, device_number)); }

extern "C" QmC qmc_transition_amplitude_multi_qubit_Pauli_operator_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number) __asm__("transition_amplitude_multi_qubit_Pauli_operator_partial_list_host");
QM_CXX_EXPORT CPPCTYPE transition_amplitude_multi_qubit_Pauli_operator_partial_list_host(const UINT* target_qubit_index_list, const UINT* Pauli_operator_type_list, UINT target_qubit_index_count, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number) { return cc(qmc_transition_amplitude_multi_qubit_Pauli_operator_partial_list_host(target_qubit_index_list, Pauli_operator_type_list, target_qubit_index_count, state_bra, state_ket, dim, stream, device_number)); }

extern "C" QmC qmc_transition_amplitude_multi_qubit_Pauli_operator_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number) __asm__("transition_amplitude_multi_qubit_Pauli_operator_whole_list_host");
QM_CXX_EXPORT CPPCTYPE transition_amplitude_multi_qubit_Pauli_operator_whole_list_host(const UINT* Pauli_operator_type_list, UINT qubit_count, void* state_bra, void* state_ket, ITYPE dim, void* stream, unsigned int device_number) { return cc(qmc_transition_amplitude_multi_qubit_Pauli_operator_whole_list_host(Pauli_operator_type_list, qubit_count, state_bra, state_ket, dim, stream, device_number)); }

extern "C" void* qmc_allocate_cuda_stream_host(unsigned int max_cuda_stream, unsigned int device_number) __asm__("allocate_cuda_stream_host");
QM_CXX_EXPORT void* allocate_cuda_stream_host(unsigned int max_cuda_stream, unsigned int device_number) { return qmc_allocate_cuda_stream_host(max_cuda_stream, device_number); }

extern "C" void qmc_release_cuda_stream_host(void* cuda_stream, unsigned int max_cuda_stream, unsigned int device_number) __asm__("release_cuda_stream_host");
QM_CXX_EXPORT void release_cuda_stream_host(void* cuda_stream, unsigned int max_cuda_stream, unsigned int device_number) { qmc_release_cuda_stream_host(cuda_stream, max_cuda_stream, device_number); }

extern "C" void* qmc_allocate_quantum_state_host(ITYPE dim, unsigned int device_number) __asm__("allocate_quantum_state_host");
QM_CXX_EXPORT void* allocate_quantum_state_host(ITYPE dim, unsigned int device_number) { return qmc_allocate_quantum_state_host(dim, device_number); }

extern "C" void qmc_initialize_quantum_state_host(void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("initialize_quantum_state_host");
QM_CXX_EXPORT void initialize_quantum_state_host(void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_initialize_quantum_state_host(state, dim, stream, device_number); }

extern "C" void qmc_release_quantum_state_host(void* state, unsigned int device_number) __asm__("release_quantum_state_host");
QM_CXX_EXPORT void release_quantum_state_host(void* state, unsigned int device_number) { qmc_release_quantum_state_host(state, device_number); }

extern "C" void qmc_initialize_Haar_random_state_host(void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("initialize_Haar_random_state_host");
QM_CXX_EXPORT void initialize_Haar_random_state_host(void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_initialize_Haar_random_state_host(state, dim, stream, device_number); }

extern "C" void qmc_initialize_Haar_random_state_with_seed_host(void* state, ITYPE dim, UINT seed, void* stream, unsigned int device_number) __asm__("initialize_Haar_random_state_with_seed_host");
QM_CXX_EXPORT void initialize_Haar_random_state_with_seed_host(void* state, ITYPE dim, UINT seed, void* stream, unsigned int device_number) { qmc_initialize_Haar_random_state_with_seed_host(state, dim, seed, stream, device_number); }

extern "C" int qmc_get_num_device() __asm__("get_num_device");
QM_CXX_EXPORT int get_num_device() { return qmc_get_num_device(); }

extern "C" void qmc_set_device(unsigned int device_num) __asm__("set_device");
QM_CXX_EXPORT void set_device(unsigned int device_num) { qmc_set_device(device_num); }

extern "C" int qmc_get_current_device() __asm__("get_current_device");
QM_CXX_EXPORT int get_current_device() { return qmc_get_current_device(); }

extern "C" void qmc_set_computational_basis_host(ITYPE comp_basis, void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("set_computational_basis_host");
QM_CXX_EXPORT void set_computational_basis_host(ITYPE comp_basis, void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_set_computational_basis_host(comp_basis, state, dim, stream, device_number); }

extern "C" void qmc_copy_quantum_state_from_device_to_device(void* state_gpu_copy, const void* state_gpu, ITYPE dim, void* stream, unsigned int device_number) __asm__("copy_quantum_state_from_device_to_device");
QM_CXX_EXPORT void copy_quantum_state_from_device_to_device(void* state_gpu_copy, const void* state_gpu, ITYPE dim, void* stream, unsigned int device_number) { qmc_copy_quantum_state_from_device_to_device(state_gpu_copy, state_gpu, dim, stream, device_number); }

extern "C" void qmc_copy_quantum_state_from_host_to_device(void* state_gpu_copy, const void* state, ITYPE dim, void* stream, unsigned int device_number) __asm__("copy_quantum_state_from_host_to_device");
QM_CXX_EXPORT void copy_quantum_state_from_host_to_device(void* state_gpu_copy, const void* state, ITYPE dim, void* stream, unsigned int device_number) { qmc_copy_quantum_state_from_host_to_device(state_gpu_copy, state, dim, stream, device_number); }

extern "C" void qmc_copy_quantum_state_from_cppstate_host(void* state_gpu_copy, const QmC* cppstate, ITYPE dim, void* stream, UINT device_number) __asm__("copy_quantum_state_from_cppstate_host");
QM_CXX_EXPORT void copy_quantum_state_from_cppstate_host(void* state_gpu_copy, const CPPCTYPE* cppstate, ITYPE dim, void* stream, UINT device_number) { qmc_copy_quantum_state_from_cppstate_host(state_gpu_copy, qc(cppstate), dim, stream, device_number); }

extern "C" void qmc_copy_quantum_state_from_device_to_host(void* state_cpu_copy, const void* state_gpu_original, ITYPE dim, void* stream, unsigned int device_number) __asm__("copy_quantum_state_from_device_to_host");
QM_CXX_EXPORT void copy_quantum_state_from_device_to_host(void* state_cpu_copy, const void* state_gpu_original, ITYPE dim, void* stream, unsigned int device_number) { qmc_copy_quantum_state_from_device_to_host(state_cpu_copy, state_gpu_original, dim, stream, device_number); }

extern "C" void qmc_get_quantum_state_host(void* state_gpu, void* psi_cpu_copy, ITYPE dim, void* stream, unsigned int device_number) __asm__("get_quantum_state_host");
QM_CXX_EXPORT void get_quantum_state_host(void* state_gpu, void* psi_cpu_copy, ITYPE dim, void* stream, unsigned int device_number) { qmc_get_quantum_state_host(state_gpu, psi_cpu_copy, dim, stream, device_number); }

extern "C" void qmc_print_quantum_state_host(void* state, ITYPE dim, unsigned int device_number) __asm__("print_quantum_state_host");
QM_CXX_EXPORT void print_quantum_state_host(void* state, ITYPE dim, unsigned int device_number) { qmc_print_quantum_state_host(state, dim, device_number); }
