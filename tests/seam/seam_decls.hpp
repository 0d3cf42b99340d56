// Test infrastructure: C++ declarations of part of the reference's GPU seam,
// restated with the reference's parameter types (src/gpusim/update_ops_cuda.h:10-140,
// stat_ops.h:7-62, memory_ops.h:7-20, util_func.h:9-26).  They have C++ linkage,
// so a program written against them references the same mangled names that
// cppsim's GPU build links against; seam_driver.cpp calls them.
#pragma once
#include <complex>

typedef unsigned int UINT;
typedef unsigned long long ITYPE;
typedef std::complex<double> CPPCTYPE;

// memory_ops.h
void* allocate_cuda_stream_host(unsigned int max_cuda_device, unsigned int device_number);
void release_cuda_stream_host(void* cuda_stream, unsigned int max_cuda_device, unsigned int device_number);
void* allocate_quantum_state_host(ITYPE dim, unsigned int device_number);
void initialize_quantum_state_host(void* state, ITYPE dim, void* stream, unsigned int device_number);
void release_quantum_state_host(void* state, unsigned int device_number);
// util_func.h
int get_num_device();
void get_quantum_state_host(void* state_gpu, void* psi_cpu_copy, ITYPE dim, void* stream, unsigned int device_number);
// update_ops_cuda.h
void H_gate_host(unsigned int target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);
void S_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);
void T_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);
void sqrtX_gate_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);
void CZ_gate_host(unsigned int control_qubit_index, unsigned int target_qubit_index, void* state, ITYPE dim,
                  void* stream, unsigned int device_number);
void CNOT_gate_host(unsigned int control_qubit_index, unsigned int target_qubit_index, void* state, ITYPE dim,
                    void* stream, unsigned int device_number);
void SWAP_gate_host(unsigned int target_qubit_index0, unsigned int target_qubit_index1, void* state, ITYPE dim,
                    void* stream, unsigned int device_number);
void RX_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream,
                  unsigned int device_number);
void RY_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream,
                  unsigned int device_number);
void RZ_gate_host(UINT target_qubit_index, double angle, void* state, ITYPE dim, void* stream,
                  unsigned int device_number);
void single_qubit_dense_matrix_gate_host(unsigned int target_qubit_index, const CPPCTYPE matrix[4], void* state,
                                         ITYPE dim, void* stream, unsigned int device_number);
void single_qubit_diagonal_matrix_gate_host(unsigned int target_qubit_index, const CPPCTYPE diagonal_matrix[2],
                                            void* state, ITYPE dim, void* stream, unsigned int device_number);
void single_qubit_control_single_qubit_dense_matrix_gate_host(unsigned int control_qubit_index,
                                                              unsigned int control_value,
                                                              unsigned int target_qubit_index,
                                                              const CPPCTYPE matrix[4], void* state, ITYPE dim,
                                                              void* stream, unsigned int device_number);
void double_qubit_dense_matrix_gate_host(unsigned int target0_qubit_index, unsigned int target1_qubit_index,
                                         const CPPCTYPE matrix[16], void* state, ITYPE dim, void* stream,
                                         unsigned int device_number);
void multi_qubit_dense_matrix_gate_host(const UINT* target_qubit_index_list, UINT target_qubit_index_count,
                                        const CPPCTYPE* matrix, void* state, ITYPE dim, void* stream,
                                        unsigned int device_number);
void multi_qubit_Pauli_rotation_gate_partial_list_host(const UINT* target_qubit_index_list,
                                                       const UINT* Pauli_operator_type_list,
                                                       UINT target_qubit_index_count, double angle, void* state,
                                                       ITYPE dim, void* stream, unsigned int device_number);
// stat_ops.h
double state_norm_squared_host(void* state, ITYPE dim, void* stream, unsigned int device_number);
double M0_prob_host(UINT target_qubit_index, void* state, ITYPE dim, void* stream, unsigned int device_number);
CPPCTYPE inner_product_host(const void* bra_state, const void* ket_state, ITYPE dim, void* stream,
                            unsigned int device_number);
double expectation_value_multi_qubit_Pauli_operator_partial_list_host(const UINT* target_qubit_index_list,
                                                                      const UINT* Pauli_operator_type_list,
                                                                      UINT target_qubit_index_count, void* state,
                                                                      ITYPE dim, void* stream,
                                                                      unsigned int device_number);
