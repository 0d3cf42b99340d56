// extern "C" entry points over the reference csim kernels, compiled together
// with the reference sources by oracle/Makefile into oracle/_ref/libcsim_ref.so.
//
// TEST INFRASTRUCTURE ONLY: this wraps the reference (compiled from
// src/csim/*.cpp where it lies, never copied) so tests and the CPU-baseline leg
// of bench.py can call it through ctypes.  The csim functions have C++ linkage
// (mangled names, std::complex arguments); these wrappers only re-type the
// pointers.  Declarations come from src/csim/update_ops.hpp, stat_ops.hpp,
// update_ops_dm.hpp, stat_ops_dm.hpp and memory_ops_dm.hpp.
#include <complex>
#include <vector>

#include "csim/memory_ops_dm.hpp"
#include "csim/stat_ops.hpp"
#include "csim/stat_ops_dm.hpp"
#include "csim/update_ops.hpp"
#include "csim/update_ops_dm.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {
inline CTYPE* C(double* p) { return reinterpret_cast<CTYPE*>(p); }
inline const CTYPE* C(const double* p) { return reinterpret_cast<const CTYPE*>(p); }
}  // namespace

#define REF_NAMED1(name)                                              \
    void csimref_##name(UINT t, double* s, ITYPE dim) { name(t, C(s), dim); }
#define REF_DM_NAMED1(name)                                           \
    void csimref_##name(UINT t, double* s, ITYPE dim) { name(t, C(s), dim); }

extern "C" {

int csimref_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void csimref_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

REF_NAMED1(X_gate)
REF_NAMED1(Y_gate)
REF_NAMED1(Z_gate)
REF_NAMED1(H_gate)
REF_NAMED1(S_gate)
REF_NAMED1(Sdag_gate)
REF_NAMED1(T_gate)
REF_NAMED1(Tdag_gate)
REF_NAMED1(sqrtX_gate)
REF_NAMED1(sqrtXdag_gate)
REF_NAMED1(sqrtY_gate)
REF_NAMED1(sqrtYdag_gate)
REF_NAMED1(P0_gate)
REF_NAMED1(P1_gate)

void csimref_RX_gate(UINT t, double a, double* s, ITYPE dim) { RX_gate(t, a, C(s), dim); }
void csimref_RY_gate(UINT t, double a, double* s, ITYPE dim) { RY_gate(t, a, C(s), dim); }
void csimref_RZ_gate(UINT t, double a, double* s, ITYPE dim) { RZ_gate(t, a, C(s), dim); }
void csimref_CNOT_gate(UINT c, UINT t, double* s, ITYPE dim) { CNOT_gate(c, t, C(s), dim); }
void csimref_CZ_gate(UINT c, UINT t, double* s, ITYPE dim) { CZ_gate(c, t, C(s), dim); }
void csimref_SWAP_gate(UINT a, UINT b, double* s, ITYPE dim) { SWAP_gate(a, b, C(s), dim); }

void csimref_single_qubit_dense_matrix_gate(UINT t, const double* m, double* s, ITYPE dim) {
    single_qubit_dense_matrix_gate(t, C(m), C(s), dim);
}
void csimref_double_qubit_dense_matrix_gate(UINT t0, UINT t1, const double* m, double* s, ITYPE dim) {
    double_qubit_dense_matrix_gate_c(t0, t1, C(m), C(s), dim);
}
void csimref_multi_qubit_dense_matrix_gate(const UINT* t, UINT n, const double* m, double* s, ITYPE dim) {
    multi_qubit_dense_matrix_gate(t, n, C(m), C(s), dim);
}
void csimref_single_qubit_control_single_qubit_dense_matrix_gate(UINT c, UINT cv, UINT t, const double* m,
                                                                 double* s, ITYPE dim) {
    single_qubit_control_single_qubit_dense_matrix_gate(c, cv, t, C(m), C(s), dim);
}
void csimref_multi_qubit_control_single_qubit_dense_matrix_gate(const UINT* c, const UINT* cv, UINT nc, UINT t,
                                                                const double* m, double* s, ITYPE dim) {
    multi_qubit_control_single_qubit_dense_matrix_gate(c, cv, nc, t, C(m), C(s), dim);
}
void csimref_single_qubit_control_multi_qubit_dense_matrix_gate(UINT c, UINT cv, const UINT* t, UINT nt,
                                                                const double* m, double* s, ITYPE dim) {
    single_qubit_control_multi_qubit_dense_matrix_gate(c, cv, t, nt, C(m), C(s), dim);
}
void csimref_multi_qubit_control_multi_qubit_dense_matrix_gate(const UINT* c, const UINT* cv, UINT nc,
                                                               const UINT* t, UINT nt, const double* m,
                                                               double* s, ITYPE dim) {
    multi_qubit_control_multi_qubit_dense_matrix_gate(c, cv, nc, t, nt, C(m), C(s), dim);
}
void csimref_single_qubit_diagonal_matrix_gate(UINT t, const double* d, double* s, ITYPE dim) {
    single_qubit_diagonal_matrix_gate(t, C(d), C(s), dim);
}
void csimref_single_qubit_phase_gate(UINT t, const double* p, double* s, ITYPE dim) {
    single_qubit_phase_gate(t, *C(p), C(s), dim);
}
void csimref_multi_qubit_diagonal_matrix_gate(const UINT* t, UINT n, const double* d, double* s, ITYPE dim) {
    multi_qubit_diagonal_matrix_gate(t, n, C(d), C(s), dim);
}
void csimref_multi_qubit_control_multi_qubit_diagonal_matrix_gate(const UINT* c, const UINT* cv, UINT nc,
                                                                  const UINT* t, UINT nt, const double* d,
                                                                  double* s, ITYPE dim) {
    multi_qubit_control_multi_qubit_diagonal_matrix_gate(c, cv, nc, t, nt, C(d), C(s), dim);
}
void csimref_multi_qubit_Pauli_gate_partial_list(const UINT* t, const UINT* p, UINT n, double* s, ITYPE dim) {
    multi_qubit_Pauli_gate_partial_list(t, p, n, C(s), dim);
}
void csimref_multi_qubit_Pauli_rotation_gate_partial_list(const UINT* t, const UINT* p, UINT n, double a,
                                                          double* s, ITYPE dim) {
    multi_qubit_Pauli_rotation_gate_partial_list(t, p, n, a, C(s), dim);
}
void csimref_normalize(double sq, double* s, ITYPE dim) { normalize(sq, C(s), dim); }
void csimref_state_add_with_coef(const double* coef, const double* a, double* s, ITYPE dim) {
    state_add_with_coef(*C(coef), C(a), C(s), dim);
}
void csimref_state_multiply(const double* coef, double* s, ITYPE dim) { state_multiply(*C(coef), C(s), dim); }

double csimref_state_norm_squared(const double* s, ITYPE dim) { return state_norm_squared(C(s), dim); }
void csimref_state_inner_product(const double* b, const double* k, ITYPE dim, double* out) {
    const CTYPE v = state_inner_product(C(b), C(k), dim);
    out[0] = v.real();
    out[1] = v.imag();
}
double csimref_M0_prob(UINT t, const double* s, ITYPE dim) { return M0_prob(t, C(s), dim); }
double csimref_M1_prob(UINT t, const double* s, ITYPE dim) { return M1_prob(t, C(s), dim); }
double csimref_marginal_prob(const UINT* t, const UINT* v, UINT n, const double* s, ITYPE dim) {
    return marginal_prob(t, v, n, C(s), dim);
}
double csimref_measurement_distribution_entropy(const double* s, ITYPE dim) {
    return measurement_distribution_entropy(C(s), dim);
}
double csimref_expectation_value_multi_qubit_Pauli_operator_partial_list(const UINT* t, const UINT* p, UINT n,
                                                                        const double* s, ITYPE dim) {
    return expectation_value_multi_qubit_Pauli_operator_partial_list(t, p, n, C(s), dim);
}
void csimref_transition_amplitude_multi_qubit_Pauli_operator_partial_list(const UINT* t, const UINT* p, UINT n,
                                                                         const double* bra, const double* ket,
                                                                         ITYPE dim, double* out) {
    const CTYPE v = transition_amplitude_multi_qubit_Pauli_operator_partial_list(t, p, n, C(bra), C(ket), dim);
    out[0] = v.real();
    out[1] = v.imag();
}

// density matrix
void csimref_dm_initialize_with_pure_state(double* rho, const double* psi, ITYPE dim) {
    dm_initialize_with_pure_state(C(rho), C(psi), dim);
}
void csimref_dm_single_qubit_dense_matrix_gate(UINT t, const double* m, double* s, ITYPE dim) {
    dm_single_qubit_dense_matrix_gate(t, C(m), C(s), dim);
}
void csimref_dm_multi_qubit_dense_matrix_gate(const UINT* t, UINT n, const double* m, double* s, ITYPE dim) {
    dm_multi_qubit_dense_matrix_gate(t, n, C(m), C(s), dim);
}
void csimref_dm_multi_qubit_control_single_qubit_dense_matrix_gate(const UINT* c, const UINT* cv, UINT nc, UINT t,
                                                                   const double* m, double* s, ITYPE dim) {
    dm_multi_qubit_control_single_qubit_dense_matrix_gate(c, cv, nc, t, C(m), C(s), dim);
}
void csimref_dm_multi_qubit_control_multi_qubit_dense_matrix_gate(const UINT* c, const UINT* cv, UINT nc,
                                                                  const UINT* t, UINT nt, const double* m,
                                                                  double* s, ITYPE dim) {
    dm_multi_qubit_control_multi_qubit_dense_matrix_gate(c, cv, nc, t, nt, C(m), C(s), dim);
}
REF_DM_NAMED1(dm_X_gate)
REF_DM_NAMED1(dm_Y_gate)
REF_DM_NAMED1(dm_Z_gate)
REF_DM_NAMED1(dm_H_gate)
REF_DM_NAMED1(dm_S_gate)
REF_DM_NAMED1(dm_T_gate)
REF_DM_NAMED1(dm_P0_gate)
REF_DM_NAMED1(dm_P1_gate)
void csimref_dm_CNOT_gate(UINT c, UINT t, double* s, ITYPE dim) { dm_CNOT_gate(c, t, C(s), dim); }
void csimref_dm_CZ_gate(UINT c, UINT t, double* s, ITYPE dim) { dm_CZ_gate(c, t, C(s), dim); }
void csimref_dm_SWAP_gate(UINT a, UINT b, double* s, ITYPE dim) { dm_SWAP_gate(a, b, C(s), dim); }
void csimref_dm_RX_gate(UINT t, double a, double* s, ITYPE dim) { dm_RX_gate(t, a, C(s), dim); }
void csimref_dm_RY_gate(UINT t, double a, double* s, ITYPE dim) { dm_RY_gate(t, a, C(s), dim); }
void csimref_dm_RZ_gate(UINT t, double a, double* s, ITYPE dim) { dm_RZ_gate(t, a, C(s), dim); }
void csimref_dm_multi_qubit_Pauli_gate_partial_list(const UINT* t, const UINT* p, UINT n, double* s, ITYPE dim) {
    dm_multi_qubit_Pauli_gate_partial_list(t, p, n, C(s), dim);
}
void csimref_dm_multi_qubit_Pauli_rotation_gate_partial_list(const UINT* t, const UINT* p, UINT n, double a,
                                                             double* s, ITYPE dim) {
    dm_multi_qubit_Pauli_rotation_gate_partial_list(t, p, n, a, C(s), dim);
}
void csimref_dm_state_add_with_coef(const double* coef, const double* a, double* s, ITYPE dim) {
    dm_state_add_with_coef(*C(coef), C(a), C(s), dim);
}
void csimref_dm_state_multiply(const double* coef, double* s, ITYPE dim) { dm_state_multiply(*C(coef), C(s), dim); }
double csimref_dm_state_norm_squared(const double* s, ITYPE dim) { return dm_state_norm_squared(C(s), dim); }
double csimref_dm_M0_prob(UINT t, const double* s, ITYPE dim) { return dm_M0_prob(t, C(s), dim); }
double csimref_dm_marginal_prob(const UINT* t, const UINT* v, UINT n, const double* s, ITYPE dim) {
    return dm_marginal_prob(t, v, n, C(s), dim);
}
double csimref_dm_expectation_value_multi_qubit_Pauli_operator_partial_list(const UINT* t, const UINT* p, UINT n,
                                                                           const double* s, ITYPE dim) {
    return dm_expectation_value_multi_qubit_Pauli_operator_partial_list(t, p, n, C(s), dim);
}

}  // extern "C"
