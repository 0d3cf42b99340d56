// CPU oracle for the qulacs state-vector / density-matrix hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, loop for loop, the csim kernels
// of qulacs (src/csim/*.cpp) and the cppsim sampling / Kraus-mixture loops, so
// that the HIP library can be checked against it.  It is loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, and by nothing
// else: the product path (qulacs_amd/) never links or calls it.
//
// Conventions (all from the reference):
//   * state vector: std::complex<double>[dim], qubit q <-> index bit q
//     (src/csim/type.hpp:18-29, test/util/util.hpp:84-93)
//   * dense k-qubit matrix: row-major 2^k x 2^k, bit i of the row/column index
//     belongs to target_list[i] (src/csim/utility.cpp:77-93)
//   * density matrix: row-major rho[y*dim + x] (src/csim/update_ops_dm.cpp:73-76)
//   * rotations are exp(+i theta P / 2) (src/csim/update_ops_pauli_single.cpp:54-116)
//
// Every complex array crosses the C ABI as interleaved doubles (re, im).
//
// Pinning status of each function (what checks it, besides the restated loop):
//   gates, stats, dm ops   -> golden vectors written by the compiled reference
//                              csim (tests/golden/) and closed forms in tests/
//   oracle_sampling        -> known answers (basis state, monotone histogram,
//                              test/cppsim/test_state.cpp:38-81) and brute-force
//                              lower_bound over an exact rational CDF
//   oracle_dm_probabilistic-> closed form (1-4p/3) rho + (2p/3) I (x) tr_q rho
//                              (test/cppsim/test_noise_dm.cpp:274-309)
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using CTYPE = std::complex<double>;
using ITYPE = unsigned long long;
using UINT = unsigned int;

namespace {

// src/csim/utility.hpp:14-18
inline ITYPE insert_zero(ITYPE basis_index, UINT qubit_index) {
    const ITYPE mask = 1ULL << qubit_index;
    return ((basis_index >> qubit_index) << (qubit_index + 1)) + (basis_index & (mask - 1));
}

// src/csim/utility.hpp:60-68 (population count parity)
inline UINT parity(ITYPE x) { return (UINT)(__builtin_popcountll(x) & 1); }

// src/csim/constant.cpp:36-37
const CTYPE PHASE_90ROT[4] = {CTYPE(1, 0), CTYPE(0, 1), CTYPE(-1, 0), CTYPE(0, -1)};
const CTYPE PHASE_M90ROT[4] = {CTYPE(1, 0), CTYPE(0, -1), CTYPE(-1, 0), CTYPE(0, 1)};

// src/csim/utility.cpp:12-43
void pauli_masks(const UINT* targets, const UINT* paulis, UINT count, ITYPE* bit_flip,
                 ITYPE* phase_flip, UINT* rot90, UINT* pivot) {
    *bit_flip = 0; *phase_flip = 0; *rot90 = 0; *pivot = 0;
    for (UINT c = 0; c < count; ++c) {
        const UINT t = targets[c];
        switch (paulis[c]) {
            case 0: break;
            case 1: *bit_flip ^= 1ULL << t; *pivot = t; break;
            case 2: *bit_flip ^= 1ULL << t; *phase_flip ^= 1ULL << t; (*rot90)++; *pivot = t; break;
            case 3: *phase_flip ^= 1ULL << t; break;
            default: abort();
        }
    }
}

// src/csim/utility.cpp:77-93
std::vector<ITYPE> matrix_mask_list(const UINT* targets, UINT count) {
    const ITYPE mdim = 1ULL << count;
    std::vector<ITYPE> masks(mdim, 0);
    for (ITYPE c = 0; c < mdim; ++c)
        for (UINT b = 0; b < count; ++b)
            if ((c >> b) & 1) masks[c] ^= 1ULL << targets[b];
    return masks;
}

// src/csim/utility.cpp:95-103 create_control_mask
ITYPE control_mask(const UINT* idx, const UINT* val, UINT n) {
    ITYPE m = 0;
    for (UINT i = 0; i < n; ++i) m ^= (1ULL << idx[i]) * val[i];
    return m;
}

std::vector<UINT> sorted_union(const UINT* a, UINT na, const UINT* b, UINT nb) {
    std::vector<UINT> v(a, a + na);
    v.insert(v.end(), b, b + nb);
    std::sort(v.begin(), v.end());
    return v;
}

inline CTYPE* C(double* p) { return reinterpret_cast<CTYPE*>(p); }
inline const CTYPE* C(const double* p) { return reinterpret_cast<const CTYPE*>(p); }

}  // namespace

extern "C" {

int oracle_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// ---------------------------------------------------------------- memory / init
// Fills a freshly allocated buffer with one value using the same static
// schedule as the kernels' parallel loops, so each page is first touched (and
// placed) by the thread that later works on it.
void oracle_parallel_fill(double* state, ITYPE dim, double re, double im) {
    CTYPE* s = C(state);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)dim; ++i) s[i] = CTYPE(re, im);
}

// src/csim/init_ops_fill.cpp:10-23
void oracle_initialize_quantum_state(double* state, ITYPE dim) {
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)dim; ++i) s[i] = 0;
    s[0] = 1.0;
}

// src/cppsim/state.hpp:494-515 set_computational_basis
void oracle_set_computational_basis(ITYPE basis, double* state, ITYPE dim) {
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)dim; ++i) s[i] = 0;
    s[basis] = 1.0;
}

// ---------------------------------------------------------------- dense gates
// src/csim/update_ops_matrix_dense_single.cpp:44-68
void oracle_single_qubit_dense_matrix_gate(UINT t, const double* matrix, double* state, ITYPE dim) {
    const CTYPE* m = C(matrix);
    CTYPE* s = C(state);
    const ITYPE loop_dim = dim / 2, mask = 1ULL << t;
#pragma omp parallel for
    for (long long i = 0; i < (long long)loop_dim; ++i) {
        const ITYPE b0 = insert_zero((ITYPE)i, t), b1 = b0 + mask;
        const CTYPE v0 = s[b0], v1 = s[b1];
        s[b0] = m[0] * v0 + m[1] * v1;
        s[b1] = m[2] * v0 + m[3] * v1;
    }
}

// src/csim/update_ops_matrix_dense_double.cpp:64-113 (matrix bit0 <-> t1, bit1 <-> t2)
void oracle_double_qubit_dense_matrix_gate(UINT t1, UINT t2, const double* matrix, double* state, ITYPE dim) {
    const CTYPE* m = C(matrix);
    CTYPE* s = C(state);
    const UINT lo = std::min(t1, t2), hi = std::max(t1, t2);
    const ITYPE m1 = 1ULL << t1, m2 = 1ULL << t2;
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 4); ++i) {
        const ITYPE b0 = insert_zero(insert_zero((ITYPE)i, lo), hi);
        const ITYPE b1 = b0 + m1, b2 = b0 + m2, b3 = b1 + m2;
        const CTYPE v0 = s[b0], v1 = s[b1], v2 = s[b2], v3 = s[b3];
        s[b0] = m[0] * v0 + m[1] * v1 + m[2] * v2 + m[3] * v3;
        s[b1] = m[4] * v0 + m[5] * v1 + m[6] * v2 + m[7] * v3;
        s[b2] = m[8] * v0 + m[9] * v1 + m[10] * v2 + m[11] * v3;
        s[b3] = m[12] * v0 + m[13] * v1 + m[14] * v2 + m[15] * v3;
    }
}

// src/csim/update_ops_control_multi_target_multi.cpp:18-79 (general form; the
// uncontrolled case is src/csim/update_ops_matrix_dense_multi.cpp:59-128)
void oracle_multi_qubit_control_multi_qubit_dense_matrix_gate(
    const UINT* ctrl, const UINT* cval, UINT nc, const UINT* targets, UINT nt,
    const double* matrix, double* state, ITYPE dim) {
    const CTYPE* m = C(matrix);
    CTYPE* s = C(state);
    const ITYPE mdim = 1ULL << nt;
    const std::vector<ITYPE> masks = matrix_mask_list(targets, nt);
    const std::vector<UINT> ins = sorted_union(targets, nt, ctrl, nc);
    const ITYPE cmask = control_mask(ctrl, cval, nc);
    const ITYPE loop_dim = dim >> (nt + nc);
#pragma omp parallel
    {
        std::vector<CTYPE> buf(mdim);
#pragma omp for
        for (long long i = 0; i < (long long)loop_dim; ++i) {
            ITYPE b0 = (ITYPE)i;
            for (UINT q : ins) b0 = insert_zero(b0, q);
            b0 ^= cmask;
            for (ITYPE y = 0; y < mdim; ++y) {
                CTYPE acc = 0;
                for (ITYPE x = 0; x < mdim; ++x) acc += m[y * mdim + x] * s[b0 ^ masks[x]];
                buf[y] = acc;
            }
            for (ITYPE y = 0; y < mdim; ++y) s[b0 ^ masks[y]] = buf[y];
        }
    }
}

// src/csim/update_ops_matrix_dense_multi.cpp:29-48 (dispatch) and :59-128
void oracle_multi_qubit_dense_matrix_gate(const UINT* targets, UINT nt, const double* matrix,
                                          double* state, ITYPE dim) {
    if (nt == 1) { oracle_single_qubit_dense_matrix_gate(targets[0], matrix, state, dim); return; }
    if (nt == 2) { oracle_double_qubit_dense_matrix_gate(targets[0], targets[1], matrix, state, dim); return; }
    oracle_multi_qubit_control_multi_qubit_dense_matrix_gate(nullptr, nullptr, 0, targets, nt, matrix, state, dim);
}

// ---------------------------------------------------------------- named gates
// src/csim/update_ops_named_CNOT.cpp:42-101
void oracle_CNOT_gate(UINT c, UINT t, double* state, ITYPE dim) {
    CTYPE* s = C(state);
    const UINT lo = std::min(c, t), hi = std::max(c, t);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 4); ++i) {
        const ITYPE b0 = insert_zero(insert_zero((ITYPE)i, lo), hi) + (1ULL << c);
        const ITYPE b1 = b0 + (1ULL << t);
        std::swap(s[b0], s[b1]);
    }
}
// src/csim/update_ops_named_CZ.cpp:35-75
void oracle_CZ_gate(UINT c, UINT t, double* state, ITYPE dim) {
    CTYPE* s = C(state);
    const UINT lo = std::min(c, t), hi = std::max(c, t);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 4); ++i) {
        const ITYPE b = insert_zero(insert_zero((ITYPE)i, lo), hi) + (1ULL << c) + (1ULL << t);
        s[b] *= -1;
    }
}
// src/csim/update_ops_named_SWAP.cpp:42-90
void oracle_SWAP_gate(UINT t0, UINT t1, double* state, ITYPE dim) {
    CTYPE* s = C(state);
    const UINT lo = std::min(t0, t1), hi = std::max(t0, t1);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 4); ++i) {
        const ITYPE b = insert_zero(insert_zero((ITYPE)i, lo), hi);
        std::swap(s[b + (1ULL << t0)], s[b + (1ULL << t1)]);
    }
}
// src/csim/update_ops_named_X.cpp (X swaps the pair)
void oracle_X_gate(UINT t, double* state, ITYPE dim) {
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 2); ++i) {
        const ITYPE b0 = insert_zero((ITYPE)i, t);
        std::swap(s[b0], s[b0 + (1ULL << t)]);
    }
}
// src/csim/update_ops_named_Y.cpp: (a0, a1) -> (-i a1, i a0)
void oracle_Y_gate(UINT t, double* state, ITYPE dim) {
    CTYPE* s = C(state);
    const CTYPE I(0, 1);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 2); ++i) {
        const ITYPE b0 = insert_zero((ITYPE)i, t), b1 = b0 + (1ULL << t);
        const CTYPE v0 = s[b0];
        s[b0] = -I * s[b1];
        s[b1] = I * v0;
    }
}
// src/csim/update_ops_named_Z.cpp: negate the bit-1 half
void oracle_Z_gate(UINT t, double* state, ITYPE dim) {
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 2); ++i) s[insert_zero((ITYPE)i, t) + (1ULL << t)] *= -1;
}
// src/csim/update_ops_named_H.cpp: (a0, a1) -> ((a0+a1)/sqrt2, (a0-a1)/sqrt2)
void oracle_H_gate(UINT t, double* state, ITYPE dim) {
    CTYPE* s = C(state);
    const double r = 1.0 / std::sqrt(2.0);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 2); ++i) {
        const ITYPE b0 = insert_zero((ITYPE)i, t), b1 = b0 + (1ULL << t);
        const CTYPE v0 = s[b0], v1 = s[b1];
        s[b0] = (v0 + v1) * r;
        s[b1] = (v0 - v1) * r;
    }
}
// src/csim/update_ops_named_projection.cpp:38-72: P0 zeroes the bit-1 half, P1 the bit-0 half
void oracle_P0_gate(UINT t, double* state, ITYPE dim) {
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 2); ++i) s[insert_zero((ITYPE)i, t) + (1ULL << t)] = 0;
}
void oracle_P1_gate(UINT t, double* state, ITYPE dim) {
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 2); ++i) s[insert_zero((ITYPE)i, t)] = 0;
}

// ---------------------------------------------------------------- diagonal / phase
// src/csim/update_ops_matrix_diagonal_single.cpp:44-66
void oracle_single_qubit_diagonal_matrix_gate(UINT t, const double* diag, double* state, ITYPE dim) {
    const CTYPE* d = C(diag);
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)dim; ++i) s[i] *= d[((ITYPE)i >> t) & 1];
}
// src/csim/update_ops_matrix_phase_single.cpp:14-60
void oracle_single_qubit_phase_gate(UINT t, const double* phase, double* state, ITYPE dim) {
    const CTYPE p = *C(phase);
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 2); ++i) s[insert_zero((ITYPE)i, t) + (1ULL << t)] *= p;
}
// src/csim/update_ops_matrix_diagonal_multi.cpp:60-112 (uncontrolled form :22-58)
void oracle_multi_qubit_control_multi_qubit_diagonal_matrix_gate(
    const UINT* ctrl, const UINT* cval, UINT nc, const UINT* targets, UINT nt,
    const double* diag, double* state, ITYPE dim) {
    const CTYPE* d = C(diag);
    CTYPE* s = C(state);
    const ITYPE mdim = 1ULL << nt;
    const std::vector<ITYPE> masks = matrix_mask_list(targets, nt);
    const std::vector<UINT> ins = sorted_union(targets, nt, ctrl, nc);
    const ITYPE cmask = control_mask(ctrl, cval, nc);
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim >> (nt + nc)); ++i) {
        ITYPE b0 = (ITYPE)i;
        for (UINT q : ins) b0 = insert_zero(b0, q);
        b0 ^= cmask;
        for (ITYPE y = 0; y < mdim; ++y) s[b0 ^ masks[y]] *= d[y];
    }
}
// src/gpusim/update_ops_multi.cu:1610-1685 (whole-state diagonal, length dim)
void oracle_whole_state_diagonal_gate(const double* diag, double* state, ITYPE dim) {
    const CTYPE* d = C(diag);
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)dim; ++i) s[i] *= d[i];
}

// ---------------------------------------------------------------- Pauli gates
// src/csim/update_ops_pauli_multi.cpp:51-90 (XZ mask) and :141-163 (Z mask), dispatch :192-209
void oracle_multi_qubit_Pauli_gate_partial_list(const UINT* targets, const UINT* paulis, UINT n,
                                                double* state, ITYPE dim) {
    CTYPE* s = C(state);
    ITYPE bf, pf; UINT rot, pivot;
    pauli_masks(targets, paulis, n, &bf, &pf, &rot, &pivot);
    if (bf == 0) {
#pragma omp parallel for
        for (long long i = 0; i < (long long)dim; ++i)
            if (parity((ITYPE)i & pf)) s[i] *= -1;
        return;
    }
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 2); ++i) {
        const ITYPE b0 = insert_zero((ITYPE)i, pivot), b1 = b0 ^ bf;
        const UINT s0 = parity(b0 & pf), s1 = parity(b1 & pf);
        const CTYPE v0 = s[b0], v1 = s[b1];
        s[b0] = v1 * PHASE_M90ROT[(rot + s0 * 2) % 4];
        s[b1] = v0 * PHASE_M90ROT[(rot + s1 * 2) % 4];
    }
}

// src/csim/update_ops_pauli_multi.cpp:92-139 (XZ mask), :165-190 (Z mask), dispatch :229-247
void oracle_multi_qubit_Pauli_rotation_gate_partial_list(const UINT* targets, const UINT* paulis,
                                                         UINT n, double angle, double* state, ITYPE dim) {
    CTYPE* s = C(state);
    ITYPE bf, pf; UINT rot, pivot;
    pauli_masks(targets, paulis, n, &bf, &pf, &rot, &pivot);
    const double c = std::cos(angle / 2), sn = std::sin(angle / 2);
    const CTYPE I(0, 1);
    if (bf == 0) {
#pragma omp parallel for
        for (long long i = 0; i < (long long)dim; ++i) {
            const int sign = 1 - 2 * (int)parity((ITYPE)i & pf);
            s[i] *= CTYPE(c, sign * sn);
        }
        return;
    }
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim / 2); ++i) {
        const ITYPE b0 = insert_zero((ITYPE)i, pivot), b1 = b0 ^ bf;
        const UINT p0 = parity(b0 & pf), p1 = parity(b1 & pf);
        const CTYPE v0 = s[b0], v1 = s[b1];
        s[b0] = c * v0 + I * sn * v1 * PHASE_M90ROT[(rot + p0 * 2) % 4];
        s[b1] = c * v1 + I * sn * v0 * PHASE_M90ROT[(rot + p1 * 2) % 4];
    }
}

// ---------------------------------------------------------------- state arithmetic
// src/csim/update_ops_named_state.cpp:21-35
void oracle_normalize(double squared_norm, double* state, ITYPE dim) {
    CTYPE* s = C(state);
    const double f = std::sqrt(1. / squared_norm);
#pragma omp parallel for
    for (long long i = 0; i < (long long)dim; ++i) s[i] *= f;
}
// src/csim/update_ops_named_state.cpp:62-98
void oracle_state_add_with_coef(const double* coef, const double* added, double* state, ITYPE dim) {
    const CTYPE cf = *C(coef);
    const CTYPE* a = C(added);
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)dim; ++i) s[i] += cf * a[i];
}
void oracle_state_multiply(const double* coef, double* state, ITYPE dim) {
    const CTYPE cf = *C(coef);
    CTYPE* s = C(state);
#pragma omp parallel for
    for (long long i = 0; i < (long long)dim; ++i) s[i] *= cf;
}

// ---------------------------------------------------------------- reductions
// src/csim/stat_ops.cpp:13-29
double oracle_state_norm_squared(const double* state, ITYPE dim) {
    const CTYPE* s = C(state);
    double norm = 0;
#pragma omp parallel for reduction(+ : norm)
    for (long long i = 0; i < (long long)dim; ++i) norm += std::norm(s[i]);
    return norm;
}
// src/csim/stat_ops.cpp:65-86
void oracle_state_inner_product(const double* bra, const double* ket, ITYPE dim, double* out) {
    const CTYPE* b = C(bra);
    const CTYPE* k = C(ket);
    double re = 0, im = 0;
#pragma omp parallel for reduction(+ : re, im)
    for (long long i = 0; i < (long long)dim; ++i) {
        const CTYPE v = std::conj(b[i]) * k[i];
        re += v.real(); im += v.imag();
    }
    out[0] = re; out[1] = im;
}
// src/csim/stat_ops_probability.cpp:11-54
double oracle_M0_prob(UINT t, const double* state, ITYPE dim) {
    const CTYPE* s = C(state);
    double sum = 0;
#pragma omp parallel for reduction(+ : sum)
    for (long long i = 0; i < (long long)(dim / 2); ++i) sum += std::norm(s[insert_zero((ITYPE)i, t)]);
    return sum;
}
double oracle_M1_prob(UINT t, const double* state, ITYPE dim) {
    const CTYPE* s = C(state);
    double sum = 0;
#pragma omp parallel for reduction(+ : sum)
    for (long long i = 0; i < (long long)(dim / 2); ++i) sum += std::norm(s[insert_zero((ITYPE)i, t) ^ (1ULL << t)]);
    return sum;
}
// src/csim/stat_ops_probability.cpp:56-80
double oracle_marginal_prob(const UINT* sorted_targets, const UINT* values, UINT n, const double* state, ITYPE dim) {
    const CTYPE* s = C(state);
    double sum = 0;
#pragma omp parallel for reduction(+ : sum)
    for (long long i = 0; i < (long long)(dim >> n); ++i) {
        ITYPE b = (ITYPE)i;
        for (UINT c = 0; c < n; ++c) {
            b = insert_zero(b, sorted_targets[c]);
            b ^= (1ULL << sorted_targets[c]) * values[c];
        }
        sum += std::norm(s[b]);
    }
    return sum;
}
// src/csim/stat_ops_probability.cpp:82-100
double oracle_measurement_distribution_entropy(const double* state, ITYPE dim) {
    const CTYPE* s = C(state);
    double ent = 0;
    const double eps = 1e-15;
#pragma omp parallel for reduction(+ : ent)
    for (long long i = 0; i < (long long)dim; ++i) {
        double p = std::norm(s[i]);
        p = p > eps ? p : eps;
        ent += -1.0 * p * std::log(p);
    }
    return ent;
}
// src/csim/stat_ops_expectation_value.cpp:352-373 (XZ), :467-482 (Z), dispatch :539-588
double oracle_expectation_value_multi_qubit_Pauli_operator_partial_list(
    const UINT* targets, const UINT* paulis, UINT n, const double* state, ITYPE dim) {
    const CTYPE* s = C(state);
    ITYPE bf, pf; UINT rot, pivot;
    pauli_masks(targets, paulis, n, &bf, &pf, &rot, &pivot);
    double sum = 0;
    if (bf == 0) {
#pragma omp parallel for reduction(+ : sum)
        for (long long i = 0; i < (long long)dim; ++i) {
            const int sign = 1 - 2 * (int)parity((ITYPE)i & pf);
            sum += std::norm(s[i]) * sign;
        }
        return sum;
    }
#pragma omp parallel for reduction(+ : sum)
    for (long long i = 0; i < (long long)(dim / 2); ++i) {
        const ITYPE b0 = insert_zero((ITYPE)i, pivot), b1 = b0 ^ bf;
        const UINT s0 = parity(b0 & pf);
        sum += std::real(s[b0] * std::conj(s[b1]) * PHASE_90ROT[(rot + s0 * 2) % 4] * 2.0);
    }
    return sum;
}
// src/csim/stat_ops_transition_amplitude.cpp:19-105 ; result <bra| P |ket>
void oracle_transition_amplitude_multi_qubit_Pauli_operator_partial_list(
    const UINT* targets, const UINT* paulis, UINT n, const double* bra, const double* ket,
    ITYPE dim, double* out) {
    const CTYPE* b = C(bra);
    const CTYPE* k = C(ket);
    ITYPE bf, pf; UINT rot, pivot;
    pauli_masks(targets, paulis, n, &bf, &pf, &rot, &pivot);
    double re = 0, im = 0;
    if (bf == 0) {
#pragma omp parallel for reduction(+ : re, im)
        for (long long i = 0; i < (long long)dim; ++i) {
            const double sign = 1 - 2 * (int)parity((ITYPE)i & pf);
            const CTYPE v = sign * k[i] * std::conj(b[i]);
            re += v.real(); im += v.imag();
        }
    } else {
#pragma omp parallel for reduction(+ : re, im)
        for (long long i = 0; i < (long long)(dim / 2); ++i) {
            const ITYPE b0 = insert_zero((ITYPE)i, pivot), b1 = b0 ^ bf;
            const UINT s0 = parity(b0 & pf), s1 = parity(b1 & pf);
            const CTYPE v1 = k[b0] * std::conj(b[b1]) * PHASE_90ROT[(rot + s0 * 2) % 4];
            const CTYPE v2 = k[b1] * std::conj(b[b0]) * PHASE_90ROT[(rot + s1 * 2) % 4];
            re += v1.real() + v2.real();
            im += v1.imag() + v2.imag();
        }
    }
    out[0] = re; out[1] = im;
}

// ---------------------------------------------------------------- sampling
// src/cppsim/state.hpp:909-930: stacked_prob[0] = 0, stacked_prob[i+1] =
// stacked_prob[i] + |psi_i|^2 summed sequentially (not normalised), and each
// sample is lower_bound(stacked_prob, u) - 1.  The per-amplitude probability is
// fixed here to fma(re, re, im*im) so that the result is a function of the
// amplitudes alone (GCC may or may not contract std::norm's x*x + y*y).  The
// uniforms are inputs; the reference draws them from std::mt19937_64 through
// uniform_real_distribution<double>(0, 1) (src/cppsim/utility.hpp:54-102).
void oracle_sampling(const double* state, ITYPE dim, const double* uniforms, UINT count,
                     unsigned long long* out) {
    std::vector<double> cdf(dim + 1);
    double sum = 0.;
    cdf[0] = 0.;
    for (ITYPE i = 0; i < dim; ++i) {
        const double re = state[2 * i], im = state[2 * i + 1];
        sum += std::fma(re, re, im * im);
        cdf[i + 1] = sum;
    }
    for (UINT c = 0; c < count; ++c) {
        auto it = std::lower_bound(cdf.begin(), cdf.end(), uniforms[c]);
        out[c] = (unsigned long long)(std::distance(cdf.begin(), it) - 1);
    }
}

// The reference's Random (src/cppsim/utility.hpp:54-102): std::mt19937_64
// seeded with `seed`, uniforms from uniform_real_distribution<double>(0, 1) of
// the same C++ library.  The raw stream is pinned by the C++ standard's check
// value (10000th output of the default seed 5489 is 9981545732273789042).
void oracle_mt64_uniforms(unsigned long long seed, UINT count, double* out) {
    std::mt19937_64 mt(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (UINT c = 0; c < count; ++c) out[c] = u(mt);
}
void oracle_mt64_raw(unsigned long long seed, UINT count, unsigned long long* out) {
    std::mt19937_64 mt(seed);
    for (UINT c = 0; c < count; ++c) out[c] = mt();
}

// src/cppsim/state_dm.hpp:383-408: same loop over |rho_ii| (std::abs of the
// complex diagonal element, which libstdc++ evaluates with libm's cabs/hypot).
void oracle_dm_sampling_cdf(const double* rho, ITYPE dim, double* cdf) {
    const CTYPE* r = C(rho);
    double sum = 0.;
    cdf[0] = 0.;
    for (ITYPE i = 0; i < dim; ++i) {
        sum += std::abs(r[i * dim + i]);
        cdf[i + 1] = sum;
    }
}
void oracle_dm_sampling(const double* rho, ITYPE dim, const double* uniforms, UINT count,
                        unsigned long long* out) {
    std::vector<double> cdf(dim + 1);
    oracle_dm_sampling_cdf(rho, dim, cdf.data());
    for (UINT c = 0; c < count; ++c) {
        auto it = std::lower_bound(cdf.begin(), cdf.end(), uniforms[c]);
        out[c] = (unsigned long long)(std::distance(cdf.begin(), it) - 1);
    }
}

// ---------------------------------------------------------------- density matrix
// src/csim/memory_ops_dm.cpp:39-51
void oracle_dm_initialize_with_pure_state(double* rho, const double* pure, ITYPE dim) {
    CTYPE* r = C(rho);
    const CTYPE* p = C(pure);
#pragma omp parallel for
    for (long long y = 0; y < (long long)dim; ++y)
        for (ITYPE x = 0; x < dim; ++x) r[y * dim + x] = p[y] * std::conj(p[x]);
}

// src/csim/update_ops_dm.cpp:33-102
void oracle_dm_single_qubit_dense_matrix_gate(UINT t, const double* matrix, double* rho, ITYPE dim) {
    const CTYPE* m = C(matrix);
    CTYPE* s = C(rho);
    CTYPE e[16];
    for (int y = 0; y < 4; ++y)
        for (int x = 0; x < 4; ++x) e[y * 4 + x] = m[(y / 2) * 2 + x / 2] * std::conj(m[(y % 2) * 2 + x % 2]);
    const ITYPE mask = 1ULL << t;
#pragma omp parallel for
    for (long long iy = 0; iy < (long long)(dim / 2); ++iy) {
        const ITYPE y0 = insert_zero((ITYPE)iy, t), y1 = y0 ^ mask;
        for (ITYPE ix = 0; ix < dim / 2; ++ix) {
            const ITYPE x0 = insert_zero(ix, t), x1 = x0 ^ mask;
            const ITYPE i00 = y0 * dim + x0, i01 = y0 * dim + x1, i10 = y1 * dim + x0, i11 = y1 * dim + x1;
            const CTYPE v00 = s[i00], v01 = s[i01], v10 = s[i10], v11 = s[i11];
            s[i00] = e[0] * v00 + e[1] * v01 + e[2] * v10 + e[3] * v11;
            s[i01] = e[4] * v00 + e[5] * v01 + e[6] * v10 + e[7] * v11;
            s[i10] = e[8] * v00 + e[9] * v01 + e[10] * v10 + e[11] * v11;
            s[i11] = e[12] * v00 + e[13] * v01 + e[14] * v10 + e[15] * v11;
        }
    }
}

// src/csim/update_ops_dm.cpp:500-688 (two passes: rows by U, then columns by U^dagger)
void oracle_dm_multi_qubit_control_multi_qubit_dense_matrix_gate(
    const UINT* ctrl, const UINT* cval, UINT nc, const UINT* targets, UINT nt,
    const double* matrix, double* rho, ITYPE dim) {
    const CTYPE* m = C(matrix);
    CTYPE* s = C(rho);
    const ITYPE mdim = 1ULL << nt;
    const std::vector<ITYPE> masks = matrix_mask_list(targets, nt);
    const std::vector<UINT> ins = sorted_union(targets, nt, ctrl, nc);
    const ITYPE cmask = control_mask(ctrl, cval, nc);
    const ITYPE loop_dim = dim >> (nt + nc);
    std::vector<CTYPE> adj(mdim * mdim);
    for (ITYPE y = 0; y < mdim; ++y)
        for (ITYPE x = 0; x < mdim; ++x) adj[y * mdim + x] = std::conj(m[x * mdim + y]);
#pragma omp parallel
    {
        std::vector<CTYPE> buf(mdim);
#pragma omp for
        for (long long x = 0; x < (long long)dim; ++x)
            for (ITYPE iy = 0; iy < loop_dim; ++iy) {
                ITYPE b0 = iy;
                for (UINT q : ins) b0 = insert_zero(b0, q);
                b0 ^= cmask;
                for (ITYPE y = 0; y < mdim; ++y) {
                    CTYPE acc = 0;
                    for (ITYPE k = 0; k < mdim; ++k) acc += m[y * mdim + k] * s[(b0 ^ masks[k]) * dim + x];
                    buf[y] = acc;
                }
                for (ITYPE y = 0; y < mdim; ++y) s[(b0 ^ masks[y]) * dim + x] = buf[y];
            }
#pragma omp barrier
#pragma omp for
        for (long long y = 0; y < (long long)dim; ++y)
            for (ITYPE ix = 0; ix < loop_dim; ++ix) {
                ITYPE b0 = ix;
                for (UINT q : ins) b0 = insert_zero(b0, q);
                b0 ^= cmask;
                for (ITYPE x = 0; x < mdim; ++x) {
                    CTYPE acc = 0;
                    for (ITYPE k = 0; k < mdim; ++k) acc += s[y * dim + (b0 ^ masks[k])] * adj[k * mdim + x];
                    buf[x] = acc;
                }
                for (ITYPE x = 0; x < mdim; ++x) s[y * dim + (b0 ^ masks[x])] = buf[x];
            }
    }
}

// src/csim/update_ops_dm.cpp:342-498 (per block: B = U rho_blk, rho_blk = B U^dagger)
void oracle_dm_multi_qubit_dense_matrix_gate(const UINT* targets, UINT nt, const double* matrix,
                                             double* rho, ITYPE dim) {
    const CTYPE* m = C(matrix);
    CTYPE* s = C(rho);
    const ITYPE mdim = 1ULL << nt;
    const std::vector<ITYPE> masks = matrix_mask_list(targets, nt);
    std::vector<UINT> ins(targets, targets + nt);
    std::sort(ins.begin(), ins.end());
    std::vector<CTYPE> adj(mdim * mdim);
    for (ITYPE y = 0; y < mdim; ++y)
        for (ITYPE x = 0; x < mdim; ++x) adj[y * mdim + x] = std::conj(m[x * mdim + y]);
    const ITYPE loop_dim = dim >> nt;
#pragma omp parallel
    {
        std::vector<CTYPE> buf(mdim * mdim);
#pragma omp for
        for (long long iy = 0; iy < (long long)loop_dim; ++iy) {
            ITYPE by = (ITYPE)iy;
            for (UINT q : ins) by = insert_zero(by, q);
            for (ITYPE ix = 0; ix < loop_dim; ++ix) {
                ITYPE bx = ix;
                for (UINT q : ins) bx = insert_zero(bx, q);
                for (ITYPE y = 0; y < mdim; ++y)
                    for (ITYPE x = 0; x < mdim; ++x) {
                        CTYPE acc = 0;
                        for (ITYPE k = 0; k < mdim; ++k)
                            acc += m[y * mdim + k] * s[(by ^ masks[k]) * dim + (bx ^ masks[x])];
                        buf[y * mdim + x] = acc;
                    }
                for (ITYPE y = 0; y < mdim; ++y)
                    for (ITYPE x = 0; x < mdim; ++x) {
                        CTYPE acc = 0;
                        for (ITYPE k = 0; k < mdim; ++k) acc += buf[y * mdim + k] * adj[k * mdim + x];
                        s[(by ^ masks[y]) * dim + (bx ^ masks[x])] = acc;
                    }
            }
        }
    }
}

// src/csim/update_ops_dm.cpp:14-31 (divides by the trace; no square root)
void oracle_dm_normalize(double squared_norm, double* rho, ITYPE dim) {
    CTYPE* s = C(rho);
    const double f = 1. / squared_norm;
#pragma omp parallel for
    for (long long i = 0; i < (long long)(dim * dim); ++i) s[i] *= f;
}

// src/csim/stat_ops_dm.cpp:12-26 (trace)
double oracle_dm_state_norm_squared(const double* rho, ITYPE dim) {
    const CTYPE* s = C(rho);
    double n = 0;
    for (ITYPE i = 0; i < dim; ++i) n += s[i * dim + i].real();
    return n;
}
// src/csim/stat_ops_dm.cpp:49-63
double oracle_dm_M0_prob(UINT t, const double* rho, ITYPE dim) {
    const CTYPE* s = C(rho);
    double sum = 0;
    for (ITYPE i = 0; i < dim / 2; ++i) {
        const ITYPE b = insert_zero(i, t);
        sum += s[b * dim + b].real();
    }
    return sum;
}
// src/csim/stat_ops_dm.cpp:89-114
double oracle_dm_marginal_prob(const UINT* sorted_targets, const UINT* values, UINT n, const double* rho, ITYPE dim) {
    const CTYPE* s = C(rho);
    double sum = 0;
    for (ITYPE i = 0; i < (dim >> n); ++i) {
        ITYPE b = i;
        for (UINT c = 0; c < n; ++c) {
            b = insert_zero(b, sorted_targets[c]);
            b ^= (1ULL << sorted_targets[c]) * values[c];
        }
        sum += s[b * dim + b].real();
    }
    return sum;
}
// src/csim/stat_ops_dm.cpp:164-191
double oracle_dm_expectation_value_multi_qubit_Pauli_operator_partial_list(
    const UINT* targets, const UINT* paulis, UINT n, const double* rho, ITYPE dim) {
    const CTYPE* s = C(rho);
    CTYPE sum = 0;
    for (ITYPE i = 0; i < dim; ++i) {
        CTYPE coef = 1.0;
        ITYPE j = i;
        for (UINT c = 0; c < n; ++c) {
            const UINT p = paulis[c], t = targets[c];
            if (p == 1) {
                j ^= 1ULL << t;
            } else if (p == 2) {
                coef *= CTYPE(0, 1);
                if (j & (1ULL << t)) coef *= -1.;
                j ^= 1ULL << t;
            } else if (p == 3) {
                if (j & (1ULL << t)) coef *= -1.;
            }
        }
        sum += coef * s[i * dim + j];
    }
    return sum.real();
}

// src/cppsim/gate_general.hpp:265-285 (QuantumGate_Probabilistic on a density
// matrix) with the gate list of DepolarizingNoise, src/cppsim/gate_factory.cpp:454-467:
// probabilities {p/3, p/3, p/3, 1-p} on {X, Z, Y, I}.  The same loop order:
// rho <- (1 - sum(dist)) rho; then for each gate G_i: rho += dist_i * G_i rho G_i^dagger.
void oracle_dm_depolarizing(UINT t, double p, double* rho, ITYPE dim) {
    const ITYPE n2 = dim * dim;
    std::vector<CTYPE> org(C(rho), C(rho) + n2), tmp(n2);
    const double dist[4] = {p / 3, p / 3, p / 3, 1 - p};
    double cum = 0;
    for (double d : dist) cum += d;
    const CTYPE X[4] = {0, 1, 1, 0};
    const CTYPE Z[4] = {1, 0, 0, -1};
    const CTYPE Y[4] = {0, CTYPE(0, -1), CTYPE(0, 1), 0};
    const CTYPE* gates[3] = {X, Z, Y};
    CTYPE* s = C(rho);
    for (ITYPE i = 0; i < n2; ++i) s[i] *= (1.0 - cum);
    for (int g = 0; g < 4; ++g) {
        if (g < 3) {
            tmp = org;
            oracle_dm_single_qubit_dense_matrix_gate(t, reinterpret_cast<const double*>(gates[g]),
                                                     reinterpret_cast<double*>(tmp.data()), dim);
            for (ITYPE i = 0; i < n2; ++i) s[i] += tmp[i] * dist[g];
        } else {
            for (ITYPE i = 0; i < n2; ++i) s[i] += org[i] * dist[g];
        }
    }
}

}  // extern "C"

extern "C" {
// The sequential CDF itself (cdf[0..dim]) with the same formula as
// oracle_sampling, for tests that place uniforms exactly on CDF values.
void oracle_sampling_cdf(const double* state, unsigned long long dim, double* cdf) {
    double sum = 0.;
    cdf[0] = 0.;
    for (unsigned long long i = 0; i < dim; ++i) {
        const double re = state[2 * i], im = state[2 * i + 1];
        sum += std::fma(re, re, im * im);
        cdf[i + 1] = sum;
    }
}
}
