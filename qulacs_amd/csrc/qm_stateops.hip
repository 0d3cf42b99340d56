// State utilities on the device: tensor product, qubit permutation, projection
// onto fixed values of some qubits (drop), and partial traces.  One thread per
// output element (grid-stride, 64-bit indices); every output is written once.
//
// Reference loops restated: src/csim/stat_ops.cpp:111-160 (state_tensor_product,
// state_permutate_qubit, state_drop_qubits) and src/csim/stat_ops_dm.cpp:193-298
// (dm_state_tensor_product, dm_state_permutate_qubit,
// dm_state_partial_trace_from_density_matrix / _from_state_vector).
#include <algorithm>

#include "qm_internal.hpp"

namespace qm {
namespace {

constexpr unsigned kMaxQubits = 64;

struct QubitList {  // up to 64 qubit positions, one byte each
    unsigned char q[kMaxQubits];
    unsigned n;
};

QubitList make_list(const unsigned* q, unsigned n) {
    QubitList l{};
    l.n = n;
    for (unsigned i = 0; i < n; ++i) l.q[i] = (unsigned char)q[i];
    return l;
}

// dst[l * dim_r + r] = left[l] * right[r]
__global__ __launch_bounds__(256) void k_tensor(const double2* __restrict__ left, const double2* __restrict__ right,
                                                double2* __restrict__ dst, u64 dim_r, u64 total) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride)
        dst[i] = cmul(left[i / dim_r], right[i % dim_r]);
}

// (y, x) of the product of two dim_l x dim_l and dim_r x dim_r matrices
__global__ __launch_bounds__(256) void k_dm_tensor(const double2* __restrict__ left, const double2* __restrict__ right,
                                                   double2* __restrict__ dst, u64 dim_l, u64 dim_r) {
    const u64 dim = dim_l * dim_r, total = dim * dim;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const u64 y = i / dim, x = i % dim;
        dst[i] = cmul(right[(y % dim_r) * dim_r + (x % dim_r)], left[(y / dim_r) * dim_l + (x / dim_r)]);
    }
}

// source index of destination index `index`: bit order[k] is set for every set bit k
__device__ __forceinline__ u64 permuted(u64 index, const QubitList& order) {
    u64 src = 0;
    for (unsigned k = 0; k < order.n; ++k) src |= ((index >> k) & 1ULL) << order.q[k];
    return src;
}

__global__ __launch_bounds__(256) void k_permute(const double2* __restrict__ src, double2* __restrict__ dst, u64 dim,
                                                 QubitList order) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) dst[i] = src[permuted(i, order)];
}

__global__ __launch_bounds__(256) void k_dm_permute(const double2* __restrict__ src, double2* __restrict__ dst, u64 dim,
                                                    QubitList order) {
    const u64 stride = (u64)gridDim.x * blockDim.x, total = dim * dim;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride)
        dst[i] = src[permuted(i / dim, order) * dim + permuted(i % dim, order)];
}

// insert zero bits at the ascending positions of `sorted`
__device__ __forceinline__ u64 spread(u64 i, const QubitList& sorted) {
    for (unsigned k = 0; k < sorted.n; ++k) i = insert_zero(i, sorted.q[k]);
    return i;
}

__global__ __launch_bounds__(256) void k_drop(const double2* __restrict__ src, double2* __restrict__ dst, u64 dst_dim,
                                              QubitList sorted, u64 projection_mask) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dst_dim; i += stride)
        dst[i] = src[spread(i, sorted) ^ projection_mask];
}

// rho'[y][x] = sum_t rho[y ^ t][x ^ t] (FROM_SV: psi[y ^ t] conj(psi[x ^ t])) over the 2^k
// patterns t of the traced qubits, in ascending t (the reference's mask-list order)
template <bool FROM_SV>
__global__ __launch_bounds__(256) void k_partial_trace(const double2* __restrict__ src, double2* __restrict__ dst,
                                                       u64 dim, u64 dst_dim, QubitList sorted, QubitList targets) {
    const u64 stride = (u64)gridDim.x * blockDim.x, total = dst_dim * dst_dim;
    const u64 trace_dim = 1ULL << targets.n;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const u64 by = spread(i / dst_dim, sorted), bx = spread(i % dst_dim, sorted);
        double2 acc = make_double2(0.0, 0.0);
        for (u64 t = 0; t < trace_dim; ++t) {
            u64 m = 0;  // bit j of t on target j (the given order)
            for (unsigned j = 0; j < targets.n; ++j) m |= ((t >> j) & 1ULL) << targets.q[j];
            if (FROM_SV) {
                const double2 a = src[by ^ m], b = src[bx ^ m];
                acc = cadd(acc, make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y));
            } else {
                acc = cadd(acc, src[(by ^ m) * dim + (bx ^ m)]);
            }
        }
        dst[i] = acc;
    }
}

bool check_positions(const unsigned* q, unsigned n, unsigned nqubits, bool distinct, const char* what) {
    if (n > kMaxQubits || (n && !q)) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: bad qubit list", what);
        return false;
    }
    u64 seen = 0;
    for (unsigned i = 0; i < n; ++i) {
        if (q[i] >= nqubits || (distinct && ((seen >> q[i]) & 1))) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: qubit %u out of range or repeated", what, q[i]);
            return false;
        }
        seen |= 1ULL << q[i];
    }
    return true;
}

int qubits_of(u64 dim, const char* what) {
    const int n = log2_dim(dim);
    if (n < 0) set_error(QM_ERR_INVALID_ARGUMENT, "%s: dim %llu is not a power of two >= 2", what, dim);
    return n;
}

}  // namespace
}  // namespace qm

using namespace qm;

extern "C" {

void state_tensor_product_host(const void* left, ITYPE dim_left, const void* right, ITYPE dim_right, void* dst,
                               void* stream, unsigned dev) {
    const char* what = "state_tensor_product_host";
    if (!left || !right || !dst || qubits_of(dim_left, what) < 0 || qubits_of(dim_right, what) < 0 ||
        !use_device(dev))
        return;
    hipStream_t st = stream_of(stream);
    const u64 total = dim_left * dim_right;
    k_tensor<<<grid_for(total, 256 * 4), 256, 0, st>>>((const double2*)left, (const double2*)right, (double2*)dst,
                                                       dim_right, total);
    finish_update(st, what);
}

void dm_state_tensor_product_host(const void* left, ITYPE dim_left, const void* right, ITYPE dim_right, void* dst,
                                  void* stream, unsigned dev) {
    const char* what = "dm_state_tensor_product_host";
    if (!left || !right || !dst || qubits_of(dim_left, what) < 0 || qubits_of(dim_right, what) < 0 ||
        !use_device(dev))
        return;
    hipStream_t st = stream_of(stream);
    const u64 dim = dim_left * dim_right;
    k_dm_tensor<<<grid_for(dim * dim, 256 * 4), 256, 0, st>>>((const double2*)left, (const double2*)right,
                                                               (double2*)dst, dim_left, dim_right);
    finish_update(st, what);
}

void state_permutate_qubit_host(const unsigned* qubit_order, const void* src, void* dst, unsigned qubit_count,
                                ITYPE dim, void* stream, unsigned dev) {
    const char* what = "state_permutate_qubit_host";
    const int n = qubits_of(dim, what);
    if (n < 0 || !src || !dst) return;
    if ((unsigned)n != qubit_count || !check_positions(qubit_order, qubit_count, qubit_count, true, what) ||
        !use_device(dev)) {
        if ((unsigned)n != qubit_count) set_error(QM_ERR_INVALID_ARGUMENT, "%s: qubit count / dim mismatch", what);
        return;
    }
    hipStream_t st = stream_of(stream);
    k_permute<<<grid_for(dim, 256 * 4), 256, 0, st>>>((const double2*)src, (double2*)dst, dim,
                                                       make_list(qubit_order, qubit_count));
    finish_update(st, what);
}

void dm_state_permutate_qubit_host(const unsigned* qubit_order, const void* src, void* dst, unsigned qubit_count,
                                   ITYPE dim, void* stream, unsigned dev) {
    const char* what = "dm_state_permutate_qubit_host";
    const int n = qubits_of(dim, what);
    if (n < 0 || !src || !dst) return;
    if ((unsigned)n != qubit_count || !check_positions(qubit_order, qubit_count, qubit_count, true, what) ||
        !use_device(dev)) {
        if ((unsigned)n != qubit_count) set_error(QM_ERR_INVALID_ARGUMENT, "%s: qubit count / dim mismatch", what);
        return;
    }
    hipStream_t st = stream_of(stream);
    k_dm_permute<<<grid_for(dim * dim, 256 * 4), 256, 0, st>>>((const double2*)src, (double2*)dst, dim,
                                                                make_list(qubit_order, qubit_count));
    finish_update(st, what);
}

void state_drop_qubits_host(const unsigned* target, const unsigned* projection, unsigned target_count,
                            const void* src, void* dst, ITYPE dim, void* stream, unsigned dev) {
    const char* what = "state_drop_qubits_host";
    const int n = qubits_of(dim, what);
    if (n < 0 || !src || !dst) return;
    if (target_count >= (unsigned)n || !check_positions(target, target_count, (unsigned)n, true, what)) {
        if (target_count >= (unsigned)n) set_error(QM_ERR_INVALID_ARGUMENT, "%s: must keep at least one qubit", what);
        return;
    }
    u64 mask = 0;
    for (unsigned i = 0; i < target_count; ++i) {
        if (!projection || projection[i] > 1) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: projection values must be 0 or 1", what);
            return;
        }
        mask ^= (u64)projection[i] << target[i];
    }
    if (!use_device(dev)) return;
    std::vector<unsigned> sorted(target, target + target_count);
    std::sort(sorted.begin(), sorted.end());
    hipStream_t st = stream_of(stream);
    const u64 dst_dim = dim >> target_count;
    k_drop<<<grid_for(dst_dim, 256 * 4), 256, 0, st>>>((const double2*)src, (double2*)dst, dst_dim,
                                                        make_list(sorted.data(), target_count), mask);
    finish_update(st, what);
}

static void partial_trace(bool from_sv, const unsigned* target, unsigned target_count, const void* src, void* dst,
                          ITYPE dim, void* stream, unsigned dev, const char* what) {
    const int n = qubits_of(dim, what);
    if (n < 0 || !src || !dst) return;
    if (target_count >= (unsigned)n || !check_positions(target, target_count, (unsigned)n, true, what)) {
        if (target_count >= (unsigned)n) set_error(QM_ERR_INVALID_ARGUMENT, "%s: must keep at least one qubit", what);
        return;
    }
    if (!use_device(dev)) return;
    std::vector<unsigned> sorted(target, target + target_count);
    std::sort(sorted.begin(), sorted.end());
    hipStream_t st = stream_of(stream);
    const u64 dst_dim = dim >> target_count;
    const QubitList s = make_list(sorted.data(), target_count), t = make_list(target, target_count);
    if (from_sv)
        k_partial_trace<true><<<grid_for(dst_dim * dst_dim, 256 * 4), 256, 0, st>>>((const double2*)src,
                                                                                   (double2*)dst, dim, dst_dim, s, t);
    else
        k_partial_trace<false><<<grid_for(dst_dim * dst_dim, 256 * 4), 256, 0, st>>>((const double2*)src,
                                                                                    (double2*)dst, dim, dst_dim, s, t);
    finish_update(st, what);
}

void dm_state_partial_trace_from_density_matrix_host(const unsigned* target, unsigned target_count, const void* src,
                                                     void* dst, ITYPE dim, void* stream, unsigned dev) {
    partial_trace(false, target, target_count, src, dst, dim, stream, dev,
                  "dm_state_partial_trace_from_density_matrix_host");
}

void dm_state_partial_trace_from_state_vector_host(const unsigned* target, unsigned target_count, const void* src,
                                                   void* dst, ITYPE dim, void* stream, unsigned dev) {
    partial_trace(true, target, target_count, src, dst, dim, stream, dev,
                  "dm_state_partial_trace_from_state_vector_host");
}

}  // extern "C"
