// Internal helpers shared by the HIP translation units of libqulacs_mi355.
//
// Layout facts every kernel relies on:
//   * a state vector is 2^n complex128 values, 16 B each, stored as double2
//     (x = real, y = imag); qubit q is bit q of the index
//     (src/csim/type.hpp:18-29 in the reference);
//   * a density matrix is row-major rho[y * 2^n + x], so as a flat array it is
//     a 2n-bit index whose low n bits are the column (ket^dagger) qubits and
//     whose high n bits are the row (ket) qubits (src/csim/update_ops_dm.cpp:73-76).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qulacs_mi355.h"

namespace qm {

using u64 = unsigned long long;
using u32 = unsigned int;

constexpr int kWave = 64;  // CDNA wavefront width

// ------------------------------------------------------------------ errors
void set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
bool hip_ok(hipError_t e, const char* what);

#define QM_CHECK_ARG(cond, ...)                            \
    do {                                                   \
        if (!(cond)) {                                     \
            ::qm::set_error(QM_ERR_INVALID_ARGUMENT, __VA_ARGS__); \
            return;                                        \
        }                                                  \
    } while (0)

#define QM_CHECK_ARG_RET(cond, ret, ...)                   \
    do {                                                   \
        if (!(cond)) {                                     \
            ::qm::set_error(QM_ERR_INVALID_ARGUMENT, __VA_ARGS__); \
            return ret;                                    \
        }                                                  \
    } while (0)

// ------------------------------------------------------------------ context
// The `stream` argument of every entry point follows the reference contract:
// a pointer to a hipStream_t (the array made by allocate_cuda_stream_host), or
// NULL for the null stream.
inline hipStream_t stream_of(void* stream) {
    return stream ? *reinterpret_cast<hipStream_t*>(stream) : (hipStream_t)0;
}

// Selects `device` if it is not current.  Returns false (error recorded) on failure.
bool use_device(unsigned device);

// Ends an update entry point: checks the launch and, in synchronous mode (the
// reference's behaviour, src/gpusim/update_ops_single.cu:130), waits for the stream.
void finish_update(hipStream_t s, const char* what);

bool synchronous_mode();

// log2 of dim if dim is a power of two >= 2, else -1.
inline int log2_dim(u64 dim) {
    if (dim < 2 || (dim & (dim - 1))) return -1;
    return __builtin_ctzll(dim);
}

// Per (device, stream) scratch: reduction partials, matrix uploads and a small
// pinned host buffer for scalar results.  Calls that share a stream are
// serialised by the stream itself; different streams get different scratch.
struct Scratch {
    int device = -1;
    hipStream_t stream = nullptr;
    double* partials = nullptr;     // device, kPartialsCap doubles
    double* host_result = nullptr;  // pinned, kHostResultCap doubles
    char* ring = nullptr;           // device ring for matrices / tables
    size_t ring_head = 0;
    void* big = nullptr;            // grown on demand (sampling CDF etc.)
    size_t big_bytes = 0;
};
// doubles: a 256-term reduction over 2048 blocks (2^20 partials) plus its
// per-term results
constexpr size_t kPartialsCap = (1 << 20) + 4096;
constexpr size_t kHostResultCap = 1 << 16;
constexpr size_t kRingBytes = 8 << 20;

// Returns the scratch for the current device and stream `s` (created on first
// use); nullptr with an error recorded if allocation fails.
Scratch* scratch_for(hipStream_t s);

// Copies `bytes` of host memory into the stream's device ring and returns the
// device pointer.  The host buffer may be reused as soon as this returns.
void* upload(Scratch* sc, const void* host, size_t bytes);

// Ensures sc->big holds at least `bytes`; returns it or nullptr.
void* big_buffer(Scratch* sc, size_t bytes);

// ------------------------------------------------------------------ device math
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc + a * b
__device__ __forceinline__ double2 cmac(double2 acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// Streaming (non-temporal) 16-byte accesses for data touched once per pass:
// measured on MI355X they raise an in-place read-modify-write sweep from
// ~5.5 to ~6.0 TB/s by not keeping the stream resident in the caches.
typedef double qm_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_stream(const double2* p) {
    const qm_d2v v = __builtin_nontemporal_load(reinterpret_cast<const qm_d2v*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(double2* p, double2 v) {
    const qm_d2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<qm_d2v*>(p));
}
// NT selects streaming accesses.  They pay off only when each 128-byte line is
// consumed by a single wave instruction, i.e. when no target/control qubit is
// below bit 3; otherwise the lines are revisited by the partner loads and must
// stay cached.
template <bool NT>
__device__ __forceinline__ double2 ld(const double2* p) {
    if (NT) return ld_stream(p);
    return *p;
}
template <bool NT>
__device__ __forceinline__ void st(double2* p, double2 v) {
    if (NT) st_stream(p, v);
    else *p = v;
}
constexpr unsigned kLineQubits = 3;  // 128-byte line = 8 amplitudes
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// Insert a zero bit at position q (src/csim/utility.hpp:14-18).
__host__ __device__ __forceinline__ u64 insert_zero(u64 i, u32 q) {
    const u64 low = (1ULL << q) - 1;
    return (i & low) | ((i & ~low) << 1);
}

// Insert zeros at the ascending positions pos[0..n).
__host__ __device__ __forceinline__ u64 insert_zeros(u64 i, const unsigned char* pos, int n) {
    for (int k = 0; k < n; ++k) i = insert_zero(i, pos[k]);
    return i;
}

// Wave-level FP64 sum (all lanes get the result).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-level FP64 sum over blockDim.x threads (multiple of 64, <= 1024).
// Deterministic: fixed shuffle tree, then a fixed-order sum over waves.
__device__ __forceinline__ double block_sum(double v, double* lds /* >= 16 */) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int w = 0; w < nw; ++w) r += lds[w];
        lds[0] = r;
    }
    __syncthreads();
    r = lds[0];
    __syncthreads();
    return r;
}

// Grid sizing for streaming kernels: enough blocks to cover `work` items with
// `per_block` items each, capped so grid-stride loops take the rest.
inline unsigned grid_for(u64 work, u64 per_block, unsigned cap = 1u << 20) {
    u64 g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Sorted, de-duplicated insert positions plus validation of a qubit list.
// Returns false (error recorded) on out-of-range or duplicated qubits.
bool check_qubits(const u32* q, u32 n, int nqubits, const char* what);

// A k-qubit gather plan: the insert positions for enumerating bases and the
// offset of each of the 2^k tuple members (bit i of the tuple index <-> targets[i]).
struct GatherPlan {
    unsigned char ins[64];
    int n_ins = 0;
    u64 control_mask = 0;
};
GatherPlan make_plan(const u32* targets, u32 nt, const u32* controls, const u32* cvals, u32 nc);

// Dense (optionally controlled) gate on an arbitrary flat array of `dim`
// complex values (used for state vectors and, through the row/column bit
// split, for density matrices).  Defined in qm_gates.hip.
void dense_any(const u32* targets, u32 nt, const u32* controls, const u32* cvals, u32 nc, const qm_complex* matrix,
               void* state, u64 dim, void* stream, unsigned device, const char* what);
// A 1..4-qubit operator with one nonzero per row and column (a permutation of
// the tuple members with factors): one pass that reads and writes only the
// members that change.  is_monomial tests a row-major 2^k x 2^k matrix.
bool is_monomial(const qm_complex* matrix, u32 k);
void monomial_any(const u32* targets, u32 nt, const qm_complex* matrix, void* state, u64 dim, void* stream,
                  unsigned device, const char* what);

// ------------------------------------------------------------------ enqueue-only
// Stream-ordered building blocks that never wait for the stream (the request
// server chains them and waits on events).  The current device must already
// be selected.  Host pointers named h_* should be pinned for the copies to be
// asynchronous.

// state <- |basis> (qm_runtime.hip)
void enqueue_basis(void* state, u64 dim, u64 basis, hipStream_t st);
// sum_k coef_k <P_k>: the Z-only terms' total is copied to h_out[0..1] and the
// flip terms' total to h_out[2..3] (re, im); the caller adds the two.
// (qm_reduce.hip)
bool enqueue_observable(const u64* bf, const u64* pf, const u32* rot, const double* coef, u32 nterms,
                        const void* state, u64 dim, hipStream_t st, double* h_out);
// lower_bound sampling of one state vector with h_uniforms -> h_out (qm_sampling.hip)
// With h_flags, also the boundary check of qm_sampling_checked_host (count bytes).
// The CDF and its work tables live in the stream's grow-on-demand scratch.
bool enqueue_sampling(const void* state, u64 dim, const double* h_uniforms, u32 count, u64* h_out, hipStream_t st,
                      double margin = 0.0, unsigned char* h_flags = nullptr);
// The same in a caller-provided device buffer of sampling_buffer_bytes bytes
// (nothing is allocated).
size_t sampling_buffer_bytes(u64 dim, u32 count, bool flags);
bool enqueue_sampling_in(void* buffer, const void* state, u64 dim, const double* h_uniforms, u32 count, u64* h_out,
                         hipStream_t st, double margin, unsigned char* h_flags);
// Uploads a fused circuit's program to `dev` ahead of its first launch.
bool circuit_prepare(qm_circuit* c, unsigned dev);
// Many small state vectors (dim <= 2^16 each), one work-group per CDF.
struct SampleJob {
    const double2* state;
    u64 dim;
    const double* h_uniforms;
    u64* h_out;
    u32 count;
    double margin = 0.0;               // boundary check (qm_sampling_checked_host)
    unsigned char* h_flags = nullptr;  // count bytes, or nullptr for no check
};
bool enqueue_sampling_jobs(const SampleJob* jobs, u32 njobs, hipStream_t st);
// qm_circuit_apply without the final synchronisation (qm_batch.hip)
bool circuit_enqueue(qm_circuit* c, void* state, u64 dim, hipStream_t st, unsigned dev);

}  // namespace qm
