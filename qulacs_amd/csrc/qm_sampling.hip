// Sampling: out[c] = lower_bound(cdf, u[c]) - 1 over the sequential,
// un-normalised CDF of QuantumStateCpu::sampling (src/cppsim/state.hpp:909-930):
//     cdf[0] = 0,  cdf[i+1] = fl(cdf[i] + p_i)
// with p_i = fma(re, re, im*im) for a state vector, |rho_ii| for a density
// matrix (src/cppsim/state_dm.hpp:383-408).
//
// A parallel scan would round differently from that left-to-right sum and can
// move a sample across a CDF boundary.  This kernel reproduces the sequential
// roundings exactly instead.  While the running sum s stays in one binade
// [2^e, 2^(e+1)) every partial sum is a multiple of u = 2^(e-52), so
//     fl(s + p) = s + u * round(p / u)
// (s is an integer S in units of u; S + round(p/u) is exact in 64-bit integers),
// as long as p/u is not exactly half-way between integers (ties round to even
// and so depend on the parity of S) and S + round(p/u) stays below 2^53 (the
// next binade has a coarser grid).  One work-group therefore walks the array in
// chunks: integer increments for the whole chunk, an integer block scan, and at
// the first element that is a tie or crosses the binade a single IEEE addition
// exactly as the CPU performs it, after which the chunk continues from there.
// Zero prefixes (s = 0, where fl(0 + p) = p) are handled the same way.
#include <cmath>

#include "qm_internal.hpp"

namespace qm {

constexpr int kCdfThreads = 1024;
constexpr int kCdfPer = 4;                       // elements per thread per chunk
constexpr int kCdfChunk = kCdfThreads * kCdfPer; // 4096 (32 KiB of LDS)

__device__ __forceinline__ double prob_sv(const double2* __restrict__ a, u64 i) {
    const double2 v = a[i];
    return fma(v.x, v.x, v.y * v.y);
}
__device__ __forceinline__ double prob_dm(const double2* __restrict__ a, u64 i, u64 dim) {
    const double2 v = a[i * dim + i];
    return v.y == 0.0 ? fabs(v.x) : hypot(v.x, v.y);
}

// Block-wide exclusive scan of one u64 per thread (1024 threads, 16 waves).
__device__ __forceinline__ u64 block_excl_scan(u64 v, u64* lds /* >= 16 */, u64* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    u64 x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const u64 y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 run = 0;
        for (int w = 0; w < kCdfThreads / 64; ++w) {
            const u64 t = lds[w];
            lds[w] = run;
            run += t;
        }
        lds[16] = run;
    }
    __syncthreads();
    const u64 excl = lds[wid] + (x - v);
    *total = lds[16];
    __syncthreads();
    return excl;
}

__device__ __forceinline__ int block_min_int(int v, int* lds /* >= 16 */) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = lds[0];
        for (int w = 1; w < kCdfThreads / 64; ++w) m = min(m, lds[w]);
        lds[16] = m;
    }
    __syncthreads();
    const int r = lds[16];
    __syncthreads();
    return r;
}

template <bool DM>
__global__ __launch_bounds__(kCdfThreads) void k_cdf_exact(const double2* __restrict__ a, u64 dim,
                                                           double* __restrict__ cdf) {
    __shared__ double P[kCdfChunk];
    __shared__ u64 scan_lds[17];
    __shared__ int min_lds[17];
    __shared__ double s_bcast;
    double s = 0.0;  // exact running sum, identical in every thread
    if (threadIdx.x == 0) cdf[0] = 0.0;
    for (u64 c0 = 0; c0 < dim; c0 += kCdfChunk) {
        const int len = (int)min((u64)kCdfChunk, dim - c0);
        for (int j = threadIdx.x; j < len; j += kCdfThreads)
            P[j] = DM ? prob_dm(a, c0 + j, dim) : prob_sv(a, c0 + j);
        __syncthreads();
        int pos = 0;  // first element of the chunk not yet summed
        while (pos < len) {
            const int lo = threadIdx.x * kCdfPer;  // this thread's elements: lo .. lo+kCdfPer-1
            if (!isfinite(s)) {
                // inf/NaN: finish the chunk with plain sequential additions
                // (only reachable from a non-finite amplitude).
                if (threadIdx.x == 0) {
                    double t = s;
                    for (int j = pos; j < len; ++j) {
                        t = t + P[j];
                        cdf[c0 + j + 1] = t;
                    }
                    s_bcast = t;
                }
                __syncthreads();
                s = s_bcast;
                __syncthreads();
                break;
            }
            if (s == 0.0) {
                // fl(0 + p) = p: the sum stays 0 up to the first positive p.
                int first = INT_MAX;
                for (int k = 0; k < kCdfPer; ++k) {
                    const int j = lo + k;
                    if (j >= pos && j < len && P[j] != 0.0) { first = j; break; }
                }
                first = block_min_int(first, min_lds);
                const int stop = first == INT_MAX ? len : first;
                for (int k = 0; k < kCdfPer; ++k) {
                    const int j = lo + k;
                    if (j >= pos && j < stop) cdf[c0 + j + 1] = 0.0;
                }
                if (first == INT_MAX) { pos = len; break; }
                s = 0.0 + P[first];
                if (threadIdx.x == 0) cdf[c0 + first + 1] = s;
                pos = first + 1;
                continue;
            }
            int e = ilogb(s);
            if (e < -1022) e = -1022;
            const double u = ldexp(1.0, e - 52);
            const u64 S = (u64)(s / u);  // exact: s is a multiple of u below 2^53 u
            const u64 LIM = 1ULL << 53;
            // per-element integer increments; ties are marked special
            u64 R[kCdfPer];
            bool tie[kCdfPer];
            u64 mine = 0;
#pragma unroll
            for (int k = 0; k < kCdfPer; ++k) {
                const int j = lo + k;
                R[k] = 0;
                tie[k] = false;
                if (j >= pos && j < len) {
                    double q = ldexp(P[j], 52 - e);  // exact scaling by a power of two (or +inf)
                    if (!(q < 0x1p62)) q = 0x1p62;   // certain binade crossing; keeps sums in range
                    const double fl = floor(q);
                    const double fr = q - fl;        // exact for q < 2^53; 0 above
                    R[k] = (u64)fl + (fr > 0.5 ? 1 : 0);
                    tie[k] = (fr == 0.5);
                }
                mine += R[k];
            }
            u64 total;
            const u64 excl = block_excl_scan(mine, scan_lds, &total);
            // first special element: a tie or the first partial sum reaching 2^53
            int first = INT_MAX;
            u64 run = S + excl;
#pragma unroll
            for (int k = 0; k < kCdfPer; ++k) {
                const int j = lo + k;
                if (j >= pos && j < len && first == INT_MAX) {
                    if (tie[k] || run + R[k] >= LIM) first = j;
                    run += R[k];
                }
            }
            first = block_min_int(first, min_lds);
            const int stop = first == INT_MAX ? len : first;
            run = S + excl;
            u64 before_first = 0;  // S + increments strictly before `first` (owned by its thread)
#pragma unroll
            for (int k = 0; k < kCdfPer; ++k) {
                const int j = lo + k;
                if (j >= pos && j < len) {
                    if (j == first) before_first = run;
                    run += R[k];
                    if (j < stop) cdf[c0 + j + 1] = (double)run * u;
                }
            }
            if (first == INT_MAX) {
                s = (double)(S + total) * u;
                pos = len;
            } else {
                // The owning thread performs the one IEEE addition exactly as the CPU does.
                if (lo <= first && first < lo + kCdfPer) {
                    const double sb = (double)before_first * u;
                    const double sn = sb + P[first];
                    cdf[c0 + first + 1] = sn;
                    s_bcast = sn;
                }
                __syncthreads();
                s = s_bcast;
                __syncthreads();
                pos = first + 1;
            }
        }
        __syncthreads();
    }
}

// lower_bound(cdf[0..dim], u) - 1, as an unsigned 64-bit index (the reference's
// -1 for u <= 0 becomes 2^64 - 1 through the same conversion).
__global__ void k_search(const double* __restrict__ cdf, u64 dim, const double* __restrict__ us, unsigned count,
                         u64* __restrict__ out) {
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= count) return;
    const double u = us[c];
    u64 lo = 0, hi = dim + 1;  // search [lo, hi) for the first cdf[i] >= u
    while (lo < hi) {
        const u64 mid = lo + (hi - lo) / 2;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    out[c] = lo - 1;
}

template <bool DM>
void sample(const double* uniforms, unsigned count, ITYPE* out, void* state, u64 dim, void* stream, unsigned dev,
            const char* what) {
    if (!state || log2_dim(dim) < 0) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null state or bad dim", what);
        return;
    }
    if (count == 0) return;
    if (!uniforms || !out) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null uniforms/out", what);
        return;
    }
    if (!use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    Scratch* sc = scratch_for(st);
    if (!sc) return;
    const size_t cdf_bytes = (size_t)(dim + 1) * sizeof(double);
    const size_t io_bytes = (size_t)count * 8;
    char* buf = (char*)big_buffer(sc, cdf_bytes + 2 * io_bytes + 256);
    if (!buf) return;
    double* cdf = (double*)buf;
    double* dus = (double*)(buf + ((cdf_bytes + 255) & ~size_t(255)));
    u64* dout = (u64*)((char*)dus + io_bytes);
    if (!hip_ok(hipMemcpyAsync(dus, uniforms, io_bytes, hipMemcpyHostToDevice, st), what)) return;
    k_cdf_exact<DM><<<1, kCdfThreads, 0, st>>>((const double2*)state, dim, cdf);
    k_search<<<(count + 255) / 256, 256, 0, st>>>(cdf, dim, dus, count, dout);
    if (!hip_ok(hipMemcpyAsync(out, dout, io_bytes, hipMemcpyDeviceToHost, st), what)) return;
    hip_ok(hipStreamSynchronize(st), what);
}

}  // namespace qm

using namespace qm;

extern "C" {

void qm_sampling_host(const double* uniforms, UINT count, ITYPE* out, void* state, ITYPE dim, void* stream,
                      unsigned dev) {
    sample<false>(uniforms, count, out, state, dim, stream, dev, "qm_sampling_host");
}

void dm_sampling_host(const double* uniforms, UINT count, ITYPE* out, void* state, ITYPE dim, void* stream,
                      unsigned dev) {
    sample<true>(uniforms, count, out, state, dim, stream, dev, "dm_sampling_host");
}

}  // extern "C"
