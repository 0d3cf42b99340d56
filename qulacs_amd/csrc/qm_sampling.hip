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
#include <climits>
#include <cmath>
#include <vector>

#include "qm_internal.hpp"

namespace qm {

constexpr int kCdfThreads = 1024;
constexpr int kCdfPer = 4;                       // elements per thread per chunk
constexpr int kCdfChunk = kCdfThreads * kCdfPer; // 4096 (32 KiB of LDS)
// A block scan that stops within kSerialTrigger elements of where it started is
// followed by kSerialRun one-by-one additions (see exact_chunk).
constexpr int kSerialTrigger = 64;
constexpr int kSerialRun = 512;

__device__ __forceinline__ double prob_sv(const double2* __restrict__ a, u64 i) {
    const double2 v = a[i];
    return fma(v.x, v.x, v.y * v.y);
}
// |x + iy| rounded as the reference's std::abs(std::complex<double>) rounds it.
// libstdc++ lowers that call to cabs = glibc's hypot, which is not correctly
// rounded (about 1 result in 1000 differs from the correctly rounded one).
// glibc 2.35's hypot is the non-FMA form of C. F. Borges, "An Improved
// Algorithm for hypot(a, b)", arXiv:1904.09481 (2019): h = sqrt(x^2 + y^2)
// followed by one correction term, with a power-of-two rescale for huge or tiny
// operands and a shortcut for widely varying ones.  It is restated here with
// FMA contraction off, so every product and sum rounds as in the host library.
__device__ __noinline__ double hypot_host(double x, double y) {
#pragma clang fp contract(off)
    double ax = fabs(x), ay = fabs(y);
    if (ax < ay) {
        const double t = ax;
        ax = ay;
        ay = t;
    }
    if (!(ax <= 0x1.fffffffffffffp+1023)) return (isinf(ax) || isinf(ay)) ? INFINITY : ax + ay;  // inf / NaN
    constexpr double kEps = 0x1p-54, kScale = 0x1p-600, kLarge = 0x1p+511, kTiny = 0x1p-511;
    double scale = 1.0;
    if (ax > kLarge) {
        if (ay <= ax * kEps) return ax + ay;
        ax *= kScale;
        ay *= kScale;
        scale = 1.0 / kScale;
    } else if (ay < kTiny) {
        if (ax >= ay / kEps) return ax + ay;
        ax /= kScale;
        ay /= kScale;
        scale = kScale;
    } else if (ay <= ax * kEps) {
        return ax + ay;
    }
    double h = sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        const double delta = h - ay;
        t1 = ax * (2.0 * delta - ax);
        t2 = (delta - 2.0 * (ax - ay)) * delta;
    } else {
        const double delta = h - ax;
        t1 = 2.0 * delta * (ax - 2.0 * ay);
        t2 = (4.0 * delta - ay) * ay + delta * delta;
    }
    h -= (t1 + t2) / (2.0 * h);
    return scale == 1.0 ? h : (scale < 1.0 ? h * scale : h / kScale);
}
__device__ __forceinline__ double prob_dm(const double2* __restrict__ a, u64 i, u64 dim) {
    const double2 v = a[i * dim + i];
    return v.y == 0.0 ? fabs(v.x) : hypot_host(v.x, v.y);
}

// Block-wide exclusive scan of one u64 per thread (1024 threads, 16 waves).
__device__ __forceinline__ u64 block_excl_scan(u64 v, u64* lds /* >= 16 */, u64* total) {
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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

// One exchange of per-wave partials across the 16 waves of a 1024-thread block
// with a single barrier.  Exchanges alternate between two LDS slots (r & 1): a
// wave that writes a slot again, two exchanges later, has passed the barrier of
// the exchange in between, which every wave reaches only after it has read the
// slot.  Every thread of the block runs the same sequence of exchanges.
constexpr int kCdfWaves = kCdfThreads / 64;
struct Xchg {
    u64 sum[kCdfWaves];
    u64 sum2[kCdfWaves];
    u64 val[kCdfWaves];
    double dv[kCdfWaves];
    int mn[kCdfWaves];
    int ok[kCdfWaves];
    double s;
    int first;
};

struct XScan {
    u64 excl;       // sum of v over the threads before this one
    u64 total;      // sum of v over the block
    u64 total2;     // sum of v2 over the block (when asked for)
    u64 wave_incl;  // in lane w < 16: sum of v over waves 0..w
    int first;      // m of the first thread with m != INT_MAX (INT_MAX if none)
};

// Cross-lane steps on the VALU (DPP) and the scalar unit (ballot, readlane)
// rather than through the LDS crossbar: the exchanges sit on the serial path
// of the walk, where their latency is the cost.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ u64 dpp_u64(u64 x) {
    // lanes without a source (or in rows left out by ROW_MASK) read 0
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(u32)x, CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u32)(x >> 32), CTRL, ROW_MASK, 0xf, false);
    return ((u64)(u32)hi << 32) | (u32)lo;
}
// inclusive scan within each row of 16 lanes
__device__ __forceinline__ u64 row_incl_scan(u64 x) {
    x += dpp_u64<0x111>(x);  // row_shr:1
    x += dpp_u64<0x112>(x);  // row_shr:2
    x += dpp_u64<0x114>(x);  // row_shr:4
    x += dpp_u64<0x118>(x);  // row_shr:8
    return x;
}
// inclusive scan over the 64 lanes of a wave
__device__ __forceinline__ u64 wave_incl_scan(u64 x) {
    x = row_incl_scan(x);
    x += dpp_u64<0x142, 0xa>(x);  // row_bcast:15: lane 15 into row 1, lane 47 into row 3
    x += dpp_u64<0x143, 0xc>(x);  // row_bcast:31: lane 31 into rows 2 and 3
    return x;
}
__device__ __forceinline__ u64 readlane_u64(u64 x, int l) {
    const u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)x, l);
    const u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(x >> 32), l);
    return ((u64)hi << 32) | lo;
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double x) {
    return __longlong_as_double((long long)dpp_u64<CTRL, ROW_MASK>((u64)__double_as_longlong(x)));
}
__device__ __forceinline__ double row_incl_scan_f64(double x) {
    x += dpp_f64<0x111>(x);
    x += dpp_f64<0x112>(x);
    x += dpp_f64<0x114>(x);
    x += dpp_f64<0x118>(x);
    return x;
}
__device__ __forceinline__ double wave_incl_scan_f64(double x) {
    x = row_incl_scan_f64(x);
    x += dpp_f64<0x142, 0xa>(x);
    x += dpp_f64<0x143, 0xc>(x);
    return x;
}
__device__ __forceinline__ double readlane_f64(double x, int l) {
    return __longlong_as_double((long long)readlane_u64((u64)__double_as_longlong(x), l));
}
// m of the lowest lane with m != INT_MAX (INT_MAX if none).  Every use passes
// an element or thread index that grows with the lane, so this is the minimum.
__device__ __forceinline__ int wave_first_set(int m) {
    const u64 b = __ballot(m != INT_MAX);
    return b ? __builtin_amdgcn_readlane(m, __ffsll((unsigned long long)b) - 1) : INT_MAX;
}

// Exclusive scan of v and the first m != INT_MAX over a block of NW waves (m
// must grow with the thread index where it is set), and with TWO the block sum
// of v2.  The sums must not exceed 2^64 - 1.
template <int NW = kCdfWaves, bool TWO = false>
__device__ __forceinline__ XScan xchg_scan_first(u64 v, int m, Xchg* X, int& r, u64 v2 = 0) {
    static_assert(NW >= 1 && NW <= kCdfWaves, "at most 16 waves");
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u64 x = wave_incl_scan(v);
    const int mf = wave_first_set(m);
    Xchg& slot = X[r & 1];
    if (lane == 63) slot.sum[wid] = x;
    if (lane == 0) slot.mn[wid] = mf;
    if constexpr (TWO) {
        const u64 x2 = wave_incl_scan(v2);
        if (lane == 63) slot.sum2[wid] = x2;
    }
    __syncthreads();
    ++r;
    const u64 t = row_incl_scan(lane < NW ? slot.sum[lane] : 0);  // lanes 0..NW-1: over waves 0..lane
    const int tm = lane < NW ? slot.mn[lane] : INT_MAX;
    XScan o;
    o.total = readlane_u64(t, NW - 1);
    o.total2 = 0;
    if constexpr (TWO) o.total2 = readlane_u64(row_incl_scan(lane < NW ? slot.sum2[lane] : 0), NW - 1);
    o.excl = (wid == 0 ? 0 : readlane_u64(t, wid - 1)) + (x - v);
    o.wave_incl = t;
    o.first = wave_first_set(tm);
    return o;
}

// One thread (owner) hands (s, first) to the whole block.
__device__ __forceinline__ void xchg_bcast(bool owner, double& s, int& first, Xchg* X, int& r) {
    Xchg& slot = X[r & 1];
    if (owner) {
        slot.s = s;
        slot.first = first;
    }
    __syncthreads();
    ++r;
    s = slot.s;
    first = slot.first;
}

// Sums chunk [c0, c0 + len) onto the running sum s exactly as the sequential
// loop does and writes cdf[c0 + 1 .. c0 + len].  Called by every thread of a
// 1024-thread block with the same s; returns with the same s in every thread.
template <bool DM>
__device__ __forceinline__ void exact_chunk(const double2* __restrict__ a, u64 dim, u64 c0, int len, double& s,
                                            double* __restrict__ cdf, double* P, Xchg* X, int& r) {
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int j = threadIdx.x; j < len; j += kCdfThreads)
        P[j] = DM ? prob_dm(a, c0 + j, dim) : prob_sv(a, c0 + j);
    __syncthreads();
    const u64 LIM = 1ULL << 53;
    const int lo = threadIdx.x * kCdfPer;  // this thread's elements: lo .. lo+kCdfPer-1
    int pos = 0;     // first element of the chunk not yet summed
    int serial = 0;  // elements to add one by one before the next block scan
    while (pos < len) {
        if (serial > 0 || !isfinite(s)) {
            // One thread adds the next elements one at a time, the same IEEE
            // additions as the sequential loop, and leaves the sums in P for the
            // block to store.  Used after crossings close together (a small
            // running sum doubling every few elements), where it is cheaper than
            // a block scan per crossing, and to finish a chunk after an inf/NaN
            // (only reachable from a non-finite amplitude).
            const int end = isfinite(s) ? min(len, pos + serial) : len;
            if (threadIdx.x == 0) {
                double t = s;
                int j = pos;
                double v[16];
                if (j + 16 <= end) {
#pragma unroll
                    for (int k = 0; k < 16; ++k) v[k] = P[j + k];
                }
                for (; j + 16 <= end; j += 16) {
                    double w[16];  // the next 16, loaded before this block's sums are stored
                    if (j + 32 <= end) {
#pragma unroll
                        for (int k = 0; k < 16; ++k) w[k] = P[j + 16 + k];
                    }
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        t = t + v[k];
                        v[k] = t;
                    }
#pragma unroll
                    for (int k = 0; k < 16; ++k) P[j + k] = v[k];
#pragma unroll
                    for (int k = 0; k < 16; ++k) v[k] = w[k];
                }
                for (; j < end; ++j) {
                    t = t + P[j];
                    P[j] = t;
                }
                s = t;
            }
            int unused = 0;
            xchg_bcast(threadIdx.x == 0, s, unused, X, r);
            for (int j = pos + threadIdx.x; j < end; j += kCdfThreads) cdf[c0 + j + 1] = P[j];
            pos = end;
            serial = 0;
            continue;
        }
        if (s == 0.0) {
            // fl(0 + p) = p: the sum stays 0 up to the first positive p.
            int first = INT_MAX;
            for (int k = 0; k < kCdfPer; ++k) {
                const int j = lo + k;
                if (j >= pos && j < len && P[j] != 0.0) { first = j; break; }
            }
            first = xchg_scan_first<>(0, first, X, r).first;
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
        // per-element integer increments, at most 2^53 each (a larger one
        // certainly crosses the binade); the first tie
        u64 R[kCdfPer];
        u64 mine = 0;
        int tie = INT_MAX;
#pragma unroll
        for (int k = 0; k < kCdfPer; ++k) {
            const int j = lo + k;
            R[k] = 0;
            if (j >= pos && j < len) {
                double q = ldexp(P[j], 52 - e);  // exact scaling by a power of two (or +inf)
                if (!(q < 0x1p53)) q = 0x1p53;
                const double fl = floor(q);
                const double fr = q - fl;        // exact for q < 2^53
                R[k] = (u64)fl + (fr > 0.5 ? 1 : 0);
                if (fr == 0.5 && tie == INT_MAX) tie = j;
            }
            mine += R[k];
        }
        // A thread whose sum reaches 2^53 holds a crossing; clamping keeps the
        // block sum below 2^63 and changes no sum before the first crossing.
        if (mine > LIM) mine = LIM;
        const XScan xs = xchg_scan_first<>(mine, tie, X, r);
        const int first_tie = xs.first;
        // The partial sums S + (increments up to element j) never decrease, so
        // the first element reaching 2^53 (the first crossing) lies in the
        // first wave whose inclusive sum reaches it.
        int wcross = kCdfWaves;
        if (S + xs.total >= LIM) wcross = __ffsll((long long)__ballot(lane < kCdfWaves && S + xs.wave_incl >= LIM)) - 1;
        // first crossing as seen from this thread's wave: after all of its
        // elements (INT_MAX), before all of them (-1), or the element itself
        int cross = wid < wcross ? INT_MAX : -1;
        u64 run = S + xs.excl;
        if (wid == wcross) {  // wave-uniform
            int mc = INT_MAX;
#pragma unroll
            for (int k = 0; k < kCdfPer; ++k) {
                const int j = lo + k;
                if (j >= pos && j < len && mc == INT_MAX) {
                    if (run + R[k] >= LIM) mc = j;
                    run += R[k];
                }
            }
            cross = wave_first_set(mc);
        }
        // the first special element (tie or crossing) as seen from this
        // thread: exact for the thread that owns it and for every thread with
        // elements before it, and before all elements of any other thread
        const int first = min(first_tie, cross);
        run = S + xs.excl;
        u64 before_first = 0;  // S + increments strictly before `first` (owned by its thread)
#pragma unroll
        for (int k = 0; k < kCdfPer; ++k) {
            const int j = lo + k;
            if (j >= pos && j < len) {
                if (j == first) before_first = run;
                run += R[k];
                if (j < first) cdf[c0 + j + 1] = (double)run * u;
            }
        }
        if (first_tie == INT_MAX && wcross == kCdfWaves) {
            s = (double)(S + xs.total) * u;
            pos = len;
        } else {
            // The owning thread performs the one IEEE addition exactly as the CPU does.
            const bool owner = lo <= first && first < lo + kCdfPer;
            int f = first;
            if (owner) {
                s = (double)before_first * u + P[first];
                cdf[c0 + first + 1] = s;
            }
            xchg_bcast(owner, s, f, X, r);
            if (f - pos < kSerialTrigger) serial = kSerialRun;
            pos = f + 1;
        }
    }
    __syncthreads();
}

template <bool DM>
__global__ __launch_bounds__(kCdfThreads) void k_cdf_exact(const double2* __restrict__ a, u64 dim,
                                                           double* __restrict__ cdf) {
    __shared__ double P[kCdfChunk];
    __shared__ Xchg X[2];
    int r = 0;
    double s = 0.0;  // exact running sum, identical in every thread
    if (threadIdx.x == 0) cdf[0] = 0.0;
    for (u64 c0 = 0; c0 < dim; c0 += kCdfChunk) {
        const int len = (int)min((u64)kCdfChunk, dim - c0);
        exact_chunk<DM>(a, dim, c0, len, s, cdf, P, X, r);
    }
}

// ------------------------------------------------------------------ multi-block CDF
// The same sums for large states, spread over the whole GPU:
//   A (grid): approximate chunk totals T[c] (used only to predict);
//   B (1 block): predicted running sum in front of each chunk and its binade E[c];
//   C (grid): per chunk, the integer increments in the predicted binade and
//      their total N[c], and the chunk's kind: plain (no tie, no crossing
//      predicted), one tie (its position and the parity that decides its
//      rounding), one predicted binade crossing (its position, the increments
//      before it, its own probability and the increments after it in the next
//      binade), or anything else;
//   D (1 block): the exact running sum walks the chunks 1024 at a time: plain
//      chunks in the running sum's binade are fast while the sum stays below
//      2^53 units and only record their exact start value; at the first other
//      chunk the walk settles a single tie or a single crossing from the
//      precomputed values (one IEEE addition for the crossing) when the exact
//      start confirms them, and sums the chunk element by element
//      (exact_chunk) otherwise;
//   E (grid): every chunk not summed by the walk writes its CDF values from its
//      exact start.
// A wrong prediction only sends a chunk to exact_chunk; every value written
// equals the sequential loop's.
constexpr int kPredictThreads = 1024;
constexpr int kIncrThreads = 1024;
constexpr int kIncrWaves = kIncrThreads / 64;
constexpr int kIncrPer = kCdfChunk / kIncrThreads;  // elements per thread
constexpr u64 kMultiBlockMinDim = 1ULL << 16;

// chunk kinds (stage C)
constexpr int kChunkPlain = 0;
constexpr int kChunkTie = 1;   // 1..kMaxTies ties, no crossing predicted
constexpr int kMaxTies = 4;
constexpr int kChunkCross = 2;
constexpr int kChunkExact = 3;
// how the walk left a chunk (stage D)
constexpr int kStartPlain = 0;  // fill from the start value in the predicted binade
constexpr int kStartExact = 1;  // summed by the walk
constexpr int kStartCross = 2;  // fill around the crossing settled by the walk
constexpr int kStartTie = 3;    // fill with the tie rounded from the start's parity

struct ChunkSums {
    double* T;      // approximate chunk totals
    double* pst;    // predicted running sum in front of each chunk
    double* ep;     // crossing chunk: probability of the crossing element; after the walk, the sum through it
    double* start;  // exact running sum in front of each chunk (walk)
    u64* N;         // increments in the predicted binade (ties counted rounded down)
    u64* eA;        // ties: bit k = parity of the increments before tie k (ties rounded down) plus its floor;
                    // crossing: increments before it
    u64* eR;        // crossing: the crossing element's increment
    u64* eB;        // ties: element index of tie k in bits 12k..12k+11; crossing: increments after it, next binade
    int* E;         // predicted binade of the sum in front (INT_MIN: zero)
    int* kind;      // kChunk*
    int* ej;        // ties: their number; crossing: element index within the chunk
    int* exact;     // kStart*
};

template <bool DM>
__device__ __forceinline__ double prob_at(const double2* __restrict__ a, u64 i, u64 dim) {
    return DM ? prob_dm(a, i, dim) : prob_sv(a, i);
}

template <bool DM>
__global__ __launch_bounds__(kIncrThreads) void k_chunk_total(const double2* __restrict__ a, u64 dim,
                                                              double* __restrict__ T) {
    __shared__ double lds[32];
    const u64 c0 = (u64)blockIdx.x * kCdfChunk;
    const int len = (int)min((u64)kCdfChunk, dim - c0);
    double acc = 0.0;
    for (int j = threadIdx.x; j < len; j += blockDim.x) acc += prob_at<DM>(a, c0 + j, dim);
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) T[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kPredictThreads) void k_chunk_predict(ChunkSums cs, u64 M) {
    static_assert(kPredictThreads == kCdfThreads, "the block scan below assumes 16 waves");
    __shared__ double wsum[kCdfWaves];
    const u64 seg = (M + kPredictThreads - 1) / kPredictThreads;
    const u64 lo = (u64)threadIdx.x * seg, hi = min(lo + seg, M);
    double local = 0.0;
    for (u64 c = lo; c < hi; ++c) local += cs.T[c];
    // Exclusive block scan of the segment sums.  The order of these additions
    // only moves a prediction: a chunk predicted wrongly is summed exactly by
    // the walk.
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double x = wave_incl_scan_f64(local);
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    const double t = row_incl_scan_f64(lane < kCdfWaves ? wsum[lane] : 0.0);
    double run = (wid == 0 ? 0.0 : readlane_f64(t, wid - 1)) + dpp_f64<0x138>(x);  // wave_shr:1
    for (u64 c = lo; c < hi; ++c) {
        cs.pst[c] = run;
        cs.E[c] = run == 0.0 ? INT_MIN : max(ilogb(run), -1022);
        run += cs.T[c];
    }
}

template <bool DM>
__global__ __launch_bounds__(kIncrThreads) void k_chunk_incr(const double2* __restrict__ a, u64 dim, ChunkSums cs) {
    __shared__ Xchg X[2];
    // probabilities, loaded with consecutive lanes on consecutive elements and
    // read back as kIncrPer consecutive elements per thread (padded: element i
    // at i + i / kIncrPer, which spreads those reads over the LDS banks)
    __shared__ double Pl[kCdfChunk + kCdfChunk / kIncrPer];
    int r = 0;
    const u64 c = blockIdx.x;
    const double tc = cs.T[c];
    const int e = cs.E[c];
    if (tc == 0.0 || !isfinite(tc) || e == INT_MIN) {
        if (threadIdx.x == 0) {
            cs.N[c] = 0;
            cs.kind[c] = tc == 0.0 ? kChunkPlain : kChunkExact;
        }
        return;
    }
    const u64 c0 = c * kCdfChunk;
    const int len = (int)min((u64)kCdfChunk, dim - c0);
#pragma unroll
    for (int k = 0; k < kIncrPer; ++k) {
        const int j = k * kIncrThreads + threadIdx.x;
        Pl[j + j / kIncrPer] = j < len ? prob_at<DM>(a, c0 + j, dim) : 0.0;
    }
    __syncthreads();
    const u64 LIM = 1ULL << 53;
    // predicted sum in front of the chunk in units of binade e (only used to
    // place a crossing, which the walk confirms from the exact sum)
    const double sp = ldexp(cs.pst[c], 52 - e);
    const u64 S0 = sp < 0x1p53 ? (u64)sp : LIM;
    const int lo = threadIdx.x * kIncrPer;
    double p[kIncrPer];
    u64 R[kIncrPer];
    bool tk[kIncrPer];
    u64 mine = 0, nties = 0, nhuge = 0;
    int tie = INT_MAX;
#pragma unroll
    for (int k = 0; k < kIncrPer; ++k) {
        const int j = lo + k;
        p[k] = 0.0;
        R[k] = 0;
        tk[k] = false;
        if (j < len) {
            p[k] = Pl[j + j / kIncrPer];
            const double q = ldexp(p[k], 52 - e);
            if (!(q < 0x1p53)) {  // certainly crosses (or NaN)
                R[k] = LIM;
                ++nhuge;
            } else {
                const double fl = floor(q), fr = q - fl;
                R[k] = (u64)fl + (fr > 0.5 ? 1 : 0);
                if (fr == 0.5) {
                    tk[k] = true;
                    ++nties;
                    if (tie == INT_MAX) tie = j;
                }
            }
        }
        mine += R[k];
    }
    // clamped per thread: block sums stay below 2^61, no sum before the first
    // crossing changes
    if (mine > LIM) mine = LIM;
    const XScan xa = xchg_scan_first<kIncrWaves, true>(mine, tie, X, r, (nties << 32) | nhuge);
    const u64 N = xa.total;
    const u64 ties = xa.total2 >> 32, huge = xa.total2 & 0xffffffffu;
    int kind = kChunkExact;
    if (S0 + N < LIM) {
        if (huge == 0 && ties == 0) {
            kind = kChunkPlain;
        } else if (huge == 0 && ties <= kMaxTies) {
            kind = kChunkTie;
            // rank of each tie, its position and the parity of the increments
            // before it plus its floor (R[k] of a tie is its floor); the fields
            // of different ties are disjoint, so their sums are their union
            const XScan xt = xchg_scan_first<kIncrWaves>(nties, INT_MAX, X, r);
            u64 run = xa.excl;  // exact: every partial sum is below 2^53
            u64 rank = xt.excl, bits = 0, posw = 0;
#pragma unroll
            for (int k = 0; k < kIncrPer; ++k) {
                if (tk[k]) {
                    bits |= ((run + R[k]) & 1) << rank;
                    posw |= (u64)(lo + k) << (12 * rank);
                    ++rank;
                }
                run += R[k];
            }
            const XScan xo = xchg_scan_first<kIncrWaves, true>(bits, INT_MAX, X, r, posw);
            if (threadIdx.x == 0) {
                cs.ej[c] = (int)ties;
                cs.eA[c] = xo.total;
                cs.eB[c] = xo.total2;
            }
        }
    } else {
        // a crossing is predicted: the first element whose partial sum reaches 2^53 units
        u64 run = S0 + xa.excl;
        int mc = INT_MAX;
        u64 A = 0, Rj = 0;
        double pj = 0.0;
#pragma unroll
        for (int k = 0; k < kIncrPer; ++k) {
            const int j = lo + k;
            if (j < len && mc == INT_MAX) {
                if (run + R[k] >= LIM) {
                    mc = j;
                    A = run - S0;
                    Rj = R[k];
                    pj = p[k];
                }
                run += R[k];
            }
        }
        const int jx = xchg_scan_first<kIncrWaves>(0, mc, X, r).first;
        // after it, increments in the next binade (units of 2u); a tie there or
        // a second crossing leaves the chunk to the walk's element-wise sum
        u64 B = 0, bad = 0;
#pragma unroll
        for (int k = 0; k < kIncrPer; ++k) {
            const int j = lo + k;
            if (j < len && j > jx) {
                const double q2 = ldexp(p[k], 51 - e);
                if (!(q2 < 0x1p53)) {
                    ++bad;
                } else {
                    const double fl = floor(q2), fr = q2 - fl;
                    B += (u64)fl + (fr > 0.5 ? 1 : 0);
                    if (fr == 0.5) ++bad;
                }
            }
        }
        if (B > LIM) B = LIM;
        const XScan xb = xchg_scan_first<kIncrWaves, true>(B, INT_MAX, X, r, bad);
        // a tie at or before the crossing element also leaves it to the walk
        if (xa.first > jx && xb.total2 == 0 && xb.total < LIM) kind = kChunkCross;
        if (mc != INT_MAX && mc == jx) {
            cs.ej[c] = jx;
            cs.eA[c] = A;
            cs.eR[c] = Rj;
            cs.ep[c] = pj;
        }
        if (threadIdx.x == 0) cs.eB[c] = xb.total;
    }
    if (threadIdx.x == 0) {
        cs.N[c] = N;
        cs.kind[c] = kind;
    }
}

template <bool DM>
__global__ __launch_bounds__(kCdfThreads) void k_cdf_walk(const double2* __restrict__ a, u64 dim, u64 M,
                                                          ChunkSums cs, double* __restrict__ cdf) {
    __shared__ double P[kCdfChunk];
    __shared__ Xchg X[2];
    int r = 0;
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u64 LIM = 1ULL << 53;
    double s = 0.0;  // exact running sum, identical in every thread
    if (threadIdx.x == 0) cdf[0] = 0.0;
    // summaries of this thread's chunk in the window, loaded one window ahead
    // (the next window's loads are in flight during an exact chunk)
    double tT = 0.0, tp = 0.0;
    int tkind = 0, tE = 0, tJ = 0;
    u64 tN = 0, tA = 0, tR = 0, tB = 0;
    auto load = [&](u64 cc) {
        const u64 ci = cc + threadIdx.x;
        tT = 0.0;
        tkind = tE = 0;
        tN = 0;
        if (ci < M) {
            tT = cs.T[ci];
            tkind = cs.kind[ci];
            tE = cs.E[ci];
            tN = cs.N[ci];
            tJ = cs.ej[ci];  // meaningful for tie and crossing chunks only
            tA = cs.eA[ci];
            tR = cs.eR[ci];
            tB = cs.eB[ci];
            tp = cs.ep[ci];
        }
    };
    u64 c = 0;
    load(c);
    while (c < M) {
        const u64 ci = c + threadIdx.x;
        int es = INT_MIN;
        double u = 0.0;
        u64 S = 0;
        if (s != 0.0 && isfinite(s)) {
            es = max(ilogb(s), -1022);
            u = ldexp(1.0, es - 52);
            S = (u64)(s / u);
        }
        bool ok = false;
        u64 v = 0;  // below 2^53 (a chunk whose sum reaches it is not plain)
        if (ci < M && isfinite(s)) {
            if (tT == 0.0) {
                ok = true;
            } else if (s != 0.0 && tkind == kChunkPlain && tE == es) {
                ok = true;
                v = tN;
            }
        }
        const XScan xs = xchg_scan_first<>(v, INT_MAX, X, r);
        // The window is walked in runs: a run ends at its first chunk that is
        // not fast; a settled tie chunk there adds its increments (delta) to
        // every later chunk of the window, and the next run starts after it.
        u64 delta = 0;
        int pos = 0;
        int stop = kCdfThreads;
        int settled = 0;  // 0, kStartTie or kStartCross for the chunk at stop
        u64 Sstop = 0;
        double s_after = 0.0;
        double rsj = 0.0;  // this thread's sum through its crossing element (last run)
        for (;;) {  // wave-uniform and block-uniform
            const int t = (int)threadIdx.x;
            const u64 S0 = S + xs.excl + delta;  // exact for the chunks of the current run and its stop
            // Chunk ci is fast when every chunk of the run up to it is ok and
            // the sum stays below 2^53 units through it; the sums never
            // decrease, so the fast chunks are the first of the run.
            const bool fast = t >= pos && ok && S0 + v < LIM;
            if (fast) {
                cs.start[ci] = s == 0.0 ? 0.0 : (double)S0 * u;
                cs.exact[ci] = kStartPlain;
            }
            // the settlement of this chunk, should it be the run's stop
            int rcode = 0;
            double rend = 0.0;
            u64 rdelta = 0;
            if (t >= pos && !fast && ci < M && s != 0.0 && isfinite(s) && tE == es && tT != 0.0) {
                if (tkind == kChunkTie) {
                    // fl(x + (F + 1/2) u) rounds to the even neighbour: tie k
                    // adds F + ((sum in front of it) mod 2)
                    u64 add = 0;
                    for (int k = 0; k < tJ; ++k) add += (S0 + ((tA >> k) & 1) + add) & 1;
                    const u64 endS = S0 + tN + add;
                    if (endS < LIM) {
                        rcode = kStartTie;
                        rdelta = tN + add;
                        rend = (double)endS * u;
                    }
                } else if (tkind == kChunkCross && S0 + tA < LIM && S0 + tA + tR >= LIM) {
                    const double sj = (double)(S0 + tA) * u + tp;  // the one IEEE addition
                    if (isfinite(sj) && ilogb(sj) == es + 1) {
                        const double u2 = 2.0 * u;
                        const u64 S1 = (u64)(sj / u2);  // exact: sj is a multiple of 2u below 2^53 * 2u
                        if (S1 + tB < LIM) {
                            rcode = kStartCross;
                            rend = (double)(S1 + tB) * u2;
                            rsj = sj;
                        }
                    }
                }
            }
            // the run's stop, the sum in front of it and its settlement: from
            // the first wave holding a chunk of the run that is not fast
            {
                const u64 nf = __ballot(t >= pos && !fast);
                const int fl = nf ? __ffsll((unsigned long long)nf) - 1 : 0;
                Xchg& slot = X[r & 1];
                const u64 at = readlane_u64(S0, fl);
                const int rk = __builtin_amdgcn_readlane(rcode, fl);
                const double re = readlane_f64(rend, fl);
                const u64 rd = readlane_u64(rdelta, fl);
                if (lane == 0) {
                    slot.mn[wid] = nf ? wid * 64 + fl : INT_MAX;
                    slot.val[wid] = at;
                    slot.ok[wid] = rk;
                    slot.dv[wid] = re;
                    slot.sum2[wid] = rd;
                }
                __syncthreads();
                ++r;
            }
            const Xchg& got = X[(r - 1) & 1];
            const int tm = lane < kCdfWaves ? got.mn[lane] : INT_MAX;
            const u64 has = __ballot(tm != INT_MAX);
            stop = kCdfThreads;
            Sstop = S + xs.total + delta;
            settled = 0;
            if (has) {
                const int w = __ffsll((unsigned long long)has) - 1;
                stop = __builtin_amdgcn_readlane(tm, w);
                Sstop = readlane_u64(lane < kCdfWaves ? got.val[lane] : 0, w);
                settled = __builtin_amdgcn_readlane(lane < kCdfWaves ? got.ok[lane] : 0, w);
                s_after = readlane_f64(lane < kCdfWaves ? got.dv[lane] : 0.0, w);
                const u64 dstop = readlane_u64(lane < kCdfWaves ? got.sum2[lane] : 0, w);
                if (settled == kStartTie && c + (u64)stop < M) {
                    if (t == stop) {
                        cs.start[ci] = (double)Sstop * u;
                        cs.exact[ci] = kStartTie;
                    }
                    delta += dstop;
                    pos = stop + 1;
                    if (pos < kCdfThreads) continue;
                    // the window ends with the settled tie
                    stop = kCdfThreads;
                    Sstop = S + xs.total + delta;
                    settled = 0;
                }
            }
            break;
        }
        if (s != 0.0 && isfinite(s)) s = (double)Sstop * u;  // the sum in front of chunk c + stop
        c += (u64)stop;
        const bool handle = stop < kCdfThreads && c < M;
        if (handle && settled == kStartCross) {
            if ((int)threadIdx.x == stop) {
                cs.start[c] = s;
                cs.exact[c] = kStartCross;
                cs.ep[c] = rsj;  // the sum through the crossing element
            }
            s = s_after;
        }
        load(c + (handle ? 1 : 0));
        if (handle) {
            if (settled != kStartCross) {
                if (threadIdx.x == 0) {
                    cs.start[c] = s;
                    cs.exact[c] = kStartExact;
                }
                const int len = (int)min((u64)kCdfChunk, dim - c * kCdfChunk);
                exact_chunk<DM>(a, dim, c * kCdfChunk, len, s, cdf, P, X, r);
            }
            c += 1;
        }
    }
}

template <bool DM>
__global__ __launch_bounds__(kCdfThreads) void k_cdf_fill(const double2* __restrict__ a, u64 dim, ChunkSums cs,
                                                          double* __restrict__ cdf) {
    __shared__ u64 scan_lds[17];
    // probabilities in, CDF values out, both moved between HBM and LDS with
    // consecutive lanes on consecutive elements (padded: element i at i + i / 4)
    __shared__ double Pl[kCdfChunk + kCdfChunk / kCdfPer];
    const u64 c = blockIdx.x;
    const int code = cs.exact[c];
    if (code == kStartExact) return;
    const u64 c0 = c * kCdfChunk;
    const int len = (int)min((u64)kCdfChunk, dim - c0);
    const int lo = threadIdx.x * kCdfPer;
    const double s0 = cs.start[c];
    if (cs.T[c] == 0.0) {
        for (int k = 0; k < kCdfPer; ++k) {
            const int j = k * kCdfThreads + threadIdx.x;
            if (j < len) cdf[c0 + j + 1] = s0;
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < kCdfPer; ++k) {
        const int j = k * kCdfThreads + threadIdx.x;
        Pl[j + j / kCdfPer] = j < len ? prob_at<DM>(a, c0 + j, dim) : 0.0;
    }
    __syncthreads();
    const int e = cs.E[c];
    const double u = ldexp(1.0, e - 52);
    const u64 S = (u64)(s0 / u);
    const int jx = code == kStartCross ? cs.ej[c] : -1;
    // ties: the rounding bit of tie k follows from the parity of the sum in
    // front of it (start, increments before it rounded down, earlier bits)
    int nt = 0;
    u64 tpos = 0, tbit = 0;
    if (code == kStartTie) {
        nt = cs.ej[c];
        tpos = cs.eB[c];
        const u64 A = cs.eA[c];
        u64 add = 0;
        for (int k = 0; k < nt; ++k) {
            const u64 b = (S + ((A >> k) & 1) + add) & 1;
            tbit |= b << k;
            add += b;
        }
    }
    u64 R[kCdfPer];
    u64 mine = 0;
#pragma unroll
    for (int k = 0; k < kCdfPer; ++k) {
        const int j = lo + k;
        R[k] = 0;
        if (j < len) {
            const double pr = Pl[j + j / kCdfPer];
            if (code == kStartCross && j > jx) {
                const double q2 = ldexp(pr, 51 - e);  // next binade: units of 2u
                const double fl = floor(q2), fr = q2 - fl;
                R[k] = (u64)fl + (fr > 0.5 ? 1 : 0);
            } else if (!(code == kStartCross && j == jx)) {
                const double q = ldexp(pr, 52 - e);
                const double fl = floor(q), fr = q - fl;
                R[k] = (u64)fl + (fr > 0.5 ? 1 : 0);
                for (int k2 = 0; k2 < nt; ++k2)
                    if (j == (int)((tpos >> (12 * k2)) & 0xfff)) R[k] = (u64)fl + ((tbit >> k2) & 1);
            }
        }
        mine += R[k];
    }
    u64 total;
    u64 run = S + block_excl_scan(mine, scan_lds, &total);  // its barriers also end the reads of Pl above
    double v[kCdfPer];
    if (code == kStartCross) {
        const double sj = cs.ep[c];
        const double u2 = 2.0 * u;
        const u64 S1A = (u64)(sj / u2) - cs.eA[c] - S;  // units of 2u: S1 - A - S, added to run after the crossing
#pragma unroll
        for (int k = 0; k < kCdfPer; ++k) {
            const int j = lo + k;
            run += R[k];
            v[k] = j < jx ? (double)run * u : j == jx ? sj : (double)(run + S1A) * u2;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kCdfPer; ++k) {
            run += R[k];
            v[k] = (double)run * u;
        }
    }
#pragma unroll
    for (int k = 0; k < kCdfPer; ++k) Pl[(lo + k) + (lo + k) / kCdfPer] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kCdfPer; ++k) {
        const int j = k * kCdfThreads + threadIdx.x;
        if (j < len) cdf[c0 + j + 1] = Pl[j + j / kCdfPer];
    }
}

// Enqueues the CDF of `a` into cdf[0..dim] on st.  `work` is device scratch of
// cdf_work_bytes(dim) bytes.
size_t cdf_work_bytes(u64 dim) {
    const u64 M = (dim + kCdfChunk - 1) / kCdfChunk;
    return (size_t)M * (8 * sizeof(double) + 4 * sizeof(int)) + 1024;
}

template <bool DM>
void enqueue_cdf(const double2* a, u64 dim, double* cdf, void* work, hipStream_t st) {
    if (dim < kMultiBlockMinDim) {
        k_cdf_exact<DM><<<1, kCdfThreads, 0, st>>>(a, dim, cdf);
        return;
    }
    const u64 M = (dim + kCdfChunk - 1) / kCdfChunk;
    ChunkSums cs;
    double* d = (double*)work;
    cs.T = d;
    cs.pst = d + M;
    cs.ep = d + 2 * M;
    cs.start = d + 3 * M;
    u64* q = (u64*)(d + 4 * M);
    cs.N = q;
    cs.eA = q + M;
    cs.eR = q + 2 * M;
    cs.eB = q + 3 * M;
    int* w = (int*)(q + 4 * M);
    cs.E = w;
    cs.kind = w + M;
    cs.ej = w + 2 * M;
    cs.exact = w + 3 * M;
    k_chunk_total<DM><<<(unsigned)M, kIncrThreads, 0, st>>>(a, dim, cs.T);
    k_chunk_predict<<<1, kPredictThreads, 0, st>>>(cs, M);
    k_chunk_incr<DM><<<(unsigned)M, kIncrThreads, 0, st>>>(a, dim, cs);
    k_cdf_walk<DM><<<1, kCdfThreads, 0, st>>>(a, dim, M, cs, cdf);
    k_cdf_fill<DM><<<(unsigned)M, kCdfThreads, 0, st>>>(a, dim, cs, cdf);
}

// lower_bound(cdf[0..dim], u) - 1, as an unsigned 64-bit index (the reference's
// -1 for u <= 0 becomes 2^64 - 1 through the same conversion).
//
// Boundary check: j = lower_bound(cdf, u) is the first index with cdf[j] >= u.
// The sample j - 1 is the same for every CDF whose entries lie within delta of
// these unless u - cdf[j-1] <= delta (j >= 2; cdf[0] = 0 is exact in every CDF)
// or cdf[j] - u <= delta (1 <= j <= dim).  delta = margin * cdf[dim]; such a
// sample is flagged (flag 1), any other one is certain (flag 0).  With margin
// <= 0 nothing is flagged.
__device__ __forceinline__ unsigned char boundary_flag(const double* __restrict__ cdf, u64 dim, double u, u64 j,
                                                       double margin) {
    if (!(margin > 0.0)) return 0;
    const double delta = margin * cdf[dim];
    if (j >= 2 && u - cdf[j - 1] <= delta) return 1;
    if (j >= 1 && j <= dim && cdf[j] - u <= delta) return 1;
    return 0;
}

__device__ __forceinline__ u64 search_checked(const double* __restrict__ cdf, u64 dim, double u, double margin,
                                              unsigned char* flag) {
    u64 lo = 0, hi = dim + 1;  // search [lo, hi) for the first cdf[i] >= u
    while (lo < hi) {
        const u64 mid = lo + (hi - lo) / 2;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    if (flag) *flag = boundary_flag(cdf, dim, u, lo, margin);
    return lo - 1;
}

// The same search with one wave per uniform: each round the 64 lanes probe 64
// evenly spaced entries of the candidate range at once, so a 2^20-entry CDF
// takes four dependent rounds of loads instead of twenty-one.  Sums built by
// the sequential loop never decrease, and once a NaN appears every later sum
// is NaN, so "cdf[i] < u" holds on a prefix of the array and the first index
// where it fails is the binary search's answer.
constexpr int kSearchWaves = 4;
__global__ __launch_bounds__(64 * kSearchWaves) void k_search(const double* __restrict__ cdf, u64 dim,
                                                              const double* __restrict__ us, unsigned count,
                                                              double margin, u64* __restrict__ out,
                                                              unsigned char* __restrict__ flags) {
    const unsigned lane = threadIdx.x & 63;
    const unsigned c = blockIdx.x * kSearchWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (c >= count) return;  // wave-uniform
    const double u = us[c];
    // invariant: every index below lo has cdf < u; cdf[hi] >= u or hi = dim + 1
    u64 lo = 0, hi = dim + 1;
    while (hi - lo > 64) {
        const u64 step = (hi - lo + 63) / 64;
        const u64 idx = lo + lane * step;
        const u64 b = __ballot(idx < hi && cdf[idx] < u);
        const int k = __popcll(b);  // probes with cdf < u: lanes 0 .. k-1
        if (k == 0) {
            hi = lo;
            break;
        }
        const u64 last = lo + (u64)(k - 1) * step;
        lo = last + 1;
        if (last + step < hi) hi = last + step;
    }
    if (hi > lo) {
        const u64 idx = lo + lane;
        lo += __popcll(__ballot(idx < hi && cdf[idx] < u));
    }
    if (lane == 0) {
        out[c] = lo - 1;
        if (flags) flags[c] = boundary_flag(cdf, dim, u, lo, margin);
    }
}
inline unsigned search_blocks(unsigned count) { return (count + kSearchWaves - 1) / kSearchWaves; }

size_t sampling_buffer_bytes(u64 dim, u32 count, bool flags) {
    const size_t cdf_bytes = ((size_t)(dim + 1) * sizeof(double) + 255) & ~size_t(255);
    const size_t work_bytes = (cdf_work_bytes(dim) + 255) & ~size_t(255);
    const size_t io_bytes = ((size_t)count * 8 + 255) & ~size_t(255);
    const size_t flag_bytes = flags ? ((size_t)count + 255) & ~size_t(255) : 0;
    return cdf_bytes + work_bytes + 2 * io_bytes + flag_bytes;
}

bool enqueue_sampling_in(void* buffer, const void* state, u64 dim, const double* h_uniforms, u32 count, u64* h_out,
                         hipStream_t st, double margin, unsigned char* h_flags) {
    if (count == 0) return true;
    const size_t cdf_bytes = ((size_t)(dim + 1) * sizeof(double) + 255) & ~size_t(255);
    const size_t work_bytes = (cdf_work_bytes(dim) + 255) & ~size_t(255);
    const size_t io_bytes = ((size_t)count * 8 + 255) & ~size_t(255);
    char* buf = (char*)buffer;
    double* cdf = (double*)buf;
    double* dus = (double*)(buf + cdf_bytes + work_bytes);
    u64* dout = (u64*)((char*)dus + io_bytes);
    unsigned char* dflags = h_flags ? (unsigned char*)dout + io_bytes : nullptr;
    if (!hip_ok(hipMemcpyAsync(dus, h_uniforms, (size_t)count * 8, hipMemcpyHostToDevice, st), "enqueue_sampling"))
        return false;
    enqueue_cdf<false>((const double2*)state, dim, cdf, buf + cdf_bytes, st);
    k_search<<<search_blocks(count), 64 * kSearchWaves, 0, st>>>(cdf, dim, dus, count, margin, dout, dflags);
    if (dflags &&
        !hip_ok(hipMemcpyAsync(h_flags, dflags, count, hipMemcpyDeviceToHost, st), "enqueue_sampling"))
        return false;
    return hip_ok(hipMemcpyAsync(h_out, dout, (size_t)count * 8, hipMemcpyDeviceToHost, st), "enqueue_sampling");
}

bool enqueue_sampling(const void* state, u64 dim, const double* h_uniforms, u32 count, u64* h_out, hipStream_t st,
                      double margin, unsigned char* h_flags) {
    if (count == 0) return true;
    Scratch* sc = scratch_for(st);
    if (!sc) return false;
    void* buf = big_buffer(sc, sampling_buffer_bytes(dim, count, h_flags != nullptr));
    return buf && enqueue_sampling_in(buf, state, dim, h_uniforms, count, h_out, st, margin, h_flags);
}

// One work-group per job: the exact sequential CDF of a small state (the
// single-block walk of k_cdf_exact), then the job's searches.
struct DevSampleJob {
    const double2* state;
    u64 dim;
    double* cdf;
    const double* uniforms;
    u64* out;
    unsigned char* flags;  // nullptr: no boundary check
    double margin;
    u32 count, pad;
};

__global__ __launch_bounds__(kCdfThreads) void k_cdf_jobs(const DevSampleJob* __restrict__ jobs) {
    __shared__ double P[kCdfChunk];
    __shared__ Xchg X[2];
    int r = 0;
    const DevSampleJob jb = jobs[blockIdx.x];
    double s = 0.0;
    if (threadIdx.x == 0) jb.cdf[0] = 0.0;
    for (u64 c0 = 0; c0 < jb.dim; c0 += kCdfChunk) {
        const int len = (int)min((u64)kCdfChunk, jb.dim - c0);
        exact_chunk<false>(jb.state, jb.dim, c0, len, s, jb.cdf, P, X, r);
    }
    __syncthreads();
    for (u32 c = threadIdx.x; c < jb.count; c += blockDim.x)
        jb.out[c] = search_checked(jb.cdf, jb.dim, jb.uniforms[c], jb.margin, jb.flags ? jb.flags + c : nullptr);
}

bool enqueue_sampling_jobs(const SampleJob* jobs, u32 njobs, hipStream_t st) {
    if (njobs == 0) return true;
    Scratch* sc = scratch_for(st);
    if (!sc) return false;
    size_t cdf_bytes = 0, total = 0;
    bool contiguous = true;  // host uniforms, outputs and flags laid out job after job
    bool any_flags = false;
    for (u32 j = 0; j < njobs; ++j) {
        if (jobs[j].dim > kMultiBlockMinDim) {
            set_error(QM_ERR_INVALID_ARGUMENT, "enqueue_sampling_jobs: state of %llu amplitudes is too large", jobs[j].dim);
            return false;
        }
        cdf_bytes += ((size_t)(jobs[j].dim + 1) * 8 + 255) & ~size_t(255);
        if (j && (jobs[j].h_uniforms != jobs[j - 1].h_uniforms + jobs[j - 1].count ||
                  jobs[j].h_out != jobs[j - 1].h_out + jobs[j - 1].count ||
                  (jobs[j].h_flags != nullptr) != (jobs[j - 1].h_flags != nullptr) ||
                  (jobs[j].h_flags && jobs[j].h_flags != jobs[j - 1].h_flags + jobs[j - 1].count)))
            contiguous = false;
        any_flags = any_flags || jobs[j].h_flags != nullptr;
        total += jobs[j].count;
    }
    const size_t desc_bytes = ((size_t)njobs * sizeof(DevSampleJob) + 255) & ~size_t(255);
    const size_t io_bytes = (total * 8 + 255) & ~size_t(255);
    const size_t flag_bytes = any_flags ? (total + 255) & ~size_t(255) : 0;
    char* buf = (char*)big_buffer(sc, desc_bytes + cdf_bytes + 2 * io_bytes + flag_bytes);
    if (!buf) return false;
    double* dus = (double*)(buf + desc_bytes + cdf_bytes);
    u64* dout = (u64*)((char*)dus + io_bytes);
    unsigned char* dflags = any_flags ? (unsigned char*)dout + io_bytes : nullptr;
    std::vector<DevSampleJob> d(njobs);
    size_t off = desc_bytes, io = 0;
    for (u32 j = 0; j < njobs; ++j) {
        d[j].state = jobs[j].state;
        d[j].dim = jobs[j].dim;
        d[j].count = jobs[j].count;
        d[j].cdf = (double*)(buf + off);
        off += ((size_t)(jobs[j].dim + 1) * 8 + 255) & ~size_t(255);
        d[j].uniforms = dus + io;
        d[j].out = dout + io;
        d[j].flags = jobs[j].h_flags ? dflags + io : nullptr;
        d[j].margin = jobs[j].margin;
        if (!contiguous && jobs[j].count &&
            !hip_ok(hipMemcpyAsync(dus + io, jobs[j].h_uniforms, (size_t)jobs[j].count * 8, hipMemcpyHostToDevice, st),
                    "enqueue_sampling_jobs"))
            return false;
        io += jobs[j].count;
    }
    if (contiguous && total &&
        !hip_ok(hipMemcpyAsync(dus, jobs[0].h_uniforms, total * 8, hipMemcpyHostToDevice, st), "enqueue_sampling_jobs"))
        return false;
    if (!hip_ok(hipMemcpyAsync(buf, d.data(), (size_t)njobs * sizeof(DevSampleJob), hipMemcpyHostToDevice, st),
                "enqueue_sampling_jobs"))
        return false;
    k_cdf_jobs<<<njobs, kCdfThreads, 0, st>>>((const DevSampleJob*)buf);
    if (contiguous) {
        if (!total) return true;
        if (jobs[0].h_flags &&
            !hip_ok(hipMemcpyAsync(jobs[0].h_flags, dflags, total, hipMemcpyDeviceToHost, st), "enqueue_sampling_jobs"))
            return false;
        return hip_ok(hipMemcpyAsync(jobs[0].h_out, dout, total * 8, hipMemcpyDeviceToHost, st), "enqueue_sampling_jobs");
    }
    for (u32 j = 0; j < njobs; ++j) {
        if (!jobs[j].count) continue;
        if (jobs[j].h_flags && !hip_ok(hipMemcpyAsync(jobs[j].h_flags, d[j].flags, jobs[j].count,
                                                      hipMemcpyDeviceToHost, st),
                                       "enqueue_sampling_jobs"))
            return false;
        if (!hip_ok(hipMemcpyAsync(jobs[j].h_out, d[j].out, (size_t)jobs[j].count * 8, hipMemcpyDeviceToHost, st),
                    "enqueue_sampling_jobs"))
            return false;
    }
    return true;
}

template <bool DM>
void sample(const double* uniforms, unsigned count, ITYPE* out, void* state, u64 dim, void* stream, unsigned dev,
            const char* what, double margin = 0.0, unsigned char* flags = nullptr, UINT* flagged = nullptr) {
    if (flagged) *flagged = 0;
    if (!state || log2_dim(dim) < 0) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null state or bad dim", what);
        return;
    }
    if (count == 0) return;
    if (!uniforms || !out) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null uniforms/out", what);
        return;
    }
    if (!(margin >= 0.0) || margin >= 1.0) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: margin %g outside [0, 1)", what, margin);
        return;
    }
    if (!use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    Scratch* sc = scratch_for(st);
    if (!sc) return;
    const size_t cdf_bytes = ((size_t)(dim + 1) * sizeof(double) + 255) & ~size_t(255);
    const size_t work_bytes = (cdf_work_bytes(dim) + 255) & ~size_t(255);
    const size_t io_bytes = ((size_t)count * 8 + 255) & ~size_t(255);
    const size_t flag_bytes = flags ? ((size_t)count + 255) & ~size_t(255) : 0;
    char* buf = (char*)big_buffer(sc, cdf_bytes + work_bytes + 2 * io_bytes + flag_bytes);
    if (!buf) return;
    double* cdf = (double*)buf;
    void* work = buf + cdf_bytes;
    double* dus = (double*)(buf + cdf_bytes + work_bytes);
    u64* dout = (u64*)((char*)dus + io_bytes);
    unsigned char* dflags = flags ? (unsigned char*)dout + io_bytes : nullptr;
    if (!hip_ok(hipMemcpyAsync(dus, uniforms, (size_t)count * 8, hipMemcpyHostToDevice, st), what)) return;
    enqueue_cdf<DM>((const double2*)state, dim, cdf, work, st);
    k_search<<<search_blocks(count), 64 * kSearchWaves, 0, st>>>(cdf, dim, dus, count, margin, dout, dflags);
    if (!hip_ok(hipMemcpyAsync(out, dout, (size_t)count * 8, hipMemcpyDeviceToHost, st), what)) return;
    if (dflags && !hip_ok(hipMemcpyAsync(flags, dflags, count, hipMemcpyDeviceToHost, st), what)) return;
    if (!hip_ok(hipStreamSynchronize(st), what)) return;
    if (flagged && flags) {
        UINT k = 0;
        for (unsigned c = 0; c < count; ++c) k += flags[c] ? 1u : 0u;
        *flagged = k;
    }
}

template <bool DM>
void cdf_to_host(double* cdf_out, void* state, u64 dim, void* stream, unsigned dev, const char* what) {
    if (!state || !cdf_out || log2_dim(dim) < 0) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null state/output or bad dim", what);
        return;
    }
    if (!use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    Scratch* sc = scratch_for(st);
    if (!sc) return;
    const size_t cdf_bytes = ((size_t)(dim + 1) * sizeof(double) + 255) & ~size_t(255);
    char* buf = (char*)big_buffer(sc, cdf_bytes + cdf_work_bytes(dim));
    if (!buf) return;
    enqueue_cdf<DM>((const double2*)state, dim, (double*)buf, buf + cdf_bytes, st);
    if (!hip_ok(hipMemcpyAsync(cdf_out, buf, (size_t)(dim + 1) * sizeof(double), hipMemcpyDeviceToHost, st), what))
        return;
    hip_ok(hipStreamSynchronize(st), what);
}

}  // namespace qm

using namespace qm;

extern "C" {

void qm_sampling_host(const double* uniforms, UINT count, ITYPE* out, void* state, ITYPE dim, void* stream,
                      unsigned dev) {
    sample<false>(uniforms, count, out, state, dim, stream, dev, "qm_sampling_host");
}

void qm_sampling_checked_host(const double* uniforms, UINT count, double margin, ITYPE* out, unsigned char* flags,
                              UINT* flagged, void* state, ITYPE dim, void* stream, unsigned dev) {
    if (!flags) {
        set_error(QM_ERR_INVALID_ARGUMENT, "qm_sampling_checked_host: null flags");
        return;
    }
    sample<false>(uniforms, count, out, state, dim, stream, dev, "qm_sampling_checked_host", margin, flags, flagged);
}

void qm_sampling_cdf_host(double* cdf_out, void* state, ITYPE dim, void* stream, unsigned dev) {
    cdf_to_host<false>(cdf_out, state, dim, stream, dev, "qm_sampling_cdf_host");
}

void dm_sampling_host(const double* uniforms, UINT count, ITYPE* out, void* state, ITYPE dim, void* stream,
                      unsigned dev) {
    sample<true>(uniforms, count, out, state, dim, stream, dev, "dm_sampling_host");
}

void dm_sampling_checked_host(const double* uniforms, UINT count, double margin, ITYPE* out, unsigned char* flags,
                              UINT* flagged, void* state, ITYPE dim, void* stream, unsigned dev) {
    if (!flags) {
        set_error(QM_ERR_INVALID_ARGUMENT, "dm_sampling_checked_host: null flags");
        return;
    }
    sample<true>(uniforms, count, out, state, dim, stream, dev, "dm_sampling_checked_host", margin, flags, flagged);
}

// A gate's update of an amplitude pair (a 2x2 complex product in FMAs) rounds
// each result to within a few eps of its magnitude, and unitary gates carry
// earlier errors forward without growth, so two evaluations of G gates differ
// by ||psi_a - psi_b|| = O(sqrt(G) eps) when the roundings are independent
// (O(G eps) in the worst case).  A CDF entry moves by at most
// sum_j | |a_j|^2 - |b_j|^2 | <= 2 ||psi_a - psi_b|| plus a few units in the
// last place of the running sum.  Measured on brickwork circuits (fused
// executor, gate-by-gate device path and the CPU loop against each other, 16 to
// 26 qubits, 900 to 4000 gates), max_i |cdf_a[i] - cdf_b[i]| stays below
// 25 sqrt(G) eps; the margin is 512 sqrt(G) eps + 64 eps.
double qm_sampling_margin(UINT gate_count) {
    const double eps = 0x1p-53;
    return 512.0 * std::sqrt((double)gate_count) * eps + 64.0 * eps;
}

void dm_sampling_cdf_host(double* cdf_out, void* state, ITYPE dim, void* stream, unsigned dev) {
    cdf_to_host<true>(cdf_out, state, dim, stream, dev, "dm_sampling_cdf_host");
}

}  // extern "C"
