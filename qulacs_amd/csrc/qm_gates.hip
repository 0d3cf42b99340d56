// State-vector gate kernels for gfx950.
//
// One kernel shape covers every gate that acts on a fixed set of k <= 5 target
// qubits (dense, controlled, diagonal, named permutations): each thread owns a
// "base" index b whose target and control bits are zero (control bits then set
// to their required values), gathers the 2^k amplitudes b + spread(x) into
// registers, applies the gate's operation and scatters back only what changed.
// Consecutive lanes take consecutive bases, i.e. they vary the lowest bits that
// are neither target nor control: when no target is below bit 6 every wave
// load is a contiguous 1 KiB run; with a or two low targets a load touches 2^a
// KiB with 1/2^a use and its partner loads (issued back to back by the same
// wave) use the rest of the same cache lines, so HBM sees each byte once.
//
// Reference loops restated: src/csim/update_ops_matrix_dense_single.cpp:44-68,
// update_ops_matrix_dense_double.cpp:64-113, update_ops_matrix_dense_multi.cpp:59-128,
// update_ops_control_*.cpp, update_ops_named_{X,Y,Z,H,CNOT,CZ,SWAP}.cpp,
// update_ops_named_projection.cpp:38-72, update_ops_matrix_diagonal_*.cpp,
// update_ops_matrix_phase_single.cpp:14-60.
#include <cmath>
#include <complex>

#include "qm_internal.hpp"

namespace qm {

// ------------------------------------------------------------------ launch plumbing
constexpr int kMaxIns = 24;  // targets + controls per gather gate

template <int K>
struct GatherArgs {
    u64 low[kMaxIns];       // (1 << p) - 1 for the ascending insert positions p (targets and controls)
    unsigned char ins0;     // lowest insert position (decides streaming vs cached accesses)
    int n_ins;
    u64 cmask;              // control bits that must be 1
    u64 offs[1 << K];       // offset of tuple member x: sum of 2^targets[i] over set bits i of x
};

template <int K>
bool fill_args(GatherArgs<K>& a, const u32* targets, const u32* controls, const u32* cvals, u32 nc) {
    const GatherPlan p = make_plan(targets, K, controls, cvals, nc);
    if (p.n_ins > kMaxIns) {
        set_error(QM_ERR_UNSUPPORTED, "more than %d target+control qubits", kMaxIns);
        return false;
    }
    for (int i = 0; i < kMaxIns; ++i) a.low[i] = i < p.n_ins ? (1ULL << p.ins[i]) - 1 : 0;
    a.ins0 = p.n_ins ? p.ins[0] : 63;
    a.n_ins = p.n_ins;
    a.cmask = p.control_mask;
    for (int x = 0; x < (1 << K); ++x) {
        u64 o = 0;
        for (int i = 0; i < K; ++i)
            if ((x >> i) & 1) o |= 1ULL << targets[i];
        a.offs[x] = o;
    }
    return true;
}

// Constant indices under a wave-uniform guard keep the masks in scalar registers.
template <int K>
__device__ __forceinline__ u64 base_of(u64 i, const GatherArgs<K>& a) {
#pragma unroll
    for (int k = 0; k < kMaxIns; ++k)
        if (k < a.n_ins) {
            const u64 lo = a.low[k];
            i = (i & lo) | ((i & ~lo) << 1);
        }
    return i | a.cmask;
}

// Op interface: READ / WRITE are bit masks over tuple members; apply() maps
// the registers in place.
template <int K, int U, bool NT, class Op>
__global__ __launch_bounds__(512) void k_gather(double2* __restrict__ s, u64 n_bases, GatherArgs<K> a, Op op) {
    constexpr int N = 1 << K;
    const u64 stride = (u64)gridDim.x * blockDim.x * U;
    for (u64 i0 = ((u64)blockIdx.x * blockDim.x) * U + threadIdx.x; i0 < n_bases; i0 += stride) {
        double2 v[U][N];
        u64 b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const u64 i = i0 + (u64)u * blockDim.x;
            b[u] = i < n_bases ? base_of<K>(i, a) : ~0ULL;
#pragma unroll
            for (int x = 0; x < N; ++x)
                if ((Op::READ >> x) & 1)
                    if (b[u] != ~0ULL) v[u][x] = ld<NT>(s + b[u] + a.offs[x]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (b[u] == ~0ULL) continue;
            op.apply(v[u]);
#pragma unroll
            for (int x = 0; x < N; ++x)
                if ((Op::WRITE >> x) & 1) st<NT>(s + b[u] + a.offs[x], v[u][x]);
        }
    }
}

// ------------------------------------------------------------------ operations
template <int K>
struct MatArg {
    double2 m[1 << (2 * K)];
};

template <int K>
struct OpDenseArg {  // matrix passed in the kernel arguments (K <= 3)
    static constexpr unsigned READ = (1u << (1 << K)) - 1;
    static constexpr unsigned WRITE = READ;
    MatArg<K> mat;
    __device__ __forceinline__ void apply(double2* v) const {
        constexpr int N = 1 << K;
        double2 out[N];
#pragma unroll
        for (int y = 0; y < N; ++y) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int x = 0; x < N; ++x) acc = cmac(acc, mat.m[y * N + x], v[x]);
            out[y] = acc;
        }
#pragma unroll
        for (int y = 0; y < N; ++y) v[y] = out[y];
    }
};

// Matrix sources for the streaming dense kernel.
template <int K>
struct MatInArg {  // K <= 3: the whole matrix rides in the kernel arguments
    MatArg<K> mat;
    __device__ __forceinline__ double2 at(int y, int x) const { return mat.m[y * (1 << K) + x]; }
};
template <int K>
struct MatInPtr {  // K = 4, 5: device copy, uniform across the wave (scalar loads)
    const double2* __restrict__ p;
    __device__ __forceinline__ double2 at(int y, int x) const { return p[y * (1 << K) + x]; }
};

// Dense k-qubit gate (k = 3..5): the tuple sits in registers and each output
// row is written as soon as it is formed, so live registers stay near 4 * 2^k.
template <int K, bool NT, class M>
__global__ __launch_bounds__(256) void k_dense_stream(double2* __restrict__ s, u64 n_bases, GatherArgs<K> a, M mat) {
    constexpr int N = 1 << K;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n_bases; i += stride) {
        const u64 b = base_of<K>(i, a);
        double2 v[N];
#pragma unroll
        for (int x = 0; x < N; ++x) v[x] = ld<NT>(s + b + a.offs[x]);
#pragma unroll
        for (int y = 0; y < N; ++y) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int x = 0; x < N; ++x) acc = cmac(acc, mat.at(y, x), v[x]);
            st<NT>(s + b + a.offs[y], acc);
        }
    }
}

// Dense k-qubit gate (k = 4, 5) on the FP64 matrix cores.
//
// For 16 consecutive tuples (the columns of X, 2^k amplitudes each) the update
// Y = M X is the real product [Yr; Yi] = [[Mr, -Mi], [Mi, Mr]] [Xr; Xi], issued
// as v_mfma_f64_16x16x4f64 blocks: 16 output rows x 16 tuples per block, 4 rows
// of X per step.  Lane layout of that instruction (checked on gfx950): lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16] and receives
// D[l / 16 + 4 v][l % 16] in item v.  So lane l owns tuple l % 16 of the group,
// loads its amplitudes x = 4 q + l / 16 (q < 2^k / 4) and stores outputs
// y = 16 r + l / 16 + 4 v.  The matrix fragments stay in VGPRs for the whole
// kernel; each wave walks groups of 16 tuples.
typedef double qm_d4 __attribute__((ext_vector_type(4)));

// Compact arguments (few scalar registers): insert positions packed one per
// byte, target positions, control bits.  Member offsets are formed from the
// target positions with scalar shifts where they are needed.
template <int K>
struct PackedArgs {
    u64 pos[3];              // ascending insert positions, byte k of pos[k / 8]
    int n_ins;
    u64 cmask;
    unsigned char tg[K];     // target positions, bit i of a member index <-> tg[i]
};

template <int K>
__device__ __forceinline__ u64 base_of(u64 i, const PackedArgs<K>& a) {
#pragma unroll
    for (int k = 0; k < kMaxIns; ++k)
        if (k < a.n_ins) {
            const unsigned p = (unsigned)(a.pos[k >> 3] >> (8 * (k & 7))) & 0xffu;
            const u64 lo = (1ULL << p) - 1;
            i = (i & lo) | ((i & ~lo) << 1);
        }
    return i | a.cmask;
}
// offset of member x (bits of x spread over the targets)
template <int K>
__device__ __forceinline__ u64 member_off(unsigned x, const PackedArgs<K>& a) {
    u64 o = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) o |= (u64)((x >> i) & 1u) << a.tg[i];
    return o;
}

template <int K, bool NT>
__global__ __launch_bounds__(256) void k_dense_mfma(double2* __restrict__ s, u64 n_bases, PackedArgs<K> a,
                                                    const double2* __restrict__ mat) {
    constexpr int N = 1 << K;
    constexpr int RT = N / 16;  // 16-row output tiles
    constexpr int KS = N / 4;   // 4-row steps over one real half of X
    const int lane = threadIdx.x & 63;
    const int col = lane & 15;  // A row / tuple column
    const int kq = lane >> 4;   // k within a step
    // A fragments in LDS, stored in lane order so each wave's read of one
    // fragment is 64 consecutive double2 (no bank conflicts); this keeps the
    // matrix out of VGPRs and leaves room for more resident waves.
    __shared__ double2 frag[RT * KS * 64];
    for (int e = threadIdx.x; e < RT * KS * 64; e += blockDim.x) {
        const int l = e & 63, q = (e >> 6) % KS, rt = (e >> 6) / KS;
        frag[e] = mat[(16 * rt + (l & 15)) * N + 4 * q + (l >> 4)];
    }
    __syncthreads();
    const u64 groups = (n_bases + 15) >> 4;
    const u64 wave = (u64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const u64 nwaves = (u64)gridDim.x * (blockDim.x >> 6);
    // Member x = 4 q + kq has offset offs[4 q] + offs[kq] (disjoint target bits):
    // the first term is wave-uniform (scalar), the second is one per-lane value.
    const u64 lane_off = member_off<K>((unsigned)kq, a);
    // Software pipeline: the next group's loads are in flight while this
    // group's MFMAs run (waves are persistent, so the matrix is loaded once).
    // Lanes past the end load a valid duplicate tuple and do not store.
    double nr[KS], ni[KS];
    u64 nb = 0;
    bool nvalid = false;
    auto fetch = [&](u64 g) {
        const u64 i = (g << 4) + col;
        nvalid = i < n_bases;
        nb = base_of<K>(nvalid ? i : n_bases - 1, a) + lane_off;
#pragma unroll
        for (int q = 0; q < KS; ++q) {
            const double2 v = ld<NT>(s + nb + member_off<K>(4u * q, a));
            nr[q] = v.x;
            ni[q] = v.y;
        }
    };
    if (wave < groups) fetch(wave);
    for (u64 g = wave; g < groups; g += nwaves) {  // wave-uniform loop: every lane runs every MFMA
        double br[KS], bi[KS];
#pragma unroll
        for (int q = 0; q < KS; ++q) {
            br[q] = nr[q];
            bi[q] = ni[q];
        }
        const u64 b = nb;
        const bool valid = nvalid;
        if (g + nwaves < groups) fetch(g + nwaves);
        qm_d4 dr[RT], di[RT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            dr[rt] = qm_d4{0.0, 0.0, 0.0, 0.0};
            di[rt] = qm_d4{0.0, 0.0, 0.0, 0.0};
        }
        // 2 * RT independent accumulators between dependent MFMAs
#pragma unroll
        for (int q = 0; q < KS; ++q) {
            const double nbi = -bi[q];  // Yr -= Mi Xi, negated on the B side
            double2 m[RT];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) m[rt] = frag[(rt * KS + q) * 64 + lane];  // (Mr, Mi) fragment
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                dr[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(m[rt].x, br[q], dr[rt], 0, 0, 0);
                di[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(m[rt].y, br[q], di[rt], 0, 0, 0);
            }
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                dr[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(m[rt].y, nbi, dr[rt], 0, 0, 0);
                di[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(m[rt].x, bi[q], di[rt], 0, 0, 0);
            }
        }
        if (valid) {
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    st<NT>(s + b + member_off<K>(16u * rt + 4u * v, a), make_double2(dr[rt][v], di[rt][v]));
        }
    }
}

struct OpSwap01 {  // X, CNOT
    static constexpr unsigned READ = 3, WRITE = 3;
    __device__ __forceinline__ void apply(double2* v) const {
        const double2 t = v[0];
        v[0] = v[1];
        v[1] = t;
    }
};
struct OpSwap12 {  // SWAP (members 01 <-> 10)
    static constexpr unsigned READ = 6, WRITE = 6;
    __device__ __forceinline__ void apply(double2* v) const {
        const double2 t = v[1];
        v[1] = v[2];
        v[2] = t;
    }
};
struct OpY {  // (a0, a1) -> (-i a1, i a0)
    static constexpr unsigned READ = 3, WRITE = 3;
    __device__ __forceinline__ void apply(double2* v) const {
        const double2 a0 = v[0], a1 = v[1];
        v[0] = make_double2(a1.y, -a1.x);
        v[1] = make_double2(-a0.y, a0.x);
    }
};
struct OpH {
    static constexpr unsigned READ = 3, WRITE = 3;
    double r;
    __device__ __forceinline__ void apply(double2* v) const {
        const double2 a0 = v[0], a1 = v[1];
        v[0] = make_double2((a0.x + a1.x) * r, (a0.y + a1.y) * r);
        v[1] = make_double2((a0.x - a1.x) * r, (a0.y - a1.y) * r);
    }
};
struct OpNeg1 {  // Z, CZ: negate member 1
    static constexpr unsigned READ = 2, WRITE = 2;
    __device__ __forceinline__ void apply(double2* v) const { v[1] = make_double2(-v[1].x, -v[1].y); }
};
struct OpPhase1 {  // S, T, phase gate: member 1 *= phase
    static constexpr unsigned READ = 2, WRITE = 2;
    double2 ph;
    __device__ __forceinline__ void apply(double2* v) const { v[1] = cmul(v[1], ph); }
};
struct OpZero0 {  // P1 projection zeroes member 0
    static constexpr unsigned READ = 0, WRITE = 1;
    __device__ __forceinline__ void apply(double2* v) const { v[0] = make_double2(0.0, 0.0); }
};
struct OpZero1 {  // P0 projection zeroes member 1
    static constexpr unsigned READ = 0, WRITE = 2;
    __device__ __forceinline__ void apply(double2* v) const { v[1] = make_double2(0.0, 0.0); }
};
template <int K>
struct OpDiag {  // member x *= d[x]
    static constexpr unsigned READ = (K >= 5) ? 0xFFFFFFFFu : ((1u << (1 << K)) - 1);
    static constexpr unsigned WRITE = READ;
    double2 d[1 << K];
    __device__ __forceinline__ void apply(double2* v) const {
#pragma unroll
        for (int x = 0; x < (1 << K); ++x) v[x] = cmul(v[x], d[x]);
    }
};

// ------------------------------------------------------------------ low-qubit gates
// Gates with a target or control below bit 3, i.e. inside one 128-byte line.
// The gather kernel gives each lane one tuple, so there each access is 16 bytes
// at a 32..64-byte stride, a line is visited by several instructions and only
// the changed 16-byte pieces are written back.  Here consecutive lanes take
// consecutive amplitudes instead: every wave instruction reads or writes whole
// lines (eight lanes per line), a lane fetches the members of its tuple that
// share its line from the lanes holding them (__shfl), and every line the gate
// touches is written back whole, unchanged amplitudes included, with streaming
// accesses.  Tuple bits in LM are the targets below bit 3 (lane bits); the
// other targets are held in registers as in the gather kernel.  Controls below
// bit 3 select per lane between the new and the old value; controls from bit 3
// up are fixed bits of the enumeration, as are the high targets (inserted
// zeros).
template <int K>
struct LaneArgs {
    u64 low[kMaxIns];      // (1 << p) - 1 for the ascending positions p >= 3 of high targets and controls
    int n_ins;
    u64 cmask;             // controls >= 3 that must be 1
    u64 offs[1 << K];      // member offsets from the high targets (entries with no LM bit)
    unsigned char pos[K];  // target position of each tuple bit
    u32 lc_mask, lc_val;   // controls below bit 3: a lane takes the new value iff (j & lc_mask) == lc_val
};

__device__ __forceinline__ double2 shfl2(double2 v, int src) {
    return make_double2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// Whether the lines of high member combination x (no LM bit) are touched:
// some member sharing them is read or written.
template <int K, unsigned LM, class Op>
__host__ __device__ constexpr bool lane_touched(int x) {
    for (int l = 0; l < (1 << K); ++l)
        if ((l & ~(int)LM) == 0 && ((((Op::READ | Op::WRITE) >> (l | x)) & 1) != 0)) return true;
    return false;
}

// U items per thread, all loaded before any is used.  Requires n_items to be a
// multiple of 64: whether an item exists is then uniform across a wave, so
// every lane takes part in every shuffle.
template <int K, unsigned LM, int U, class Op>
__global__ __launch_bounds__(512) void k_lane(double2* __restrict__ s, u64 n_items, LaneArgs<K> a, Op op) {
    constexpr int N = 1 << K;
    const u64 stride = (u64)gridDim.x * blockDim.x * U;
    const int lane = threadIdx.x & 63;
    for (u64 i0 = (u64)blockIdx.x * blockDim.x * U + threadIdx.x; i0 < n_items; i0 += stride) {
        u64 j[U];
        unsigned own[U];
        double2 h[U][N];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            u64 jj = i0 + (u64)u * blockDim.x;
            if (jj >= n_items) continue;  // wave-uniform
#pragma unroll
            for (int k = 0; k < kMaxIns; ++k)
                if (k < a.n_ins) {
                    const u64 lo = a.low[k];
                    jj = (jj & lo) | ((jj & ~lo) << 1);
                }
            jj |= a.cmask;
            j[u] = jj;
            // this lane's tuple bits below bit 3: its own member within the line group
            unsigned o = 0;
#pragma unroll
            for (int t = 0; t < K; ++t)
                if ((LM >> t) & 1) o |= (unsigned)((jj >> a.pos[t]) & 1) << t;
            own[u] = o;
#pragma unroll
            for (int x = 0; x < N; ++x)
                if ((x & LM) == 0 && lane_touched<K, LM, Op>(x)) h[u][x] = ld_stream(s + jj + a.offs[x]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (i0 + (u64)u * blockDim.x >= n_items) continue;  // wave-uniform
            double2 v[N];
#pragma unroll
            for (int x = 0; x < N; ++x) {
                v[x] = make_double2(0.0, 0.0);
                if (!((Op::READ >> x) & 1)) continue;
                // member x sits in the lane whose low tuple bits are x & LM
                const unsigned d = own[u] ^ (x & LM);
                int dl = 0;
#pragma unroll
                for (int t = 0; t < K; ++t)
                    if ((LM >> t) & 1) dl |= (int)((d >> t) & 1) << a.pos[t];
                v[x] = shfl2(h[u][x & ~LM], lane ^ dl);
            }
            op.apply(v);
            const bool ctrl = ((u32)j[u] & a.lc_mask) == a.lc_val;
#pragma unroll
            for (int x = 0; x < N; ++x) {
                if ((x & LM) != 0 || !lane_touched<K, LM, Op>(x)) continue;
                double2 out = v[x];
                bool wr = (Op::WRITE >> x) & 1;
#pragma unroll
                for (int l = 1; l < N; ++l)
                    if ((l & ~LM) == 0 && own[u] == (unsigned)l) {
                        out = v[l | x];
                        wr = (Op::WRITE >> (l | x)) & 1;
                    }
                st_stream(s + j[u] + a.offs[x], (ctrl && wr) ? out : h[u][x]);
            }
        }
    }
}

// Sets up k_lane for a gate with K targets; returns false when the gate has no
// qubit below bit 3 or the state is too small (fewer than 64 work items).
template <int K>
bool fill_lane_args(LaneArgs<K>& a, unsigned& lm, u64& n_items, u64 dim, const u32* targets, const u32* controls,
                    const u32* cvals, u32 nc) {
    u64 hi_bits = 0;
    bool any_low = false;
    lm = 0;
    a.cmask = 0;
    a.lc_mask = a.lc_val = 0;
    for (int t = 0; t < K; ++t) {
        a.pos[t] = (unsigned char)targets[t];
        if (targets[t] < kLineQubits) {
            lm |= 1u << t;
            any_low = true;
        } else {
            hi_bits |= 1ULL << targets[t];
        }
    }
    for (u32 c = 0; c < nc; ++c) {
        if (controls[c] < kLineQubits) {
            any_low = true;
            a.lc_mask |= 1u << controls[c];
            if (cvals[c]) a.lc_val |= 1u << controls[c];
        } else {
            hi_bits |= 1ULL << controls[c];
            if (cvals[c]) a.cmask |= 1ULL << controls[c];
        }
    }
    if (!any_low) return false;
    a.n_ins = 0;
    for (int b = 0; b < 64; ++b)
        if ((hi_bits >> b) & 1) {
            if (a.n_ins == kMaxIns) return false;
            a.low[a.n_ins++] = (1ULL << b) - 1;
        }
    for (int k = a.n_ins; k < kMaxIns; ++k) a.low[k] = 0;
    n_items = dim >> a.n_ins;
    if (n_items < 64) return false;
    for (int x = 0; x < (1 << K); ++x) {
        u64 o = 0;
        for (int t = 0; t < K; ++t)
            if (((x >> t) & 1) && !((lm >> t) & 1)) o |= 1ULL << targets[t];
        a.offs[x] = o;
    }
    return true;
}

template <int K, class Op>
bool launch_lane(double2* s, u64 dim, const u32* targets, const u32* controls, const u32* cvals, u32 nc,
                 const Op& op, hipStream_t st) {
    LaneArgs<K> a;
    unsigned lm = 0;
    u64 n_items = 0;
    if (!fill_lane_args<K>(a, lm, n_items, dim, targets, controls, cvals, nc)) return false;
    constexpr int U = 2;
    const unsigned block = 512;
    const unsigned grid = grid_for(n_items, (u64)block * U);
    switch (lm) {
        case 0: k_lane<K, 0, U, Op><<<grid, block, 0, st>>>(s, n_items, a, op); break;
        case 1: k_lane<K, 1, U, Op><<<grid, block, 0, st>>>(s, n_items, a, op); break;
        case 2: if constexpr (K >= 2) k_lane<K, 2, U, Op><<<grid, block, 0, st>>>(s, n_items, a, op); break;
        default: if constexpr (K >= 2) k_lane<K, 3, U, Op><<<grid, block, 0, st>>>(s, n_items, a, op); break;
    }
    return true;
}

// ------------------------------------------------------------------ generic driver
template <int K, int U, class Op>
void launch_gather(double2* s, u64 dim, const u32* targets, const u32* controls, const u32* cvals, u32 nc,
                   const Op& op, hipStream_t st) {
    // The line kernel pays off where the gather kernel would touch lines
    // partially: operations that read or write only some tuple members (CZ,
    // phases, projections) and controls below bit 3.  A full dense update with
    // a low target streams whole lines through the gather kernel's partner loads
    // already, and the line kernel would repeat its arithmetic per lane.
    constexpr unsigned kAll = (1u << (1 << K)) - 1;
    bool low_control = false;
    for (u32 c = 0; c < nc; ++c) low_control |= controls[c] < kLineQubits;
    if (K <= 2 && (Op::READ != kAll || Op::WRITE != kAll || low_control) &&
        launch_lane<K>(s, dim, targets, controls, cvals, nc, op, st))
        return;
    GatherArgs<K> a;
    if (!fill_args<K>(a, targets, controls, cvals, nc)) return;
    // An operation that reads one member per tuple (Z, CZ, phases, projections)
    // takes two tuples per thread, so a thread keeps as many loads in flight as
    // a single-qubit dense update does (30 qubits: CZ 1.85 -> 1.54-1.59 ms, Z and
    // S 3.6-3.7 -> 3.0-3.1 ms; four or eight tuples measured no better).
    constexpr int UU = __builtin_popcount(Op::READ) == 1 ? 2 * U : U;
    const u64 n_bases = dim >> (K + nc);
    const unsigned block = 512;
    const unsigned grid = grid_for(n_bases, (u64)block * UU);
    if (a.ins0 >= kLineQubits)
        k_gather<K, UU, true, Op><<<grid, block, 0, st>>>(s, n_bases, a, op);
    else
        k_gather<K, UU, false, Op><<<grid, block, 0, st>>>(s, n_bases, a, op);
}

template <int K, class M>
void launch_dense_stream(double2* s, u64 dim, const u32* targets, const u32* controls, const u32* cvals, u32 nc,
                         const M& m, hipStream_t st) {
    GatherArgs<K> a;
    if (!fill_args<K>(a, targets, controls, cvals, nc)) return;
    const u64 n_bases = dim >> (K + nc);
    if (a.ins0 >= kLineQubits)
        k_dense_stream<K, true, M><<<grid_for(n_bases, 256), 256, 0, st>>>(s, n_bases, a, m);
    else
        k_dense_stream<K, false, M><<<grid_for(n_bases, 256), 256, 0, st>>>(s, n_bases, a, m);
}

template <int K>
void launch_dense_mfma(double2* s, u64 dim, const u32* targets, const u32* controls, const u32* cvals, u32 nc,
                       const double2* dmat, hipStream_t st) {
    const GatherPlan p = make_plan(targets, K, controls, cvals, nc);
    if (p.n_ins > kMaxIns) {
        set_error(QM_ERR_UNSUPPORTED, "more than %d target+control qubits", kMaxIns);
        return;
    }
    PackedArgs<K> a;
    a.pos[0] = a.pos[1] = a.pos[2] = 0;
    for (int k = 0; k < p.n_ins; ++k) a.pos[k >> 3] |= (u64)p.ins[k] << (8 * (k & 7));
    a.n_ins = p.n_ins;
    a.cmask = p.control_mask;
    for (int i = 0; i < K; ++i) a.tg[i] = (unsigned char)targets[i];
    const u64 n_bases = dim >> (K + nc);
    const u64 groups = (n_bases + 15) >> 4;
    // persistent waves, about one resident set (256 CUs x 4 SIMDs, 4 waves per block)
    const unsigned grid = grid_for(groups, 4, K == 5 ? 768 : 1792);
    if (p.ins[0] >= kLineQubits)
        k_dense_mfma<K, true><<<grid, 256, 0, st>>>(s, n_bases, a, dmat);
    else
        k_dense_mfma<K, false><<<grid, 256, 0, st>>>(s, n_bases, a, dmat);
}

// 4-qubit dense: the register-tuple kernel (one tuple per thread, FP64 FMAs)
// or the matrix-core kernel.  Measured on a 30-qubit state (34 target sets):
// the register-tuple kernel is 0-20 % faster on every scattered target set
// (5.6-6.9 ms against 6.7-8.1 ms) and on runs of four adjacent qubits from 4
// to 17, but slower when all four targets lie below qubit 6 (the members of a
// tuple share 128-byte lines and each lane's loads are 256 bytes apart:
// 7.1-9.9 ms against 6.8-7.5 ms) or when three or more lie in 18..25 (tuple
// members 4-512 MiB apart: up to 9.1 ms against 6.9 ms).
bool dense4_on_valu(const u32* targets) {
    u32 mid = 0, hi = 0;
    for (int i = 0; i < 4; ++i) {
        mid += targets[i] >= 18 && targets[i] <= 25;
        hi = std::max(hi, targets[i]);
    }
    return hi >= 6 && mid < 3;
}

// Validates state/dim/targets/controls; returns n (qubit count) or -1.
int check_gate(void* state, u64 dim, const u32* targets, u32 nt, const u32* controls, u32 nc, const char* what) {
    const int n = log2_dim(dim);
    if (!state || n < 0) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null state or dim %llu is not a power of two >= 2", what, dim);
        return -1;
    }
    u32 all[64];
    if (nt + nc > 64 || (int)(nt + nc) > n) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: %u targets + %u controls exceed %d qubits", what, nt, nc, n);
        return -1;
    }
    for (u32 i = 0; i < nt; ++i) all[i] = targets[i];
    for (u32 i = 0; i < nc; ++i) all[nt + i] = controls[i];
    if (!check_qubits(all, nt + nc, n, what)) return -1;
    return n;
}

template <int K>
MatArg<K> mat_arg(const qm_complex* m) {
    MatArg<K> a;
    for (int i = 0; i < (1 << (2 * K)); ++i) a.m[i] = make_double2(m[i].re, m[i].im);
    return a;
}

// Dense k-qubit matrix with c controls, k >= 6: one block per tuple, tuple in
// LDS, matrix read transposed from device memory so a wave's loads of one
// column are contiguous.
__global__ __launch_bounds__(256) void k_dense_wide(double2* __restrict__ s, u64 n_bases, int K,
                                                     const unsigned char* __restrict__ ins_dev, int n_ins,
                                                     u64 cmask, const u64* __restrict__ offs,
                                                     const double2* __restrict__ mt) {
    extern __shared__ double2 tup[];
    const int N = 1 << K;
    for (u64 t = blockIdx.x; t < n_bases; t += gridDim.x) {
        u64 b = t;
        for (int k = 0; k < n_ins; ++k) b = insert_zero(b, ins_dev[k]);
        b |= cmask;
        for (int x = threadIdx.x; x < N; x += blockDim.x) tup[x] = s[b + offs[x]];
        __syncthreads();
        for (int y = threadIdx.x; y < N; y += blockDim.x) {
            double2 acc = make_double2(0.0, 0.0);
            for (int x = 0; x < N; ++x) acc = cmac(acc, mt[(u64)x * N + y], tup[x]);
            s[b + offs[y]] = acc;
        }
        __syncthreads();
    }
}

void dense_any(const u32* targets, u32 nt, const u32* controls, const u32* cvals, u32 nc, const qm_complex* matrix,
               void* state, u64 dim, void* stream, unsigned device, const char* what) {
    if (check_gate(state, dim, targets, nt, controls, nc, what) < 0) return;
    QM_CHECK_ARG(nt >= 1 && matrix, "%s: no targets or null matrix", what);
    if (!use_device(device)) return;
    hipStream_t st = stream_of(stream);
    double2* s = (double2*)state;
    switch (nt) {
        case 1: launch_gather<1, 1>(s, dim, targets, controls, cvals, nc, OpDenseArg<1>{mat_arg<1>(matrix)}, st); break;
        case 2: launch_gather<2, 1>(s, dim, targets, controls, cvals, nc, OpDenseArg<2>{mat_arg<2>(matrix)}, st); break;
        case 3: launch_dense_stream<3>(s, dim, targets, controls, cvals, nc, MatInArg<3>{mat_arg<3>(matrix)}, st); break;
        case 4:
        case 5: {
            Scratch* sc = scratch_for(st);
            if (!sc) return;
            const size_t nel = (size_t)1 << (2 * nt);
            const double2* dm = (const double2*)upload(sc, matrix, nel * sizeof(double2));
            if (!dm) return;
            if (nt == 4 && dense4_on_valu(targets))
                launch_dense_stream<4>(s, dim, targets, controls, cvals, nc, MatInPtr<4>{dm}, st);
            else if (nt == 4)
                launch_dense_mfma<4>(s, dim, targets, controls, cvals, nc, dm, st);
            else
                launch_dense_mfma<5>(s, dim, targets, controls, cvals, nc, dm, st);
            break;
        }
        default: {
            if (nt > 12) {
                set_error(QM_ERR_UNSUPPORTED, "%s: %u targets (at most 12 supported)", what, nt);
                return;
            }
            Scratch* sc = scratch_for(st);
            if (!sc) return;
            const u64 N = 1ULL << nt;
            std::vector<double2> mt(N * N);
            for (u64 y = 0; y < N; ++y)
                for (u64 x = 0; x < N; ++x) mt[x * N + y] = make_double2(matrix[y * N + x].re, matrix[y * N + x].im);
            const GatherPlan p = make_plan(targets, nt, controls, cvals, nc);
            std::vector<u64> offs(N);
            for (u64 x = 0; x < N; ++x) {
                u64 o = 0;
                for (u32 i = 0; i < nt; ++i)
                    if ((x >> i) & 1) o |= 1ULL << targets[i];
                offs[x] = o;
            }
            // Matrix beyond the ring goes to the grow-on-demand buffer.
            const size_t mbytes = N * N * sizeof(double2);
            double2* dmt;
            if (mbytes + 8192 + N * 8 <= kRingBytes / 2) {
                dmt = (double2*)upload(sc, mt.data(), mbytes);
            } else {
                dmt = (double2*)big_buffer(sc, mbytes);
                if (dmt && !hip_ok(hipMemcpyAsync(dmt, mt.data(), mbytes, hipMemcpyHostToDevice, st), "matrix upload"))
                    dmt = nullptr;
            }
            const u64* doffs = (const u64*)upload(sc, offs.data(), N * sizeof(u64));
            const unsigned char* dins = (const unsigned char*)upload(sc, p.ins, 64);
            if (!dmt || !doffs || !dins) return;
            const u64 n_bases = dim >> (nt + nc);
            k_dense_wide<<<grid_for(n_bases, 1, 1u << 16), 256, N * sizeof(double2), st>>>(
                s, n_bases, (int)nt, dins, p.n_ins, p.control_mask, doffs, dmt);
            if (mbytes + 8192 + N * 8 > kRingBytes / 2) hip_ok(hipStreamSynchronize(st), "wide gate");
            break;
        }
    }
    finish_update(st, what);
}

// Monomial k-qubit operator (one nonzero per row and column: a permutation of
// the tuple members with a factor each), e.g. the superoperator of a CNOT or CZ
// on a density matrix.  Only members that change are read and written:
// member y receives ph[y] * member src[y].  The member tables are wave-uniform
// kernel arguments, so every access address is a scalar offset plus the base.
template <int K>
struct MonoArgs {
    double2 ph[1 << K];
    unsigned char src[1 << K];
    u32 write;     // members that change
    u32 unit;      // members whose factor is exactly 1 (no multiply)
};

template <int K, bool NT>
__global__ __launch_bounds__(512) void k_monomial(double2* __restrict__ s, u64 n_bases, GatherArgs<K> a,
                                                   MonoArgs<K> m) {
    constexpr int N = 1 << K;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n_bases; i += stride) {
        const u64 b = base_of<K>(i, a);
        double2 v[N];
#pragma unroll
        for (int y = 0; y < N; ++y)
            if ((m.write >> y) & 1) v[y] = ld<NT>(s + b + a.offs[m.src[y]]);
#pragma unroll
        for (int y = 0; y < N; ++y)
            if ((m.write >> y) & 1) st<NT>(s + b + a.offs[y], ((m.unit >> y) & 1) ? v[y] : cmul(v[y], m.ph[y]));
    }
}

template <int K>
void launch_monomial(double2* s, u64 dim, const u32* targets, const qm_complex* matrix, hipStream_t st) {
    constexpr int N = 1 << K;
    MonoArgs<K> m{};
    for (int y = 0; y < N; ++y) {
        int src = -1;
        for (int x = 0; x < N; ++x)
            if (matrix[y * N + x].re != 0.0 || matrix[y * N + x].im != 0.0) src = x;
        m.src[y] = (unsigned char)src;
        m.ph[y] = make_double2(matrix[y * N + src].re, matrix[y * N + src].im);
        const bool unit = m.ph[y].x == 1.0 && m.ph[y].y == 0.0;
        if (unit) m.unit |= 1u << y;
        if (!(unit && src == y)) m.write |= 1u << y;
    }
    if (!m.write) return;
    GatherArgs<K> a;
    if (!fill_args<K>(a, targets, nullptr, nullptr, 0)) return;
    const u64 n_bases = dim >> K;
    const unsigned grid = grid_for(n_bases, 512);
    if (a.ins0 >= kLineQubits)
        k_monomial<K, true><<<grid, 512, 0, st>>>(s, n_bases, a, m);
    else
        k_monomial<K, false><<<grid, 512, 0, st>>>(s, n_bases, a, m);
}

// True when every row and every column of the 2^k x 2^k matrix has exactly one
// nonzero entry.
bool is_monomial(const qm_complex* matrix, u32 k) {
    const u32 N = 1u << k;
    std::vector<int> col_count(N, 0);
    for (u32 y = 0; y < N; ++y) {
        int nz = 0;
        for (u32 x = 0; x < N; ++x)
            if (matrix[y * N + x].re != 0.0 || matrix[y * N + x].im != 0.0) {
                ++nz;
                ++col_count[x];
            }
        if (nz != 1) return false;
    }
    for (u32 x = 0; x < N; ++x)
        if (col_count[x] != 1) return false;
    return true;
}

void monomial_any(const u32* targets, u32 nt, const qm_complex* matrix, void* state, u64 dim, void* stream,
                  unsigned device, const char* what) {
    if (check_gate(state, dim, targets, nt, nullptr, 0, what) < 0) return;
    QM_CHECK_ARG(nt >= 1 && nt <= 4 && matrix && is_monomial(matrix, nt), "%s: not a monomial 1..4-qubit operator",
                 what);
    if (!use_device(device)) return;
    hipStream_t st = stream_of(stream);
    double2* s = (double2*)state;
    switch (nt) {
        case 1: launch_monomial<1>(s, dim, targets, matrix, st); break;
        case 2: launch_monomial<2>(s, dim, targets, matrix, st); break;
        case 3: launch_monomial<3>(s, dim, targets, matrix, st); break;
        default: launch_monomial<4>(s, dim, targets, matrix, st); break;
    }
    finish_update(st, what);
}

// Named single-target gate through the gather kernel.
template <class Op>
void named1(u32 t, const Op& op, void* state, u64 dim, void* stream, unsigned device, const char* what) {
    if (check_gate(state, dim, &t, 1, nullptr, 0, what) < 0) return;
    if (!use_device(device)) return;
    hipStream_t st = stream_of(stream);
    launch_gather<1, 1>((double2*)state, dim, &t, nullptr, nullptr, 0, op, st);
    finish_update(st, what);
}

// Single target with one control bit set (CNOT, CZ).
template <class Op>
void controlled1(u32 c, u32 t, const Op& op, void* state, u64 dim, void* stream, unsigned device, const char* what) {
    const u32 one = 1;
    if (check_gate(state, dim, &t, 1, &c, 1, what) < 0) return;
    if (!use_device(device)) return;
    hipStream_t st = stream_of(stream);
    launch_gather<1, 1>((double2*)state, dim, &t, &c, &one, 1, op, st);
    finish_update(st, what);
}

// Elementwise diagonal over the controlled subspace for any number of targets:
// s[b] *= d[index of b's target bits in target-list order].
__global__ __launch_bounds__(256) void k_diag_any(double2* __restrict__ s, u64 n, const unsigned char* __restrict__ cins,
                                                   int n_cins, u64 cmask, const unsigned char* __restrict__ tg, int nt,
                                                   const double2* __restrict__ d) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u64 b = i;
        for (int k = 0; k < n_cins; ++k) b = insert_zero(b, cins[k]);
        b |= cmask;
        unsigned y = 0;
        for (int k = 0; k < nt; ++k) y |= (unsigned)((b >> tg[k]) & 1) << k;
        s[b] = cmul(s[b], d[y]);
    }
}

__global__ __launch_bounds__(256) void k_diag_whole(double2* __restrict__ s, u64 dim, const double2* __restrict__ d) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) st_stream(s + i, cmul(ld_stream(s + i), d[i]));
}

// ------------------------------------------------------------------ Pauli gates
// src/csim/update_ops_pauli_multi.cpp:51-190.  PHASE_M90ROT[j] = (-i)^j.
__device__ __forceinline__ double2 mul_mi_pow(double2 v, unsigned j) {
    switch (j & 3) {
        case 0: return v;
        case 1: return make_double2(v.y, -v.x);    // * (-i)
        case 2: return make_double2(-v.x, -v.y);   // * (-1)
        default: return make_double2(-v.y, v.x);   // * (+i)
    }
}
__device__ __forceinline__ unsigned par(u64 x) { return __popcll(x) & 1; }

template <bool NT>
__global__ __launch_bounds__(256) void k_pauli_xz(double2* __restrict__ s, u64 half, u64 bf, u64 pf, unsigned rot,
                                                   unsigned pivot) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < half; i += stride) {
        const u64 b0 = insert_zero(i, pivot), b1 = b0 ^ bf;
        const double2 v0 = ld<NT>(s + b0), v1 = ld<NT>(s + b1);
        st<NT>(s + b0, mul_mi_pow(v1, rot + 2 * par(b0 & pf)));
        st<NT>(s + b1, mul_mi_pow(v0, rot + 2 * par(b1 & pf)));
    }
}
__global__ __launch_bounds__(256) void k_pauli_z(double2* __restrict__ s, u64 dim, u64 pf) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride)
        if (par(i & pf)) {
            const double2 v = s[i];
            s[i] = make_double2(-v.x, -v.y);
        }
}
// exp(+i angle P / 2): v0' = c v0 + i s (phase(b0) v1), v1' = c v1 + i s (phase(b1) v0)
template <bool NT>
__global__ __launch_bounds__(256) void k_pauli_rot_xz(double2* __restrict__ s, u64 half, u64 bf, u64 pf,
                                                       unsigned rot, unsigned pivot, double c, double sn) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < half; i += stride) {
        const u64 b0 = insert_zero(i, pivot), b1 = b0 ^ bf;
        const double2 v0 = ld<NT>(s + b0), v1 = ld<NT>(s + b1);
        const double2 p1 = mul_mi_pow(v1, rot + 2 * par(b0 & pf));
        const double2 p0 = mul_mi_pow(v0, rot + 2 * par(b1 & pf));
        // i * sn * p = (-sn p.y, sn p.x)
        st<NT>(s + b0, make_double2(fma(c, v0.x, -sn * p1.y), fma(c, v0.y, sn * p1.x)));
        st<NT>(s + b1, make_double2(fma(c, v1.x, -sn * p0.y), fma(c, v1.y, sn * p0.x)));
    }
}
__global__ __launch_bounds__(256) void k_pauli_rot_z(double2* __restrict__ s, u64 dim, u64 pf, double c, double sn) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        const double2 f = make_double2(c, par(i & pf) ? -sn : sn);
        st_stream(s + i, cmul(ld_stream(s + i), f));
    }
}

void pauli_masks(const u32* t, const u32* p, u32 n, u64* bf, u64* pf, u32* rot, u32* pivot, bool* ok) {
    *bf = 0; *pf = 0; *rot = 0; *pivot = 0; *ok = true;
    for (u32 c = 0; c < n; ++c) {
        switch (p[c]) {
            case 0: break;
            case 1: *bf ^= 1ULL << t[c]; *pivot = t[c]; break;
            case 2: *bf ^= 1ULL << t[c]; *pf ^= 1ULL << t[c]; (*rot)++; *pivot = t[c]; break;
            case 3: *pf ^= 1ULL << t[c]; break;
            default: *ok = false;
        }
    }
}

void pauli_gate_masks(u64 bf, u64 pf, u32 rot, u32 pivot, void* state, u64 dim, void* stream, unsigned device,
                      const char* what) {
    const int n = log2_dim(dim);
    QM_CHECK_ARG(state && n >= 0, "%s: null state or bad dim", what);
    QM_CHECK_ARG(((bf | pf) >> n) == 0, "%s: mask beyond %d qubits", what, n);
    QM_CHECK_ARG(bf == 0 || ((bf >> pivot) & 1), "%s: pivot %u not in the bit-flip mask", what, pivot);
    if (!use_device(device)) return;
    hipStream_t st = stream_of(stream);
    if (bf == 0)
        k_pauli_z<<<grid_for(dim, 256 * 4), 256, 0, st>>>((double2*)state, dim, pf);
    else
        (pivot >= kLineQubits ? k_pauli_xz<true> : k_pauli_xz<false>)<<<grid_for(dim / 2, 256 * 4), 256, 0, st>>>((double2*)state, dim / 2, bf, pf, rot, pivot);
    finish_update(st, what);
}

void pauli_rot_masks(u64 bf, u64 pf, u32 rot, u32 pivot, double angle, void* state, u64 dim, void* stream,
                     unsigned device, const char* what) {
    const int n = log2_dim(dim);
    QM_CHECK_ARG(state && n >= 0, "%s: null state or bad dim", what);
    QM_CHECK_ARG(((bf | pf) >> n) == 0, "%s: mask beyond %d qubits", what, n);
    QM_CHECK_ARG(bf == 0 || ((bf >> pivot) & 1), "%s: pivot %u not in the bit-flip mask", what, pivot);
    if (!use_device(device)) return;
    hipStream_t st = stream_of(stream);
    const double c = cos(angle / 2), sn = sin(angle / 2);
    if (bf == 0)
        k_pauli_rot_z<<<grid_for(dim, 256 * 4), 256, 0, st>>>((double2*)state, dim, pf, c, sn);
    else
        (pivot >= kLineQubits ? k_pauli_rot_xz<true> : k_pauli_rot_xz<false>)<<<grid_for(dim / 2, 256 * 4), 256, 0, st>>>((double2*)state, dim / 2, bf, pf, rot, pivot, c, sn);
    finish_update(st, what);
}

}  // namespace qm

using namespace qm;

// ------------------------------------------------------------------ C ABI
extern "C" {

static const qm_complex kSqrtX[4] = {{0.5, 0.5}, {0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}};
static const qm_complex kSqrtXdag[4] = {{0.5, -0.5}, {0.5, 0.5}, {0.5, 0.5}, {0.5, -0.5}};
static const qm_complex kSqrtY[4] = {{0.5, 0.5}, {-0.5, -0.5}, {0.5, 0.5}, {0.5, 0.5}};
static const qm_complex kSqrtYdag[4] = {{0.5, -0.5}, {0.5, -0.5}, {-0.5, 0.5}, {0.5, -0.5}};

void X_gate_host(unsigned t, void* state, ITYPE dim, void* stream, unsigned dev) {
    named1(t, OpSwap01{}, state, dim, stream, dev, "X_gate_host");
}
void Y_gate_host(unsigned t, void* state, ITYPE dim, void* stream, unsigned dev) {
    named1(t, OpY{}, state, dim, stream, dev, "Y_gate_host");
}
void Z_gate_host(unsigned t, void* state, ITYPE dim, void* stream, unsigned dev) {
    named1(t, OpNeg1{}, state, dim, stream, dev, "Z_gate_host");
}
void H_gate_host(unsigned t, void* state, ITYPE dim, void* stream, unsigned dev) {
    named1(t, OpH{1.0 / std::sqrt(2.0)}, state, dim, stream, dev, "H_gate_host");
}
// src/csim/update_ops_named.cpp:13-26: S/Sdag/T/Tdag are phase gates.
void S_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    named1(t, OpPhase1{make_double2(0.0, 1.0)}, state, dim, stream, dev, "S_gate_host");
}
void Sdag_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    named1(t, OpPhase1{make_double2(0.0, -1.0)}, state, dim, stream, dev, "Sdag_gate_host");
}
void T_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    const double r = 1.0 / std::sqrt(2.0);
    named1(t, OpPhase1{make_double2(r, r)}, state, dim, stream, dev, "T_gate_host");
}
void Tdag_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    const double r = 1.0 / std::sqrt(2.0);
    named1(t, OpPhase1{make_double2(r, -r)}, state, dim, stream, dev, "Tdag_gate_host");
}
void sqrtX_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    dense_any(&t, 1, nullptr, nullptr, 0, kSqrtX, state, dim, stream, dev, "sqrtX_gate_host");
}
void sqrtXdag_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    dense_any(&t, 1, nullptr, nullptr, 0, kSqrtXdag, state, dim, stream, dev, "sqrtXdag_gate_host");
}
void sqrtY_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    dense_any(&t, 1, nullptr, nullptr, 0, kSqrtY, state, dim, stream, dev, "sqrtY_gate_host");
}
void sqrtYdag_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    dense_any(&t, 1, nullptr, nullptr, 0, kSqrtYdag, state, dim, stream, dev, "sqrtYdag_gate_host");
}
void P0_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    named1(t, OpZero1{}, state, dim, stream, dev, "P0_gate_host");
}
void P1_gate_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    named1(t, OpZero0{}, state, dim, stream, dev, "P1_gate_host");
}
void CNOT_gate_host(unsigned c, unsigned t, void* state, ITYPE dim, void* stream, unsigned dev) {
    controlled1(c, t, OpSwap01{}, state, dim, stream, dev, "CNOT_gate_host");
}
void CZ_gate_host(unsigned c, unsigned t, void* state, ITYPE dim, void* stream, unsigned dev) {
    controlled1(c, t, OpNeg1{}, state, dim, stream, dev, "CZ_gate_host");
}
void SWAP_gate_host(unsigned t0, unsigned t1, void* state, ITYPE dim, void* stream, unsigned dev) {
    const u32 tg[2] = {t0, t1};
    if (check_gate(state, dim, tg, 2, nullptr, 0, "SWAP_gate_host") < 0) return;
    if (!use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    launch_gather<2, 1>((double2*)state, dim, tg, nullptr, nullptr, 0, OpSwap12{}, st);
    finish_update(st, "SWAP_gate_host");
}

// src/csim/update_ops_pauli_single.cpp:54-122 (exp(+i angle P / 2))
void RX_gate_host(UINT t, double angle, void* state, ITYPE dim, void* stream, unsigned dev) {
    const double c = std::cos(angle / 2), s = std::sin(angle / 2);
    const qm_complex m[4] = {{c, 0}, {0, s}, {0, s}, {c, 0}};
    dense_any(&t, 1, nullptr, nullptr, 0, m, state, dim, stream, dev, "RX_gate_host");
}
void RY_gate_host(UINT t, double angle, void* state, ITYPE dim, void* stream, unsigned dev) {
    const double c = std::cos(angle / 2), s = std::sin(angle / 2);
    const qm_complex m[4] = {{c, 0}, {s, 0}, {-s, 0}, {c, 0}};
    dense_any(&t, 1, nullptr, nullptr, 0, m, state, dim, stream, dev, "RY_gate_host");
}
void RZ_gate_host(UINT t, double angle, void* state, ITYPE dim, void* stream, unsigned dev) {
    const double c = std::cos(angle / 2), s = std::sin(angle / 2);
    const qm_complex d[2] = {{c, s}, {c, -s}};
    single_qubit_diagonal_matrix_gate_host(t, d, state, dim, stream, dev);
}

void single_qubit_Pauli_gate_host(UINT t, UINT op, void* state, ITYPE dim, void* stream, unsigned dev) {
    switch (op) {
        case 0: return;
        case 1: X_gate_host(t, state, dim, stream, dev); return;
        case 2: Y_gate_host(t, state, dim, stream, dev); return;
        case 3: Z_gate_host(t, state, dim, stream, dev); return;
        default: set_error(QM_ERR_INVALID_ARGUMENT, "single_qubit_Pauli_gate_host: Pauli id %u", op);
    }
}
void single_qubit_Pauli_rotation_gate_host(unsigned t, unsigned op, double angle, void* state, ITYPE dim,
                                           void* stream, unsigned dev) {
    switch (op) {
        case 0: return;
        case 1: RX_gate_host(t, angle, state, dim, stream, dev); return;
        case 2: RY_gate_host(t, angle, state, dim, stream, dev); return;
        case 3: RZ_gate_host(t, angle, state, dim, stream, dev); return;
        default: set_error(QM_ERR_INVALID_ARGUMENT, "single_qubit_Pauli_rotation_gate_host: Pauli id %u", op);
    }
}

void single_qubit_dense_matrix_gate_host(unsigned t, const qm_complex matrix[4], void* state, ITYPE dim,
                                         void* stream, unsigned dev) {
    dense_any(&t, 1, nullptr, nullptr, 0, matrix, state, dim, stream, dev, "single_qubit_dense_matrix_gate_host");
}
void single_qubit_diagonal_matrix_gate_host(unsigned t, const qm_complex d[2], void* state, ITYPE dim,
                                            void* stream, unsigned dev) {
    QM_CHECK_ARG(d, "single_qubit_diagonal_matrix_gate_host: null diagonal");
    OpDiag<1> op;
    op.d[0] = make_double2(d[0].re, d[0].im);
    op.d[1] = make_double2(d[1].re, d[1].im);
    named1(t, op, state, dim, stream, dev, "single_qubit_diagonal_matrix_gate_host");
}
void single_qubit_phase_gate_host(unsigned t, qm_complex phase, void* state, ITYPE dim, void* stream,
                                  unsigned dev) {
    named1(t, OpPhase1{make_double2(phase.re, phase.im)}, state, dim, stream, dev, "single_qubit_phase_gate_host");
}
void single_qubit_control_single_qubit_dense_matrix_gate_host(unsigned c, unsigned cv, unsigned t,
                                                              const qm_complex matrix[4], void* state, ITYPE dim,
                                                              void* stream, unsigned dev) {
    QM_CHECK_ARG(cv <= 1, "single_qubit_control_single_qubit_dense_matrix_gate_host: control value %u", cv);
    dense_any(&t, 1, &c, &cv, 1, matrix, state, dim, stream, dev,
              "single_qubit_control_single_qubit_dense_matrix_gate_host");
}
void multi_qubit_control_single_qubit_dense_matrix_gate_host(const UINT* cl, const UINT* cvl, UINT nc, UINT t,
                                                             const qm_complex matrix[4], void* state, ITYPE dim,
                                                             void* stream, unsigned dev) {
    for (UINT i = 0; i < nc; ++i)
        QM_CHECK_ARG(cvl[i] <= 1, "multi_qubit_control_single_qubit_dense_matrix_gate_host: control value %u", cvl[i]);
    dense_any(&t, 1, cl, cvl, nc, matrix, state, dim, stream, dev,
              "multi_qubit_control_single_qubit_dense_matrix_gate_host");
}

void double_qubit_dense_matrix_gate_host(unsigned t0, unsigned t1, const qm_complex matrix[16], void* state,
                                         ITYPE dim, void* stream, unsigned dev) {
    const u32 tg[2] = {t0, t1};
    dense_any(tg, 2, nullptr, nullptr, 0, matrix, state, dim, stream, dev, "double_qubit_dense_matrix_gate_host");
}
void triple_qubit_dense_matrix_gate_host(unsigned t0, unsigned t1, unsigned t2, const qm_complex matrix[64],
                                         void* state, ITYPE dim, void* stream, unsigned dev) {
    const u32 tg[3] = {t0, t1, t2};
    dense_any(tg, 3, nullptr, nullptr, 0, matrix, state, dim, stream, dev, "triple_qubit_dense_matrix_gate_host");
}
void quad_qubit_dense_matrix_gate_host(const unsigned tg[4], const qm_complex matrix[256], void* state, ITYPE dim,
                                       void* stream, unsigned dev) {
    dense_any(tg, 4, nullptr, nullptr, 0, matrix, state, dim, stream, dev, "quad_qubit_dense_matrix_gate_host");
}
void penta_qubit_dense_matrix_gate_host(const unsigned tg[5], const qm_complex matrix[1024], void* state, ITYPE dim,
                                        void* stream, UINT dev) {
    dense_any(tg, 5, nullptr, nullptr, 0, matrix, state, dim, stream, dev, "penta_qubit_dense_matrix_gate_host");
}
void multi_qubit_dense_matrix_gate_host(const UINT* tl, UINT nt, const qm_complex* matrix, void* state, ITYPE dim,
                                        void* stream, unsigned dev) {
    dense_any(tl, nt, nullptr, nullptr, 0, matrix, state, dim, stream, dev, "multi_qubit_dense_matrix_gate_host");
}
void single_qubit_control_multi_qubit_dense_matrix_gate_host(UINT c, UINT cv, const UINT* tl, UINT nt,
                                                             const qm_complex* matrix, void* state, ITYPE dim,
                                                             void* stream, unsigned dev) {
    QM_CHECK_ARG(cv <= 1, "single_qubit_control_multi_qubit_dense_matrix_gate_host: control value %u", cv);
    dense_any(tl, nt, &c, &cv, 1, matrix, state, dim, stream, dev,
              "single_qubit_control_multi_qubit_dense_matrix_gate_host");
}
void multi_qubit_control_multi_qubit_dense_matrix_gate_host(const UINT* cl, const UINT* cvl, UINT nc,
                                                            const UINT* tl, UINT nt, const qm_complex* matrix,
                                                            void* state, ITYPE dim, void* stream, unsigned dev) {
    for (UINT i = 0; i < nc; ++i)
        QM_CHECK_ARG(cvl[i] <= 1, "multi_qubit_control_multi_qubit_dense_matrix_gate_host: control value %u", cvl[i]);
    dense_any(tl, nt, cl, cvl, nc, matrix, state, dim, stream, dev,
              "multi_qubit_control_multi_qubit_dense_matrix_gate_host");
}

void qm_multi_qubit_control_multi_qubit_diagonal_matrix_gate_host(const UINT* cl, const UINT* cvl, UINT nc,
                                                                  const UINT* tl, UINT nt, const qm_complex* d,
                                                                  void* state, ITYPE dim, void* stream,
                                                                  unsigned dev) {
    const char* what = "qm_multi_qubit_control_multi_qubit_diagonal_matrix_gate_host";
    if (check_gate(state, dim, tl, nt, cl, nc, what) < 0) return;
    QM_CHECK_ARG(d && nt >= 1 && nt <= 24, "%s: null diagonal or %u targets", what, nt);
    for (UINT i = 0; i < nc; ++i) QM_CHECK_ARG(cvl[i] <= 1, "%s: control value %u", what, cvl[i]);
    if (!use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    Scratch* sc = scratch_for(st);
    if (!sc) return;
    const GatherPlan cp = make_plan(nullptr, 0, cl, cvl, nc);
    unsigned char tg[64];
    for (UINT i = 0; i < nt; ++i) tg[i] = (unsigned char)tl[i];
    const size_t nd = (size_t)1 << nt;
    const double2* dd = (const double2*)upload(sc, d, nd * sizeof(double2));
    const unsigned char* dcins = (const unsigned char*)upload(sc, cp.ins, 64);
    const unsigned char* dtg = (const unsigned char*)upload(sc, tg, 64);
    if (!dd || !dcins || !dtg) return;
    const u64 n = dim >> nc;
    k_diag_any<<<grid_for(n, 256 * 4), 256, 0, st>>>((double2*)state, n, dcins, cp.n_ins, cp.control_mask, dtg,
                                                      (int)nt, dd);
    finish_update(st, what);
}

void multi_qubit_diagonal_matrix_gate_host(const qm_complex* d, void* state, ITYPE dim, void* stream, UINT dev) {
    const char* what = "multi_qubit_diagonal_matrix_gate_host";
    QM_CHECK_ARG(state && d && log2_dim(dim) >= 0, "%s: bad arguments", what);
    if (!use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    Scratch* sc = scratch_for(st);
    if (!sc) return;
    void* buf = big_buffer(sc, (size_t)dim * 16);
    if (!buf) return;
    if (!hip_ok(hipMemcpyAsync(buf, d, (size_t)dim * 16, hipMemcpyHostToDevice, st), what)) return;
    k_diag_whole<<<grid_for(dim, 256 * 4), 256, 0, st>>>((double2*)state, dim, (const double2*)buf);
    if (!hip_ok(hipGetLastError(), what)) return;
    // The diagonal was copied from the caller's pageable memory through the
    // shared buffer; wait so a later call cannot overwrite it early.
    hip_ok(hipStreamSynchronize(st), what);
}

void multi_qubit_Pauli_gate_Z_mask_host(ITYPE pf, void* state, ITYPE dim, void* stream, unsigned dev) {
    pauli_gate_masks(0, pf, 0, 0, state, dim, stream, dev, "multi_qubit_Pauli_gate_Z_mask_host");
}
void multi_qubit_Pauli_gate_XZ_mask_host(ITYPE bf, ITYPE pf, UINT rot, UINT pivot, void* state, ITYPE dim,
                                         void* stream, unsigned dev) {
    QM_CHECK_ARG(bf != 0, "multi_qubit_Pauli_gate_XZ_mask_host: empty bit-flip mask");
    pauli_gate_masks(bf, pf, rot, pivot, state, dim, stream, dev, "multi_qubit_Pauli_gate_XZ_mask_host");
}
void multi_qubit_Pauli_rotation_gate_XZ_mask_host(ITYPE bf, ITYPE pf, UINT rot, UINT pivot, double angle,
                                                  void* state, ITYPE dim, void* stream, unsigned dev) {
    QM_CHECK_ARG(bf != 0, "multi_qubit_Pauli_rotation_gate_XZ_mask_host: empty bit-flip mask");
    pauli_rot_masks(bf, pf, rot, pivot, angle, state, dim, stream, dev, "multi_qubit_Pauli_rotation_gate_XZ_mask_host");
}
void multi_qubit_Pauli_rotation_gate_Z_mask_host(ITYPE pf, double angle, void* state, ITYPE dim, void* stream,
                                                 unsigned dev) {
    pauli_rot_masks(0, pf, 0, 0, angle, state, dim, stream, dev, "multi_qubit_Pauli_rotation_gate_Z_mask_host");
}
void multi_qubit_Pauli_gate_partial_list_host(const UINT* tl, const UINT* pl, UINT n, void* state, ITYPE dim,
                                              void* stream, unsigned dev) {
    u64 bf, pf; u32 rot, pivot; bool ok;
    pauli_masks(tl, pl, n, &bf, &pf, &rot, &pivot, &ok);
    QM_CHECK_ARG(ok, "multi_qubit_Pauli_gate_partial_list_host: Pauli id out of 0..3");
    pauli_gate_masks(bf, pf, rot, pivot, state, dim, stream, dev, "multi_qubit_Pauli_gate_partial_list_host");
}
void multi_qubit_Pauli_gate_whole_list_host(const UINT* pl, UINT n, void* state, ITYPE dim, void* stream,
                                            unsigned dev) {
    std::vector<u32> tl(n);
    for (UINT i = 0; i < n; ++i) tl[i] = i;
    multi_qubit_Pauli_gate_partial_list_host(tl.data(), pl, n, state, dim, stream, dev);
}
void multi_qubit_Pauli_rotation_gate_partial_list_host(const UINT* tl, const UINT* pl, UINT n, double angle,
                                                       void* state, ITYPE dim, void* stream, unsigned dev) {
    u64 bf, pf; u32 rot, pivot; bool ok;
    pauli_masks(tl, pl, n, &bf, &pf, &rot, &pivot, &ok);
    QM_CHECK_ARG(ok, "multi_qubit_Pauli_rotation_gate_partial_list_host: Pauli id out of 0..3");
    pauli_rot_masks(bf, pf, rot, pivot, angle, state, dim, stream, dev,
                    "multi_qubit_Pauli_rotation_gate_partial_list_host");
}
void multi_qubit_Pauli_rotation_gate_whole_list_host(const UINT* pl, UINT n, double angle, void* state, ITYPE dim,
                                                     void* stream, unsigned dev) {
    std::vector<u32> tl(n);
    for (UINT i = 0; i < n; ++i) tl[i] = i;
    multi_qubit_Pauli_rotation_gate_partial_list_host(tl.data(), pl, n, angle, state, dim, stream, dev);
}

}  // extern "C"
