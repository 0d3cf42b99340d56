// Adjoint differentiation of <psi(theta)| H |psi(theta)> for circuits of fixed
// gates and parametric Pauli rotations, batched over parameter sets.
//
// Reference: ParametricQuantumCircuit::backprop / backprop_inner_product
// (src/vqcsim/parametric_circuit.cpp:205-335): run the circuit, set the
// "bistate" to 2 H|psi>, then walk the gates backwards applying each inverse to
// both states; at a parametric rotation R(theta) = exp(i theta P / 2) the
// gradient is Re <bistate| R(theta + pi) |state> / 2 = -Im <lambda|P|psi> with
// lambda = H|psi> (the reference's bistate is 2 lambda).  The reference spends
// a state copy, a gate and an inner product per parameter (about three passes)
// plus two more passes for the inverses.  Here the gates are grouped into
// passes over blocks of K = 3 qubits: a pass keeps the 8 amplitudes of psi
// (and of lambda on the way back) that differ only in its block qubits in
// registers and runs a contiguous run of gates on them -- every gate whose
// non-diagonal qubits lie in the block (rotations with their X/Y qubits in it,
// dense gates on block qubits) and every diagonal gate (CZ, Z rotations, diagonal
// matrices) wherever its qubits are, since an outside bit is fixed per thread.
// Forward, a pass applies its gates in order; backward, in reverse order it
// adds each rotation's term -Im <lambda|P|psi> and applies the inverse to both
// states.  A hardware-efficient layer on 20 qubits (RY, RZ on every qubit, a CZ
// chain) is 7 passes instead of 59 per-gate passes.  H|psi> with <psi|H|psi>
// is one kernel over all terms.
//
// Layout: `batch` states of 2^n complex128 each, contiguous (state b at
// b * 2^n); theta is batch x params (row-major).  Every reduction is written
// as per-block partial sums and summed in a fixed order at the end, so the
// results do not depend on scheduling.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qm_internal.hpp"

namespace qm {
namespace {

constexpr int kAdjThreads = 256;
constexpr unsigned kAdjBlocks = 64;  // blocks per state per launch (grid-stride)

struct AdjGate {
    u32 kind;  // 0 fixed dense (k <= 2), 1 rotation
    u32 k;
    u32 q[2];
    double2 m[16];      // fixed: matrix (row-major, bit i <-> q[i])
    u64 bf, pf;         // rotation: Pauli masks
    u32 rot, pivot;
    u32 param, occ;     // rotation: parameter index, occurrence slot
    double coef;
};

struct AdjTerm {
    u64 bf, pf;
    u32 rot, pad;
    double coef;
};

constexpr int kPassQubits = 3;            // block qubits per pass
constexpr int kPassRots = 16;             // rotations per pass (LDS accumulators)

// One gate inside a pass, in the pass's terms.  Member bit j of a tuple is
// block qubit blk[j] (ascending); bits outside the block come from the tuple's
// base index.
struct PassOp {
    u32 type;          // 0 rotation, 1 dense on one block bit, 2 dense on two block bits, 3 diagonal
    u32 lbf, lpf;      // rotation: flip / phase masks over member bits
    u64 opf;           // rotation: phase mask over qubits outside the block
    u32 rot;           // rotation: number of Y factors (P = i^rot X^bf Z^pf)
    u32 slot;          // rotation: index among the pass's rotations
    u32 occ;           // rotation: occurrence (partials row)
    u32 a, b;          // dense: member bits (matrix bit 0 <-> a, bit 1 <-> b)
    u32 dk;            // diagonal: number of qubits (1 or 2)
    u32 dloc[2];       // diagonal: member bit, or 255 for a qubit outside the block
    u32 dpos[2];       // diagonal: qubit position when outside the block
    double2 m[16];     // dense matrix (row-major) or the diagonal's entries
};

struct PassArgs {
    u64 offs[1 << kPassQubits];  // offset of member m
    unsigned char blk[kPassQubits];
    u32 op0, nops;               // ops[op0 .. op0 + nops)
    u32 nrot;
    u32 param[kPassRots];        // rotation slot -> parameter
    double coef[kPassRots];      // rotation slot -> coefficient
};

__device__ __forceinline__ double2 ipow(double2 z, u32 k) {  // z * i^k
    switch (k & 3u) {
        case 0: return z;
        case 1: return make_double2(-z.y, z.x);
        case 2: return make_double2(-z.x, -z.y);
        default: return make_double2(z.y, -z.x);
    }
}

__global__ __launch_bounds__(kAdjThreads) void k_adj_zero(double2* __restrict__ s, u64 dim) {
    double2* st = s + (u64)blockIdx.y * dim;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += (u64)gridDim.x * blockDim.x)
        st[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
}

// Fixed gate (dense on 1 or 2 qubits) applied as M (INV = false) or M^dagger,
// to psi and, when TWO, to lambda as well.
template <bool INV, bool TWO>
__global__ __launch_bounds__(kAdjThreads) void k_adj_fixed(double2* __restrict__ psi, double2* __restrict__ lam,
                                                           u64 dim, AdjGate g) {
    const u64 off = (u64)blockIdx.y * dim;
    double2 m[16];
    const u32 d = 1u << g.k;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if ((u32)r >= d || (u32)c >= d) continue;
            m[r * 4 + c] = INV ? cconj(g.m[c * d + r]) : g.m[r * d + c];
        }
    const u64 groups = dim >> g.k;
    for (u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x; t < groups; t += (u64)gridDim.x * blockDim.x) {
        if (g.k == 1) {
            const u64 b0 = insert_zero(t, g.q[0]), b1 = b0 | (1ULL << g.q[0]);
#pragma unroll
            for (int w = 0; w < (TWO ? 2 : 1); ++w) {
                double2* s = (w ? lam : psi) + off;
                const double2 x = s[b0], y = s[b1];
                s[b0] = cmac(cmul(m[0], x), m[1], y);
                s[b1] = cmac(cmul(m[4], x), m[5], y);
            }
        } else {
            const u32 lo = min(g.q[0], g.q[1]), hi = max(g.q[0], g.q[1]);
            const u64 b = insert_zero(insert_zero(t, lo), hi);
            const u64 i0 = b, i1 = b | (1ULL << g.q[0]), i2 = b | (1ULL << g.q[1]), i3 = i1 | i2;
#pragma unroll
            for (int w = 0; w < (TWO ? 2 : 1); ++w) {
                double2* s = (w ? lam : psi) + off;
                const double2 a0 = s[i0], a1 = s[i1], a2 = s[i2], a3 = s[i3];
                s[i0] = cmac(cmac(cmac(cmul(m[0], a0), m[1], a1), m[2], a2), m[3], a3);
                s[i1] = cmac(cmac(cmac(cmul(m[4], a0), m[5], a1), m[6], a2), m[7], a3);
                s[i2] = cmac(cmac(cmac(cmul(m[8], a0), m[9], a1), m[10], a2), m[11], a3);
                s[i3] = cmac(cmac(cmac(cmul(m[12], a0), m[13], a1), m[14], a2), m[15], a3);
            }
        }
    }
}

// Rotation exp(i a P / 2), a = coef * theta[b][param].  FORWARD applies it to
// psi.  Otherwise (backward) the block adds -Im <lambda|P|psi> over its
// amplitudes to partials[(occ * batch + b) * gridDim.x + blockIdx.x] and then
// applies exp(-i a P / 2) to psi and lambda.
template <bool FORWARD>
__global__ __launch_bounds__(kAdjThreads) void k_adj_rot(double2* __restrict__ psi, double2* __restrict__ lam,
                                                         u64 dim, AdjGate g, const double* __restrict__ theta,
                                                         u32 nparams, double* __restrict__ partials) {
    __shared__ double red[32];
    const u32 b = blockIdx.y;
    const u64 off = (u64)b * dim;
    const double a = g.coef * theta[(u64)b * nparams + g.param];
    const double c = cos(0.5 * a), sn = FORWARD ? sin(0.5 * a) : -sin(0.5 * a);
    double acc = 0.0;
    double2* P = psi + off;
    double2* Lm = lam + off;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    if (g.bf == 0) {
        // diagonal: (P psi)_x = s_x psi_x with s_x = i^rot (-1)^popcount(x & pf)
        for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < dim; x += stride) {
            const double2 ph = ipow(make_double2((__popcll(x & g.pf) & 1) ? -1.0 : 1.0, 0.0), g.rot);
            const double2 v = P[x];
            const double2 pv = cmul(ph, v);
            if (!FORWARD) {
                const double2 l = Lm[x];
                acc -= l.x * pv.y - l.y * pv.x;  // -Im(conj(l) * pv)
            }
            // (c + i sn ph) v
            const double2 f = make_double2(c - sn * ph.y, sn * ph.x);
            P[x] = cmul(f, v);
            if (!FORWARD) Lm[x] = cmul(f, Lm[x]);
        }
    } else {
        const u64 half = dim >> 1;
        for (u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x; t < half; t += stride) {
            const u64 x = insert_zero(t, g.pivot), y = x ^ g.bf;
            // P|y> = i^rot (-1)^popcount(y & pf) |x>, P|x> = i^rot (-1)^popcount(x & pf) |y>
            const double sy = (__popcll(y & g.pf) & 1) ? -1.0 : 1.0, sx = (__popcll(x & g.pf) & 1) ? -1.0 : 1.0;
            const double2 px = ipow(make_double2(sy, 0.0), g.rot);  // coefficient of psi_y in (P psi)_x
            const double2 py = ipow(make_double2(sx, 0.0), g.rot);  // coefficient of psi_x in (P psi)_y
            const double2 vx = P[x], vy = P[y];
            const double2 ppx = cmul(px, vy), ppy = cmul(py, vx);
            if (!FORWARD) {
                const double2 lx = Lm[x], ly = Lm[y];
                acc -= (lx.x * ppx.y - lx.y * ppx.x) + (ly.x * ppy.y - ly.y * ppy.x);
                // inverse on lambda: c l + i sn (P l)
                const double2 qx = cmul(px, ly), qy = cmul(py, lx);
                Lm[x] = make_double2(c * lx.x - sn * qx.y, c * lx.y + sn * qx.x);
                Lm[y] = make_double2(c * ly.x - sn * qy.y, c * ly.y + sn * qy.x);
            }
            P[x] = make_double2(c * vx.x - sn * ppx.y, c * vx.y + sn * ppx.x);
            P[y] = make_double2(c * vy.x - sn * ppy.y, c * vy.y + sn * ppy.x);
        }
    }
    if (!FORWARD) {
        acc = block_sum(acc * g.coef, red);  // d/dtheta = coef * d/da
        if (threadIdx.x == 0) partials[((u64)g.occ * gridDim.y + b) * gridDim.x + blockIdx.x] = acc;
    }
}

// ---------------------------------------------------------------- pass kernel
constexpr int kPassN = 1 << kPassQubits;

// exp(+-i a P / 2) on the tuple, P's flip part LBF known at compile time so
// every register index is a constant.  (P v)_y = i^rot (-1)^(so + |x & lpf|) v_x
// with x = y ^ LBF.  BACK: first acc += -Im sum_y conj(l_y) (P v)_y, then the
// inverse on v and l.
template <int LBF, bool BACK>
__device__ __forceinline__ void rot_apply(double2* v, double2* l, u32 lpf, u32 rot, u32 so, double c, double sn,
                                          double& acc) {
#pragma unroll
    for (int y = 0; y < kPassN; ++y) {
        const int x = y ^ LBF;
        if (LBF != 0 && x < y) continue;  // pairs handled once, from their lower member
        // factors of P: (P v)_y = fy * v_x, (P v)_x = fx * v_y
        const u32 sy = (so + __popc((u32)x & lpf)) & 1u, sx = (so + __popc((u32)y & lpf)) & 1u;
        const double2 fy = ipow(make_double2(sy ? -1.0 : 1.0, 0.0), rot);
        const double2 fx = ipow(make_double2(sx ? -1.0 : 1.0, 0.0), rot);
        if (LBF == 0) {
            const double2 pv = cmul(fy, v[y]);
            if (BACK) acc -= l[y].x * pv.y - l[y].y * pv.x;
            // (c + i sn fy) v_y
            const double2 f = make_double2(c - sn * fy.y, sn * fy.x);
            v[y] = cmul(f, v[y]);
            if (BACK) l[y] = cmul(f, l[y]);
        } else {
            const double2 vy = v[y], vx = v[x];
            const double2 py = cmul(fy, vx), px = cmul(fx, vy);  // (P v)_y, (P v)_x
            if (BACK) acc -= (l[y].x * py.y - l[y].y * py.x) + (l[x].x * px.y - l[x].y * px.x);
            v[y] = make_double2(c * vy.x - sn * py.y, c * vy.y + sn * py.x);
            v[x] = make_double2(c * vx.x - sn * px.y, c * vx.y + sn * px.x);
            if (BACK) {
                const double2 ly = l[y], lx = l[x];
                const double2 qy = cmul(fy, lx), qx = cmul(fx, ly);
                l[y] = make_double2(c * ly.x - sn * qy.y, c * ly.y + sn * qy.x);
                l[x] = make_double2(c * lx.x - sn * qx.y, c * lx.y + sn * qx.x);
            }
        }
    }
}

template <bool BACK>
__device__ __forceinline__ void rot_dispatch(u32 lbf, double2* v, double2* l, u32 lpf, u32 rot, u32 so, double c,
                                             double sn, double& acc) {
    switch (lbf) {
#define QM_ROT_CASE(B)                                                        \
    case B:                                                                   \
        if constexpr (B < kPassN) rot_apply<B, BACK>(v, l, lpf, rot, so, c, sn, acc); \
        break;
        QM_ROT_CASE(0) QM_ROT_CASE(1) QM_ROT_CASE(2) QM_ROT_CASE(3) QM_ROT_CASE(4) QM_ROT_CASE(5) QM_ROT_CASE(6)
        QM_ROT_CASE(7) QM_ROT_CASE(8) QM_ROT_CASE(9) QM_ROT_CASE(10) QM_ROT_CASE(11) QM_ROT_CASE(12)
        QM_ROT_CASE(13) QM_ROT_CASE(14) QM_ROT_CASE(15)
#undef QM_ROT_CASE
        default: break;
    }
}

// Dense matrix on member bit A (INV: its conjugate transpose) applied to v.
template <int A, bool INV>
__device__ __forceinline__ void dense1_apply(double2* v, const double2* m) {
    const double2 m00 = INV ? cconj(m[0]) : m[0], m01 = INV ? cconj(m[2]) : m[1];
    const double2 m10 = INV ? cconj(m[1]) : m[2], m11 = INV ? cconj(m[3]) : m[3];
#pragma unroll
    for (int y = 0; y < kPassN; ++y) {
        if (y & (1 << A)) continue;
        const int z = y | (1 << A);
        const double2 a0 = v[y], a1 = v[z];
        v[y] = cmac(cmul(m00, a0), m01, a1);
        v[z] = cmac(cmul(m10, a0), m11, a1);
    }
}

template <int A, int B, bool INV>
__device__ __forceinline__ void dense2_apply(double2* v, const double2* m) {
    double2 mm[16];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) mm[r * 4 + c] = INV ? cconj(m[c * 4 + r]) : m[r * 4 + c];
#pragma unroll
    for (int y = 0; y < kPassN; ++y) {
        if ((y & (1 << A)) || (y & (1 << B))) continue;
        const int i1 = y | (1 << A), i2 = y | (1 << B), i3 = i1 | i2;
        const double2 a0 = v[y], a1 = v[i1], a2 = v[i2], a3 = v[i3];
        v[y] = cmac(cmac(cmac(cmul(mm[0], a0), mm[1], a1), mm[2], a2), mm[3], a3);
        v[i1] = cmac(cmac(cmac(cmul(mm[4], a0), mm[5], a1), mm[6], a2), mm[7], a3);
        v[i2] = cmac(cmac(cmac(cmul(mm[8], a0), mm[9], a1), mm[10], a2), mm[11], a3);
        v[i3] = cmac(cmac(cmac(cmul(mm[12], a0), mm[13], a1), mm[14], a2), mm[15], a3);
    }
}

template <bool INV>
__device__ __forceinline__ void dense1_dispatch(u32 a, double2* v, const double2* m) {
    switch (a) {
        case 0: dense1_apply<0, INV>(v, m); break;
        case 1: dense1_apply<1, INV>(v, m); break;
        case 2: dense1_apply<2, INV>(v, m); break;
        default:
            if constexpr (kPassQubits > 3) dense1_apply<3, INV>(v, m);
            break;
    }
}

template <bool INV>
__device__ __forceinline__ void dense2_dispatch(u32 a, u32 b, double2* v, const double2* m) {
    switch (a * 4 + b) {
#define QM_D2_CASE(A, B)                                                       \
    case A * 4 + B:                                                            \
        if constexpr (A < kPassQubits && B < kPassQubits) dense2_apply<A, B, INV>(v, m); \
        break;
        QM_D2_CASE(0, 1) QM_D2_CASE(0, 2) QM_D2_CASE(0, 3) QM_D2_CASE(1, 0) QM_D2_CASE(1, 2) QM_D2_CASE(1, 3)
        QM_D2_CASE(2, 0) QM_D2_CASE(2, 1) QM_D2_CASE(2, 3) QM_D2_CASE(3, 0) QM_D2_CASE(3, 1) QM_D2_CASE(3, 2)
#undef QM_D2_CASE
        default: break;
    }
}

// Diagonal gate: member y of the tuple with base index `base` gets entry d[idx],
// idx collecting the gate's qubit bits (member bits or base bits).
template <bool INV>
__device__ __forceinline__ void diag_apply(double2* v, const PassOp& op, u64 base) {
    u32 outside = 0;
    for (u32 j = 0; j < op.dk; ++j)
        if (op.dloc[j] == 255u) outside |= (u32)((base >> op.dpos[j]) & 1ULL) << j;
#pragma unroll
    for (int y = 0; y < kPassN; ++y) {
        u32 idx = outside;
        for (u32 j = 0; j < op.dk; ++j)
            if (op.dloc[j] != 255u) idx |= (((u32)y >> op.dloc[j]) & 1u) << j;
        const double2 d = INV ? cconj(op.m[idx]) : op.m[idx];
        v[y] = cmul(d, v[y]);
    }
}

// One pass over `batch` states (blockIdx.y = state).  BACK: psi and lambda,
// gates in reverse order, rotation terms summed per block into
// partials[(occ * batch + b) * gridDim.x + blockIdx.x].
template <bool BACK>
__global__ __launch_bounds__(kAdjThreads) void k_adj_pass(double2* __restrict__ psi, double2* __restrict__ lam,
                                                          u64 dim, PassArgs pa, const PassOp* __restrict__ ops,
                                                          const double* __restrict__ theta, u32 nparams,
                                                          double* __restrict__ partials) {
    __shared__ double acc_lds[BACK ? kPassRots * kAdjThreads : 1];
    __shared__ double cs_lds[kPassRots][2];
    __shared__ double red[32];
    const u32 bs = blockIdx.y;
    const u64 off = (u64)bs * dim;
    if (threadIdx.x < pa.nrot) {
        const double ang = pa.coef[threadIdx.x] * theta[(u64)bs * nparams + pa.param[threadIdx.x]];
        cs_lds[threadIdx.x][0] = cos(0.5 * ang);
        cs_lds[threadIdx.x][1] = sin(0.5 * ang);
    }
    if (BACK)
        for (u32 r = 0; r < pa.nrot; ++r) acc_lds[r * kAdjThreads + threadIdx.x] = 0.0;
    __syncthreads();
    double2* P = psi + off;
    double2* Lm = lam + off;
    const u64 tuples = dim >> kPassQubits;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x; t < tuples; t += stride) {
        u64 base = t;
#pragma unroll
        for (int j = 0; j < kPassQubits; ++j) base = insert_zero(base, pa.blk[j]);
        double2 v[kPassN], l[kPassN];
#pragma unroll
        for (int y = 0; y < kPassN; ++y) v[y] = P[base + pa.offs[y]];
        if (BACK) {
#pragma unroll
            for (int y = 0; y < kPassN; ++y) l[y] = Lm[base + pa.offs[y]];
        }
        for (u32 k = 0; k < pa.nops; ++k) {
            const PassOp& op = ops[pa.op0 + (BACK ? pa.nops - 1 - k : k)];
            switch (op.type) {
                case 0: {
                    const u32 so = (u32)__popcll(base & op.opf) & 1u;
                    const double c = cs_lds[op.slot][0], sn = BACK ? -cs_lds[op.slot][1] : cs_lds[op.slot][1];
                    double acc = 0.0;
                    rot_dispatch<BACK>(op.lbf, v, l, op.lpf, op.rot, so, c, sn, acc);
                    if (BACK) acc_lds[op.slot * kAdjThreads + threadIdx.x] += acc;
                    break;
                }
                case 1:
                    dense1_dispatch<BACK>(op.a, v, op.m);
                    if (BACK) dense1_dispatch<true>(op.a, l, op.m);
                    break;
                case 2:
                    dense2_dispatch<BACK>(op.a, op.b, v, op.m);
                    if (BACK) dense2_dispatch<true>(op.a, op.b, l, op.m);
                    break;
                default:
                    diag_apply<BACK>(v, op, base);
                    if (BACK) diag_apply<true>(l, op, base);
                    break;
            }
        }
#pragma unroll
        for (int y = 0; y < kPassN; ++y) P[base + pa.offs[y]] = v[y];
        if (BACK) {
#pragma unroll
            for (int y = 0; y < kPassN; ++y) Lm[base + pa.offs[y]] = l[y];
        }
    }
    if (BACK) {
        for (u32 k = 0; k < pa.nops; ++k) {
            const PassOp& op = ops[pa.op0 + k];
            if (op.type != 0) continue;
            const double a = block_sum(acc_lds[op.slot * kAdjThreads + threadIdx.x], red);
            // d/dtheta = coef * d/dangle
            if (threadIdx.x == 0)
                partials[((u64)op.occ * gridDim.y + bs) * gridDim.x + blockIdx.x] = a * pa.coef[op.slot];
        }
    }
}

// lambda = sum_k c_k P_k psi and the block's share of Re <psi|lambda>.
__global__ __launch_bounds__(kAdjThreads) void k_adj_hpsi(const double2* __restrict__ psi, double2* __restrict__ lam,
                                                          u64 dim, const AdjTerm* __restrict__ terms, u32 nterms,
                                                          double* __restrict__ epart) {
    __shared__ double red[32];
    const u64 off = (u64)blockIdx.y * dim;
    const double2* P = psi + off;
    double acc = 0.0;
    for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < dim; x += (u64)gridDim.x * blockDim.x) {
        double2 s = make_double2(0.0, 0.0);
        for (u32 k = 0; k < nterms; ++k) {
            const AdjTerm tm = terms[k];
            const u64 y = x ^ tm.bf;
            const double sg = (__popcll(y & tm.pf) & 1) ? -tm.coef : tm.coef;
            const double2 v = ipow(P[y], tm.rot);
            s.x = fma(sg, v.x, s.x);
            s.y = fma(sg, v.y, s.y);
        }
        lam[off + x] = s;
        const double2 p = P[x];
        acc += p.x * s.x + p.y * s.y;
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) epart[(u64)blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// grad[b][p] = sum over occurrences of p (in order) and blocks (in order).
__global__ void k_adj_finish(const double* __restrict__ partials, const double* __restrict__ epart,
                             const u32* __restrict__ occ_param, u32 nocc, u32 nparams, u32 batch, u32 nblk,
                             double* __restrict__ grad, double* __restrict__ energy) {
    const u64 idx = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < (u64)batch * nparams) {
        const u32 b = (u32)(idx / nparams), p = (u32)(idx % nparams);
        double s = 0.0;
        for (u32 o = 0; o < nocc; ++o) {
            if (occ_param[o] != p) continue;
            const double* src = partials + ((u64)o * batch + b) * nblk;
            for (u32 k = 0; k < nblk; ++k) s += src[k];
        }
        grad[idx] = s;
    }
    if (idx < batch) {
        double e = 0.0;
        for (u32 k = 0; k < nblk; ++k) e += epart[idx * nblk + k];
        energy[idx] = e;
    }
}

}  // namespace
}  // namespace qm

using namespace qm;

struct qm_adjoint {
    u32 n = 0, nparams = 0;
    std::vector<AdjGate> gates;
    std::vector<u32> occ_param;  // parameter of each rotation occurrence
    std::vector<PassOp> pops;    // the gates as pass operations, pass after pass
    std::vector<PassArgs> passes;
    int device = -1;
    u32* d_occ = nullptr;
    PassOp* d_ops = nullptr;
    void release() {
        if (device >= 0) {
            (void)hipSetDevice(device);
            if (d_occ) (void)hipFree(d_occ);
            if (d_ops) (void)hipFree(d_ops);
        }
        d_occ = nullptr;
        d_ops = nullptr;
        device = -1;
    }
    void plan();
};

// Groups the gates into passes (see the file comment): a gate joins the open
// pass when the qubits it needs in the block (a rotation's X/Y qubits, a
// non-diagonal fixed gate's qubits; none for a diagonal gate) fit in
// kPassQubits together with the pass's, and the pass has a free rotation slot.
void qm_adjoint::plan() {
    pops.clear();
    passes.clear();
    if (n < (u32)kPassQubits) return;  // tiny states keep one kernel per gate
    auto diagonal = [](const AdjGate& g) {
        const u32 d = 1u << g.k;
        for (u32 r = 0; r < d; ++r)
            for (u32 c = 0; c < d; ++c)
                if (r != c && (g.m[r * d + c].x != 0.0 || g.m[r * d + c].y != 0.0)) return false;
        return true;
    };
    auto needs = [&](const AdjGate& g) -> u64 {
        if (g.kind == 1) return g.bf;
        if (diagonal(g)) return 0;
        u64 m = 1ULL << g.q[0];
        if (g.k == 2) m |= 1ULL << g.q[1];
        return m;
    };
    for (const AdjGate& g : gates)
        if (__builtin_popcountll(needs(g)) > kPassQubits) return;  // keeps one kernel per gate
    size_t i = 0;
    while (i < gates.size()) {
        u64 blk = 0;
        size_t j = i;
        u32 nrot = 0;
        for (; j < gates.size(); ++j) {
            const u64 nb = blk | needs(gates[j]);
            if (__builtin_popcountll(nb) > kPassQubits) break;
            if (gates[j].kind == 1 && nrot == (u32)kPassRots) break;
            blk = nb;
            nrot += gates[j].kind == 1 ? 1 : 0;
        }
        // pad the block with the highest free qubits (consecutive lanes then read
        // consecutive amplitudes of every member)
        for (int q = (int)n - 1; q >= 0 && __builtin_popcountll(blk) < kPassQubits; --q) blk |= 1ULL << q;
        PassArgs pa{};
        int pos[64];
        int k = 0;
        for (u32 q = 0; q < n; ++q) {
            pos[q] = -1;
            if ((blk >> q) & 1) {
                pa.blk[k] = (unsigned char)q;
                pos[q] = k++;
            }
        }
        for (int y = 0; y < kPassN; ++y) {
            u64 o = 0;
            for (int b = 0; b < kPassQubits; ++b)
                if ((y >> b) & 1) o |= 1ULL << pa.blk[b];
            pa.offs[y] = o;
        }
        pa.op0 = (u32)pops.size();
        pa.nops = (u32)(j - i);
        pa.nrot = 0;
        auto local = [&](u64 mask) {
            u32 l = 0;
            for (u32 q = 0; q < n; ++q)
                if (((mask >> q) & 1) && pos[q] >= 0) l |= 1u << pos[q];
            return l;
        };
        for (size_t g = i; g < j; ++g) {
            const AdjGate& ag = gates[g];
            PassOp op{};
            if (ag.kind == 1) {
                op.type = 0;
                op.lbf = local(ag.bf);
                op.lpf = local(ag.pf);
                op.opf = ag.pf & ~blk;
                op.rot = ag.rot;
                op.slot = pa.nrot;
                op.occ = ag.occ;
                pa.param[pa.nrot] = ag.param;
                pa.coef[pa.nrot] = ag.coef;
                ++pa.nrot;
            } else if (diagonal(ag)) {
                op.type = 3;
                op.dk = ag.k;
                const u32 d = 1u << ag.k;
                for (u32 t = 0; t < ag.k; ++t) {
                    const u32 q = ag.q[t];
                    op.dloc[t] = pos[q] >= 0 ? (u32)pos[q] : 255u;
                    op.dpos[t] = q;
                }
                for (u32 x = 0; x < d; ++x) op.m[x] = ag.m[x * d + x];
            } else {
                op.type = ag.k == 1 ? 1 : 2;
                op.a = (u32)pos[ag.q[0]];
                op.b = ag.k == 2 ? (u32)pos[ag.q[1]] : 0;
                for (u32 x = 0; x < (1u << (2 * ag.k)); ++x) op.m[x] = ag.m[x];
            }
            pops.push_back(op);
        }
        passes.push_back(pa);
        i = j;
    }
}

extern "C" {

qm_adjoint* qm_adjoint_create(UINT n, const qm_adjoint_gate* gates, UINT gate_count, UINT nparams) {
    const char* what = "qm_adjoint_create";
    if (n < 1 || n > 34 || (gate_count && !gates)) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: qubit count %u outside 1..34 or null gates", what, n);
        return nullptr;
    }
    auto* a = new qm_adjoint();
    a->n = n;
    a->nparams = nparams;
    for (UINT i = 0; i < gate_count; ++i) {
        const qm_adjoint_gate& in = gates[i];
        AdjGate g{};
        const u32 k = in.qubit_count;
        u64 mask = 0;
        bool ok = k >= 1 && k <= 4;
        for (u32 j = 0; ok && j < k; ++j) {
            ok = in.qubits[j] < n && !((mask >> in.qubits[j]) & 1);
            mask |= 1ULL << in.qubits[j];
        }
        if (ok && in.kind == QM_ADJ_ROT) {
            g.kind = 1;
            u32 rot = 0;
            for (u32 j = 0; ok && j < k; ++j) {
                const u64 bit = 1ULL << in.qubits[j];
                switch (in.paulis[j]) {
                    case 1: g.bf |= bit; break;
                    case 2: g.bf |= bit; g.pf |= bit; ++rot; break;
                    case 3: g.pf |= bit; break;
                    default: ok = false;
                }
            }
            ok = ok && in.param < nparams;
            g.rot = rot;
            g.pivot = g.bf ? 63 - __builtin_clzll(g.bf) : 0;
            g.param = in.param;
            g.coef = in.coef;
            g.occ = (u32)a->occ_param.size();
            if (ok) a->occ_param.push_back(in.param);
        } else if (ok) {
            g.kind = 0;
            const bool two = in.kind == QM_BATCH_CZ || in.kind == QM_BATCH_CNOT;
            ok = (in.kind <= QM_BATCH_CNOT) && (two ? k == 2 : k <= 2) && (two || in.matrix);
            g.k = k;
            g.q[0] = in.qubits[0];
            g.q[1] = k > 1 ? in.qubits[1] : 0;
            const u32 d = 1u << k;
            for (u32 r = 0; ok && r < d; ++r)
                for (u32 c = 0; c < d; ++c) {
                    double2 v = make_double2(0, 0);
                    if (in.kind == QM_BATCH_DENSE) v = make_double2(in.matrix[r * d + c].re, in.matrix[r * d + c].im);
                    else if (in.kind == QM_BATCH_DIAG && r == c) v = make_double2(in.matrix[r].re, in.matrix[r].im);
                    else if (in.kind == QM_BATCH_CZ && r == c) v = make_double2(r == 3 ? -1.0 : 1.0, 0.0);
                    else if (in.kind == QM_BATCH_CNOT) {
                        // bit 0 <-> control, bit 1 <-> target
                        const u32 cb = c & 1, tb = (c >> 1) & 1;
                        v = make_double2(r == (cb | ((tb ^ cb) << 1)) ? 1.0 : 0.0, 0.0);
                    }
                    g.m[r * d + c] = v;
                }
        }
        if (!ok) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: gate %u is not a 1..2-qubit fixed gate or a valid rotation", what, i);
            delete a;
            return nullptr;
        }
        a->gates.push_back(g);
    }
    a->plan();
    return a;
}

void qm_adjoint_destroy(qm_adjoint* a) {
    if (!a) return;
    a->release();
    delete a;
}

void qm_adjoint_stats(const qm_adjoint* a, ITYPE* passes, ITYPE* gates) {
    if (passes) *passes = a ? (ITYPE)a->passes.size() : 0;
    if (gates) *gates = a ? (ITYPE)a->gates.size() : 0;
}

ITYPE qm_adjoint_workspace_bytes(const qm_adjoint* a, UINT batch) {
    if (!a) return 0;
    const u64 states = 2ULL * batch * (16ULL << a->n);
    const u64 parts = ((u64)a->occ_param.size() * batch + batch) * kAdjBlocks * sizeof(double);
    return states + parts + 256;
}

int qm_adjoint_run(qm_adjoint* a, UINT batch, const double* theta, const ITYPE* bf, const ITYPE* pf, const UINT* rot,
                   const double* coefs, UINT nterms, double* energy, double* grad, void* workspace, void* stream,
                   unsigned dev) {
    const char* what = "qm_adjoint_run";
    if (!a || batch == 0 || batch > 65535 || !theta || !energy || !workspace || (nterms && (!bf || !pf || !rot || !coefs)) ||
        (a->nparams && !grad)) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null argument or batch outside 1..65535", what);
        return QM_ERR_INVALID_ARGUMENT;
    }
    for (UINT k = 0; k < nterms; ++k)
        if (((bf[k] | pf[k]) >> a->n) != 0) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: term %u acts beyond %u qubits", what, k, a->n);
            return QM_ERR_INVALID_ARGUMENT;
        }
    if (!use_device(dev)) return QM_ERR_HIP;
    hipStream_t st = stream_of(stream);
    if (a->device != (int)dev) {
        a->release();
        if (!a->occ_param.empty()) {
            if (!hip_ok(hipMalloc((void**)&a->d_occ, a->occ_param.size() * sizeof(u32)), what) ||
                !hip_ok(hipMemcpy(a->d_occ, a->occ_param.data(), a->occ_param.size() * sizeof(u32),
                                  hipMemcpyHostToDevice),
                        what))
                return QM_ERR_HIP;
        }
        if (!a->pops.empty()) {
            if (!hip_ok(hipMalloc((void**)&a->d_ops, a->pops.size() * sizeof(PassOp)), what) ||
                !hip_ok(hipMemcpy(a->d_ops, a->pops.data(), a->pops.size() * sizeof(PassOp), hipMemcpyHostToDevice),
                        what))
                return QM_ERR_HIP;
        }
        a->device = (int)dev;
    }
    Scratch* sc = scratch_for(st);
    if (!sc) return QM_ERR_HIP;
    const u64 dim = 1ULL << a->n;
    double2* psi = (double2*)workspace;
    double2* lam = psi + (u64)batch * dim;
    double* partials = (double*)(lam + (u64)batch * dim);
    double* epart = partials + (u64)a->occ_param.size() * batch * kAdjBlocks;
    std::vector<AdjTerm> ht(nterms);
    for (UINT k = 0; k < nterms; ++k) ht[k] = AdjTerm{bf[k], pf[k], rot[k], 0, coefs[k]};
    const AdjTerm* dterms = nterms ? (const AdjTerm*)upload(sc, ht.data(), sizeof(AdjTerm) * nterms) : nullptr;
    if (nterms && !dterms) return QM_ERR_HIP;
    const unsigned nblk = kAdjBlocks;
    const dim3 grid(nblk, batch);
    k_adj_zero<<<grid, kAdjThreads, 0, st>>>(psi, dim);
    if (!a->passes.empty()) {
        for (const PassArgs& pa : a->passes)
            k_adj_pass<false><<<grid, kAdjThreads, 0, st>>>(psi, lam, dim, pa, a->d_ops, theta, a->nparams, partials);
    } else {
        for (const AdjGate& g : a->gates) {
            if (g.kind == 1) k_adj_rot<true><<<grid, kAdjThreads, 0, st>>>(psi, lam, dim, g, theta, a->nparams, partials);
            else k_adj_fixed<false, false><<<grid, kAdjThreads, 0, st>>>(psi, lam, dim, g);
        }
    }
    k_adj_hpsi<<<grid, kAdjThreads, 0, st>>>(psi, lam, dim, dterms, nterms, epart);
    if (!a->passes.empty()) {
        for (size_t i = a->passes.size(); i-- > 0;)
            k_adj_pass<true><<<grid, kAdjThreads, 0, st>>>(psi, lam, dim, a->passes[i], a->d_ops, theta, a->nparams,
                                                            partials);
    } else {
        for (size_t i = a->gates.size(); i-- > 0;) {
            const AdjGate& g = a->gates[i];
            if (g.kind == 1) k_adj_rot<false><<<grid, kAdjThreads, 0, st>>>(psi, lam, dim, g, theta, a->nparams, partials);
            else k_adj_fixed<true, true><<<grid, kAdjThreads, 0, st>>>(psi, lam, dim, g);
        }
    }
    const u64 outs = std::max<u64>((u64)batch * a->nparams, batch);
    k_adj_finish<<<(unsigned)((outs + 255) / 256), 256, 0, st>>>(partials, epart, a->d_occ, (u32)a->occ_param.size(),
                                                                a->nparams, batch, nblk, grad, energy);
    if (!hip_ok(hipGetLastError(), what)) return QM_ERR_HIP;
    finish_update(st, what);
    return QM_OK;
}

}  // extern "C"
