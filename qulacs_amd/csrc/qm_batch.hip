// Batched executor: many small circuits, each run from |0...0> and reduced to
// the expectation value of its observable, in one launch.
//
// What it replaces: for each circuit, QuantumCircuit::update_quantum_state
// (src/cppsim/circuit.cpp:46-57, one kernel launch and synchronisation per gate
// on the GPU backend) and Observable::get_expectation_value
// (src/cppsim/observable.cpp:56-58 -> general_quantum_operator.cpp:105-141, one
// reduction per term).
//
// Design:
//   * one work-group (1024 threads, 512 at K = 4) per circuit, circuits taken from a queue
//     (an atomic counter) in descending-cost order, so one launch covers a batch
//     of uneven sizes; every work-group leaves the loop once the queue is empty;
//   * the state lives in LDS: 2^lq complex128 with lq = min(n, 13) (128 KiB at
//     13 qubits); circuits of 14..20 qubits keep the state in device memory and
//     are processed in 2^13-amplitude tiles, one "stage" at a time: a stage is a
//     set of gates whose qubits all lie in one 13-qubit window (qubits 0..3 are
//     always in the window, so a tile is made of 256-byte contiguous runs);
//   * inside a stage the gates are grouped into passes: each pass names K
//     "block" qubits; a thread loads the 2^K amplitudes that differ in those
//     bits into registers, applies every gate of the pass (dense gates on block
//     qubits, diagonal gates and CZ on any local qubits) and writes them back,
//     so one LDS round trip serves several gates;
//   * LDS slots are XOR-swizzled (slot = i ^ fold3(i)) so that the 16-byte
//     accesses of a wave spread over the banks whichever block bits a pass uses;
//     fold3 is linear over GF(2), so slot(base | off) = slot(base) ^ slot(off);
//   * the observable sum_k c_k <P_k> is reduced per work-group in a fixed order
//     (deterministic), with P_k in the mask form of the reference
//     (src/csim/utility.cpp:12-43): P|x> = i^rot (-1)^popcount(x & pf) |x ^ bf>.
//
// Scheduling (host): gates are scanned in order; a gate joins the current
// stage / pass unless it touches a qubit of a gate skipped earlier in the same
// scan (then it is skipped too and its qubits are blocked) or it does not fit
// the window / block.  A gate is therefore only moved ahead of gates it
// commutes with trivially (disjoint qubits), and every scan schedules at least
// its first gate.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "qm_internal.hpp"

namespace qm {
namespace {

using cd = std::complex<double>;

constexpr u32 kMaxLocal = 13;   // qubits resident in LDS
constexpr u32 kLowFixed = 4;    // qubits 0..3 are local in every stage
// Work-group size per register-block width: 8 waves at K = 4 (the register
// block limits a SIMD to 2 waves), 16 waves below, so a CU holding one
// 128 KiB tile still has 4 waves per SIMD to hide LDS latency.
template <int K>
constexpr u32 block_threads() { return K >= 4 ? 512u : 1024u; }
constexpr u32 kMaxBlockThreads = 1024;
constexpr u32 kMinQubits = 4;   // smaller circuits run on 4 qubits (extra qubits stay |0>)

// Sub-operation kinds (low byte of DSub::type).
// LAYER: CZ gates, then one 2x2 matrix on each block bit of a mask.  type bits
//   8..11: the block-bit mask; bits 16..31: the sign pattern over the register
//   block of the CZs between two block bits; bit 15: a and m hold, for block
//   bit i, the mask of local positions outside the block that share a CZ with
//   it (16 bits each: a = M0 | M1 << 16, m = M2 | M3 << 16), so the sign of
//   register slot j changes with the parity of base & M_i where bit i of j is
//   set; bits 12..14: the number (<= 4) of CZs between two positions outside
//   the block, byte c of pairs = p | q << 4.
//   data: the matrices (4 complex each) of the mask's bits in ascending order.
// D2: 4x4 on block bits byte 0 / byte 1 of a; DK: 2^K x 2^K on the block;
// DIAG: diagonal on the m local positions in the bytes of a.
enum : u32 { SUB_LAYER = 0, SUB_D2 = 1, SUB_DK = 2, SUB_DIAG = 3 };
// bits 16 i .. 16 i + 15: the register-block indices j (< 16) with bit i set
constexpr u64 kBlockPatterns = 0xFF00F0F0CCCCAAAAull;
constexpr u32 kCzPairsMax = 4;

struct DCirc {
    u32 n, lq, nstage, stage0, term0, nterm;
    u64 state_off;  // complex128 offset of the circuit's state in the scratch buffer
};
// Qubit lists are packed one byte per entry into 64-bit words so that the
// kernel indexes them with shifts (a runtime index into an array would put it
// in private memory).
struct DStage {
    u64 win[2];   // local bit 4 + i <-> qubit byte i of win (i < lq - 4 <= 9)
    u64 comp[4];  // tile bit i <-> qubit byte i of comp (i < n - lq <= 32)
    u32 npass, pass0;
};
__host__ __device__ __forceinline__ u32 byte_at(u64 w, u32 i) { return (u32)(w >> (8 * i)) & 0xffu; }
__host__ __forceinline__ void set_byte(u64& w, u32 i, u32 v) { w = (w & ~(0xffULL << (8 * i))) | ((u64)v << (8 * i)); }
struct DPass {
    unsigned char pos[4];  // block bits (local positions, ascending)
    u32 nsub, sub0;
    u32 flags;             // kPassHasDiag: a DIAG sub-operation (one register block per thread)
    unsigned short sbit[4];   // slot_of(1 << pos[i]), computed on the host
};
constexpr u32 kPassHasDiag = 1;
struct DSub {
    u32 type;
    u32 a;      // D2: block-bit indices (bytes); DIAG: local positions (bytes); LAYER: CZ masks
    u32 m;      // DIAG: qubit count; LAYER: CZ masks
    u32 data;   // offset into the complex pool
    u32 pairs;  // LAYER: CZs between two positions outside the block
    u32 pad[3];
};
struct DTerm {
    u64 bf, pf;
    u32 rot, pad;
    double coef;
};

struct HOp {
    u32 type;  // 0 dense, 1 diag, 2 cz
    std::vector<u32> q;
    std::vector<cd> m;
    u64 mask() const {
        u64 r = 0;
        for (u32 x : q) r |= 1ULL << x;
        return r;
    }
};
struct HCirc {
    u32 n = 0;  // as given
    u32 ne = 0; // as run (>= kMinQubits)
    std::vector<HOp> ops;
    std::vector<DTerm> terms;
};


__host__ __device__ __forceinline__ u32 fold3(u32 x) { return ((x >> 3) ^ (x >> 6) ^ (x >> 9) ^ (x >> 12)) & 7u; }
__host__ __device__ __forceinline__ u32 slot_of(u32 x) { return x ^ fold3(x); }

// (x, y) <- (m0 x + m1 y, m2 x + m3 y), complex, updated in place; the matrix
// entries are wave-uniform (scalar registers).  Written as one asm block so the
// four results land in the registers of the four inputs: the sub-operation loop
// of run_passes then keeps the register block at fixed registers across all
// sub-operation kinds instead of copying it at every merge point.  The last
// reader of each input writes a result into it (x.re holds the x.re chain from
// its 4th op on), so three temporaries suffice.
__device__ __forceinline__ void pair_update(double2& x, double2& y, const double2 m0, const double2 m1,
                                            const double2 m2, const double2 m3) {
    double tb, tc, td;
    asm("v_mul_f64 %[tb], %[m0i], %[xr]\n\t"          // B = x'.im
        "v_mul_f64 %[tc], %[m2r], %[xr]\n\t"          // C = y'.re
        "v_mul_f64 %[td], %[m2i], %[xr]\n\t"          // D = y'.im
        "v_mul_f64 %[xr], %[m0r], %[xr]\n\t"          // A = x'.re
        "v_fma_f64 %[xr], -%[m0i], %[xi], %[xr]\n\t"
        "v_fma_f64 %[tb], %[m0r], %[xi], %[tb]\n\t"
        "v_fma_f64 %[tc], -%[m2i], %[xi], %[tc]\n\t"
        "v_fma_f64 %[td], %[m2r], %[xi], %[td]\n\t"
        "v_fma_f64 %[xr], %[m1r], %[yr], %[xr]\n\t"
        "v_fma_f64 %[tb], %[m1i], %[yr], %[tb]\n\t"
        "v_fma_f64 %[tc], %[m3r], %[yr], %[tc]\n\t"
        "v_fma_f64 %[td], %[m3i], %[yr], %[td]\n\t"
        "v_fma_f64 %[xi], %[m1r], %[yi], %[tb]\n\t"
        "v_fma_f64 %[yr], -%[m3i], %[yi], %[tc]\n\t"
        "v_fma_f64 %[xr], -%[m1i], %[yi], %[xr]\n\t"
        "v_fma_f64 %[yi], %[m3r], %[yi], %[td]"
        : [xr] "+v"(x.x), [xi] "+v"(x.y), [yr] "+v"(y.x), [yi] "+v"(y.y), [tb] "=&v"(tb), [tc] "=&v"(tc),
          [td] "=&v"(td)
        : [m0r] "s"(m0.x), [m0i] "s"(m0.y), [m1r] "s"(m1.x), [m1i] "s"(m1.y), [m2r] "s"(m2.x), [m2i] "s"(m2.y),
          [m3r] "s"(m3.x), [m3i] "s"(m3.y));
}

// v <- v * z, complex, in place.
__device__ __forceinline__ void cmul_inplace(double2& v, const double2 z) {
    double t;
    asm("v_mul_f64 %[t], %[vr], %[zi]\n\t"
        "v_mul_f64 %[vr], %[vr], %[zr]\n\t"
        "v_fma_f64 %[vr], -%[vi], %[zi], %[vr]\n\t"
        "v_fma_f64 %[vi], %[vi], %[zr], %[t]"
        : [vr] "+v"(v.x), [vi] "+v"(v.y), [t] "=&v"(t)
        : [zr] "v"(z.x), [zi] "v"(z.y));
}

// Two independent pair updates interleaved instruction by instruction, so a
// dependent FMA follows its producer 8 instructions later (pair_update alone:
// 4), which keeps the FP64 pipeline busy with only two waves per SIMD.
__device__ __forceinline__ void pair_update2(double2& x0, double2& y0, double2& x1, double2& y1, const double2 m0,
                                             const double2 m1, const double2 m2, const double2 m3) {
    double tb0, tc0, td0, tb1, tc1, td1;
    asm("v_mul_f64 %[tb0], %[m0i], %[xr0]\n\t"
        "v_mul_f64 %[tb1], %[m0i], %[xr1]\n\t"
        "v_mul_f64 %[tc0], %[m2r], %[xr0]\n\t"
        "v_mul_f64 %[tc1], %[m2r], %[xr1]\n\t"
        "v_mul_f64 %[td0], %[m2i], %[xr0]\n\t"
        "v_mul_f64 %[td1], %[m2i], %[xr1]\n\t"
        "v_mul_f64 %[xr0], %[m0r], %[xr0]\n\t"
        "v_mul_f64 %[xr1], %[m0r], %[xr1]\n\t"
        "v_fma_f64 %[xr0], -%[m0i], %[xi0], %[xr0]\n\t"
        "v_fma_f64 %[xr1], -%[m0i], %[xi1], %[xr1]\n\t"
        "v_fma_f64 %[tb0], %[m0r], %[xi0], %[tb0]\n\t"
        "v_fma_f64 %[tb1], %[m0r], %[xi1], %[tb1]\n\t"
        "v_fma_f64 %[tc0], -%[m2i], %[xi0], %[tc0]\n\t"
        "v_fma_f64 %[tc1], -%[m2i], %[xi1], %[tc1]\n\t"
        "v_fma_f64 %[td0], %[m2r], %[xi0], %[td0]\n\t"
        "v_fma_f64 %[td1], %[m2r], %[xi1], %[td1]\n\t"
        "v_fma_f64 %[xr0], %[m1r], %[yr0], %[xr0]\n\t"
        "v_fma_f64 %[xr1], %[m1r], %[yr1], %[xr1]\n\t"
        "v_fma_f64 %[tb0], %[m1i], %[yr0], %[tb0]\n\t"
        "v_fma_f64 %[tb1], %[m1i], %[yr1], %[tb1]\n\t"
        "v_fma_f64 %[tc0], %[m3r], %[yr0], %[tc0]\n\t"
        "v_fma_f64 %[tc1], %[m3r], %[yr1], %[tc1]\n\t"
        "v_fma_f64 %[td0], %[m3i], %[yr0], %[td0]\n\t"
        "v_fma_f64 %[td1], %[m3i], %[yr1], %[td1]\n\t"
        "v_fma_f64 %[xi0], %[m1r], %[yi0], %[tb0]\n\t"
        "v_fma_f64 %[xi1], %[m1r], %[yi1], %[tb1]\n\t"
        "v_fma_f64 %[yr0], -%[m3i], %[yi0], %[tc0]\n\t"
        "v_fma_f64 %[yr1], -%[m3i], %[yi1], %[tc1]\n\t"
        "v_fma_f64 %[xr0], -%[m1i], %[yi0], %[xr0]\n\t"
        "v_fma_f64 %[xr1], -%[m1i], %[yi1], %[xr1]\n\t"
        "v_fma_f64 %[yi0], %[m3r], %[yi0], %[td0]\n\t"
        "v_fma_f64 %[yi1], %[m3r], %[yi1], %[td1]"
        : [xr0] "+v"(x0.x), [xi0] "+v"(x0.y), [yr0] "+v"(y0.x), [yi0] "+v"(y0.y), [xr1] "+v"(x1.x),
          [xi1] "+v"(x1.y), [yr1] "+v"(y1.x), [yi1] "+v"(y1.y), [tb0] "=&v"(tb0), [tc0] "=&v"(tc0),
          [td0] "=&v"(td0), [tb1] "=&v"(tb1), [tc1] "=&v"(tc1), [td1] "=&v"(td1)
        : [m0r] "s"(m0.x), [m0i] "s"(m0.y), [m1r] "s"(m1.x), [m1i] "s"(m1.y), [m2r] "s"(m2.x), [m2i] "s"(m2.y),
          [m3r] "s"(m3.x), [m3i] "s"(m3.y));
}

// One 2x2 matrix on block bit P of each of the G register blocks of a thread.
// With two blocks, the pair updates interleave pair q of block 0 with pair q of
// block 1; with one block, pairs q and q + 1.
template <int K, int P, int G>
__device__ __forceinline__ void sub_d1(double2 (&v)[G][1 << K], const double2* mp) {
    const double2 m0 = mp[0], m1 = mp[1], m2 = mp[2], m3 = mp[3];
    // the 2^(K-1) pairs (j, j | 2^P)
    constexpr int NP = 1 << (K - 1);
    auto pair_j = [](int q) { return ((q >> P) << (P + 1)) | (q & ((1 << P) - 1)); };
    if constexpr (G == 2) {
#pragma unroll
        for (int q = 0; q < NP; ++q)
            pair_update2(v[0][pair_j(q)], v[0][pair_j(q) | (1 << P)], v[1][pair_j(q)], v[1][pair_j(q) | (1 << P)], m0,
                         m1, m2, m3);
    } else {
#pragma unroll
        for (int q = 0; q < NP; q += 2)
            pair_update2(v[0][pair_j(q)], v[0][pair_j(q) | (1 << P)], v[0][pair_j(q + 1)],
                         v[0][pair_j(q + 1) | (1 << P)], m0, m1, m2, m3);
    }
}

template <int K, int P0, int P1>
__device__ __forceinline__ void sub_d2(double2* v, const double2* mp) {
    double2 m[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = mp[i];
#pragma unroll
    for (int j = 0; j < (1 << K); ++j) {
        if (j & ((1 << P0) | (1 << P1))) continue;
        const double2 i0 = v[j], i1 = v[j | (1 << P0)], i2 = v[j | (1 << P1)], i3 = v[j | (1 << P0) | (1 << P1)];
        v[j] = cmac(cmac(cmac(cmul(m[0], i0), m[1], i1), m[2], i2), m[3], i3);
        v[j | (1 << P0)] = cmac(cmac(cmac(cmul(m[4], i0), m[5], i1), m[6], i2), m[7], i3);
        v[j | (1 << P1)] = cmac(cmac(cmac(cmul(m[8], i0), m[9], i1), m[10], i2), m[11], i3);
        v[j | (1 << P0) | (1 << P1)] = cmac(cmac(cmac(cmul(m[12], i0), m[13], i1), m[14], i2), m[15], i3);
    }
}

template <int K>
__device__ __forceinline__ void sub_dk(double2* v, const double2* mp) {
    constexpr int D = 1 << K;
    double2 out[D];
#pragma unroll
    for (int r = 0; r < D; ++r) {
        double2 acc = cmul(mp[r * D], v[0]);
#pragma unroll
        for (int c = 1; c < D; ++c) acc = cmac(acc, mp[r * D + c], v[c]);
        out[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < D; ++r) v[r] = out[r];
}

#define QM_D2_CASE(A, B) \
    case (A) * 4 + (B):  \
        if constexpr (K > (A) && K > (B)) sub_d2<K, A, B>(v, mp); \
        break;
template <int K>
__device__ __forceinline__ void apply_d2(double2* v, u32 p0, u32 p1, const double2* mp) {
    switch (p0 * 4 + p1) {
        QM_D2_CASE(0, 1) QM_D2_CASE(0, 2) QM_D2_CASE(0, 3)
        QM_D2_CASE(1, 0) QM_D2_CASE(1, 2) QM_D2_CASE(1, 3)
        QM_D2_CASE(2, 0) QM_D2_CASE(2, 1) QM_D2_CASE(2, 3)
        QM_D2_CASE(3, 0) QM_D2_CASE(3, 1) QM_D2_CASE(3, 2)
        default: break;
    }
}
#undef QM_D2_CASE

// This thread's share of sum_k c_k Re <psi|P_k|psi>, reading the state from LDS
// (swizzled slots) or from device memory.
template <bool FROM_LDS>
__device__ __forceinline__ double energy_sum(const double2* __restrict__ gs, const double2* st, const DCirc& c,
                                             const DTerm* __restrict__ terms) {
    double acc = 0.0;
    const u64 dim = 1ULL << c.n;
    for (u64 x = threadIdx.x; x < dim; x += blockDim.x) {
        const double2 px = FROM_LDS ? st[slot_of((u32)x)] : gs[x];
        for (u32 k = 0; k < c.nterm; ++k) {
            const DTerm tm = terms[c.term0 + k];
            const u64 y = x ^ tm.bf;
            const double2 py = FROM_LDS ? st[slot_of((u32)y)] : gs[y];  // y == x when bf == 0
            // psi_x * conj(psi_y) * i^rot * (-1)^popcount(x & pf)
            const double re0 = px.x * py.x + px.y * py.y, im0 = px.y * py.x - px.x * py.y;
            double re;
            switch (tm.rot & 3u) {
                case 0: re = re0; break;
                case 1: re = -im0; break;
                case 2: re = -re0; break;
                default: re = im0; break;
            }
            if (__popcll(x & tm.pf) & 1) re = -re;
            acc = fma(tm.coef, re, acc);
        }
    }
    return acc;
}

__global__ void k_batch_reset(int* counter) {
    if (threadIdx.x == 0) *counter = 0;
}

// hi_tab[j] = global offset of local index j << 4 (the window qubits 4.. of the
// stage).  TI = u32 holds offsets below 2^32 (states of at most 32 qubits);
// larger states use u64.
template <typename TI>
__device__ __forceinline__ void build_hi_tab(TI* hi_tab, u32 lq, const DStage& sg) {
    const u64 win0 = sg.win[0], win1 = sg.win[1];
    for (u32 j = threadIdx.x; j < (1u << (lq - kLowFixed)); j += blockDim.x) {
        TI o = 0;
        for (u32 i = 0; i + kLowFixed < lq; ++i) o |= (TI)((j >> i) & 1u) << byte_at(i < 8 ? win0 : win1, i & 7);
        hi_tab[j] = o;
    }
}

// global index of the first amplitude of tile t (tile bits deposited on the non-window qubits)
__device__ __forceinline__ u64 tile_base(u64 t, u32 tl, const DStage& sg) {
    const u64 c0 = sg.comp[0], c1 = sg.comp[1], c2 = sg.comp[2], c3 = sg.comp[3];
    u64 tb = 0;
    for (u32 i = 0; i < tl; ++i) {
        const u64 w = i < 8 ? c0 : i < 16 ? c1 : i < 24 ? c2 : c3;
        tb |= ((t >> i) & 1ULL) << byte_at(w, i & 7);
    }
    return tb;
}

template <typename TI>
__device__ __forceinline__ u64 tile_index(u64 tb, const TI* hi_tab, u32 i) {
    return tb + hi_tab[i >> kLowFixed] + (i & ((1u << kLowFixed) - 1));
}

// One LAYER sub-operation on the G register blocks of a thread, block g with
// block base base[g] (local index with zeros at the block positions).  The
// sub-operation is decoded once for all blocks.
template <int K, int G>
__device__ __forceinline__ void apply_layer(double2 (&v)[G][1 << K], const DSub& sub, const u32 (&base)[G],
                                            const double2* __restrict__ pool) {
    constexpr int D = 1 << K;
    const u32 npair = (sub.type >> 12) & 7u, d1 = (sub.type >> 8) & 15u;
    if (sub.type >> 12) {
        // bit j of sgn[g]: amplitude j of block g changes sign; base is zero at
        // the block positions, so a local position p outside the block is
        // (base >> p) & 1 for the whole block
        u32 sgn[G];
#pragma unroll
        for (int g = 0; g < G; ++g) sgn[g] = sub.type >> 16;
        if (sub.type & 0x8000u) {
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const u32 mi = ((i < 2 ? sub.a : sub.m) >> (16 * (i & 1))) & 0xffffu;
                constexpr u32 kPat[4] = {(u32)(kBlockPatterns & 0xffffu), (u32)((kBlockPatterns >> 16) & 0xffffu),
                                         (u32)((kBlockPatterns >> 32) & 0xffffu), (u32)(kBlockPatterns >> 48)};
#pragma unroll
                for (int g = 0; g < G; ++g) sgn[g] ^= (0u - ((u32)__popc(base[g] & mi) & 1u)) & kPat[i];
            }
        }
        for (u32 c = 0; c < npair; ++c) {
            const u32 w = sub.pairs >> (8 * c), p0 = w & 15u, p1 = (w >> 4) & 15u;
#pragma unroll
            for (int g = 0; g < G; ++g) sgn[g] ^= 0u - ((base[g] >> p0) & (base[g] >> p1) & 1u);
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                // flip the sign bits of both components where bit j of sgn is set
                const int f = (int)((sgn[g] << (31 - j)) & 0x80000000u);
                v[g][j].x = __hiloint2double(__double2hiint(v[g][j].x) ^ f, __double2loint(v[g][j].x));
                v[g][j].y = __hiloint2double(__double2hiint(v[g][j].y) ^ f, __double2loint(v[g][j].y));
            }
    }
    const double2* mp = pool + sub.data;
    if (d1 & 1u) {
        sub_d1<K, 0, G>(v, mp);
        mp += 4;
    }
    if (d1 & 2u) {
        sub_d1<K, 1, G>(v, mp);
        mp += 4;
    }
    if constexpr (K > 2)
        if (d1 & 4u) {
            sub_d1<K, 2, G>(v, mp);
            mp += 4;
        }
    if constexpr (K > 3)
        if (d1 & 8u) sub_d1<K, 3, G>(v, mp);
}

// The sub-operations of one pass on the G register blocks of groups gi0,
// gi0 + T, ..., read from and written back to LDS.  slot_of is linear over
// GF(2), so the slot of base | off[j] is slot_of(base) ^ (XOR of sbit[i] over
// the bits i of j): each lane forms its 2^K slots from the K per-bit offsets
// (the empty asm keeps every partial XOR in a vector register, so the compiler
// does not form the 2^K uniform combinations in scalar registers, which spill).
template <int K, u32 T, int G>
__device__ __forceinline__ void pass_groups(double2* st, const DPass& ps, const u32 (&sbit)[K], u32 gi0,
                                            const DSub* __restrict__ subs, const double2* __restrict__ pool) {
    constexpr int D = 1 << K;
    u32 base[G], slot[G][D];
    double2 v[G][D];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        u32 b = gi0 + (u32)g * T;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const u32 q = ps.pos[i];
            b = (b & ((1u << q) - 1)) | ((b >> q) << (q + 1));
        }
        base[g] = b;
        slot[g][0] = slot_of(b);
#pragma unroll
        for (int j = 1; j < D; ++j) {
            const int lb = __builtin_ctz(j);
            slot[g][j] = slot[g][j & (j - 1)] ^ sbit[lb];
            // one block per thread: keep the chain per lane (see above); two
            // blocks: the 2^K uniform offsets are formed once and shared, which
            // costs fewer vector registers where both blocks are live
            if constexpr (G == 1) asm volatile("" : "+v"(slot[g][j]));
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int j = 0; j < D; ++j) v[g][j] = st[slot[g][j]];
    // LAYER sub-operations run in an inner loop of their own and update the
    // register blocks in place, so the blocks stay at fixed registers there;
    // the other kinds are handled between runs of LAYERs.
    u32 k = 0;
    while (k < ps.nsub) {
        DSub sub = subs[ps.sub0 + k];
        while ((sub.type & 0xffu) == SUB_LAYER) {
            apply_layer<K, G>(v, sub, base, pool);
            if (++k == ps.nsub) break;
            sub = subs[ps.sub0 + k];
        }
        if (k == ps.nsub) break;
        const u32 kind = sub.type & 0xffu;
        const double2* mp = pool + sub.data;
        if (kind == SUB_D2) {
#pragma unroll
            for (int g = 0; g < G; ++g) apply_d2<K>(v[g], byte_at(sub.a, 0), byte_at(sub.a, 1), mp);
        } else if (kind == SUB_DK) {
#pragma unroll
            for (int g = 0; g < G; ++g) sub_dk<K>(v[g], mp);
        } else if (G == 1 && kind == SUB_DIAG) {
            // (passes with a DIAG run with one block per thread: the per-lane
            // reads of the diagonal for two blocks exhaust the registers)
            // bit i of the diagonal index: local position p_i, i.e. bit p_i of
            // base or, for a block position, bit j of the block pattern
            u32 pat[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const u32 p = byte_at(sub.a, i);
                u32 pt = 0;
#pragma unroll
                for (int b = 0; b < K; ++b) pt |= (ps.pos[b] == p) ? (u32)(kBlockPatterns >> (16 * b)) & 0xffffu : 0u;
                pat[i] = i < (int)sub.m ? pt : 0u;
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                u32 tbit[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) tbit[i] = i < (int)sub.m ? (base[g] >> byte_at(sub.a, i)) & 1u : 0u;
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    u32 idx = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i) idx |= ((tbit[i] | (pat[i] >> j)) & 1u) << i;
                    cmul_inplace(v[g][j], mp[idx]);
                }
            }
        }
        ++k;
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int j = 0; j < D; ++j) st[slot[g][j]] = v[g][j];
}

// Applies passes [pass0, pass0 + npass) to the 2^lq amplitudes in LDS with a
// work-group of T threads; ends with a barrier after every pass.  With G = 2
// and exactly 2 T groups of 2^K amplitudes, every thread holds two register
// blocks and decodes each sub-operation once for both.
template <int K, u32 T, int G>
__device__ __forceinline__ void run_passes(double2* st, u32 lq, const DPass* __restrict__ passes, u32 pass0,
                                           u32 npass, const DSub* __restrict__ subs,
                                           const double2* __restrict__ pool) {
    constexpr int D = 1 << K;
    const u32 ngroups = (1u << lq) >> K;
    for (u32 p = 0; p < npass; ++p) {
        const DPass ps = passes[pass0 + p];
        u32 sbit[K];
#pragma unroll
        for (int i = 0; i < K; ++i) sbit[i] = ps.sbit[i];
        if (G == 2 && ngroups == 2 * T && !(ps.flags & kPassHasDiag)) {
            pass_groups<K, T, G>(st, ps, sbit, threadIdx.x, subs, pool);
        } else if (ngroups == T) {
            pass_groups<K, T, 1>(st, ps, sbit, threadIdx.x, subs, pool);
        } else {
            for (u32 g = threadIdx.x; g < ngroups; g += T) pass_groups<K, T, 1>(st, ps, sbit, g, subs, pool);
        }
        __syncthreads();
    }
}

// Tiles of one stage of a batched circuit held in device memory (every stage
// after the first), the next tile loaded into registers while the passes run
// on the current one.  P amplitudes per thread, T threads.  The loop is laid
// out as in stage_tiles: unconditional loads of the next tile at the top, the
// LDS fill from them after the stores of the current tile.
constexpr u32 kBatchTile = 1u << kMaxLocal;
template <int K, u32 P, u32 T, int G>
__device__ __forceinline__ void batch_tiles(double2* st, double2* __restrict__ gs, const u32* hi_tab, u32 lq, u32 tl,
                                            const DStage& sg, const DPass* __restrict__ passes,
                                            const DSub* __restrict__ subs, const double2* __restrict__ pool) {
    const u32 tid = threadIdx.x, ntiles = 1u << tl;
    {
        const u64 tb = tile_base(0, tl, sg);
#pragma unroll
        for (u32 k = 0; k < P; ++k) st[slot_of(tid + k * T)] = gs[tile_index(tb, hi_tab, tid + k * T)];
        __syncthreads();
    }
    for (u32 t = 0;; ++t) {
        double rx[P], ry[P];
        const bool more = t + 1 < ntiles;
        {
            const u64 tb = tile_base(t + 1, tl, sg);
#pragma unroll
            for (u32 k = 0; k < P; ++k) {
                const double2 a = gs[more ? tile_index(tb, hi_tab, tid + k * T) : 0];
                rx[k] = a.x;
                ry[k] = a.y;
            }
        }
        run_passes<K, T, G>(st, lq, passes, sg.pass0, sg.npass, subs, pool);
        const u64 tb = tile_base(t, tl, sg);
#pragma unroll
        for (u32 k = 0; k < P; ++k) gs[tile_index(tb, hi_tab, tid + k * T)] = st[slot_of(tid + k * T)];
        __syncthreads();
        if (!more) break;
#pragma unroll
        for (u32 k = 0; k < P; ++k) st[slot_of(tid + k * T)] = make_double2(rx[k], ry[k]);
        __syncthreads();
    }
}

// The program arrays arrive as separate read-only, non-aliasing arguments so the
// compiler reads them through the scalar cache (inside a by-value struct they
// became per-lane vector loads).
template <int K>
__global__ __launch_bounds__(block_threads<K>()) void k_batch(
    const DCirc* __restrict__ circs, const DStage* __restrict__ stages, const DPass* __restrict__ passes,
    const DSub* __restrict__ subs, const double2* __restrict__ pool, const DTerm* __restrict__ terms,
    const u32* __restrict__ order, double2* __restrict__ scratch, double* __restrict__ energies, int* counter,
    u32 ncirc, int keep) {
    extern __shared__ double2 st[];
    __shared__ u32 hi_tab[1u << (kMaxLocal - kLowFixed)];
    __shared__ double red[16];
    __shared__ int s_slot;
    const u32 tid = threadIdx.x, bd = blockDim.x;

    for (;;) {
        if (tid == 0) s_slot = atomicAdd(counter, 1);
        __syncthreads();
        const int slot = s_slot;
        __syncthreads();
        if (slot >= (int)ncirc) break;  // uniform: every thread read the same slot
        const u32 ci = order[slot];
        const DCirc c = circs[ci];
        const u32 lq = c.lq, L = 1u << lq, tl = c.n - lq;
        double2* gs = scratch + c.state_off;

        for (u32 s = 0; s < c.nstage; ++s) {
            const DStage sg = stages[c.stage0 + s];
            if (tl > 0) {
                build_hi_tab(hi_tab, lq, sg);
                __syncthreads();
            }
            if (tl > 0 && s > 0 && L == kBatchTile) {
                // one register block per thread (two measured slower here)
                batch_tiles<K, kBatchTile / block_threads<K>(), block_threads<K>(), 1>(st, gs, hi_tab, lq, tl,
                                                                                        sg, passes, subs, pool);
                continue;
            }
            for (u32 t = 0; t < (1u << tl); ++t) {
                const u64 tb = tile_base(t, tl, sg);
                if (s == 0) {
                    // |0...0>: slot_of(0) == 0 and amplitude 0 lies in tile 0
                    for (u32 i = tid; i < L; i += bd) st[i] = make_double2((i == 0 && t == 0) ? 1.0 : 0.0, 0.0);
                } else {
                    for (u32 i = tid; i < L; i += bd) st[slot_of(i)] = gs[tile_index(tb, hi_tab, i)];
                }
                __syncthreads();
                run_passes<K, block_threads<K>(), 1>(st, lq, passes, sg.pass0, sg.npass, subs, pool);
                if (tl > 0) {
                    for (u32 i = tid; i < L; i += bd) gs[tile_index(tb, hi_tab, i)] = st[slot_of(i)];
                } else if (keep) {
                    for (u32 i = tid; i < L; i += bd) gs[i] = st[slot_of(i)];
                }
                __syncthreads();
            }
        }

        // sum_k c_k Re <psi| P_k |psi>
        const double acc = tl ? energy_sum<false>(gs, st, c, terms) : energy_sum<true>(gs, st, c, terms);
        const double total = block_sum(acc, red);
        if (tid == 0) energies[ci] = total;
        __syncthreads();
    }
}

// One stage of a single-state circuit: every work-group takes tiles of 2^lq
// amplitudes (grid-stride), runs the stage's passes on each in LDS and writes
// it back in place.  Tiles of one stage are disjoint, so work-groups never
// touch the same amplitudes.  The next tile is loaded into registers while the
// passes run on the current one (a barrier does not wait for global loads), so
// the HBM traffic of consecutive tiles overlaps the arithmetic.
//
// Work-group size and register blocks per thread: at K = 3, 512 threads with
// two 8-amplitude blocks each (a 13-qubit tile is 1024 groups), so every
// sub-operation is decoded once per two blocks and the FP64 chains of the two
// interleave; K = 2: 1024 threads with two 4-amplitude blocks; K = 4: 512
// threads with one 16-amplitude block.
template <int K>
constexpr u32 stage_threads() { return K == 2 ? 1024u : 512u; }
template <int K>
constexpr int stage_blocks() { return K <= 3 ? 2 : 1; }

// Tiles of exactly P * T amplitudes, P per thread, with prefetch.  Each
// iteration issues the loads of the next tile, runs the passes on the tile in
// LDS, stores it and only then moves the prefetched registers into LDS, so the
// wait for the loads follows the stores in program order and lets the stores
// stay in flight (the vector memory counter returns in order); with the LDS
// fill at the top of the loop the compiler's wait there has to hold for the
// first entry too and would drain the previous tile's stores.  The loads are
// unconditional (after the last tile every lane reads amplitude 0, one cache
// line, and the values are not used): a conditional load merged with the old
// values makes the compiler copy the loaded registers, i.e. wait, at once.
template <int K, u32 P, u32 T = stage_threads<K>()>
__device__ __forceinline__ void stage_tiles(double2* __restrict__ state, double2* st, const u64* hi_tab, u32 lq, u32 tl,
                                            const DStage& sg, const DPass* __restrict__ passes,
                                            const DSub* __restrict__ subs, const double2* __restrict__ pool) {
    const u32 tid = threadIdx.x;
    const u64 ntiles = 1ULL << tl;
    u64 t = blockIdx.x;
    if (t >= ntiles) return;
    {
        const u64 tb = tile_base(t, tl, sg);
#pragma unroll
        for (u32 k = 0; k < P; ++k) st[slot_of(tid + k * T)] = state[tile_index(tb, hi_tab, tid + k * T)];
        __syncthreads();
    }
    for (;;) {
        double rx[P], ry[P];  // the prefetched tile (separate scalars keep it in registers)
        const u64 tn = t + gridDim.x;
        const bool more = tn < ntiles;
        {
            const u64 tb = tile_base(tn, tl, sg);
#pragma unroll
            for (u32 k = 0; k < P; ++k) {
                const double2 x = state[more ? tile_index(tb, hi_tab, tid + k * T) : 0];
                rx[k] = x.x;
                ry[k] = x.y;
            }
        }
        run_passes<K, T, stage_blocks<K>()>(st, lq, passes, sg.pass0, sg.npass, subs, pool);
        const u64 tb = tile_base(t, tl, sg);
#pragma unroll
        for (u32 k = 0; k < P; ++k) state[tile_index(tb, hi_tab, tid + k * T)] = st[slot_of(tid + k * T)];
        if (!more) break;
        __syncthreads();  // the tile has been read out of LDS
#pragma unroll
        for (u32 k = 0; k < P; ++k) st[slot_of(tid + k * T)] = make_double2(rx[k], ry[k]);
        __syncthreads();
        t = tn;
    }
}

template <int K>
__global__ __launch_bounds__(stage_threads<K>()) void k_stage(double2* __restrict__ state, u32 lq, u32 tl, DStage sg,
                                                         const DPass* __restrict__ passes,
                                                         const DSub* __restrict__ subs,
                                                         const double2* __restrict__ pool) {
    extern __shared__ double2 st[];
    __shared__ u64 hi_tab[1u << (kMaxLocal - kLowFixed)];  // offsets of states above 32 qubits exceed 32 bits
    const u32 L = 1u << lq;
    build_hi_tab(hi_tab, lq, sg);
    __syncthreads();
    constexpr u32 T = stage_threads<K>();
    // the prefetched tile needs 4 P registers per thread on top of the register
    // blocks (at K = 4: 64 beside ~165, within the 256 of a 512-thread group)
    if constexpr (K >= 3) {
        if (L == 16 * T) {
            stage_tiles<K, 16>(state, st, hi_tab, lq, tl, sg, passes, subs, pool);
            return;
        }
    }
    if constexpr (K <= 3) {
        if (L == 8 * T) {
            stage_tiles<K, 8>(state, st, hi_tab, lq, tl, sg, passes, subs, pool);
            return;
        }
        if (L == 4 * T) {
            stage_tiles<K, 4>(state, st, hi_tab, lq, tl, sg, passes, subs, pool);
            return;
        }
    }
    const u64 ntiles = 1ULL << tl;
    for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 tb = tile_base(t, tl, sg);
        for (u32 i = threadIdx.x; i < L; i += T) st[slot_of(i)] = state[tile_index(tb, hi_tab, i)];
        __syncthreads();
        run_passes<K, T, stage_blocks<K>()>(st, lq, passes, sg.pass0, sg.npass, subs, pool);
        for (u32 i = threadIdx.x; i < L; i += T) state[tile_index(tb, hi_tab, i)] = st[slot_of(i)];
        __syncthreads();
    }
}

// ------------------------------------------------------------------ host side
void mat_mul(const std::vector<cd>& a, const std::vector<cd>& b, std::vector<cd>& out, int d) {
    out.assign((size_t)d * d, 0.0);
    for (int i = 0; i < d; ++i)
        for (int k = 0; k < d; ++k)
            for (int j = 0; j < d; ++j) out[i * d + j] += a[i * d + k] * b[k * d + j];
}

template <class T>
bool dev_upload(T** dst, const std::vector<T>& src) {
    *dst = nullptr;
    if (src.empty()) return true;
    if (!hip_ok(hipMalloc((void**)dst, src.size() * sizeof(T)), "qm_batch_compile: hipMalloc")) return false;
    return hip_ok(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice),
                  "qm_batch_compile: hipMemcpy");
}

// Grows *dst to hold `bytes` (contents not kept); capacity in *cap.
bool dev_reserve(void** dst, size_t* cap, size_t bytes, const char* what) {
    if (*cap >= bytes && *dst) return true;
    if (*dst) {
        if (!hip_ok(hipFree(*dst), what)) return false;
        *dst = nullptr;
        *cap = 0;
    }
    if (!hip_ok(hipMalloc(dst, std::max<size_t>(bytes, 256)), what)) return false;
    *cap = std::max<size_t>(bytes, 256);
    return true;
}

template <class T>
bool dev_put(T** dst, size_t* cap, const std::vector<T>& src, const char* what) {
    if (!dev_reserve((void**)dst, cap, src.size() * sizeof(T), what)) return false;
    return src.empty() || hip_ok(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), what);
}

}  // namespace
}  // namespace qm

struct qm_batch {
    std::vector<qm::HCirc> circs;
    // compiled
    int device = -1;
    unsigned K = 0;
    int keep = 0;
    unsigned lq_max = 0;
    std::vector<qm::DCirc> h_circs;
    qm::u64 n_stages = 0, n_passes = 0, n_subs = 0;
    qm::DCirc* d_circs = nullptr;
    qm::DStage* d_stages = nullptr;
    qm::DPass* d_passes = nullptr;
    qm::DSub* d_subs = nullptr;
    double2* d_pool = nullptr;
    qm::DTerm* d_terms = nullptr;
    qm::u32* d_order = nullptr;
    double2* d_scratch = nullptr;
    double* d_energies = nullptr;
    int* d_counter = nullptr;
    size_t cap[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // bytes held by each device buffer above
    int grid = 0;                       // work-groups of the launch (computed once per compile)
    unsigned lq_cap = 13;               // qubits held in LDS per tile (qm_batch_set_local_qubits)
    hipGraph_t graph = nullptr;         // captured launch (qm_batch_capture)
    hipGraphExec_t graph_exec = nullptr;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }

    void release() {
        drop_graph();
        grid = 0;
        if (device >= 0) {
            (void)hipSetDevice(device);
            void* ptrs[] = {d_circs, d_stages, d_passes, d_subs, d_pool, d_terms, d_order, d_scratch, d_energies, d_counter};
            for (void* p : ptrs)
                if (p) (void)hipFree(p);
        }
        d_circs = nullptr; d_stages = nullptr; d_passes = nullptr; d_subs = nullptr; d_pool = nullptr;
        d_terms = nullptr; d_order = nullptr; d_scratch = nullptr; d_energies = nullptr; d_counter = nullptr;
        for (size_t& c : cap) c = 0;
        device = -1;
        h_circs.clear();
    }
};

struct qm_circuit {
    unsigned n = 0, K = 4, lq = 13;
    std::vector<qm::DStage> stages;
    std::vector<qm::DPass> passes;
    std::vector<qm::DSub> subs;
    std::vector<double2> pool;
    int device = -1;
    int per_cu = 0, cus = 0;
    qm::DPass* d_passes = nullptr;
    qm::DSub* d_subs = nullptr;
    double2* d_pool = nullptr;
    void release() {
        if (device >= 0) {
            (void)hipSetDevice(device);
            void* ptrs[] = {d_passes, d_subs, d_pool};
            for (void* p : ptrs)
                if (p) (void)hipFree(p);
        }
        d_passes = nullptr; d_subs = nullptr; d_pool = nullptr;
        device = -1;
        per_cu = 0;
    }
};

namespace qm {
namespace {

// Schedules one circuit; appends to the flat arrays.
struct Flat {
    std::vector<DStage> stages;
    std::vector<DPass> passes;
    std::vector<DSub> subs;
    std::vector<double2> pool;
};

u32 pool_put(Flat& f, const std::vector<cd>& m) {
    const u32 off = (u32)f.pool.size();
    for (const cd& z : m) f.pool.push_back(make_double2(z.real(), z.imag()));
    return off;
}

// Passes of one stage: ops (already in local positions) -> DPass/DSub.
void schedule_passes(Flat& f, std::vector<HOp> ops, u32 lq, u32 K, DStage& sg) {
    sg.pass0 = (u32)f.passes.size();
    sg.npass = 0;
    std::vector<bool> done(ops.size(), false);
    size_t left = ops.size();
    while (left > 0) {
        u64 blocked = 0, block = 0;
        std::vector<size_t> sel;
        for (size_t i = 0; i < ops.size(); ++i) {
            if (done[i]) continue;
            const HOp& o = ops[i];
            const u64 m = o.mask();
            if (m & blocked) {
                blocked |= m;
                continue;
            }
            if (o.type != 0) {
                sel.push_back(i);
                continue;
            }
            if ((u32)__builtin_popcountll(block | m) <= K) {
                block |= m;
                sel.push_back(i);
            } else {
                blocked |= m;
            }
        }
        // pad the block to K local positions
        for (u32 p = 0; p < lq && (u32)__builtin_popcountll(block) < K; ++p) block |= 1ULL << p;
        DPass ps{};
        u32 bpos[4] = {0, 0, 0, 0};
        {
            u32 k = 0;
            for (u32 p = 0; p < lq; ++p)
                if ((block >> p) & 1) bpos[k++] = p;
        }
        for (u32 i = 0; i < 4; ++i) ps.pos[i] = (unsigned char)bpos[i];
        for (u32 i = 0; i < K; ++i) ps.sbit[i] = (unsigned short)slot_of(1u << bpos[i]);
        ps.sub0 = (u32)f.subs.size();
        auto bit_of = [&](u32 p) -> u32 {
            for (u32 i = 0; i < K; ++i)
                if (bpos[i] == p) return i;
            return 0;
        };
        // Consecutive CZ gates, then consecutive single-qubit matrices on distinct
        // block bits (these commute with each other), become one LAYER.
        std::vector<const std::vector<cd>*> pend_d1(K, nullptr);
        u32 d1mask = 0, cz_const = 0, npair = 0, pairs = 0;
        u32 czmask[4] = {0, 0, 0, 0};
        bool cz_any = false;
        auto flush = [&]() {
            if (d1mask == 0 && !cz_any) return;
            DSub sb{};
            const bool masks = (czmask[0] | czmask[1] | czmask[2] | czmask[3]) != 0;
            sb.type = SUB_LAYER | (d1mask << 8) | (npair << 12) | (masks ? 0x8000u : 0u) | (cz_const << 16);
            sb.a = czmask[0] | (czmask[1] << 16);
            sb.m = czmask[2] | (czmask[3] << 16);
            sb.pairs = pairs;
            sb.data = (u32)f.pool.size();
            for (u32 b = 0; b < K; ++b)
                if (pend_d1[b]) pool_put(f, *pend_d1[b]);
            f.subs.push_back(sb);
            std::fill(pend_d1.begin(), pend_d1.end(), nullptr);
            d1mask = 0;
            cz_const = 0;
            npair = 0;
            pairs = 0;
            std::fill(czmask, czmask + 4, 0u);
            cz_any = false;
        };
        const u32 jmask = (1u << (1u << K)) - 1;
        auto add_cz = [&](u32 p0, u32 p1) {
            const bool in0 = (block >> p0) & 1, in1 = (block >> p1) & 1;
            if (in0 && in1) {
                cz_const ^= (u32)(kBlockPatterns >> (16 * bit_of(p0))) & (u32)(kBlockPatterns >> (16 * bit_of(p1))) &
                            jmask;
            } else if (in0 || in1) {
                czmask[bit_of(in0 ? p0 : p1)] ^= 1u << (in0 ? p1 : p0);
            } else {
                if (npair == kCzPairsMax) flush();
                pairs |= (p0 | (p1 << 4)) << (8 * npair);
                ++npair;
            }
            cz_any = true;
        };
        for (size_t i : sel) {
            const HOp& o = ops[i];
            done[i] = true;
            --left;
            if (o.type == 2) {
                if (d1mask) flush();
                add_cz(o.q[0], o.q[1]);
                continue;
            }
            if (o.type == 0 && o.q.size() == 1) {
                const u32 b = bit_of(o.q[0]);
                if ((d1mask >> b) & 1) flush();
                pend_d1[b] = &o.m;
                d1mask |= 1u << b;
                continue;
            }
            flush();
            DSub sb{};
            if (o.type == 1) {
                ps.flags |= kPassHasDiag;
                sb.type = SUB_DIAG;
                sb.m = (u32)o.q.size();
                for (size_t k = 0; k < o.q.size(); ++k) sb.a |= o.q[k] << (8 * k);
                sb.data = pool_put(f, o.m);
            } else if (o.q.size() == 2) {
                sb.type = SUB_D2;
                sb.a = bit_of(o.q[0]) | (bit_of(o.q[1]) << 8);
                sb.data = pool_put(f, o.m);
            } else {
                // embed into the K block bits (ascending) with the identity elsewhere
                const u32 k = (u32)o.q.size(), D = 1u << K;
                std::vector<u32> bits(k);
                u32 omask = 0;
                for (u32 t = 0; t < k; ++t) {
                    bits[t] = bit_of(o.q[t]);
                    omask |= 1u << bits[t];
                }
                std::vector<cd> M((size_t)D * D, 0.0);
                for (u32 col = 0; col < D; ++col) {
                    u32 x = 0;
                    for (u32 t = 0; t < k; ++t) x |= ((col >> bits[t]) & 1u) << t;
                    for (u32 y = 0; y < (1u << k); ++y) {
                        u32 row = col & ~omask;
                        for (u32 t = 0; t < k; ++t) row |= ((y >> t) & 1u) << bits[t];
                        M[(size_t)row * D + col] += o.m[(size_t)y * (1u << k) + x];
                    }
                }
                sb.type = SUB_DK;
                sb.data = pool_put(f, M);
            }
            f.subs.push_back(sb);
        }
        flush();
        ps.nsub = (u32)f.subs.size() - ps.sub0;
        f.passes.push_back(ps);
        ++sg.npass;
    }
}

// Stages (windows of lq_cap qubits including 0..3) for circuits above lq_cap
// qubits; a single identity stage otherwise.  Returns the number of stages appended.
u32 schedule_circuit(Flat& f, const HCirc& c, u32 K, u32 lq_cap = kMaxLocal) {
    const u32 n = c.ne, lq = std::min(n, lq_cap);
    if (n <= lq_cap) {
        DStage sg{};
        for (u32 i = 0; i + kLowFixed < lq; ++i) set_byte(sg.win[i >> 3], i & 7, kLowFixed + i);
        schedule_passes(f, c.ops, lq, K, sg);
        f.stages.push_back(sg);
        return 1;
    }
    std::vector<bool> done(c.ops.size(), false);
    size_t left = c.ops.size();
    u32 count = 0;
    const u64 low = (1ULL << kLowFixed) - 1;
    while (left > 0) {
        u64 win = low, blocked = 0;
        std::vector<size_t> sel;
        for (size_t i = 0; i < c.ops.size(); ++i) {
            if (done[i]) continue;
            const u64 m = c.ops[i].mask();
            if (m & blocked) {
                blocked |= m;
                continue;
            }
            if ((u32)__builtin_popcountll(win | m) <= lq) {
                win |= m;
                sel.push_back(i);
            } else {
                blocked |= m;
            }
        }
        for (u32 q = 0; q < n && (u32)__builtin_popcountll(win) < lq; ++q) win |= 1ULL << q;
        DStage sg{};
        u32 local_of[64];
        u32 wi = 0, ci = 0;
        for (u32 q = 0; q < n; ++q) {
            if (q < kLowFixed) {
                local_of[q] = q;
            } else if ((win >> q) & 1) {
                local_of[q] = kLowFixed + wi;
                set_byte(sg.win[wi >> 3], wi & 7, q);
                ++wi;
            } else {
                set_byte(sg.comp[ci >> 3], ci & 7, q);
            ++ci;
            }
        }
        std::vector<HOp> ops;
        ops.reserve(sel.size());
        for (size_t i : sel) {
            HOp o = c.ops[i];
            for (u32& q : o.q) q = local_of[q];
            ops.push_back(std::move(o));
            done[i] = true;
            --left;
        }
        schedule_passes(f, ops, lq, K, sg);
        f.stages.push_back(sg);
        ++count;
    }
    if (count == 0) {  // a circuit without gates still needs one stage to initialise |0>
        DStage sg{};
        u32 wi = 0, ci = 0;
        for (u32 q = kLowFixed; q < n; ++q) {
            if (wi + kLowFixed < lq) {
                set_byte(sg.win[wi >> 3], wi & 7, q);
                ++wi;
            } else {
                set_byte(sg.comp[ci >> 3], ci & 7, q);
            ++ci;
            }
        }
        sg.npass = 0;
        sg.pass0 = (u32)f.passes.size();
        f.stages.push_back(sg);
        count = 1;
    }
    return count;
}

// Gates in the C ABI form -> internal ops, with consecutive single-qubit dense /
// diagonal gates on the same qubit multiplied together.
// Appends o; a single-qubit dense / diagonal gate is fused into the previous
// op on its qubit when that is one too (last[q]: index of the last op on q).
void push_op(HCirc& c, std::vector<long>& last, HOp o) {
    if (o.q.size() == 1 && o.type != 2) {
        const long p = last[o.q[0]];
        if (p >= 0 && c.ops[p].q.size() == 1 && c.ops[p].type != 2) {
            HOp& prev = c.ops[p];
            auto full = [](const HOp& h) {
                if (h.type == 0) return h.m;
                return std::vector<cd>{h.m[0], 0.0, 0.0, h.m[1]};
            };
            if (prev.type == 1 && o.type == 1) {
                prev.m[0] *= o.m[0];
                prev.m[1] *= o.m[1];
            } else {
                std::vector<cd> prod;
                mat_mul(full(o), full(prev), prod, 2);  // new after old
                prev.m = prod;
                prev.type = 0;
            }
            return;
        }
    }
    c.ops.push_back(std::move(o));
    for (u32 q : c.ops.back().q) last[q] = (long)c.ops.size() - 1;
}

bool convert_gates(u32 n, const qm_batch_gate* gates, UINT gate_count, HCirc& c, const char* what) {
    c.n = n;
    c.ne = std::max<u32>(n, kMinQubits);
    std::vector<long> last(c.ne, -1);  // last op touching each qubit
    for (UINT gi = 0; gi < gate_count; ++gi) {
        const qm_batch_gate& g = gates[gi];
        const u32 k = g.qubit_count;
        const bool two = g.kind == QM_BATCH_CZ || g.kind == QM_BATCH_CNOT;
        if (g.kind > QM_BATCH_CNOT || k < 1 || k > 4 || (two && k != 2)) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: gate %u: kind %u with %u qubits is not supported", what, gi, g.kind, k);
            return false;
        }
        u64 mask = 0;
        for (u32 i = 0; i < k; ++i) {
            if (g.qubits[i] >= n || ((mask >> g.qubits[i]) & 1)) {
                set_error(QM_ERR_INVALID_ARGUMENT, "%s: gate %u: qubit %u out of range or repeated", what, gi, g.qubits[i]);
                return false;
            }
            mask |= 1ULL << g.qubits[i];
        }
        if ((g.kind == QM_BATCH_DENSE || g.kind == QM_BATCH_DIAG) && !g.matrix) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: gate %u: null matrix", what, gi);
            return false;
        }
        if (g.kind == QM_BATCH_CNOT) {
            // CNOT(c, t) = H_t CZ(c, t) H_t; the Hadamards fuse into the
            // single-qubit gates next to them on t
            const double h = 0.70710678118654752440;
            HOp hd;
            hd.type = 0;
            hd.q = {g.qubits[1]};
            hd.m = {h, h, h, -h};
            HOp cz;
            cz.type = 2;
            cz.q = {g.qubits[0], g.qubits[1]};
            push_op(c, last, hd);
            push_op(c, last, std::move(cz));
            push_op(c, last, std::move(hd));
            continue;
        }
        HOp o;
        o.q.assign(g.qubits, g.qubits + k);
        const u32 d = 1u << k;
        if (g.kind == QM_BATCH_CZ) {
            o.type = 2;
        } else if (g.kind == QM_BATCH_DIAG) {
            o.type = 1;
            for (u32 i = 0; i < d; ++i) o.m.push_back(cd(g.matrix[i].re, g.matrix[i].im));
        } else {
            o.type = 0;
            for (u32 i = 0; i < d * d; ++i) o.m.push_back(cd(g.matrix[i].re, g.matrix[i].im));
        }
        push_op(c, last, std::move(o));
    }
    return true;
}

// Sets the kernel's LDS limit and sizes the grid (resident work-groups, at most
// one per circuit).  Kept out of the launch path so a graph capture only sees
// stream operations.
bool batch_prepare_launch(qm_batch* b, unsigned dev, const char* what) {
    const size_t lds = sizeof(double2) << b->lq_max;
    const void* fn = b->K == 2 ? (const void*)k_batch<2> : b->K == 3 ? (const void*)k_batch<3> : (const void*)k_batch<4>;
    if (!hip_ok(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), what)) return false;
    int per_cu = 0, cus = 0;
    const int threads = (int)(b->K == 2 ? block_threads<2>() : b->K == 3 ? block_threads<3>() : block_threads<4>());
    if (!hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds), what)) return false;
    if (!hip_ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, (int)dev), what)) return false;
    if (per_cu < 1) {
        set_error(QM_ERR_HIP, "%s: the kernel does not fit a compute unit with %zu bytes of LDS", what, lds);
        return false;
    }
    b->grid = (int)std::min<u64>(b->h_circs.size(), (u64)per_cu * cus);
    return true;
}

bool batch_ok(qm_batch* b, const char* what) {
    if (!b) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null batch", what);
        return false;
    }
    return true;
}

}  // namespace
}  // namespace qm

using namespace qm;

extern "C" {

qm_batch* qm_batch_create(void) { return new qm_batch(); }

void qm_batch_destroy(qm_batch* b) {
    if (!b) return;
    b->release();
    delete b;
}

int qm_batch_add_circuit(qm_batch* b, UINT n, const qm_batch_gate* gates, UINT gate_count, const ITYPE* bf,
                         const ITYPE* pf, const UINT* rot, const double* coefs, UINT term_count) {
    const char* what = "qm_batch_add_circuit";
    if (!batch_ok(b, what)) return -1;
    if (n < 1 || n > QM_BATCH_MAX_QUBITS) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: qubit count %u outside 1..%u", what, n, QM_BATCH_MAX_QUBITS);
        return -1;
    }
    if ((gate_count && !gates) || (term_count && (!bf || !pf || !rot || !coefs))) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null gate or term array", what);
        return -1;
    }
    HCirc c;
    if (!convert_gates(n, gates, gate_count, c, what)) return -1;
    for (UINT t = 0; t < term_count; ++t) {
        if ((bf[t] >> n) || (pf[t] >> n)) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: term %u acts beyond %u qubits", what, t, n);
            return -1;
        }
        DTerm tm{};
        tm.bf = bf[t];
        tm.pf = pf[t];
        tm.rot = rot[t];
        tm.coef = coefs[t];
        c.terms.push_back(tm);
    }
    b->circs.push_back(std::move(c));
    return (int)b->circs.size() - 1;
}

int qm_batch_compile(qm_batch* b, UINT block_qubits, int keep_states, unsigned dev) {
    const char* what = "qm_batch_compile";
    if (!batch_ok(b, what)) return QM_ERR_INVALID_ARGUMENT;
    const u32 K = block_qubits ? block_qubits : 4;
    if (K < 2 || K > 4) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: block_qubits %u outside 2..4", what, K);
        return QM_ERR_INVALID_ARGUMENT;
    }
    for (const HCirc& c : b->circs)
        for (const HOp& o : c.ops)
            if (o.type == 0 && o.q.size() > K) {
                set_error(QM_ERR_INVALID_ARGUMENT, "%s: a %zu-qubit dense gate needs block_qubits >= %zu", what,
                          o.q.size(), o.q.size());
                return QM_ERR_INVALID_ARGUMENT;
            }
    if (b->device >= 0 && b->device != (int)dev) b->release();
    b->drop_graph();  // device buffers and the queue length may change
    b->grid = 0;
    b->h_circs.clear();
    if (!use_device(dev)) return QM_ERR_HIP;
    Flat f;    std::vector<DTerm> terms;
    std::vector<DCirc> dc;
    u64 state_off = 0;
    u32 lq_max = kMinQubits;
    for (const HCirc& c : b->circs) {
        DCirc d{};
        d.n = c.ne;
        d.lq = std::min(c.ne, (u32)b->lq_cap);
        d.stage0 = (u32)f.stages.size();
        d.nstage = schedule_circuit(f, c, K, b->lq_cap);
        d.term0 = (u32)terms.size();
        d.nterm = (u32)c.terms.size();
        terms.insert(terms.end(), c.terms.begin(), c.terms.end());
        d.state_off = state_off;
        if (c.ne > b->lq_cap || keep_states) state_off += 1ULL << c.ne;
        lq_max = std::max(lq_max, d.lq);
        dc.push_back(d);
    }
    // descending cost: amplitudes x (passes + terms)
    std::vector<u32> order(dc.size());
    std::vector<double> cost(dc.size());
    for (size_t i = 0; i < dc.size(); ++i) {
        u64 np = 0;
        for (u32 s = 0; s < dc[i].nstage; ++s) np += f.stages[dc[i].stage0 + s].npass;
        cost[i] = std::ldexp(1.0, (int)dc[i].n) * (double)(np + dc[i].nterm + 1);
        order[i] = (u32)i;
    }
    std::stable_sort(order.begin(), order.end(), [&](u32 x, u32 y) { return cost[x] > cost[y]; });
    if (f.pool.empty()) f.pool.push_back(make_double2(0, 0));
    if (f.subs.empty()) f.subs.push_back(DSub{});
    if (f.passes.empty()) f.passes.push_back(DPass{});
    if (terms.empty()) terms.push_back(DTerm{});
    b->device = (int)dev;
    bool ok = dev_put(&b->d_circs, &b->cap[0], dc, what) && dev_put(&b->d_stages, &b->cap[1], f.stages, what) &&
              dev_put(&b->d_passes, &b->cap[2], f.passes, what) && dev_put(&b->d_subs, &b->cap[3], f.subs, what) &&
              dev_put(&b->d_pool, &b->cap[4], f.pool, what) && dev_put(&b->d_terms, &b->cap[5], terms, what) &&
              dev_put(&b->d_order, &b->cap[6], order, what);
    if (ok && state_off) {
        ok = dev_reserve((void**)&b->d_scratch, &b->cap[7], state_off * sizeof(double2), what);
        if (!ok) set_error(QM_ERR_OUT_OF_MEMORY, "%s: %llu bytes of state memory", what, state_off * 16ULL);
    }
    if (ok) ok = dev_reserve((void**)&b->d_energies, &b->cap[8], std::max<size_t>(1, dc.size()) * sizeof(double), what);
    if (ok) ok = dev_reserve((void**)&b->d_counter, &b->cap[9], sizeof(int), what);
    if (!ok) {
        b->release();
        return QM_ERR_HIP;
    }
    b->K = K;
    b->keep = keep_states ? 1 : 0;
    b->lq_max = lq_max;
    b->h_circs = dc;
    b->n_stages = f.stages.size();
    b->n_passes = 0;
    for (const DStage& s : f.stages) b->n_passes += s.npass;
    b->n_subs = 0;
    for (const DPass& p : f.passes) b->n_subs += p.nsub;
    return QM_OK;
}

const double* qm_batch_run_async(qm_batch* b, void* stream, unsigned dev) {
    const char* what = "qm_batch_run";
    if (!batch_ok(b, what)) return nullptr;
    if (b->device < 0 || b->device != (int)dev) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: batch not compiled for device %u", what, dev);
        return nullptr;
    }
    if (!use_device(dev)) return nullptr;
    hipStream_t st = stream_of(stream);
    const u32 ncirc = (u32)b->h_circs.size();
    if (ncirc == 0) return b->d_energies;
    const size_t lds = sizeof(double2) << b->lq_max;
    if (b->grid == 0 && !batch_prepare_launch(b, dev, what)) return nullptr;
    const u32 grid = (u32)b->grid;
#define QM_BATCH_ARGS                                                                                       \
    b->d_circs, b->d_stages, b->d_passes, b->d_subs, b->d_pool, b->d_terms, b->d_order, b->d_scratch, b->d_energies, \
        b->d_counter, ncirc, b->keep
    k_batch_reset<<<1, 64, 0, st>>>(b->d_counter);
    switch (b->K) {
        case 2: k_batch<2><<<grid, block_threads<2>(), lds, st>>>(QM_BATCH_ARGS); break;
        case 3: k_batch<3><<<grid, block_threads<3>(), lds, st>>>(QM_BATCH_ARGS); break;
        default: k_batch<4><<<grid, block_threads<4>(), lds, st>>>(QM_BATCH_ARGS); break;
    }
#undef QM_BATCH_ARGS
    if (!hip_ok(hipGetLastError(), what)) return nullptr;
    return b->d_energies;
}

int qm_batch_run(qm_batch* b, double* energies, void* stream, unsigned dev) {
    const double* d = qm_batch_run_async(b, stream, dev);
    if (!d) return QM_ERR_HIP;
    hipStream_t st = stream_of(stream);
    const size_t n = b->h_circs.size();
    if (n && energies) {
        if (!hip_ok(hipMemcpyAsync(energies, d, n * sizeof(double), hipMemcpyDeviceToHost, st), "qm_batch_run"))
            return QM_ERR_HIP;
    }
    if (!hip_ok(hipStreamSynchronize(st), "qm_batch_run")) return QM_ERR_HIP;
    return QM_OK;
}

int qm_batch_get_state(qm_batch* b, UINT index, qm_complex* out, void* stream, unsigned dev) {
    const char* what = "qm_batch_get_state";
    if (!batch_ok(b, what)) return QM_ERR_INVALID_ARGUMENT;
    if (!b->keep || index >= b->h_circs.size() || !out || b->device != (int)dev) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: no kept state %u (compile with keep_states, then run)", what, index);
        return QM_ERR_INVALID_ARGUMENT;
    }
    if (!use_device(dev)) return QM_ERR_HIP;
    hipStream_t st = stream_of(stream);
    const DCirc& c = b->h_circs[index];
    const u32 n = b->circs[index].n;  // the circuit's own size: extra qubits stay |0>, so the state is the prefix
    if (!hip_ok(hipMemcpyAsync(out, b->d_scratch + c.state_off, sizeof(double2) << n, hipMemcpyDeviceToHost, st), what))
        return QM_ERR_HIP;
    if (!hip_ok(hipStreamSynchronize(st), what)) return QM_ERR_HIP;
    return QM_OK;
}

int qm_batch_capture(qm_batch* b, void* stream, unsigned dev) {
    const char* what = "qm_batch_capture";
    if (!batch_ok(b, what)) return QM_ERR_INVALID_ARGUMENT;
    if (!stream || b->device != (int)dev || b->h_circs.empty()) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: needs a compiled, non-empty batch and a created stream", what);
        return QM_ERR_INVALID_ARGUMENT;
    }
    if (!use_device(dev)) return QM_ERR_HIP;
    if (b->grid == 0 && !batch_prepare_launch(b, dev, what)) return QM_ERR_HIP;
    b->drop_graph();
    hipStream_t st = stream_of(stream);
    if (!hip_ok(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), what)) return QM_ERR_HIP;
    const double* d = qm_batch_run_async(b, stream, dev);
    hipGraph_t g = nullptr;
    const bool ended = hip_ok(hipStreamEndCapture(st, &g), what);
    if (!d || !ended || !g) {
        if (g) (void)hipGraphDestroy(g);
        return QM_ERR_HIP;
    }
    b->graph = g;
    if (!hip_ok(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0), what)) {
        b->drop_graph();
        return QM_ERR_HIP;
    }
    return QM_OK;
}

int qm_batch_run_graph(qm_batch* b, double* energies, void* stream, unsigned dev) {
    const char* what = "qm_batch_run_graph";
    if (!batch_ok(b, what)) return QM_ERR_INVALID_ARGUMENT;
    if (!b->graph_exec || b->device != (int)dev) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: capture the batch first (qm_batch_capture)", what);
        return QM_ERR_INVALID_ARGUMENT;
    }
    if (!use_device(dev)) return QM_ERR_HIP;
    hipStream_t st = stream_of(stream);
    if (!hip_ok(hipGraphLaunch(b->graph_exec, st), what)) return QM_ERR_HIP;
    const size_t n = b->h_circs.size();
    if (energies && n &&
        !hip_ok(hipMemcpyAsync(energies, b->d_energies, n * sizeof(double), hipMemcpyDeviceToHost, st), what))
        return QM_ERR_HIP;
    if (!hip_ok(hipStreamSynchronize(st), what)) return QM_ERR_HIP;
    return QM_OK;
}

int qm_batch_set_local_qubits(qm_batch* b, UINT local_qubits) {
    if (!batch_ok(b, "qm_batch_set_local_qubits")) return QM_ERR_INVALID_ARGUMENT;
    if (local_qubits < 8 || local_qubits > kMaxLocal) {
        set_error(QM_ERR_INVALID_ARGUMENT, "qm_batch_set_local_qubits: %u outside 8..%u", local_qubits, kMaxLocal);
        return QM_ERR_INVALID_ARGUMENT;
    }
    b->lq_cap = local_qubits;
    return QM_OK;
}

void qm_batch_clear(qm_batch* b) {
    if (!b) return;
    b->circs.clear();
    b->h_circs.clear();
}

void* qm_batch_state_device_ptr(qm_batch* b, UINT index) {
    if (!b || !b->keep || index >= b->h_circs.size()) {
        set_error(QM_ERR_INVALID_ARGUMENT, "qm_batch_state_device_ptr: no kept state %u", index);
        return nullptr;
    }
    return b->d_scratch + b->h_circs[index].state_off;
}

qm_circuit* qm_circuit_create(UINT n, const qm_batch_gate* gates, UINT gate_count, UINT block_qubits,
                              UINT local_qubits) {
    const char* what = "qm_circuit_create";
    // Defaults (0): up to 19 qubits a 13-qubit window leaves at most 64 tiles
    // per stage for 256 CUs; 12-qubit windows with 3-qubit blocks were 11-28 %
    // faster there on brickwork circuits (13 q: 0.198 -> 0.149 ms, 14 q: 0.130
    // -> 0.093 ms, 18 q: 0.179 -> 0.133 ms; 10-12 q, one stage: 15-20 % from the
    // block width alone), equal at 20 and slower from 21 qubits up.
    const bool small_state = n <= 19;
    u32 K = block_qubits ? block_qubits : (small_state ? 3 : 4);
    const u32 lq = local_qubits ? local_qubits : (small_state ? 12 : kMaxLocal);
    if (n < kMinQubits || n > QM_CIRCUIT_MAX_QUBITS || K < 2 || K > 4 || lq < 8 || lq > kMaxLocal ||
        (gate_count && !gates)) {
        set_error(QM_ERR_INVALID_ARGUMENT,
                  "%s: need %u <= qubits (%u) <= %u, block_qubits (%u) in 2..4, local_qubits (%u) in 8..%u and a gate array",
                  what, kMinQubits, n, QM_CIRCUIT_MAX_QUBITS, K, lq, kMaxLocal);
        return nullptr;
    }
    auto* c = new qm_circuit();
    HCirc hc;
    if (!convert_gates(n, gates, gate_count, hc, what)) {
        delete c;
        return nullptr;
    }
    if (!block_qubits)  // the default width widens to the largest dense gate (at most 4 qubits)
        for (const HOp& o : hc.ops)
            if (o.type == 0 && o.q.size() > K && o.q.size() <= 4) K = (u32)o.q.size();
    for (const HOp& o : hc.ops)
        if (o.type == 0 && o.q.size() > K) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: a %zu-qubit dense gate needs block_qubits >= %zu", what, o.q.size(),
                      o.q.size());
            delete c;
            return nullptr;
        }
    c->n = n;
    c->K = K;
    c->lq = std::min(n, lq);
    Flat f;
    schedule_circuit(f, hc, K, lq);
    c->stages = f.stages;
    c->passes = f.passes;
    c->subs = f.subs;
    c->pool = f.pool;
    if (c->pool.empty()) c->pool.push_back(make_double2(0, 0));
    if (c->subs.empty()) c->subs.push_back(DSub{});
    if (c->passes.empty()) c->passes.push_back(DPass{});
    return c;
}

void qm_circuit_destroy(qm_circuit* c) {
    if (!c) return;
    c->release();
    delete c;
}

int qm_circuit_apply(qm_circuit* c, void* state, ITYPE dim, void* stream, unsigned dev) {
    const char* what = "qm_circuit_apply";
    if (!c || !state || dim != (1ULL << c->n)) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null circuit/state or dim is not 2^%u", what, c ? c->n : 0u);
        return QM_ERR_INVALID_ARGUMENT;
    }
    if (!use_device(dev)) return QM_ERR_HIP;
    hipStream_t st = stream_of(stream);
    if (!qm::circuit_enqueue(c, state, dim, st, dev)) return QM_ERR_HIP;
    finish_update(st, what);
    return QM_OK;
}

}  // extern "C"

namespace qm {
bool circuit_prepare(qm_circuit* c, unsigned dev) {
    if (c->device != (int)dev) {
        c->release();
        bool ok = dev_upload(&c->d_passes, c->passes) && dev_upload(&c->d_subs, c->subs) &&
                  dev_upload(&c->d_pool, c->pool);
        if (!ok) {
            c->release();
            return false;
        }
        c->device = (int)dev;
    }
    return true;
}

bool circuit_enqueue(qm_circuit* c, void* state, u64 dim, hipStream_t st, unsigned dev) {
    const char* what = "qm_circuit_apply";
    if (!c || !state || dim != (1ULL << c->n)) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null circuit/state or dim is not 2^%u", what, c ? c->n : 0u);
        return false;
    }
    if (!circuit_prepare(c, dev)) return false;
    const size_t lds = sizeof(double2) << c->lq;
    const void* fn = c->K == 2 ? (const void*)k_stage<2> : c->K == 3 ? (const void*)k_stage<3> : (const void*)k_stage<4>;
    if (c->per_cu == 0) {
        int per_cu = 0, cus = 0;
        if (!hip_ok(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), what) ||
            !hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)(c->K == 2 ? stage_threads<2>() : c->K == 3 ? stage_threads<3>() : stage_threads<4>()), lds), what) ||
            !hip_ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, (int)dev), what))
            return false;
        if (per_cu < 1) {
            set_error(QM_ERR_HIP, "%s: the kernel does not fit a compute unit with %zu bytes of LDS", what, lds);
            return false;
        }
        c->per_cu = per_cu;
        c->cus = cus;
    }
    const u32 tl = c->n - c->lq;
    const u64 ntiles = 1ULL << tl;
    const u32 grid = (u32)std::min<u64>(ntiles, (u64)c->per_cu * c->cus);
    for (const DStage& sg : c->stages) {
        if (sg.npass == 0) continue;
        switch (c->K) {
            case 2: k_stage<2><<<grid, stage_threads<2>(), lds, st>>>((double2*)state, c->lq, tl, sg, c->d_passes, c->d_subs, c->d_pool); break;
            case 3: k_stage<3><<<grid, stage_threads<3>(), lds, st>>>((double2*)state, c->lq, tl, sg, c->d_passes, c->d_subs, c->d_pool); break;
            default: k_stage<4><<<grid, stage_threads<4>(), lds, st>>>((double2*)state, c->lq, tl, sg, c->d_passes, c->d_subs, c->d_pool); break;
        }
    }
    return hip_ok(hipGetLastError(), what);
}
}  // namespace qm

extern "C" {

void qm_circuit_stats(const qm_circuit* c, ITYPE* stages, ITYPE* passes, ITYPE* subops) {
    ITYPE np = 0, ns = 0;
    if (c)
        for (const DStage& sg : c->stages) {
            np += sg.npass;
            for (u32 p = 0; p < sg.npass; ++p) ns += c->passes[sg.pass0 + p].nsub;
        }
    if (stages) *stages = c ? c->stages.size() : 0;
    if (passes) *passes = np;
    if (subops) *subops = ns;
}

void qm_batch_stats(const qm_batch* b, ITYPE* circuits, ITYPE* stages, ITYPE* passes, ITYPE* subops) {
    if (circuits) *circuits = b ? b->circs.size() : 0;
    if (stages) *stages = b ? b->n_stages : 0;
    if (passes) *passes = b ? b->n_passes : 0;
    if (subops) *subops = b ? b->n_subs : 0;
}

}  // extern "C"
