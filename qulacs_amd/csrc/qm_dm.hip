// Density-matrix kernels.
//
// rho is row-major rho[y * 2^n + x] (src/csim/update_ops_dm.cpp:73-76), so the
// flat index has 2n bits: bit q is column qubit q and bit q + n is row qubit q.
// With that view:
//   * U rho U^dagger is U applied to the row bits followed by conj(U) applied
//     to the column bits -- two passes of the state-vector gather kernels over
//     the 2n-bit flat array, each as fast as a state-vector gate of 2n qubits;
//     a unitary on at most two qubits (controls included) is one pass with
//     E = U (x) conj(U) on the column and row bits of its qubits (for one qubit
//     the reference's ext_matrix, update_ops_dm.cpp:41-50; the reference applies
//     larger ones block by block, update_ops_dm.cpp:342-498), a monomial E
//     (CNOT, CZ, SWAP) through a pass that moves only the entries that change;
//   * a k-qubit CPTP map sum_i K_i rho K_i^dagger is one pass of a 2k-bit dense
//     gate whose matrix is the superoperator sum_i K_i (x) conj(K_i), replacing
//     the copy / gate / scale / add loop of QuantumGate_Probabilistic on a
//     density matrix (src/cppsim/gate_general.hpp:265-285, about ten passes).
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "qm_internal.hpp"

namespace qm {

using cd = std::complex<double>;

static bool dm_ok(void* state, u64 dim, int* n, const char* what) {
    *n = log2_dim(dim);
    if (!state || *n < 0 || *n > 31) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null state or dim %llu is not a power of two in [2, 2^31]", what, dim);
        return false;
    }
    return true;
}

// sum over the diagonal rho_ii (real parts) with (i & mask) == value;
// ENT selects -p log p over entries with p > 1e-15 (stat_ops_dm.cpp:28-47).
template <bool ENT>
__global__ __launch_bounds__(256) void k_dm_diag(const double2* __restrict__ r, u64 dim, u64 mask, u64 value,
                                                  double* __restrict__ partials) {
    __shared__ double lds[32];
    double acc = 0.0;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        if ((i & mask) != value) continue;
        const double p = r[i * dim + i].x;
        if (ENT) {
            if (p > 1e-15) acc += -p * log(p);
        } else {
            acc += p;
        }
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// Re sum_i i^rot (-1)^popcount(i & pf) rho[i, i ^ bf]  (stat_ops_dm.cpp:164-191)
__global__ __launch_bounds__(256) void k_dm_pauli(const double2* __restrict__ r, u64 dim, u64 bf, u64 pf,
                                                   unsigned rot, double* __restrict__ partials) {
    __shared__ double lds[32];
    double acc = 0.0;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        const double2 v = r[i * dim + (i ^ bf)];
        double re;
        switch (rot & 3) {
            case 0: re = v.x; break;
            case 1: re = -v.y; break;
            case 2: re = -v.x; break;
            default: re = v.y; break;
        }
        acc += (__popcll(i & pf) & 1) ? -re : re;
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ p, unsigned n, double* __restrict__ out) {
    __shared__ double lds[32];
    double acc = 0.0;
    for (unsigned k = threadIdx.x; k < n; k += blockDim.x) acc += p[k];
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) out[0] = acc;
}

static double finish_scalar(Scratch* sc, unsigned nb, hipStream_t st, const char* what) {
    double* dout = sc->partials + nb;
    k_sum_partials<<<1, 256, 0, st>>>(sc->partials, nb, dout);
    if (!hip_ok(hipMemcpyAsync(sc->host_result, dout, sizeof(double), hipMemcpyDeviceToHost, st), what)) return 0.0;
    if (!hip_ok(hipStreamSynchronize(st), what)) return 0.0;
    return sc->host_result[0];
}

static double dm_diag_sum(bool ent, u64 mask, u64 value, void* state, u64 dim, void* stream, unsigned dev,
                          const char* what) {
    int n;
    if (!dm_ok(state, dim, &n, what) || !use_device(dev)) return 0.0;
    hipStream_t st = stream_of(stream);
    Scratch* sc = scratch_for(st);
    if (!sc) return 0.0;
    const unsigned nb = grid_for(dim, 256 * 4, 1024);
    if (ent)
        k_dm_diag<true><<<nb, 256, 0, st>>>((const double2*)state, dim, mask, value, sc->partials);
    else
        k_dm_diag<false><<<nb, 256, 0, st>>>((const double2*)state, dim, mask, value, sc->partials);
    return finish_scalar(sc, nb, st, what);
}

__global__ __launch_bounds__(256) void k_outer(double2* __restrict__ r, const double2* __restrict__ psi, u64 dim) {
    const u64 total = dim * dim, stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const double2 a = psi[i / dim], b = psi[i % dim];
        // a * conj(b)
        r[i] = make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y));
    }
}

__global__ __launch_bounds__(256) void k_dm_scale(double2* __restrict__ r, u64 total, double f) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) r[i] = cscale(r[i], f);
}

// Row/column split of a (controlled) k-qubit unitary on an n-qubit density matrix.
static void dm_unitary(const u32* cl, const u32* cvl, u32 nc, const u32* tl, u32 nt, const qm_complex* m,
                       void* state, u64 dim, void* stream, unsigned dev, const char* what) {
    int n;
    if (!dm_ok(state, dim, &n, what)) return;
    QM_CHECK_ARG(tl && m && nt >= 1 && nt + nc <= (u32)n, "%s: bad target/control counts", what);
    u32 all[64];
    for (u32 i = 0; i < nt; ++i) all[i] = tl[i];
    for (u32 i = 0; i < nc; ++i) all[nt + i] = cl[i];
    if (!check_qubits(all, nt + nc, n, what)) return;
    for (u32 i = 0; i < nc; ++i) QM_CHECK_ARG(cvl[i] <= 1, "%s: control value %u", what, cvl[i]);
    const u64 flat = dim * dim;
    if (nt == 1 && nc == 0) {
        // one pass: E[(oy,ox),(iy,ix)] = U[oy,iy] conj(U[ox,ix]); member bit0 = column t, bit1 = row t+n
        qm_complex e[16];
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                const cd a(m[(y >> 1) * 2 + (x >> 1)].re, m[(y >> 1) * 2 + (x >> 1)].im);
                const cd b(m[(y & 1) * 2 + (x & 1)].re, m[(y & 1) * 2 + (x & 1)].im);
                const cd v = a * std::conj(b);
                e[y * 4 + x] = qm_complex{v.real(), v.imag()};
            }
        const u32 ft[2] = {tl[0], tl[0] + (u32)n};
        dense_any(ft, 2, nullptr, nullptr, 0, e, state, flat, stream, dev, what);
        return;
    }
    if (nt + nc <= 2) {
        // one pass with the superoperator F (x) conj(F) of the full (controlled)
        // unitary F on the qubits q = targets then controls; member bits 0..k-1
        // are the column bits of q, bits k..2k-1 the row bits.  A monomial F (CNOT,
        // CZ, SWAP, ...) has a monomial superoperator: only changed entries move.
        const u32 k = nt + nc;
        const u32 d = 1u << k;
        u32 q[2];
        for (u32 i = 0; i < nt; ++i) q[i] = tl[i];
        for (u32 i = 0; i < nc; ++i) q[nt + i] = cl[i];
        u32 cpat = 0;
        for (u32 i = 0; i < nc; ++i) cpat |= cvl[i] << i;
        const u32 tmask = (1u << nt) - 1;
        std::vector<cd> F((size_t)d * d, cd(0.0, 0.0));
        for (u32 x = 0; x < d; ++x) {
            if ((x >> nt) != cpat) {
                F[(size_t)x * d + x] = 1.0;
                continue;
            }
            for (u32 yt = 0; yt <= tmask; ++yt) {
                const u32 y = (x & ~tmask) | yt;
                const qm_complex e = m[(size_t)yt * (tmask + 1) + (x & tmask)];
                F[(size_t)y * d + x] = cd(e.re, e.im);
            }
        }
        const u32 D = d * d;
        std::vector<qm_complex> E((size_t)D * D);
        for (u32 yr = 0; yr < d; ++yr)
            for (u32 yc = 0; yc < d; ++yc)
                for (u32 xr = 0; xr < d; ++xr)
                    for (u32 xc = 0; xc < d; ++xc) {
                        const cd v = F[(size_t)yr * d + xr] * std::conj(F[(size_t)yc * d + xc]);
                        E[(size_t)(yr * d + yc) * D + (xr * d + xc)] = qm_complex{v.real(), v.imag()};
                    }
        u32 ft[4];
        for (u32 i = 0; i < k; ++i) {
            ft[i] = q[i];
            ft[k + i] = q[i] + (u32)n;
        }
        if (is_monomial(E.data(), 2 * k))
            monomial_any(ft, 2 * k, E.data(), state, flat, stream, dev, what);
        else
            dense_any(ft, 2 * k, nullptr, nullptr, 0, E.data(), state, flat, stream, dev, what);
        return;
    }
    u32 rt[32], rc[32], ct[32];
    for (u32 i = 0; i < nt; ++i) { rt[i] = tl[i] + n; ct[i] = tl[i]; }
    for (u32 i = 0; i < nc; ++i) rc[i] = cl[i] + n;
    const size_t N = (size_t)1 << nt;
    std::vector<qm_complex> mc(N * N);
    for (size_t i = 0; i < N * N; ++i) mc[i] = qm_complex{m[i].re, -m[i].im};
    dense_any(rt, nt, nc ? rc : nullptr, cvl, nc, m, state, flat, stream, dev, what);
    dense_any(ct, nt, nc ? cl : nullptr, cvl, nc, mc.data(), state, flat, stream, dev, what);
}

// ------------------------------------------------------------------ single-qubit channels in one pass
// A single-qubit channel whose superoperator S (4 x 4 over member = row bit * 2
// + column bit) maps {00, 11} onto itself and {01, 10} onto itself -- Pauli
// channels (depolarizing, dephasing, bit/phase flips), amplitude and phase
// damping -- only mixes rho[y][x] with rho[y ^ 2^q][x ^ 2^q].  For channels on
// distinct qubits q_0..q_{K-1} the 2^K entries reached from one entry by such
// flips ("an orbit") are closed under all K channels, so one thread loads an
// orbit, applies the K channels as butterflies in registers and stores it: K
// channels in one pass instead of one pass each.
//
// Each orbit is named by one representative entry.  For a channel on a qubit q
// >= 3 the representative has column bit q = 0 and its row bit q is the
// channel's "class" (row bit == column bit or not), taken from the top bits of
// the work index so that it is the same across a wave.  For q < 3 (inside a
// 128-byte line) the representative has row bit q = 0 and the column bit q is a
// free bit, so the two halves of every line are handled by neighbouring lanes;
// the class is then that column bit.  Every wave instruction covers whole
// lines, and the accesses stream.
constexpr int kOrbitMax = 5;  // an orbit of 2^K complex values stays in registers
template <int K>
struct OrbitArgs {
    u64 ins[2 * K];         // (1 << p) - 1 for the ascending inserted positions
    int n_ins;
    u64 colbit[K], rowbit[K];
    u64 free_items;         // work items per class word
    double2 blk[K][2][4];   // [channel][class][2 x 2 block: representative's member first]
};

// Channels 0..L-1 are on qubits below 3 ("low"), L..K-1 on qubits >= 3; the
// class word holds the high channels' classes, bit k - L for channel k.
template <int K, int L>
__global__ __launch_bounds__(256) void k_dm_orbit(double2* __restrict__ s, u64 n_orbits, OrbitArgs<K> a) {
    constexpr int N = 1 << K;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n_orbits; i += stride) {
        // free_items is a multiple of 64, so the class word is uniform across the wave
        const unsigned word = __builtin_amdgcn_readfirstlane((unsigned)(i / a.free_items));
        u64 b = i % a.free_items;
#pragma unroll
        for (int k = 0; k < 2 * K; ++k)
            if (k < a.n_ins) b = (b & a.ins[k]) | ((b & ~a.ins[k]) << 1);
#pragma unroll
        for (int k = L; k < K; ++k)
            if ((word >> (k - L)) & 1) b |= a.rowbit[k];
        // Member x is b ^ (XOR over set bits k of x of colbit | rowbit).  The
        // flipped bits of b are the same in every lane of the wave except the
        // column bits of low channels, so member x = lane 0's b + a uniform delta
        // + a small per-lane offset that depends on x's low-channel bits only.
        const u64 b0 = ((u64)__builtin_amdgcn_readfirstlane((unsigned)(b >> 32)) << 32) |
                       (u64)__builtin_amdgcn_readfirstlane((unsigned)b);
        unsigned lane_bytes[1 << L];
#pragma unroll
        for (int xl = 0; xl < (1 << L); ++xl) {
            long long adj = (long long)(b - b0);
#pragma unroll
            for (int k = 0; k < L; ++k)
                if ((xl >> k) & 1) adj += (b & a.colbit[k]) ? -(long long)a.colbit[k] : (long long)a.colbit[k];
            lane_bytes[xl] = (unsigned)(adj * (long long)sizeof(double2));
        }
        auto member = [&](int x) {
            long long uni = 0;
#pragma unroll
            for (int k = 0; k < K; ++k)
                if ((x >> k) & 1) {
                    if (k < L)
                        uni += (long long)a.rowbit[k];
                    else
                        uni += (long long)a.colbit[k] +
                               (((word >> (k - L)) & 1) ? -(long long)a.rowbit[k] : (long long)a.rowbit[k]);
                }
            return (double2*)((char*)(s + (long long)b0 + uni) + lane_bytes[x & ((1 << L) - 1)]);
        };
        double2 v[N];
#pragma unroll
        for (int x = 0; x < N; ++x) v[x] = ld_stream(member(x));
#pragma unroll
        for (int k = 0; k < K; ++k) {
            // class: for a low channel the representative's column bit, else the word bit
            const int c = k < L ? ((b & a.colbit[k]) ? 1 : 0) : (int)((word >> (k - L)) & 1);
            const double2 m0 = a.blk[k][c][0], m1 = a.blk[k][c][1], m2 = a.blk[k][c][2], m3 = a.blk[k][c][3];
#pragma unroll
            for (int x = 0; x < N; ++x) {
                if ((x >> k) & 1) continue;
                const double2 p = v[x], q = v[x | (1 << k)];
                v[x] = cmac(cmul(m0, p), m1, q);
                v[x | (1 << k)] = cmac(cmul(m2, p), m3, q);
            }
        }
#pragma unroll
        for (int x = 0; x < N; ++x) st_stream(member(x), v[x]);
    }
}

// Channels in `q` (any order); returns false when the matrix is too small for
// whole-wave classes (the caller then applies the channels one by one).
template <int K>
bool launch_orbit(double2* s, u64 flat, int n, const u32* q, const qm_complex* S, hipStream_t st) {
    // low channels first (the channels act on distinct qubits and commute)
    int order[K], L = 0;
    for (int k = 0; k < K; ++k)
        if (q[k] < kLineQubits) order[L++] = k;
    int h = L;
    for (int k = 0; k < K; ++k)
        if (q[k] >= kLineQubits) order[h++] = k;
    OrbitArgs<K> a;
    u32 pos[2 * K];
    int np = 0;
    for (int j = 0; j < K; ++j) {
        const u32 qq = q[order[j]];
        a.colbit[j] = 1ULL << qq;
        a.rowbit[j] = 1ULL << (qq + n);
        pos[np++] = qq + n;
        if (j >= L) pos[np++] = qq;
        const qm_complex* m = S + 16 * order[j];
        auto c = [&](int r, int col) { return make_double2(m[4 * r + col].re, m[4 * r + col].im); };
        // class 0: representative 00, partner 11
        a.blk[j][0][0] = c(0, 0); a.blk[j][0][1] = c(0, 3); a.blk[j][0][2] = c(3, 0); a.blk[j][0][3] = c(3, 3);
        if (j >= L) {  // class 1: representative 10 (row 1, column 0), partner 01
            a.blk[j][1][0] = c(2, 2); a.blk[j][1][1] = c(2, 1); a.blk[j][1][2] = c(1, 2); a.blk[j][1][3] = c(1, 1);
        } else {       // class 1: representative 01 (row 0, column 1), partner 10
            a.blk[j][1][0] = c(1, 1); a.blk[j][1][1] = c(1, 2); a.blk[j][1][2] = c(2, 1); a.blk[j][1][3] = c(2, 2);
        }
    }
    std::sort(pos, pos + np);
    a.n_ins = np;
    for (int k = 0; k < 2 * K; ++k) a.ins[k] = k < np ? (1ULL << pos[k]) - 1 : 0;
    const u64 n_orbits = flat >> K;
    a.free_items = n_orbits >> (K - L);
    if (a.free_items % 64 != 0) return false;
    const unsigned grid = grid_for(n_orbits, 256);
    switch (L) {
        case 0: k_dm_orbit<K, 0><<<grid, 256, 0, st>>>(s, n_orbits, a); break;
        case 1: k_dm_orbit<K, 1><<<grid, 256, 0, st>>>(s, n_orbits, a); break;
        case 2: if constexpr (K >= 2) k_dm_orbit<K, 2><<<grid, 256, 0, st>>>(s, n_orbits, a); break;
        default: if constexpr (K >= 3) k_dm_orbit<K, 3><<<grid, 256, 0, st>>>(s, n_orbits, a); break;
    }
    return true;
}

}  // namespace qm

using namespace qm;

extern "C" {

void dm_initialize_quantum_state_host(void* state, ITYPE dim, void* stream, unsigned dev) {
    int n;
    if (!dm_ok(state, dim, &n, "dm_initialize_quantum_state_host")) return;
    set_computational_basis_host(0, state, dim * dim, stream, dev);
}

void dm_initialize_with_pure_state_host(void* state, const void* pure, ITYPE dim, void* stream, unsigned dev) {
    int n;
    if (!dm_ok(state, dim, &n, "dm_initialize_with_pure_state_host")) return;
    QM_CHECK_ARG(pure, "dm_initialize_with_pure_state_host: null pure state");
    if (!use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    k_outer<<<grid_for(dim * dim, 256 * 4), 256, 0, st>>>((double2*)state, (const double2*)pure, dim);
    finish_update(st, "dm_initialize_with_pure_state_host");
}

void dm_single_qubit_dense_matrix_gate_host(UINT t, const qm_complex m[4], void* state, ITYPE dim, void* stream,
                                            unsigned dev) {
    dm_unitary(nullptr, nullptr, 0, &t, 1, m, state, dim, stream, dev, "dm_single_qubit_dense_matrix_gate_host");
}
void dm_multi_qubit_dense_matrix_gate_host(const UINT* tl, UINT nt, const qm_complex* m, void* state, ITYPE dim,
                                           void* stream, unsigned dev) {
    dm_unitary(nullptr, nullptr, 0, tl, nt, m, state, dim, stream, dev, "dm_multi_qubit_dense_matrix_gate_host");
}
void dm_multi_qubit_control_multi_qubit_dense_matrix_gate_host(const UINT* cl, const UINT* cvl, UINT nc,
                                                               const UINT* tl, UINT nt, const qm_complex* m,
                                                               void* state, ITYPE dim, void* stream, unsigned dev) {
    dm_unitary(cl, cvl, nc, tl, nt, m, state, dim, stream, dev,
               "dm_multi_qubit_control_multi_qubit_dense_matrix_gate_host");
}

void dm_superoperator_host(const UINT* tl, UINT nt, const qm_complex* superop, void* state, ITYPE dim, void* stream,
                           unsigned dev) {
    const char* what = "dm_superoperator_host";
    int n;
    if (!dm_ok(state, dim, &n, what)) return;
    QM_CHECK_ARG(tl && superop && nt >= 1 && nt <= 6 && nt <= (u32)n, "%s: bad target count %u", what, nt);
    if (!check_qubits(tl, nt, n, what)) return;
    // member bits 0..k-1: column bits t_i; bits k..2k-1: row bits t_i + n
    u32 ft[12];
    for (u32 i = 0; i < nt; ++i) {
        ft[i] = tl[i];
        ft[nt + i] = tl[i] + n;
    }
    dense_any(ft, 2 * nt, nullptr, nullptr, 0, superop, state, dim * dim, stream, dev, what);
}

void dm_depolarizing_host(UINT t, double p, void* state, ITYPE dim, void* stream, unsigned dev) {
    // S = (1-p) I(x)I + p/3 (X(x)X + Y(x)conj(Y) + Z(x)Z) on (row, column) bits of t:
    // rho00' = (1-2p/3) rho00 + (2p/3) rho11, rho01' = (1-4p/3) rho01, and symmetric.
    const double a = 1.0 - 2.0 * p / 3.0, b = 2.0 * p / 3.0, c = 1.0 - 4.0 * p / 3.0;
    // member index = row_bit * 2 + column_bit
    const qm_complex S[16] = {
        {a, 0}, {0, 0}, {0, 0}, {b, 0},
        {0, 0}, {c, 0}, {0, 0}, {0, 0},
        {0, 0}, {0, 0}, {c, 0}, {0, 0},
        {b, 0}, {0, 0}, {0, 0}, {a, 0},
    };
    dm_superoperator_host(&t, 1, S, state, dim, stream, dev);
}

void qm_dm_single_qubit_channels_host(const UINT* qubits, UINT count, const qm_complex* superops, void* state,
                                      ITYPE dim, void* stream, unsigned dev) {
    const char* what = "qm_dm_single_qubit_channels_host";
    int n;
    if (!dm_ok(state, dim, &n, what)) return;
    QM_CHECK_ARG(qubits && superops && count >= 1 && count <= (UINT)kOrbitMax && count <= (UINT)n,
                 "%s: %u channels (1..%d, at most the qubit count)", what, count, kOrbitMax);
    if (!check_qubits(qubits, count, n, what)) return;
    for (UINT k = 0; k < count; ++k) {
        const qm_complex* m = superops + 16 * k;
        // entries linking {00, 11} with {01, 10} must vanish
        static const int cross[8][2] = {{0, 1}, {0, 2}, {3, 1}, {3, 2}, {1, 0}, {1, 3}, {2, 0}, {2, 3}};
        for (const auto& rc : cross) {
            const qm_complex z = m[4 * rc[0] + rc[1]];
            QM_CHECK_ARG(z.re == 0.0 && z.im == 0.0, "%s: channel %u mixes diagonal and off-diagonal entries", what, k);
        }
    }
    if (!use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    double2* s = (double2*)state;
    const u64 flat = dim * dim;
    bool done = false;
    switch (count) {
        case 1: done = launch_orbit<1>(s, flat, n, qubits, superops, st); break;
        case 2: done = launch_orbit<2>(s, flat, n, qubits, superops, st); break;
        case 3: done = launch_orbit<3>(s, flat, n, qubits, superops, st); break;
        case 4: done = launch_orbit<4>(s, flat, n, qubits, superops, st); break;
        default: done = launch_orbit<5>(s, flat, n, qubits, superops, st); break;
    }
    if (!done) {  // small matrices: one superoperator pass per channel
        for (UINT k = 0; k < count; ++k)
            dm_superoperator_host(qubits + k, 1, superops + 16 * k, state, dim, stream, dev);
        return;
    }
    finish_update(st, what);
}

double dm_state_norm_squared_host(void* state, ITYPE dim, void* stream, unsigned dev) {
    return dm_diag_sum(false, 0, 0, state, dim, stream, dev, "dm_state_norm_squared_host");
}
double dm_measurement_distribution_entropy_host(void* state, ITYPE dim, void* stream, unsigned dev) {
    return dm_diag_sum(true, 0, 0, state, dim, stream, dev, "dm_measurement_distribution_entropy_host");
}
double dm_M0_prob_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    QM_CHECK_ARG_RET(log2_dim(dim) > (int)t, 0.0, "dm_M0_prob_host: qubit %u out of range", t);
    return dm_diag_sum(false, 1ULL << t, 0, state, dim, stream, dev, "dm_M0_prob_host");
}
double dm_marginal_prob_host(const UINT* tl, const UINT* vl, UINT nt, void* state, ITYPE dim, void* stream,
                             unsigned dev) {
    const int nq = log2_dim(dim);
    u64 mask = 0, value = 0;
    for (UINT i = 0; i < nt; ++i) {
        QM_CHECK_ARG_RET((int)tl[i] < nq && vl[i] <= 1 && !((mask >> tl[i]) & 1), 0.0,
                         "dm_marginal_prob_host: bad qubit %u / value %u", tl[i], vl[i]);
        mask |= 1ULL << tl[i];
        value |= (u64)vl[i] << tl[i];
    }
    return dm_diag_sum(false, mask, value, state, dim, stream, dev, "dm_marginal_prob_host");
}
double dm_expectation_value_multi_qubit_Pauli_operator_partial_list_host(const UINT* tl, const UINT* pl, UINT nt,
                                                                         void* state, ITYPE dim, void* stream,
                                                                         unsigned dev) {
    const char* what = "dm_expectation_value_multi_qubit_Pauli_operator_partial_list_host";
    int n;
    if (!dm_ok(state, dim, &n, what)) return 0.0;
    u64 bf = 0, pf = 0;
    unsigned rot = 0;
    for (UINT c = 0; c < nt; ++c) {
        QM_CHECK_ARG_RET((int)tl[c] < n && pl[c] <= 3, 0.0, "%s: bad qubit %u or Pauli id %u", what, tl[c], pl[c]);
        switch (pl[c]) {
            case 1: bf ^= 1ULL << tl[c]; break;
            case 2: bf ^= 1ULL << tl[c]; pf ^= 1ULL << tl[c]; rot++; break;
            case 3: pf ^= 1ULL << tl[c]; break;
            default: break;
        }
    }
    if (!use_device(dev)) return 0.0;
    hipStream_t st = stream_of(stream);
    Scratch* sc = scratch_for(st);
    if (!sc) return 0.0;
    const unsigned nb = grid_for(dim, 256 * 4, 1024);
    k_dm_pauli<<<nb, 256, 0, st>>>((const double2*)state, dim, bf, pf, rot, sc->partials);
    return finish_scalar(sc, nb, st, what);
}

void dm_normalize_host(double squared_norm, void* state, ITYPE dim, void* stream, unsigned dev) {
    int n;
    if (!dm_ok(state, dim, &n, "dm_normalize_host") || !use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    k_dm_scale<<<grid_for(dim * dim, 256 * 4), 256, 0, st>>>((double2*)state, dim * dim, 1.0 / squared_norm);
    finish_update(st, "dm_normalize_host");
}

}  // extern "C"
