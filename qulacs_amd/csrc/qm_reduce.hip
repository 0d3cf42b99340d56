// Reductions over a state vector, in FP64, deterministic: every block reduces
// a fixed slice (wave shuffle tree, then a fixed-order sum over waves) into a
// partials array, and one block sums the partials in index order.  No atomics,
// no per-call allocation (the reference allocates, memsets and atomically adds
// per call: src/gpusim/stat_ops.cu:140-198).
//
// Reference loops restated: src/csim/stat_ops.cpp:13-86,
// stat_ops_probability.cpp:11-100, stat_ops_expectation_value.cpp:352-482,
// stat_ops_transition_amplitude.cpp:19-128, update_ops_named_state.cpp:21-98.
#include <cmath>

#include "qm_internal.hpp"

namespace qm {

constexpr int kRBlock = 256;
constexpr unsigned kMaxRBlocks = 2048;

// real part of z * i^k
__device__ __forceinline__ double re_rot(double2 z, unsigned k) {
    switch (k & 3) {
        case 0: return z.x;
        case 1: return -z.y;
        case 2: return -z.x;
        default: return z.y;
    }
}
__device__ __forceinline__ double2 rot(double2 z, unsigned k) {  // z * i^k
    switch (k & 3) {
        case 0: return z;
        case 1: return make_double2(-z.y, z.x);
        case 2: return make_double2(-z.x, -z.y);
        default: return make_double2(z.y, -z.x);
    }
}
__device__ __forceinline__ unsigned par(u64 x) { return __popcll(x) & 1; }
__device__ __forceinline__ double norm2(double2 v) { return fma(v.x, v.x, v.y * v.y); }

// One Pauli term (or a plain reduction) as seen by the kernels.
struct Term {
    u64 bf, pf;
    unsigned rot, pivot;
};

// Kinds of per-element functionals.
enum RKind : int {
    R_NORM = 0,      // sum |psi_i|^2
    R_ENTROPY = 1,   // sum -p log p with p clamped to >= 1e-15
    R_MASKED = 2,    // sum |psi_b|^2 over b with (b & mask) == value (M0, M1, marginal)
    R_PAULI = 3,     // <psi|P|psi> (XZ or Z form per term)
    R_TRANS = 4,     // <bra|P|ket> (complex)
    R_INNER = 5,     // <bra|ket> (complex)
};

struct RArgs {
    const double2* __restrict__ a;  // state / ket
    const double2* __restrict__ b;  // bra (R_TRANS, R_INNER)
    u64 dim;
    u64 mask, value;                // R_MASKED
    const Term* __restrict__ terms; // R_PAULI / R_TRANS: one per blockIdx.y
    double* __restrict__ partials;  // [gridDim.y][gridDim.x][2]
};

template <int KIND>
__global__ __launch_bounds__(kRBlock) void k_reduce(RArgs r) {
    __shared__ double lds[32];
    double acc_re = 0.0, acc_im = 0.0;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 t0 = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (KIND == R_NORM) {
        for (u64 i = t0; i < r.dim; i += stride) acc_re += norm2(r.a[i]);
    } else if (KIND == R_ENTROPY) {
        for (u64 i = t0; i < r.dim; i += stride) {
            double p = norm2(r.a[i]);
            p = p > 1e-15 ? p : 1e-15;
            acc_re += -p * log(p);
        }
    } else if (KIND == R_MASKED) {
        for (u64 i = t0; i < r.dim; i += stride)
            if ((i & r.mask) == r.value) acc_re += norm2(r.a[i]);
    } else if (KIND == R_PAULI) {
        const Term tm = r.terms[blockIdx.y];
        if (tm.bf == 0) {
            for (u64 i = t0; i < r.dim; i += stride) {
                const double p = norm2(r.a[i]);
                acc_re += par(i & tm.pf) ? -p : p;
            }
        } else {
            const u64 half = r.dim >> 1;
            for (u64 i = t0; i < half; i += stride) {
                const u64 b0 = insert_zero(i, tm.pivot), b1 = b0 ^ tm.bf;
                const double2 x = r.a[b0], y = r.a[b1];
                // x * conj(y)
                const double2 z = make_double2(fma(x.x, y.x, x.y * y.y), fma(x.y, y.x, -x.x * y.y));
                acc_re += 2.0 * re_rot(z, tm.rot + 2 * par(b0 & tm.pf));
            }
        }
    } else if (KIND == R_TRANS) {
        const Term tm = r.terms[blockIdx.y];
        const double2* ket = r.a;
        const double2* bra = r.b;
        if (tm.bf == 0) {
            for (u64 i = t0; i < r.dim; i += stride) {
                const double2 k = ket[i], b = bra[i];
                // k * conj(b)
                double2 z = make_double2(fma(k.x, b.x, k.y * b.y), fma(k.y, b.x, -k.x * b.y));
                if (par(i & tm.pf)) z = make_double2(-z.x, -z.y);
                acc_re += z.x;
                acc_im += z.y;
            }
        } else {
            const u64 half = r.dim >> 1;
            for (u64 i = t0; i < half; i += stride) {
                const u64 b0 = insert_zero(i, tm.pivot), b1 = b0 ^ tm.bf;
                const double2 k0 = ket[b0], k1 = ket[b1], c0 = bra[b0], c1 = bra[b1];
                const double2 z1 = rot(make_double2(fma(k0.x, c1.x, k0.y * c1.y), fma(k0.y, c1.x, -k0.x * c1.y)),
                                       tm.rot + 2 * par(b0 & tm.pf));
                const double2 z2 = rot(make_double2(fma(k1.x, c0.x, k1.y * c0.y), fma(k1.y, c0.x, -k1.x * c0.y)),
                                       tm.rot + 2 * par(b1 & tm.pf));
                acc_re += z1.x + z2.x;
                acc_im += z1.y + z2.y;
            }
        }
    } else if (KIND == R_INNER) {
        for (u64 i = t0; i < r.dim; i += stride) {
            const double2 b = r.b[i], k = r.a[i];
            // conj(b) * k
            acc_re += fma(b.x, k.x, b.y * k.y);
            acc_im += fma(b.x, k.y, -b.y * k.x);
        }
    }
    const double s_re = block_sum(acc_re, lds);
    double s_im = 0.0;
    if (KIND == R_TRANS || KIND == R_INNER) s_im = block_sum(acc_im, lds);
    if (threadIdx.x == 0) {
        double* out = r.partials + ((u64)blockIdx.y * gridDim.x + blockIdx.x) * 2;
        out[0] = s_re;
        out[1] = s_im;
    }
}

// Sums partials[term][0..nb) per term, multiplies by the term's coefficient
// and writes per-term values plus their total.  One wave per term (16 terms at
// a time): lane-strided sums in index order, then a fixed shuffle tree; the
// per-term values are added in term order.  Deterministic, and with no block
// barriers inside the term loop.
constexpr int kFinishThreads = 1024;
__global__ __launch_bounds__(kFinishThreads) void k_finish(const double* __restrict__ partials, unsigned nb,
                                                           unsigned nterms, const double2* __restrict__ coefs,
                                                           double* __restrict__ out) {
    const unsigned lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (unsigned t = wid; t < nterms; t += nw) {
        const double* p = partials + (u64)t * nb * 2;
        double re = 0.0, im = 0.0;
        for (unsigned k = lane; k < nb; k += 64) {
            re += p[2 * k];
            im += p[2 * k + 1];
        }
        re = wave_sum(re);
        im = wave_sum(im);
        if (coefs) {
            const double2 c = coefs[t];
            const double pr = c.x * re - c.y * im, pi = c.x * im + c.y * re;
            re = pr;
            im = pi;
        }
        if (lane == 0) {
            out[2 + 2 * t] = re;
            out[3 + 2 * t] = im;
        }
    }
    __syncthreads();  // the per-term values written above are visible to the whole block
    if (threadIdx.x == 0) {
        double total_re = 0.0, total_im = 0.0;
        for (unsigned t = 0; t < nterms; ++t) {
            total_re += out[2 + 2 * t];
            total_im += out[3 + 2 * t];
        }
        out[0] = total_re;
        out[1] = total_im;
    }
}

unsigned reduce_blocks(u64 work) { return grid_for(work, (u64)kRBlock * 16, kMaxRBlocks); }

// Enqueues one reduction family; on success *dout points at the device result
// (2 + 2 * max(nterms, 1) doubles: the ordered total, then per-term values),
// valid until the next reduction on the same stream.
template <int KIND>
bool enqueue_reduce(RArgs r, u64 work, const Term* host_terms, unsigned nterms, const qm_complex* coefs, hipStream_t st,
                    double** dout, unsigned* ny_out) {
    Scratch* sc = scratch_for(st);
    if (!sc) return false;
    const unsigned nb = reduce_blocks(work);
    const unsigned ny = nterms ? nterms : 1;
    if ((u64)nb * ny * 2 + 2 + 2 * (u64)ny > kPartialsCap || 2 + 2 * (u64)ny > kHostResultCap) {
        set_error(QM_ERR_UNSUPPORTED, "too many terms in one reduction (%u)", ny);
        return false;
    }
    if (nterms) {
        r.terms = (const Term*)upload(sc, host_terms, sizeof(Term) * nterms);
        if (!r.terms) return false;
    }
    const double2* dcoefs = nullptr;
    if (coefs) {
        dcoefs = (const double2*)upload(sc, coefs, sizeof(double2) * ny);
        if (!dcoefs) return false;
    }
    r.partials = sc->partials;
    double* out = sc->partials + (u64)nb * ny * 2;
    k_reduce<KIND><<<dim3(nb, ny), kRBlock, 0, st>>>(r);
    k_finish<<<1, kFinishThreads, 0, st>>>(sc->partials, nb, ny, dcoefs, out);
    *dout = out;
    *ny_out = ny;
    return true;
}

// Runs one reduction family and waits for it.  `nterms` > 0 for term-indexed
// kinds.  Returns false on error; result (re, im) of the ordered total in res.
template <int KIND>
bool run_reduce(RArgs r, u64 work, const Term* host_terms, unsigned nterms, const qm_complex* coefs, hipStream_t st,
                double res[2], std::vector<double>* per_term = nullptr) {
    double* dout = nullptr;
    unsigned ny = 0;
    if (!enqueue_reduce<KIND>(r, work, host_terms, nterms, coefs, st, &dout, &ny)) return false;
    Scratch* sc = scratch_for(st);
    const size_t nout = 2 + 2 * (size_t)ny;
    if (!hip_ok(hipMemcpyAsync(sc->host_result, dout, nout * sizeof(double), hipMemcpyDeviceToHost, st), "reduce"))
        return false;
    if (!hip_ok(hipStreamSynchronize(st), "reduce sync")) return false;
    res[0] = sc->host_result[0];
    res[1] = sc->host_result[1];
    if (per_term) per_term->assign(sc->host_result + 2, sc->host_result + nout);
    return true;
}

bool enqueue_observable(const u64* bf, const u64* pf, const u32* rot, const double* coef, u32 nterms,
                        const void* state, u64 dim, hipStream_t st, double* h_out) {
    std::vector<Term> zt, xt;
    std::vector<qm_complex> zc, xc;
    for (u32 k = 0; k < nterms; ++k) {
        const unsigned pivot = bf[k] ? 63 - __builtin_clzll(bf[k]) : 0;
        Term t{bf[k], pf[k], rot[k], pivot};
        if (bf[k]) {
            xt.push_back(t);
            xc.push_back(qm_complex{coef[k], 0.0});
        } else {
            zt.push_back(t);
            zc.push_back(qm_complex{coef[k], 0.0});
        }
    }
    RArgs r{};
    r.a = (const double2*)state;
    r.dim = dim;
    for (int g = 0; g < 2; ++g) {
        const std::vector<Term>& tv = g ? xt : zt;
        const std::vector<qm_complex>& cv = g ? xc : zc;
        if (tv.empty()) {
            h_out[2 * g] = 0.0;
            h_out[2 * g + 1] = 0.0;
            continue;
        }
        double* dout = nullptr;
        unsigned ny = 0;
        if (!enqueue_reduce<R_PAULI>(r, g ? dim / 2 : dim, tv.data(), (unsigned)tv.size(), cv.data(), st, &dout, &ny))
            return false;
        // copied before the next reduction on this stream can reuse the scratch
        if (!hip_ok(hipMemcpyAsync(h_out + 2 * g, dout, 2 * sizeof(double), hipMemcpyDeviceToHost, st),
                    "enqueue_observable"))
            return false;
    }
    return true;
}

bool state_ok(const void* s, u64 dim, const char* what) {
    if (!s || log2_dim(dim) < 0) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null state or dim %llu is not a power of two >= 2", what, dim);
        return false;
    }
    return true;
}

Term make_term(u64 bf, u64 pf, unsigned rot, unsigned pivot) {
    Term t;
    t.bf = bf;
    t.pf = pf;
    t.rot = rot & 3;
    t.pivot = pivot;
    return t;
}

bool pauli_term_from_list(const u32* tl, const u32* pl, u32 n, int nq, Term* t, const char* what) {
    u64 bf = 0, pf = 0;
    unsigned rot = 0, pivot = 0;
    for (u32 c = 0; c < n; ++c) {
        if ((int)tl[c] >= nq) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: qubit %u out of range", what, tl[c]);
            return false;
        }
        switch (pl[c]) {
            case 0: break;
            case 1: bf ^= 1ULL << tl[c]; pivot = tl[c]; break;
            case 2: bf ^= 1ULL << tl[c]; pf ^= 1ULL << tl[c]; rot++; pivot = tl[c]; break;
            case 3: pf ^= 1ULL << tl[c]; break;
            default:
                set_error(QM_ERR_INVALID_ARGUMENT, "%s: Pauli id %u out of 0..3", what, pl[c]);
                return false;
        }
    }
    *t = make_term(bf, pf, rot, pivot);
    return true;
}

double pauli_expectation(const Term& t, void* state, u64 dim, void* stream, unsigned dev, const char* what) {
    if (!state_ok(state, dim, what)) return 0.0;
    const int n = log2_dim(dim);
    if (((t.bf | t.pf) >> n) != 0 || (t.bf && !((t.bf >> t.pivot) & 1))) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: masks beyond %d qubits or pivot outside the flip mask", what, n);
        return 0.0;
    }
    if (!use_device(dev)) return 0.0;
    RArgs r{};
    r.a = (const double2*)state;
    r.dim = dim;
    double res[2] = {0, 0};
    run_reduce<R_PAULI>(r, t.bf ? dim / 2 : dim, &t, 1, nullptr, stream_of(stream), res);
    return res[0];
}

qm_complex transition(const Term& t, void* bra, void* ket, u64 dim, void* stream, unsigned dev, const char* what) {
    qm_complex z{0, 0};
    if (!state_ok(bra, dim, what) || !state_ok(ket, dim, what)) return z;
    const int n = log2_dim(dim);
    if (((t.bf | t.pf) >> n) != 0 || (t.bf && !((t.bf >> t.pivot) & 1))) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: masks beyond %d qubits or pivot outside the flip mask", what, n);
        return z;
    }
    if (!use_device(dev)) return z;
    RArgs r{};
    r.a = (const double2*)ket;
    r.b = (const double2*)bra;
    r.dim = dim;
    double res[2] = {0, 0};
    if (run_reduce<R_TRANS>(r, t.bf ? dim / 2 : dim, &t, 1, nullptr, stream_of(stream), res)) {
        z.re = res[0];
        z.im = res[1];
    }
    return z;
}

// ------------------------------------------------------------------ elementwise state arithmetic
__global__ __launch_bounds__(256) void k_axpy(double2* __restrict__ s, const double2* __restrict__ a, u64 dim,
                                              double2 coef) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) st_stream(s + i, cmac(ld_stream(s + i), coef, ld_stream(a + i)));
}
__global__ __launch_bounds__(256) void k_add(double2* __restrict__ s, const double2* __restrict__ a, u64 dim) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) st_stream(s + i, cadd(ld_stream(s + i), ld_stream(a + i)));
}
// Elementwise in-place passes: from 128 amplitudes up, each thread takes two
// amplitudes 64 apart (both loads issued before either store), so a wave reads
// 2 KiB of consecutive lines per iteration; smaller states take one per thread.
template <class F>
__device__ __forceinline__ void inplace_pass(double2* __restrict__ s, u64 dim, F f) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 t0 = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (dim < 128) {
        for (u64 i = t0; i < dim; i += stride) s[i] = f(s[i]);
        return;
    }
    for (u64 i = t0; i < (dim >> 1); i += stride) {
        const u64 a = ((i >> 6) << 7) | (i & 63);
        const double2 x = ld_stream(s + a), y = ld_stream(s + a + 64);
        st_stream(s + a, f(x));
        st_stream(s + a + 64, f(y));
    }
}
__global__ __launch_bounds__(512) void k_scale_c(double2* __restrict__ s, u64 dim, double2 coef) {
    inplace_pass(s, dim, [coef](double2 v) { return cmul(v, coef); });
}
__global__ __launch_bounds__(512) void k_scale_r(double2* __restrict__ s, u64 dim, double f) {
    inplace_pass(s, dim, [f](double2 v) { return cscale(v, f); });
}

// ------------------------------------------------------------------ allocation-free observable
// qm_observable_expectation_noalloc_host runs the same two launches per term
// group as qm_observable_expectation_host (k_reduce<R_PAULI>, then k_finish),
// but its term table, coefficients, per-block partials and group results live
// in `scratch`, a fixed device array, instead of the library's per-stream
// allocations.  Every block of k_reduce writes its partial slot and k_finish
// writes every result slot, so nothing in it has to start at zero.  All calls
// on one device share that device's copy: calls that overlap on one device
// (other threads, other streams) race.
constexpr unsigned kNoAllocMaxTerms = 128;
constexpr size_t kNoAllocTermsOff = 0;
constexpr size_t kNoAllocCoefsOff = 4096;                                        // >= 128 * sizeof(Term)
constexpr size_t kNoAllocOutOff = kNoAllocCoefsOff + 2048;                       // 128 * 16 bytes
constexpr size_t kNoAllocOutDoubles = 2 + 2 * kNoAllocMaxTerms;                  // per group
constexpr size_t kNoAllocPartialsOff = kNoAllocOutOff + 2 * kNoAllocOutDoubles * 8 + 256 - (2 * kNoAllocOutDoubles * 8) % 256;
constexpr size_t kNoAllocBytes = kNoAllocPartialsOff + (size_t)kMaxRBlocks * kNoAllocMaxTerms * 2 * sizeof(double);
static_assert(kNoAllocMaxTerms * sizeof(Term) <= kNoAllocCoefsOff, "term table overlaps the coefficients");
static_assert(kNoAllocPartialsOff % 256 == 0, "partials not 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per
// device; the array's name itself is only a host-side stand-in).
unsigned char* noalloc_scratch(unsigned dev, const char* what) {
    if (dev >= (unsigned)kNoAllocMaxDevices) {
        set_error(QM_ERR_UNSUPPORTED, "%s: device %u beyond %d", what, dev, kNoAllocMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), what)) return nullptr;
        g_noalloc_ptr[dev] = (unsigned char*)p;
    }
    return g_noalloc_ptr[dev];
}

}  // namespace qm

using namespace qm;

extern "C" {

double state_norm_squared_host(void* state, ITYPE dim, void* stream, unsigned dev) {
    if (!state_ok(state, dim, "state_norm_squared_host") || !use_device(dev)) return 0.0;
    RArgs r{};
    r.a = (const double2*)state;
    r.dim = dim;
    double res[2] = {0, 0};
    run_reduce<R_NORM>(r, dim, nullptr, 0, nullptr, stream_of(stream), res);
    return res[0];
}

double state_norm_squared_cublas_host(void* state, ITYPE dim) {
    return state_norm_squared_host(state, dim, nullptr, (unsigned)get_current_device());
}

double measurement_distribution_entropy_host(void* state, ITYPE dim, void* stream, unsigned dev) {
    if (!state_ok(state, dim, "measurement_distribution_entropy_host") || !use_device(dev)) return 0.0;
    RArgs r{};
    r.a = (const double2*)state;
    r.dim = dim;
    double res[2] = {0, 0};
    run_reduce<R_ENTROPY>(r, dim, nullptr, 0, nullptr, stream_of(stream), res);
    return res[0];
}

static double masked_prob(u64 mask, u64 value, void* state, ITYPE dim, void* stream, unsigned dev, const char* what) {
    if (!state_ok(state, dim, what) || !use_device(dev)) return 0.0;
    RArgs r{};
    r.a = (const double2*)state;
    r.dim = dim;
    r.mask = mask;
    r.value = value;
    double res[2] = {0, 0};
    run_reduce<R_MASKED>(r, dim, nullptr, 0, nullptr, stream_of(stream), res);
    return res[0];
}

double M0_prob_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    QM_CHECK_ARG_RET(log2_dim(dim) > (int)t, 0.0, "M0_prob_host: qubit %u out of range", t);
    return masked_prob(1ULL << t, 0, state, dim, stream, dev, "M0_prob_host");
}
double M1_prob_host(UINT t, void* state, ITYPE dim, void* stream, unsigned dev) {
    QM_CHECK_ARG_RET(log2_dim(dim) > (int)t, 0.0, "M1_prob_host: qubit %u out of range", t);
    return masked_prob(1ULL << t, 1ULL << t, state, dim, stream, dev, "M1_prob_host");
}
double marginal_prob_host(UINT* tl, UINT* vl, UINT n, void* state, ITYPE dim, void* stream, unsigned dev) {
    const int nq = log2_dim(dim);
    u64 mask = 0, value = 0;
    for (UINT i = 0; i < n; ++i) {
        QM_CHECK_ARG_RET((int)tl[i] < nq && vl[i] <= 1 && !((mask >> tl[i]) & 1), 0.0,
                         "marginal_prob_host: bad qubit %u / value %u", tl[i], vl[i]);
        mask |= 1ULL << tl[i];
        value |= (u64)vl[i] << tl[i];
    }
    return masked_prob(mask, value, state, dim, stream, dev, "marginal_prob_host");
}

double expectation_value_multi_qubit_Pauli_operator_XZ_mask_host(ITYPE bf, ITYPE pf, UINT rot, UINT pivot,
                                                                 void* state, ITYPE dim, void* stream, unsigned dev) {
    QM_CHECK_ARG_RET(bf != 0, 0.0, "expectation XZ mask: empty bit-flip mask");
    return pauli_expectation(make_term(bf, pf, rot, pivot), state, dim, stream, dev,
                             "expectation_value_multi_qubit_Pauli_operator_XZ_mask_host");
}
double expectation_value_multi_qubit_Pauli_operator_Z_mask_host(ITYPE pf, void* state, ITYPE dim, void* stream,
                                                                unsigned dev) {
    return pauli_expectation(make_term(0, pf, 0, 0), state, dim, stream, dev,
                             "expectation_value_multi_qubit_Pauli_operator_Z_mask_host");
}
double expectation_value_multi_qubit_Pauli_operator_partial_list_host(const UINT* tl, const UINT* pl, UINT n,
                                                                      void* state, ITYPE dim, void* stream,
                                                                      unsigned dev) {
    const char* what = "expectation_value_multi_qubit_Pauli_operator_partial_list_host";
    if (!state_ok(state, dim, what)) return 0.0;
    Term t;
    if (!pauli_term_from_list(tl, pl, n, log2_dim(dim), &t, what)) return 0.0;
    return pauli_expectation(t, state, dim, stream, dev, what);
}
double expectation_value_multi_qubit_Pauli_operator_whole_list_host(const UINT* pl, UINT n, void* state, ITYPE dim,
                                                                    void* stream, unsigned dev) {
    std::vector<u32> tl(n);
    for (UINT i = 0; i < n; ++i) tl[i] = i;
    return expectation_value_multi_qubit_Pauli_operator_partial_list_host(tl.data(), pl, n, state, dim, stream, dev);
}
double expectation_value_single_qubit_Pauli_operator_host(unsigned op, unsigned t, void* state, ITYPE dim,
                                                          void* stream, unsigned dev) {
    if (op == 0) return state_norm_squared_host(state, dim, stream, dev);
    return expectation_value_multi_qubit_Pauli_operator_partial_list_host(&t, &op, 1, state, dim, stream, dev);
}

qm_complex transition_amplitude_multi_qubit_Pauli_operator_XZ_mask_host(ITYPE bf, ITYPE pf, UINT rot, UINT pivot,
                                                                        void* bra, void* ket, ITYPE dim,
                                                                        void* stream, unsigned dev) {
    if (bf == 0) {
        set_error(QM_ERR_INVALID_ARGUMENT, "transition amplitude XZ mask: empty bit-flip mask");
        return qm_complex{0, 0};
    }
    return transition(make_term(bf, pf, rot, pivot), bra, ket, dim, stream, dev,
                      "transition_amplitude_multi_qubit_Pauli_operator_XZ_mask_host");
}
qm_complex transition_amplitude_multi_qubit_Pauli_operator_Z_mask_host(ITYPE pf, void* bra, void* ket, ITYPE dim,
                                                                       void* stream, unsigned dev) {
    return transition(make_term(0, pf, 0, 0), bra, ket, dim, stream, dev,
                      "transition_amplitude_multi_qubit_Pauli_operator_Z_mask_host");
}
qm_complex transition_amplitude_multi_qubit_Pauli_operator_partial_list_host(const UINT* tl, const UINT* pl, UINT n,
                                                                             void* bra, void* ket, ITYPE dim,
                                                                             void* stream, unsigned dev) {
    const char* what = "transition_amplitude_multi_qubit_Pauli_operator_partial_list_host";
    if (!state_ok(ket, dim, what)) return qm_complex{0, 0};
    Term t;
    if (!pauli_term_from_list(tl, pl, n, log2_dim(dim), &t, what)) return qm_complex{0, 0};
    return transition(t, bra, ket, dim, stream, dev, what);
}
qm_complex transition_amplitude_multi_qubit_Pauli_operator_whole_list_host(const UINT* pl, UINT n, void* bra,
                                                                           void* ket, ITYPE dim, void* stream,
                                                                           unsigned dev) {
    std::vector<u32> tl(n);
    for (UINT i = 0; i < n; ++i) tl[i] = i;
    return transition_amplitude_multi_qubit_Pauli_operator_partial_list_host(tl.data(), pl, n, bra, ket, dim, stream,
                                                                             dev);
}

qm_complex inner_product_host(const void* bra, const void* ket, ITYPE dim, void* stream, unsigned dev) {
    qm_complex z{0, 0};
    if (!state_ok(bra, dim, "inner_product_host") || !state_ok(ket, dim, "inner_product_host") || !use_device(dev))
        return z;
    RArgs r{};
    r.a = (const double2*)ket;
    r.b = (const double2*)bra;
    r.dim = dim;
    double res[2] = {0, 0};
    if (run_reduce<R_INNER>(r, dim, nullptr, 0, nullptr, stream_of(stream), res)) {
        z.re = res[0];
        z.im = res[1];
    }
    return z;
}

qm_complex qm_observable_expectation_host(const ITYPE* bfs, const ITYPE* pfs, const UINT* rots,
                                          const qm_complex* coefs, UINT nterms, void* state, ITYPE dim,
                                          void* stream, unsigned dev) {
    const char* what = "qm_observable_expectation_host";
    qm_complex z{0, 0};
    if (!state_ok(state, dim, what)) return z;
    if (nterms == 0) return z;
    if (!bfs || !pfs || !rots || !coefs) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null term arrays", what);
        return z;
    }
    if (!use_device(dev)) return z;
    const int n = log2_dim(dim);
    // Split into Z-only terms and flip terms: each group runs as one launch
    // (blockIdx.y = term) so a group never re-launches per term.
    std::vector<Term> zt, xt;
    std::vector<qm_complex> zc, xc;
    for (UINT k = 0; k < nterms; ++k) {
        if (((bfs[k] | pfs[k]) >> n) != 0) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: term %u acts beyond %d qubits", what, k, n);
            return z;
        }
        unsigned pivot = bfs[k] ? 63 - __builtin_clzll(bfs[k]) : 0;
        const Term t = make_term(bfs[k], pfs[k], rots[k], pivot);
        if (bfs[k]) { xt.push_back(t); xc.push_back(coefs[k]); }
        else { zt.push_back(t); zc.push_back(coefs[k]); }
    }
    hipStream_t st = stream_of(stream);
    RArgs r{};
    r.a = (const double2*)state;
    r.dim = dim;
    double total_re = 0, total_im = 0;
    const size_t chunk = 256;
    for (int g = 0; g < 2; ++g) {
        const std::vector<Term>& tv = g ? xt : zt;
        const std::vector<qm_complex>& cv = g ? xc : zc;
        for (size_t off = 0; off < tv.size(); off += chunk) {
            const unsigned cnt = (unsigned)std::min(chunk, tv.size() - off);
            double res[2] = {0, 0};
            if (!run_reduce<R_PAULI>(r, g ? dim / 2 : dim, tv.data() + off, cnt, cv.data() + off, st, res)) return z;
            total_re += res[0];
            total_im += res[1];
        }
    }
    z.re = total_re;
    z.im = total_im;
    return z;
}

// ------------------------------------------------------------------ allocation-free observable
qm_complex qm_observable_expectation_noalloc_host(const ITYPE* bfs, const ITYPE* pfs, const UINT* rots,
                                                 const qm_complex* coefs, UINT nterms, void* state, ITYPE dim,
                                                 void* stream, unsigned dev) {
    const char* what = "qm_observable_expectation_noalloc_host";
    qm_complex z{0, 0};
    if (!state_ok(state, dim, what)) return z;
    if (nterms == 0) return z;
    if (!bfs || !pfs || !rots || !coefs) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null term arrays", what);
        return z;
    }
    if (nterms > kNoAllocMaxTerms) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: %u terms exceed the %u this entry point holds", what, nterms,
                  kNoAllocMaxTerms);
        return z;
    }
    const int n = log2_dim(dim);
    for (UINT k = 0; k < nterms; ++k)
        if (((bfs[k] | pfs[k]) >> n) != 0) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: term %u acts beyond %d qubits", what, k, n);
            return z;
        }
    // Z-only terms first, then flip terms, each group in input order (the
    // grouping of the allocating entry point).
    Term ht[kNoAllocMaxTerms];
    qm_complex hc[kNoAllocMaxTerms];
    unsigned count[2] = {0, 0};
    for (int g = 0; g < 2; ++g)
        for (UINT k = 0; k < nterms; ++k) {
            if ((bfs[k] != 0) != (g == 1)) continue;
            const unsigned pivot = bfs[k] ? 63 - __builtin_clzll(bfs[k]) : 0;
            ht[count[0] + count[1]] = make_term(bfs[k], pfs[k], rots[k], pivot);
            hc[count[0] + count[1]] = coefs[k];
            ++count[g];
        }
    const unsigned first[2] = {0, count[0]};
    if (!use_device(dev)) return z;
    unsigned char* base = noalloc_scratch(dev, what);
    if (!base) return z;
    Term* dterms = (Term*)(base + kNoAllocTermsOff);
    double2* dcoefs = (double2*)(base + kNoAllocCoefsOff);
    double* dout = (double*)(base + kNoAllocOutOff);
    double* partials = (double*)(base + kNoAllocPartialsOff);
    hipStream_t st = stream_of(stream);
    if (!hip_ok(hipMemcpyAsync(dterms, ht, nterms * sizeof(Term), hipMemcpyHostToDevice, st), what) ||
        !hip_ok(hipMemcpyAsync(dcoefs, hc, nterms * sizeof(double2), hipMemcpyHostToDevice, st), what))
        return z;
    RArgs r{};
    r.a = (const double2*)state;
    r.dim = dim;
    r.partials = partials;
    for (int g = 0; g < 2; ++g) {
        if (!count[g]) continue;
        const unsigned nb = reduce_blocks(g ? dim / 2 : dim);
        r.terms = dterms + first[g];
        k_reduce<R_PAULI><<<dim3(nb, count[g]), kRBlock, 0, st>>>(r);
        k_finish<<<1, kFinishThreads, 0, st>>>(partials, nb, count[g], dcoefs + first[g], dout + g * kNoAllocOutDoubles);
    }
    double res[2][2] = {{0, 0}, {0, 0}};
    for (int g = 0; g < 2; ++g)
        if (count[g] && !hip_ok(hipMemcpyAsync(res[g], dout + g * kNoAllocOutDoubles, 2 * sizeof(double),
                                               hipMemcpyDeviceToHost, st), what))
            return z;
    if (!hip_ok(hipGetLastError(), what) || !hip_ok(hipStreamSynchronize(st), what)) return z;
    double total_re = 0, total_im = 0;
    for (int g = 0; g < 2; ++g)
        if (count[g]) {
            total_re += res[g][0];
            total_im += res[g][1];
        }
    z.re = total_re;
    z.im = total_im;
    return z;
}

// ------------------------------------------------------------------ state arithmetic
void normalize_host(double squared_norm, void* state, ITYPE dim, void* stream, unsigned dev) {
    if (!state_ok(state, dim, "normalize_host") || !use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    const double f = std::sqrt(1.0 / squared_norm);  // src/csim/update_ops_named_state.cpp:23
    k_scale_r<<<grid_for(dim >> 1, 512), 512, 0, st>>>((double2*)state, dim, f);
    finish_update(st, "normalize_host");
}
void state_add_host(void* added, void* state, ITYPE dim, void* stream, unsigned dev) {
    if (!state_ok(state, dim, "state_add_host") || !state_ok(added, dim, "state_add_host") || !use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    k_add<<<grid_for(dim, 256 * 4), 256, 0, st>>>((double2*)state, (const double2*)added, dim);
    finish_update(st, "state_add_host");
}
void qm_state_add_with_coef_host(qm_complex coef, const void* added, void* state, ITYPE dim, void* stream,
                                 unsigned dev) {
    if (!state_ok(state, dim, "qm_state_add_with_coef_host") || !state_ok(added, dim, "qm_state_add_with_coef_host") ||
        !use_device(dev))
        return;
    hipStream_t st = stream_of(stream);
    k_axpy<<<grid_for(dim, 256 * 4), 256, 0, st>>>((double2*)state, (const double2*)added, dim,
                                                    make_double2(coef.re, coef.im));
    finish_update(st, "qm_state_add_with_coef_host");
}
void state_multiply_host(qm_complex coef, void* state, ITYPE dim, void* stream, unsigned dev) {
    if (!state_ok(state, dim, "state_multiply_host") || !use_device(dev)) return;
    hipStream_t st = stream_of(stream);
    k_scale_c<<<grid_for(dim >> 1, 512), 512, 0, st>>>((double2*)state, dim, make_double2(coef.re, coef.im));
    finish_update(st, "state_multiply_host");
}

}  // extern "C"
