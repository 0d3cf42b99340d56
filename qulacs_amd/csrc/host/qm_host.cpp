// Implementation of the host-side qulacs API over the C ABI (see qm_host.hpp).
#include "qm_host.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <deque>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <numeric>
#include <sstream>

namespace qmh {

namespace {
inline const qm_complex* qc(const CPPCTYPE* p) { return reinterpret_cast<const qm_complex*>(p); }
inline qm_complex qc(CPPCTYPE z) { return qm_complex{z.real(), z.imag()}; }
inline CPPCTYPE cc(qm_complex z) { return CPPCTYPE(z.re, z.im); }
const CPPCTYPE I_(0.0, 1.0);

ComplexMatrix pauli_matrix(unsigned p) {
    ComplexMatrix m(2, 2);
    switch (p) {
        case 0: m(0, 0) = 1; m(1, 1) = 1; break;
        case 1: m(0, 1) = 1; m(1, 0) = 1; break;
        case 2: m(0, 1) = -I_; m(1, 0) = I_; break;
        case 3: m(0, 0) = 1; m(1, 1) = -1; break;
        default: throw QulacsError("Pauli id must be 0..3");
    }
    return m;
}

// Kronecker product in qulacs target order: bit i of the index <-> factor i.
ComplexMatrix pauli_product(const std::vector<unsigned>& paulis) {
    const size_t k = paulis.size(), d = size_t(1) << k;
    ComplexMatrix m(d, d);
    for (size_t y = 0; y < d; ++y)
        for (size_t x = 0; x < d; ++x) {
            CPPCTYPE v = 1.0;
            for (size_t i = 0; i < k; ++i) v *= pauli_matrix(paulis[i])((y >> i) & 1, (x >> i) & 1);
            m(y, x) = v;
        }
    return m;
}

// Operator of `m` (over `targets`) embedded into the ordered qubit list `onto`.
ComplexMatrix embed(const ComplexMatrix& m, const std::vector<unsigned>& targets, const std::vector<unsigned>& onto) {
    const size_t k = onto.size(), d = size_t(1) << k;
    std::vector<unsigned> pos(targets.size());
    for (size_t i = 0; i < targets.size(); ++i) {
        auto it = std::find(onto.begin(), onto.end(), targets[i]);
        if (it == onto.end()) throw QulacsError("embed: target not in qubit list");
        pos[i] = (unsigned)(it - onto.begin());
    }
    size_t tmask = 0;
    for (unsigned p : pos) tmask |= size_t(1) << p;
    ComplexMatrix out(d, d);
    for (size_t col = 0; col < d; ++col) {
        size_t x = 0;
        for (size_t i = 0; i < pos.size(); ++i) x |= ((col >> pos[i]) & 1) << i;
        const size_t base = col & ~tmask;
        for (size_t y = 0; y < m.rows; ++y) {
            size_t row = base;
            for (size_t i = 0; i < pos.size(); ++i) row |= ((y >> i) & 1) << pos[i];
            out(row, col) += m(y, x);
        }
    }
    return out;
}

// S = sum_i w_i K_i (x) conj(K_i); member = row bits * 2^k + column bits.
void accumulate_superop(ComplexMatrix& S, const ComplexMatrix& K, double w) {
    const size_t d = K.rows;
    for (size_t oy = 0; oy < d; ++oy)
        for (size_t ox = 0; ox < d; ++ox)
            for (size_t iy = 0; iy < d; ++iy)
                for (size_t ix = 0; ix < d; ++ix)
                    S(oy * d + ox, iy * d + ix) += w * K(oy, iy) * std::conj(K(ox, ix));
}

std::vector<unsigned> sorted_union(const std::vector<std::vector<unsigned>>& lists) {
    std::vector<unsigned> u;
    for (const auto& l : lists) u.insert(u.end(), l.begin(), l.end());
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    return u;
}

// X -> commutes with X, Y -> Y, Z -> Z, I -> none (src/cppsim/gate_named_pauli.hpp:27-43)
std::vector<unsigned> pauli_commutation(const std::vector<unsigned>& paulis) {
    std::vector<unsigned> f(paulis.size(), 0);
    for (size_t i = 0; i < paulis.size(); ++i)
        f[i] = paulis[i] == 1 ? QuantumGateBase::COMMUTE_X
             : paulis[i] == 2 ? QuantumGateBase::COMMUTE_Y
             : paulis[i] == 3 ? QuantumGateBase::COMMUTE_Z
                              : 0u;
    return f;
}
// Flags of a gate built from `gates` over `targets`: on each qubit the AND of every
// gate's flags there (a control counts as Z, an unused qubit as all three), which is
// the merge rule of src/cppsim/gate_merge.cpp:19-150.
std::vector<unsigned> merged_commutation(const std::vector<const QuantumGateBase*>& gates,
                                         const std::vector<unsigned>& targets) {
    std::vector<unsigned> f(targets.size(), 7u);
    for (size_t i = 0; i < targets.size(); ++i)
        for (const auto* g : gates) f[i] &= g->commutation_of(targets[i]);
    return f;
}

// Matrix of g over its targets followed by its controls (control qubits become
// targets; the matrix is the identity where a control value does not match).
ComplexMatrix with_controls(const QuantumGateBase* g, std::vector<unsigned>* on) {
    const auto& tl = g->get_target_index_list();
    const auto& cl = g->get_control_index_list();
    const auto& cv = g->get_control_value_list();
    ComplexMatrix m = g->get_matrix();
    *on = tl;
    if (cl.empty()) return m;
    on->insert(on->end(), cl.begin(), cl.end());
    const size_t nt = tl.size(), dt = size_t(1) << nt, dd = size_t(1) << on->size();
    ComplexMatrix full = ComplexMatrix::identity(dd);
    for (size_t col = 0; col < dd; ++col) {
        bool active = true;
        for (size_t c = 0; c < cl.size(); ++c) active = active && (((col >> (nt + c)) & 1) == cv[c]);
        if (!active) continue;
        const size_t high = col & ~(dt - 1), x = col & (dt - 1);
        for (size_t y = 0; y < dt; ++y) full(high | y, col) = m(y, x);
    }
    return full;
}

QuantumStateMI355* as_sv(QuantumStateBase* s) { return dynamic_cast<QuantumStateMI355*>(s); }
DensityMatrixMI355* as_dm(QuantumStateBase* s) { return dynamic_cast<DensityMatrixMI355*>(s); }

void check_targets(const QuantumStateBase* s, const std::vector<unsigned>& qs, const char* what) {
    for (unsigned q : qs)
        if (q >= s->qubit_count)
            throw QulacsError(std::string(what) + ": qubit index " + std::to_string(q) + " >= qubit count " +
                              std::to_string(s->qubit_count));
}
}  // namespace

void check_abi(const char* where) {
    char buf[512];
    const int code = qm_get_last_error(buf, sizeof buf);
    if (code != QM_OK) {
        qm_clear_last_error();
        throw QulacsError(std::string(where) + ": " + buf);
    }
}

ComplexMatrix ComplexMatrix::identity(size_t n) {
    ComplexMatrix m(n, n);
    for (size_t i = 0; i < n; ++i) m(i, i) = 1.0;
    return m;
}
ComplexMatrix ComplexMatrix::operator*(const ComplexMatrix& o) const {
    if (cols != o.rows) throw QulacsError("matrix shape mismatch");
    ComplexMatrix r(rows, o.cols);
    for (size_t i = 0; i < rows; ++i)
        for (size_t k = 0; k < cols; ++k) {
            const CPPCTYPE a = (*this)(i, k);
            if (a == CPPCTYPE(0)) continue;
            for (size_t j = 0; j < o.cols; ++j) r(i, j) += a * o(k, j);
        }
    return r;
}
ComplexMatrix ComplexMatrix::adjoint() const {
    ComplexMatrix r(cols, rows);
    for (size_t i = 0; i < rows; ++i)
        for (size_t j = 0; j < cols; ++j) r(j, i) = std::conj((*this)(i, j));
    return r;
}

Random::Random() {
    std::random_device rd;
    mt_.seed(rd());
}

// ------------------------------------------------------------------ QuantumStateBase
QuantumStateBase::QuantumStateBase(unsigned qc_, ITYPE dim_, unsigned device)
    : qubit_count(qc_), dim(dim_), device_number(device) {
    if (qc_ == 0 || qc_ > 40) throw QulacsError("qubit count must be in 1..40");
    stream_ = allocate_cuda_stream_host(1, device);
    check_abi("allocate_cuda_stream_host");
}

QuantumStateBase::~QuantumStateBase() {
    if (data_) release_quantum_state_host(data_, device_number);
    if (stream_) release_cuda_stream_host(stream_, 1, device_number);
    qm_clear_last_error();
}

void QuantumStateBase::load(const QuantumStateBase* o) {
    if (o->element_count() != element_count() || o->is_state_vector() != is_state_vector())
        throw QulacsError("load: state kind or size mismatch");
    qm_synchronize(o->get_cuda_stream(), o->device_number);  // other stream's pending work first
    copy_quantum_state_from_device_to_device(data_, o->data(), element_count(), stream_, device_number);
    check_abi("load");
    classical_register_ = o->classical_register_;
}

std::vector<CPPCTYPE> QuantumStateBase::duplicate_data_cpp() const {
    std::vector<CPPCTYPE> h(element_count());
    get_quantum_state_host(data_, h.data(), element_count(), stream_, device_number);
    check_abi("duplicate_data_cpp");
    return h;
}

void QuantumStateBase::add_state(const QuantumStateBase* o) { add_state_with_coef(1.0, o); }

void QuantumStateBase::add_state_with_coef(CPPCTYPE coef, const QuantumStateBase* o) {
    if (o->element_count() != element_count()) throw QulacsError("add_state: size mismatch");
    qm_synchronize(o->get_cuda_stream(), o->device_number);
    qm_state_add_with_coef_host(qc(coef), o->data(), data_, element_count(), stream_, device_number);
    check_abi("add_state_with_coef");
}

void QuantumStateBase::multiply_coef(CPPCTYPE coef) {
    state_multiply_host(qc(coef), data_, element_count(), stream_, device_number);
    check_abi("multiply_coef");
}

std::vector<ITYPE> QuantumStateBase::sampling(unsigned count, unsigned seed) {
    random_.set_seed(seed);  // src/cppsim/state.hpp:933-936
    return sampling(count);
}

unsigned QuantumStateBase::get_classical_value(unsigned index) {
    if (index >= classical_register_.size()) classical_register_.resize(index + 1, 0);
    return classical_register_[index];
}
void QuantumStateBase::set_classical_value(unsigned index, unsigned value) {
    if (index >= classical_register_.size()) classical_register_.resize(index + 1, 0);
    classical_register_[index] = value;
}

// ------------------------------------------------------------------ state vector
QuantumStateMI355::QuantumStateMI355(unsigned n, unsigned device) : QuantumStateBase(n, 1ULL << n, device) {
    data_ = allocate_quantum_state_host(dim, device);
    check_abi("allocate_quantum_state_host");
    set_zero_state();
}
void QuantumStateBase::set_zero_norm_state() {
    set_zero_state();  // every entry finite, so the product below is exactly zero
    multiply_coef(0.0);
}
void QuantumStateMI355::set_zero_state() {
    initialize_quantum_state_host(data_, dim, stream_, device_number);
    check_abi("set_zero_state");
}
void QuantumStateMI355::set_computational_basis(ITYPE b) {
    if (b >= dim) throw QulacsError("set_computational_basis: index out of range");
    set_computational_basis_host(b, data_, dim, stream_, device_number);
    check_abi("set_computational_basis");
}
void QuantumStateMI355::set_Haar_random_state() { set_Haar_random_state(random_.int32()); }
void QuantumStateMI355::set_Haar_random_state(unsigned seed) {
    initialize_Haar_random_state_with_seed_host(data_, dim, seed, stream_, device_number);
    check_abi("set_Haar_random_state");
}
double QuantumStateMI355::get_zero_probability(unsigned t) const {
    if (t >= qubit_count) throw QulacsError("get_zero_probability: qubit index out of range");
    const double v = M0_prob_host(t, data_, dim, stream_, device_number);
    check_abi("get_zero_probability");
    return v;
}
double QuantumStateMI355::get_marginal_probability(const std::vector<unsigned>& mv) const {
    if (mv.size() != qubit_count) throw QulacsError("get_marginal_probability: one value per qubit expected");
    std::vector<UINT> t, v;
    for (unsigned q = 0; q < qubit_count; ++q)
        if (mv[q] == 0 || mv[q] == 1) {
            t.push_back(q);
            v.push_back(mv[q]);
        }
    const double r = marginal_prob_host(t.data(), v.data(), (UINT)t.size(), data_, dim, stream_, device_number);
    check_abi("get_marginal_probability");
    return r;
}
double QuantumStateMI355::get_entropy() const {
    const double v = measurement_distribution_entropy_host(data_, dim, stream_, device_number);
    check_abi("get_entropy");
    return v;
}
double QuantumStateMI355::get_squared_norm() const {
    const double v = state_norm_squared_host(data_, dim, stream_, device_number);
    check_abi("get_squared_norm");
    return v;
}
void QuantumStateMI355::normalize(double sq) {
    normalize_host(sq, data_, dim, stream_, device_number);
    check_abi("normalize");
}
void QuantumStateMI355::multiply_elementwise_function(const std::function<CPPCTYPE(ITYPE)>& f) {
    std::vector<CPPCTYPE> d(dim);
    for (ITYPE i = 0; i < dim; ++i) d[i] = f(i);
    multi_qubit_diagonal_matrix_gate_host(qc(d.data()), data_, dim, stream_, device_number);
    check_abi("multiply_elementwise_function");
}
QuantumStateBase* QuantumStateMI355::allocate_buffer() const { return new QuantumStateMI355(qubit_count, device_number); }
QuantumStateBase* QuantumStateMI355::copy() const {
    auto* s = new QuantumStateMI355(qubit_count, device_number);
    s->load(this);
    return s;
}
void QuantumStateMI355::load(const std::vector<CPPCTYPE>& h) {
    if (h.size() != dim) throw QulacsError("load: vector size must equal 2^qubit_count");
    copy_quantum_state_from_host_to_device(data_, h.data(), dim, stream_, device_number);
    check_abi("load");
}
std::vector<ITYPE> QuantumStateMI355::sampling_with_uniforms(const std::vector<double>& u) const {
    std::vector<ITYPE> out(u.size());
    if (u.empty()) return out;
    qm_sampling_host(u.data(), (UINT)u.size(), out.data(), data_, dim, stream_, device_number);
    check_abi("sampling");
    return out;
}
std::pair<std::vector<ITYPE>, std::vector<unsigned char>> QuantumStateMI355::sampling_with_uniforms_checked(
    const std::vector<double>& u, double margin) const {
    std::vector<ITYPE> out(u.size());
    std::vector<unsigned char> flags(u.size());
    if (u.empty()) return {out, flags};
    qm_sampling_checked_host(u.data(), (UINT)u.size(), margin, out.data(), flags.data(), nullptr, data_, dim, stream_,
                             device_number);
    check_abi("sampling");
    return {out, flags};
}
std::vector<ITYPE> QuantumStateMI355::sampling(unsigned count) {
    std::vector<double> u(count);
    for (auto& x : u) x = random_.uniform();  // src/cppsim/state.hpp:924
    return sampling_with_uniforms(u);
}
std::string QuantumStateMI355::to_string() const {
    std::ostringstream os;
    os << " *** Quantum State ***\n * Qubit Count : " << qubit_count << "\n * Dimension   : " << dim
       << "\n * State vector : \n";
    const auto h = duplicate_data_cpp();
    for (ITYPE i = 0; i < std::min<ITYPE>(dim, 64); ++i) os << h[i] << "\n";
    if (dim > 64) os << "...\n";
    return os.str();
}

CPPCTYPE inner_product(const QuantumStateMI355* bra, const QuantumStateMI355* ket) {
    if (bra->dim != ket->dim) throw QulacsError("inner_product: dimension mismatch");
    qm_synchronize(bra->get_cuda_stream(), bra->device_number);
    const qm_complex z = inner_product_host(bra->data(), ket->data(), ket->dim, ket->get_cuda_stream(), ket->device_number);
    check_abi("inner_product");
    return cc(z);
}

namespace stateops {
namespace {
void wait_for(const QuantumStateBase* s) { qm_synchronize(s->get_cuda_stream(), s->device_number); }
}  // namespace
QuantumStateMI355* tensor_product(const QuantumStateMI355* l, const QuantumStateMI355* r) {
    std::unique_ptr<QuantumStateMI355> out(new QuantumStateMI355(l->qubit_count + r->qubit_count, l->device_number));
    wait_for(l);
    wait_for(r);
    state_tensor_product_host(l->data(), l->dim, r->data(), r->dim, out->data(), out->get_cuda_stream(),
                              out->device_number);
    check_abi("tensor_product");
    return out.release();
}
DensityMatrixMI355* tensor_product(const DensityMatrixMI355* l, const DensityMatrixMI355* r) {
    std::unique_ptr<DensityMatrixMI355> out(new DensityMatrixMI355(l->qubit_count + r->qubit_count, l->device_number));
    wait_for(l);
    wait_for(r);
    dm_state_tensor_product_host(l->data(), l->dim, r->data(), r->dim, out->data(), out->get_cuda_stream(),
                                 out->device_number);
    check_abi("tensor_product");
    return out.release();
}
QuantumStateMI355* permutate_qubit(const QuantumStateMI355* s, const std::vector<unsigned>& order) {
    if (order.size() != s->qubit_count) throw QulacsError("permutate_qubit: invalid qubit count");
    std::unique_ptr<QuantumStateMI355> out(new QuantumStateMI355(s->qubit_count, s->device_number));
    wait_for(s);
    state_permutate_qubit_host(order.data(), s->data(), out->data(), s->qubit_count, s->dim, out->get_cuda_stream(),
                               out->device_number);
    check_abi("permutate_qubit");
    return out.release();
}
DensityMatrixMI355* permutate_qubit(const DensityMatrixMI355* s, const std::vector<unsigned>& order) {
    if (order.size() != s->qubit_count) throw QulacsError("permutate_qubit: invalid qubit count");
    std::unique_ptr<DensityMatrixMI355> out(new DensityMatrixMI355(s->qubit_count, s->device_number));
    wait_for(s);
    dm_state_permutate_qubit_host(order.data(), s->data(), out->data(), s->qubit_count, s->dim,
                                  out->get_cuda_stream(), out->device_number);
    check_abi("permutate_qubit");
    return out.release();
}
QuantumStateMI355* drop_qubit(const QuantumStateMI355* s, const std::vector<unsigned>& target,
                              const std::vector<unsigned>& projection) {
    if (s->qubit_count <= target.size() || target.size() != projection.size())
        throw QulacsError("drop_qubit: invalid qubit count");
    std::unique_ptr<QuantumStateMI355> out(
        new QuantumStateMI355(s->qubit_count - (unsigned)target.size(), s->device_number));
    wait_for(s);
    state_drop_qubits_host(target.data(), projection.data(), (UINT)target.size(), s->data(), out->data(), s->dim,
                           out->get_cuda_stream(), out->device_number);
    check_abi("drop_qubit");
    return out.release();
}
DensityMatrixMI355* partial_trace(const QuantumStateBase* s, const std::vector<unsigned>& target) {
    if (s->qubit_count <= target.size()) throw QulacsError("partial_trace: invalid qubit count");
    std::unique_ptr<DensityMatrixMI355> out(
        new DensityMatrixMI355(s->qubit_count - (unsigned)target.size(), s->device_number));
    wait_for(s);
    if (s->is_state_vector())
        dm_state_partial_trace_from_state_vector_host(target.data(), (UINT)target.size(), s->data(), out->data(),
                                                      s->dim, out->get_cuda_stream(), out->device_number);
    else
        dm_state_partial_trace_from_density_matrix_host(target.data(), (UINT)target.size(), s->data(), out->data(),
                                                        s->dim, out->get_cuda_stream(), out->device_number);
    check_abi("partial_trace");
    return out.release();
}
QuantumStateMI355* make_superposition(CPPCTYPE c1, const QuantumStateMI355* s1, CPPCTYPE c2,
                                      const QuantumStateMI355* s2) {
    if (s1->qubit_count != s2->qubit_count) throw QulacsError("make_superposition: invalid qubit count");
    std::unique_ptr<QuantumStateMI355> out(new QuantumStateMI355(s1->qubit_count, s1->device_number));
    out->multiply_coef(0.0);  // the zero vector
    out->add_state_with_coef(c1, s1);
    out->add_state_with_coef(c2, s2);
    return out.release();
}
DensityMatrixMI355* make_mixture(CPPCTYPE p1, const QuantumStateBase* s1, CPPCTYPE p2, const QuantumStateBase* s2) {
    if (s1->qubit_count != s2->qubit_count) throw QulacsError("make_mixture: invalid qubit count");
    const unsigned n = s1->qubit_count;
    auto as_density = [&](const QuantumStateBase* s) {
        std::unique_ptr<DensityMatrixMI355> d(new DensityMatrixMI355(n, s->device_number));
        if (s->is_state_vector()) d->load_pure(static_cast<const QuantumStateMI355*>(s));
        else d->load(s);
        return d;
    };
    std::unique_ptr<DensityMatrixMI355> d1 = as_density(s1), d2 = as_density(s2);
    std::unique_ptr<DensityMatrixMI355> out(new DensityMatrixMI355(n, s1->device_number));
    out->multiply_coef(0.0);
    out->add_state_with_coef(p1, d1.get());
    out->add_state_with_coef(p2, d2.get());
    return out.release();
}
}  // namespace stateops

// ------------------------------------------------------------------ density matrix
DensityMatrixMI355::DensityMatrixMI355(unsigned n, unsigned device) : QuantumStateBase(n, 1ULL << n, device) {
    if (n > 20) throw QulacsError("DensityMatrix: at most 20 qubits");
    data_ = allocate_quantum_state_host(dim * dim, device);
    check_abi("allocate_quantum_state_host");
    set_zero_state();
}
void DensityMatrixMI355::set_zero_state() {
    dm_initialize_quantum_state_host(data_, dim, stream_, device_number);
    check_abi("set_zero_state");
}
void DensityMatrixMI355::set_computational_basis(ITYPE b) {
    if (b >= dim) throw QulacsError("set_computational_basis: index out of range");
    set_computational_basis_host(b * dim + b, data_, dim * dim, stream_, device_number);
    check_abi("set_computational_basis");
}
void DensityMatrixMI355::set_Haar_random_state() { set_Haar_random_state(random_.int32()); }
void DensityMatrixMI355::set_Haar_random_state(unsigned seed) {
    QuantumStateMI355 pure(qubit_count, device_number);
    pure.set_Haar_random_state(seed);
    load_pure(&pure);
}
void DensityMatrixMI355::load_pure(const QuantumStateMI355* p) {
    if (p->dim != dim) throw QulacsError("load: dimension mismatch");
    qm_synchronize(p->get_cuda_stream(), p->device_number);
    dm_initialize_with_pure_state_host(data_, p->data(), dim, stream_, device_number);
    check_abi("load");
}
double DensityMatrixMI355::get_zero_probability(unsigned t) const {
    if (t >= qubit_count) throw QulacsError("get_zero_probability: qubit index out of range");
    const double v = dm_M0_prob_host(t, data_, dim, stream_, device_number);
    check_abi("get_zero_probability");
    return v;
}
double DensityMatrixMI355::get_marginal_probability(const std::vector<unsigned>& mv) const {
    if (mv.size() != qubit_count) throw QulacsError("get_marginal_probability: one value per qubit expected");
    std::vector<UINT> t, v;
    for (unsigned q = 0; q < qubit_count; ++q)
        if (mv[q] == 0 || mv[q] == 1) {
            t.push_back(q);
            v.push_back(mv[q]);
        }
    const double r = dm_marginal_prob_host(t.data(), v.data(), (UINT)t.size(), data_, dim, stream_, device_number);
    check_abi("get_marginal_probability");
    return r;
}
double DensityMatrixMI355::get_entropy() const {
    const double v = dm_measurement_distribution_entropy_host(data_, dim, stream_, device_number);
    check_abi("get_entropy");
    return v;
}
double DensityMatrixMI355::get_squared_norm() const {
    const double v = dm_state_norm_squared_host(data_, dim, stream_, device_number);
    check_abi("get_squared_norm");
    return v;
}
void DensityMatrixMI355::normalize(double sq) {
    dm_normalize_host(sq, data_, dim, stream_, device_number);
    check_abi("normalize");
}
QuantumStateBase* DensityMatrixMI355::allocate_buffer() const { return new DensityMatrixMI355(qubit_count, device_number); }
QuantumStateBase* DensityMatrixMI355::copy() const {
    auto* s = new DensityMatrixMI355(qubit_count, device_number);
    s->load(this);
    return s;
}
void DensityMatrixMI355::load(const std::vector<CPPCTYPE>& h) {
    if (h.size() == dim) {
        QuantumStateMI355 pure(qubit_count, device_number);
        pure.load(h);
        load_pure(&pure);
    } else if (h.size() == dim * dim) {
        copy_quantum_state_from_host_to_device(data_, h.data(), dim * dim, stream_, device_number);
        check_abi("load");
    } else {
        throw QulacsError("load: expected 2^n (pure state) or 4^n (matrix) entries");
    }
}
std::vector<ITYPE> DensityMatrixMI355::sampling_with_uniforms(const std::vector<double>& u) const {
    std::vector<ITYPE> out(u.size());
    if (u.empty()) return out;
    dm_sampling_host(u.data(), (UINT)u.size(), out.data(), data_, dim, stream_, device_number);
    check_abi("sampling");
    return out;
}
std::pair<std::vector<ITYPE>, std::vector<unsigned char>> DensityMatrixMI355::sampling_with_uniforms_checked(
    const std::vector<double>& u, double margin) const {
    std::vector<ITYPE> out(u.size());
    std::vector<unsigned char> flags(u.size());
    if (u.empty()) return {out, flags};
    dm_sampling_checked_host(u.data(), (UINT)u.size(), margin, out.data(), flags.data(), nullptr, data_, dim, stream_,
                             device_number);
    check_abi("sampling");
    return {out, flags};
}
std::vector<ITYPE> DensityMatrixMI355::sampling(unsigned count) {
    std::vector<double> u(count);
    for (auto& x : u) x = random_.uniform();
    return sampling_with_uniforms(u);
}
std::string DensityMatrixMI355::to_string() const {
    std::ostringstream os;
    os << " *** Density Matrix ***\n * Qubit Count : " << qubit_count << "\n * Dimension   : " << dim
       << "\n * Density matrix : \n";
    const auto h = duplicate_data_cpp();
    const ITYPE show = std::min<ITYPE>(dim, 16);
    for (ITYPE y = 0; y < show; ++y) {
        for (ITYPE x = 0; x < show; ++x) os << h[y * dim + x] << " ";
        os << "\n";
    }
    return os.str();
}

// ------------------------------------------------------------------ gates
QuantumGateBase* QuantumGateBase::get_inverse() const {
    auto* g = new QuantumGateMatrix(targets_, get_matrix().adjoint(), controls_, cvals_, name_ + "^dagger");
    g->set_target_commutation(tcomm_);
    return g;
}

std::vector<std::pair<unsigned, unsigned>> QuantumGateBase::get_control_index_value_list() const {
    std::vector<std::pair<unsigned, unsigned>> r;
    for (size_t i = 0; i < controls_.size(); ++i) r.emplace_back(controls_[i], cvals_[i]);
    return r;
}
unsigned QuantumGateBase::commutation_of(unsigned q) const {
    for (size_t i = 0; i < targets_.size(); ++i)
        if (targets_[i] == q) return get_target_commutation(i);
    for (unsigned c : controls_)
        if (c == q) return COMMUTE_Z;
    return COMMUTE_X | COMMUTE_Y | COMMUTE_Z;
}
bool QuantumGateBase::is_diagonal() const {
    for (size_t i = 0; i < targets_.size(); ++i)
        if (!(get_target_commutation(i) & COMMUTE_Z)) return false;
    return true;
}
bool QuantumGateBase::is_commute(const QuantumGateBase* o) const {
    for (size_t i = 0; i < targets_.size(); ++i) {
        const unsigned f = get_target_commutation(i);
        for (size_t j = 0; j < o->targets_.size(); ++j)
            if (o->targets_[j] == targets_[i] && !(f & o->get_target_commutation(j))) return false;
        for (unsigned c : o->controls_)
            if (c == targets_[i] && !(f & COMMUTE_Z)) return false;
    }
    for (unsigned c : controls_)
        for (size_t j = 0; j < o->targets_.size(); ++j)
            if (o->targets_[j] == c && !(o->get_target_commutation(j) & COMMUTE_Z)) return false;
    return true;
}

QuantumGateBase* QuantumGateBase::create_gate_whose_qubit_indices_are_replaced(const std::vector<unsigned>& t,
                                                                              const std::vector<unsigned>& c) const {
    if (t.size() != targets_.size()) throw QulacsError("create_gate_whose_qubit_indices_are_replaced: target count differs");
    if (c.size() != controls_.size())
        throw QulacsError("create_gate_whose_qubit_indices_are_replaced: control count differs");
    QuantumGateBase* g = copy();
    g->targets_ = t;
    g->controls_ = c;
    return g;
}

// src/cppsim/gate.cpp:100-126
std::string QuantumGateBase::to_string() const {
    std::ostringstream os;
    os << " *** gate info *** \n * gate name : " << get_name() << "\n * target    : \n";
    for (size_t i = 0; i < targets_.size(); ++i) {
        const unsigned f = get_target_commutation(i);
        os << " " << targets_[i] << " : commute " << (f & COMMUTE_X ? "X" : " ") << " " << (f & COMMUTE_Y ? "Y" : " ")
           << " " << (f & COMMUTE_Z ? "Z" : " ") << " \n";
    }
    os << " * control   : \n";
    for (size_t i = 0; i < controls_.size(); ++i) os << " " << controls_[i] << " : value " << cvals_[i] << "\n";
    auto yn = [](bool b) { return b ? "yes" : "no"; };
    os << " * Pauli     : " << yn(is_Pauli()) << "\n * Clifford  : " << yn(is_Clifford())
       << "\n * Gaussian  : " << yn(is_Gaussian()) << "\n * Parametric: " << yn(is_parametric())
       << "\n * Diagonal  : " << yn(is_diagonal()) << "\n";
    return os.str();
}

QuantumGateMatrix::QuantumGateMatrix(std::vector<unsigned> t, ComplexMatrix m, std::vector<unsigned> c,
                                     std::vector<unsigned> v, std::string name)
    : matrix_(std::move(m)) {
    targets_ = std::move(t);
    controls_ = std::move(c);
    cvals_ = std::move(v);
    name_ = std::move(name);
    if (cvals_.size() != controls_.size()) throw QulacsError("one control value per control qubit");
    const size_t d = size_t(1) << targets_.size();
    if (matrix_.rows != d || matrix_.cols != d) throw QulacsError("matrix size must be 2^k x 2^k for k targets");
}

void QuantumGateMatrix::add_control_qubit(unsigned q, unsigned v) {
    if (std::find(targets_.begin(), targets_.end(), q) != targets_.end() ||
        std::find(controls_.begin(), controls_.end(), q) != controls_.end())
        throw QulacsError("add_control_qubit: qubit already used by the gate");
    controls_.push_back(q);
    cvals_.push_back(v);
    kind = GENERIC;
    property_ &= ~(PROPERTY_PAULI | PROPERTY_GAUSSIAN);
}
void QuantumGateMatrix::multiply_scalar(CPPCTYPE value) {
    for (auto& z : matrix_.data) z *= value;
    kind = GENERIC;
}

QuantumGateBase* QuantumGateMatrix::get_inverse() const {
    auto* g = new QuantumGateMatrix(targets_, matrix_.adjoint(), controls_, cvals_, name_);
    g->set_target_commutation(tcomm_);
    if (kind != GENERIC) g->set_gate_property(property_);  // S <-> Sdag etc. share their class
    switch (kind) {
        case RX: case RY: case RZ: g->kind = kind; g->angle = -angle; break;
        case X: case Y: case Z: case H: case CNOT: case CZ: case SWAP: case IDENTITY: g->kind = kind; break;
        case S: g->kind = SDAG; break;
        case SDAG: g->kind = S; break;
        case T: g->kind = TDAG; break;
        case TDAG: g->kind = T; break;
        case SQRTX: g->kind = SQRTXDAG; break;
        case SQRTXDAG: g->kind = SQRTX; break;
        case SQRTY: g->kind = SQRTYDAG; break;
        case SQRTYDAG: g->kind = SQRTY; break;
        default: break;
    }
    return g;
}

void QuantumGateMatrix::update_quantum_state(QuantumStateBase* state) {
    check_targets(state, targets_, "update_quantum_state");
    check_targets(state, controls_, "update_quantum_state");
    if (auto* s = as_sv(state)) update_sv(s);
    else if (auto* d = as_dm(state)) update_dm(d);
    else throw QulacsError("unsupported state type");
}

void QuantumGateMatrix::update_sv(QuantumStateMI355* s) {
    void* p = s->data();
    const ITYPE dim = s->dim;
    void* st = s->get_cuda_stream();
    const unsigned dev = s->device_number;
    const unsigned t0 = targets_.empty() ? 0 : targets_[0];
    switch (kind) {
        case IDENTITY: return;
        case X: X_gate_host(t0, p, dim, st, dev); break;
        case Y: Y_gate_host(t0, p, dim, st, dev); break;
        case Z: Z_gate_host(t0, p, dim, st, dev); break;
        case H: H_gate_host(t0, p, dim, st, dev); break;
        case S: S_gate_host(t0, p, dim, st, dev); break;
        case SDAG: Sdag_gate_host(t0, p, dim, st, dev); break;
        case T: T_gate_host(t0, p, dim, st, dev); break;
        case TDAG: Tdag_gate_host(t0, p, dim, st, dev); break;
        case SQRTX: sqrtX_gate_host(t0, p, dim, st, dev); break;
        case SQRTXDAG: sqrtXdag_gate_host(t0, p, dim, st, dev); break;
        case SQRTY: sqrtY_gate_host(t0, p, dim, st, dev); break;
        case SQRTYDAG: sqrtYdag_gate_host(t0, p, dim, st, dev); break;
        case P0: P0_gate_host(t0, p, dim, st, dev); break;
        case P1: P1_gate_host(t0, p, dim, st, dev); break;
        case RX: RX_gate_host(t0, angle, p, dim, st, dev); break;
        case RY: RY_gate_host(t0, angle, p, dim, st, dev); break;
        case RZ: RZ_gate_host(t0, angle, p, dim, st, dev); break;
        case CNOT: CNOT_gate_host(controls_[0], t0, p, dim, st, dev); break;
        case CZ: CZ_gate_host(controls_[0], t0, p, dim, st, dev); break;
        case SWAP: SWAP_gate_host(targets_[0], targets_[1], p, dim, st, dev); break;
        case GENERIC: {
            const qm_complex* m = qc(matrix_.data.data());
            const UINT nt = (UINT)targets_.size(), nc = (UINT)controls_.size();
            if (nc == 0) {
                if (nt == 1) single_qubit_dense_matrix_gate_host(t0, m, p, dim, st, dev);
                else if (nt == 2) double_qubit_dense_matrix_gate_host(targets_[0], targets_[1], m, p, dim, st, dev);
                else multi_qubit_dense_matrix_gate_host(targets_.data(), nt, m, p, dim, st, dev);
            } else if (nt == 1 && nc == 1) {
                single_qubit_control_single_qubit_dense_matrix_gate_host(controls_[0], cvals_[0], t0, m, p, dim, st, dev);
            } else if (nt == 1) {
                multi_qubit_control_single_qubit_dense_matrix_gate_host(controls_.data(), cvals_.data(), nc, t0, m, p,
                                                                        dim, st, dev);
            } else if (nc == 1) {
                single_qubit_control_multi_qubit_dense_matrix_gate_host(controls_[0], cvals_[0], targets_.data(), nt,
                                                                        m, p, dim, st, dev);
            } else {
                multi_qubit_control_multi_qubit_dense_matrix_gate_host(controls_.data(), cvals_.data(), nc,
                                                                       targets_.data(), nt, m, p, dim, st, dev);
            }
            break;
        }
    }
    check_abi(name_.c_str());
}

void QuantumGateMatrix::update_dm(DensityMatrixMI355* s) {
    if (kind == IDENTITY) return;
    void* p = s->data();
    const qm_complex* m = qc(matrix_.data.data());
    if (controls_.empty()) {
        if (targets_.size() == 1)
            dm_single_qubit_dense_matrix_gate_host(targets_[0], m, p, s->dim, s->get_cuda_stream(), s->device_number);
        else
            dm_multi_qubit_dense_matrix_gate_host(targets_.data(), (UINT)targets_.size(), m, p, s->dim,
                                                  s->get_cuda_stream(), s->device_number);
    } else {
        dm_multi_qubit_control_multi_qubit_dense_matrix_gate_host(controls_.data(), cvals_.data(), (UINT)controls_.size(),
                                                                  targets_.data(), (UINT)targets_.size(), m, p, s->dim,
                                                                  s->get_cuda_stream(), s->device_number);
    }
    check_abi(name_.c_str());
}

QuantumGateDiagonalMatrix::QuantumGateDiagonalMatrix(std::vector<unsigned> t, std::vector<CPPCTYPE> d,
                                                     std::vector<unsigned> c, std::vector<unsigned> v)
    : diag_(std::move(d)) {
    targets_ = std::move(t);
    controls_ = std::move(c);
    cvals_ = std::move(v);
    name_ = "DiagonalMatrix";
    if (diag_.size() != (size_t(1) << targets_.size())) throw QulacsError("diagonal must have 2^k entries");
}
ComplexMatrix QuantumGateDiagonalMatrix::get_matrix() const {
    ComplexMatrix m(diag_.size(), diag_.size());
    for (size_t i = 0; i < diag_.size(); ++i) m(i, i) = diag_[i];
    return m;
}
void QuantumGateDiagonalMatrix::update_quantum_state(QuantumStateBase* state) {
    check_targets(state, targets_, "DiagonalMatrix");
    check_targets(state, controls_, "DiagonalMatrix");
    if (auto* s = as_sv(state)) {
        qm_multi_qubit_control_multi_qubit_diagonal_matrix_gate_host(
            controls_.data(), cvals_.data(), (UINT)controls_.size(), targets_.data(), (UINT)targets_.size(),
            qc(diag_.data()), s->data(), s->dim, s->get_cuda_stream(), s->device_number);
        check_abi("DiagonalMatrix");
    } else {
        QuantumGateMatrix(targets_, get_matrix(), controls_, cvals_).update_quantum_state(state);
    }
}

ClsPauliGate::ClsPauliGate(std::vector<unsigned> t, std::vector<unsigned> p, bool rotation, double angle)
    : paulis_(std::move(p)), rotation_(rotation), angle_(angle) {
    targets_ = std::move(t);
    name_ = rotation ? "Pauli-rotation" : "Pauli";
    if (targets_.size() != paulis_.size()) throw QulacsError("one Pauli id per target");
    for (unsigned x : paulis_)
        if (x > 3) throw QulacsError("Pauli id must be 0..3");
    tcomm_ = pauli_commutation(paulis_);
}
ComplexMatrix ClsPauliGate::get_matrix() const {
    ComplexMatrix P = pauli_product(paulis_);
    if (!rotation_) return P;
    ComplexMatrix m = ComplexMatrix::identity(P.rows);
    for (size_t i = 0; i < m.data.size(); ++i) m.data[i] = std::cos(angle_ / 2) * m.data[i] + I_ * std::sin(angle_ / 2) * P.data[i];
    return m;
}
QuantumGateBase* ClsPauliGate::get_inverse() const {
    if (rotation_) return new ClsPauliGate(targets_, paulis_, true, -angle_);
    return new ClsPauliGate(targets_, paulis_, false, 0.0);
}
void ClsPauliGate::update_quantum_state(QuantumStateBase* state) {
    check_targets(state, targets_, name_.c_str());
    if (auto* s = as_sv(state)) {
        if (rotation_)
            multi_qubit_Pauli_rotation_gate_partial_list_host(targets_.data(), paulis_.data(), (UINT)targets_.size(),
                                                              angle_, s->data(), s->dim, s->get_cuda_stream(),
                                                              s->device_number);
        else
            multi_qubit_Pauli_gate_partial_list_host(targets_.data(), paulis_.data(), (UINT)targets_.size(), s->data(),
                                                     s->dim, s->get_cuda_stream(), s->device_number);
        check_abi(name_.c_str());
    } else {
        QuantumGateMatrix(targets_, get_matrix()).update_quantum_state(state);
    }
}

QuantumGate_Probabilistic::QuantumGate_Probabilistic(std::vector<double> dist, std::vector<QuantumGateBase*> gates)
    : dist_(std::move(dist)) {
    if (dist_.size() != gates.size()) throw QulacsError("Probabilistic: one probability per gate");
    cumulative_.push_back(0.0);
    for (double p : dist_) cumulative_.push_back(cumulative_.back() + p);
    if (cumulative_.back() > 1.0 + 1e-12) throw QulacsError("Probabilistic: probabilities sum above 1");
    std::vector<std::vector<unsigned>> lists;
    for (auto* g : gates) {
        gates_.push_back(g->copy());
        lists.push_back(g->get_target_index_list());
        lists.push_back(g->get_control_index_list());
    }
    targets_ = sorted_union(lists);
    name_ = "Probabilistic";
    std::vector<const QuantumGateBase*> cg(gates_.begin(), gates_.end());
    tcomm_ = merged_commutation(cg, targets_);
}
void QuantumGate_Probabilistic::optimize_ProbablisticGate() {
    std::vector<std::pair<double, size_t>> order;
    for (size_t i = 0; i < dist_.size(); ++i) order.emplace_back(dist_[i], i);
    std::sort(order.rbegin(), order.rend());  // descending probability, ties by descending index
    std::vector<QuantumGateBase*> gates;
    for (size_t i = 0; i < order.size(); ++i) {
        dist_[i] = order[i].first;
        gates.push_back(gates_[order[i].second]);
    }
    gates_ = gates;
    cumulative_.assign(1, 0.0);
    for (double p : dist_) cumulative_.push_back(cumulative_.back() + p);
}
QuantumGate_Probabilistic::QuantumGate_Probabilistic(const QuantumGate_Probabilistic& o)
    : QuantumGateBase(o),
      dist_(o.dist_),
      cumulative_(o.cumulative_),
      random_(o.random_),
      instrument_(o.instrument_),
      register_(o.register_) {
    gates_.clear();
    for (auto* g : o.gates_) gates_.push_back(g->copy());
}
QuantumGate_Probabilistic::~QuantumGate_Probabilistic() {
    for (auto* g : gates_) delete g;
}
ComplexMatrix QuantumGate_Probabilistic::get_matrix() const {
    throw QulacsError("Probabilistic gate has no single matrix; use superoperator()");
}
ComplexMatrix QuantumGate_Probabilistic::superoperator() const {
    const size_t d = size_t(1) << targets_.size();
    ComplexMatrix S(d * d, d * d);
    const double rest = 1.0 - cumulative_.back();
    if (rest != 0.0) accumulate_superop(S, ComplexMatrix::identity(d), rest);
    for (size_t i = 0; i < gates_.size(); ++i) {
        const QuantumGateBase* g = gates_[i];
        std::vector<unsigned> on;
        const ComplexMatrix m = with_controls(g, &on);
        accumulate_superop(S, embed(m, on, targets_), dist_[i]);
    }
    return S;
}
void QuantumGate_Probabilistic::update_quantum_state(QuantumStateBase* state) {
    if (state->is_state_vector()) {
        // src/cppsim/gate_general.hpp:253-264: draw one branch
        const double r = random_.uniform();
        const size_t idx = (size_t)(std::upper_bound(cumulative_.begin(), cumulative_.end(), r) - cumulative_.begin()) - 1;
        if (idx < gates_.size()) gates_[idx]->update_quantum_state(state);
        if (instrument_) state->set_classical_value(register_, (unsigned)idx);
        return;
    }
    auto* d = as_dm(state);
    check_targets(state, targets_, "Probabilistic");
    if (targets_.size() <= 3) {
        const ComplexMatrix S = superoperator();
        dm_superoperator_host(targets_.data(), (UINT)targets_.size(), qc(S.data.data()), d->data(), d->dim,
                              d->get_cuda_stream(), d->device_number);
        check_abi("Probabilistic");
        return;
    }
    // wide mixtures: the reference's accumulate-over-copies form
    std::unique_ptr<QuantumStateBase> org(d->copy()), tmp(d->allocate_buffer());
    d->multiply_coef(1.0 - cumulative_.back());
    for (size_t i = 0; i < gates_.size(); ++i) {
        tmp->load(org.get());
        gates_[i]->update_quantum_state(tmp.get());
        d->add_state_with_coef(dist_[i], tmp.get());
    }
}

QuantumGate_CPTP::QuantumGate_CPTP(const std::vector<const QuantumGateBase*>& kraus) {
    std::vector<std::vector<unsigned>> lists;
    for (const auto* k : kraus) {
        gates_.push_back(k->copy());
        lists.push_back(k->get_target_index_list());
        lists.push_back(k->get_control_index_list());
    }
    targets_ = sorted_union(lists);
    name_ = "CPTP";
    tcomm_ = merged_commutation(std::vector<const QuantumGateBase*>(gates_.begin(), gates_.end()), targets_);
}
QuantumGate_CPTP::QuantumGate_CPTP(const QuantumGate_CPTP& o)
    : QuantumGateBase(o),
      random_(o.random_),
      instrument_(o.instrument_),
      register_(o.register_),
      state_normalize_(o.state_normalize_),
      probability_normalize_(o.probability_normalize_),
      assign_zero_(o.assign_zero_) {
    gates_.clear();
    for (auto* g : o.gates_) gates_.push_back(g->copy());
}
QuantumGate_CPTP::~QuantumGate_CPTP() {
    for (auto* g : gates_) delete g;
}
ComplexMatrix QuantumGate_CPTP::get_matrix() const { throw QulacsError(name_ + " map has no single matrix"); }
ComplexMatrix QuantumGate_CPTP::superoperator() const {
    const size_t d = size_t(1) << targets_.size();
    ComplexMatrix S(d * d, d * d);
    for (const auto* k : gates_) {
        if (!k->has_matrix()) throw QulacsError(name_ + ": a Kraus operator has no matrix");
        std::vector<unsigned> on;
        const ComplexMatrix m = with_controls(k, &on);
        accumulate_superop(S, embed(m, on, targets_), 1.0);
    }
    return S;
}
void QuantumGate_CPTP::update_quantum_state(QuantumStateBase* state) {
    check_targets(state, targets_, name_.c_str());
    if (auto* d = as_dm(state)) {
        bool matrices = targets_.size() <= 3;
        for (const auto* k : gates_) matrices = matrices && k->has_matrix();
        if (matrices) {
            const ComplexMatrix S = superoperator();
            dm_superoperator_host(targets_.data(), (UINT)targets_.size(), qc(S.data.data()), d->data(), d->dim,
                                  d->get_cuda_stream(), d->device_number);
            check_abi(name_.c_str());
            return;
        }
        // rho <- sum_i K_i rho K_i^dagger on copies (src/cppsim/gate_general.hpp:462-478)
        std::unique_ptr<QuantumStateBase> org(d->copy()), tmp(d->allocate_buffer());
        d->multiply_coef(0.0);
        for (auto* k : gates_) {
            tmp->load(org.get());
            k->update_quantum_state(tmp.get());
            d->add_state(tmp.get());
        }
        return;
    }
    // state vector (src/cppsim/gate_general.hpp:433-461 and 600-640): branch i is
    // taken when r * (sum of branch probabilities, or 1) first falls below the running
    // sum of ||K_i psi||^2 / ||psi||^2
    const double r = random_.uniform();
    const double org = state->get_squared_norm();
    std::unique_ptr<QuantumStateBase> buf(state->allocate_buffer());
    double total = 1.0;
    if (probability_normalize_) {
        total = 0.0;
        for (auto* k : gates_) {
            buf->load(state);
            k->update_quantum_state(buf.get());
            total += buf->get_squared_norm() / org;
        }
    }
    double sum = 0.0;
    unsigned index = 0;
    bool chosen = false;
    for (auto* k : gates_) {
        buf->load(state);
        k->update_quantum_state(buf.get());
        const double rel = buf->get_squared_norm() / org;
        sum += rel;
        if (r * total < sum) {
            state->load(buf.get());
            if (state_normalize_) state->normalize(rel);
            chosen = true;
            break;
        }
        ++index;
    }
    if (!chosen && assign_zero_) state->multiply_coef(0.0);
    if (instrument_) state->set_classical_value(register_, index);
}

QuantumGate_CP::QuantumGate_CP(const std::vector<const QuantumGateBase*>& kraus, bool state_normalize,
                               bool probability_normalize, bool assign_zero_if_not_matched)
    : QuantumGate_CPTP(kraus) {
    name_ = "CP";
    state_normalize_ = state_normalize;
    probability_normalize_ = probability_normalize;
    assign_zero_ = assign_zero_if_not_matched;
}

QuantumGate_LinearCombination::QuantumGate_LinearCombination(const std::vector<CPPCTYPE>& coefs,
                                                             const std::vector<const QuantumGateBase*>& gates)
    : coefs_(coefs) {
    if (coefs.size() != gates.size()) throw QulacsError("LinearCombination: one coefficient per gate");
    std::vector<std::vector<unsigned>> lists;
    for (const auto* g : gates) {
        gates_.push_back(g->copy());
        lists.push_back(g->get_target_index_list());
        lists.push_back(g->get_control_index_list());
    }
    targets_ = sorted_union(lists);
    name_ = "LinearCombination";
    tcomm_ = merged_commutation(std::vector<const QuantumGateBase*>(gates_.begin(), gates_.end()), targets_);
}
QuantumGate_LinearCombination::QuantumGate_LinearCombination(const QuantumGate_LinearCombination& o)
    : QuantumGateBase(o), coefs_(o.coefs_) {
    gates_.clear();
    for (auto* g : o.gates_) gates_.push_back(g->copy());
}
QuantumGate_LinearCombination::~QuantumGate_LinearCombination() {
    for (auto* g : gates_) delete g;
}
ComplexMatrix QuantumGate_LinearCombination::get_matrix() const {
    const size_t d = size_t(1) << targets_.size();
    ComplexMatrix m(d, d);
    for (size_t i = 0; i < gates_.size(); ++i) {
        std::vector<unsigned> on;
        const ComplexMatrix e = embed(with_controls(gates_[i], &on), on, targets_);
        for (size_t k = 0; k < m.data.size(); ++k) m.data[k] += coefs_[i] * e.data[k];
    }
    return m;
}
void QuantumGate_LinearCombination::update_quantum_state(QuantumStateBase* state) {
    if (!state->is_state_vector()) throw QulacsError("LinearCombination: density matrices are not supported");
    check_targets(state, targets_, "LinearCombination");
    std::unique_ptr<QuantumStateBase> acc(state->allocate_buffer()), tmp(state->allocate_buffer());
    acc->multiply_coef(0.0);
    for (size_t i = 0; i < gates_.size(); ++i) {
        tmp->load(state);
        gates_[i]->update_quantum_state(tmp.get());
        acc->add_state_with_coef(coefs_[i], tmp.get());
    }
    state->load(acc.get());
}

ClsStateReflectionGate::ClsStateReflectionGate(const QuantumStateMI355* r)
    : r_(new QuantumStateMI355(r->qubit_count, r->device_number)) {
    r_->load(r);
    for (unsigned q = 0; q < r->qubit_count; ++q) targets_.push_back(q);
    name_ = "Reflection";
}
ClsStateReflectionGate::ClsStateReflectionGate(const ClsStateReflectionGate& o)
    : QuantumGateBase(o), r_(new QuantumStateMI355(o.r_->qubit_count, o.r_->device_number)) {
    r_->load(o.r_.get());
}
ComplexMatrix ClsStateReflectionGate::get_matrix() const { throw QulacsError("Reflection gate has no matrix"); }
void ClsStateReflectionGate::update_quantum_state(QuantumStateBase* state) {
    auto* s = as_sv(state);
    if (!s) throw QulacsError("Reflection: density matrices are not supported");
    if (s->qubit_count != r_->qubit_count) throw QulacsError("Reflection: qubit count must equal the reflection state's");
    if (s->device_number != r_->device_number) throw QulacsError("Reflection: states on different devices");
    const CPPCTYPE c = inner_product(r_.get(), s);
    s->multiply_coef(-1.0);
    s->add_state_with_coef(2.0 * c, r_.get());
}

namespace {
void adaptive_targets(const QuantumGateBase* g, std::vector<unsigned>* t) {
    *t = sorted_union({g->get_target_index_list(), g->get_control_index_list()});
}
}  // namespace
QuantumGate_Adaptive::QuantumGate_Adaptive(const QuantumGateBase* g, Condition c) : gate_(g->copy()), cond_(std::move(c)) {
    adaptive_targets(g, &targets_);
    tcomm_ = merged_commutation({g}, targets_);
    name_ = "Adaptive";
}
QuantumGate_Adaptive::QuantumGate_Adaptive(const QuantumGateBase* g, ConditionWithId c, unsigned id)
    : gate_(g->copy()), cond_id_(std::move(c)), id_((int)id) {
    adaptive_targets(g, &targets_);
    tcomm_ = merged_commutation({g}, targets_);
    name_ = "Adaptive";
}
QuantumGate_Adaptive::QuantumGate_Adaptive(const QuantumGate_Adaptive& o)
    : QuantumGateBase(o), gate_(o.gate_->copy()), cond_(o.cond_), cond_id_(o.cond_id_), id_(o.id_) {}
ComplexMatrix QuantumGate_Adaptive::get_matrix() const { throw QulacsError("Adaptive gate has no matrix"); }
void QuantumGate_Adaptive::update_quantum_state(QuantumStateBase* state) {
    const bool go = id_ < 0 ? cond_(state->get_classical_register()) : cond_id_(state->get_classical_register(), (unsigned)id_);
    if (go) gate_->update_quantum_state(state);
}

ClsNpairQubitGate::ClsNpairQubitGate(unsigned t1, unsigned t2, unsigned b) : block_(b) {
    if (b == 0) throw QulacsError("FusedSWAP: block size must be positive");
    if (std::min(t1, t2) + b > std::max(t1, t2)) throw QulacsError("FusedSWAP: the two blocks overlap");
    for (unsigned i = 0; i < b; ++i) targets_.push_back(t1 + i);
    for (unsigned i = 0; i < b; ++i) targets_.push_back(t2 + i);
    tcomm_.assign(targets_.size(), 0u);
    property_ = PROPERTY_CLIFFORD;
    name_ = "FusedSWAP";
}
ComplexMatrix ClsNpairQubitGate::get_matrix() const {
    if (2 * block_ > 12) throw QulacsError("FusedSWAP: matrix above 12 qubits");
    const size_t p = size_t(1) << block_, d = p * p;
    ComplexMatrix m(d, d);
    for (size_t i = 0; i < p; ++i)
        for (size_t j = 0; j < p; ++j) m(i * p + j, i + j * p) = 1.0;
    return m;
}
void ClsNpairQubitGate::update_quantum_state(QuantumStateBase* state) {
    check_targets(state, targets_, "FusedSWAP");
    for (unsigned i = 0; i < block_; ++i) {
        std::unique_ptr<QuantumGateMatrix> sw(gate::SWAP(targets_[i], targets_[block_ + i]));
        sw->update_quantum_state(state);
    }
}

namespace {
// old qubit -> new qubit from a gate's (targets, controls) and their replacements
std::vector<std::pair<unsigned, unsigned>> qubit_map(const QuantumGateBase& g, const std::vector<unsigned>& t,
                                                     const std::vector<unsigned>& c) {
    const auto &gt = g.get_target_index_list(), &gc = g.get_control_index_list();
    if (t.size() != gt.size() || c.size() != gc.size())
        throw QulacsError("create_gate_whose_qubit_indices_are_replaced: qubit count differs");
    std::vector<std::pair<unsigned, unsigned>> m;
    for (size_t i = 0; i < gt.size(); ++i) m.emplace_back(gt[i], t[i]);
    for (size_t i = 0; i < gc.size(); ++i) m.emplace_back(gc[i], c[i]);
    return m;
}
std::vector<unsigned> map_qubits(const std::vector<unsigned>& l, const std::vector<std::pair<unsigned, unsigned>>& m) {
    std::vector<unsigned> out;
    for (unsigned q : l) {
        auto it = std::find_if(m.begin(), m.end(), [q](const std::pair<unsigned, unsigned>& e) { return e.first == q; });
        if (it == m.end()) throw QulacsError("create_gate_whose_qubit_indices_are_replaced: unmapped qubit");
        out.push_back(it->second);
    }
    return out;
}
QuantumGateBase* remap_gate(const QuantumGateBase* g, const std::vector<std::pair<unsigned, unsigned>>& m) {
    return g->create_gate_whose_qubit_indices_are_replaced(map_qubits(g->get_target_index_list(), m),
                                                           map_qubits(g->get_control_index_list(), m));
}
}  // namespace

QuantumGateBase* QuantumGate_Probabilistic::create_gate_whose_qubit_indices_are_replaced(
    const std::vector<unsigned>& t, const std::vector<unsigned>& c) const {
    const auto m = qubit_map(*this, t, c);
    auto* r = new QuantumGate_Probabilistic(*this);
    for (auto*& g : r->gates_) {
        QuantumGateBase* n = remap_gate(g, m);
        delete g;
        g = n;
    }
    r->targets_ = t;
    r->controls_ = c;
    return r;
}
QuantumGateBase* QuantumGate_CPTP::create_gate_whose_qubit_indices_are_replaced(const std::vector<unsigned>& t,
                                                                               const std::vector<unsigned>& c) const {
    const auto m = qubit_map(*this, t, c);
    auto* r = static_cast<QuantumGate_CPTP*>(copy());
    for (auto*& g : r->gates_) {
        QuantumGateBase* n = remap_gate(g, m);
        delete g;
        g = n;
    }
    r->targets_ = t;
    r->controls_ = c;
    return r;
}
QuantumGateBase* QuantumGate_LinearCombination::create_gate_whose_qubit_indices_are_replaced(
    const std::vector<unsigned>& t, const std::vector<unsigned>& c) const {
    const auto m = qubit_map(*this, t, c);
    auto* r = new QuantumGate_LinearCombination(*this);
    for (auto*& g : r->gates_) {
        QuantumGateBase* n = remap_gate(g, m);
        delete g;
        g = n;
    }
    r->targets_ = t;
    r->controls_ = c;
    return r;
}
QuantumGateBase* QuantumGate_Adaptive::create_gate_whose_qubit_indices_are_replaced(const std::vector<unsigned>& t,
                                                                                   const std::vector<unsigned>& c) const {
    const auto m = qubit_map(*this, t, c);
    auto* r = new QuantumGate_Adaptive(*this);
    r->gate_.reset(remap_gate(gate_.get(), m));
    r->targets_ = t;
    r->controls_ = c;
    return r;
}

QuantumGate_SingleParameter::QuantumGate_SingleParameter(std::vector<unsigned> t, std::vector<unsigned> p, double angle,
                                                         std::string name)
    : paulis_(std::move(p)), angle_(angle) {
    targets_ = std::move(t);
    name_ = std::move(name);
    if (targets_.size() != paulis_.size()) throw QulacsError("one Pauli id per target");
    tcomm_ = pauli_commutation(paulis_);
}
ComplexMatrix QuantumGate_SingleParameter::get_matrix() const {
    return ClsPauliGate(targets_, paulis_, true, angle_).get_matrix();
}
QuantumGateBase* QuantumGate_SingleParameter::get_inverse() const {
    return new QuantumGate_SingleParameter(targets_, paulis_, -angle_, name_);
}
void QuantumGate_SingleParameter::update_quantum_state(QuantumStateBase* state) {
    check_targets(state, targets_, name_.c_str());
    if (auto* s = as_sv(state)) {
        if (targets_.size() == 1) {
            const unsigned t = targets_[0];
            switch (paulis_[0]) {
                case 1: RX_gate_host(t, angle_, s->data(), s->dim, s->get_cuda_stream(), s->device_number); break;
                case 2: RY_gate_host(t, angle_, s->data(), s->dim, s->get_cuda_stream(), s->device_number); break;
                case 3: RZ_gate_host(t, angle_, s->data(), s->dim, s->get_cuda_stream(), s->device_number); break;
                default: return;
            }
        } else {
            multi_qubit_Pauli_rotation_gate_partial_list_host(targets_.data(), paulis_.data(), (UINT)targets_.size(),
                                                              angle_, s->data(), s->dim, s->get_cuda_stream(),
                                                              s->device_number);
        }
        check_abi(name_.c_str());
    } else {
        QuantumGateMatrix(targets_, get_matrix()).update_quantum_state(state);
    }
}

// ------------------------------------------------------------------ gate factory
namespace gate {
namespace {
constexpr unsigned CX = QuantumGateBase::COMMUTE_X, CY = QuantumGateBase::COMMUTE_Y, CZf = QuantumGateBase::COMMUTE_Z;
constexpr unsigned PP = QuantumGateBase::PROPERTY_PAULI, PC = QuantumGateBase::PROPERTY_CLIFFORD,
                   PG = QuantumGateBase::PROPERTY_GAUSSIAN;
// commutation and class flags as the reference's named gates set them
// (src/cppsim/gate_named_one.hpp:80-330)
QuantumGateMatrix* named1(unsigned q, QuantumGateMatrix::Kind k, ComplexMatrix m, const char* name, unsigned comm,
                          unsigned prop) {
    auto* g = new QuantumGateMatrix({q}, std::move(m), {}, {}, name);
    g->kind = k;
    g->set_target_commutation({comm});
    g->set_gate_property(prop);
    return g;
}
ComplexMatrix m2(CPPCTYPE a, CPPCTYPE b, CPPCTYPE c, CPPCTYPE d) {
    ComplexMatrix m(2, 2);
    m(0, 0) = a; m(0, 1) = b; m(1, 0) = c; m(1, 1) = d;
    return m;
}
// exp(+i angle P / 2) (src/csim/update_ops_pauli_single.cpp:54-116)
QuantumGateMatrix* rotation(unsigned q, unsigned pauli, double angle, const char* name) {
    const double c = std::cos(angle / 2), s = std::sin(angle / 2);
    ComplexMatrix m = pauli == 1 ? m2(c, I_ * s, I_ * s, c) : pauli == 2 ? m2(c, s, -s, c) : m2(c + I_ * s, 0, 0, c - I_ * s);
    auto* g = new QuantumGateMatrix({q}, m, {}, {}, name);
    g->kind = pauli == 1 ? QuantumGateMatrix::RX : pauli == 2 ? QuantumGateMatrix::RY : QuantumGateMatrix::RZ;
    g->angle = angle;
    g->set_target_commutation({pauli == 1 ? CX : pauli == 2 ? CY : CZf});
    return g;
}
// get_IBMQ_matrix (src/cppsim/gate_factory.cpp:169-181)
ComplexMatrix ibmq(double theta, double phi, double lambda) {
    const CPPCTYPE e1 = std::exp(I_ * phi), e2 = std::exp(I_ * lambda);
    const double c = std::cos(theta / 2), s = std::sin(theta / 2);
    return m2(c, -e2 * s, e1 * s, e1 * e2 * c);
}
}  // namespace

QuantumGateMatrix* Identity(unsigned q) {
    return named1(q, QuantumGateMatrix::IDENTITY, ComplexMatrix::identity(2), "I", CX | CY | CZf, PP | PC | PG);
}
QuantumGateMatrix* X(unsigned q) { return named1(q, QuantumGateMatrix::X, m2(0, 1, 1, 0), "X", CX, PP | PC); }
QuantumGateMatrix* Y(unsigned q) { return named1(q, QuantumGateMatrix::Y, m2(0, -I_, I_, 0), "Y", CY, PP | PC); }
QuantumGateMatrix* Z(unsigned q) { return named1(q, QuantumGateMatrix::Z, m2(1, 0, 0, -1), "Z", CZf, PP | PC | PG); }
QuantumGateMatrix* H(unsigned q) {
    const double r = 1.0 / std::sqrt(2.0);
    return named1(q, QuantumGateMatrix::H, m2(r, r, r, -r), "H", 0, PC);
}
QuantumGateMatrix* S(unsigned q) { return named1(q, QuantumGateMatrix::S, m2(1, 0, 0, I_), "S", CZf, PC | PG); }
QuantumGateMatrix* Sdag(unsigned q) { return named1(q, QuantumGateMatrix::SDAG, m2(1, 0, 0, -I_), "Sdag", CZf, PC | PG); }
QuantumGateMatrix* T(unsigned q) {
    return named1(q, QuantumGateMatrix::T, m2(1, 0, 0, (1.0 + I_) / std::sqrt(2.0)), "T", CZf, PG);
}
QuantumGateMatrix* Tdag(unsigned q) {
    return named1(q, QuantumGateMatrix::TDAG, m2(1, 0, 0, (1.0 - I_) / std::sqrt(2.0)), "Tdag", CZf, PG);
}
QuantumGateMatrix* sqrtX(unsigned q) {
    return named1(q, QuantumGateMatrix::SQRTX, m2(0.5 + 0.5 * I_, 0.5 - 0.5 * I_, 0.5 - 0.5 * I_, 0.5 + 0.5 * I_), "sqrtX",
                  CX, PC);
}
QuantumGateMatrix* sqrtXdag(unsigned q) {
    return named1(q, QuantumGateMatrix::SQRTXDAG, m2(0.5 - 0.5 * I_, 0.5 + 0.5 * I_, 0.5 + 0.5 * I_, 0.5 - 0.5 * I_),
                  "sqrtXdag", CX, PC);
}
QuantumGateMatrix* sqrtY(unsigned q) {
    return named1(q, QuantumGateMatrix::SQRTY, m2(0.5 + 0.5 * I_, -0.5 - 0.5 * I_, 0.5 + 0.5 * I_, 0.5 + 0.5 * I_), "sqrtY",
                  CY, PC);
}
QuantumGateMatrix* sqrtYdag(unsigned q) {
    return named1(q, QuantumGateMatrix::SQRTYDAG, m2(0.5 - 0.5 * I_, 0.5 - 0.5 * I_, -0.5 + 0.5 * I_, 0.5 - 0.5 * I_),
                  "sqrtYdag", CY, PC);
}
QuantumGateMatrix* P0(unsigned q) { return named1(q, QuantumGateMatrix::P0, m2(1, 0, 0, 0), "Projection-0", 0, PC | PG); }
QuantumGateMatrix* P1(unsigned q) { return named1(q, QuantumGateMatrix::P1, m2(0, 0, 0, 1), "Projection-1", 0, PC | PG); }
QuantumGateMatrix* RX(unsigned q, double a) { return rotation(q, 1, a, "X-rotation"); }
QuantumGateMatrix* RY(unsigned q, double a) { return rotation(q, 2, a, "Y-rotation"); }
QuantumGateMatrix* RZ(unsigned q, double a) { return rotation(q, 3, a, "Z-rotation"); }
// RotX/RotY/RotZ use the opposite sign (src/cppsim/gate_factory.cpp:152-165)
QuantumGateMatrix* RotX(unsigned q, double a) { return rotation(q, 1, -a, "X-rotation"); }
QuantumGateMatrix* RotY(unsigned q, double a) { return rotation(q, 2, -a, "Y-rotation"); }
QuantumGateMatrix* RotZ(unsigned q, double a) { return rotation(q, 3, -a, "Z-rotation"); }
QuantumGateMatrix* RotInvX(unsigned q, double a) { return rotation(q, 1, a, "X-rotation"); }
QuantumGateMatrix* RotInvY(unsigned q, double a) { return rotation(q, 2, a, "Y-rotation"); }
QuantumGateMatrix* RotInvZ(unsigned q, double a) { return rotation(q, 3, a, "Z-rotation"); }
QuantumGateMatrix* U1(unsigned q, double l) { return new QuantumGateMatrix({q}, ibmq(0, 0, l), {}, {}, "DenseMatrix"); }
QuantumGateMatrix* U2(unsigned q, double p, double l) {
    return new QuantumGateMatrix({q}, ibmq(M_PI / 2, p, l), {}, {}, "DenseMatrix");
}
QuantumGateMatrix* U3(unsigned q, double t, double p, double l) {
    return new QuantumGateMatrix({q}, ibmq(t, p, l), {}, {}, "DenseMatrix");
}
QuantumGateMatrix* CNOT(unsigned c, unsigned t) {
    if (c == t) throw QulacsError("CNOT: control and target must differ");
    auto* g = new QuantumGateMatrix({t}, m2(0, 1, 1, 0), {c}, {1}, "CNOT");
    g->kind = QuantumGateMatrix::CNOT;
    g->set_target_commutation({CX});
    g->set_gate_property(PC);
    return g;
}
QuantumGateMatrix* CZ(unsigned c, unsigned t) {
    if (c == t) throw QulacsError("CZ: control and target must differ");
    auto* g = new QuantumGateMatrix({t}, m2(1, 0, 0, -1), {c}, {1}, "CZ");
    g->kind = QuantumGateMatrix::CZ;
    g->set_target_commutation({CZf});
    g->set_gate_property(PC);
    return g;
}
QuantumGateMatrix* SWAP(unsigned a, unsigned b) {
    if (a == b) throw QulacsError("SWAP: targets must differ");
    ComplexMatrix m(4, 4);
    m(0, 0) = 1; m(1, 2) = 1; m(2, 1) = 1; m(3, 3) = 1;
    auto* g = new QuantumGateMatrix({a, b}, m, {}, {}, "SWAP");
    g->kind = QuantumGateMatrix::SWAP;
    g->set_gate_property(PC);
    return g;
}
// Factories given a qubit list reject repeated qubits, as the reference's do
// (DuplicatedQubitIndexException, test/cppsim/test_gate.cpp:1540-1625).
static void require_distinct(const std::vector<unsigned>& t, const char* what) {
    std::vector<unsigned> s(t);
    std::sort(s.begin(), s.end());
    if (std::adjacent_find(s.begin(), s.end()) != s.end())
        throw QulacsError(std::string(what) + ": duplicated target");
}
QuantumGateMatrix* DenseMatrix(const std::vector<unsigned>& t, const ComplexMatrix& m) {
    require_distinct(t, "DenseMatrix");
    return new QuantumGateMatrix(t, m);
}
QuantumGateDiagonalMatrix* DiagonalMatrix(const std::vector<unsigned>& t, const std::vector<CPPCTYPE>& d) {
    require_distinct(t, "DiagonalMatrix");
    return new QuantumGateDiagonalMatrix(t, d);
}
ClsPauliGate* Pauli(const std::vector<unsigned>& t, const std::vector<unsigned>& p) {
    require_distinct(t, "Pauli");
    return new ClsPauliGate(t, p, false, 0.0);
}
ClsPauliGate* PauliRotation(const std::vector<unsigned>& t, const std::vector<unsigned>& p, double a) {
    require_distinct(t, "PauliRotation");
    return new ClsPauliGate(t, p, true, a);
}
// Haar-random unitary: QR of a complex Gaussian matrix (modified Gram-Schmidt)
// with the phases of R's diagonal divided out, as gate::RandomUnitary
// (src/cppsim/gate_factory.cpp:325-380); the Gaussian stream is this library's.
QuantumGateMatrix* RandomUnitary(const std::vector<unsigned>& t, uint64_t seed) {
    require_distinct(t, "RandomUnitary");
    Random rnd;
    rnd.set_seed(seed);
    const size_t d = size_t(1) << t.size();
    ComplexMatrix a(d, d);
    for (auto& z : a.data) z = CPPCTYPE(rnd.normal(), rnd.normal()) / std::sqrt(2.0);
    ComplexMatrix q(d, d);
    for (size_t j = 0; j < d; ++j) {
        std::vector<CPPCTYPE> v(d);
        for (size_t i = 0; i < d; ++i) v[i] = a(i, j);
        for (size_t k = 0; k < j; ++k) {
            CPPCTYPE dot = 0;
            for (size_t i = 0; i < d; ++i) dot += std::conj(q(i, k)) * v[i];
            for (size_t i = 0; i < d; ++i) v[i] -= dot * q(i, k);
        }
        double nrm = 0;
        for (auto& z : v) nrm += std::norm(z);
        nrm = std::sqrt(nrm);
        // R_jj = nrm (real positive here), so the phase fix is the identity
        for (size_t i = 0; i < d; ++i) q(i, j) = v[i] / nrm;
    }
    return new QuantumGateMatrix(t, q);
}
QuantumGate_Probabilistic* DepolarizingNoise(unsigned q, double p) {
    // probabilities {p/3, p/3, p/3, 1-p} on {X, Z, Y, I} (src/cppsim/gate_factory.cpp:454-467)
    std::unique_ptr<QuantumGateBase> x(X(q)), z(Z(q)), y(Y(q)), i(Identity(q));
    auto* g = new QuantumGate_Probabilistic({p / 3, p / 3, p / 3, 1 - p}, {x.get(), z.get(), y.get(), i.get()});
    return g;
}
QuantumGate_Probabilistic* TwoQubitDepolarizingNoise(unsigned a, unsigned b, double p) {
    if (a == b) throw QulacsError("TwoQubitDepolarizingNoise: target list contains duplicated values");
    std::vector<std::unique_ptr<QuantumGateBase>> own;
    std::vector<QuantumGateBase*> gates;
    std::vector<double> dist;
    for (unsigned i = 0; i < 16; ++i) {
        const unsigned pa = i % 4, pb = i / 4;
        own.emplace_back(i == 0 ? static_cast<QuantumGateBase*>(Identity(a)) : Pauli({a, b}, {pa, pb}));
        gates.push_back(own.back().get());
        dist.push_back(i == 0 ? 1 - p : p / 15);
    }
    return new QuantumGate_Probabilistic(dist, gates);
}
QuantumGate_Probabilistic* BitFlipNoise(unsigned q, double p) {
    std::unique_ptr<QuantumGateBase> x(X(q)), i(Identity(q));
    return new QuantumGate_Probabilistic({p, 1 - p}, {x.get(), i.get()});
}
QuantumGate_Probabilistic* DephasingNoise(unsigned q, double p) {
    std::unique_ptr<QuantumGateBase> z(Z(q)), i(Identity(q));
    return new QuantumGate_Probabilistic({p, 1 - p}, {z.get(), i.get()});
}
QuantumGate_Probabilistic* IndependentXZNoise(unsigned q, double p) {
    std::unique_ptr<QuantumGateBase> x(X(q)), z(Z(q)), y(Y(q)), i(Identity(q));
    const double p1 = p * (1 - p), p2 = p * p;
    return new QuantumGate_Probabilistic({p1, p1, p2, 1 - 2 * p1 - p2}, {x.get(), z.get(), y.get(), i.get()});
}
QuantumGate_CPTP* AmplitudeDampingNoise(unsigned q, double p) {
    const QuantumGateMatrix k0({q}, m2(1, 0, 0, std::sqrt(1 - p))), k1({q}, m2(0, std::sqrt(p), 0, 0));
    return new QuantumGate_CPTP({&k0, &k1});
}
QuantumGate_Probabilistic* Probabilistic(const std::vector<double>& dist, const std::vector<QuantumGateBase*>& gates) {
    return new QuantumGate_Probabilistic(dist, gates);
}
QuantumGate_CPTP* CPTP(const std::vector<QuantumGateBase*>& kraus) {
    return new QuantumGate_CPTP(std::vector<const QuantumGateBase*>(kraus.begin(), kraus.end()));
}
QuantumGate_CP* CP(const std::vector<QuantumGateBase*>& kraus, bool state_normalize, bool probability_normalize,
                   bool assign_zero_if_not_matched) {
    return new QuantumGate_CP(std::vector<const QuantumGateBase*>(kraus.begin(), kraus.end()), state_normalize,
                              probability_normalize, assign_zero_if_not_matched);
}
QuantumGate_Probabilistic* ProbabilisticInstrument(const std::vector<double>& dist,
                                                   const std::vector<QuantumGateBase*>& gates, unsigned reg) {
    auto* g = new QuantumGate_Probabilistic(dist, gates);
    g->set_instrument(reg);
    return g;
}
QuantumGate_LinearCombination* LinearCombination(const std::vector<CPPCTYPE>& coefs,
                                                 const std::vector<QuantumGateBase*>& gates) {
    return new QuantumGate_LinearCombination(coefs, std::vector<const QuantumGateBase*>(gates.begin(), gates.end()));
}
QuantumGate_CPTP* MultiQubitPauliMeasurement(const std::vector<unsigned>& t, const std::vector<unsigned>& p,
                                             unsigned reg) {
    std::unique_ptr<QuantumGateBase> id(Identity(0)), pg(Pauli(t, p));
    const QuantumGate_LinearCombination plus({0.5, 0.5}, {id.get(), pg.get()}), minus({0.5, -0.5}, {id.get(), pg.get()});
    auto* g = new QuantumGate_CPTP({&plus, &minus});
    g->set_instrument(reg);
    return g;
}
ClsStateReflectionGate* StateReflection(const QuantumStateMI355* r) { return new ClsStateReflectionGate(r); }
QuantumGate_Adaptive* Adaptive(const QuantumGateBase* g, QuantumGate_Adaptive::Condition c) {
    return new QuantumGate_Adaptive(g, std::move(c));
}
QuantumGate_Adaptive* Adaptive(const QuantumGateBase* g, QuantumGate_Adaptive::ConditionWithId c, unsigned id) {
    return new QuantumGate_Adaptive(g, std::move(c), id);
}
QuantumGateMatrix* ReversibleBoolean(const std::vector<unsigned>& t, const std::function<ITYPE(ITYPE, ITYPE)>& f) {
    std::vector<unsigned> s(t);
    std::sort(s.begin(), s.end());
    if (std::adjacent_find(s.begin(), s.end()) != s.end()) throw QulacsError("ReversibleBoolean: duplicated target");
    if (t.size() > 12) throw QulacsError("ReversibleBoolean: at most 12 targets");
    const ITYPE d = ITYPE(1) << t.size();
    ComplexMatrix m(d, d);
    std::vector<char> hit(d, 0);
    for (ITYPE x = 0; x < d; ++x) {
        const ITYPE y = f(x, d);
        if (y >= d || hit[y]) throw QulacsError("ReversibleBoolean: the function is not a permutation of 0..2^k-1");
        hit[y] = 1;
        m(y, x) = 1.0;
    }
    return new QuantumGateMatrix(t, m, {}, {}, "ReversibleBoolean");
}
ClsNpairQubitGate* FusedSWAP(unsigned t1, unsigned t2, unsigned b) { return new ClsNpairQubitGate(t1, t2, b); }
QuantumGate_CPTP* Instrument(const std::vector<QuantumGateBase*>& kraus, unsigned reg) {
    QuantumGate_CPTP* g = CPTP(kraus);
    g->set_instrument(reg);
    return g;
}
// projective measurement of one qubit into a classical register (src/cppsim/gate_factory.cpp:522-537)
QuantumGate_CPTP* Measurement(unsigned q, unsigned reg) {
    std::unique_ptr<QuantumGateMatrix> p0(P0(q)), p1(P1(q));
    return Instrument({p0.get(), p1.get()}, reg);
}
QuantumGateMatrix* TOFFOLI(unsigned c1, unsigned c2, unsigned t) {
    if (c1 == c2) throw QulacsError("TOFFOLI: controls must differ");
    QuantumGateMatrix* g = new QuantumGateMatrix({t}, m2(0, 1, 1, 0), {}, {}, "DenseMatrix");
    g->set_target_commutation({CX});
    g->add_control_qubit(c1, 1);
    g->add_control_qubit(c2, 1);
    return g;
}
QuantumGateMatrix* FREDKIN(unsigned c, unsigned t1, unsigned t2) {
    std::unique_ptr<QuantumGateMatrix> s(SWAP(t1, t2));
    QuantumGateMatrix* g = new QuantumGateMatrix(s->get_target_index_list(), s->get_matrix(), {}, {}, "DenseMatrix");
    g->add_control_qubit(c, 1);
    return g;
}
// Sum of two gates' matrices (src/cppsim/gate_merge.cpp:372-430): a control on the
// same qubit with the same value in both stays a control of the sum; other
// controls are folded into the matrices over the union of the qubits.
QuantumGateMatrix* add(const QuantumGateBase* a, const QuantumGateBase* b) {
    std::vector<unsigned> kq, kv;
    const auto &ca = a->get_control_index_list(), &va = a->get_control_value_list();
    const auto &cb = b->get_control_index_list(), &vb = b->get_control_value_list();
    for (size_t i = 0; i < ca.size(); ++i)
        for (size_t j = 0; j < cb.size(); ++j)
            if (ca[i] == cb[j] && va[i] == vb[j]) {
                kq.push_back(ca[i]);
                kv.push_back(va[i]);
            }
    auto fold = [&](const QuantumGateBase* g, std::vector<unsigned>* on) {
        std::vector<unsigned> c, v;
        const auto &gc = g->get_control_index_list(), &gv = g->get_control_value_list();
        for (size_t i = 0; i < gc.size(); ++i)
            if (std::find(kq.begin(), kq.end(), gc[i]) == kq.end()) {
                c.push_back(gc[i]);
                v.push_back(gv[i]);
            }
        QuantumGateMatrix rest(g->get_target_index_list(), g->get_matrix(), c, v);
        return with_controls(&rest, on);
    };
    std::vector<unsigned> on1, on2;
    const ComplexMatrix m1 = fold(a, &on1), mb = fold(b, &on2);
    const auto u = sorted_union({on1, on2});
    ComplexMatrix sum = embed(m1, on1, u);
    const ComplexMatrix e2 = embed(mb, on2, u);
    for (size_t i = 0; i < sum.data.size(); ++i) sum.data[i] += e2.data[i];
    auto* g = new QuantumGateMatrix(u, sum, kq, kv);
    g->set_target_commutation(merged_commutation({a, b}, u));
    return g;
}
QuantumGateMatrix* add(const std::vector<QuantumGateBase*>& gates) {
    if (gates.empty()) throw QulacsError("add: empty gate list");
    std::unique_ptr<QuantumGateMatrix> acc(to_matrix_gate(gates[0]));
    for (size_t i = 1; i < gates.size(); ++i) acc.reset(add(acc.get(), gates[i]));
    return acc.release();
}
QuantumGateMatrix* to_matrix_gate(const QuantumGateBase* g) {
    auto* m = new QuantumGateMatrix(g->get_target_index_list(), g->get_matrix(), g->get_control_index_list(),
                                    g->get_control_value_list());
    std::vector<unsigned> f(g->get_target_index_list().size());
    for (size_t i = 0; i < f.size(); ++i) f[i] = g->get_target_commutation(i);
    m->set_target_commutation(f);
    return m;
}
// second * first (src/cppsim/gate_merge.cpp:153-357): a control on the same qubit with
// the same value in both gates stays a control; every other qubit becomes a target,
// in ascending order, with the other controls folded into the matrix. The class
// flags are the AND of both gates'.
QuantumGateMatrix* merge(const QuantumGateBase* first, const QuantumGateBase* second) {
    std::vector<unsigned> kq, kv;
    const auto &ca = first->get_control_index_list(), &va = first->get_control_value_list();
    const auto &cb = second->get_control_index_list(), &vb = second->get_control_value_list();
    for (size_t i = 0; i < ca.size(); ++i)
        for (size_t j = 0; j < cb.size(); ++j)
            if (ca[i] == cb[j] && va[i] == vb[j]) {
                kq.push_back(ca[i]);
                kv.push_back(va[i]);
            }
    auto fold = [&](const QuantumGateBase* g, std::vector<unsigned>* on) {
        std::vector<unsigned> c, v;
        const auto &gc = g->get_control_index_list(), &gv = g->get_control_value_list();
        for (size_t i = 0; i < gc.size(); ++i)
            if (std::find(kq.begin(), kq.end(), gc[i]) == kq.end()) {
                c.push_back(gc[i]);
                v.push_back(gv[i]);
            }
        QuantumGateMatrix rest(g->get_target_index_list(), g->get_matrix(), c, v);
        return with_controls(&rest, on);
    };
    std::vector<unsigned> on1, on2;
    const ComplexMatrix m1 = fold(first, &on1), m2_ = fold(second, &on2);
    const auto u = sorted_union({on1, on2});
    auto* g = new QuantumGateMatrix(u, embed(m2_, on2, u) * embed(m1, on1, u), kq, kv);
    g->set_target_commutation(merged_commutation({first, second}, u));
    g->set_gate_property(first->get_property_value() & second->get_property_value() & ~QuantumGateBase::PROPERTY_PARAMETRIC);
    return g;
}
QuantumGate_SingleParameter* ParametricRX(unsigned q, double a) { return new QuantumGate_SingleParameter({q}, {1}, a, "ParametricRX"); }
QuantumGate_SingleParameter* ParametricRY(unsigned q, double a) { return new QuantumGate_SingleParameter({q}, {2}, a, "ParametricRY"); }
QuantumGate_SingleParameter* ParametricRZ(unsigned q, double a) { return new QuantumGate_SingleParameter({q}, {3}, a, "ParametricRZ"); }
QuantumGate_SingleParameter* ParametricPauliRotation(const std::vector<unsigned>& t, const std::vector<unsigned>& p, double a) {
    require_distinct(t, "ParametricPauliRotation");
    return new QuantumGate_SingleParameter(t, p, a, "ParametricPauliRotation");
}
}  // namespace gate

// ------------------------------------------------------------------ circuit
QuantumCircuit::QuantumCircuit(const QuantumCircuit& o) : qubit_count_(o.qubit_count_) {
    for (auto* g : o.gates_) gates_.push_back(g->copy());
}
QuantumCircuit::~QuantumCircuit() {
    release_plan();
    for (auto* g : gates_) delete g;
}
const QuantumGateBase* QuantumCircuit::get_gate(size_t i) const {
    if (i >= gates_.size()) throw QulacsError("get_gate: index out of range");
    return gates_[i];
}
// Every qubit a gate names must exist in the circuit (src/cppsim/circuit.cpp:144-150).
void QuantumCircuit::check_gate_indices(const QuantumGateBase& g) const {
    for (unsigned q : g.get_target_index_list())
        if (q >= qubit_count_) throw QulacsError("add_gate: target index exceeds the circuit's qubit count");
    for (unsigned q : g.get_control_index_list())
        if (q >= qubit_count_) throw QulacsError("add_gate: control index exceeds the circuit's qubit count");
}
void QuantumCircuit::add_gate(const QuantumGateBase& g) { add_gate_consume(g.copy()); }
void QuantumCircuit::add_gate(const QuantumGateBase& g, size_t pos) {
    if (pos > gates_.size()) throw QulacsError("add_gate: position out of range");
    check_gate_indices(g);
    gates_.insert(gates_.begin() + pos, g.copy());
    touched();
}
void QuantumCircuit::add_gate_consume(QuantumGateBase* g) {
    try {
        check_gate_indices(*g);
    } catch (...) {
        delete g;
        throw;
    }
    gates_.push_back(g);
    touched();
}
void QuantumCircuit::remove_gate(size_t i) {
    if (i >= gates_.size()) throw QulacsError("remove_gate: index out of range");
    delete gates_[i];
    gates_.erase(gates_.begin() + i);
    touched();
}
void QuantumCircuit::merge_circuit(const QuantumCircuit& c) {
    const size_t n = c.gates_.size();  // `c` may be this circuit
    for (size_t i = 0; i < n; ++i) add_gate(*c.gates_[i]);
}
void QuantumCircuit::add_noise_gate(const QuantumGateBase& g, const std::string& type, double p) {
    std::vector<unsigned> qs = g.get_target_index_list();
    for (unsigned c : g.get_control_index_list()) qs.push_back(c);
    auto need = [&](size_t k) {
        if (qs.size() > k || (k == 1 && qs.size() != 1))
            throw QulacsError("add_noise_gate: " + type + " noise takes " + std::to_string(k) + " qubit(s), the gate has " +
                              std::to_string(qs.size()));
    };
    QuantumGateBase* noise = nullptr;
    if (type == "Depolarizing") {
        need(2);
        if (qs.size() == 1) noise = gate::DepolarizingNoise(qs[0], p);
        else if (qs.size() == 2) noise = gate::TwoQubitDepolarizingNoise(qs[0], qs[1], p);
        else throw QulacsError("add_noise_gate: Depolarizing noise takes 1 or 2 qubits");
    } else if (type == "BitFlip") {
        need(1);
        noise = gate::BitFlipNoise(qs[0], p);
    } else if (type == "Dephasing") {
        need(1);
        noise = gate::DephasingNoise(qs[0], p);
    } else if (type == "IndependentXZ") {
        need(1);
        noise = gate::IndependentXZNoise(qs[0], p);
    } else if (type == "AmplitudeDamping") {
        need(1);
        noise = gate::AmplitudeDampingNoise(qs[0], p);
    } else {
        throw QulacsError("add_noise_gate: unknown noise type '" + type + "'");
    }
    add_gate(g);
    add_gate_consume(noise);
}
void QuantumCircuit::add_observable_rotation_gate(const GeneralQuantumOperator& o, double angle, unsigned repeats) {
    if (!o.is_hermitian()) throw QulacsError("add_observable_rotation_gate: the observable must be Hermitian");
    if (repeats == 0) repeats = (unsigned)std::ceil(angle * (double)o.get_qubit_count() * 100.0);
    for (unsigned r = 0; r < repeats; ++r)
        for (size_t k = 0; k < o.get_term_count(); ++k) {
            const PauliOperator& t = o.get_term(k);
            add_gate_consume(
                gate::PauliRotation(t.get_index_list(), t.get_pauli_id_list(), t.get_coef().real() * angle / repeats));
        }
}
void QuantumCircuit::add_diagonal_observable_rotation_gate(const GeneralQuantumOperator& o, double angle) {
    if (!o.is_hermitian()) throw QulacsError("add_diagonal_observable_rotation_gate: the observable must be Hermitian");
    for (size_t k = 0; k < o.get_term_count(); ++k) {
        const PauliOperator& t = o.get_term(k);
        std::unique_ptr<ClsPauliGate> g(
            gate::PauliRotation(t.get_index_list(), t.get_pauli_id_list(), t.get_coef().real() * angle));
        if (!g->is_diagonal()) throw QulacsError("add_diagonal_observable_rotation_gate: the observable is not diagonal");
        add_gate_consume(g.release());
    }
}

namespace {
std::atomic<bool> g_fusion{true};
// Shortest run worth a program: below 22 qubits a gate's launch costs less than building
// and uploading a program, so short runs stay per gate; from 22 qubits every gate is a
// pass over at least 64 MiB and two gates in one stage already halve the traffic.
size_t min_fused_run(unsigned n) { return n >= 22 ? 2 : 8; }

// One gate as a stage-executor descriptor; matrices go to `store` (stable addresses).
bool lower_gate(const QuantumGateBase* g, qm_batch_gate* out, std::deque<std::vector<qm_complex>>* store,
                bool* skip) {
    *skip = false;
    if (g->is_noise() || !g->has_matrix()) return false;
    if (const auto* m = dynamic_cast<const QuantumGateMatrix*>(g)) {
        if (m->kind == QuantumGateMatrix::IDENTITY) {
            *skip = true;
            return true;
        }
        if (m->kind == QuantumGateMatrix::CNOT || m->kind == QuantumGateMatrix::CZ) {
            out->kind = m->kind == QuantumGateMatrix::CNOT ? QM_BATCH_CNOT : QM_BATCH_CZ;
            out->qubit_count = 2;
            out->qubits[0] = m->get_control_index_list()[0];
            out->qubits[1] = m->get_target_index_list()[0];
            out->matrix = nullptr;
            return true;
        }
    }
    if (g->get_target_index_list().size() + g->get_control_index_list().size() > 4) return false;
    std::vector<unsigned> on;
    const ComplexMatrix m = with_controls(g, &on);
    const size_t d = m.rows;
    bool diagonal = true;
    for (size_t y = 0; y < d && diagonal; ++y)
        for (size_t x = 0; x < d; ++x)
            if (x != y && m(y, x) != CPPCTYPE(0.0)) {
                diagonal = false;
                break;
            }
    store->emplace_back();
    auto& v = store->back();
    if (diagonal) {
        for (size_t i = 0; i < d; ++i) v.push_back(qc(m(i, i)));
    } else {
        for (const auto& z : m.data) v.push_back(qc(z));
    }
    out->kind = diagonal ? QM_BATCH_DIAG : QM_BATCH_DENSE;
    out->qubit_count = (UINT)on.size();
    for (size_t i = 0; i < on.size(); ++i) out->qubits[i] = on[i];
    out->matrix = v.data();
    return true;
}

// Superoperator of g on the ascending qubit list q (4^|q| x 4^|q|, member = row bits *
// 2^|q| + column bits): sum_i w_i K_i (x) conj(K_i) over g's Kraus form. False when a part
// of g has no matrix.
bool superop_on(const QuantumGateBase* g, const std::vector<unsigned>& q, ComplexMatrix* S) {
    const size_t d = size_t(1) << q.size();
    *S = ComplexMatrix(d * d, d * d);
    auto add = [&](const QuantumGateBase* k, double w) {
        if (k->is_noise() || !k->has_matrix()) return false;
        std::vector<unsigned> on;
        const ComplexMatrix m = with_controls(k, &on);
        accumulate_superop(*S, embed(m, on, q), w);
        return true;
    };
    if (const auto* p = dynamic_cast<const QuantumGate_Probabilistic*>(g)) {
        const double rest = 1.0 - p->get_cumulative_distribution().back();
        if (rest != 0.0) accumulate_superop(*S, ComplexMatrix::identity(d), rest);
        for (size_t i = 0; i < p->get_gate_list().size(); ++i)
            if (!add(p->get_gate_list()[i], p->get_distribution()[i])) return false;
        return true;
    }
    if (const auto* c = dynamic_cast<const QuantumGate_CPTP*>(g)) {
        for (const auto* k : c->get_gate_list())
            if (!add(k, 1.0)) return false;
        return true;
    }
    return add(g, 1.0);
}
bool dm_fusable(const QuantumGateBase* g) {
    const std::vector<QuantumGateBase*>* parts = nullptr;
    if (const auto* p = dynamic_cast<const QuantumGate_Probabilistic*>(g)) parts = &p->get_gate_list();
    else if (const auto* c = dynamic_cast<const QuantumGate_CPTP*>(g)) parts = &c->get_gate_list();
    if (!parts) return !g->is_noise() && g->has_matrix();
    for (const auto* k : *parts)
        if (k->is_noise() || !k->has_matrix()) return false;
    return true;
}
}  // namespace

// Density matrices: consecutive gates and channels whose qubits fit in 2 together are
// composed on the host into one superoperator and applied in one pass
// (dm_superoperator_host); a lone gate keeps its own entry point.  Three to five
// single-qubit channels in a row on distinct qubits (the noise after a
// multi-qubit gate) are one pass of qm_dm_single_qubit_channels_host when each
// maps {00, 11} and {01, 10} onto themselves.
namespace {
constexpr size_t kDmFuseQubits = 2;
constexpr size_t kDmChannelRun = 5;  // channels per qm_dm_single_qubit_channels_host call

// Single-qubit channel whose 4 x 4 superoperator only links 00 with 11 and 01
// with 10 (row bit * 2 + column bit); its superoperator goes to *S.
bool orbit_channel(const QuantumGateBase* g, ComplexMatrix* S) {
    if (!g->is_noise() || g->get_target_index_list().size() != 1 || !g->get_control_index_list().empty() ||
        !dm_fusable(g))
        return false;
    if (!superop_on(g, {g->get_target_index_list()[0]}, S)) return false;
    static const int cross[8][2] = {{0, 1}, {0, 2}, {3, 1}, {3, 2}, {1, 0}, {1, 3}, {2, 0}, {2, 3}};
    for (const auto& rc : cross)
        if ((*S)(rc[0], rc[1]) != CPPCTYPE(0.0, 0.0)) return false;
    return true;
}
}  // namespace
void QuantumCircuit::update_dm_fused(DensityMatrixMI355* d) {
    std::vector<unsigned> q;
    size_t first = 0, count = 0;
    auto flush = [&]() {
        if (count == 1) gates_[first]->update_quantum_state(d);
        if (count >= 2) {
            const size_t dd = size_t(1) << (2 * q.size());
            ComplexMatrix S = ComplexMatrix::identity(dd), Sg;
            for (size_t k = first; k < first + count; ++k) {
                if (!superop_on(gates_[k], q, &Sg)) throw QulacsError("update_quantum_state: no superoperator");
                S = Sg * S;
            }
            dm_superoperator_host(q.data(), (UINT)q.size(), qc(S.data.data()), d->data(), d->dim, d->get_cuda_stream(),
                                  d->device_number);
            check_abi("update_quantum_state");
        }
        q.clear();
        count = 0;
    };
    for (size_t k = 0; k < gates_.size(); ++k) {
        {
            std::vector<unsigned> cq;
            std::vector<CPPCTYPE> sops;
            ComplexMatrix S;
            while (k + cq.size() < gates_.size() && cq.size() < kDmChannelRun &&
                   orbit_channel(gates_[k + cq.size()], &S)) {
                const unsigned t = gates_[k + cq.size()]->get_target_index_list()[0];
                if (std::find(cq.begin(), cq.end(), t) != cq.end()) break;
                cq.push_back(t);
                for (size_t r = 0; r < 4; ++r)
                    for (size_t c = 0; c < 4; ++c) sops.push_back(S(r, c));
            }
            if (cq.size() >= 3) {
                flush();
                qm_dm_single_qubit_channels_host(cq.data(), (UINT)cq.size(), qc(sops.data()), d->data(), d->dim,
                                                 d->get_cuda_stream(), d->device_number);
                check_abi("update_quantum_state");
                k += cq.size() - 1;
                continue;
            }
        }
        const QuantumGateBase* g = gates_[k];
        const auto qs = sorted_union({g->get_target_index_list(), g->get_control_index_list()});
        if (qs.size() > kDmFuseQubits || !dm_fusable(g)) {
            flush();
            gates_[k]->update_quantum_state(d);
            continue;
        }
        auto nq = sorted_union({q, qs});
        if (nq.size() > kDmFuseQubits) {
            flush();
            nq = qs;
        }
        if (count == 0) first = k;
        q = nq;
        ++count;
    }
    flush();
}

void set_circuit_fusion(bool enabled) { g_fusion = enabled; }
bool get_circuit_fusion() { return g_fusion; }

void QuantumCircuit::release_plan() {
    for (auto& r : plan_) qm_circuit_destroy(r.program);
    plan_.clear();
    plan_version_ = ~0ULL;
}
void QuantumCircuit::build_plan() {
    release_plan();
    const size_t n = gates_.size();
    size_t i = 0;
    while (i < n) {
        std::deque<std::vector<qm_complex>> store;
        std::vector<qm_batch_gate> lowered;
        size_t j = i, members = 0;
        for (; j < n; ++j) {
            qm_batch_gate bg{};
            bool skip = false;
            if (!lower_gate(gates_[j], &bg, &store, &skip)) break;
            ++members;
            if (skip) continue;
            lowered.push_back(bg);
        }
        if (members >= min_fused_run(qubit_count_)) {
            qm_circuit* prog = qm_circuit_create(qubit_count_, lowered.data(), (UINT)lowered.size(), 0, 0);
            if (prog) {
                plan_.push_back({i, j, prog});
                i = j;
                continue;
            }
            qm_clear_last_error();  // fall back to per-gate execution
        }
        const size_t stop = (j == i) ? i + 1 : j;  // an unfusable gate, or a short run
        if (!plan_.empty() && plan_.back().program == nullptr) plan_.back().end = stop;
        else plan_.push_back({i, stop, nullptr});
        i = stop;
    }
    plan_version_ = version_;
}

// src/cppsim/circuit.cpp:46-57
void QuantumCircuit::update_quantum_state(QuantumStateBase* s) {
    if (auto* dm = as_dm(s); dm && g_fusion) {
        if (s->qubit_count != qubit_count_) throw QulacsError("update_quantum_state: qubit count mismatch");
        for (const auto* g : gates_) {
            check_targets(s, g->get_target_index_list(), "update_quantum_state");
            check_targets(s, g->get_control_index_list(), "update_quantum_state");
        }
        update_dm_fused(dm);
        return;
    }
    auto* sv = as_sv(s);
    if (!g_fusion || !sv || qubit_count_ < 4 || qubit_count_ > QM_CIRCUIT_MAX_QUBITS ||
        gates_.size() < min_fused_run(qubit_count_)) {
        update_quantum_state(s, 0, gates_.size());
        return;
    }
    if (s->qubit_count != qubit_count_) throw QulacsError("update_quantum_state: qubit count mismatch");
    std::lock_guard<std::mutex> lock(plan_mutex_);
    if (plan_version_ != version_) build_plan();
    for (const auto& r : plan_) {
        if (!r.program) {
            for (size_t k = r.begin; k < r.end; ++k) gates_[k]->update_quantum_state(s);
            continue;
        }
        qm_circuit_apply(r.program, sv->data(), sv->dim, sv->get_cuda_stream(), sv->device_number);
        check_abi("update_quantum_state");
    }
}
void QuantumCircuit::update_quantum_state(QuantumStateBase* s, size_t start, size_t end) {
    if (s->qubit_count != qubit_count_) throw QulacsError("update_quantum_state: qubit count mismatch");
    if (start > end || end > gates_.size()) throw QulacsError("update_quantum_state: bad gate range");
    for (size_t i = start; i < end; ++i) gates_[i]->update_quantum_state(s);
}
unsigned QuantumCircuit::calculate_depth() const {
    std::vector<unsigned> depth(qubit_count_, 0);
    unsigned best = 0;
    for (auto* g : gates_) {
        unsigned d = 0;
        auto qs = g->get_target_index_list();
        const auto& cl = g->get_control_index_list();
        qs.insert(qs.end(), cl.begin(), cl.end());
        for (unsigned q : qs) d = std::max(d, depth[q]);
        for (unsigned q : qs) depth[q] = d + 1;
        best = std::max(best, d + 1);
    }
    return best;
}
std::string QuantumCircuit::to_string() const {
    std::ostringstream os;
    os << "*** Quantum Circuit Info ***\n# of qubit: " << qubit_count_ << "\n# of step : " << calculate_depth()
       << "\n# of gate : " << gates_.size() << "\n";
    return os.str();
}
QuantumCircuit* QuantumCircuit::get_inverse() const {
    auto* c = new QuantumCircuit(qubit_count_);
    for (auto it = gates_.rbegin(); it != gates_.rend(); ++it) c->add_gate_consume((*it)->get_inverse());
    return c;
}

// ------------------------------------------------------------------ operators
// src/cppsim/pauli_operator.cpp:16-46
PauliOperator::PauliOperator(const std::string& s, CPPCTYPE coef) : coef_(coef) {
    std::istringstream ss(s);
    std::string p;
    unsigned idx;
    while (ss >> p) {
        if (!(ss >> idx)) throw QulacsError("Pauli string without index: " + s);
        unsigned id;
        if (p == "I" || p == "i") id = 0;
        else if (p == "X" || p == "x") id = 1;
        else if (p == "Y" || p == "y") id = 2;
        else if (p == "Z" || p == "z") id = 3;
        else throw QulacsError("invalid Pauli string is given : " + p);
        if (id != 0) add_single_Pauli(idx, id);
    }
}
// src/cppsim/pauli_operator.cpp:87-96: identity entries stay in the index list
PauliOperator::PauliOperator(const std::vector<unsigned>& t, const std::vector<unsigned>& p, CPPCTYPE coef) : coef_(coef) {
    if (t.size() != p.size()) throw QulacsError("one Pauli id per target");
    for (size_t i = 0; i < t.size(); ++i) add_single_Pauli(t[i], p[i]);
}
void PauliOperator::add_single_Pauli(unsigned q, unsigned p) {
    if (p > 3) throw QulacsError("Pauli id must be 0..3");
    targets_.push_back(q);
    paulis_.push_back(p);
}
std::string PauliOperator::get_pauli_string() const {
    std::ostringstream os;
    const char* names = "IXYZ";
    bool first = true;
    for (size_t i = 0; i < targets_.size(); ++i) {
        if (paulis_[i] == 0) continue;  // identities are not printed (src/cppsim/pauli_operator.cpp:263-285)
        os << (first ? "" : " ") << names[paulis_[i]] << " " << targets_[i];
        first = false;
    }
    return os.str();
}
void PauliOperator::xz_bits(std::vector<uint8_t>* x, std::vector<uint8_t>* z) const {
    unsigned n = 0;
    for (unsigned q : targets_) n = std::max(n, q + 1);
    x->assign(n, 0);
    z->assign(n, 0);
    for (size_t i = 0; i < targets_.size(); ++i) {
        if (paulis_[i] == 1 || paulis_[i] == 2) (*x)[targets_[i]] = 1;
        if (paulis_[i] == 2 || paulis_[i] == 3) (*z)[targets_[i]] = 1;
    }
}
bool PauliOperator::same_pauli_string(const PauliOperator& o) const {
    std::vector<uint8_t> x0, z0, x1, z1;
    xz_bits(&x0, &z0);
    o.xz_bits(&x1, &z1);
    const size_t n = std::max(x0.size(), x1.size());
    x0.resize(n), z0.resize(n), x1.resize(n), z1.resize(n);
    return x0 == x1 && z0 == z1;
}
// Single-qubit products: XY = iZ, YZ = iX, ZX = iY and the reversed orders with -i.
PauliOperator PauliOperator::operator*(const PauliOperator& t) const {
    std::vector<uint8_t> x0, z0, x1, z1;
    xz_bits(&x0, &z0);
    t.xz_bits(&x1, &z1);
    const size_t n = std::max(x0.size(), x1.size());
    x0.resize(n), z0.resize(n), x1.resize(n), z1.resize(n);
    const CPPCTYPE I(0.0, 1.0);
    CPPCTYPE phase = 1.0;
    auto id = [](uint8_t x, uint8_t z) -> unsigned { return x ? (z ? 2u : 1u) : (z ? 3u : 0u); };
    std::vector<unsigned> tl, pl;
    for (size_t q = 0; q < n; ++q) {
        const unsigned a = id(x0[q], z0[q]), b = id(x1[q], z1[q]);
        if (a != 0 && b != 0 && a != b) phase *= ((b == a % 3 + 1) ? I : -I);  // X->Y->Z->X is +i
        const unsigned c = id(x0[q] ^ x1[q], z0[q] ^ z1[q]);
        if (c != 0) {
            tl.push_back((unsigned)q);
            pl.push_back(c);
        }
    }
    return PauliOperator(tl, pl, coef_ * t.get_coef() * phase);
}
// src/csim/utility.cpp:12-43
void PauliOperator::masks(ITYPE* bf, ITYPE* pf, unsigned* rot) const {
    *bf = 0; *pf = 0; *rot = 0;
    for (size_t i = 0; i < targets_.size(); ++i) {
        const ITYPE b = 1ULL << targets_[i];
        switch (paulis_[i]) {
            case 1: *bf ^= b; break;
            case 2: *bf ^= b; *pf ^= b; ++*rot; break;
            case 3: *pf ^= b; break;
            default: break;
        }
    }
}
CPPCTYPE PauliOperator::get_expectation_value(const QuantumStateBase* state) const {
    for (unsigned q : targets_)
        if (q >= state->qubit_count) throw QulacsError("PauliOperator acts beyond the state's qubits");
    double v;
    if (state->is_state_vector())
        v = expectation_value_multi_qubit_Pauli_operator_partial_list_host(
            targets_.data(), paulis_.data(), (UINT)targets_.size(), state->data(), state->dim, state->get_cuda_stream(),
            state->device_number);
    else
        v = dm_expectation_value_multi_qubit_Pauli_operator_partial_list_host(
            targets_.data(), paulis_.data(), (UINT)targets_.size(), state->data(), state->dim, state->get_cuda_stream(),
            state->device_number);
    check_abi("PauliOperator::get_expectation_value");
    return coef_ * v;
}
CPPCTYPE PauliOperator::get_transition_amplitude(const QuantumStateBase* bra, const QuantumStateBase* ket) const {
    if (!bra->is_state_vector() || !ket->is_state_vector()) throw QulacsError("transition amplitude needs state vectors");
    qm_synchronize(bra->get_cuda_stream(), bra->device_number);
    const qm_complex z = transition_amplitude_multi_qubit_Pauli_operator_partial_list_host(
        targets_.data(), paulis_.data(), (UINT)targets_.size(), bra->data(), ket->data(), ket->dim,
        ket->get_cuda_stream(), ket->device_number);
    check_abi("PauliOperator::get_transition_amplitude");
    return coef_ * cc(z);
}

void GeneralQuantumOperator::add_operator(const PauliOperator& p) {
    for (unsigned q : p.get_index_list())
        if (q >= qubit_count_) throw QulacsError("add_operator: Pauli operator acts beyond qubit_count");
    terms_.push_back(p);
}
void Observable::add_operator(const PauliOperator& p) {
    if (p.get_coef().imag() != 0.0) throw QulacsError("Observable: PauliOperator must be Hermitian (real coefficient)");
    GeneralQuantumOperator::add_operator(p);
}
void GeneralQuantumOperator::add_operator(CPPCTYPE coef, const std::string& s) { add_operator(PauliOperator(s, coef)); }
void GeneralQuantumOperator::add_operator(const std::vector<unsigned>& t, const std::vector<unsigned>& p, CPPCTYPE coef) {
    add_operator(PauliOperator(t, p, coef));
}
bool GeneralQuantumOperator::is_hermitian() const {
    for (const auto& t : terms_)
        if (t.get_coef().imag() != 0.0) return false;
    return true;
}
CPPCTYPE GeneralQuantumOperator::get_expectation_value(const QuantumStateBase* state) const {
    if (qubit_count_ > state->qubit_count) throw QulacsError("get_expectation_value: invalid qubit count");
    if (terms_.empty()) return 0.0;
    if (!state->is_state_vector()) {
        CPPCTYPE sum = 0;
        for (const auto& t : terms_) sum += t.get_expectation_value(state);
        return sum;
    }
    std::vector<ITYPE> bf(terms_.size()), pf(terms_.size());
    std::vector<UINT> rot(terms_.size());
    std::vector<qm_complex> co(terms_.size());
    for (size_t i = 0; i < terms_.size(); ++i) {
        unsigned r;
        terms_[i].masks(&bf[i], &pf[i], &r);
        rot[i] = r;
        co[i] = qc(terms_[i].get_coef());
    }
    const qm_complex z = qm_observable_expectation_host(bf.data(), pf.data(), rot.data(), co.data(), (UINT)terms_.size(),
                                                        state->data(), state->dim, state->get_cuda_stream(),
                                                        state->device_number);
    check_abi("get_expectation_value");
    return cc(z);
}
CPPCTYPE GeneralQuantumOperator::get_transition_amplitude(const QuantumStateBase* bra, const QuantumStateBase* ket) const {
    CPPCTYPE sum = 0;
    for (const auto& t : terms_) sum += t.get_transition_amplitude(bra, ket);
    return sum;
}
// src/cppsim/general_quantum_operator.cpp:330-348
void GeneralQuantumOperator::apply_to_state(QuantumStateMI355* work, const QuantumStateMI355& src,
                                            QuantumStateMI355* dst) const {
    dst->multiply_coef(0.0);
    for (const auto& t : terms_) {
        work->load(&src);
        if (!t.get_index_list().empty()) {
            multi_qubit_Pauli_gate_partial_list_host(t.get_index_list().data(), t.get_pauli_id_list().data(),
                                                     (UINT)t.get_index_list().size(), work->data(), work->dim,
                                                     work->get_cuda_stream(), work->device_number);
            check_abi("apply_to_state");
        }
        dst->add_state_with_coef(t.get_coef(), work);
    }
}
void GeneralQuantumOperator::apply_to_state(QuantumStateMI355* state, QuantumStateMI355* dst) const {
    if (state->qubit_count != dst->qubit_count) throw QulacsError("apply_to_state: qubit counts differ");
    dst->multiply_coef(0.0);
    for (const auto& t : terms_) {
        const auto& tl = t.get_index_list();
        const auto& pl = t.get_pauli_id_list();
        if (!tl.empty()) {
            multi_qubit_Pauli_gate_partial_list_host(tl.data(), pl.data(), (UINT)tl.size(), state->data(), state->dim,
                                                     state->get_cuda_stream(), state->device_number);
            check_abi("apply_to_state");
        }
        dst->add_state_with_coef(t.get_coef(), state);
        if (!tl.empty()) {  // every Pauli string is its own inverse
            multi_qubit_Pauli_gate_partial_list_host(tl.data(), pl.data(), (UINT)tl.size(), state->data(), state->dim,
                                                     state->get_cuda_stream(), state->device_number);
            check_abi("apply_to_state");
        }
    }
}

GeneralQuantumOperator GeneralQuantumOperator::copy_base() const {
    GeneralQuantumOperator c(qubit_count_);
    c.terms_ = terms_;
    return c;
}
GeneralQuantumOperator GeneralQuantumOperator::get_dagger() const {
    GeneralQuantumOperator d(qubit_count_);
    for (const auto& t : terms_) d.add_operator(std::conj(t.get_coef()), t.get_pauli_string());
    return d;
}
std::string GeneralQuantumOperator::to_string() const {
    std::ostringstream os;
    for (size_t i = 0; i < terms_.size(); ++i)
        os << terms_[i].get_coef() << " " << terms_[i].get_pauli_string() << (i + 1 < terms_.size() ? " + " : "");
    return os.str();
}
std::string Observable::to_string() const {
    std::ostringstream os;
    for (size_t i = 0; i < terms_.size(); ++i)
        os << terms_[i].get_coef().real() << " " << terms_[i].get_pauli_string()
           << (i + 1 < terms_.size() ? " + " : "");
    return os.str();
}
// Draw order per term: one int32 per qubit (Pauli id = value % 4), then one uniform.
void GeneralQuantumOperator::add_random_operator(unsigned operator_count) {
    for (unsigned k = 0; k < operator_count; ++k) {
        std::vector<unsigned> tl(qubit_count_), pl(qubit_count_);
        for (unsigned q = 0; q < qubit_count_; ++q) {
            tl[q] = q;
            pl[q] = random_.int32() % 4;
        }
        const CPPCTYPE coef = random_.uniform() * 2 - 1.0;
        add_operator(PauliOperator(tl, pl, coef));
    }
}
void GeneralQuantumOperator::add_random_operator(unsigned operator_count, unsigned seed) {
    random_.set_seed(seed);
    add_random_operator(operator_count);
}
CPPCTYPE GeneralQuantumOperator::calculate_default_mu() const {
    double mu = 0.0;
    for (const auto& t : terms_) mu += std::abs(t.get_coef().real());
    return mu;
}

// terms_[i] += sign * (every term of `terms` with the same Pauli string); then each term
// of `terms` that matches nothing in the list (as the list grows) is appended with
// coefficient sign * coef (src/cppsim/general_quantum_operator.cpp:557-608 and 650-703).
void GeneralQuantumOperator::merge_terms(const std::vector<PauliOperator>& in, double sign) {
    const std::vector<PauliOperator> terms(in);  // `in` may be this operator's own list
    for (auto& t : terms_)
        for (const auto& a : terms)
            if (t.same_pauli_string(a)) t.change_coef(t.get_coef() + sign * a.get_coef());
    for (const auto& a : terms) {
        bool found = false;
        for (const auto& t : terms_)
            if (t.same_pauli_string(a)) {
                found = true;
                break;
            }
        if (!found) {
            PauliOperator c(a);
            c.change_coef(sign * a.get_coef());
            add_operator(c);
        }
    }
}
GeneralQuantumOperator& GeneralQuantumOperator::operator+=(const GeneralQuantumOperator& t) {
    merge_terms(t.terms_, 1.0);
    return *this;
}
GeneralQuantumOperator& GeneralQuantumOperator::operator+=(const PauliOperator& t) {
    merge_terms({t}, 1.0);
    return *this;
}
GeneralQuantumOperator& GeneralQuantumOperator::operator-=(const GeneralQuantumOperator& t) {
    merge_terms(t.terms_, -1.0);
    return *this;
}
GeneralQuantumOperator& GeneralQuantumOperator::operator-=(const PauliOperator& t) {
    merge_terms({t}, -1.0);
    return *this;
}
// Products term by term, each folded in with += (src/cppsim/general_quantum_operator.cpp:750-805).
GeneralQuantumOperator& GeneralQuantumOperator::operator*=(const GeneralQuantumOperator& target) {
    const std::vector<PauliOperator> left(terms_), right(target.terms_);
    terms_.clear();
    for (const auto& a : left)
        for (const auto& b : right) *this += a * b;
    return *this;
}
GeneralQuantumOperator& GeneralQuantumOperator::operator*=(const PauliOperator& target) {
    const std::vector<PauliOperator> left(terms_);
    terms_.clear();
    for (const auto& a : left) *this += a * target;
    return *this;
}
GeneralQuantumOperator& GeneralQuantumOperator::operator*=(CPPCTYPE c) {
    for (auto& t : terms_) t *= c;
    return *this;
}

// ------------------------------------------------------------------ eigenvalue solvers
namespace {

using CMat = std::vector<std::vector<CPPCTYPE>>;

// Eigenvalues of a complex upper Hessenberg matrix by the shifted QR iteration
// (Givens rotations, Wilkinson shift, deflation on negligible subdiagonals).
std::vector<CPPCTYPE> hessenberg_eigenvalues(CMat h) {
    const int n = (int)h.size();
    std::vector<CPPCTYPE> ev(n);
    const double eps = std::numeric_limits<double>::epsilon();
    double hnorm = 0.0;
    for (const auto& row : h)
        for (const auto& v : row) hnorm = std::max(hnorm, std::abs(v));
    int hi = n - 1, iter = 0;
    while (hi >= 0) {
        if (hi == 0) {
            ev[0] = h[0][0];
            break;
        }
        int l = hi;
        for (; l > 0; --l)
            if (std::abs(h[l][l - 1]) <=
                eps * std::max(std::abs(h[l - 1][l - 1]) + std::abs(h[l][l]), eps * hnorm)) {
                h[l][l - 1] = 0.0;
                break;
            }
        if (l == hi) {
            ev[hi] = h[hi][hi];
            --hi;
            iter = 0;
            continue;
        }
        if (++iter > 1000) throw QulacsError("eigenvalue iteration did not converge");
        // Wilkinson shift: the eigenvalue of the trailing 2x2 block nearer h[hi][hi]
        const CPPCTYPE a = h[hi - 1][hi - 1], b = h[hi - 1][hi], c = h[hi][hi - 1], d = h[hi][hi];
        const CPPCTYPE half = 0.5 * (a - d), disc = std::sqrt(half * half + b * c);
        CPPCTYPE mu = d - b * c / (half + (std::abs(half + disc) >= std::abs(half - disc) ? disc : -disc));
        if (!std::isfinite(mu.real()) || !std::isfinite(mu.imag())) mu = d;
        if (iter % 11 == 10) mu = d + std::abs(c);  // exceptional shift
        for (int k = l; k <= hi; ++k) h[k][k] -= mu;
        std::vector<CPPCTYPE> gx(hi + 1), gy(hi + 1);
        std::vector<double> gr(hi + 1);
        for (int k = l; k < hi; ++k) {  // QR: rotate rows k, k+1 so that h[k+1][k] = 0
            const CPPCTYPE x = h[k][k], y = h[k + 1][k];
            const double r = std::sqrt(std::norm(x) + std::norm(y));
            gx[k] = x, gy[k] = y, gr[k] = r;
            if (r == 0.0) continue;
            for (int j = k; j <= hi; ++j) {
                const CPPCTYPE u = h[k][j], v = h[k + 1][j];
                h[k][j] = (std::conj(x) * u + std::conj(y) * v) / r;
                h[k + 1][j] = (-y * u + x * v) / r;
            }
        }
        for (int k = l; k < hi; ++k) {  // RQ: the same rotations from the right
            const double r = gr[k];
            if (r == 0.0) continue;
            const CPPCTYPE x = gx[k], y = gy[k];
            for (int i = l; i <= std::min(k + 2, hi); ++i) {
                const CPPCTYPE u = h[i][k], v = h[i][k + 1];
                h[i][k] = (u * x + v * y) / r;
                h[i][k + 1] = (-u * std::conj(y) + v * std::conj(x)) / r;
            }
        }
        for (int k = l; k <= hi; ++k) h[k][k] += mu;
    }
    return ev;
}

// Unit eigenvector of `a` for the eigenvalue `lambda` by inverse iteration with a
// partially pivoted LU of (a - lambda I).
std::vector<CPPCTYPE> eigenvector_for(const CMat& a, CPPCTYPE lambda) {
    const int n = (int)a.size();
    double scale = 0.0;
    for (const auto& row : a)
        for (const auto& v : row) scale = std::max(scale, std::abs(v));
    const double tiny = std::max(scale, 1.0) * 1e-14;
    CMat m = a;
    for (int i = 0; i < n; ++i) m[i][i] -= lambda;
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::abs(m[i][k]) > std::abs(m[p][k])) p = i;
        std::swap(m[k], m[p]);
        piv[k] = p;
        if (std::abs(m[k][k]) < tiny) m[k][k] = tiny;
        for (int i = k + 1; i < n; ++i) {
            const CPPCTYPE f = m[i][k] / m[k][k];
            m[i][k] = f;
            for (int j = k + 1; j < n; ++j) m[i][j] -= f * m[k][j];
        }
    }
    std::vector<CPPCTYPE> v(n, 1.0);
    for (int it = 0; it < 3; ++it) {
        for (int k = 0; k < n; ++k) std::swap(v[k], v[piv[k]]);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) v[i] -= m[i][j] * v[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) v[i] -= m[i][j] * v[j];
            v[i] /= m[i][i];
        }
        double nrm = 0.0;
        for (const auto& x : v) nrm += std::norm(x);
        nrm = std::sqrt(nrm);
        for (auto& x : v) x /= nrm;
    }
    return v;
}

// A = V diag(w) V^-1 for a general complex matrix: Householder reduction to Hessenberg
// form, shifted QR to the complex Schur form T = Q^H A Q with Q accumulated, the
// eigenvectors of the triangular T by back substitution (a near-zero pivot from a
// repeated eigenvalue is replaced by a small one), V = Q X with unit columns.
void complex_eigen(const CMat& a, std::vector<CPPCTYPE>* w, CMat* vout) {
    const int n = (int)a.size();
    CMat h = a, q(n, std::vector<CPPCTYPE>(n, 0.0));
    for (int i = 0; i < n; ++i) q[i][i] = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    double hnorm = 0.0;
    for (const auto& row : a)
        for (const auto& z : row) hnorm = std::max(hnorm, std::abs(z));
    // Householder: zero h[k+2..][k]
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;
        std::vector<CPPCTYPE> v(m);
        double xn = 0.0;
        for (int i = 0; i < m; ++i) {
            v[i] = h[k + 1 + i][k];
            xn += std::norm(v[i]);
        }
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        const CPPCTYPE ph = std::abs(v[0]) > 0 ? v[0] / std::abs(v[0]) : CPPCTYPE(1.0);
        v[0] += ph * xn;  // v = x + e^{i arg x0} |x| e1
        double vn = 0.0;
        for (auto& z : v) vn += std::norm(z);
        vn = std::sqrt(vn);
        for (auto& z : v) z /= vn;
        for (int j = 0; j < n; ++j) {  // h <- (I - 2 v v^H) h on rows k+1..
            CPPCTYPE d = 0.0;
            for (int i = 0; i < m; ++i) d += std::conj(v[i]) * h[k + 1 + i][j];
            for (int i = 0; i < m; ++i) h[k + 1 + i][j] -= 2.0 * v[i] * d;
        }
        for (int i = 0; i < n; ++i) {  // h <- h (I - 2 v v^H), q <- q (I - 2 v v^H) on columns k+1..
            CPPCTYPE d = 0.0, e = 0.0;
            for (int j = 0; j < m; ++j) {
                d += h[i][k + 1 + j] * v[j];
                e += q[i][k + 1 + j] * v[j];
            }
            for (int j = 0; j < m; ++j) {
                h[i][k + 1 + j] -= 2.0 * d * std::conj(v[j]);
                q[i][k + 1 + j] -= 2.0 * e * std::conj(v[j]);
            }
        }
        for (int i = k + 2; i < n; ++i) h[i][k] = 0.0;
    }
    // shifted QR with the rotations applied to whole rows / columns and to q
    int hi = n - 1, iter = 0;
    while (hi > 0) {
        int l = hi;
        for (; l > 0; --l)
            if (std::abs(h[l][l - 1]) <= eps * std::max(std::abs(h[l - 1][l - 1]) + std::abs(h[l][l]), eps * hnorm)) {
                h[l][l - 1] = 0.0;
                break;
            }
        if (l == hi) {
            --hi;
            iter = 0;
            continue;
        }
        if (++iter > 1000) throw QulacsError("eigen decomposition did not converge");
        const CPPCTYPE A = h[hi - 1][hi - 1], B = h[hi - 1][hi], C = h[hi][hi - 1], D = h[hi][hi];
        const CPPCTYPE half = 0.5 * (A - D), disc = std::sqrt(half * half + B * C);
        CPPCTYPE mu = D - B * C / (half + (std::abs(half + disc) >= std::abs(half - disc) ? disc : -disc));
        if (!std::isfinite(mu.real()) || !std::isfinite(mu.imag())) mu = D;
        if (iter % 11 == 10) mu = D + std::abs(C);
        for (int k = l; k <= hi; ++k) h[k][k] -= mu;
        std::vector<CPPCTYPE> gx(n), gy(n);
        std::vector<double> gr(n, 0.0);
        for (int k = l; k < hi; ++k) {
            const CPPCTYPE x = h[k][k], y = h[k + 1][k];
            const double r = std::sqrt(std::norm(x) + std::norm(y));
            gx[k] = x, gy[k] = y, gr[k] = r;
            if (r == 0.0) continue;
            for (int j = k; j < n; ++j) {
                const CPPCTYPE u = h[k][j], v = h[k + 1][j];
                h[k][j] = (std::conj(x) * u + std::conj(y) * v) / r;
                h[k + 1][j] = (-y * u + x * v) / r;
            }
        }
        for (int k = l; k < hi; ++k) {
            const double r = gr[k];
            if (r == 0.0) continue;
            const CPPCTYPE x = gx[k], y = gy[k];
            for (int i = 0; i <= std::min(k + 2, hi); ++i) {
                const CPPCTYPE u = h[i][k], v = h[i][k + 1];
                h[i][k] = (u * x + v * y) / r;
                h[i][k + 1] = (-u * std::conj(y) + v * std::conj(x)) / r;
            }
            for (int i = 0; i < n; ++i) {
                const CPPCTYPE u = q[i][k], v = q[i][k + 1];
                q[i][k] = (u * x + v * y) / r;
                q[i][k + 1] = (-u * std::conj(y) + v * std::conj(x)) / r;
            }
        }
        for (int k = l; k <= hi; ++k) h[k][k] += mu;
    }
    w->resize(n);
    for (int i = 0; i < n; ++i) (*w)[i] = h[i][i];
    // eigenvectors of the upper triangular h, then V = q x
    const double smin = std::max(eps * hnorm, std::numeric_limits<double>::min());
    vout->assign(n, std::vector<CPPCTYPE>(n, 0.0));
    std::vector<CPPCTYPE> x(n);
    for (int k = 0; k < n; ++k) {
        std::fill(x.begin(), x.end(), CPPCTYPE(0.0));
        x[k] = 1.0;
        for (int i = k - 1; i >= 0; --i) {
            CPPCTYPE sum = 0.0;
            for (int j = i + 1; j <= k; ++j) sum += h[i][j] * x[j];
            CPPCTYPE den = h[i][i] - h[k][k];
            if (std::abs(den) < smin) den = smin;
            x[i] = -sum / den;
        }
        double nrm = 0.0;
        std::vector<CPPCTYPE> col(n, 0.0);
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j <= k; ++j) col[i] += q[i][j] * x[j];
            nrm += std::norm(col[i]);
        }
        nrm = std::sqrt(nrm);
        for (int i = 0; i < n; ++i) (*vout)[i][k] = col[i] / nrm;
    }
}

// inverse by Gauss-Jordan elimination with partial pivoting
CMat invert(CMat a) {
    const int n = (int)a.size();
    CMat inv(n, std::vector<CPPCTYPE>(n, 0.0));
    for (int i = 0; i < n; ++i) inv[i][i] = 1.0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::abs(a[i][k]) > std::abs(a[p][k])) p = i;
        if (std::abs(a[p][k]) == 0.0) throw QulacsError("matrix inverse: singular matrix");
        std::swap(a[k], a[p]);
        std::swap(inv[k], inv[p]);
        const CPPCTYPE d = 1.0 / a[k][k];
        for (int j = 0; j < n; ++j) {
            a[k][j] *= d;
            inv[k][j] *= d;
        }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const CPPCTYPE f = a[i][k];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) {
                a[i][j] -= f * a[k][j];
                inv[i][j] -= f * inv[k][j];
            }
        }
    }
    return inv;
}

// Eigen-decomposition of the symmetric tridiagonal matrix (alpha on the diagonal, beta
// beside it) by cyclic Jacobi rotations; column k of `vec` belongs to val[k].
void tridiagonal_eigen(const std::vector<double>& alpha, const std::vector<double>& beta, std::vector<double>* val,
                       std::vector<std::vector<double>>* vec) {
    const int n = (int)alpha.size();
    std::vector<std::vector<double>> a(n, std::vector<double>(n, 0.0)), v(n, std::vector<double>(n, 0.0));
    for (int i = 0; i < n; ++i) {
        a[i][i] = alpha[i];
        v[i][i] = 1.0;
        if (i + 1 < n) a[i][i + 1] = a[i + 1][i] = beta[i];
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, total = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) (i == j ? total : off) += a[i][j] * a[i][j];
        if (off <= 1e-30 * (total + off)) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    val->resize(n);
    for (int i = 0; i < n; ++i) (*val)[i] = a[i][i];
    *vec = v;
}

void require_terms(const GeneralQuantumOperator& o, const char* who) {
    if (o.get_term_count() == 0) throw QulacsError(std::string(who) + ": at least one PauliOperator is required");
}

}  // namespace

// src/cppsim/general_quantum_operator.cpp:290-328. The value returned adds the `mu`
// argument (not the default shift) to <A>, as the reference does.
CPPCTYPE GeneralQuantumOperator::solve_ground_state_eigenvalue_by_power_method(QuantumStateMI355* state,
                                                                               unsigned iter_count,
                                                                               CPPCTYPE mu) const {
    require_terms(*this, "solve_ground_state_eigenvalue_by_power_method");
    const CPPCTYPE shift = (mu == 0.0) ? calculate_default_mu() : mu;
    QuantumStateMI355 multiplied(state->qubit_count, state->device_number);
    QuantumStateMI355 work(state->qubit_count, state->device_number);
    for (unsigned i = 0; i < iter_count; ++i) {
        apply_to_state(&work, *state, &multiplied);  // A q
        multiplied.add_state_with_coef(-shift, state);  // (A - mu) q
        state->load(&multiplied);
        state->normalize(state->get_squared_norm());
    }
    return get_expectation_value(state) + mu;
}

// Arnoldi: orthonormal Krylov basis q_0..q_m of (A - mu), Hessenberg H = Q^* (A - mu) Q,
// ground eigenvalue = the eigenvalue of H with the smallest real part, plus mu
// (src/cppsim/general_quantum_operator.cpp:204-288).
CPPCTYPE GeneralQuantumOperator::solve_ground_state_eigenvalue_by_arnoldi_method(QuantumStateMI355* state,
                                                                                 unsigned iter_count,
                                                                                 CPPCTYPE mu) const {
    require_terms(*this, "solve_ground_state_eigenvalue_by_arnoldi_method");
    if (iter_count == 0) throw QulacsError("solve_ground_state_eigenvalue_by_arnoldi_method: iter_count must be > 0");
    const unsigned n = state->qubit_count, dev = state->device_number;
    const CPPCTYPE shift = (mu == 0.0) ? calculate_default_mu() : mu;
    std::vector<std::unique_ptr<QuantumStateMI355>> basis;
    state->normalize(state->get_squared_norm());
    basis.emplace_back(new QuantumStateMI355(n, dev));
    basis.back()->load(state);
    QuantumStateMI355 work(n, dev);
    CMat h(iter_count, std::vector<CPPCTYPE>(iter_count, 0.0));
    for (unsigned i = 0; i < iter_count; ++i) {
        auto v = std::make_unique<QuantumStateMI355>(n, dev);
        apply_to_state(&work, *basis[i], v.get());
        v->add_state_with_coef(-shift, basis[i].get());
        for (unsigned j = 0; j <= i; ++j) {
            const CPPCTYPE c = inner_product(basis[j].get(), v.get());
            h[j][i] = c;
            v->add_state_with_coef(-c, basis[j].get());
        }
        const double norm = v->get_squared_norm();
        if (i + 1 != iter_count) h[i + 1][i] = std::sqrt(norm);
        v->normalize(norm);
        basis.push_back(std::move(v));
    }
    const std::vector<CPPCTYPE> ev = hessenberg_eigenvalues(h);
    size_t best = 0;
    for (size_t i = 0; i < ev.size(); ++i)
        if (ev[i].real() < ev[best].real()) best = i;
    const std::vector<CPPCTYPE> y = eigenvector_for(h, ev[best]);
    QuantumStateMI355 acc(n, dev);
    acc.multiply_coef(0.0);
    for (unsigned i = 0; i < iter_count; ++i) acc.add_state_with_coef(y[i], basis[i].get());
    state->load(&acc);
    return ev[best] + shift;
}

// Lanczos three-term recurrence on (A - mu); the tridiagonal matrix's lowest eigenpair
// gives the eigenvalue and, by re-running the recurrence, the eigenvector
// (src/cppsim/observable.cpp:61-190).
CPPCTYPE Observable::solve_ground_state_eigenvalue_by_lanczos_method(QuantumStateMI355* state, unsigned iter_count,
                                                                     CPPCTYPE mu) const {
    require_terms(*this, "solve_ground_state_eigenvalue_by_lanczos_method");
    if (iter_count == 0) throw QulacsError("solve_ground_state_eigenvalue_by_lanczos_method: iter_count must be > 0");
    const unsigned n = state->qubit_count, dev = state->device_number;
    const CPPCTYPE shift = (mu == 0.0) ? calculate_default_mu() : mu;
    QuantumStateMI355 work(n, dev), q_prev(n, dev), q(n, dev), v(n, dev);
    state->normalize(state->get_squared_norm());
    std::vector<double> alpha(iter_count), beta(iter_count > 1 ? iter_count - 1 : 0);
    auto recurrence = [&](bool measure, unsigned i) {
        // v = (A - mu) q_i - alpha_i q_i - beta_{i-1} q_{i-1}; q_{i+1} = v / beta_i
        apply_to_state(&work, q, &v);
        v.add_state_with_coef(-shift, &q);
        if (measure) alpha[i] = inner_product(&q, &v).real();
        if (i + 1 == iter_count) return false;
        v.add_state_with_coef(-alpha[i], &q);
        if (i != 0) v.add_state_with_coef(-beta[i - 1], &q_prev);
        if (measure) beta[i] = std::sqrt(v.get_squared_norm());
        v.multiply_coef(1.0 / beta[i]);
        q_prev.load(&q);
        q.load(&v);
        return true;
    };
    q.load(state);
    for (unsigned i = 0; i < iter_count && recurrence(true, i); ++i) {
    }
    std::vector<double> val;
    std::vector<std::vector<double>> vec;
    tridiagonal_eigen(alpha, beta, &val, &vec);
    size_t best = 0;
    for (size_t i = 0; i < val.size(); ++i)
        if (val[i] < val[best]) best = i;
    // eigenvector = sum_i vec[i][best] q_i, with the q_i regenerated from the start vector
    q.load(state);
    state->multiply_coef(0.0);
    for (unsigned i = 0; i < iter_count; ++i) {
        state->add_state_with_coef(vec[i][best], &q);
        if (!recurrence(false, i)) break;
    }
    return val[best] + shift;
}

GeneralQuantumOperator* to_general_quantum_operator(const QuantumGateBase* g, unsigned qubits, double tol) {
    if (!g->get_control_index_list().empty()) throw QulacsError("to_general_quantum_operator: gate must not have controls");
    const ComplexMatrix u = g->get_matrix();
    const auto& tl = g->get_target_index_list();
    const unsigned k = (unsigned)tl.size();
    const size_t d = size_t(1) << k;
    auto op = std::make_unique<GeneralQuantumOperator>(qubits);
    const CPPCTYPE ipow[4] = {1.0, CPPCTYPE(0, 1), -1.0, CPPCTYPE(0, -1)};
    for (size_t i = 0; i < d; ++i)
        for (size_t j = 0; j < d; ++j) {
            // string: bit h of (i, j) = (0,0) I, (0,1) X, (1,0) Y, (1,1) Z
            size_t flip = 0, phase = 0;
            unsigned ny = 0;
            std::vector<unsigned> idx, pid;
            for (unsigned h = 0; h < k; ++h) {
                const bool a = (i >> h) & 1, b = (j >> h) & 1;
                if (!a && !b) continue;
                const unsigned p = a ? (b ? 3u : 2u) : 1u;
                idx.push_back(tl[h]);
                pid.push_back(p);
                if (p != 3) flip |= size_t(1) << h;
                if (p != 1) phase |= size_t(1) << h;
                ny += p == 2;
            }
            // tr(P U) = sum_y <y ^ flip| P |y> U[y][y ^ flip], P|y> = i^#Y (-1)^|y & phase| |y ^ flip>
            CPPCTYPE tr = 0.0;
            for (size_t y = 0; y < d; ++y) {
                const double sgn = (__builtin_popcountll(y & phase) & 1) ? -1.0 : 1.0;
                tr += sgn * u(y, y ^ flip);
            }
            tr *= ipow[ny % 4];
            if (std::abs(tr) <= tol) continue;
            op->add_operator(PauliOperator(idx, pid, tr / (double)d));
        }
    return op.release();
}

// ------------------------------------------------------------------ noisy evolution
ClsNoisyEvolution::ClsNoisyEvolution(const Observable& h, const std::vector<const GeneralQuantumOperator*>& c_ops,
                                     double time, double dt)
    : h_(new Observable(h.get_qubit_count())), time_(time), dt_(dt) {
    if (!(dt > 0) || !(time >= 0)) throw QulacsError("NoisyEvolution: time must be >= 0 and dt > 0");
    for (size_t i = 0; i < h.get_term_count(); ++i) h_->add_operator(h.get_term(i));
    // H_eff = H - (i/2) sum_k c_k^dagger c_k, as a general operator
    heff_.reset(new GeneralQuantumOperator(h.get_qubit_count()));
    for (size_t i = 0; i < h.get_term_count(); ++i) heff_->add_operator(h.get_term(i));
    for (const auto* c : c_ops) {
        c_ops_.emplace_back(new GeneralQuantumOperator(*c));
        *heff_ += c->get_dagger() * *c * CPPCTYPE(0.0, -0.5);
    }
    name_ = "NoisyEvolution";
}
ClsNoisyEvolution::ClsNoisyEvolution(const ClsNoisyEvolution& o)
    : QuantumGateBase(o),
      h_(new Observable(o.h_->get_qubit_count())),
      heff_(new GeneralQuantumOperator(*o.heff_)),
      time_(o.time_),
      dt_(o.dt_),
      norm_tol_(o.norm_tol_),
      max_steps_(o.max_steps_),
      random_(o.random_) {
    for (size_t i = 0; i < o.h_->get_term_count(); ++i) h_->add_operator(o.h_->get_term(i));
    for (const auto& c : o.c_ops_) c_ops_.emplace_back(new GeneralQuantumOperator(*c));
}
ComplexMatrix ClsNoisyEvolution::get_matrix() const { throw QulacsError("NoisyEvolution has no matrix"); }
// one RK4 step of d psi / dt = -i H_eff psi; `buffer` ends holding the state before the step
void ClsNoisyEvolution::evolve(QuantumStateMI355* k1, QuantumStateMI355* k2, QuantumStateMI355* k3,
                               QuantumStateMI355* k4, QuantumStateMI355* buffer, QuantumStateMI355* state,
                               double dt) const {
    const CPPCTYPE mi(0.0, -1.0);
    heff_->apply_to_state(state, k1);
    buffer->load(state);
    buffer->add_state_with_coef(mi * (dt / 2), k1);
    heff_->apply_to_state(buffer, k2);
    buffer->load(state);
    buffer->add_state_with_coef(mi * (dt / 2), k2);
    heff_->apply_to_state(buffer, k3);
    buffer->load(state);
    buffer->add_state_with_coef(mi * dt, k3);
    heff_->apply_to_state(buffer, k4);
    buffer->load(state);
    state->add_state_with_coef(mi * (dt / 6), k1);
    state->add_state_with_coef(mi * (dt / 3), k2);
    state->add_state_with_coef(mi * (dt / 3), k3);
    state->add_state_with_coef(mi * (dt / 6), k4);
}
// time in [0, dt] at which the squared norm reaches `target`, bracketed between `prev`
// (norm >= target) and `now` (norm <= target): secant steps on log(norm) for the first
// 21 tries, bisection after; `now` ends holding the state at that time
// (src/cppsim/gate_noisy_evolution.cpp:23-104)
double ClsNoisyEvolution::find_collapse(QuantumStateMI355* k1, QuantumStateMI355* k2, QuantumStateMI355* k3,
                                        QuantumStateMI355* k4, QuantumStateMI355* prev, QuantumStateMI355* now,
                                        double target, double dt) const {
    double lo_norm = prev->get_squared_norm(), hi_norm = now->get_squared_norm();
    if (std::abs(lo_norm - target) < norm_tol_) {
        now->load(prev);
        return 0.0;
    }
    if (std::abs(hi_norm - target) < norm_tol_) return dt;
    if (lo_norm < target) throw QulacsError("NoisyEvolution: the previous norm is below the jump threshold");
    if (hi_norm > target) throw QulacsError("NoisyEvolution: the current norm is above the jump threshold");
    double t_lo = 0.0, t_hi = dt;
    const double lt = std::log(target);
    double llo = std::log(lo_norm), lhi = std::log(hi_norm);
    QuantumStateMI355 buf(prev->qubit_count, prev->device_number), scratch(prev->qubit_count, prev->device_number);
    for (int step = 0; step < max_steps_; ++step) {
        const double t = step <= 20 ? t_lo + (t_hi - t_lo) * (llo - lt) / (llo - lhi) : 0.5 * (t_lo + t_hi);
        buf.load(prev);
        evolve(k1, k2, k3, k4, &scratch, &buf, t);
        const double nb = buf.get_squared_norm();
        if (std::abs(nb - target) < norm_tol_) {
            now->load(&buf);
            return t;
        }
        if (nb < target) {
            now->load(&buf);
            t_hi = t;
            lhi = std::log(nb);
        } else {
            t_lo = t;
            llo = std::log(nb);
        }
    }
    throw QulacsError("NoisyEvolution: failed to find the jump time; try a smaller dt");
}
void ClsNoisyEvolution::update_quantum_state(QuantumStateBase* state_base) {
    auto* state = as_sv(state_base);
    if (!state) throw QulacsError("NoisyEvolution: density matrices are not supported");
    if (h_->get_qubit_count() > state->qubit_count) throw QulacsError("NoisyEvolution: operator beyond the state");
    const unsigned n = state->qubit_count, dev = state->device_number;
    const double norm0 = state->get_squared_norm();
    double r = random_.uniform();
    QuantumStateMI355 k1(n, dev), k2(n, dev), k3(n, dev), k4(n, dev), buffer(n, dev);
    std::vector<double> cum(c_ops_.size());
    double t = 0.0;
    while (std::abs(t - time_) > 1e-10 * time_) {
        const double dt = (t + dt_ > time_) ? time_ - t : dt_;
        evolve(&k1, &k2, &k3, &k4, &buffer, state, dt);
        if (state->get_squared_norm() <= r) {  // a jump inside this step
            const double tj = find_collapse(&k1, &k2, &k3, &k4, &buffer, state, r, dt);
            double sum = 0.0;
            for (size_t k = 0; k < c_ops_.size(); ++k) {
                c_ops_[k]->apply_to_state(state, &buffer);
                sum += buffer.get_squared_norm();
                cum[k] = sum;
            }
            const double jr = random_.uniform() * sum;
            const size_t idx = (size_t)(std::lower_bound(cum.begin(), cum.end(), jr) - cum.begin());
            if (idx < c_ops_.size()) {
                c_ops_[idx]->apply_to_state(state, &buffer);
                buffer.normalize(buffer.get_squared_norm());
                state->load(&buffer);
            }
            t += tj;
            r = random_.uniform();
        } else {
            t += dt;
        }
    }
    state->normalize(state->get_squared_norm() / norm0);
}
ClsNoisyEvolution_fast::ClsNoisyEvolution_fast(const Observable& h,
                                               const std::vector<const GeneralQuantumOperator*>& c_ops, double time)
    : h_(new Observable(h.get_qubit_count())), time_(time) {
    if (!(time >= 0)) throw QulacsError("NoisyEvolution_fast: time must be >= 0");
    for (size_t i = 0; i < h.get_term_count(); ++i) h_->add_operator(h.get_term(i));
    heff_.reset(new GeneralQuantumOperator(h.get_qubit_count()));
    for (size_t i = 0; i < h.get_term_count(); ++i) heff_->add_operator(h.get_term(i));
    for (const auto* c : c_ops) {
        c_ops_.emplace_back(new GeneralQuantumOperator(*c));
        *heff_ += c->get_dagger() * *c * CPPCTYPE(0.0, -0.5);
    }
    name_ = "NoisyEvolution_fast";
    if (heff_->get_term_count() == 0) {
        none_ = true;
        return;
    }
    // -i H_eff as one matrix over the qubits its terms touch
    std::vector<std::unique_ptr<QuantumGateBase>> owned;
    std::vector<QuantumGateBase*> parts;
    for (size_t i = 0; i < heff_->get_term_count(); ++i) {
        const PauliOperator& t = heff_->get_term(i);
        std::unique_ptr<ClsPauliGate> pg(gate::Pauli(t.get_index_list(), t.get_pauli_id_list()));
        QuantumGateMatrix* m = gate::to_matrix_gate(pg.get());
        m->multiply_scalar(-t.get_coef() * CPPCTYPE(0.0, 1.0));
        owned.emplace_back(m);
        parts.push_back(m);
    }
    std::vector<std::vector<unsigned>> lists;
    for (auto* g : parts) lists.push_back(g->get_target_index_list());
    if (sorted_union(lists).size() > 8)
        throw QulacsError("NoisyEvolution_fast: the operators touch more than 8 qubits; split independent parts");
    std::unique_ptr<QuantumGateMatrix> a(gate::add(parts));
    targets_ = a->get_target_index_list();
    const ComplexMatrix am = a->get_matrix();
    const size_t d = am.rows;
    CMat m(d, std::vector<CPPCTYPE>(d));
    for (size_t i = 0; i < d; ++i)
        for (size_t j = 0; j < d; ++j) m[i][j] = am(i, j);
    CMat v;
    complex_eigen(m, &eig_, &v);
    const CMat vi = invert(v);
    ComplexMatrix V(d, d), Vi(d, d);
    for (size_t i = 0; i < d; ++i)
        for (size_t j = 0; j < d; ++j) V(i, j) = v[i][j], Vi(i, j) = vi[i][j];
    v_.reset(new QuantumGateMatrix(targets_, V));
    vinv_.reset(new QuantumGateMatrix(targets_, Vi));
}
ClsNoisyEvolution_fast::ClsNoisyEvolution_fast(const ClsNoisyEvolution_fast& o)
    : QuantumGateBase(o),
      h_(new Observable(o.h_->get_qubit_count())),
      heff_(new GeneralQuantumOperator(*o.heff_)),
      time_(o.time_),
      norm_tol_(o.norm_tol_),
      max_steps_(o.max_steps_),
      random_(o.random_),
      none_(o.none_),
      eig_(o.eig_),
      v_(o.v_ ? new QuantumGateMatrix(*o.v_) : nullptr),
      vinv_(o.vinv_ ? new QuantumGateMatrix(*o.vinv_) : nullptr) {
    for (size_t i = 0; i < o.h_->get_term_count(); ++i) h_->add_operator(o.h_->get_term(i));
    for (const auto& c : o.c_ops_) c_ops_.emplace_back(new GeneralQuantumOperator(*c));
}
ComplexMatrix ClsNoisyEvolution_fast::get_matrix() const { throw QulacsError("NoisyEvolution_fast has no matrix"); }
void ClsNoisyEvolution_fast::evolve(QuantumStateMI355* s, double t) const {
    vinv_->update_quantum_state(s);
    std::vector<CPPCTYPE> e(eig_.size());
    for (size_t i = 0; i < e.size(); ++i) e[i] = std::exp(eig_[i] * t);
    QuantumGateDiagonalMatrix(targets_, e).update_quantum_state(s);
    v_->update_quantum_state(s);
}
// bracketed search for the jump time as in the RK4 version; the secant estimate uses
// weights from log-norm distances, every third try from their square roots
// (src/cppsim/gate_noisy_evolution.cpp:332-420)
double ClsNoisyEvolution_fast::find_collapse(QuantumStateMI355* prev, QuantumStateMI355* now, double target,
                                             double t_step) const {
    double lo_norm = prev->get_squared_norm(), hi_norm = now->get_squared_norm();
    if (std::abs(lo_norm - target) < norm_tol_) {
        now->load(prev);
        return 0.0;
    }
    if (std::abs(hi_norm - target) < norm_tol_) return t_step;
    if (lo_norm < target) throw QulacsError("NoisyEvolution_fast: the previous norm is below the jump threshold");
    if (hi_norm > target) throw QulacsError("NoisyEvolution_fast: the current norm is above the jump threshold");
    double t_lo = 0.0, t_hi = t_step;
    const double lt = std::log(target);
    double llo = std::log(lo_norm), lhi = std::log(hi_norm);
    QuantumStateMI355 buf(prev->qubit_count, prev->device_number);
    for (int step = 0; step < max_steps_; ++step) {
        double t;
        if (step <= 20) {
            const double a = (llo - lt) / (llo - lhi), b = (lt - lhi) / (llo - lhi);
            t = ((step - 2) % 3 != 2) ? t_lo + (t_hi - t_lo) * a / (a + b)
                                      : t_lo + (t_hi - t_lo) * std::sqrt(a) / (std::sqrt(a) + std::sqrt(b));
        } else {
            t = 0.5 * (t_lo + t_hi);
        }
        buf.load(prev);
        evolve(&buf, t);
        const double nb = buf.get_squared_norm();
        if (std::abs(nb - target) < norm_tol_) {
            now->load(&buf);
            return t;
        }
        if (nb < target) {
            now->load(&buf);
            t_hi = t;
            lhi = std::log(nb);
        } else {
            t_lo = t;
            llo = std::log(nb);
        }
    }
    throw QulacsError("NoisyEvolution_fast: failed to find the jump time");
}
void ClsNoisyEvolution_fast::update_quantum_state(QuantumStateBase* state_base) {
    if (none_) return;
    auto* state = as_sv(state_base);
    if (!state) throw QulacsError("NoisyEvolution_fast: density matrices are not supported");
    check_targets(state, targets_, "NoisyEvolution_fast");
    const unsigned n = state->qubit_count, dev = state->device_number;
    const double norm0 = state->get_squared_norm();
    double r = random_.uniform();
    QuantumStateMI355 buffer(n, dev);
    std::vector<double> cum(c_ops_.size());
    double t = 0.0;
    while (std::abs(t - time_) > 1e-10 * time_) {
        const double step = time_ - t;
        buffer.load(state);
        evolve(state, step);
        if (state->get_squared_norm() <= r) {
            const double tj = find_collapse(&buffer, state, r, step);
            double sum = 0.0;
            for (size_t k = 0; k < c_ops_.size(); ++k) {
                c_ops_[k]->apply_to_state(state, &buffer);
                sum += buffer.get_squared_norm();
                cum[k] = sum;
            }
            const double jr = random_.uniform() * sum;
            const size_t idx = (size_t)(std::lower_bound(cum.begin(), cum.end(), jr) - cum.begin());
            if (idx < c_ops_.size()) {
                c_ops_[idx]->apply_to_state(state, &buffer);
                buffer.normalize(buffer.get_squared_norm());
                state->load(&buffer);
            }
            t += tj;
            r = random_.uniform();
        } else {
            t += step;
        }
    }
    state->normalize(state->get_squared_norm() / norm0);
}

namespace gate {
ClsNoisyEvolution* NoisyEvolution(const Observable& h, const std::vector<const GeneralQuantumOperator*>& c_ops,
                                  double time, double dt) {
    return new ClsNoisyEvolution(h, c_ops, time, dt);
}
ClsNoisyEvolution_fast* NoisyEvolution_fast(const Observable& h, const std::vector<const GeneralQuantumOperator*>& c_ops,
                                            double time) {
    return new ClsNoisyEvolution_fast(h, c_ops, time);
}
}  // namespace gate

// ------------------------------------------------------------------ noise simulator
NoiseSimulator::Result::~Result() {
    for (auto& p : result) delete p.first;
}
std::vector<ITYPE> NoiseSimulator::Result::sampling() const {
    std::vector<ITYPE> out;
    for (const auto& p : result) {
        const std::vector<ITYPE> s = p.first->sampling(p.second);
        out.insert(out.end(), s.begin(), s.end());
    }
    return out;
}
NoiseSimulator::NoiseSimulator(const QuantumCircuit* c, const QuantumStateMI355* init)
    : circuit_(c->copy()),
      initial_(new QuantumStateMI355(c->get_qubit_count(), init ? init->device_number : 0)) {
    if (init) {
        if (init->qubit_count != c->get_qubit_count()) throw QulacsError("NoiseSimulator: qubit counts differ");
        initial_->load(init);
    } else {
        initial_->set_zero_state();
    }
    for (auto* g : circuit_->gates_) {
        if (auto* p = dynamic_cast<QuantumGate_Probabilistic*>(g)) p->optimize_ProbablisticGate();
        else if (g->is_noise())
            throw QulacsError("NoiseSimulator: only probabilistic noise gates can be sampled (got " + g->get_name() + ")");
    }
}
// branch by lower_bound on the cumulative distribution (src/cppsim/noisesimulator.cpp:164-182);
// the index past the last gate is the implicit identity
unsigned NoiseSimulator::choose_branch(const QuantumGateBase* g) {
    const auto* p = dynamic_cast<const QuantumGate_Probabilistic*>(g);
    if (!p) return 0;
    const auto& cum = p->get_cumulative_distribution();
    const double r = random_.uniform();
    const long i = (long)(std::lower_bound(cum.begin(), cum.end(), r) - cum.begin()) - 1;
    return (unsigned)std::max(0L, i);
}
// Gates [start, end) on s; branch[c] is the drawn branch of the c-th probabilistic gate,
// `col` the column of the first probabilistic gate at or after `start`.
void NoiseSimulator::apply_from(const std::vector<unsigned>& branch, QuantumStateMI355* s, size_t start,
                                size_t col) const {
    for (size_t k = start; k < circuit_->gates_.size(); ++k) {
        const QuantumGateBase* g = circuit_->gates_[k];
        if (const auto* p = dynamic_cast<const QuantumGate_Probabilistic*>(g)) {
            const unsigned b = branch[col++];
            if (b < p->get_gate_list().size()) p->get_gate_list()[b]->update_quantum_state(s);
        } else {
            circuit_->gates_[k]->update_quantum_state(s);
        }
    }
}
std::vector<ITYPE> NoiseSimulator::execute(unsigned count) {
    std::unique_ptr<Result> r(execute_and_get_result(count));
    return r->sampling();
}
NoiseSimulator::Result* NoiseSimulator::execute_and_get_result(unsigned count) {
    const size_t ng = circuit_->gates_.size();
    std::vector<size_t> noisy;  // positions of the probabilistic gates: one column each
    for (size_t k = 0; k < ng; ++k)
        if (dynamic_cast<const QuantumGate_Probabilistic*>(circuit_->gates_[k])) noisy.push_back(k);
    std::vector<std::vector<unsigned>> pick(count, std::vector<unsigned>(noisy.size(), 0));
    for (auto& row : pick)
        for (size_t c = 0; c < noisy.size(); ++c) row[c] = choose_branch(circuit_->gates_[noisy[c]]);
    // shots with the same branches share one trajectory; trajectories in descending
    // lexicographic order, so those starting with branch 0 share the longest prefix
    std::sort(pick.begin(), pick.end());
    std::reverse(pick.begin(), pick.end());
    std::vector<std::pair<std::vector<unsigned>, unsigned>> req;
    for (size_t i = 0; i < pick.size(); ++i) {
        if (!req.empty() && req.back().first == pick[i]) ++req.back().second;
        else req.emplace_back(pick[i], 1u);
    }
    pick.clear();
    auto out = std::make_unique<Result>();
    QuantumStateMI355 common(initial_->qubit_count, initial_->device_number);
    common.load(initial_.get());
    size_t done = 0, col = 0;  // gates [0, done) are applied to `common`; col = probabilistic gates among them
    for (const auto& rq : req) {
        const auto& br = rq.first;
        while (done < ng) {
            const QuantumGateBase* g = circuit_->gates_[done];
            if (const auto* p = dynamic_cast<const QuantumGate_Probabilistic*>(g)) {
                if (br[col] != 0) break;
                if (!p->get_gate_list().empty()) p->get_gate_list()[0]->update_quantum_state(&common);
                ++col;
            } else {
                circuit_->gates_[done]->update_quantum_state(&common);
            }
            ++done;
        }
        auto* st = new QuantumStateMI355(common.qubit_count, common.device_number);
        st->load(&common);
        out->result.emplace_back(st, rq.second);
        apply_from(br, st, done, col);
    }
    return out.release();
}

// ------------------------------------------------------------------ causal cone simulator
namespace {
struct DisjointSets {
    std::vector<int> up;  // negative size at a root
    explicit DisjointSets(size_t n) : up(n, -1) {}
    int root(int a) { return up[a] < 0 ? a : (up[a] = root(up[a])); }
    int size(int a) { return -up[root(a)]; }
    void join(int a, int b) {
        a = root(a), b = root(b);
        if (a == b) return;
        if (size(a) < size(b)) std::swap(a, b);
        up[a] += up[b];
        up[b] = a;
    }
};
}  // namespace

CausalConeSimulator::CausalConeSimulator(const ParametricQuantumCircuit& c, const Observable& o)
    : circuit_(static_cast<ParametricQuantumCircuit*>(c.copy())), observable_(new Observable(o.get_qubit_count())) {
    for (size_t i = 0; i < o.get_term_count(); ++i) observable_->add_operator(o.get_term(i));
}
CausalConeSimulator::~CausalConeSimulator() {
    for (auto& row : circuits_)
        for (auto* c : row) delete c;
}
void CausalConeSimulator::build() {
    built_ = true;
    const unsigned n = circuit_->get_qubit_count();
    const size_t ng = circuit_->get_gate_count();
    for (size_t t = 0; t < observable_->get_term_count(); ++t) {
        const PauliOperator& term = observable_->get_term(t);
        DisjointSets ds(n);
        std::vector<char> qubit_used(n, 0), gate_used(ng, 0);
        for (unsigned q : term.get_index_list()) qubit_used.at(q) = 1;
        for (size_t k = ng; k-- > 0;) {  // the light cone, last gate first
            const QuantumGateBase* g = circuit_->get_gate(k);
            const auto &tl = g->get_target_index_list(), &cl = g->get_control_index_list();
            bool hit = false;
            for (unsigned q : tl) hit = hit || qubit_used[q];
            for (unsigned q : cl) hit = hit || qubit_used[q];
            if (!hit) continue;
            gate_used[k] = 1;
            for (unsigned q : tl) qubit_used[q] = 1;
            for (unsigned q : cl) qubit_used[q] = 1;
            for (size_t j = 1; j < tl.size(); ++j) ds.join(tl[j - 1], tl[j]);
            for (size_t j = 1; j < cl.size(); ++j) ds.join(cl[j - 1], cl[j]);
            if (!tl.empty() && !cl.empty()) ds.join(tl[0], cl[0]);
        }
        std::vector<int> roots;
        for (unsigned q = 0; q < n; ++q)
            if (qubit_used[q] && ds.root(q) == (int)q) roots.push_back((int)q);
        std::vector<ParametricQuantumCircuit*> row;
        std::vector<PauliOperator> prow;
        for (int root : roots) {
            std::vector<int> code(n, -1);
            unsigned m = 0;
            for (unsigned q = 0; q < n; ++q)
                if (ds.root(q) == root) code[q] = (int)m++;
            auto* sub = new ParametricQuantumCircuit(m);
            auto enc = [&](const std::vector<unsigned>& l) {
                std::vector<unsigned> r;
                for (unsigned q : l) r.push_back((unsigned)code[q]);
                return r;
            };
            for (size_t k = 0; k < ng; ++k) {
                if (!gate_used[k]) continue;
                const QuantumGateBase* g = circuit_->get_gate(k);
                if (ds.root(g->get_target_index_list()[0]) != root) continue;
                sub->add_gate_consume(g->create_gate_whose_qubit_indices_are_replaced(enc(g->get_target_index_list()),
                                                                                      enc(g->get_control_index_list())));
            }
            PauliOperator p(1.0);
            const auto &il = term.get_index_list(), &pl = term.get_pauli_id_list();
            for (size_t j = 0; j < il.size(); ++j)
                if (ds.root(il[j]) == root) p.add_single_Pauli((unsigned)code[il[j]], pl[j]);
            row.push_back(sub);
            prow.push_back(p);
        }
        circuits_.push_back(row);
        paulis_.push_back(prow);
        coefs_.push_back(term.get_coef());
    }
}
CPPCTYPE CausalConeSimulator::get_expectation_value() {
    if (!built_) build();
    CPPCTYPE sum = 0.0;
    for (size_t t = 0; t < circuits_.size(); ++t) {
        CPPCTYPE e = 1.0;
        for (size_t j = 0; j < circuits_[t].size(); ++j) {
            QuantumStateMI355 st(circuits_[t][j]->get_qubit_count());
            st.set_zero_state();
            circuits_[t][j]->update_quantum_state(&st);
            e *= paulis_[t][j].get_expectation_value(&st);
        }
        sum += e * coefs_[t];
    }
    return sum;
}

// ------------------------------------------------------------------ OpenFermion text
namespace {

// One line "coef [P0 P1 ...]": coef is "(re+imj)", "(re-imj)", "re" or "imj"; a
// missing or empty bracket is the identity "I0" (src/cppsim/utility.cpp:63-106).
// Returns false for a blank line.
bool parse_openfermion_line(const std::string& line, CPPCTYPE* coef, std::string* ops) {
    size_t p = line.find_first_not_of(" \t\r");
    if (p == std::string::npos) return false;
    const char* s = line.c_str() + p;
    char* end = nullptr;
    double re = 0.0, im = 0.0;
    auto bad = [&]() { return QulacsError("invalid OpenFermion term: " + line); };
    if (*s == '(') {
        re = std::strtod(s + 1, &end);
        if (end == s + 1) throw bad();
        const char* t = end;
        im = std::strtod(t, &end);
        if (end == t || *end != 'j' || end[1] != ')') throw bad();
        s = end + 2;
    } else {
        const double v = std::strtod(s, &end);
        if (end == s) throw bad();
        if (*end == 'j') {
            im = v;
            s = end + 1;
        } else {
            re = v;
            s = end;
        }
    }
    *coef = CPPCTYPE(re, im);
    std::string body;
    const char* lb = std::strchr(s, '[');
    if (lb) {
        const char* rb = std::strchr(lb, ']');
        if (!rb) throw bad();
        body.assign(lb + 1, rb);
    }
    if (body.find_first_not_of(" ") == std::string::npos) body = "I0";
    // "X0 Y1" -> "X 0 Y 1" (src/cppsim/utility.cpp:56-62)
    ops->clear();
    for (char ch : body) {
        ops->push_back(ch);
        if (ch == 'X' || ch == 'Y' || ch == 'Z' || ch == 'I') ops->push_back(' ');
    }
    return true;
}

struct OpenFermionTerms {
    unsigned qubit_count = 0;
    std::vector<CPPCTYPE> coefs;
    std::vector<std::string> ops;
};

OpenFermionTerms parse_openfermion_lines(std::istream& in) {
    OpenFermionTerms r;
    std::string line, ops;
    CPPCTYPE coef;
    while (std::getline(in, line)) {
        if (!parse_openfermion_line(line, &coef, &ops)) continue;
        std::string idx;
        auto flush = [&]() {
            if (!idx.empty()) r.qubit_count = std::max(r.qubit_count, (unsigned)std::stoul(idx) + 1);
            idx.clear();
        };
        for (char ch : ops) {
            if (std::strchr("IXYZ ", ch)) flush();
            else idx.push_back(ch);
        }
        flush();
        r.coefs.push_back(coef);
        r.ops.push_back(ops);
    }
    return r;
}

OpenFermionTerms read_openfermion_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw QulacsError("cannot open OpenFermion file: " + path);
    return parse_openfermion_lines(f);
}

template <typename Op>
Op* build_operator(const OpenFermionTerms& t, int select) {  // select: -1 all, 0 Z/I only, 1 with X or Y
    auto op = std::make_unique<Op>(t.qubit_count);
    for (size_t i = 0; i < t.ops.size(); ++i) {
        const bool offdiag = t.ops[i].find_first_of("XY") != std::string::npos;
        if (select < 0 || (select == 1) == offdiag) op->add_operator(t.coefs[i], t.ops[i]);
    }
    return op.release();
}

}  // namespace

namespace quantum_operator {
GeneralQuantumOperator* create_general_quantum_operator_from_openfermion_text(const std::string& text) {
    std::istringstream in(text);
    return build_operator<GeneralQuantumOperator>(parse_openfermion_lines(in), -1);
}
GeneralQuantumOperator* create_general_quantum_operator_from_openfermion_file(const std::string& path) {
    return build_operator<GeneralQuantumOperator>(read_openfermion_file(path), -1);
}
std::pair<GeneralQuantumOperator*, GeneralQuantumOperator*> create_split_general_quantum_operator(
    const std::string& path) {
    const OpenFermionTerms t = read_openfermion_file(path);
    return {build_operator<GeneralQuantumOperator>(t, 0), build_operator<GeneralQuantumOperator>(t, 1)};
}
}  // namespace quantum_operator
namespace observable {
Observable* create_observable_from_openfermion_text(const std::string& text) {
    std::istringstream in(text);
    return build_operator<Observable>(parse_openfermion_lines(in), -1);
}
Observable* create_observable_from_openfermion_file(const std::string& path) {
    return build_operator<Observable>(read_openfermion_file(path), -1);
}
std::pair<Observable*, Observable*> create_split_observable(const std::string& path) {
    const OpenFermionTerms t = read_openfermion_file(path);
    return {build_operator<Observable>(t, 0), build_operator<Observable>(t, 1)};
}
}  // namespace observable

}  // namespace qmh

namespace qmh {

// ------------------------------------------------------------------ parametric circuit
QuantumGate_SingleParameter* ParametricQuantumCircuit::param(size_t i) const {
    if (i >= param_pos_.size()) throw QulacsError("parameter index out of range");
    return static_cast<QuantumGate_SingleParameter*>(gates_[param_pos_[i]]);
}
void ParametricQuantumCircuit::add_parametric_gate(const QuantumGate_SingleParameter& g) {
    add_gate_consume(g.copy());
    param_pos_.push_back(gates_.size() - 1);
}
void ParametricQuantumCircuit::add_parametric_gate(const QuantumGate_SingleParameter& g, size_t position) {
    add_gate(g, position);
    param_pos_.push_back(position);
}
void ParametricQuantumCircuit::add_gate(const QuantumGateBase& g, size_t position) {
    QuantumCircuit::add_gate(g, position);
    for (auto& p : param_pos_)
        if (p >= position) ++p;
}
void ParametricQuantumCircuit::merge_circuit(const ParametricQuantumCircuit& c) {
    const size_t base = gates_.size();
    const std::vector<size_t> pos(c.param_pos_);  // `c` may be this circuit
    QuantumCircuit::merge_circuit(c);
    for (size_t p : pos) param_pos_.push_back(base + p);
}
void ParametricQuantumCircuit::remove_gate(size_t i) {
    QuantumCircuit::remove_gate(i);
    std::vector<size_t> kept;
    for (size_t p : param_pos_) {
        if (p == i) continue;
        kept.push_back(p > i ? p - 1 : p);
    }
    param_pos_ = kept;
}
void ParametricQuantumCircuit::add_parametric_RX_gate(unsigned q, double a) {
    std::unique_ptr<QuantumGate_SingleParameter> g(gate::ParametricRX(q, a));
    add_parametric_gate(*g);
}
void ParametricQuantumCircuit::add_parametric_RY_gate(unsigned q, double a) {
    std::unique_ptr<QuantumGate_SingleParameter> g(gate::ParametricRY(q, a));
    add_parametric_gate(*g);
}
void ParametricQuantumCircuit::add_parametric_RZ_gate(unsigned q, double a) {
    std::unique_ptr<QuantumGate_SingleParameter> g(gate::ParametricRZ(q, a));
    add_parametric_gate(*g);
}
void ParametricQuantumCircuit::add_parametric_multi_Pauli_rotation_gate(const std::vector<unsigned>& t,
                                                                        const std::vector<unsigned>& p, double a) {
    std::unique_ptr<QuantumGate_SingleParameter> g(gate::ParametricPauliRotation(t, p, a));
    add_parametric_gate(*g);
}
double ParametricQuantumCircuit::get_parameter(size_t i) const { return param(i)->get_parameter_value(); }
void ParametricQuantumCircuit::set_parameter(size_t i, double v) {
    param(i)->set_parameter_value(v);
    touched();
}
size_t ParametricQuantumCircuit::get_parametric_gate_position(size_t i) const {
    if (i >= param_pos_.size()) throw QulacsError("parameter index out of range");
    return param_pos_[i];
}

// Walks the circuit backwards holding psi_i (state after gate i) and the
// adjoint-propagated bistate; d/d theta of exp(i theta P / 2) is
// exp(i (theta + pi) P / 2) / 2, so each parameter costs one extra rotation
// and one inner product on the device.
std::vector<double> ParametricQuantumCircuit::backprop_inner_product(QuantumStateMI355* bistate) {
    if (bistate->qubit_count != qubit_count_) throw QulacsError("backprop: qubit count mismatch");
    QuantumStateMI355 state(qubit_count_, bistate->device_number), work(qubit_count_, bistate->device_number);
    update_quantum_state(&state);
    std::vector<long> which(gates_.size(), -1);
    for (size_t i = 0; i < param_pos_.size(); ++i) which[param_pos_[i]] = (long)i;
    std::vector<double> grad(param_pos_.size(), 0.0);
    for (size_t gi = gates_.size(); gi-- > 0;) {
        QuantumGateBase* g = gates_[gi];
        if (which[gi] >= 0) {
            auto* pg = static_cast<QuantumGate_SingleParameter*>(g);
            work.load(&state);
            QuantumGate_SingleParameter shift(pg->get_target_index_list(), pg->get_pauli_id_list(), M_PI, "shift");
            shift.update_quantum_state(&work);
            grad[which[gi]] = inner_product(bistate, &work).real() / 2.0;
        }
        std::unique_ptr<QuantumGateBase> inv(g->get_inverse());
        inv->update_quantum_state(bistate);
        inv->update_quantum_state(&state);
    }
    return grad;
}

std::vector<double> ParametricQuantumCircuit::backprop(const GeneralQuantumOperator& obs) {
    QuantumStateMI355 state(qubit_count_), bistate(qubit_count_), work(qubit_count_);
    update_quantum_state(&state);
    obs.apply_to_state(&work, state, &bistate);
    bistate.multiply_coef(2.0);
    return backprop_inner_product(&bistate);
}

std::vector<CPPCTYPE> GradCalculator::calculate_grad(ParametricQuantumCircuit& c, const GeneralQuantumOperator& obs,
                                                     const std::vector<double>& theta) {
    const size_t np = c.get_parameter_count();
    if (theta.size() != np) throw QulacsError("calculate_grad: one angle per parameter expected");
    std::unique_ptr<ParametricQuantumCircuit> cc(static_cast<ParametricQuantumCircuit*>(c.copy()));
    QuantumStateMI355 state(c.get_qubit_count());
    std::vector<CPPCTYPE> grad(np);
    for (size_t k = 0; k < np; ++k) {
        CPPCTYPE e[2];
        for (int s = 0; s < 2; ++s) {
            for (size_t q = 0; q < np; ++q) cc->set_parameter(q, theta[q] + (q == k ? (s ? -M_PI_2 : M_PI_2) : 0.0));
            state.set_zero_state();
            cc->update_quantum_state(&state);
            e[s] = obs.get_expectation_value(&state);
        }
        grad[k] = (e[0] - e[1]) / 2.0;
    }
    return grad;
}
std::vector<CPPCTYPE> GradCalculator::calculate_grad(ParametricQuantumCircuit& c, const GeneralQuantumOperator& obs) {
    std::vector<double> theta;
    for (size_t i = 0; i < c.get_parameter_count(); ++i) theta.push_back(c.get_parameter(i));
    return calculate_grad(c, obs, theta);
}

}  // namespace qmh

namespace qmh {

// ------------------------------------------------------------------ circuit optimizer
namespace {
bool mergeable(const QuantumGateBase* g) { return !g->is_parametric() && !g->is_noise(); }

uint64_t qubit_mask(const QuantumGateBase* g) {
    uint64_t m = 0;
    for (unsigned q : g->get_target_index_list()) m |= 1ULL << q;
    for (unsigned q : g->get_control_index_list()) m |= 1ULL << q;
    return m;
}
}  // namespace

// the gates' own commutation flags decide (src/cppsim/circuit_optimizer.cpp:35-54)
bool QuantumCircuitOptimizer::commute(size_t a, size_t b) const { return c_->gates_[a]->is_commute(c_->gates_[b]); }
// src/cppsim/circuit_optimizer.cpp:35-54
size_t QuantumCircuitOptimizer::rightmost_commute(size_t i) const {
    size_t cur = i + 1;
    for (; cur < c_->gates_.size(); ++cur)
        if (!commute(i, cur)) break;
    return cur - 1;
}
size_t QuantumCircuitOptimizer::leftmost_commute(size_t i) const {
    long cur = (long)i - 1;
    for (; cur >= 0; --cur)
        if (!commute(i, (size_t)cur)) break;
    return (size_t)(cur + 1);
}

// src/cppsim/circuit_optimizer.cpp:113-164
void QuantumCircuitOptimizer::optimize(QuantumCircuit* circuit, unsigned max_block_size, unsigned swap_level) {
    if (swap_level != 0) throw QulacsError("optimize: swap_level > 0 needs a distributed state");
    if (circuit->get_qubit_count() > 64) throw QulacsError("optimize: at most 64 qubits");
    c_ = circuit;
    auto& gl = circuit->gates_;
    bool merged = true;
    while (merged) {
        merged = false;
        for (size_t i1 = 0; i1 < gl.size(); ++i1) {
            for (size_t i2 = i1 + 1; i2 < gl.size(); ++i2) {
                if (!mergeable(gl[i1]) || !mergeable(gl[i2])) continue;
                if ((unsigned)__builtin_popcountll(qubit_mask(gl[i1]) | qubit_mask(gl[i2])) > max_block_size) continue;
                if (leftmost_commute(i2) > rightmost_commute(i1) + 1) continue;  // not neighbouring
                QuantumGateMatrix* m = gate::merge(gl[i1], gl[i2]);
                const size_t insert = std::max(leftmost_commute(i2), i1 + 1);
                delete gl[i2];
                gl.erase(gl.begin() + i2);
                gl.insert(gl.begin() + insert, m);
                delete gl[i1];
                gl.erase(gl.begin() + i1);
                --i2;
                merged = true;
            }
        }
    }
    circuit->touched();
    c_ = nullptr;
}

// src/cppsim/circuit_optimizer.cpp:166-216
void QuantumCircuitOptimizer::optimize_light(QuantumCircuit* circuit, unsigned swap_level) {
    if (swap_level != 0) throw QulacsError("optimize_light: swap_level > 0 needs a distributed state");
    auto& gl = circuit->gates_;
    const unsigned n = circuit->get_qubit_count();
    std::vector<std::pair<long, std::vector<unsigned>>> step(n, {-1, {}});
    for (size_t i = 0; i < gl.size(); ++i) {
        QuantumGateBase* g = gl[i];
        std::vector<unsigned> qs = g->get_target_index_list();
        for (unsigned c : g->get_control_index_list()) qs.push_back(c);
        std::sort(qs.begin(), qs.end());
        long pos = -1;
        long hit = -1;
        for (unsigned qb : qs)
            if (step[qb].first > pos) {
                pos = step[qb].first;
                hit = qb;
            }
        std::vector<unsigned> parent;
        if (hit != -1) parent = step[hit].second;
        // A parametric or noise gate that cannot merge starts a new step here; the
        // reference skips it without recording it, which lets a later gate merge
        // across it (src/cppsim/circuit_optimizer.cpp:196-200).
        if (std::includes(parent.begin(), parent.end(), qs.begin(), qs.end()) && mergeable(gl[pos]) &&
            mergeable(g)) {
            QuantumGateMatrix* m = gate::merge(gl[pos], g);
            delete gl[i];
            gl.erase(gl.begin() + i);
            gl.insert(gl.begin() + pos + 1, m);
            delete gl[pos];
            gl.erase(gl.begin() + pos);
            --i;
        } else {
            for (unsigned qb : qs) step[qb] = {(long)i, qs};
        }
    }
    circuit->touched();
}

// src/cppsim/circuit_optimizer.cpp:218-235
QuantumGateMatrix* QuantumCircuitOptimizer::merge_all(const QuantumCircuit* circuit) {
    std::unique_ptr<QuantumGateBase> cur(gate::to_matrix_gate(std::unique_ptr<QuantumGateBase>(gate::Identity(0)).get()));
    for (auto* g : circuit->gates_) cur.reset(gate::merge(cur.get(), g));
    return static_cast<QuantumGateMatrix*>(cur.release());
}

}  // namespace qmh

namespace qmh {
CPPCTYPE QuantumCircuitSimulator::get_expectation_value(const GeneralQuantumOperator& observable) {
    return observable.get_expectation_value(state_);
}
void QuantumCircuitSimulator::swap_state_and_buffer() {
    if (!buffer_) {
        buffer_ = state_->allocate_buffer();
        buffer_->set_zero_state();
    }
    std::unique_ptr<QuantumStateBase> tmp(state_->copy());
    state_->load(buffer_);
    buffer_->load(tmp.get());
}
}  // namespace qmh
