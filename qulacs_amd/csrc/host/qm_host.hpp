// Host-side C++ API over libqulacs_mi355: the qulacs cppsim interface for the
// state-vector / density-matrix hot path, with every state resident on the GPU.
//
// Mirrors (names, argument meaning, conventions):
//   QuantumStateBase / QuantumState       src/cppsim/state.hpp:18-417, state_gpu.hpp:13-415
//   DensityMatrix                         src/cppsim/state_dm.hpp
//   QuantumGateBase::update_quantum_state src/cppsim/gate.hpp:83-200
//   named / rotation / matrix / diagonal / Pauli gates
//                                         src/cppsim/gate_named_*.hpp, gate_matrix*.cpp
//   QuantumGate_Probabilistic / CPTP      src/cppsim/gate_general.hpp
//   QuantumCircuit                        src/cppsim/circuit.hpp:29
//   PauliOperator / GeneralQuantumOperator / Observable
//                                         src/cppsim/pauli_operator.hpp, general_quantum_operator.hpp,
//                                         observable.hpp
// Every device operation goes through the C ABI in include/qulacs_mi355.h.
#pragma once

#include <complex>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/qulacs_mi355.h"

namespace qmh {

using CPPCTYPE = std::complex<double>;

// Raised for invalid arguments and for errors recorded by the C ABI.
class QulacsError : public std::runtime_error {
   public:
    using std::runtime_error::runtime_error;
};

// Throws if the last C ABI call on this thread recorded an error.
void check_abi(const char* where);

// Row-major dense matrix.
struct ComplexMatrix {
    size_t rows = 0, cols = 0;
    std::vector<CPPCTYPE> data;
    ComplexMatrix() = default;
    ComplexMatrix(size_t r, size_t c) : rows(r), cols(c), data(r * c) {}
    CPPCTYPE& operator()(size_t r, size_t c) { return data[r * cols + c]; }
    const CPPCTYPE& operator()(size_t r, size_t c) const { return data[r * cols + c]; }
    static ComplexMatrix identity(size_t n);
    ComplexMatrix operator*(const ComplexMatrix& o) const;
    ComplexMatrix adjoint() const;
};

// std::mt19937_64 + uniform_real_distribution<double>(0, 1), as the reference's
// Random (src/cppsim/utility.hpp:54-102), so a given seed draws the same uniforms.
class Random {
   public:
    Random();
    void set_seed(uint64_t seed) { mt_.seed(seed); }
    double uniform() { return uniform_(mt_); }
    double normal() { return normal_(mt_); }
    uint64_t int64() { return mt_(); }
    uint32_t int32() { return (uint32_t)(mt_() % 0xffffffffULL); }

   private:
    std::uniform_real_distribution<double> uniform_{0.0, 1.0};
    std::normal_distribution<double> normal_{0.0, 1.0};
    std::mt19937_64 mt_;
};

// ------------------------------------------------------------------ states
class QuantumStateBase {
   public:
    const unsigned qubit_count;
    const ITYPE dim;
    const unsigned device_number;

    QuantumStateBase(unsigned qubit_count, ITYPE dim, unsigned device);
    virtual ~QuantumStateBase();
    QuantumStateBase(const QuantumStateBase&) = delete;
    QuantumStateBase& operator=(const QuantumStateBase&) = delete;

    virtual bool is_state_vector() const = 0;
    std::string get_device_name() const { return "gpu"; }
    // Number of complex entries in device memory (dim for a state vector, dim^2 for a DM).
    virtual ITYPE element_count() const = 0;

    virtual void set_zero_state() = 0;
    // Every entry 0 (src/cppsim/state.hpp:136; the reference's GPU state throws here).
    void set_zero_norm_state();
    virtual void set_computational_basis(ITYPE comp_basis) = 0;
    virtual void set_Haar_random_state() = 0;
    virtual void set_Haar_random_state(unsigned seed) = 0;
    virtual double get_zero_probability(unsigned target_qubit_index) const = 0;
    // measured_values[q] in {0, 1, 2}; 2 means "not measured" (src/cppsim/state.hpp).
    virtual double get_marginal_probability(const std::vector<unsigned>& measured_values) const = 0;
    virtual double get_entropy() const = 0;
    virtual double get_squared_norm() const = 0;
    virtual void normalize(double squared_norm) = 0;
    virtual QuantumStateBase* allocate_buffer() const = 0;
    virtual QuantumStateBase* copy() const = 0;
    void load(const QuantumStateBase* other);
    virtual void load(const std::vector<CPPCTYPE>& host) = 0;
    void* data() const { return data_; }
    std::vector<CPPCTYPE> duplicate_data_cpp() const;
    void add_state(const QuantumStateBase* other);
    void add_state_with_coef(CPPCTYPE coef, const QuantumStateBase* other);
    void multiply_coef(CPPCTYPE coef);
    virtual std::vector<ITYPE> sampling(unsigned count) = 0;
    std::vector<ITYPE> sampling(unsigned count, unsigned seed);
    virtual std::string to_string() const = 0;

    void* get_cuda_stream() const { return stream_; }
    unsigned get_classical_value(unsigned index);
    void set_classical_value(unsigned index, unsigned value);
    const std::vector<unsigned>& get_classical_register() const { return classical_register_; }

    Random& random() { return random_; }

   protected:
    void* data_ = nullptr;
    void* stream_ = nullptr;  // hipStream_t[1] from allocate_cuda_stream_host
    std::vector<unsigned> classical_register_;
    Random random_;
};

// State vector of `qubit_count` qubits in GPU memory; get_device_name() == "gpu".
class QuantumStateMI355 : public QuantumStateBase {
   public:
    explicit QuantumStateMI355(unsigned qubit_count, unsigned device = 0);
    bool is_state_vector() const override { return true; }
    ITYPE element_count() const override { return dim; }
    void set_zero_state() override;
    void set_computational_basis(ITYPE comp_basis) override;
    void set_Haar_random_state() override;
    void set_Haar_random_state(unsigned seed) override;
    double get_zero_probability(unsigned target_qubit_index) const override;
    double get_marginal_probability(const std::vector<unsigned>& measured_values) const override;
    double get_entropy() const override;
    double get_squared_norm() const override;
    void normalize(double squared_norm) override;
    QuantumStateBase* allocate_buffer() const override;
    QuantumStateBase* copy() const override;
    using QuantumStateBase::load;
    void load(const std::vector<CPPCTYPE>& host) override;
    using QuantumStateBase::sampling;
    std::vector<ITYPE> sampling(unsigned count) override;
    // Sampling with caller-supplied uniforms (same indices as the sequential CPU loop).
    std::vector<ITYPE> sampling_with_uniforms(const std::vector<double>& uniforms) const;
    // samples and their boundary flags (qm_sampling_checked_host)
    std::pair<std::vector<ITYPE>, std::vector<unsigned char>> sampling_with_uniforms_checked(
        const std::vector<double>& uniforms, double margin) const;
    std::string to_string() const override;
    // psi[i] *= f(i): f is evaluated on the host for every index, then one device
    // pass multiplies (src/cppsim/state.hpp:885-891, state_gpu.hpp:353).
    void multiply_elementwise_function(const std::function<CPPCTYPE(ITYPE)>& f);
};

// Density matrix of `qubit_count` qubits (dim x dim, row-major) in GPU memory.
class DensityMatrixMI355 : public QuantumStateBase {
   public:
    explicit DensityMatrixMI355(unsigned qubit_count, unsigned device = 0);
    bool is_state_vector() const override { return false; }
    ITYPE element_count() const override { return dim * dim; }
    void set_zero_state() override;
    void set_computational_basis(ITYPE comp_basis) override;
    void set_Haar_random_state() override;
    void set_Haar_random_state(unsigned seed) override;
    double get_zero_probability(unsigned target_qubit_index) const override;
    double get_marginal_probability(const std::vector<unsigned>& measured_values) const override;
    double get_entropy() const override;
    double get_squared_norm() const override;
    void normalize(double squared_norm) override;
    QuantumStateBase* allocate_buffer() const override;
    QuantumStateBase* copy() const override;
    using QuantumStateBase::load;
    void load(const std::vector<CPPCTYPE>& host) override;  // a dim vector (pure) or dim^2 matrix
    void load_pure(const QuantumStateMI355* pure);
    using QuantumStateBase::sampling;
    std::vector<ITYPE> sampling(unsigned count) override;
    std::vector<ITYPE> sampling_with_uniforms(const std::vector<double>& uniforms) const;
    // samples and their boundary flags (qm_sampling_checked_host)
    std::pair<std::vector<ITYPE>, std::vector<unsigned char>> sampling_with_uniforms_checked(
        const std::vector<double>& uniforms, double margin) const;
    std::string to_string() const override;
};

CPPCTYPE inner_product(const QuantumStateMI355* bra, const QuantumStateMI355* ket);

// New states built from existing ones on the device (src/cppsim/state.cpp:32-100,
// src/cppsim/state_dm.cpp:9-90).
namespace stateops {
QuantumStateMI355* tensor_product(const QuantumStateMI355* left, const QuantumStateMI355* right);
DensityMatrixMI355* tensor_product(const DensityMatrixMI355* left, const DensityMatrixMI355* right);
QuantumStateMI355* permutate_qubit(const QuantumStateMI355* s, const std::vector<unsigned>& order);
DensityMatrixMI355* permutate_qubit(const DensityMatrixMI355* s, const std::vector<unsigned>& order);
QuantumStateMI355* drop_qubit(const QuantumStateMI355* s, const std::vector<unsigned>& target,
                              const std::vector<unsigned>& projection);
DensityMatrixMI355* partial_trace(const QuantumStateBase* s, const std::vector<unsigned>& target);
QuantumStateMI355* make_superposition(CPPCTYPE c1, const QuantumStateMI355* s1, CPPCTYPE c2,
                                      const QuantumStateMI355* s2);
DensityMatrixMI355* make_mixture(CPPCTYPE p1, const QuantumStateBase* s1, CPPCTYPE p2, const QuantumStateBase* s2);
}  // namespace stateops

// ------------------------------------------------------------------ gates
class QuantumGateBase {
   public:
    virtual ~QuantumGateBase() = default;
    virtual void update_quantum_state(QuantumStateBase* state) = 0;
    virtual QuantumGateBase* copy() const = 0;
    // Dense matrix over the target list (bit i <-> target i); throws for non-unitary channels.
    virtual ComplexMatrix get_matrix() const = 0;
    virtual QuantumGateBase* get_inverse() const;
    virtual std::string get_name() const { return name_; }
    virtual bool is_parametric() const { return false; }
    virtual bool is_noise() const { return false; }
    // false when get_matrix() has no meaning for the gate (channels, reflections, ...)
    virtual bool has_matrix() const { return true; }
    // A copy acting on other qubits: targets[i] replaces target i and controls[i]
    // control i (src/cppsim/gate.hpp:283-310); gates made of gates remap those too.
    virtual QuantumGateBase* create_gate_whose_qubit_indices_are_replaced(const std::vector<unsigned>& targets,
                                                                          const std::vector<unsigned>& controls) const;
    std::string to_string() const;
    const std::vector<unsigned>& get_target_index_list() const { return targets_; }
    const std::vector<unsigned>& get_control_index_list() const { return controls_; }
    const std::vector<unsigned>& get_control_value_list() const { return cvals_; }
    std::vector<std::pair<unsigned, unsigned>> get_control_index_value_list() const;

    // Which single-qubit Paulis the gate commutes with on each target: bit 0 X,
    // bit 1 Y, bit 2 Z (src/cppsim/qubit_info.hpp:7-11); a control commutes with Z only.
    enum : unsigned { COMMUTE_X = 1, COMMUTE_Y = 2, COMMUTE_Z = 4 };
    // Gate class flags (src/cppsim/gate.hpp:69-75).
    enum : unsigned { PROPERTY_PAULI = 1, PROPERTY_CLIFFORD = 2, PROPERTY_GAUSSIAN = 4, PROPERTY_PARAMETRIC = 8 };
    unsigned get_target_commutation(size_t i) const { return i < tcomm_.size() ? tcomm_[i] : 0u; }
    // flags of qubit q: its target flags, COMMUTE_Z for a control, all three when unused
    unsigned commutation_of(unsigned q) const;
    void set_target_commutation(std::vector<unsigned> flags) { tcomm_ = std::move(flags); }
    unsigned get_property_value() const { return property_ | (is_parametric() ? PROPERTY_PARAMETRIC : 0u); }
    void set_gate_property(unsigned p) { property_ = p; }
    bool is_Pauli() const { return property_ & PROPERTY_PAULI; }
    bool is_Clifford() const { return property_ & PROPERTY_CLIFFORD; }
    bool is_Gaussian() const { return property_ & PROPERTY_GAUSSIAN; }
    // every target commutes with Z (src/cppsim/gate.cpp:46-60)
    bool is_diagonal() const;
    // target-target pairs on one qubit share a Pauli they commute with; a target on
    // another gate's control commutes with Z (src/cppsim/gate.cpp:12-31)
    bool is_commute(const QuantumGateBase* other) const;

   protected:
    std::string name_;
    std::vector<unsigned> targets_, controls_, cvals_;
    std::vector<unsigned> tcomm_;  // parallel to targets_ (missing entries are 0)
    unsigned property_ = 0;
};

// Dense (optionally controlled) matrix gate; also the base of every unitary
// gate with a fast path (named, rotation, CNOT/CZ/SWAP).
class QuantumGateMatrix : public QuantumGateBase {
   public:
    QuantumGateMatrix(std::vector<unsigned> targets, ComplexMatrix m, std::vector<unsigned> controls = {},
                      std::vector<unsigned> cvals = {}, std::string name = "DenseMatrix");
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new QuantumGateMatrix(*this); }
    ComplexMatrix get_matrix() const override { return matrix_; }
    QuantumGateBase* get_inverse() const override;
    // clears the Pauli and Gaussian flags (src/cppsim/gate_matrix.cpp:296-302)
    void add_control_qubit(unsigned qubit, unsigned value);
    // matrix *= value; the gate becomes a generic dense gate
    void multiply_scalar(CPPCTYPE value);
    // kind selects a dedicated entry point (0 = generic dense)
    enum Kind { GENERIC, X, Y, Z, H, S, SDAG, T, TDAG, SQRTX, SQRTXDAG, SQRTY, SQRTYDAG, P0, P1, RX, RY, RZ,
                CNOT, CZ, SWAP, IDENTITY };
    Kind kind = GENERIC;
    double angle = 0.0;  // RX/RY/RZ in the reference's exp(+i angle P / 2) convention

   protected:
    ComplexMatrix matrix_;
    void update_sv(QuantumStateMI355* s);
    void update_dm(DensityMatrixMI355* s);
};

class QuantumGateDiagonalMatrix : public QuantumGateBase {
   public:
    QuantumGateDiagonalMatrix(std::vector<unsigned> targets, std::vector<CPPCTYPE> diag,
                              std::vector<unsigned> controls = {}, std::vector<unsigned> cvals = {});
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new QuantumGateDiagonalMatrix(*this); }
    ComplexMatrix get_matrix() const override;
    const std::vector<CPPCTYPE>& get_diagonal() const { return diag_; }

   private:
    std::vector<CPPCTYPE> diag_;
};

// Pauli product (rotation = false) or exp(+i angle P / 2) (rotation = true).
class ClsPauliGate : public QuantumGateBase {
   public:
    ClsPauliGate(std::vector<unsigned> targets, std::vector<unsigned> paulis, bool rotation, double angle);
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new ClsPauliGate(*this); }
    ComplexMatrix get_matrix() const override;
    QuantumGateBase* get_inverse() const override;
    const std::vector<unsigned>& get_pauli_id_list() const { return paulis_; }
    bool is_rotation() const { return rotation_; }
    double get_angle() const { return angle_; }
    void set_angle(double a) { angle_ = a; }

   protected:
    std::vector<unsigned> paulis_;
    bool rotation_;
    double angle_;
};

// Mixture sum_i p_i G_i rho G_i^dagger (src/cppsim/gate_general.hpp:230-290).
// On a state vector one G_i is drawn; on a density matrix the whole map is applied
// as one superoperator pass when every G_i is a matrix gate.
class QuantumGate_Probabilistic : public QuantumGateBase {
   public:
    QuantumGate_Probabilistic(std::vector<double> distribution, std::vector<QuantumGateBase*> gates);
    QuantumGate_Probabilistic(const QuantumGate_Probabilistic& o);
    ~QuantumGate_Probabilistic() override;
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new QuantumGate_Probabilistic(*this); }
    ComplexMatrix get_matrix() const override;
    bool is_noise() const override { return true; }
    void set_seed(uint64_t seed) { random_.set_seed(seed); }
    QuantumGateBase* create_gate_whose_qubit_indices_are_replaced(const std::vector<unsigned>& targets,
                                                                  const std::vector<unsigned>& controls) const override;
    const std::vector<double>& get_distribution() const { return dist_; }
    const std::vector<double>& get_cumulative_distribution() const { return cumulative_; }
    const std::vector<QuantumGateBase*>& get_gate_list() const { return gates_; }
    // gates in descending order of probability, so the branch search ends early
    // (src/cppsim/gate_general.hpp:350-372)
    void optimize_ProbablisticGate();
    bool has_matrix() const override { return false; }
    // ProbabilisticInstrument: on a state vector the drawn branch index (the gate count
    // for the implicit identity) goes to this classical register address
    // (src/cppsim/gate_general.hpp:230-262)
    void set_instrument(unsigned classical_register_address) {
        instrument_ = true;
        register_ = classical_register_address;
    }
    bool is_instrument() const { return instrument_; }
    unsigned get_classical_register_address() const { return register_; }
    // superoperator over targets_ (4^k x 4^k, member = row bits * 2^k + column bits)
    ComplexMatrix superoperator() const;

   private:
    std::vector<double> dist_, cumulative_;
    std::vector<QuantumGateBase*> gates_;
    Random random_;
    bool instrument_ = false;
    unsigned register_ = 0;
};

// Kraus channel sum_i K_i rho K_i^dagger (src/cppsim/gate_general.hpp:380-470); the K_i
// may be any gates. On a state vector one K_i is chosen with probability
// ||K_i psi||^2 / ||psi||^2 and the state rescaled to its old norm (identity when the
// map is not trace preserving and no branch is hit); on a density matrix it is one
// superoperator pass when every K_i has a matrix and at most 3 qubits are involved,
// else the sum of the K_i applied to copies.
class QuantumGate_CPTP : public QuantumGateBase {
   public:
    explicit QuantumGate_CPTP(const std::vector<const QuantumGateBase*>& kraus);
    QuantumGate_CPTP(const QuantumGate_CPTP& o);
    QuantumGate_CPTP& operator=(const QuantumGate_CPTP&) = delete;
    ~QuantumGate_CPTP() override;
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new QuantumGate_CPTP(*this); }
    ComplexMatrix get_matrix() const override;
    bool has_matrix() const override { return false; }
    bool is_noise() const override { return true; }
    // superoperator over targets_ (4^k x 4^k); throws when a Kraus operator has no matrix
    ComplexMatrix superoperator() const;
    void set_seed(uint64_t seed) { random_.set_seed(seed); }
    QuantumGateBase* create_gate_whose_qubit_indices_are_replaced(const std::vector<unsigned>& targets,
                                                                  const std::vector<unsigned>& controls) const override;
    const std::vector<QuantumGateBase*>& get_gate_list() const { return gates_; }
    // Instrument: on a state vector, the index of the chosen Kraus operator (the
    // operator count when none was chosen) is written to this classical register
    // address.
    void set_instrument(unsigned classical_register_address) {
        instrument_ = true;
        register_ = classical_register_address;
    }
    bool is_instrument() const { return instrument_; }
    unsigned get_classical_register_address() const { return register_; }

   protected:
    std::vector<QuantumGateBase*> gates_;
    Random random_;
    bool instrument_ = false;
    unsigned register_ = 0;
    // Branch rules (a CPTP map: rescale the chosen branch, raw probabilities, leave the
    // state alone when no branch is chosen).
    bool state_normalize_ = true, probability_normalize_ = false, assign_zero_ = false;
};

// Completely positive map (src/cppsim/gate_general.hpp:560-690): a CPTP map whose
// branch rules are chosen: state_normalize rescales the chosen branch,
// probability_normalize divides the branch probabilities by their sum, and
// assign_zero_if_not_matched zeroes the state when no branch is chosen.
class QuantumGate_CP : public QuantumGate_CPTP {
   public:
    QuantumGate_CP(const std::vector<const QuantumGateBase*>& kraus, bool state_normalize, bool probability_normalize,
                   bool assign_zero_if_not_matched);
    QuantumGateBase* copy() const override { return new QuantumGate_CP(*this); }
    bool get_state_normalize() const { return state_normalize_; }
    bool get_probability_normalize() const { return probability_normalize_; }
    bool get_assign_zero_if_not_matched() const { return assign_zero_; }
};

// sum_i c_i G_i on a state vector (src/cppsim/gate_general.hpp:17-150): each G_i acts
// on a copy and the copies are summed with their coefficients.
class QuantumGate_LinearCombination : public QuantumGateBase {
   public:
    QuantumGate_LinearCombination(const std::vector<CPPCTYPE>& coefs, const std::vector<const QuantumGateBase*>& gates);
    QuantumGate_LinearCombination(const QuantumGate_LinearCombination& o);
    QuantumGate_LinearCombination& operator=(const QuantumGate_LinearCombination&) = delete;
    ~QuantumGate_LinearCombination() override;
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new QuantumGate_LinearCombination(*this); }
    // sum_i c_i (matrix of G_i) over targets_
    ComplexMatrix get_matrix() const override;
    const std::vector<CPPCTYPE>& get_coef_list() const { return coefs_; }
    const std::vector<QuantumGateBase*>& get_gate_list() const { return gates_; }
    QuantumGateBase* create_gate_whose_qubit_indices_are_replaced(const std::vector<unsigned>& targets,
                                                                  const std::vector<unsigned>& controls) const override;

   private:
    std::vector<CPPCTYPE> coefs_;
    std::vector<QuantumGateBase*> gates_;
};

// psi <- 2 <r|psi> r - psi for a fixed state r (src/cppsim/gate_reflect.hpp,
// src/csim/update_ops_reflection.cpp:14-27); state vectors only.
class ClsStateReflectionGate : public QuantumGateBase {
   public:
    explicit ClsStateReflectionGate(const QuantumStateMI355* reflection_state);
    ClsStateReflectionGate(const ClsStateReflectionGate& o);
    ClsStateReflectionGate& operator=(const ClsStateReflectionGate&) = delete;
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new ClsStateReflectionGate(*this); }
    ComplexMatrix get_matrix() const override;
    bool has_matrix() const override { return false; }
    const QuantumStateMI355* get_reflection_state() const { return r_.get(); }

   private:
    std::unique_ptr<QuantumStateMI355> r_;
};

// Applies the gate when a condition on the classical register holds
// (src/cppsim/gate_general.hpp:693-790).
class QuantumGate_Adaptive : public QuantumGateBase {
   public:
    using Condition = std::function<bool(const std::vector<unsigned>&)>;
    using ConditionWithId = std::function<bool(const std::vector<unsigned>&, unsigned)>;
    QuantumGate_Adaptive(const QuantumGateBase* g, Condition c);
    QuantumGate_Adaptive(const QuantumGateBase* g, ConditionWithId c, unsigned id);
    QuantumGate_Adaptive(const QuantumGate_Adaptive& o);
    QuantumGate_Adaptive& operator=(const QuantumGate_Adaptive&) = delete;
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new QuantumGate_Adaptive(*this); }
    ComplexMatrix get_matrix() const override;
    bool has_matrix() const override { return false; }
    QuantumGateBase* create_gate_whose_qubit_indices_are_replaced(const std::vector<unsigned>& targets,
                                                                  const std::vector<unsigned>& controls) const override;

   private:
    std::unique_ptr<QuantumGateBase> gate_;
    Condition cond_;
    ConditionWithId cond_id_;
    int id_ = -1;
};

// FusedSWAP: swaps the qubit block [t1, t1 + b) with [t2, t2 + b)
// (src/cppsim/gate_named_npair.hpp:12-140); its matrix index takes the first block in
// the low bits.
class ClsNpairQubitGate : public QuantumGateBase {
   public:
    ClsNpairQubitGate(unsigned t1, unsigned t2, unsigned block_size);
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new ClsNpairQubitGate(*this); }
    ComplexMatrix get_matrix() const override;
    QuantumGateBase* get_inverse() const override { return copy(); }
    unsigned get_block_size() const { return block_; }

   private:
    unsigned block_;
};

// Parametric rotation gates (src/vqcsim/parametric_gate.hpp:31-212).
class QuantumGate_SingleParameter : public QuantumGateBase {
   public:
    // kind: 1 = RX, 2 = RY, 3 = RZ (single target) or Pauli rotation when paulis given
    QuantumGate_SingleParameter(std::vector<unsigned> targets, std::vector<unsigned> paulis, double angle,
                                std::string name);
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new QuantumGate_SingleParameter(*this); }
    ComplexMatrix get_matrix() const override;
    QuantumGateBase* get_inverse() const override;
    bool is_parametric() const override { return true; }
    double get_parameter_value() const { return angle_; }
    void set_parameter_value(double v) { angle_ = v; }
    const std::vector<unsigned>& get_pauli_id_list() const { return paulis_; }

   private:
    std::vector<unsigned> paulis_;
    double angle_;
};

namespace gate {
QuantumGateMatrix* Identity(unsigned q);
QuantumGateMatrix* X(unsigned q);
QuantumGateMatrix* Y(unsigned q);
QuantumGateMatrix* Z(unsigned q);
QuantumGateMatrix* H(unsigned q);
QuantumGateMatrix* S(unsigned q);
QuantumGateMatrix* Sdag(unsigned q);
QuantumGateMatrix* T(unsigned q);
QuantumGateMatrix* Tdag(unsigned q);
QuantumGateMatrix* sqrtX(unsigned q);
QuantumGateMatrix* sqrtXdag(unsigned q);
QuantumGateMatrix* sqrtY(unsigned q);
QuantumGateMatrix* sqrtYdag(unsigned q);
QuantumGateMatrix* P0(unsigned q);
QuantumGateMatrix* P1(unsigned q);
QuantumGateMatrix* RX(unsigned q, double angle);
QuantumGateMatrix* RY(unsigned q, double angle);
QuantumGateMatrix* RZ(unsigned q, double angle);
QuantumGateMatrix* RotX(unsigned q, double angle);
QuantumGateMatrix* RotY(unsigned q, double angle);
QuantumGateMatrix* RotZ(unsigned q, double angle);
QuantumGateMatrix* RotInvX(unsigned q, double angle);
QuantumGateMatrix* RotInvY(unsigned q, double angle);
QuantumGateMatrix* RotInvZ(unsigned q, double angle);
QuantumGateMatrix* U1(unsigned q, double lambda);
QuantumGateMatrix* U2(unsigned q, double phi, double lambda);
QuantumGateMatrix* U3(unsigned q, double theta, double phi, double lambda);
QuantumGateMatrix* CNOT(unsigned control, unsigned target);
QuantumGateMatrix* CZ(unsigned control, unsigned target);
QuantumGateMatrix* SWAP(unsigned a, unsigned b);
QuantumGateMatrix* DenseMatrix(const std::vector<unsigned>& targets, const ComplexMatrix& m);
QuantumGateDiagonalMatrix* DiagonalMatrix(const std::vector<unsigned>& targets, const std::vector<CPPCTYPE>& diag);
ClsPauliGate* Pauli(const std::vector<unsigned>& targets, const std::vector<unsigned>& paulis);
ClsPauliGate* PauliRotation(const std::vector<unsigned>& targets, const std::vector<unsigned>& paulis, double angle);
QuantumGateMatrix* RandomUnitary(const std::vector<unsigned>& targets, uint64_t seed);
QuantumGate_Probabilistic* DepolarizingNoise(unsigned q, double p);
QuantumGate_Probabilistic* TwoQubitDepolarizingNoise(unsigned a, unsigned b, double p);
QuantumGate_Probabilistic* BitFlipNoise(unsigned q, double p);
QuantumGate_Probabilistic* DephasingNoise(unsigned q, double p);
QuantumGate_Probabilistic* IndependentXZNoise(unsigned q, double p);
QuantumGate_CPTP* AmplitudeDampingNoise(unsigned q, double p);
QuantumGate_Probabilistic* Probabilistic(const std::vector<double>& dist, const std::vector<QuantumGateBase*>& gates);
QuantumGate_CPTP* CPTP(const std::vector<QuantumGateBase*>& kraus);
QuantumGate_CPTP* Instrument(const std::vector<QuantumGateBase*>& kraus, unsigned classical_register_address);
QuantumGate_CPTP* Measurement(unsigned q, unsigned classical_register_address);
QuantumGate_CP* CP(const std::vector<QuantumGateBase*>& kraus, bool state_normalize, bool probability_normalize,
                   bool assign_zero_if_not_matched);
QuantumGate_Probabilistic* ProbabilisticInstrument(const std::vector<double>& dist,
                                                   const std::vector<QuantumGateBase*>& gates,
                                                   unsigned classical_register_address);
QuantumGate_LinearCombination* LinearCombination(const std::vector<CPPCTYPE>& coefs,
                                                 const std::vector<QuantumGateBase*>& gates);
// Instrument with the projectors (I + P) / 2 and (I - P) / 2 of a Pauli string
// (src/cppsim/gate_factory.cpp:539-570): 0 is written for the +1 outcome, 1 for -1.
QuantumGate_CPTP* MultiQubitPauliMeasurement(const std::vector<unsigned>& targets, const std::vector<unsigned>& paulis,
                                             unsigned classical_register_address);
ClsStateReflectionGate* StateReflection(const QuantumStateMI355* reflection_state);
QuantumGate_Adaptive* Adaptive(const QuantumGateBase* g, QuantumGate_Adaptive::Condition c);
QuantumGate_Adaptive* Adaptive(const QuantumGateBase* g, QuantumGate_Adaptive::ConditionWithId c, unsigned id);
// Basis permutation |x> -> |f(x, 2^k)> on the targets (src/cppsim/gate_reversible.hpp,
// src/csim/update_ops_reversible_boolean.cpp:5-46), evaluated once into a permutation
// matrix; f must be a bijection of 0..2^k-1.
QuantumGateMatrix* ReversibleBoolean(const std::vector<unsigned>& targets, const std::function<ITYPE(ITYPE, ITYPE)>& f);
ClsNpairQubitGate* FusedSWAP(unsigned t1, unsigned t2, unsigned block_size);
QuantumGateMatrix* TOFFOLI(unsigned control1, unsigned control2, unsigned target);
QuantumGateMatrix* FREDKIN(unsigned control, unsigned target1, unsigned target2);
QuantumGateMatrix* add(const QuantumGateBase* a, const QuantumGateBase* b);
QuantumGateMatrix* add(const std::vector<QuantumGateBase*>& gates);
QuantumGateMatrix* to_matrix_gate(const QuantumGateBase* g);
QuantumGateMatrix* merge(const QuantumGateBase* first, const QuantumGateBase* second);
QuantumGate_SingleParameter* ParametricRX(unsigned q, double angle);
QuantumGate_SingleParameter* ParametricRY(unsigned q, double angle);
QuantumGate_SingleParameter* ParametricRZ(unsigned q, double angle);
QuantumGate_SingleParameter* ParametricPauliRotation(const std::vector<unsigned>& targets,
                                                     const std::vector<unsigned>& paulis, double angle);
}  // namespace gate

// ------------------------------------------------------------------ circuits
class GeneralQuantumOperator;
class NoiseSimulator;

// Whether QuantumCircuit::update_quantum_state may run runs of fusable gates through
// the stage executor (default on; off gives one device call per gate, as the
// reference's loop).
void set_circuit_fusion(bool enabled);
bool get_circuit_fusion();

class QuantumCircuit {
   public:
    explicit QuantumCircuit(unsigned qubit_count) : qubit_count_(qubit_count) {}
    QuantumCircuit(const QuantumCircuit& o);
    QuantumCircuit& operator=(const QuantumCircuit&) = delete;
    virtual ~QuantumCircuit();
    unsigned get_qubit_count() const { return qubit_count_; }
    size_t get_gate_count() const { return gates_.size(); }
    const QuantumGateBase* get_gate(size_t i) const;
    void add_gate(const QuantumGateBase& g);
    void add_gate(const QuantumGateBase& g, size_t position);
    void add_gate_consume(QuantumGateBase* g);
    // throws QulacsError when g names a qubit outside the circuit
    void check_gate_indices(const QuantumGateBase& g) const;
    void remove_gate(size_t i);
    virtual QuantumCircuit* copy() const { return new QuantumCircuit(*this); }
    void update_quantum_state(QuantumStateBase* state);
    void update_quantum_state(QuantumStateBase* state, size_t start, size_t end);
    unsigned calculate_depth() const;
    std::string to_string() const;
    QuantumCircuit* get_inverse() const;

    // add_<name>_gate helpers (src/cppsim/circuit.hpp)
    void add_X_gate(unsigned q) { add_gate_consume(gate::X(q)); }
    void add_Y_gate(unsigned q) { add_gate_consume(gate::Y(q)); }
    void add_Z_gate(unsigned q) { add_gate_consume(gate::Z(q)); }
    void add_H_gate(unsigned q) { add_gate_consume(gate::H(q)); }
    void add_S_gate(unsigned q) { add_gate_consume(gate::S(q)); }
    void add_Sdag_gate(unsigned q) { add_gate_consume(gate::Sdag(q)); }
    void add_T_gate(unsigned q) { add_gate_consume(gate::T(q)); }
    void add_Tdag_gate(unsigned q) { add_gate_consume(gate::Tdag(q)); }
    void add_sqrtX_gate(unsigned q) { add_gate_consume(gate::sqrtX(q)); }
    void add_sqrtXdag_gate(unsigned q) { add_gate_consume(gate::sqrtXdag(q)); }
    void add_sqrtY_gate(unsigned q) { add_gate_consume(gate::sqrtY(q)); }
    void add_sqrtYdag_gate(unsigned q) { add_gate_consume(gate::sqrtYdag(q)); }
    void add_P0_gate(unsigned q) { add_gate_consume(gate::P0(q)); }
    void add_P1_gate(unsigned q) { add_gate_consume(gate::P1(q)); }
    void add_CNOT_gate(unsigned c, unsigned t) { add_gate_consume(gate::CNOT(c, t)); }
    void add_CZ_gate(unsigned c, unsigned t) { add_gate_consume(gate::CZ(c, t)); }
    void add_SWAP_gate(unsigned a, unsigned b) { add_gate_consume(gate::SWAP(a, b)); }
    void add_RX_gate(unsigned q, double a) { add_gate_consume(gate::RX(q, a)); }
    void add_RY_gate(unsigned q, double a) { add_gate_consume(gate::RY(q, a)); }
    void add_RZ_gate(unsigned q, double a) { add_gate_consume(gate::RZ(q, a)); }
    void add_RotX_gate(unsigned q, double a) { add_gate_consume(gate::RotX(q, a)); }
    void add_RotY_gate(unsigned q, double a) { add_gate_consume(gate::RotY(q, a)); }
    void add_RotZ_gate(unsigned q, double a) { add_gate_consume(gate::RotZ(q, a)); }
    void add_RotInvX_gate(unsigned q, double a) { add_gate_consume(gate::RotInvX(q, a)); }
    void add_RotInvY_gate(unsigned q, double a) { add_gate_consume(gate::RotInvY(q, a)); }
    void add_RotInvZ_gate(unsigned q, double a) { add_gate_consume(gate::RotInvZ(q, a)); }
    void add_U1_gate(unsigned q, double l) { add_gate_consume(gate::U1(q, l)); }
    void add_U2_gate(unsigned q, double p, double l) { add_gate_consume(gate::U2(q, p, l)); }
    void add_U3_gate(unsigned q, double t, double p, double l) { add_gate_consume(gate::U3(q, t, p, l)); }
    void add_dense_matrix_gate(const std::vector<unsigned>& t, const ComplexMatrix& m) {
        add_gate_consume(gate::DenseMatrix(t, m));
    }
    void add_diagonal_matrix_gate(const std::vector<unsigned>& t, const std::vector<CPPCTYPE>& d) {
        add_gate_consume(gate::DiagonalMatrix(t, d));
    }
    void add_multi_Pauli_gate(const std::vector<unsigned>& t, const std::vector<unsigned>& p) {
        add_gate_consume(gate::Pauli(t, p));
    }
    void add_multi_Pauli_rotation_gate(const std::vector<unsigned>& t, const std::vector<unsigned>& p, double a) {
        add_gate_consume(gate::PauliRotation(t, p, a));
    }
    void add_random_unitary_gate(const std::vector<unsigned>& t, uint64_t seed) {
        add_gate_consume(gate::RandomUnitary(t, seed));
    }
    // copies of every gate of `circuit`, appended (src/cppsim/circuit.hpp:150-155)
    void merge_circuit(const QuantumCircuit& circuit);
    // the gate, then the named noise channel ("Depolarizing", "BitFlip", "Dephasing",
    // "IndependentXZ", "AmplitudeDamping") on its targets and controls
    // (src/cppsim/circuit.cpp:167-248)
    void add_noise_gate(const QuantumGateBase& g, const std::string& noise_type, double noise_prob);
    // num_repeats rounds of one Pauli rotation by Re(coef_k) * angle / num_repeats per
    // term; num_repeats == 0 picks ceil(angle * qubit_count * 100) (src/cppsim/circuit.cpp:500-522)
    void add_observable_rotation_gate(const GeneralQuantumOperator& observable, double angle, unsigned num_repeats);
    // one Pauli rotation per term; every term must be diagonal (src/cppsim/circuit.cpp:481-499)
    void add_diagonal_observable_rotation_gate(const GeneralQuantumOperator& observable, double angle);

   protected:
    friend class QuantumCircuitOptimizer;
    friend class NoiseSimulator;
    void touched() { ++version_; }
    unsigned qubit_count_;
    std::vector<QuantumGateBase*> gates_;

   private:
    // Fused execution of a whole-circuit update on a state vector: maximal runs of
    // gates that lower to the stage executor's kinds (dense / diagonal on <= 4 qubits
    // counting controls, CZ, CNOT) become one compiled program each; the rest run
    // gate by gate. The plan is rebuilt when the gate list or a parameter changes.
    struct FusedRun {
        size_t begin = 0, end = 0;  // gates [begin, end)
        qm_circuit* program = nullptr;  // null: run the gates one by one
    };
    void release_plan();
    void build_plan();
    void update_dm_fused(DensityMatrixMI355* d);
    std::vector<FusedRun> plan_;
    uint64_t version_ = 0, plan_version_ = ~0ULL;
    std::mutex plan_mutex_;  // one fused update at a time per circuit (the plan uploads lazily per device)
};

// Host-side gate fusion (src/cppsim/circuit_optimizer.cpp:113-235).  Two gates
// commute as QuantumGateBase::is_commute says (their per-qubit Pauli commutation
// flags); parametric and noise gates are never merged.  Only swap_level 0 (no distributed-qubit swap
// insertion) exists on one device.
class QuantumCircuitOptimizer {
   public:
    void optimize(QuantumCircuit* circuit, unsigned max_block_size, unsigned swap_level = 0);
    void optimize_light(QuantumCircuit* circuit, unsigned swap_level = 0);
    QuantumGateMatrix* merge_all(const QuantumCircuit* circuit);

   private:
    QuantumCircuit* c_ = nullptr;
    bool commute(size_t a, size_t b) const;
    size_t rightmost_commute(size_t i) const;
    size_t leftmost_commute(size_t i) const;
};

class GeneralQuantumOperator;

// Circuit + state + one buffer (src/cppsim/simulator.hpp).
class QuantumCircuitSimulator {
   public:
    QuantumCircuitSimulator(QuantumCircuit* circuit, QuantumStateBase* state)
        : circuit_(circuit), state_(state) {}
    ~QuantumCircuitSimulator() { delete buffer_; }
    void initialize_state(ITYPE computational_basis = 0) { state_->set_computational_basis(computational_basis); }
    void initialize_random_state() { state_->set_Haar_random_state(); }
    void initialize_random_state(unsigned seed) { state_->set_Haar_random_state(seed); }
    void simulate() { circuit_->update_quantum_state(state_); }
    void simulate_range(size_t start, size_t end) { circuit_->update_quantum_state(state_, start, end); }
    CPPCTYPE get_expectation_value(const GeneralQuantumOperator& observable);
    size_t get_gate_count() const { return circuit_->get_gate_count(); }
    void copy_state_to_buffer() {
        if (!buffer_) buffer_ = state_->allocate_buffer();
        buffer_->load(state_);
    }
    void copy_state_from_buffer() {
        if (!buffer_) throw QulacsError("copy_state_from_buffer: the buffer is empty");
        state_->load(buffer_);
    }
    void swap_state_and_buffer();

   private:
    QuantumCircuit* circuit_;
    QuantumStateBase* state_;
    QuantumStateBase* buffer_ = nullptr;
};

// Circuit with registered parametric rotations (src/vqcsim/parametric_circuit.hpp).
// A parametric gate added through add_gate is a plain gate; only the
// add_parametric_* calls register a parameter.
class ParametricQuantumCircuit : public QuantumCircuit {
   public:
    using QuantumCircuit::QuantumCircuit;
    QuantumCircuit* copy() const override { return new ParametricQuantumCircuit(*this); }
    void add_parametric_gate(const QuantumGate_SingleParameter& g);
    void add_parametric_gate(const QuantumGate_SingleParameter& g, size_t position);
    void add_parametric_RX_gate(unsigned q, double a);
    void add_parametric_RY_gate(unsigned q, double a);
    void add_parametric_RZ_gate(unsigned q, double a);
    void add_parametric_multi_Pauli_rotation_gate(const std::vector<unsigned>& t, const std::vector<unsigned>& p,
                                                  double a);
    void add_gate(const QuantumGateBase& g) { QuantumCircuit::add_gate(g); }
    void add_gate(const QuantumGateBase& g, size_t position);
    // gates appended with their parameters registered after this circuit's
    // (src/vqcsim/parametric_circuit.cpp:167-181)
    using QuantumCircuit::merge_circuit;
    void merge_circuit(const ParametricQuantumCircuit& circuit);
    void remove_gate(size_t i);
    size_t get_parameter_count() const { return param_pos_.size(); }
    double get_parameter(size_t i) const;
    void set_parameter(size_t i, double v);
    size_t get_parametric_gate_position(size_t i) const;
    // d<O>/d(theta_i) by the adjoint sweep (src/vqcsim/parametric_circuit.cpp:205-335)
    std::vector<double> backprop(const GeneralQuantumOperator& obs);
    // Re <bistate|d psi/d theta_i> with psi = circuit |0> (bistate is overwritten)
    std::vector<double> backprop_inner_product(QuantumStateMI355* bistate);

   private:
    QuantumGate_SingleParameter* param(size_t i) const;
    std::vector<size_t> param_pos_;  // gate index of parameter i
};

// Parameter-shift gradient (src/vqcsim/GradCalculator.cpp:8-53).
struct GradCalculator {
    std::vector<CPPCTYPE> calculate_grad(ParametricQuantumCircuit& c, const GeneralQuantumOperator& obs,
                                         const std::vector<double>& theta);
    std::vector<CPPCTYPE> calculate_grad(ParametricQuantumCircuit& c, const GeneralQuantumOperator& obs);
};

// ------------------------------------------------------------------ operators
class PauliOperator {
   public:
    explicit PauliOperator(CPPCTYPE coef = 1.0) : coef_(coef) {}
    PauliOperator(const std::string& pauli_string, CPPCTYPE coef = 1.0);
    PauliOperator(const std::vector<unsigned>& targets, const std::vector<unsigned>& paulis, CPPCTYPE coef = 1.0);
    void add_single_Pauli(unsigned qubit, unsigned pauli);
    CPPCTYPE get_coef() const { return coef_; }
    void change_coef(CPPCTYPE c) { coef_ = c; }
    const std::vector<unsigned>& get_index_list() const { return targets_; }
    const std::vector<unsigned>& get_pauli_id_list() const { return paulis_; }
    std::string get_pauli_string() const;
    // coef * <psi|P|psi> (complex because coef is)
    CPPCTYPE get_expectation_value(const QuantumStateBase* state) const;
    CPPCTYPE get_transition_amplitude(const QuantumStateBase* bra, const QuantumStateBase* ket) const;
    void masks(ITYPE* bit_flip, ITYPE* phase_flip, unsigned* rot90) const;
    PauliOperator* copy() const { return new PauliOperator(*this); }
    // x and z bit of every qubit up to the highest target; adding a Pauli sets bits
    // (X: x, Y: x and z, Z: z), never clears them (src/cppsim/pauli_operator.cpp:116-131).
    void xz_bits(std::vector<uint8_t>* x, std::vector<uint8_t>* z) const;
    bool same_pauli_string(const PauliOperator& other) const;
    // Product with the i^k phase of the single-qubit products; the result lists its
    // Paulis in ascending qubit order (src/cppsim/pauli_operator.cpp:300-413).
    PauliOperator operator*(const PauliOperator& target) const;
    PauliOperator operator*(CPPCTYPE c) const { return PauliOperator(targets_, paulis_, coef_ * c); }
    PauliOperator& operator*=(const PauliOperator& target) { return *this = *this * target; }
    PauliOperator& operator*=(CPPCTYPE c) {
        coef_ *= c;
        return *this;
    }

   private:
    std::vector<unsigned> targets_, paulis_;
    CPPCTYPE coef_;
};

class GeneralQuantumOperator {
   public:
    explicit GeneralQuantumOperator(unsigned qubit_count) : qubit_count_(qubit_count) {}
    virtual ~GeneralQuantumOperator() = default;
    // Every term enters through this one (an Observable rejects complex coefficients).
    virtual void add_operator(const PauliOperator& p);
    void add_operator(CPPCTYPE coef, const std::string& pauli_string);
    void add_operator(const std::vector<unsigned>& targets, const std::vector<unsigned>& paulis, CPPCTYPE coef);
    size_t get_term_count() const { return terms_.size(); }
    unsigned get_qubit_count() const { return qubit_count_; }
    ITYPE get_state_dim() const { return 1ULL << qubit_count_; }
    const PauliOperator& get_term(size_t i) const { return terms_.at(i); }
    bool is_hermitian() const;
    // sum_k coef_k <P_k>: one device call for a state vector
    CPPCTYPE get_expectation_value(const QuantumStateBase* state) const;
    CPPCTYPE get_transition_amplitude(const QuantumStateBase* bra, const QuantumStateBase* ket) const;
    // dst = (sum_k coef_k P_k) src, using `work` as scratch (same size)
    void apply_to_state(QuantumStateMI355* work, const QuantumStateMI355& src, QuantumStateMI355* dst) const;
    // Same, with `state` itself as scratch: each Pauli is applied and undone
    // (src/cppsim/general_quantum_operator.cpp:350-368).
    void apply_to_state(QuantumStateMI355* state, QuantumStateMI355* dst) const;
    // Terms with conjugated coefficients (src/cppsim/general_quantum_operator.cpp:526-533).
    GeneralQuantumOperator get_dagger() const;
    // "coef pauli_string + ..." with coef printed as a std::complex
    // (src/cppsim/general_quantum_operator.cpp:1011-1022).
    virtual std::string to_string() const;
    // operator_count terms on every qubit, uniform Pauli ids and coefficients in [-1, 1)
    // (src/cppsim/general_quantum_operator.cpp:179-202).
    void add_random_operator(unsigned operator_count);
    void add_random_operator(unsigned operator_count, unsigned seed);
    // sum of |Re coef| (src/cppsim/general_quantum_operator.cpp:442-450)
    CPPCTYPE calculate_default_mu() const;
    // Ground-state eigenvalue of the operator; `state` is the start vector and is left
    // holding the solver's eigenvector estimate. mu == 0 selects calculate_default_mu()
    // (src/cppsim/general_quantum_operator.cpp:204-328).
    CPPCTYPE solve_ground_state_eigenvalue_by_power_method(QuantumStateMI355* state, unsigned iter_count,
                                                           CPPCTYPE mu = 0.0) const;
    CPPCTYPE solve_ground_state_eigenvalue_by_arnoldi_method(QuantumStateMI355* state, unsigned iter_count,
                                                             CPPCTYPE mu = 0.0) const;

    // Sums merge terms with the same Pauli string; products multiply every pair of
    // terms (src/cppsim/general_quantum_operator.cpp:547-805).
    GeneralQuantumOperator& operator+=(const GeneralQuantumOperator& target);
    GeneralQuantumOperator& operator+=(const PauliOperator& target);
    GeneralQuantumOperator& operator-=(const GeneralQuantumOperator& target);
    GeneralQuantumOperator& operator-=(const PauliOperator& target);
    GeneralQuantumOperator& operator*=(const GeneralQuantumOperator& target);
    GeneralQuantumOperator& operator*=(const PauliOperator& target);
    GeneralQuantumOperator& operator*=(CPPCTYPE target);
    GeneralQuantumOperator operator+(const GeneralQuantumOperator& t) const { return copy_base() += t; }
    GeneralQuantumOperator operator+(const PauliOperator& t) const { return copy_base() += t; }
    GeneralQuantumOperator operator-(const GeneralQuantumOperator& t) const { return copy_base() -= t; }
    GeneralQuantumOperator operator-(const PauliOperator& t) const { return copy_base() -= t; }
    GeneralQuantumOperator operator*(const GeneralQuantumOperator& t) const { return copy_base() *= t; }
    GeneralQuantumOperator operator*(const PauliOperator& t) const { return copy_base() *= t; }
    GeneralQuantumOperator operator*(CPPCTYPE t) const { return copy_base() *= t; }

   protected:
    GeneralQuantumOperator copy_base() const;
    void merge_terms(const std::vector<PauliOperator>& terms, double sign);
    unsigned qubit_count_;
    std::vector<PauliOperator> terms_;
    Random random_;
};

// HermitianQuantumOperator: expectation values are real and every coefficient is real
// (src/cppsim/observable.cpp:19-58).
class Observable : public GeneralQuantumOperator {
   public:
    using GeneralQuantumOperator::GeneralQuantumOperator;
    using GeneralQuantumOperator::add_operator;
    void add_operator(const PauliOperator& p) override;
    double get_expectation_value_real(const QuantumStateBase* state) const {
        return GeneralQuantumOperator::get_expectation_value(state).real();
    }
    // "coef pauli_string + ..." with the real part of coef (src/cppsim/observable.cpp:192-203)
    std::string to_string() const override;
    // Lanczos on the operator shifted by -mu (src/cppsim/observable.cpp:61-190).
    CPPCTYPE solve_ground_state_eigenvalue_by_lanczos_method(QuantumStateMI355* state, unsigned iter_count,
                                                             CPPCTYPE mu = 0.0) const;
};

// Pauli decomposition of a gate's matrix over GQO_qubits qubits: coefficient of the
// string P on the gate's targets is tr(P U) / 2^k, terms with |tr(P U)| <= tol are
// dropped, strings in the order (row bits, column bits) of the transform in
// src/cppsim/gate_to_gqo.hpp:9-65 (row and column bit h of the string index:
// 00 I, 01 X, 10 Y, 11 Z on target h). Controlled gates are rejected.
GeneralQuantumOperator* to_general_quantum_operator(const QuantumGateBase* gate, unsigned GQO_qubits,
                                                    double tol = 1e-6);

// Quantum-trajectory evolution under a Lindblad master equation with Hamiltonian H and
// collapse operators c_k for a given time (src/cppsim/gate_noisy_evolution.cpp:23-310):
// RK4 steps of dt under H_eff = H - (i/2) sum_k c_k^dagger c_k until the squared norm
// drops to a uniform r, the jump time found by a secant search on log(norm), then
// c_k applied with probability ||c_k psi||^2 / sum and a new r drawn. State vectors
// only; every step runs on the device.
class ClsNoisyEvolution : public QuantumGateBase {
   public:
    ClsNoisyEvolution(const Observable& hamiltonian, const std::vector<const GeneralQuantumOperator*>& c_ops,
                      double time, double dt = 1e-6);
    ClsNoisyEvolution(const ClsNoisyEvolution& o);
    ClsNoisyEvolution& operator=(const ClsNoisyEvolution&) = delete;
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new ClsNoisyEvolution(*this); }
    ComplexMatrix get_matrix() const override;
    bool has_matrix() const override { return false; }
    bool is_noise() const override { return true; }
    void set_seed(uint64_t seed) { random_.set_seed(seed); }
    GeneralQuantumOperator get_effective_hamiltonian() const { return *heff_; }
    void set_find_collapse_max_steps(int n) { max_steps_ = n; }
    const Observable& get_hamiltonian() const { return *h_; }
    const std::vector<std::unique_ptr<GeneralQuantumOperator>>& get_c_ops() const { return c_ops_; }
    double get_time() const { return time_; }
    double get_dt() const { return dt_; }

   private:
    void evolve(QuantumStateMI355* k1, QuantumStateMI355* k2, QuantumStateMI355* k3, QuantumStateMI355* k4,
                QuantumStateMI355* buffer, QuantumStateMI355* state, double dt) const;
    double find_collapse(QuantumStateMI355* k1, QuantumStateMI355* k2, QuantumStateMI355* k3, QuantumStateMI355* k4,
                         QuantumStateMI355* prev, QuantumStateMI355* now, double target, double dt) const;
    std::unique_ptr<Observable> h_;
    std::vector<std::unique_ptr<GeneralQuantumOperator>> c_ops_;
    std::unique_ptr<GeneralQuantumOperator> heff_;
    double time_, dt_;
    double norm_tol_ = 1e-6;
    int max_steps_ = 200;
    Random random_;
};
// The same trajectories with exact propagation: -i H_eff on the qubits its terms touch
// (at most 8) is diagonalised once, V diag(w) V^-1, and exp(-i H_eff t) is applied as
// V^-1, diag(exp(w t)), V on the device; each step runs to the end time or the next
// jump (src/cppsim/gate_noisy_evolution.cpp:330-616).
class ClsNoisyEvolution_fast : public QuantumGateBase {
   public:
    ClsNoisyEvolution_fast(const Observable& hamiltonian, const std::vector<const GeneralQuantumOperator*>& c_ops,
                           double time);
    ClsNoisyEvolution_fast(const ClsNoisyEvolution_fast& o);
    ClsNoisyEvolution_fast& operator=(const ClsNoisyEvolution_fast&) = delete;
    void update_quantum_state(QuantumStateBase* state) override;
    QuantumGateBase* copy() const override { return new ClsNoisyEvolution_fast(*this); }
    ComplexMatrix get_matrix() const override;
    bool has_matrix() const override { return false; }
    bool is_noise() const override { return true; }
    void set_seed(uint64_t seed) { random_.set_seed(seed); }
    GeneralQuantumOperator get_effective_hamiltonian() const { return *heff_; }
    void set_find_collapse_max_steps(int n) { max_steps_ = n; }
    const Observable& get_hamiltonian() const { return *h_; }
    const std::vector<std::unique_ptr<GeneralQuantumOperator>>& get_c_ops() const { return c_ops_; }
    double get_time() const { return time_; }

   private:
    void evolve(QuantumStateMI355* state, double t) const;
    double find_collapse(QuantumStateMI355* prev, QuantumStateMI355* now, double target, double t_step) const;
    std::unique_ptr<Observable> h_;
    std::vector<std::unique_ptr<GeneralQuantumOperator>> c_ops_;
    std::unique_ptr<GeneralQuantumOperator> heff_;
    double time_;
    double norm_tol_ = 1e-6;
    int max_steps_ = 200;
    Random random_;
    bool none_ = false;  // no Hamiltonian terms: the gate does nothing
    std::vector<CPPCTYPE> eig_;
    std::unique_ptr<QuantumGateMatrix> v_, vinv_;
};
namespace gate {
ClsNoisyEvolution* NoisyEvolution(const Observable& hamiltonian, const std::vector<const GeneralQuantumOperator*>& c_ops,
                                  double time, double dt);
ClsNoisyEvolution_fast* NoisyEvolution_fast(const Observable& hamiltonian,
                                            const std::vector<const GeneralQuantumOperator*>& c_ops, double time);
}  // namespace gate

// Samples a circuit whose noise gates are probabilistic mixtures by drawing one branch
// per noise gate per shot, grouping shots with the same branches, and simulating each
// distinct trajectory once from the longest prefix they share
// (src/cppsim/noisesimulator.cpp:18-198). Other noise gates (Kraus maps, whose branch
// depends on the state) are rejected.
class NoiseSimulator {
   public:
    struct Result {
        std::vector<std::pair<QuantumStateMI355*, unsigned>> result;  // (final state, shot count)
        Result() = default;
        Result(const Result&) = delete;
        Result& operator=(const Result&) = delete;
        ~Result();
        // each state sampled its shot count times, in order
        std::vector<ITYPE> sampling() const;
    };
    NoiseSimulator(const QuantumCircuit* circuit, const QuantumStateMI355* initial_state = nullptr);
    std::vector<ITYPE> execute(unsigned sample_count);
    Result* execute_and_get_result(unsigned sample_count);
    void set_seed(uint64_t seed) { random_.set_seed(seed); }

   private:
    std::unique_ptr<QuantumCircuit> circuit_;
    std::unique_ptr<QuantumStateMI355> initial_;
    Random random_;
    unsigned choose_branch(const QuantumGateBase* g);
    void apply_from(const std::vector<unsigned>& branches, QuantumStateMI355* s, size_t start, size_t col) const;
};

// Expectation values from light cones: for each term, the gates that can reach its
// qubits (scanning the circuit backwards), split into independent qubit groups by
// union-find, each group simulated on its own smaller state
// (src/vqcsim/causalcone_simulator.hpp:58-212).
class CausalConeSimulator {
   public:
    CausalConeSimulator(const ParametricQuantumCircuit& circuit, const Observable& observable);
    ~CausalConeSimulator();
    CausalConeSimulator(const CausalConeSimulator&) = delete;
    CausalConeSimulator& operator=(const CausalConeSimulator&) = delete;
    void build();
    CPPCTYPE get_expectation_value();
    const std::vector<std::vector<ParametricQuantumCircuit*>>& get_circuit_list() const { return circuits_; }
    const std::vector<std::vector<PauliOperator>>& get_pauli_operator_list() const { return paulis_; }
    const std::vector<CPPCTYPE>& get_coef_list() const { return coefs_; }

   private:
    std::unique_ptr<ParametricQuantumCircuit> circuit_;
    std::unique_ptr<Observable> observable_;
    std::vector<std::vector<ParametricQuantumCircuit*>> circuits_;
    std::vector<std::vector<PauliOperator>> paulis_;
    std::vector<CPPCTYPE> coefs_;
    bool built_ = false;
};

// Operators from OpenFermion's text form, one "coef [P0 P1 ...]" term per line
// (src/cppsim/general_quantum_operator.cpp:806-952, src/cppsim/observable.cpp:205-350,
// src/cppsim/utility.cpp:63-106). A blank line is skipped; a term without Paulis is
// the identity.
namespace quantum_operator {
GeneralQuantumOperator* create_general_quantum_operator_from_openfermion_text(const std::string& text);
GeneralQuantumOperator* create_general_quantum_operator_from_openfermion_file(const std::string& path);
// (terms with only Z/I, terms with an X or Y)
std::pair<GeneralQuantumOperator*, GeneralQuantumOperator*> create_split_general_quantum_operator(
    const std::string& path);
}  // namespace quantum_operator
namespace observable {
Observable* create_observable_from_openfermion_text(const std::string& text);
Observable* create_observable_from_openfermion_file(const std::string& path);
std::pair<Observable*, Observable*> create_split_observable(const std::string& path);
}  // namespace observable

}  // namespace qmh
