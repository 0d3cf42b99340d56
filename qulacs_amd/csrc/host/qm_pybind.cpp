// Python module qulacs_amd._core: the qulacs Python API (python/cppsim_wrapper.cpp)
// for the GPU-resident state-vector / density-matrix path, over the host classes
// in qm_host.hpp. Names, argument names and return kinds follow the reference's
// bindings so that `from qulacs_amd import QuantumState, gate, ...` replaces
// `from qulacs import ...` for this path.
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "qm_host.hpp"
#include "qm_json.hpp"

namespace py = pybind11;
using namespace qmh;

namespace {

using carray = py::array_t<CPPCTYPE, py::array::c_style | py::array::forcecast>;

ComplexMatrix to_matrix(const carray& a) {
    if (a.ndim() != 2) throw QulacsError("matrix must be two-dimensional");
    ComplexMatrix m((size_t)a.shape(0), (size_t)a.shape(1));
    std::copy(a.data(), a.data() + a.size(), m.data.begin());
    return m;
}

py::array_t<CPPCTYPE> from_matrix(const ComplexMatrix& m) {
    py::array_t<CPPCTYPE> a({(py::ssize_t)m.rows, (py::ssize_t)m.cols});
    std::copy(m.data.begin(), m.data.end(), a.mutable_data());
    return a;
}

py::array_t<CPPCTYPE> from_vector(std::vector<CPPCTYPE>&& v) {
    py::array_t<CPPCTYPE> a((py::ssize_t)v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

std::vector<CPPCTYPE> to_vector(const carray& a) {
    if (a.ndim() != 1 && a.ndim() != 2) throw QulacsError("state data must be one- or two-dimensional");
    return std::vector<CPPCTYPE>(a.data(), a.data() + a.size());
}

// sum_k coef_k P_k as a scipy.sparse.csr_matrix (python/cppsim_wrapper.cpp get_matrix):
// P|x> = i^(#Y) (-1)^popcount(x & phase_flip) |x ^ bit_flip>.
py::object operator_matrix(const GeneralQuantumOperator& o) {
    const ITYPE dim = o.get_state_dim();
    std::vector<ITYPE> rows, cols;
    std::vector<CPPCTYPE> vals;
    const CPPCTYPE ipow[4] = {1.0, CPPCTYPE(0, 1), -1.0, CPPCTYPE(0, -1)};
    for (size_t k = 0; k < o.get_term_count(); ++k) {
        const PauliOperator& p = o.get_term(k);
        ITYPE bf, pf;
        unsigned rot;
        p.masks(&bf, &pf, &rot);
        for (ITYPE x = 0; x < dim; ++x) {
            const CPPCTYPE v = p.get_coef() * ipow[rot % 4] * (__builtin_popcountll(x & pf) & 1 ? -1.0 : 1.0);
            rows.push_back(x ^ bf);
            cols.push_back(x);
            vals.push_back(v);
        }
    }
    auto sp = py::module_::import("scipy.sparse");
    py::tuple shape = py::make_tuple(dim, dim);
    return sp.attr("csr_matrix")(py::make_tuple(py::cast(vals), py::make_tuple(py::cast(rows), py::cast(cols))),
                                 py::arg("shape") = shape);
}

std::vector<unsigned> as_targets(const py::object& o) {
    if (py::isinstance<py::int_>(o)) return {o.cast<unsigned>()};
    return o.cast<std::vector<unsigned>>();
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "qulacs state-vector / density-matrix API on MI355X (HIP, complex128)";
    py::register_exception<QulacsError>(m, "QulacsError", PyExc_RuntimeError);
    m.attr("__version__") = qm_version();
    // one process, one GPU: no MPI-distributed states
    m.def("check_build_for_mpi", []() { return false; });
    m.def("set_circuit_fusion", &set_circuit_fusion, py::arg("enabled"));
    m.def("get_circuit_fusion", &get_circuit_fusion);

    // ---------------------------------------------------------------- operators
    py::class_<PauliOperator>(m, "PauliOperator")
        .def(py::init<CPPCTYPE>(), py::arg("coef") = CPPCTYPE(1.0))
        .def(py::init<const std::string&, CPPCTYPE>(), py::arg("pauli_string"), py::arg("coef") = CPPCTYPE(1.0))
        .def(py::init<const std::vector<unsigned>&, const std::vector<unsigned>&, CPPCTYPE>(),
             py::arg("target_qubit_index_list"), py::arg("pauli_operator_type_list"), py::arg("coef") = CPPCTYPE(1.0))
        .def("get_index_list", &PauliOperator::get_index_list)
        .def("get_pauli_id_list", &PauliOperator::get_pauli_id_list)
        .def("get_coef", &PauliOperator::get_coef)
        .def("change_coef", &PauliOperator::change_coef, py::arg("new_coef"))
        .def("add_single_Pauli", &PauliOperator::add_single_Pauli, py::arg("index"), py::arg("pauli_type"))
        .def("get_pauli_string", &PauliOperator::get_pauli_string)
        .def("get_expectation_value", &PauliOperator::get_expectation_value, py::arg("state"))
        // one device reduction either way
        .def("get_expectation_value_single_thread", &PauliOperator::get_expectation_value, py::arg("state"))
        .def("get_transition_amplitude", &PauliOperator::get_transition_amplitude, py::arg("state_bra"),
             py::arg("state_ket"))
        .def("copy", [](const PauliOperator& p) { return PauliOperator(p); })
        .def("__mul__", [](const PauliOperator& a, const PauliOperator& b) { return a * b; }, py::is_operator())
        .def("__mul__", [](const PauliOperator& a, CPPCTYPE c) { return a * c; }, py::is_operator())
        .def("__imul__", [](PauliOperator& a, const PauliOperator& b) -> PauliOperator& { return a *= b; },
             py::is_operator())
        .def("__imul__", [](PauliOperator& a, CPPCTYPE c) -> PauliOperator& { return a *= c; }, py::is_operator())
        .def("__IMUL__", [](PauliOperator& a, CPPCTYPE c) -> PauliOperator& { return a *= c; })
        .def(py::pickle([](const PauliOperator& p) { return ptree::to_json(to_ptree(p)); },
                        [](const std::string& j) { return pauli_operator_from_ptree(ptree::from_json(j)); }));

    py::class_<GeneralQuantumOperator>(m, "GeneralQuantumOperator")
        .def(py::init<unsigned>(), py::arg("qubit_count"))
        .def("add_operator", py::overload_cast<const PauliOperator&>(&GeneralQuantumOperator::add_operator),
             py::arg("pauli_operator"))
        .def("add_operator", py::overload_cast<CPPCTYPE, const std::string&>(&GeneralQuantumOperator::add_operator),
             py::arg("coef"), py::arg("pauli_string"))
        .def("add_operator",
             py::overload_cast<const std::vector<unsigned>&, const std::vector<unsigned>&, CPPCTYPE>(
                 &GeneralQuantumOperator::add_operator),
             py::arg("target_qubit_index_list"), py::arg("target_qubit_pauli_list"), py::arg("coef"))
        .def("add_operator_move", py::overload_cast<const PauliOperator&>(&GeneralQuantumOperator::add_operator),
             py::arg("pauli_operator"))
        .def("add_operator_copy", py::overload_cast<const PauliOperator&>(&GeneralQuantumOperator::add_operator),
             py::arg("pauli_operator"))
        .def("is_hermitian", &GeneralQuantumOperator::is_hermitian)
        .def("get_qubit_count", &GeneralQuantumOperator::get_qubit_count)
        .def("get_state_dim", &GeneralQuantumOperator::get_state_dim)
        .def("get_term_count", &GeneralQuantumOperator::get_term_count)
        .def("get_term", [](const GeneralQuantumOperator& o, size_t i) { return o.get_term(i); }, py::arg("index"))
        .def("get_expectation_value", &GeneralQuantumOperator::get_expectation_value, py::arg("state"))
        .def("get_expectation_value_single_thread", &GeneralQuantumOperator::get_expectation_value, py::arg("state"))
        .def("get_matrix", &operator_matrix)
        .def("get_transition_amplitude", &GeneralQuantumOperator::get_transition_amplitude, py::arg("state_bra"),
             py::arg("state_ket"))
        .def("apply_to_state",
             [](const GeneralQuantumOperator& o, QuantumStateMI355* work, const QuantumStateMI355& src,
                QuantumStateMI355* dst) { o.apply_to_state(work, src, dst); },
             py::arg("work_state"), py::arg("state_to_be_multiplied"), py::arg("dst_state"))
        .def("apply_to_state",
             [](const GeneralQuantumOperator& o, const QuantumStateMI355& src, QuantumStateMI355* dst) {
                 QuantumStateMI355 work(src.qubit_count, src.device_number);
                 o.apply_to_state(&work, src, dst);
             },
             py::arg("state_to_be_multiplied"), py::arg("dst_state"))
        .def("get_dagger", &GeneralQuantumOperator::get_dagger)
        .def("copy", [](const GeneralQuantumOperator& o) { return GeneralQuantumOperator(o); })
        .def("__str__", &GeneralQuantumOperator::to_string)
        .def("add_random_operator", py::overload_cast<unsigned>(&GeneralQuantumOperator::add_random_operator),
             py::arg("operator_count"))
        .def("add_random_operator", py::overload_cast<unsigned, unsigned>(&GeneralQuantumOperator::add_random_operator),
             py::arg("operator_count"), py::arg("seed"))
        .def("solve_ground_state_eigenvalue_by_arnoldi_method",
             &GeneralQuantumOperator::solve_ground_state_eigenvalue_by_arnoldi_method, py::arg("state"),
             py::arg("iter_count"), py::arg("mu") = CPPCTYPE(0.0))
        .def("solve_ground_state_eigenvalue_by_power_method",
             &GeneralQuantumOperator::solve_ground_state_eigenvalue_by_power_method, py::arg("state"),
             py::arg("iter_count"), py::arg("mu") = CPPCTYPE(0.0))
        .def("to_json", [](const GeneralQuantumOperator& o) { return ptree::to_json(to_ptree(o)); })
        .def("__add__", [](const GeneralQuantumOperator& a, const GeneralQuantumOperator& b) { return a + b; },
             py::is_operator())
        .def("__add__", [](const GeneralQuantumOperator& a, const PauliOperator& b) { return a + b; },
             py::is_operator())
        .def("__sub__", [](const GeneralQuantumOperator& a, const GeneralQuantumOperator& b) { return a - b; },
             py::is_operator())
        .def("__sub__", [](const GeneralQuantumOperator& a, const PauliOperator& b) { return a - b; },
             py::is_operator())
        .def("__mul__", [](const GeneralQuantumOperator& a, const GeneralQuantumOperator& b) { return a * b; },
             py::is_operator())
        .def("__mul__", [](const GeneralQuantumOperator& a, const PauliOperator& b) { return a * b; },
             py::is_operator())
        .def("__mul__", [](const GeneralQuantumOperator& a, CPPCTYPE c) { return a * c; }, py::is_operator())
        .def("__iadd__",
             [](GeneralQuantumOperator& a, const GeneralQuantumOperator& b) -> GeneralQuantumOperator& {
                 return a += b;
             },
             py::is_operator())
        .def("__iadd__", [](GeneralQuantumOperator& a, const PauliOperator& b) -> GeneralQuantumOperator& { return a += b; },
             py::is_operator())
        .def("__isub__",
             [](GeneralQuantumOperator& a, const GeneralQuantumOperator& b) -> GeneralQuantumOperator& {
                 return a -= b;
             },
             py::is_operator())
        .def("__isub__", [](GeneralQuantumOperator& a, const PauliOperator& b) -> GeneralQuantumOperator& { return a -= b; },
             py::is_operator())
        .def("__imul__",
             [](GeneralQuantumOperator& a, const GeneralQuantumOperator& b) -> GeneralQuantumOperator& {
                 return a *= b;
             },
             py::is_operator())
        .def("__imul__", [](GeneralQuantumOperator& a, const PauliOperator& b) -> GeneralQuantumOperator& { return a *= b; },
             py::is_operator())
        .def("__imul__", [](GeneralQuantumOperator& a, CPPCTYPE c) -> GeneralQuantumOperator& { return a *= c; },
             py::is_operator())
        // spellings the reference also binds (python/cppsim_wrapper.cpp:170-215)
        .def("__IADD__", [](GeneralQuantumOperator& a, const PauliOperator& b) -> GeneralQuantumOperator& { return a += b; })
        .def("__ISUB__", [](GeneralQuantumOperator& a, const PauliOperator& b) -> GeneralQuantumOperator& { return a -= b; })
        .def("__IMUL__", [](GeneralQuantumOperator& a, const PauliOperator& b) -> GeneralQuantumOperator& { return a *= b; })
        .def("__IMUL__", [](GeneralQuantumOperator& a, CPPCTYPE c) -> GeneralQuantumOperator& { return a *= c; })
        .def(py::pickle([](const GeneralQuantumOperator& o) { return ptree::to_json(to_ptree(o)); },
                        [](const std::string& j) {
                            std::unique_ptr<GeneralQuantumOperator> p(operator_from_ptree(ptree::from_json(j)));
                            return GeneralQuantumOperator(*p);
                        }));

    py::class_<Observable, GeneralQuantumOperator>(m, "Observable")
        .def(py::init<unsigned>(), py::arg("qubit_count"))
        .def(
            "add_operator",
            [](Observable& o, CPPCTYPE coef, const std::string& s) {
                if (coef.imag() != 0.0) throw QulacsError("Observable: coefficient must be real");
                o.add_operator(coef, s);
            },
            py::arg("coef"), py::arg("pauli_string"))
        .def(
            "add_operator",
            [](Observable& o, const PauliOperator& p) {
                if (p.get_coef().imag() != 0.0) throw QulacsError("Observable: coefficient must be real");
                o.add_operator(p);
            },
            py::arg("pauli_operator"))
        .def("get_expectation_value", &Observable::get_expectation_value_real, py::arg("state"))
        .def("get_expectation_value_single_thread", &Observable::get_expectation_value_real, py::arg("state"))
        .def("add_operator_move", [](Observable& o, const PauliOperator& p) { o.add_operator(p); },
             py::arg("pauli_operator"))
        .def("add_operator_copy", [](Observable& o, const PauliOperator& p) { o.add_operator(p); },
             py::arg("pauli_operator"))
        .def("solve_ground_state_eigenvalue_by_lanczos_method",
             &Observable::solve_ground_state_eigenvalue_by_lanczos_method, py::arg("state"), py::arg("iter_count"),
             py::arg("mu") = CPPCTYPE(0.0))
        .def("__str__", &Observable::to_string)
        .def("copy",
             [](const Observable& o) {
                 Observable c(o.get_qubit_count());
                 for (size_t i = 0; i < o.get_term_count(); ++i) c.add_operator(o.get_term(i));
                 return c;
             })
        .def(py::pickle([](const Observable& o) { return ptree::to_json(to_ptree(o)); },
                        [](const std::string& j) {
                            std::unique_ptr<Observable> p(observable_from_ptree(ptree::from_json(j)));
                            Observable c(p->get_qubit_count());
                            for (size_t i = 0; i < p->get_term_count(); ++i) c.add_operator(p->get_term(i));
                            return c;
                        }));

    m.def("to_general_quantum_operator", &to_general_quantum_operator, py::return_value_policy::take_ownership,
          py::arg("gate"), py::arg("qubits"), py::arg("tol"));

    auto mquantum_operator = m.def_submodule("quantum_operator");
    mquantum_operator.def(
        "from_json", [](const std::string& json) { return operator_from_ptree(ptree::from_json(json)); },
        py::return_value_policy::take_ownership, py::arg("json"));
    mquantum_operator.def("create_quantum_operator_from_openfermion_file",
                          &quantum_operator::create_general_quantum_operator_from_openfermion_file,
                          py::return_value_policy::take_ownership, py::arg("file_path"));
    mquantum_operator.def("create_quantum_operator_from_openfermion_text",
                          &quantum_operator::create_general_quantum_operator_from_openfermion_text,
                          py::return_value_policy::take_ownership, py::arg("text"));
    mquantum_operator.def(
        "create_split_quantum_operator",
        [](const std::string& path) {
            auto pr = quantum_operator::create_split_general_quantum_operator(path);
            std::unique_ptr<GeneralQuantumOperator> a(pr.first), b(pr.second);
            return py::make_tuple(GeneralQuantumOperator(*a), GeneralQuantumOperator(*b));
        },
        py::arg("file_path"));
    auto mobservable = m.def_submodule("observable");
    mobservable.def(
        "from_json", [](const std::string& json) { return observable_from_ptree(ptree::from_json(json)); },
        py::return_value_policy::take_ownership, py::arg("json"));
    mobservable.def("create_observable_from_openfermion_file", &observable::create_observable_from_openfermion_file,
                    py::return_value_policy::take_ownership, py::arg("file_path"));
    mobservable.def("create_observable_from_openfermion_text", &observable::create_observable_from_openfermion_text,
                    py::return_value_policy::take_ownership, py::arg("text"));
    mobservable.def(
        "create_split_observable",
        [](const std::string& path) {
            auto pr = observable::create_split_observable(path);
            return py::make_tuple(py::cast(pr.first, py::return_value_policy::take_ownership),
                                  py::cast(pr.second, py::return_value_policy::take_ownership));
        },
        py::arg("file_path"));

    // ---------------------------------------------------------------- states
    py::class_<QuantumStateBase>(m, "QuantumStateBase")
        .def("get_qubit_count", [](const QuantumStateBase& s) { return s.qubit_count; })
        .def("get_device_name", &QuantumStateBase::get_device_name)
        .def("set_zero_state", &QuantumStateBase::set_zero_state)
        .def("set_zero_norm_state", &QuantumStateBase::set_zero_norm_state)
        .def("set_computational_basis", &QuantumStateBase::set_computational_basis, py::arg("comp_basis"))
        .def("set_Haar_random_state", py::overload_cast<>(&QuantumStateBase::set_Haar_random_state))
        .def("set_Haar_random_state", py::overload_cast<unsigned>(&QuantumStateBase::set_Haar_random_state),
             py::arg("seed"))
        .def("get_zero_probability", &QuantumStateBase::get_zero_probability, py::arg("target_qubit_index"))
        .def("get_marginal_probability", &QuantumStateBase::get_marginal_probability, py::arg("measured_values"))
        .def("get_entropy", &QuantumStateBase::get_entropy)
        .def("get_squared_norm", &QuantumStateBase::get_squared_norm)
        .def("normalize", &QuantumStateBase::normalize, py::arg("squared_norm"))
        .def("allocate_buffer", &QuantumStateBase::allocate_buffer, py::return_value_policy::take_ownership)
        .def("copy", &QuantumStateBase::copy, py::return_value_policy::take_ownership)
        .def("load", [](QuantumStateBase& s, const QuantumStateBase& o) { s.load(&o); }, py::arg("state"))
        .def("load", [](QuantumStateBase& s, const carray& a) { s.load(to_vector(a)); }, py::arg("state"))
        .def("add_state", [](QuantumStateBase& s, const QuantumStateBase& o) { s.add_state(&o); }, py::arg("state"))
        .def("add_state_with_coef",
             [](QuantumStateBase& s, CPPCTYPE c, const QuantumStateBase& o) { s.add_state_with_coef(c, &o); },
             py::arg("coef"), py::arg("state"))
        .def("multiply_coef", &QuantumStateBase::multiply_coef, py::arg("coef"))
        .def("get_classical_value", &QuantumStateBase::get_classical_value, py::arg("index"))
        .def("set_classical_value", &QuantumStateBase::set_classical_value, py::arg("index"), py::arg("value"))
        .def("sampling", py::overload_cast<unsigned>(&QuantumStateBase::sampling), py::arg("sampling_count"))
        .def("sampling", py::overload_cast<unsigned, unsigned>(&QuantumStateBase::sampling), py::arg("sampling_count"),
             py::arg("random_seed"))
        .def("to_string", &QuantumStateBase::to_string)
        .def("__str__", &QuantumStateBase::to_string)
        .def("to_json", [](const QuantumStateBase& s) { return ptree::to_json(to_ptree(s)); });

    py::class_<QuantumStateMI355, QuantumStateBase>(m, "QuantumStateMI355")
        .def(py::init<unsigned, unsigned>(), py::arg("qubit_count"), py::arg("gpu_id") = 0)
        .def("get_vector", [](const QuantumStateMI355& s) { return from_vector(s.duplicate_data_cpp()); })
        .def("multiply_elementwise_function", &QuantumStateMI355::multiply_elementwise_function, py::arg("func"))
        .def("get_amplitude",
             [](const QuantumStateMI355& s, ITYPE i) {
                 if (i >= s.dim) throw QulacsError("get_amplitude: index out of range");
                 CPPCTYPE z;
                 copy_quantum_state_from_device_to_host(&z, static_cast<const CPPCTYPE*>(s.data()) + i, 1,
                                                        s.get_cuda_stream(), s.device_number);
                 check_abi("get_amplitude");
                 return z;
             },
             py::arg("index"))
        .def("sampling_with_uniforms", &QuantumStateMI355::sampling_with_uniforms, py::arg("uniforms"))
        .def("sampling_with_uniforms_checked", &QuantumStateMI355::sampling_with_uniforms_checked, py::arg("uniforms"),
             py::arg("margin"))
        .def("data_ptr", [](const QuantumStateMI355& s) { return reinterpret_cast<uintptr_t>(s.data()); })
        .def("stream_ptr", [](const QuantumStateMI355& s) { return reinterpret_cast<uintptr_t>(s.get_cuda_stream()); })
        .def(py::pickle([](const QuantumStateMI355& s) { return ptree::to_json(to_ptree(s)); },
                        [](const std::string& json) {
                            std::unique_ptr<QuantumStateBase> b(state_from_ptree(ptree::from_json(json)));
                            if (!b->is_state_vector()) throw QulacsError("pickle: not a state vector");
                            return std::unique_ptr<QuantumStateMI355>(static_cast<QuantumStateMI355*>(b.release()));
                        }));

    py::class_<DensityMatrixMI355, QuantumStateBase>(m, "DensityMatrixMI355")
        .def(py::init<unsigned, unsigned>(), py::arg("qubit_count"), py::arg("gpu_id") = 0)
        .def("get_matrix",
             [](const DensityMatrixMI355& s) {
                 auto v = s.duplicate_data_cpp();
                 py::array_t<CPPCTYPE> a({(py::ssize_t)s.dim, (py::ssize_t)s.dim});
                 std::copy(v.begin(), v.end(), a.mutable_data());
                 return a;
             })
        .def("load", [](DensityMatrixMI355& s, const QuantumStateBase& o) {
                 if (o.is_state_vector()) s.load_pure(static_cast<const QuantumStateMI355*>(&o));
                 else s.load(&o);
             }, py::arg("state"))
        .def("load", [](DensityMatrixMI355& s, const carray& a) { s.load(to_vector(a)); }, py::arg("state"))
        .def("sampling_with_uniforms", &DensityMatrixMI355::sampling_with_uniforms, py::arg("uniforms"))
        .def("sampling_with_uniforms_checked", &DensityMatrixMI355::sampling_with_uniforms_checked,
             py::arg("uniforms"), py::arg("margin"))
        .def("data_ptr", [](const DensityMatrixMI355& s) { return reinterpret_cast<uintptr_t>(s.data()); })
        .def(py::pickle([](const DensityMatrixMI355& s) { return ptree::to_json(to_ptree(s)); },
                        [](const std::string& json) {
                            std::unique_ptr<QuantumStateBase> b(state_from_ptree(ptree::from_json(json)));
                            if (b->is_state_vector()) throw QulacsError("pickle: not a density matrix");
                            return std::unique_ptr<DensityMatrixMI355>(static_cast<DensityMatrixMI355*>(b.release()));
                        }));

    auto mstate = m.def_submodule("state");
    mstate.def("inner_product", &inner_product, py::arg("state_bra"), py::arg("state_ket"));
    mstate.def(
        "from_json", [](const std::string& json) { return state_from_ptree(ptree::from_json(json)); },
        py::return_value_policy::take_ownership, py::arg("json"));
    const auto own_s = py::return_value_policy::take_ownership;
    mstate.def("tensor_product", py::overload_cast<const QuantumStateMI355*, const QuantumStateMI355*>(
                                     &stateops::tensor_product), own_s, py::arg("state_left"), py::arg("state_right"));
    mstate.def("tensor_product", py::overload_cast<const DensityMatrixMI355*, const DensityMatrixMI355*>(
                                     &stateops::tensor_product), own_s, py::arg("state_left"), py::arg("state_right"));
    mstate.def("permutate_qubit",
               py::overload_cast<const QuantumStateMI355*, const std::vector<unsigned>&>(&stateops::permutate_qubit),
               own_s, py::arg("state"), py::arg("qubit_order"));
    mstate.def("permutate_qubit",
               py::overload_cast<const DensityMatrixMI355*, const std::vector<unsigned>&>(&stateops::permutate_qubit),
               own_s, py::arg("state"), py::arg("qubit_order"));
    mstate.def("drop_qubit", &stateops::drop_qubit, own_s, py::arg("state"), py::arg("target"), py::arg("projection"));
    mstate.def("partial_trace", &stateops::partial_trace, own_s, py::arg("state"), py::arg("target_traceout"));
    mstate.def("make_superposition", &stateops::make_superposition, own_s, py::arg("coef1"), py::arg("state1"),
               py::arg("coef2"), py::arg("state2"));
    mstate.def("make_mixture", &stateops::make_mixture, own_s, py::arg("prob1"), py::arg("state1"), py::arg("prob2"),
               py::arg("state2"));

    // ---------------------------------------------------------------- gates
    py::class_<QuantumGateBase>(m, "QuantumGateBase")
        .def("update_quantum_state", &QuantumGateBase::update_quantum_state, py::arg("state"))
        .def("copy", &QuantumGateBase::copy, py::return_value_policy::take_ownership)
        .def("get_inverse", &QuantumGateBase::get_inverse, py::return_value_policy::take_ownership)
        .def("get_matrix", [](const QuantumGateBase& g) { return from_matrix(g.get_matrix()); })
        .def("get_name", &QuantumGateBase::get_name)
        .def("get_target_index_list", &QuantumGateBase::get_target_index_list)
        .def("get_control_index_list", &QuantumGateBase::get_control_index_list)
        .def("get_control_value_list", &QuantumGateBase::get_control_value_list)
        .def("get_control_index_value_list", &QuantumGateBase::get_control_index_value_list)
        .def("is_parametric", &QuantumGateBase::is_parametric)
        .def("is_noise", &QuantumGateBase::is_noise)
        .def("is_Pauli", &QuantumGateBase::is_Pauli)
        .def("is_Clifford", &QuantumGateBase::is_Clifford)
        .def("is_Gaussian", &QuantumGateBase::is_Gaussian)
        .def("is_diagonal", &QuantumGateBase::is_diagonal)
        .def("is_commute", &QuantumGateBase::is_commute, py::arg("gate"))
        .def("get_property_value", &QuantumGateBase::get_property_value)
        .def("to_string", &QuantumGateBase::to_string)
        .def("__str__", &QuantumGateBase::to_string)
        .def("to_json", [](const QuantumGateBase& g) { return ptree::to_json(to_ptree(g)); });
    py::class_<QuantumGateMatrix, QuantumGateBase>(m, "QuantumGateMatrix")
        .def("add_control_qubit", &QuantumGateMatrix::add_control_qubit, py::arg("index"), py::arg("control_value"))
        .def("multiply_scalar", &QuantumGateMatrix::multiply_scalar, py::arg("value"));
    py::class_<QuantumGateDiagonalMatrix, QuantumGateBase>(m, "QuantumGateDiagonalMatrix");
    py::class_<ClsPauliGate, QuantumGateBase>(m, "ClsPauliGate")
        .def("get_pauli_id_list", &ClsPauliGate::get_pauli_id_list)
        .def("get_angle", &ClsPauliGate::get_angle);
    py::class_<QuantumGate_Probabilistic, QuantumGateBase>(m, "QuantumGate_Probabilistic")
        .def("get_distribution", &QuantumGate_Probabilistic::get_distribution)
        .def("get_cumulative_distribution", &QuantumGate_Probabilistic::get_cumulative_distribution)
        .def("is_instrument", &QuantumGate_Probabilistic::is_instrument)
        .def("optimize_ProbablisticGate", &QuantumGate_Probabilistic::optimize_ProbablisticGate)
        .def("get_gate_list",
             [](const QuantumGate_Probabilistic& g) {
                 std::vector<QuantumGateBase*> out;
                 for (const QuantumGateBase* x : g.get_gate_list()) out.push_back(x->copy());
                 return out;
             },
             py::return_value_policy::take_ownership)
        .def("set_seed", &QuantumGate_Probabilistic::set_seed, py::arg("seed"))
        .def("get_superoperator", [](const QuantumGate_Probabilistic& g) { return from_matrix(g.superoperator()); });
    py::class_<QuantumGate_LinearCombination, QuantumGateBase>(m, "QuantumGate_LinearCombination")
        .def("get_coef_list", &QuantumGate_LinearCombination::get_coef_list)
        .def("get_gate_list",
             [](const QuantumGate_LinearCombination& g) {
                 std::vector<QuantumGateBase*> out;
                 for (const QuantumGateBase* x : g.get_gate_list()) out.push_back(x->copy());
                 return out;
             },
             py::return_value_policy::take_ownership);
    py::class_<ClsStateReflectionGate, QuantumGateBase>(m, "ClsStateReflectionGate");
    py::class_<QuantumGate_Adaptive, QuantumGateBase>(m, "QuantumGate_Adaptive");
    py::class_<ClsNpairQubitGate, QuantumGateBase>(m, "ClsNpairQubitGate");
    py::class_<ClsNoisyEvolution, QuantumGateBase>(m, "ClsNoisyEvolution")
        .def("set_seed", &ClsNoisyEvolution::set_seed, py::arg("seed"))
        .def("get_effective_hamiltonian", &ClsNoisyEvolution::get_effective_hamiltonian)
        .def("set_find_collapse_max_steps", &ClsNoisyEvolution::set_find_collapse_max_steps, py::arg("n"));
    py::class_<ClsNoisyEvolution_fast, QuantumGateBase>(m, "ClsNoisyEvolution_fast")
        .def("set_seed", &ClsNoisyEvolution_fast::set_seed, py::arg("seed"))
        .def("get_effective_hamiltonian", &ClsNoisyEvolution_fast::get_effective_hamiltonian)
        .def("set_find_collapse_max_steps", &ClsNoisyEvolution_fast::set_find_collapse_max_steps, py::arg("n"));
    py::class_<QuantumGate_CPTP, QuantumGateBase>(m, "QuantumGate_CPTP")
        .def("set_seed", &QuantumGate_CPTP::set_seed, py::arg("seed"))
        .def("get_gate_list",
             [](const QuantumGate_CPTP& g) {
                 std::vector<QuantumGateBase*> out;
                 for (const QuantumGateBase* k : g.get_gate_list()) out.push_back(k->copy());
                 return out;
             },
             py::return_value_policy::take_ownership)
        .def("get_superoperator", [](const QuantumGate_CPTP& g) { return from_matrix(g.superoperator()); })
        .def("is_instrument", &QuantumGate_CPTP::is_instrument);
    py::class_<QuantumGate_CP, QuantumGate_CPTP>(m, "QuantumGate_CP");
    py::class_<QuantumGate_SingleParameter, QuantumGateBase>(m, "QuantumGate_SingleParameter")
        .def("get_parameter_value", &QuantumGate_SingleParameter::get_parameter_value)
        .def("set_parameter_value", &QuantumGate_SingleParameter::set_parameter_value, py::arg("value"));

    auto mg = m.def_submodule("gate");
    const auto own_ = py::return_value_policy::take_ownership;
    mg.def("Identity", &gate::Identity, own_, py::arg("index"));
    mg.def("X", &gate::X, own_, py::arg("index"));
    mg.def("Y", &gate::Y, own_, py::arg("index"));
    mg.def("Z", &gate::Z, own_, py::arg("index"));
    mg.def("H", &gate::H, own_, py::arg("index"));
    mg.def("S", &gate::S, own_, py::arg("index"));
    mg.def("Sdag", &gate::Sdag, own_, py::arg("index"));
    mg.def("T", &gate::T, own_, py::arg("index"));
    mg.def("Tdag", &gate::Tdag, own_, py::arg("index"));
    mg.def("sqrtX", &gate::sqrtX, own_, py::arg("index"));
    mg.def("sqrtXdag", &gate::sqrtXdag, own_, py::arg("index"));
    mg.def("sqrtY", &gate::sqrtY, own_, py::arg("index"));
    mg.def("sqrtYdag", &gate::sqrtYdag, own_, py::arg("index"));
    mg.def("P0", &gate::P0, own_, py::arg("index"));
    mg.def("P1", &gate::P1, own_, py::arg("index"));
    mg.def("U1", &gate::U1, own_, py::arg("index"), py::arg("lambda_"));
    mg.def("U2", &gate::U2, own_, py::arg("index"), py::arg("phi"), py::arg("lambda_"));
    mg.def("U3", &gate::U3, own_, py::arg("index"), py::arg("theta"), py::arg("phi"), py::arg("lambda_"));
    mg.def("RX", &gate::RX, own_, py::arg("index"), py::arg("angle"));
    mg.def("RY", &gate::RY, own_, py::arg("index"), py::arg("angle"));
    mg.def("RZ", &gate::RZ, own_, py::arg("index"), py::arg("angle"));
    mg.def("RotInvX", &gate::RotInvX, own_, py::arg("index"), py::arg("angle"));
    mg.def("RotInvY", &gate::RotInvY, own_, py::arg("index"), py::arg("angle"));
    mg.def("RotInvZ", &gate::RotInvZ, own_, py::arg("index"), py::arg("angle"));
    mg.def("RotX", &gate::RotX, own_, py::arg("index"), py::arg("angle"));
    mg.def("RotY", &gate::RotY, own_, py::arg("index"), py::arg("angle"));
    mg.def("RotZ", &gate::RotZ, own_, py::arg("index"), py::arg("angle"));
    mg.def("CNOT", &gate::CNOT, own_, py::arg("control"), py::arg("target"));
    mg.def("CZ", &gate::CZ, own_, py::arg("control"), py::arg("target"));
    mg.def("SWAP", &gate::SWAP, own_, py::arg("target1"), py::arg("target2"));
    mg.def(
        "DenseMatrix", [](const py::object& t, const carray& mat) { return gate::DenseMatrix(as_targets(t), to_matrix(mat)); },
        own_, py::arg("index_list"), py::arg("matrix"));
    mg.def(
        "DiagonalMatrix",
        [](const py::object& t, const carray& d) { return gate::DiagonalMatrix(as_targets(t), to_vector(d)); }, own_,
        py::arg("index_list"), py::arg("diagonal_element"));
    mg.def("Pauli", &gate::Pauli, own_, py::arg("index_list"), py::arg("pauli_ids"));
    mg.def("PauliRotation", &gate::PauliRotation, own_, py::arg("index_list"), py::arg("pauli_ids"), py::arg("angle"));
    mg.def(
        "RandomUnitary",
        [](const std::vector<unsigned>& t, py::object seed) {
            const uint64_t s = seed.is_none() ? std::random_device{}() : seed.cast<uint64_t>();
            return gate::RandomUnitary(t, s);
        },
        own_, py::arg("index_list"), py::arg("seed") = py::none());
    // noise factories take an optional seed for the branch draw on state vectors
    // (src/cppsim/gate_factory.cpp:403-520)
    auto seeded = [](auto* g, const py::object& seed) {
        if (!seed.is_none()) g->set_seed(seed.cast<uint64_t>());
        return g;
    };
    mg.def("DepolarizingNoise", [seeded](unsigned t, double p, py::object s) { return seeded(gate::DepolarizingNoise(t, p), s); },
           own_, py::arg("index"), py::arg("prob"), py::arg("seed") = py::none());
    mg.def("TwoQubitDepolarizingNoise",
           [seeded](unsigned a, unsigned b, double p, py::object s) {
               return seeded(gate::TwoQubitDepolarizingNoise(a, b, p), s);
           },
           own_, py::arg("index1"), py::arg("index2"), py::arg("prob"), py::arg("seed") = py::none());
    mg.def("BitFlipNoise", [seeded](unsigned t, double p, py::object s) { return seeded(gate::BitFlipNoise(t, p), s); },
           own_, py::arg("index"), py::arg("prob"), py::arg("seed") = py::none());
    mg.def("DephasingNoise", [seeded](unsigned t, double p, py::object s) { return seeded(gate::DephasingNoise(t, p), s); },
           own_, py::arg("index"), py::arg("prob"), py::arg("seed") = py::none());
    mg.def("IndependentXZNoise",
           [seeded](unsigned t, double p, py::object s) { return seeded(gate::IndependentXZNoise(t, p), s); }, own_,
           py::arg("index"), py::arg("prob"), py::arg("seed") = py::none());
    mg.def("AmplitudeDampingNoise",
           [seeded](unsigned t, double p, py::object s) { return seeded(gate::AmplitudeDampingNoise(t, p), s); }, own_,
           py::arg("index"), py::arg("prob"), py::arg("seed") = py::none());
    mg.def("Probabilistic", &gate::Probabilistic, own_, py::arg("distribution"), py::arg("gate_list"));
    mg.def("CPTP", &gate::CPTP, own_, py::arg("gate_list"));
    mg.def("to_matrix_gate", &gate::to_matrix_gate, own_, py::arg("gate"));
    mg.def("merge", &gate::merge, own_, py::arg("gate1"), py::arg("gate2"));
    mg.def(
        "merge",
        [](const std::vector<QuantumGateBase*>& gs) {
            if (gs.empty()) throw QulacsError("merge: empty gate list");
            std::unique_ptr<QuantumGateBase> acc(gate::to_matrix_gate(gs[0]));
            for (size_t i = 1; i < gs.size(); ++i) acc.reset(gate::merge(acc.get(), gs[i]));
            return static_cast<QuantumGateMatrix*>(acc.release());
        },
        own_, py::arg("gate_list"));
    mg.def("ParametricRX", &gate::ParametricRX, own_, py::arg("index"), py::arg("angle"));
    mg.def("ParametricRY", &gate::ParametricRY, own_, py::arg("index"), py::arg("angle"));
    mg.def("ParametricRZ", &gate::ParametricRZ, own_, py::arg("index"), py::arg("angle"));
    mg.def("ParametricPauliRotation", &gate::ParametricPauliRotation, own_, py::arg("index_list"),
           py::arg("pauli_ids"), py::arg("angle"));
    mg.def(
        "from_json", [](const std::string& json) { return gate_from_ptree(ptree::from_json(json)); }, own_,
        py::arg("json"));
    mg.def("TOFFOLI", &gate::TOFFOLI, own_, py::arg("control1"), py::arg("control2"), py::arg("target"));
    mg.def("FREDKIN", &gate::FREDKIN, own_, py::arg("control"), py::arg("target1"), py::arg("target2"));
    mg.def("add", py::overload_cast<const QuantumGateBase*, const QuantumGateBase*>(&gate::add), own_,
           py::arg("gate1"), py::arg("gate2"));
    mg.def("add", py::overload_cast<const std::vector<QuantumGateBase*>&>(&gate::add), own_, py::arg("gate_list"));
    mg.def("Instrument", &gate::Instrument, own_, py::arg("kraus_list"), py::arg("register"));
    mg.def("CP", &gate::CP, own_, py::arg("kraus_list"), py::arg("state_normalize"), py::arg("probability_normalize"),
           py::arg("assign_zero_if_not_matched"));
    mg.def("ProbabilisticInstrument", &gate::ProbabilisticInstrument, own_, py::arg("prob_list"), py::arg("gate_list"),
           py::arg("register"));
    mg.def("LinearCombination", &gate::LinearCombination, own_, py::arg("coefs"), py::arg("gate_list"));
    mg.def("MultiQubitPauliMeasurement", &gate::MultiQubitPauliMeasurement, own_, py::arg("target_index_list"),
           py::arg("pauli_id_list"), py::arg("classical_register_address"));
    mg.def(
        "MultiQubitPauliMeasurement",
        [](const std::vector<unsigned>& t, const std::vector<unsigned>& p, unsigned reg, uint64_t seed) {
            QuantumGate_CPTP* g = gate::MultiQubitPauliMeasurement(t, p, reg);
            g->set_seed(seed);
            return g;
        },
        own_, py::arg("target_index_list"), py::arg("pauli_id_list"), py::arg("classical_register_address"),
        py::arg("seed"));
    mg.def("StateReflection", &gate::StateReflection, own_, py::arg("state"));
    mg.def("Adaptive", py::overload_cast<const QuantumGateBase*, QuantumGate_Adaptive::Condition>(&gate::Adaptive), own_,
           py::arg("gate"), py::arg("condition"));
    mg.def("Adaptive",
           py::overload_cast<const QuantumGateBase*, QuantumGate_Adaptive::ConditionWithId, unsigned>(&gate::Adaptive),
           own_, py::arg("gate"), py::arg("condition"), py::arg("id"));
    mg.def("ReversibleBoolean", &gate::ReversibleBoolean, own_, py::arg("index_list"), py::arg("func"));
    mg.def("FusedSWAP", &gate::FusedSWAP, own_, py::arg("target1"), py::arg("target2"), py::arg("block_size"));
    mg.def(
        "NoisyEvolution",
        [](const Observable& h, const std::vector<GeneralQuantumOperator*>& c, double time, double dt) {
            return gate::NoisyEvolution(h, std::vector<const GeneralQuantumOperator*>(c.begin(), c.end()), time, dt);
        },
        own_, py::arg("hamiltonian"), py::arg("c_ops"), py::arg("time"), py::arg("dt"));
    mg.def(
        "NoisyEvolution_fast",
        [](const Observable& h, const std::vector<GeneralQuantumOperator*>& c, double time) {
            return gate::NoisyEvolution_fast(h, std::vector<const GeneralQuantumOperator*>(c.begin(), c.end()), time);
        },
        own_, py::arg("hamiltonian"), py::arg("c_ops"), py::arg("time"));
    mg.def("Measurement", &gate::Measurement, own_, py::arg("index"), py::arg("register"));
    mg.def(
        "Measurement",
        [](unsigned q, unsigned reg, uint64_t seed) {
            QuantumGate_CPTP* g = gate::Measurement(q, reg);
            g->set_seed(seed);
            return g;
        },
        own_, py::arg("index"), py::arg("register"), py::arg("seed"));
    mg.def(
        "SparseMatrix",
        [](const std::vector<unsigned>& targets, const py::object& sparse) {
            // a scipy.sparse matrix, applied through its dense form
            const carray dense = sparse.attr("toarray")().cast<carray>();
            return gate::DenseMatrix(targets, to_matrix(dense));
        },
        own_, py::arg("index_list"), py::arg("matrix"));

    // ---------------------------------------------------------------- circuits
    py::class_<QuantumCircuit>(m, "QuantumCircuit")
        .def(py::init<unsigned>(), py::arg("qubit_count"))
        .def("merge_circuit", &QuantumCircuit::merge_circuit, py::arg("circuit"))
        .def("add_FusedSWAP_gate",
             [](QuantumCircuit& c, unsigned t1, unsigned t2, unsigned b) { c.add_gate_consume(gate::FusedSWAP(t1, t2, b)); },
             py::arg("target1"), py::arg("target2"), py::arg("block_size"))
        .def("add_noise_gate", &QuantumCircuit::add_noise_gate, py::arg("gate"), py::arg("NoiseType"),
             py::arg("NoiseProbability"))
        .def("add_observable_rotation_gate", &QuantumCircuit::add_observable_rotation_gate, py::arg("observable"),
             py::arg("angle"), py::arg("repeat"))
        .def("add_diagonal_observable_rotation_gate", &QuantumCircuit::add_diagonal_observable_rotation_gate,
             py::arg("observable"), py::arg("angle"))
        .def("to_json", [](const QuantumCircuit& c) { return ptree::to_json(to_ptree(c)); })
        .def(py::pickle([](const QuantumCircuit& c) { return ptree::to_json(to_ptree(c)); },
                        [](const std::string& json) {
                            const PTree pt = ptree::from_json(json);
                            if (pt.get("name") != "QuantumCircuit")
                                throw QulacsError("pickle: not a QuantumCircuit: " + pt.get("name"));
                            return std::unique_ptr<QuantumCircuit>(circuit_from_ptree(pt));
                        }))
        .def("copy", &QuantumCircuit::copy, py::return_value_policy::take_ownership)
        .def("add_gate", py::overload_cast<const QuantumGateBase&>(&QuantumCircuit::add_gate), py::arg("gate"))
        .def("add_gate", py::overload_cast<const QuantumGateBase&, size_t>(&QuantumCircuit::add_gate), py::arg("gate"),
             py::arg("position"))
        .def("remove_gate", &QuantumCircuit::remove_gate, py::arg("position"))
        .def("get_gate", [](const QuantumCircuit& c, size_t i) { return c.get_gate(i)->copy(); },
             py::return_value_policy::take_ownership, py::arg("position"))
        .def("get_gate_count", &QuantumCircuit::get_gate_count)
        .def("get_qubit_count", &QuantumCircuit::get_qubit_count)
        .def("update_quantum_state", py::overload_cast<QuantumStateBase*>(&QuantumCircuit::update_quantum_state),
             py::arg("state"))
        .def("update_quantum_state",
             py::overload_cast<QuantumStateBase*, size_t, size_t>(&QuantumCircuit::update_quantum_state),
             py::arg("state"), py::arg("start"), py::arg("end"))
        .def("calculate_depth", &QuantumCircuit::calculate_depth)
        .def("get_inverse", &QuantumCircuit::get_inverse, py::return_value_policy::take_ownership)
        .def("to_string", &QuantumCircuit::to_string)
        .def("__str__", &QuantumCircuit::to_string)
        .def("add_X_gate", &QuantumCircuit::add_X_gate, py::arg("index"))
        .def("add_Y_gate", &QuantumCircuit::add_Y_gate, py::arg("index"))
        .def("add_Z_gate", &QuantumCircuit::add_Z_gate, py::arg("index"))
        .def("add_H_gate", &QuantumCircuit::add_H_gate, py::arg("index"))
        .def("add_S_gate", &QuantumCircuit::add_S_gate, py::arg("index"))
        .def("add_Sdag_gate", &QuantumCircuit::add_Sdag_gate, py::arg("index"))
        .def("add_T_gate", &QuantumCircuit::add_T_gate, py::arg("index"))
        .def("add_Tdag_gate", &QuantumCircuit::add_Tdag_gate, py::arg("index"))
        .def("add_sqrtX_gate", &QuantumCircuit::add_sqrtX_gate, py::arg("index"))
        .def("add_sqrtXdag_gate", &QuantumCircuit::add_sqrtXdag_gate, py::arg("index"))
        .def("add_sqrtY_gate", &QuantumCircuit::add_sqrtY_gate, py::arg("index"))
        .def("add_sqrtYdag_gate", &QuantumCircuit::add_sqrtYdag_gate, py::arg("index"))
        .def("add_P0_gate", &QuantumCircuit::add_P0_gate, py::arg("index"))
        .def("add_P1_gate", &QuantumCircuit::add_P1_gate, py::arg("index"))
        .def("add_CNOT_gate", &QuantumCircuit::add_CNOT_gate, py::arg("control"), py::arg("target"))
        .def("add_CZ_gate", &QuantumCircuit::add_CZ_gate, py::arg("control"), py::arg("target"))
        .def("add_SWAP_gate", &QuantumCircuit::add_SWAP_gate, py::arg("target1"), py::arg("target2"))
        .def("add_RX_gate", &QuantumCircuit::add_RX_gate, py::arg("index"), py::arg("angle"))
        .def("add_RY_gate", &QuantumCircuit::add_RY_gate, py::arg("index"), py::arg("angle"))
        .def("add_RZ_gate", &QuantumCircuit::add_RZ_gate, py::arg("index"), py::arg("angle"))
        .def("add_RotX_gate", &QuantumCircuit::add_RotX_gate, py::arg("index"), py::arg("angle"))
        .def("add_RotY_gate", &QuantumCircuit::add_RotY_gate, py::arg("index"), py::arg("angle"))
        .def("add_RotZ_gate", &QuantumCircuit::add_RotZ_gate, py::arg("index"), py::arg("angle"))
        .def("add_RotInvX_gate", &QuantumCircuit::add_RotInvX_gate, py::arg("index"), py::arg("angle"))
        .def("add_RotInvY_gate", &QuantumCircuit::add_RotInvY_gate, py::arg("index"), py::arg("angle"))
        .def("add_RotInvZ_gate", &QuantumCircuit::add_RotInvZ_gate, py::arg("index"), py::arg("angle"))
        .def("add_U1_gate", &QuantumCircuit::add_U1_gate, py::arg("index"), py::arg("lambda_"))
        .def("add_U2_gate", &QuantumCircuit::add_U2_gate, py::arg("index"), py::arg("phi"), py::arg("lambda_"))
        .def("add_U3_gate", &QuantumCircuit::add_U3_gate, py::arg("index"), py::arg("theta"), py::arg("phi"),
             py::arg("lambda_"))
        .def("add_dense_matrix_gate",
             [](QuantumCircuit& c, const py::object& t, const carray& mat) {
                 c.add_dense_matrix_gate(as_targets(t), to_matrix(mat));
             },
             py::arg("index_list"), py::arg("matrix"))
        .def("add_diagonal_matrix_gate",
             [](QuantumCircuit& c, const py::object& t, const carray& d) {
                 c.add_diagonal_matrix_gate(as_targets(t), to_vector(d));
             },
             py::arg("index_list"), py::arg("diagonal_element"))
        .def("add_multi_Pauli_gate", &QuantumCircuit::add_multi_Pauli_gate, py::arg("index_list"),
             py::arg("pauli_ids"))
        .def("add_multi_Pauli_rotation_gate", &QuantumCircuit::add_multi_Pauli_rotation_gate, py::arg("index_list"),
             py::arg("pauli_ids"), py::arg("angle"))
        .def(
            "add_random_unitary_gate",
            [](QuantumCircuit& c, const std::vector<unsigned>& t, py::object seed) {
                const uint64_t s = seed.is_none() ? std::random_device{}() : seed.cast<uint64_t>();
                c.add_random_unitary_gate(t, s);
            },
            py::arg("index_list"), py::arg("seed") = py::none());

    py::class_<ParametricQuantumCircuit, QuantumCircuit>(m, "ParametricQuantumCircuit")
        .def(py::init<unsigned>(), py::arg("qubit_count"))
        .def("merge_circuit", py::overload_cast<const ParametricQuantumCircuit&>(&ParametricQuantumCircuit::merge_circuit),
             py::arg("circuit"))
        .def("copy", &ParametricQuantumCircuit::copy, py::return_value_policy::take_ownership)
        .def("add_parametric_gate",
             py::overload_cast<const QuantumGate_SingleParameter&>(&ParametricQuantumCircuit::add_parametric_gate),
             py::arg("gate"))
        .def("add_parametric_gate",
             py::overload_cast<const QuantumGate_SingleParameter&, size_t>(
                 &ParametricQuantumCircuit::add_parametric_gate),
             py::arg("gate"), py::arg("position"))
        .def("add_gate", py::overload_cast<const QuantumGateBase&>(&ParametricQuantumCircuit::add_gate),
             py::arg("gate"))
        .def("add_gate", py::overload_cast<const QuantumGateBase&, size_t>(&ParametricQuantumCircuit::add_gate),
             py::arg("gate"), py::arg("position"))
        .def("remove_gate", &ParametricQuantumCircuit::remove_gate, py::arg("position"))
        .def("add_parametric_RX_gate", &ParametricQuantumCircuit::add_parametric_RX_gate, py::arg("index"),
             py::arg("angle"))
        .def("add_parametric_RY_gate", &ParametricQuantumCircuit::add_parametric_RY_gate, py::arg("index"),
             py::arg("angle"))
        .def("add_parametric_RZ_gate", &ParametricQuantumCircuit::add_parametric_RZ_gate, py::arg("index"),
             py::arg("angle"))
        .def("add_parametric_multi_Pauli_rotation_gate",
             &ParametricQuantumCircuit::add_parametric_multi_Pauli_rotation_gate, py::arg("index_list"),
             py::arg("pauli_ids"), py::arg("angle"))
        .def("get_parameter_count", &ParametricQuantumCircuit::get_parameter_count)
        .def("get_parameter", &ParametricQuantumCircuit::get_parameter, py::arg("index"))
        .def("set_parameter", &ParametricQuantumCircuit::set_parameter, py::arg("index"), py::arg("parameter"))
        .def("get_parametric_gate_position", &ParametricQuantumCircuit::get_parametric_gate_position,
             py::arg("index"))
        .def("backprop", &ParametricQuantumCircuit::backprop, py::arg("obs"))
        .def("backprop_inner_product", &ParametricQuantumCircuit::backprop_inner_product, py::arg("state"))
        .def(py::pickle([](const ParametricQuantumCircuit& c) { return ptree::to_json(to_ptree(c)); },
                        [](const std::string& json) {
                            return std::unique_ptr<ParametricQuantumCircuit>(
                                parametric_circuit_from_ptree(ptree::from_json(json)));
                        }));

    py::class_<GradCalculator>(m, "GradCalculator")
        .def(py::init<>())
        .def("calculate_grad",
             py::overload_cast<ParametricQuantumCircuit&, const GeneralQuantumOperator&>(
                 &GradCalculator::calculate_grad),
             py::arg("parametric_circuit"), py::arg("observable"))
        .def("calculate_grad",
             py::overload_cast<ParametricQuantumCircuit&, const GeneralQuantumOperator&, const std::vector<double>&>(
                 &GradCalculator::calculate_grad),
             py::arg("parametric_circuit"), py::arg("observable"), py::arg("angles_of_gates"));

    auto mcircuit = m.def_submodule("circuit");
    mcircuit.def(
        "from_json", [](const std::string& json) { return circuit_from_ptree(ptree::from_json(json)); },
        py::return_value_policy::take_ownership, py::arg("json"));
    py::class_<QuantumCircuitOptimizer>(mcircuit, "QuantumCircuitOptimizer")
        .def(py::init([](unsigned mpi_size) {
                 if (mpi_size > 1) throw QulacsError("QuantumCircuitOptimizer: one device only");
                 return new QuantumCircuitOptimizer();
             }),
             py::arg("mpi_size") = 0)
        .def("optimize", &QuantumCircuitOptimizer::optimize, py::arg("circuit"), py::arg("block_size"),
             py::arg("swap_level") = 0)
        .def("optimize_light", &QuantumCircuitOptimizer::optimize_light, py::arg("circuit"), py::arg("swap_level") = 0)
        .def("merge_all", &QuantumCircuitOptimizer::merge_all, py::return_value_policy::take_ownership,
             py::arg("circuit"));

    py::class_<NoiseSimulator::Result>(m, "SimulationResult")
        .def("get_count", [](const NoiseSimulator::Result& r) { return (unsigned)r.result.size(); })
        .def(
            "get_state",
            [](const NoiseSimulator::Result& r, unsigned i) {
                const QuantumStateMI355* s = r.result.at(i).first;
                auto* c = new QuantumStateMI355(s->qubit_count, s->device_number);
                c->load(s);
                return c;
            },
            py::return_value_policy::take_ownership, py::arg("i"))
        .def("get_frequency", [](const NoiseSimulator::Result& r, unsigned i) { return r.result.at(i).second; },
             py::arg("i"))
        .def("sampling", &NoiseSimulator::Result::sampling);
    py::class_<NoiseSimulator>(m, "NoiseSimulator")
        .def(py::init<const QuantumCircuit*, const QuantumStateMI355*>(), py::arg("circuit"),
             py::arg("init_state") = nullptr)
        .def("execute", &NoiseSimulator::execute, py::arg("sample_count"))
        .def("execute_and_get_result", &NoiseSimulator::execute_and_get_result, py::return_value_policy::take_ownership,
             py::arg("execution_count"))
        .def("set_seed", &NoiseSimulator::set_seed, py::arg("seed"));
    py::class_<CausalConeSimulator>(m, "CausalConeSimulator")
        .def(py::init<const ParametricQuantumCircuit&, const Observable&>(), py::arg("circuit"), py::arg("observable"))
        .def("build", &CausalConeSimulator::build)
        .def("get_expectation_value", &CausalConeSimulator::get_expectation_value)
        .def("get_circuit_list",
             [](const CausalConeSimulator& c) {
                 py::list out;
                 for (const auto& row : c.get_circuit_list()) {
                     py::list r;
                     for (const auto* x : row)
                         r.append(py::cast(static_cast<ParametricQuantumCircuit*>(x->copy()),
                                           py::return_value_policy::take_ownership));
                     out.append(r);
                 }
                 return out;
             })
        .def("get_pauli_operator_list", &CausalConeSimulator::get_pauli_operator_list)
        .def("get_coef_list", &CausalConeSimulator::get_coef_list);
    py::class_<QuantumCircuitSimulator>(m, "QuantumCircuitSimulator")
        .def(py::init<QuantumCircuit*, QuantumStateBase*>(), py::arg("circuit"), py::arg("state"),
             py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
        .def("initialize_state", &QuantumCircuitSimulator::initialize_state, py::arg("computational_basis") = 0)
        .def("initialize_random_state", py::overload_cast<>(&QuantumCircuitSimulator::initialize_random_state))
        .def("initialize_random_state", py::overload_cast<unsigned>(&QuantumCircuitSimulator::initialize_random_state),
             py::arg("seed"))
        .def("simulate", &QuantumCircuitSimulator::simulate)
        .def("simulate_range", &QuantumCircuitSimulator::simulate_range, py::arg("start"), py::arg("end"))
        .def("get_expectation_value", &QuantumCircuitSimulator::get_expectation_value, py::arg("observable"))
        .def("get_gate_count", &QuantumCircuitSimulator::get_gate_count)
        .def("copy_state_to_buffer", &QuantumCircuitSimulator::copy_state_to_buffer)
        .def("copy_state_from_buffer", &QuantumCircuitSimulator::copy_state_from_buffer)
        .def("swap_state_and_buffer", &QuantumCircuitSimulator::swap_state_and_buffer);

    // ---------------------------------------------------------------- runtime controls
    m.def("set_synchronous", [](bool on) { qm_set_synchronous(on ? 1 : 0); }, py::arg("on"));
    m.def("get_synchronous", []() { return qm_get_synchronous() != 0; });
    m.def("synchronize", [](const QuantumStateBase& s) {
        qm_synchronize(s.get_cuda_stream(), s.device_number);
        check_abi("synchronize");
    });
    m.def("get_num_device", &get_num_device);
    m.def("sampling_margin", &qm_sampling_margin, py::arg("gate_count"));
}
