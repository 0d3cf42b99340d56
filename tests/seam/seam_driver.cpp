// Test infrastructure: a program written against the reference's C++ seam
// declarations (seam_decls.hpp), as cppsim's GPU build is.  It reads a gate
// list from argv[1], runs it on a fresh |0...0> state through the mangled
// *_host entry points, and writes the final state (complex128, binary) to
// argv[2] and one line per requested scalar to stdout.
//
// Input lines:  n <qubits> | d1 t <8 doubles> | d2 t0 t1 <32 doubles> |
//   dk k t... <2*4^k doubles> | cd c v t <8 doubles> | diag1 t <4 doubles> |
//   cnot c t | cz c t | swap a b | rx|ry|rz t angle | h|s|t|sqrtx t |
//   prot k t... p... angle | exp k t... p... | norm | m0 t | inner
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "seam_decls.hpp"

static std::vector<CPPCTYPE> read_mat(std::istringstream& in, size_t count) {
    std::vector<CPPCTYPE> m(count);
    for (auto& z : m) {
        double re, im;
        in >> re >> im;
        z = CPPCTYPE(re, im);
    }
    return m;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: seam_driver OPS OUT\n");
        return 2;
    }
    if (get_num_device() < 1) {
        std::fprintf(stderr, "no device\n");
        return 3;
    }
    std::ifstream f(argv[1]);
    std::string line;
    unsigned n = 0;
    ITYPE dim = 0;
    void* stream = allocate_cuda_stream_host(1, 0);
    void* state = nullptr;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string op;
        in >> op;
        if (op.empty()) continue;
        if (op == "n") {
            in >> n;
            dim = 1ULL << n;
            state = allocate_quantum_state_host(dim, 0);
            initialize_quantum_state_host(state, dim, stream, 0);
            continue;
        }
        if (!state) return 4;
        unsigned a = 0, b = 0, c = 0;
        if (op == "d1") {
            in >> a;
            const auto m = read_mat(in, 4);
            single_qubit_dense_matrix_gate_host(a, m.data(), state, dim, stream, 0);
        } else if (op == "d2") {
            in >> a >> b;
            const auto m = read_mat(in, 16);
            double_qubit_dense_matrix_gate_host(a, b, m.data(), state, dim, stream, 0);
        } else if (op == "dk") {
            unsigned k;
            in >> k;
            std::vector<UINT> t(k);
            for (auto& x : t) in >> x;
            const auto m = read_mat(in, (size_t)1 << (2 * k));
            multi_qubit_dense_matrix_gate_host(t.data(), k, m.data(), state, dim, stream, 0);
        } else if (op == "cd") {
            in >> a >> b >> c;
            const auto m = read_mat(in, 4);
            single_qubit_control_single_qubit_dense_matrix_gate_host(a, b, c, m.data(), state, dim, stream, 0);
        } else if (op == "diag1") {
            in >> a;
            const auto d = read_mat(in, 2);
            single_qubit_diagonal_matrix_gate_host(a, d.data(), state, dim, stream, 0);
        } else if (op == "cnot" || op == "cz" || op == "swap") {
            in >> a >> b;
            if (op == "cnot") CNOT_gate_host(a, b, state, dim, stream, 0);
            else if (op == "cz") CZ_gate_host(a, b, state, dim, stream, 0);
            else SWAP_gate_host(a, b, state, dim, stream, 0);
        } else if (op == "rx" || op == "ry" || op == "rz") {
            double ang;
            in >> a >> ang;
            if (op == "rx") RX_gate_host(a, ang, state, dim, stream, 0);
            else if (op == "ry") RY_gate_host(a, ang, state, dim, stream, 0);
            else RZ_gate_host(a, ang, state, dim, stream, 0);
        } else if (op == "h" || op == "s" || op == "t" || op == "sqrtx") {
            in >> a;
            if (op == "h") H_gate_host(a, state, dim, stream, 0);
            else if (op == "s") S_gate_host(a, state, dim, stream, 0);
            else if (op == "t") T_gate_host(a, state, dim, stream, 0);
            else sqrtX_gate_host(a, state, dim, stream, 0);
        } else if (op == "prot" || op == "exp") {
            unsigned k;
            in >> k;
            std::vector<UINT> t(k), p(k);
            for (auto& x : t) in >> x;
            for (auto& x : p) in >> x;
            if (op == "prot") {
                double ang;
                in >> ang;
                multi_qubit_Pauli_rotation_gate_partial_list_host(t.data(), p.data(), k, ang, state, dim, stream, 0);
            } else {
                const double e = expectation_value_multi_qubit_Pauli_operator_partial_list_host(t.data(), p.data(), k,
                                                                                                state, dim, stream, 0);
                std::printf("exp %.17g\n", e);
            }
        } else if (op == "norm") {
            std::printf("norm %.17g\n", state_norm_squared_host(state, dim, stream, 0));
        } else if (op == "m0") {
            in >> a;
            std::printf("m0 %.17g\n", M0_prob_host(a, state, dim, stream, 0));
        } else if (op == "inner") {
            const CPPCTYPE z = inner_product_host(state, state, dim, stream, 0);
            std::printf("inner %.17g %.17g\n", z.real(), z.imag());
        } else {
            std::fprintf(stderr, "unknown op %s\n", op.c_str());
            return 5;
        }
    }
    std::vector<CPPCTYPE> host(dim);
    get_quantum_state_host(state, host.data(), dim, stream, 0);
    std::ofstream out(argv[2], std::ios::binary);
    out.write(reinterpret_cast<const char*>(host.data()), (std::streamsize)(dim * sizeof(CPPCTYPE)));
    release_quantum_state_host(state, 0);
    release_cuda_stream_host(stream, 1, 0);
    return 0;
}
