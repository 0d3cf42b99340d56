// JSON wire format of states, gates, circuits and operators.
//
// The layout is the reference's property-tree layout (to_ptree / from_ptree:
// src/cppsim/state.hpp:994-1003, state_dm.hpp:424-434, state_dm.cpp:89-118,
// gate_factory.cpp:696-900, circuit.cpp:559-595, vqcsim/parametric_circuit.cpp:337-378,
// vqcsim/parametric_gate_factory.cpp:96-123, pauli_operator.cpp:290-300,
// general_quantum_operator.cpp:535-545 and 953-998, observable.cpp:351-366), written
// the way boost::property_tree::write_json writes it (src/cppsim/utility.cpp:298-308):
// every leaf is a JSON string, a node whose children all have empty keys is a JSON
// array, an empty node is "".  So a document written by the reference loads here and
// one written here loads there.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "qm_host.hpp"

namespace qmh {

// A property tree: a leaf value or ordered (key, child) pairs.
struct PTree {
    std::string value;
    std::vector<std::pair<std::string, PTree>> children;

    PTree() = default;
    explicit PTree(std::string v) : value(std::move(v)) {}
    PTree& put(const std::string& key, const std::string& v);
    PTree& put(const std::string& key, PTree child);
    PTree& push_back(PTree child);  // array element (empty key)
    const PTree& child(const std::string& key) const;  // throws QulacsError when absent
    const std::string& get(const std::string& key) const;
    unsigned get_uint(const std::string& key) const;
    double get_double(const std::string& key) const;
    bool get_bool(const std::string& key) const;
};

namespace ptree {
std::string to_json(const PTree& pt);
PTree from_json(const std::string& json);
std::string format_double(double v);  // shortest round-trip text (17 significant digits)
}  // namespace ptree

PTree to_ptree(const QuantumStateBase& state);
QuantumStateBase* state_from_ptree(const PTree& pt, unsigned device = 0);

PTree to_ptree(const QuantumGateBase& gate);
QuantumGateBase* gate_from_ptree(const PTree& pt);
QuantumGate_SingleParameter* parametric_gate_from_ptree(const PTree& pt);

PTree to_ptree(const QuantumCircuit& circuit);  // "ParametricQuantumCircuit" for that subclass
QuantumCircuit* circuit_from_ptree(const PTree& pt);
ParametricQuantumCircuit* parametric_circuit_from_ptree(const PTree& pt);

PTree to_ptree(const PauliOperator& op);
PauliOperator pauli_operator_from_ptree(const PTree& pt);
PTree to_ptree(const GeneralQuantumOperator& op);
GeneralQuantumOperator* operator_from_ptree(const PTree& pt);
Observable* observable_from_ptree(const PTree& pt);  // throws for a non-Hermitian term list

}  // namespace qmh
