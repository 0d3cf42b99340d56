// JSON wire format: property tree, writer / reader, and the to_ptree /
// from_ptree of every serialisable object.  Key names and nesting follow the
// reference files cited in qm_json.hpp; the code is this library's.
#include "qm_json.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>

namespace qmh {

// ------------------------------------------------------------------ PTree
PTree& PTree::put(const std::string& key, const std::string& v) { return put(key, PTree(v)); }
PTree& PTree::put(const std::string& key, PTree child) {
    children.emplace_back(key, std::move(child));
    return children.back().second;
}
PTree& PTree::push_back(PTree child) { return put("", std::move(child)); }
const PTree& PTree::child(const std::string& key) const {
    for (const auto& kv : children)
        if (kv.first == key) return kv.second;
    throw QulacsError("JSON: missing property \"" + key + "\"");
}
const std::string& PTree::get(const std::string& key) const { return child(key).value; }
unsigned PTree::get_uint(const std::string& key) const {
    const std::string& s = get(key);
    char* end = nullptr;
    const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
    if (s.empty() || *end != '\0' || v > 0xffffffffULL || s[0] == '-')
        throw QulacsError("JSON: property \"" + key + "\" is not an unsigned integer: " + s);
    return (unsigned)v;
}
double PTree::get_double(const std::string& key) const {
    const std::string& s = get(key);
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (s.empty() || *end != '\0') throw QulacsError("JSON: property \"" + key + "\" is not a number: " + s);
    return v;
}
bool PTree::get_bool(const std::string& key) const {
    const std::string& s = get(key);
    if (s == "true" || s == "1") return true;
    if (s == "false" || s == "0") return false;
    throw QulacsError("JSON: property \"" + key + "\" is not a boolean: " + s);
}

namespace ptree {
namespace {
void quote(std::string& out, const std::string& s) {
    out += '"';
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", c);
                    out += buf;
                } else {
                    out += (char)c;
                }
        }
    }
    out += '"';
}
void write(std::string& out, const PTree& pt, int indent) {
    if (pt.children.empty()) {
        quote(out, pt.value);
        return;
    }
    bool array = true;
    for (const auto& kv : pt.children) array = array && kv.first.empty();
    out += array ? "[\n" : "{\n";
    const std::string pad(indent + 4, ' ');
    for (size_t i = 0; i < pt.children.size(); ++i) {
        out += pad;
        if (!array) {
            quote(out, pt.children[i].first);
            out += ": ";
        }
        write(out, pt.children[i].second, indent + 4);
        out += i + 1 < pt.children.size() ? ",\n" : "\n";
    }
    out += std::string(indent, ' ');
    out += array ? "]" : "}";
}

struct Reader {
    const std::string& s;
    size_t i = 0;
    [[noreturn]] void fail(const char* what) const {
        throw QulacsError(std::string("from_json: ") + what + " at offset " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    static void utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xc0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3f));
        } else {
            out += (char)(0xe0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3f));
            out += (char)(0x80 | (cp & 0x3f));
        }
    }
    std::string string() {
        if (!eat('"')) fail("expected a string");
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size()) fail("unterminated escape");
            c = s[i++];
            switch (c) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    if (i + 4 > s.size()) fail("short \\u escape");
                    const unsigned cp = (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
                    i += 4;
                    utf8(out, cp);
                    break;
                }
                default: fail("unknown escape");
            }
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
    PTree value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        const char c = s[i];
        PTree pt;
        if (c == '{') {
            ++i;
            if (eat('}')) return pt;
            do {
                ws();
                std::string key = string();
                if (!eat(':')) fail("expected ':'");
                pt.put(key, value());
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
        } else if (c == '[') {
            ++i;
            if (eat(']')) return pt;
            do pt.push_back(value());
            while (eat(','));
            if (!eat(']')) fail("expected ']'");
        } else if (c == '"') {
            pt.value = string();
        } else {
            // number, true, false or null: kept as text (the tree stores strings)
            const size_t b = i;
            while (i < s.size() && s[i] != ',' && s[i] != '}' && s[i] != ']' && s[i] != ' ' && s[i] != '\n' &&
                   s[i] != '\t' && s[i] != '\r')
                ++i;
            pt.value = s.substr(b, i - b);
            if (pt.value == "null") pt.value.clear();
            if (pt.value.empty() && s.compare(b, 4, "null") != 0) fail("expected a value");
        }
        return pt;
    }
};
}  // namespace

std::string to_json(const PTree& pt) {
    std::string out;
    write(out, pt, 0);
    out += '\n';
    return out;
}
PTree from_json(const std::string& json) {
    Reader r{json};
    PTree pt = r.value();
    r.ws();
    if (r.i != json.size()) r.fail("trailing characters");
    return pt;
}
std::string format_double(double v) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}
}  // namespace ptree

// ------------------------------------------------------------------ helpers
namespace {
std::string u2s(unsigned v) { return std::to_string(v); }
PTree complex_pt(CPPCTYPE z) {
    PTree pt;
    pt.put("real", ptree::format_double(z.real()));
    pt.put("imag", ptree::format_double(z.imag()));
    return pt;
}
CPPCTYPE complex_from(const PTree& pt) { return CPPCTYPE(pt.get_double("real"), pt.get_double("imag")); }
PTree complex_array_pt(const std::vector<CPPCTYPE>& v) {
    PTree pt;
    for (const CPPCTYPE& z : v) pt.push_back(complex_pt(z));
    return pt;
}
std::vector<CPPCTYPE> complex_array_from(const PTree& pt) {
    std::vector<CPPCTYPE> v;
    v.reserve(pt.children.size());
    for (const auto& kv : pt.children) v.push_back(complex_from(kv.second));
    return v;
}
PTree uint_array_pt(const std::vector<unsigned>& v) {
    PTree pt;
    for (unsigned x : v) pt.push_back(PTree(u2s(x)));
    return pt;
}
std::vector<unsigned> uint_array_from(const PTree& pt) {
    std::vector<unsigned> v;
    for (const auto& kv : pt.children) {
        PTree leaf;
        leaf.put("v", kv.second.value);
        v.push_back(leaf.get_uint("v"));
    }
    return v;
}
PTree control_list_pt(const std::vector<unsigned>& controls, const std::vector<unsigned>& values) {
    PTree pt;
    for (size_t i = 0; i < controls.size(); ++i) {
        PTree c;
        c.put("index", u2s(controls[i]));
        c.put("value", u2s(values[i]));
        pt.push_back(std::move(c));
    }
    return pt;
}
void control_list_from(const PTree& pt, std::vector<unsigned>& controls, std::vector<unsigned>& values) {
    for (const auto& kv : pt.children) {
        controls.push_back(kv.second.get_uint("index"));
        values.push_back(kv.second.get_uint("value"));
    }
}
PTree matrix_pt(const ComplexMatrix& m) {  // row-major data, as the reference's ComplexMatrix
    PTree pt;
    pt.put("rows", u2s((unsigned)m.rows));
    pt.put("cols", u2s((unsigned)m.cols));
    pt.put("data", complex_array_pt(m.data));
    return pt;
}
ComplexMatrix matrix_from(const PTree& pt) {
    ComplexMatrix m(pt.get_uint("rows"), pt.get_uint("cols"));
    std::vector<CPPCTYPE> d = complex_array_from(pt.child("data"));
    if (d.size() != m.rows * m.cols) throw QulacsError("JSON: matrix data size does not match rows x cols");
    m.data = std::move(d);
    return m;
}
ComplexMatrix sparse_matrix_from(const PTree& pt) {  // triplets (row, col, value) into a dense matrix
    ComplexMatrix m(pt.get_uint("rows"), pt.get_uint("cols"));
    for (const auto& kv : pt.child("data").children) {
        const unsigned r = kv.second.get_uint("row"), c = kv.second.get_uint("col");
        if (r >= m.rows || c >= m.cols) throw QulacsError("JSON: sparse matrix entry out of range");
        m(r, c) += complex_from(kv.second.child("value"));
    }
    return m;
}
PTree pauli_pt(const std::vector<unsigned>& targets, const std::vector<unsigned>& paulis, CPPCTYPE coef) {
    PTree pt;
    pt.put("name", "PauliOperator");
    PTree list;
    for (size_t i = 0; i < targets.size(); ++i) {
        PTree single;
        single.put("name", "SinglePauliOperator");
        single.put("index", u2s(targets[i]));
        single.put("pauli_id", u2s(paulis[i]));
        list.push_back(std::move(single));
    }
    pt.put("pauli_list", std::move(list));
    pt.put("coef", complex_pt(coef));
    return pt;
}
void expect_name(const PTree& pt, const char* want) {
    const std::string& name = pt.get("name");
    if (name != want) throw QulacsError("unknown value for property \"name\":" + name);
}
}  // namespace

// ------------------------------------------------------------------ states
PTree to_ptree(const QuantumStateBase& state) {
    PTree pt;
    pt.put("name", state.is_state_vector() ? "QuantumState" : "DensityMatrix");
    pt.put("qubit_count", u2s(state.qubit_count));
    pt.put("classical_register", uint_array_pt(state.get_classical_register()));
    pt.put(state.is_state_vector() ? "state_vector" : "density_matrix", complex_array_pt(state.duplicate_data_cpp()));
    return pt;
}
QuantumStateBase* state_from_ptree(const PTree& pt, unsigned device) {
    const std::string& name = pt.get("name");
    const unsigned n = pt.get_uint("qubit_count");
    std::unique_ptr<QuantumStateBase> s;
    std::vector<CPPCTYPE> data;
    if (name == "QuantumState") {
        s.reset(new QuantumStateMI355(n, device));
        data = complex_array_from(pt.child("state_vector"));
    } else if (name == "DensityMatrix") {
        s.reset(new DensityMatrixMI355(n, device));
        data = complex_array_from(pt.child("density_matrix"));
    } else {
        throw QulacsError("unknown value for property \"name\":" + name);
    }
    const std::vector<unsigned> reg = uint_array_from(pt.child("classical_register"));
    for (unsigned i = 0; i < reg.size(); ++i) s->set_classical_value(i, reg[i]);
    s->load(data);
    return s.release();
}

// ------------------------------------------------------------------ gates
PTree to_ptree(const QuantumGateBase& g) {
    PTree pt;
    const auto& targets = g.get_target_index_list();
    const auto& controls = g.get_control_index_list();
    const auto& cvals = g.get_control_value_list();
    if (const auto* m = dynamic_cast<const QuantumGateMatrix*>(&g)) {
        using K = QuantumGateMatrix;
        switch (m->kind) {
            case K::IDENTITY: case K::X: case K::Y: case K::Z: case K::H: case K::S: case K::SDAG: case K::T:
            case K::TDAG: case K::SQRTX: case K::SQRTXDAG: case K::SQRTY: case K::SQRTYDAG: case K::P0: case K::P1:
                pt.put("name", m->get_name() + "Gate");
                pt.put("target_qubit", u2s(targets[0]));
                return pt;
            case K::RX: case K::RY: case K::RZ:
                pt.put("name", m->get_name() + "Gate");
                pt.put("target_qubit", u2s(targets[0]));
                pt.put("angle", ptree::format_double(m->angle));
                return pt;
            case K::CNOT: case K::CZ:
                pt.put("name", m->get_name() + "Gate");
                pt.put("control_qubit", u2s(controls[0]));
                pt.put("target_qubit", u2s(targets[0]));
                return pt;
            case K::SWAP:
                pt.put("name", "SWAPGate");
                pt.put("target_qubit_list", uint_array_pt(targets));
                return pt;
            default:
                pt.put("name", "DenseMatrixGate");
                pt.put("target_qubit_list", uint_array_pt(targets));
                pt.put("control_qubit_list", control_list_pt(controls, cvals));
                pt.put("matrix", matrix_pt(m->get_matrix()));
                return pt;
        }
    }
    if (const auto* d = dynamic_cast<const QuantumGateDiagonalMatrix*>(&g)) {
        pt.put("name", "DiagonalMatrixGate");
        pt.put("target_qubit_list", uint_array_pt(targets));
        pt.put("control_qubit_list", control_list_pt(controls, cvals));
        PTree vec;
        vec.put("size", u2s((unsigned)d->get_diagonal().size()));
        vec.put("data", complex_array_pt(d->get_diagonal()));
        pt.put("vector", std::move(vec));
        return pt;
    }
    if (const auto* p = dynamic_cast<const ClsPauliGate*>(&g)) {
        pt.put("name", p->is_rotation() ? "PauliRotationGate" : "PauliGate");
        if (p->is_rotation()) pt.put("angle", ptree::format_double(p->get_angle()));
        pt.put("pauli", pauli_pt(targets, p->get_pauli_id_list(), 1.0));
        return pt;
    }
    if (const auto* p = dynamic_cast<const QuantumGate_SingleParameter*>(&g)) {
        pt.put("name", p->get_name() + "Gate");
        if (p->get_name() == "ParametricPauliRotation") {
            pt.put("angle", ptree::format_double(p->get_parameter_value()));
            pt.put("pauli", pauli_pt(targets, p->get_pauli_id_list(), 1.0));
        } else {
            pt.put("target_qubit", u2s(targets[0]));
            pt.put("angle", ptree::format_double(p->get_parameter_value()));
        }
        return pt;
    }
    if (const auto* p = dynamic_cast<const QuantumGate_Probabilistic*>(&g)) {
        pt.put("name", "ProbabilisticGate");
        PTree dist, gates;
        for (double x : p->get_distribution()) dist.push_back(PTree(ptree::format_double(x)));
        for (const QuantumGateBase* x : p->get_gate_list()) gates.push_back(to_ptree(*x));
        pt.put("distribution", std::move(dist));
        pt.put("gate_list", std::move(gates));
        pt.put("is_instrument", p->is_instrument() ? "true" : "false");
        if (p->is_instrument()) pt.put("classical_register_address", u2s(p->get_classical_register_address()));
        return pt;
    }
    // src/cppsim/gate_general.hpp:676-690
    if (const auto* c = dynamic_cast<const QuantumGate_CP*>(&g)) {
        pt.put("name", "CPMapGate");
        PTree gates;
        for (const QuantumGateBase* k : c->get_gate_list()) gates.push_back(to_ptree(*k));
        pt.put("gate_list", std::move(gates));
        pt.put("state_normalize", c->get_state_normalize() ? "true" : "false");
        pt.put("probability_normalize", c->get_probability_normalize() ? "true" : "false");
        pt.put("assign_zero_if_not_matched", c->get_assign_zero_if_not_matched() ? "true" : "false");
        return pt;
    }
    if (const auto* c = dynamic_cast<const QuantumGate_CPTP*>(&g)) {
        pt.put("name", "CPTPMapGate");
        PTree gates;
        for (const QuantumGateBase* k : c->get_gate_list()) gates.push_back(to_ptree(*k));
        pt.put("gate_list", std::move(gates));
        pt.put("is_instrument", c->is_instrument() ? "true" : "false");
        if (c->is_instrument()) pt.put("classical_register_address", u2s(c->get_classical_register_address()));
        return pt;
    }
    // src/cppsim/gate_general.hpp:131-146: each coefficient as the text "(re,im)"
    if (const auto* c = dynamic_cast<const QuantumGate_LinearCombination*>(&g)) {
        pt.put("name", "LinearCombinationGate");
        PTree coefs, gates;
        for (CPPCTYPE z : c->get_coef_list())
            coefs.push_back(PTree("(" + ptree::format_double(z.real()) + "," + ptree::format_double(z.imag()) + ")"));
        for (const QuantumGateBase* k : c->get_gate_list()) gates.push_back(to_ptree(*k));
        pt.put("coefs", std::move(coefs));
        pt.put("gate_list", std::move(gates));
        return pt;
    }
    // src/cppsim/gate_reflect.hpp:98-103
    if (const auto* r = dynamic_cast<const ClsStateReflectionGate*>(&g)) {
        pt.put("name", "StateReflectionGate");
        pt.put("reflection_state", to_ptree(*r->get_reflection_state()));
        return pt;
    }
    // src/cppsim/gate_noisy_evolution.cpp:297-310
    if (const auto* e = dynamic_cast<const ClsNoisyEvolution*>(&g)) {
        pt.put("name", "NoisyEvolutionGate");
        pt.put("hamiltonian", to_ptree(static_cast<const GeneralQuantumOperator&>(e->get_hamiltonian())));
        PTree cops;
        for (const auto& c : e->get_c_ops()) cops.push_back(to_ptree(*c));
        pt.put("c_ops", std::move(cops));
        pt.put("time", ptree::format_double(e->get_time()));
        pt.put("dt", ptree::format_double(e->get_dt()));
        return pt;
    }
    // src/cppsim/gate_noisy_evolution.cpp:605-616
    if (const auto* e = dynamic_cast<const ClsNoisyEvolution_fast*>(&g)) {
        pt.put("name", "NoisyEvolutionFastGate");
        pt.put("hamiltonian", to_ptree(static_cast<const GeneralQuantumOperator&>(e->get_hamiltonian())));
        PTree cops;
        for (const auto& c : e->get_c_ops()) cops.push_back(to_ptree(*c));
        pt.put("c_ops", std::move(cops));
        pt.put("time", ptree::format_double(e->get_time()));
        return pt;
    }
    // src/cppsim/gate_named_npair.hpp:117-128
    if (dynamic_cast<const ClsNpairQubitGate*>(&g)) {
        pt.put("name", "FusedSWAPGate");
        pt.put("target_qubit_list", uint_array_pt(g.get_target_index_list()));
        return pt;
    }
    throw QulacsError("to_json: gate " + g.get_name() + " has no JSON form");
}

QuantumGate_SingleParameter* parametric_gate_from_ptree(const PTree& pt) {
    const std::string& name = pt.get("name");
    if (name == "ParametricRXGate") return gate::ParametricRX(pt.get_uint("target_qubit"), pt.get_double("angle"));
    if (name == "ParametricRYGate") return gate::ParametricRY(pt.get_uint("target_qubit"), pt.get_double("angle"));
    if (name == "ParametricRZGate") return gate::ParametricRZ(pt.get_uint("target_qubit"), pt.get_double("angle"));
    if (name == "ParametricPauliRotationGate") {
        const PauliOperator p = pauli_operator_from_ptree(pt.child("pauli"));
        return gate::ParametricPauliRotation(p.get_index_list(), p.get_pauli_id_list(), pt.get_double("angle"));
    }
    throw QulacsError("unknown value for property \"name\":" + name);
}

QuantumGateBase* gate_from_ptree(const PTree& pt) {
    const std::string& name = pt.get("name");
    if (name.compare(0, 10, "Parametric") == 0) return parametric_gate_from_ptree(pt);
    using F1 = QuantumGateMatrix* (*)(unsigned);
    static const std::pair<const char*, F1> named[] = {
        {"IGate", gate::Identity}, {"XGate", gate::X}, {"YGate", gate::Y}, {"ZGate", gate::Z},
        {"HGate", gate::H}, {"SGate", gate::S}, {"SdagGate", gate::Sdag}, {"TGate", gate::T},
        {"TdagGate", gate::Tdag}, {"sqrtXGate", gate::sqrtX}, {"sqrtXdagGate", gate::sqrtXdag},
        {"sqrtYGate", gate::sqrtY}, {"sqrtYdagGate", gate::sqrtYdag}, {"Projection-0Gate", gate::P0},
        {"Projection-1Gate", gate::P1}};
    for (const auto& nf : named)
        if (name == nf.first) return nf.second(pt.get_uint("target_qubit"));
    if (name == "X-rotationGate") return gate::RX(pt.get_uint("target_qubit"), pt.get_double("angle"));
    if (name == "Y-rotationGate") return gate::RY(pt.get_uint("target_qubit"), pt.get_double("angle"));
    if (name == "Z-rotationGate") return gate::RZ(pt.get_uint("target_qubit"), pt.get_double("angle"));
    if (name == "CNOTGate") return gate::CNOT(pt.get_uint("control_qubit"), pt.get_uint("target_qubit"));
    if (name == "CZGate") return gate::CZ(pt.get_uint("control_qubit"), pt.get_uint("target_qubit"));
    if (name == "SWAPGate") {
        const std::vector<unsigned> t = uint_array_from(pt.child("target_qubit_list"));
        if (t.size() != 2) throw QulacsError("JSON: SWAPGate needs two targets");
        return gate::SWAP(t[0], t[1]);
    }
    if (name == "DenseMatrixGate" || name == "SparseMatrixGate" || name == "DiagonalMatrixGate") {
        const std::vector<unsigned> t = uint_array_from(pt.child("target_qubit_list"));
        std::vector<unsigned> c, v;
        control_list_from(pt.child("control_qubit_list"), c, v);
        if (name == "DiagonalMatrixGate") {
            const std::vector<CPPCTYPE> d = complex_array_from(pt.child("vector").child("data"));
            return new QuantumGateDiagonalMatrix(t, d, c, v);
        }
        const ComplexMatrix m = name == "DenseMatrixGate" ? matrix_from(pt.child("matrix"))
                                                          : sparse_matrix_from(pt.child("matrix"));
        return new QuantumGateMatrix(t, m, c, v);
    }
    if (name == "PauliGate" || name == "PauliRotationGate") {
        const PauliOperator p = pauli_operator_from_ptree(pt.child("pauli"));
        if (name == "PauliGate") return gate::Pauli(p.get_index_list(), p.get_pauli_id_list());
        return gate::PauliRotation(p.get_index_list(), p.get_pauli_id_list(), pt.get_double("angle"));
    }
    if (name == "ProbabilisticGate" || name == "CPTPMapGate" || name == "CPMapGate" || name == "LinearCombinationGate") {
        std::vector<std::unique_ptr<QuantumGateBase>> owned;
        std::vector<QuantumGateBase*> gates;
        for (const auto& kv : pt.child("gate_list").children) {
            owned.emplace_back(gate_from_ptree(kv.second));
            gates.push_back(owned.back().get());
        }
        if (name == "CPMapGate")
            return gate::CP(gates, pt.get_bool("state_normalize"), pt.get_bool("probability_normalize"),
                            pt.get_bool("assign_zero_if_not_matched"));
        if (name == "LinearCombinationGate") {
            std::vector<CPPCTYPE> coefs;
            for (const auto& kv : pt.child("coefs").children) {
                double re = 0, im = 0;
                if (std::sscanf(kv.second.value.c_str(), "(%lf,%lf)", &re, &im) != 2)
                    throw QulacsError("JSON: bad LinearCombination coefficient " + kv.second.value);
                coefs.emplace_back(re, im);
            }
            return gate::LinearCombination(coefs, gates);
        }
        const bool instrument = pt.get_bool("is_instrument");
        if (name == "CPTPMapGate")
            return instrument ? gate::Instrument(gates, pt.get_uint("classical_register_address")) : gate::CPTP(gates);
        std::vector<double> dist;
        for (const auto& kv : pt.child("distribution").children) {
            PTree leaf;
            leaf.put("p", kv.second.value);
            dist.push_back(leaf.get_double("p"));
        }
        if (instrument) return gate::ProbabilisticInstrument(dist, gates, pt.get_uint("classical_register_address"));
        return gate::Probabilistic(dist, gates);
    }
    if (name == "StateReflectionGate") {
        std::unique_ptr<QuantumStateBase> r(state_from_ptree(pt.child("reflection_state")));
        auto* sv = dynamic_cast<QuantumStateMI355*>(r.get());
        if (!sv) throw QulacsError("JSON: a reflection state must be a state vector");
        return gate::StateReflection(sv);
    }
    if (name == "NoisyEvolutionGate" || name == "NoisyEvolutionFastGate") {
        std::unique_ptr<Observable> h(observable_from_ptree(pt.child("hamiltonian")));
        std::vector<std::unique_ptr<GeneralQuantumOperator>> owned;
        std::vector<const GeneralQuantumOperator*> cops;
        for (const auto& kv : pt.child("c_ops").children) {
            owned.emplace_back(operator_from_ptree(kv.second));
            cops.push_back(owned.back().get());
        }
        if (name == "NoisyEvolutionFastGate") return gate::NoisyEvolution_fast(*h, cops, pt.get_double("time"));
        return gate::NoisyEvolution(*h, cops, pt.get_double("time"), pt.get_double("dt"));
    }
    if (name == "FusedSWAPGate") {
        const std::vector<unsigned> t = uint_array_from(pt.child("target_qubit_list"));
        if (t.empty() || t.size() % 2) throw QulacsError("JSON: FusedSWAP needs two equal blocks");
        return gate::FusedSWAP(t[0], t[t.size() / 2], (unsigned)(t.size() / 2));
    }
    throw QulacsError("unknown value for property \"name\":" + name);
}

// ------------------------------------------------------------------ circuits
PTree to_ptree(const QuantumCircuit& c) {
    PTree pt;
    pt.put("name", dynamic_cast<const ParametricQuantumCircuit*>(&c) ? "ParametricQuantumCircuit" : "QuantumCircuit");
    pt.put("qubit_count", u2s(c.get_qubit_count()));
    PTree gates;
    for (size_t i = 0; i < c.get_gate_count(); ++i) gates.push_back(to_ptree(*c.get_gate(i)));
    pt.put("gate_list", std::move(gates));
    return pt;
}
QuantumCircuit* circuit_from_ptree(const PTree& pt) {
    if (pt.get("name") == "ParametricQuantumCircuit") return parametric_circuit_from_ptree(pt);
    expect_name(pt, "QuantumCircuit");
    std::unique_ptr<QuantumCircuit> c(new QuantumCircuit(pt.get_uint("qubit_count")));
    for (const auto& kv : pt.child("gate_list").children) c->add_gate_consume(gate_from_ptree(kv.second));
    return c.release();
}
ParametricQuantumCircuit* parametric_circuit_from_ptree(const PTree& pt) {
    expect_name(pt, "ParametricQuantumCircuit");
    std::unique_ptr<ParametricQuantumCircuit> c(new ParametricQuantumCircuit(pt.get_uint("qubit_count")));
    for (const auto& kv : pt.child("gate_list").children) {
        // gates named Parametric* are registered parameters (parametric_circuit.cpp:360-376)
        if (kv.second.get("name").compare(0, 10, "Parametric") == 0) {
            std::unique_ptr<QuantumGate_SingleParameter> g(parametric_gate_from_ptree(kv.second));
            c->add_parametric_gate(*g);
        } else {
            std::unique_ptr<QuantumGateBase> g(gate_from_ptree(kv.second));
            c->add_gate(*g);
        }
    }
    return c.release();
}

// ------------------------------------------------------------------ operators
PTree to_ptree(const PauliOperator& op) { return pauli_pt(op.get_index_list(), op.get_pauli_id_list(), op.get_coef()); }
PauliOperator pauli_operator_from_ptree(const PTree& pt) {
    expect_name(pt, "PauliOperator");
    PauliOperator p(complex_from(pt.child("coef")));
    for (const auto& kv : pt.child("pauli_list").children) {
        expect_name(kv.second, "SinglePauliOperator");
        p.add_single_Pauli(kv.second.get_uint("index"), kv.second.get_uint("pauli_id"));
    }
    return p;
}
PTree to_ptree(const GeneralQuantumOperator& op) {
    PTree pt;
    pt.put("name", "GeneralQuantumOperator");  // an Observable is written under the same name
    pt.put("qubit_count", u2s(op.get_qubit_count()));
    PTree list;
    for (size_t i = 0; i < op.get_term_count(); ++i) list.push_back(to_ptree(op.get_term(i)));
    pt.put("operator_list", std::move(list));
    return pt;
}
GeneralQuantumOperator* operator_from_ptree(const PTree& pt) {
    expect_name(pt, "GeneralQuantumOperator");
    std::unique_ptr<GeneralQuantumOperator> op(new GeneralQuantumOperator(pt.get_uint("qubit_count")));
    for (const auto& kv : pt.child("operator_list").children) op->add_operator(pauli_operator_from_ptree(kv.second));
    return op.release();
}
Observable* observable_from_ptree(const PTree& pt) {
    expect_name(pt, "GeneralQuantumOperator");
    std::unique_ptr<Observable> op(new Observable(pt.get_uint("qubit_count")));
    for (const auto& kv : pt.child("operator_list").children) {
        const PauliOperator p = pauli_operator_from_ptree(kv.second);
        if (p.get_coef().imag() != 0.0) throw QulacsError("Observable: coefficient must be real");
        op->add_operator(p);
    }
    return op.release();
}

}  // namespace qmh
