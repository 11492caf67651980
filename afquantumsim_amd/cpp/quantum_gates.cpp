// Circuit builders (src/quantum_gates.cpp:16-338) over the aqs:: API.
//
// They only compose Gate / ControlGate / Swap; the device work is what those
// gates emit (one MATK pass per sub-circuit of at most 5 qubits, flattened
// passes beyond).
#include "aqs/quantum_gates.h"

#include <algorithm>
#include <stdexcept>
#include <unordered_set>

namespace aqs {

QCircuit Group_Gate(uint32_t qubits, std::vector<uint32_t> target_qubits, QCircuit gate, bool compile) {
    if (gate.qubit_count() != 1) throw std::invalid_argument{"Gate not supported"};
    std::sort(target_qubits.begin(), target_qubits.end());
    if (target_qubits.size() > qubits)
        throw std::invalid_argument{"Cannot add more target qubits than there are total qubits"};
    gate.compile();
    QCircuit qc{qubits};
    for (const auto& qubit : target_qubits) qc << Gate{gate, qubit};
    if (compile) qc.compile();
    return qc;
}

// Targets below the control form one controlled sub-circuit, targets above
// it another.  Each sub-circuit holds one copy of `gate` per target in that
// group, on consecutive qubits, and is placed at the group's first target
// (src/quantum_gates.cpp:55-72): a group whose targets are not consecutive,
// e.g. {0, 2} with the control at 3, acts on qubits 0 and 1.
QCircuit Control_Group_Gate(uint32_t qubits, uint32_t control_qubit, std::vector<uint32_t> target_qubits,
                            const QCircuit& gate, bool compile) {
    if (control_qubit >= qubits) throw std::invalid_argument{"Invalid control qubit position"};
    if (gate.qubit_count() != 1) throw std::invalid_argument{"Gate not supported"};
    std::sort(target_qubits.begin(), target_qubits.end());
    if (target_qubits.size() >= qubits) throw std::invalid_argument{"Cannot add control gate at the given position"};
    auto pivot = std::lower_bound(target_qubits.cbegin(), target_qubits.cend(), control_qubit);
    if (pivot != target_qubits.cend() && *pivot == control_qubit)
        throw std::invalid_argument{"Cannot add control gate at the target qubit positions"};
    QCircuit qc(qubits);
    auto add_group = [&](std::vector<uint32_t>::const_iterator first, std::vector<uint32_t>::const_iterator last) {
        const auto width = (uint32_t)std::distance(first, last);
        if (width == 0) return;
        QCircuit group(width);
        for (uint32_t j = 0; j < width; ++j) group << Gate(gate, j);
        qc << ControlGate(group, control_qubit, *first);
    };
    add_group(target_qubits.cbegin(), pivot);
    add_group(pivot, target_qubits.cend());
    if (compile) qc.compile();
    return qc;
}

// The controls are consecutive, directly above the target range.  The
// innermost ControlGate takes the last control; each further one wraps the
// circuit built so far with a control on its new qubit 0, and the result is
// placed at control_qubit_begin (src/quantum_gates.cpp:96-111).
QCircuit NControl_Gate(uint32_t qubits, uint32_t control_qubit_begin, uint32_t control_qubit_count,
                       uint32_t target_qubit_begin, const QCircuit& gate, bool compile) {
    if (gate.qubit_count() >= qubits) throw std::invalid_argument{"Gate not supported"};
    if (control_qubit_count == 0) throw std::invalid_argument{"The number of control qubits must be at least 1"};
    if (target_qubit_begin + gate.qubit_count() > qubits)
        throw std::invalid_argument{"Invalid target qubit_begin position"};
    if (control_qubit_begin + control_qubit_count > target_qubit_begin)
        throw std::invalid_argument{"Invalid control_qubit position"};
    QCircuit qc(qubits);
    if (control_qubit_count == 1) {
        qc << ControlGate(gate, control_qubit_begin, target_qubit_begin);
    } else {
        // qubit 0 of `inner` is the last control; the gap up to the targets stays idle
        const uint32_t gap = target_qubit_begin - (control_qubit_begin + control_qubit_count - 1);
        QCircuit inner(gate.qubit_count() + gap);
        inner << ControlGate(gate, 0, gap);
        for (uint32_t added = 1; added < control_qubit_count; ++added) {
            QCircuit wider(inner.qubit_count() + 1);
            wider << ControlGate(inner, 0, 1);
            inner = std::move(wider);
        }
        qc << Gate(inner, control_qubit_begin);
    }
    if (compile) qc.compile();
    return qc;
}

// The reference's construction, step for step (src/quantum_gates.cpp:116-178),
// including the width and placement it gives:
//   - controls below the targets are added innermost first; the last of them
//     widens the circuit to qubits - target_qubit_begin qubits, counted from
//     target_qubit_begin;
//   - controls above the targets are then added from the nearest to the
//     farthest; the farthest places the circuit built so far at
//     qubits - (its width).
// So with controls only below the targets the result has
// qubits - target_qubit_begin qubits (its qubit 0 is target_qubit_begin), and
// with one control above and none below the gate lands on the last
// gate.qubit_count() qubits whatever target_qubit_begin says.
QCircuit NControl_Gate(uint32_t qubits, std::vector<uint32_t> control_qubits, uint32_t target_qubit_begin,
                       const QCircuit& gate, bool compile) {
    if (control_qubits.empty()) throw std::invalid_argument{"Number of control qubits must be at least one"};
    if (gate.qubit_count() + control_qubits.size() > qubits)
        throw std::invalid_argument{"Invalid number of qubits for number of control of qubits and gate count"};
    if (target_qubit_begin + gate.qubit_count() > qubits)
        throw std::invalid_argument{"Invalid target qubit begin position"};
    std::sort(control_qubits.begin(), control_qubits.end());
    if (control_qubits.back() >= qubits) throw std::invalid_argument{"Cannot add control gate at the given position"};
    auto pivot = std::lower_bound(control_qubits.begin(), control_qubits.end(), target_qubit_begin);
    if (pivot != control_qubits.end() && *pivot < target_qubit_begin + gate.qubit_count())
        throw std::invalid_argument{"Cannot add control gate at the target qubit positions"};
    const std::vector<uint32_t> above(control_qubits.begin(), pivot);
    const std::vector<uint32_t> below(pivot, control_qubits.end());

    QCircuit current = gate;
    for (size_t i = 0; i < below.size(); ++i) {
        // relative to target_qubit_begin; the last control widens to the end
        const uint32_t c = below[i] - target_qubit_begin;
        const uint32_t width = (i + 1 < below.size()) ? c + 1 : qubits - target_qubit_begin;
        QCircuit wider(width);
        wider << ControlGate(current, c, 0);
        current = std::move(wider);
    }
    if (!above.empty()) {
        uint32_t nearer = target_qubit_begin;
        for (size_t i = above.size() - 1; i > 0; --i) {
            const uint32_t c = above[i];
            QCircuit wider(qubits - c);
            wider << ControlGate{current, 0, nearer - c};
            nearer = c;
            current = std::move(wider);
        }
        QCircuit whole(qubits);
        whole << ControlGate(current, above.front(), qubits - current.qubit_count());
        current = std::move(whole);
    }
    if (compile) current.compile();
    return current;
}

QCircuit Rewire_Gate(uint32_t qubits, const std::vector<uint32_t>& new_qubit_positions, const QCircuit& gate,
                     bool compile) {
    if (new_qubit_positions.size() != gate.qubit_count())
        throw std::invalid_argument{"New qubit positions must map all the qubits in the gate"};
    if (gate.qubit_count() > qubits) throw std::domain_error{"Cannot rewire circuit to a lower number of qubits"};
    std::vector<uint32_t> sorted = new_qubit_positions;
    std::sort(sorted.begin(), sorted.end());
    for (size_t i = 1; i < sorted.size(); ++i)
        if (sorted[i - 1] == sorted[i]) throw std::invalid_argument{"Cannot rewire multiple qubits to the same qubit"};
    // follow each cycle of the mapping with swaps, run the gate, undo the swaps
    std::vector<Swap> swaps;
    std::unordered_set<uint32_t> done;
    for (uint32_t i = 0; i < new_qubit_positions.size(); ++i) {
        if (done.count(i)) continue;
        done.insert(i);
        uint32_t current = i;
        while (i != new_qubit_positions[current]) {
            swaps.emplace_back(current, new_qubit_positions[current]);
            current = new_qubit_positions[current];
            done.insert(current);
        }
    }
    QCircuit qc{qubits};
    for (const auto& s : swaps) qc << s;
    qc << Gate{gate, 0};
    for (auto it = swaps.rbegin(); it != swaps.rend(); ++it) qc << *it;
    if (compile) qc.compile();
    return qc;
}

namespace {

template <typename G>
void negate_angle(QGate& g) {
    auto& t = static_cast<G&>(g);
    t.angle = -t.angle;
}

// Turns one gate into its inverse in place (src/quantum_gates.cpp:244-313).
// Self-inverse gates stay; rotation and phase angles are negated; a
// sub-circuit gate gets the adjoint of its circuit.  CH and any gate outside
// this list throw, as in the reference.  U3 (no reference counterpart)
// becomes U3(-theta, -lambda, -phi).
void invert_in_place(QGate& g) {
    const uint32_t t = g.type();
    if (t == Barrier::static_type() || t == X::static_type() || t == Y::static_type() || t == Z::static_type() ||
        t == H::static_type() || t == Swap::static_type() || t == CSwap::static_type() || t == CX::static_type() ||
        t == CY::static_type() || t == CZ::static_type() || t == CCNot::static_type() || t == Or::static_type())
        return;
    if (t == RotX::static_type()) return negate_angle<RotX>(g);
    if (t == RotY::static_type()) return negate_angle<RotY>(g);
    if (t == RotZ::static_type()) return negate_angle<RotZ>(g);
    if (t == Phase::static_type()) return negate_angle<Phase>(g);
    if (t == CRotX::static_type()) return negate_angle<CRotX>(g);
    if (t == CRotY::static_type()) return negate_angle<CRotY>(g);
    if (t == CRotZ::static_type()) return negate_angle<CRotZ>(g);
    if (t == CPhase::static_type()) return negate_angle<CPhase>(g);
    if (t == U3::static_type()) {
        auto& u = static_cast<U3&>(g);
        const float phi = u.phi;
        u.theta = -u.theta;
        u.phi = -u.lambda;
        u.lambda = -phi;
        return;
    }
    if (t == Gate::static_type()) {
        auto& sub = static_cast<Gate&>(g);
        sub.internal_circuit = std::make_shared<QCircuit>(Adjoint_Gate(*sub.internal_circuit));
        return;
    }
    if (t == ControlGate::static_type()) {
        auto& sub = static_cast<ControlGate&>(g);
        sub.internal_circuit = std::make_shared<QCircuit>(Adjoint_Gate(*sub.internal_circuit));
        return;
    }
    throw std::runtime_error{"Unknown unsupported gate cannot be adjoint"};
}

}  // namespace

// A copy of the circuit shares its gate objects (QCircuit's copy keeps the
// shared pointers), and the reference inverts those objects in place
// (src/quantum_gates.cpp:233-313): after the call the input circuit's
// rotation angles are negated too, and a CH part way through leaves the gates
// after it (in program order) already inverted when it throws.  The returned
// circuit is compiled: it runs the conjugate transpose of the input's
// operator (the reference transposes its compiled matrix, :335; compiling
// the inverted gates in reverse order gives the same operator).
QCircuit Adjoint_Gate(const QCircuit& gate) {
    QCircuit qc = gate;
    auto& list = qc.gate_list();
    std::reverse(list.begin(), list.end());
    for (auto& g : list) invert_in_place(*g);
    // representation entries in reverse order (src/quantum_gates.cpp:322-331)
    const std::string& rep = qc.representation();
    std::string reversed;
    size_t prev = 0, mark = rep.find(';', prev);
    while (mark != std::string::npos) {
        reversed = rep.substr(prev, mark - prev + 1) + reversed;
        prev = mark + 1;
        mark = rep.find(';', prev);
    }
    qc.representation() = reversed;
    qc.clear_cache();
    qc.compile();
    return qc;
}

}  // namespace aqs
