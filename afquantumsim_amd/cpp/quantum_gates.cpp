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
    // targets above the control form one controlled sub-circuit, targets below another
    const auto top = (uint32_t)std::distance(target_qubits.cbegin(), pivot);
    const auto bottom = (uint32_t)std::distance(pivot, target_qubits.cend());
    if (top) {
        QCircuit temp(target_qubits[top - 1] - target_qubits.front() + 1);
        for (auto it = target_qubits.cbegin(); it != pivot; ++it) temp << Gate(gate, *it - target_qubits.front());
        qc << ControlGate(temp, control_qubit, target_qubits.front());
    }
    if (bottom) {
        QCircuit temp(target_qubits.back() - *pivot + 1);
        for (auto it = pivot; it != target_qubits.cend(); ++it) temp << Gate(gate, *it - *pivot);
        qc << ControlGate(temp, control_qubit, *pivot);
    }
    if (compile) qc.compile();
    return qc;
}

QCircuit NControl_Gate(uint32_t qubits, uint32_t control_qubit_begin, uint32_t control_qubit_count,
                       uint32_t target_qubit_begin, const QCircuit& gate, bool compile) {
    if (gate.qubit_count() >= qubits) throw std::invalid_argument{"Gate not supported"};
    if (control_qubit_count == 0) throw std::invalid_argument{"The number of control qubits must be at least 1"};
    if (target_qubit_begin + gate.qubit_count() > qubits)
        throw std::invalid_argument{"Invalid target qubit_begin position"};
    if (control_qubit_begin + control_qubit_count > target_qubit_begin)
        throw std::invalid_argument{"Invalid control_qubit position"};
    std::vector<uint32_t> controls;
    for (uint32_t i = 0; i < control_qubit_count; ++i) controls.push_back(control_qubit_begin + i);
    return NControl_Gate(qubits, controls, target_qubit_begin, gate, compile);
}

// Controls are added one at a time, each wrapping the circuit built so far in
// a ControlGate, first those below the target range, then those above it.
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
    std::vector<uint32_t> above(control_qubits.begin(), pivot);
    std::vector<uint32_t> below(pivot, control_qubits.end());
    // `current` spans qubits [lo, lo + current.qubit_count()) of the result
    QCircuit current = gate;
    uint32_t lo = target_qubit_begin;
    for (uint32_t c : below) {
        QCircuit wider(c - lo + 1);
        wider << ControlGate(current, c - lo, 0);
        current = std::move(wider);
    }
    for (auto it = above.rbegin(); it != above.rend(); ++it) {
        const uint32_t c = *it;
        QCircuit wider(lo + current.qubit_count() - c);
        wider << ControlGate(current, 0, lo - c);
        current = std::move(wider);
        lo = c;
    }
    QCircuit qc(qubits);
    qc << Gate(current, lo);
    if (compile) qc.compile();
    return qc;
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
std::shared_ptr<QGate> copy_as(const QGate& g) {
    return std::make_shared<G>(static_cast<const G&>(g));
}

// A copy of the inverse of one gate.
std::shared_ptr<QGate> adjoint_of(const QGate& g) {
    const uint32_t t = g.type();
    if (t == Barrier::static_type()) return copy_as<Barrier>(g);
    if (t == X::static_type()) return copy_as<X>(g);
    if (t == Y::static_type()) return copy_as<Y>(g);
    if (t == Z::static_type()) return copy_as<Z>(g);
    if (t == H::static_type()) return copy_as<H>(g);
    if (t == Swap::static_type()) return copy_as<Swap>(g);
    if (t == CSwap::static_type()) return copy_as<CSwap>(g);
    if (t == CX::static_type()) return copy_as<CX>(g);
    if (t == CY::static_type()) return copy_as<CY>(g);
    if (t == CZ::static_type()) return copy_as<CZ>(g);
    if (t == CH::static_type()) return copy_as<CH>(g);
    if (t == CCNot::static_type()) return copy_as<CCNot>(g);
    if (t == Or::static_type()) return copy_as<Or>(g);
    auto negate = [](auto p) {
        p->angle = -p->angle;
        return p;
    };
    if (t == RotX::static_type()) return negate(std::make_shared<RotX>(static_cast<const RotX&>(g)));
    if (t == RotY::static_type()) return negate(std::make_shared<RotY>(static_cast<const RotY&>(g)));
    if (t == RotZ::static_type()) return negate(std::make_shared<RotZ>(static_cast<const RotZ&>(g)));
    if (t == Phase::static_type()) return negate(std::make_shared<Phase>(static_cast<const Phase&>(g)));
    if (t == CRotX::static_type()) return negate(std::make_shared<CRotX>(static_cast<const CRotX&>(g)));
    if (t == CRotY::static_type()) return negate(std::make_shared<CRotY>(static_cast<const CRotY&>(g)));
    if (t == CRotZ::static_type()) return negate(std::make_shared<CRotZ>(static_cast<const CRotZ&>(g)));
    if (t == CPhase::static_type()) return negate(std::make_shared<CPhase>(static_cast<const CPhase&>(g)));
    if (t == U3::static_type()) {
        auto p = std::make_shared<U3>(static_cast<const U3&>(g));
        const float phi = p->phi;
        p->theta = -p->theta;
        p->phi = -p->lambda;
        p->lambda = -phi;
        return p;
    }
    if (t == Gate::static_type()) {
        auto p = std::make_shared<Gate>(static_cast<const Gate&>(g));
        p->internal_circuit = std::make_shared<QCircuit>(Adjoint_Gate(*p->internal_circuit));
        return p;
    }
    if (t == ControlGate::static_type()) {
        auto p = std::make_shared<ControlGate>(static_cast<const ControlGate&>(g));
        p->internal_circuit = std::make_shared<QCircuit>(Adjoint_Gate(*p->internal_circuit));
        return p;
    }
    throw std::runtime_error{"Unknown unsupported gate cannot be adjoint"};
}

}  // namespace

QCircuit Adjoint_Gate(const QCircuit& gate) {
    QCircuit qc{gate.qubit_count()};
    for (auto it = gate.gate_list().rbegin(); it != gate.gate_list().rend(); ++it)
        qc.gate_list().push_back(adjoint_of(**it));
    // representation entries in reverse order (src/quantum_gates.cpp:322-331)
    const std::string& rep = gate.representation();
    std::string reversed;
    size_t prev = 0, mark = rep.find(';', prev);
    while (mark != std::string::npos) {
        reversed = rep.substr(prev, mark - prev + 1) + reversed;
        prev = mark + 1;
        mark = rep.find(';', prev);
    }
    qc.representation() = reversed;
    qc.compile();
    return qc;
}

}  // namespace aqs
