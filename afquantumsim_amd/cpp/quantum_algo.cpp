// Circuit generators (src/quantum_algo.cpp:15-129, 315-346).
#include "aqs/quantum_algo.h"

#include <stdexcept>

#include "aqs/quantum_gates.h"

namespace aqs {

namespace {

// Z on qubit n-1 controlled by qubits 0..n-2, wrapped as one Gate
// (src/quantum_algo.cpp:28-30).
Gate all_controlled_z(uint32_t n) { return Gate{NControl_Gate(n, 0, n - 1, n - 1, Z::gate()), 0}; }

// One amplification step after the oracle (src/quantum_algo.cpp:59-73).
void diffusion(QCircuit& qc, uint32_t n) {
    for (uint32_t j = 0; j < n; ++j) qc << H{j};
    for (uint32_t j = 0; j < n; ++j) qc << X{j};
    qc << all_controlled_z(n);
    for (uint32_t j = 0; j < n; ++j) {
        qc << X{j};
        qc << H{j};
    }
}

}  // namespace

QCircuit grover_oracle(uint32_t search_qubits, uint32_t marked_state, bool compile) {
    if (search_qubits >= 32 || marked_state >= fast_pow2(search_qubits))
        throw std::invalid_argument{"Marked state should be in the range [0, 2^search_qubits)"};
    QCircuit qc{search_qubits};
    for (uint32_t i = 0; i < search_qubits; ++i)
        if (!(marked_state & (1u << i))) qc << X{i};
    qc << all_controlled_z(search_qubits);
    for (uint32_t i = 0; i < search_qubits; ++i)
        if (!(marked_state & (1u << i))) qc << X{i};
    if (compile) qc.compile();
    return qc;
}

QCircuit grover_search(uint32_t search_qubits, const QCircuit& oracle, uint32_t iterations, std::string oracle_name,
                       bool compile) {
    if (oracle.qubit_count() < search_qubits)
        throw std::invalid_argument{"Cannot use given oracle for this qubit circuit"};
    QCircuit qc{oracle.qubit_count()};
    for (uint32_t i = 0; i < search_qubits; ++i) qc << H{i};
    for (uint32_t it = 0; it < iterations; ++it) {
        qc << Barrier{false};
        qc << Gate{oracle, 0, oracle_name};
        qc << Barrier{false};
        diffusion(qc, search_qubits);
    }
    if (compile) qc.compile();
    return qc;
}

QCircuit grover_iteration(uint32_t search_qubits, const QCircuit& oracle, uint32_t iterations, bool compile) {
    QCircuit qc{search_qubits};
    for (uint32_t it = 0; it < iterations; ++it) {
        qc << Gate{oracle, 0};
        diffusion(qc, search_qubits);
    }
    if (compile) qc.compile();
    return qc;
}

QCircuit fourier_transform(uint32_t qubits, bool compile) {
    QCircuit qc(qubits);
    for (int32_t i = (int32_t)qubits - 1; i >= 0; --i) {
        qc << H{static_cast<uint32_t>(i)};
        for (uint32_t j = 0; j < (uint32_t)i; ++j) qc << CPhase{j, static_cast<uint32_t>(i), pi / (1 << (i - j))};
    }
    if (compile) qc.compile();
    return qc;
}

QCircuit inverse_fourier_transform(uint32_t qubits, bool compile) {
    QCircuit qc(qubits);
    for (uint32_t i = 0; i < qubits; ++i) {
        for (int32_t j = (int32_t)i - 1; j >= 0; --j) qc << CPhase{static_cast<uint32_t>(j), i, -pi / (1 << (i - j))};
        qc << H{i};
    }
    if (compile) qc.compile();
    return qc;
}

QCircuit linear_entanglement_varstate(uint32_t qubits, uint32_t depth, const std::vector<float>& values,
                                      bool compile) {
    if (qubits == 0) throw std::invalid_argument{"Number of qubits must be greater than zero"};
    if (depth == 0) throw std::invalid_argument{"Depth of circuit must be greater than zero"};
    if (qubits * depth * 2 != values.size()) throw std::invalid_argument{"Parameter count passed is not the expected received"};
    QCircuit qc{qubits};
    for (uint32_t i = 0; i < qubits; ++i) qc << RotY{i, values[i]};
    for (uint32_t i = 0; i < qubits; ++i) qc << RotZ{i, values[qubits + i]};
    for (uint32_t j = 0; j < depth - 1; ++j) {
        for (uint32_t i = 0; i < qubits - 1; ++i) qc << CX{i, i + 1};
        // the reference's parameter indices (consecutive layers share values)
        for (uint32_t i = 0; i < qubits; ++i) qc << RotY{i, values[qubits * (j + 2) + i]};
        for (uint32_t i = 0; i < qubits; ++i) qc << RotZ{i, values[qubits * (j + 3) + i]};
    }
    if (compile) qc.compile();
    return qc;
}

}  // namespace aqs
