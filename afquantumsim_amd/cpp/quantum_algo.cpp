// Circuit generators (src/quantum_algo.cpp:103-129, 315-346).
#include "aqs/quantum_algo.h"

#include <stdexcept>

namespace aqs {

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
