/*
 * aqs/quantum_gates.h -- circuit builders of the reference's
 * include/quantum_gates.h (src/quantum_gates.cpp:16-338), over the aqs::
 * gate classes of this engine.
 */
#ifndef AQS_QUANTUM_GATES_H
#define AQS_QUANTUM_GATES_H

#include <vector>

#include "aqs/quantum.h"

namespace aqs {

/* The one-qubit `gate` on every listed qubit. */
QCircuit Group_Gate(uint32_t qubits, std::vector<uint32_t> target_qubits, QCircuit gate, bool compile = false);

/* The one-qubit `gate` on every listed qubit, controlled by control_qubit. */
QCircuit Control_Group_Gate(uint32_t qubits, uint32_t control_qubit, std::vector<uint32_t> target_qubits,
                            const QCircuit& gate, bool compile = false);

/* `gate` at target_qubit_begin controlled by a contiguous range of qubits
 * above it. */
QCircuit NControl_Gate(uint32_t qubits, uint32_t control_qubit_begin, uint32_t control_qubit_count,
                       uint32_t target_qubit_begin, const QCircuit& gate, bool compile = false);

/* `gate` at target_qubit_begin controlled by any set of qubits outside it. */
QCircuit NControl_Gate(uint32_t qubits, std::vector<uint32_t> control_qubits, uint32_t target_qubit_begin,
                       const QCircuit& gate, bool compile = false);

/* `gate` with its qubit i moved to new_qubit_positions[i] (swap network). */
QCircuit Rewire_Gate(uint32_t qubits, const std::vector<uint32_t>& new_qubit_positions, const QCircuit& gate,
                     bool compile = false);

/* The inverse circuit: gates reversed, rotation angles negated, sub-circuits
 * inverted.  Unlike the reference, whose copy shares (and so negates) the
 * input's gate objects, the input circuit is left unchanged. */
QCircuit Adjoint_Gate(const QCircuit& gate);

}  // namespace aqs

#endif
