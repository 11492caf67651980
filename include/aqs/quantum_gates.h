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

/* The one-qubit `gate` controlled by control_qubit, as the reference builds
 * it: the targets below the control get `gate` on as many consecutive qubits
 * from the first of them, the targets above likewise; non-consecutive
 * targets are therefore packed together (targets {0, 2}, control 3 act on
 * qubits 0 and 1). */
QCircuit Control_Group_Gate(uint32_t qubits, uint32_t control_qubit, std::vector<uint32_t> target_qubits,
                            const QCircuit& gate, bool compile = false);

/* `gate` at target_qubit_begin controlled by a contiguous range of qubits
 * above it. */
QCircuit NControl_Gate(uint32_t qubits, uint32_t control_qubit_begin, uint32_t control_qubit_count,
                       uint32_t target_qubit_begin, const QCircuit& gate, bool compile = false);

/* `gate` at target_qubit_begin controlled by a set of qubits outside it, built
 * as the reference builds it.  With controls only below the targets the
 * result has qubits - target_qubit_begin qubits (its qubit 0 stands for
 * target_qubit_begin); with exactly one control above the targets and none
 * below, the gate sits on the last gate.qubit_count() qubits. */
QCircuit NControl_Gate(uint32_t qubits, std::vector<uint32_t> control_qubits, uint32_t target_qubit_begin,
                       const QCircuit& gate, bool compile = false);

/* `gate` with its qubit i moved to new_qubit_positions[i] (swap network). */
QCircuit Rewire_Gate(uint32_t qubits, const std::vector<uint32_t>& new_qubit_positions, const QCircuit& gate,
                     bool compile = false);

/* The inverse circuit: gates reversed, rotation angles negated, sub-circuits
 * inverted, compiled.  As in the reference the gate objects are shared with
 * the input and inverted in place, so the input's angles are negated too;
 * CH throws std::runtime_error. */
QCircuit Adjoint_Gate(const QCircuit& gate);

}  // namespace aqs

#endif
