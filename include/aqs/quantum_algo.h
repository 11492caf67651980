/*
 * aqs/quantum_algo.h -- circuit generators of the reference's
 * include/quantum_algo.h that feed the measured workloads and the
 * end-to-end regression circuits.  `compile` defaults to false, as in the
 * reference (include/quantum_algo.h:38-169).
 */
#ifndef AQS_QUANTUM_ALGO_H
#define AQS_QUANTUM_ALGO_H

#include <string>
#include <vector>

#include "aqs/quantum.h"

namespace aqs {

/* src/quantum_algo.cpp:15-38: X on every qubit i whose bit i of
 * marked_state is 0, an N-controlled Z over all qubits (controls 0..n-2,
 * target n-1), and the X layer again.  The marked basis index therefore has
 * qubit i (index bit n-1-i) equal to bit i of marked_state.
 * std::invalid_argument when marked_state >= 2^search_qubits. */
QCircuit grover_oracle(uint32_t search_qubits, uint32_t marked_state, bool compile = false);

/* src/quantum_algo.cpp:40-77: H on the search qubits, then `iterations` x
 * [Barrier, Gate(oracle, 0, oracle_name), Barrier, H layer, X layer,
 * N-controlled Z, (X, H) per qubit].  std::invalid_argument when the oracle
 * has fewer qubits than the search space. */
QCircuit grover_search(uint32_t search_qubits, const QCircuit& oracle, uint32_t iterations,
                       std::string oracle_name = "", bool compile = false);

/* src/quantum_algo.cpp:79-101: the iterations of grover_search without the
 * initial H layer and barriers. */
QCircuit grover_iteration(uint32_t search_qubits, const QCircuit& oracle, uint32_t iterations,
                          bool compile = false);

/* src/quantum_algo.cpp:103-114: for i = n-1..0, H(i) then CPhase(j, i, pi/2^(i-j)), j < i.
 * No final swaps (as in the reference). */
QCircuit fourier_transform(uint32_t qubits, bool compile = false);

/* src/quantum_algo.cpp:116-129. */
QCircuit inverse_fourier_transform(uint32_t qubits, bool compile = false);

/* src/quantum_algo.cpp:315-346: RY layer, RZ layer, then depth-1 x
 * [CX ladder, RY layer, RZ layer], parameters indexed as the reference does. */
QCircuit linear_entanglement_varstate(uint32_t qubits, uint32_t depth, const std::vector<float>& values,
                                      bool compile = false);

}  // namespace aqs

#endif
