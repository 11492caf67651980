/*
 * aqs/quantum_algo.h -- circuit generators of the reference's
 * include/quantum_algo.h that feed the measured workloads.
 */
#ifndef AQS_QUANTUM_ALGO_H
#define AQS_QUANTUM_ALGO_H

#include <vector>

#include "aqs/quantum.h"

namespace aqs {

/* src/quantum_algo.cpp:103-114: for i = n-1..0, H(i) then CPhase(j, i, pi/2^(i-j)), j < i.
 * No final swaps (as in the reference). */
QCircuit fourier_transform(uint32_t qubits, bool compile = true);

/* src/quantum_algo.cpp:116-129. */
QCircuit inverse_fourier_transform(uint32_t qubits, bool compile = true);

/* src/quantum_algo.cpp:315-346: RY layer, RZ layer, then depth-1 x
 * [CX ladder, RY layer, RZ layer], parameters indexed as the reference does. */
QCircuit linear_entanglement_varstate(uint32_t qubits, uint32_t depth, const std::vector<float>& values,
                                      bool compile = true);

}  // namespace aqs

#endif
