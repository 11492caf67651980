/*
 * aqs_oracle.h -- FP64 CPU restatement of the afQuantumSim state-vector path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and nowhere else.  The product (libaqs_hip.so and the aqs:: C++ layer)
 * never links, loads or calls it.
 *
 * State layout: interleaved double (re, im), 2^n complex entries, qubit q at
 * bit position n-1-q (qubit 0 is the MOST significant bit), as the reference
 * fixes in include/quantum.h:102-106.
 *
 * Every gate is restated at index granularity from the CSR rows / Kronecker
 * products the reference builds in src/quantum.cpp (line citations are on
 * each function in aqs_oracle.c).
 */
#ifndef AQS_ORACLE_H
#define AQS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate kinds: the reference's GateTypes order (include/quantum.h:799-823),
 * followed by U3, which has no reference counterpart. */
enum orc_kind {
    ORC_BARRIER = 0,
    ORC_X,
    ORC_Y,
    ORC_Z,
    ORC_H,
    ORC_PHASE,
    ORC_SWAP,
    ORC_ROTX,
    ORC_ROTY,
    ORC_ROTZ,
    ORC_CX,
    ORC_CY,
    ORC_CZ,
    ORC_CH,
    ORC_CPHASE,
    ORC_CSWAP,
    ORC_CROTX,
    ORC_CROTY,
    ORC_CROTZ,
    ORC_CCX,
    ORC_OR,
    ORC_CIRCUIT,
    ORC_CONTROL_CIRCUIT,
    ORC_U3
};

/* One gate.  q0,q1,q2 follow the reference constructors' argument order:
 *   X/Y/Z/H/RotX/RotY/RotZ/Phase/U3 : q0 = target
 *   Swap                            : q0 = A, q1 = B
 *   CX/CY/CZ/CH/CPhase/CRot*        : q0 = control, q1 = target
 *   CSwap                           : q0 = control, q1 = A, q2 = B
 *   CCNot/Or                        : q0 = control A, q1 = control B, q2 = target
 *   Circuit                         : q0 = target_qubit_begin, k = sub-circuit qubits
 *   ControlCircuit                  : q0 = control, q1 = target_qubit_begin, k
 * Angles are the reference's float arguments; the oracle evaluates the
 * trigonometry in double.  u is a row-major 2^k x 2^k complex matrix
 * (interleaved re, im) for the circuit kinds. */
typedef struct orc_gate {
    int32_t kind;
    int32_t q0, q1, q2;
    float angle; /* theta for U3 */
    float phi, lam; /* U3 only */
    int32_t k;
    const double* u;
} orc_gate;

/* Gate application (index-level restatement). */
void orc_apply(double* sv, int n, const orc_gate* g);
void orc_run(double* sv, int n, const orc_gate* gates, int count);

/* The 2x2 matrix (row-major, interleaved) the oracle uses for a one-qubit
 * kind; returns 0 when the kind is not a dense one-qubit gate. */
int orc_matrix_1q(const orc_gate* g, double m[8]);

/* Reference method: the explicit 2^n x 2^n operator (row-major, interleaved),
 * built as the reference does -- a Kronecker product I (x) U (x) I for dense
 * gates (src/utils.cpp:169-195) and the CSR rows expanded for sparse ones. */
void orc_dense_operator(int n, const orc_gate* g, double* op);
/* out = op * in for an N x N operator. */
void orc_gemv(int n, const double* op, const double* in, double* out);

/* State preparation. */
void orc_basis_state(double* sv, int n, uint64_t index);
/* Kronecker product of per-qubit 2-vectors, qubit 0 first
 * (src/quantum.cpp:261-275).  qamps: n x 2 complex, interleaved. */
void orc_product_state(double* sv, int n, const double* qamps);
/* Seeded Gaussian state: re, im ~ N(0,1) from a counter-based generator,
 * normalised in FP64. */
void orc_gaussian_state(double* sv, int n, uint64_t seed);
double orc_gaussian_component(uint64_t seed, uint64_t index, int imag_part);

/* Measurement. */
void orc_probabilities(const double* sv, int n, double* p);
double orc_norm2(const double* sv, int n);
double orc_prob1(const double* sv, int n, int qubit);
double orc_expect_sum_z(const double* sv, int n);

/* Sampling rule of the reference (first index whose inclusive cumulative
 * probability exceeds u, src/quantum.cpp:353-356,486) evaluated on an exact
 * fixed-point cumulative sum: p_i = fma(re, re, im*im) in FP64, quantised
 * to rint(p_i * 2^60) in uint64.  amps_c64 are the float amplitudes the GPU
 * holds.  When u lies at or above the total, the last index with non-zero
 * quantised probability is returned. */
void orc_sample_c64(const float* amps_c64, int n, const double* u, int shots,
                    uint32_t* out);
void orc_sample_c128(const double* sv, int n, const double* u, int shots,
                     uint32_t* out);
/* Same rule with a plain sequential FP64 cumulative sum. */
void orc_sample_fp64_cdf(const double* sv, int n, const double* u, int shots,
                         uint32_t* out);

/* Multi-threaded plain state-vector pass (OpenMP) used to time the CPU
 * baseline; identical results to orc_run. */
void orc_run_threads(double* sv, int n, const orc_gate* gates, int count,
                     int threads);
int orc_max_threads(void);

/* Timing restatement of the reference's simulate() on an uncompiled circuit
 * (identity N x N product, then a dense N x N product per dense gate and an
 * O(N) pass per CSR gate); same result as orc_run up to rounding.  Allocates
 * one N x N operator; returns -1 when that fails. */
int orc_refmethod_run(double* sv, int n, const orc_gate* gates, int count, int threads);

/* ---- plain complex64 CPU state vector (cpu_sv.c): the timed CPU baseline --
 * sv: 2^n interleaved FP32 pairs, updated in place with `threads` OpenMP
 * threads.  csv_apply returns 0 for a gate kind it does not implement;
 * csv_run returns the number of gates applied. */
int csv_apply(float* sv, int n, const orc_gate* g, int threads);
int csv_run(float* sv, int n, const orc_gate* gates, int count, int threads);
/* cdf: caller's scratch of 2^n doubles. */
void csv_sample(const float* sv, int n, double* cdf, const double* u, int shots, uint32_t* out, int threads);

#ifdef __cplusplus
}
#endif

#endif
