/*
 * aqs_hip.h -- C ABI of the MI355X state-vector engine (libaqs_hip.so).
 *
 * This is the device boundary of the aqs:: API.  In afQuantumSim every
 * operation below is an ArrayFire call on an af::array holding the state;
 * here the state is a raw HBM buffer owned by an opaque aqs_state handle and
 * every operation is a hand-written HIP kernel for gfx950.
 *
 * Conventions
 *  - Qubit q of an n-qubit state is bit n-1-q of the amplitude index
 *    (qubit 0 is the most significant bit), as in the reference
 *    (include/quantum.h:102-106).
 *  - Every entry point returns an aqs_status; C++ exceptions never cross this
 *    boundary.  The aqs:: C++ layer rethrows AQS_ERR_* as the reference's
 *    exception types (std::invalid_argument, std::out_of_range,
 *    std::domain_error, std::runtime_error) with aqs_last_error() as message.
 *  - Work is enqueued on the state's HIP stream; functions that return host
 *    data synchronise that stream.
 *  - Matrices cross the boundary in FP64 (row-major, interleaved re/im); the
 *    complex64 kernels round them to FP32 once, at upload.
 */
#ifndef AQS_HIP_H
#define AQS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum aqs_status {
    AQS_OK = 0,
    AQS_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    AQS_ERR_OUT_OF_RANGE = 2,     /* std::out_of_range */
    AQS_ERR_DOMAIN = 3,           /* std::domain_error */
    AQS_ERR_RUNTIME = 4,          /* std::runtime_error */
    AQS_ERR_DEVICE = 5,           /* a HIP runtime error */
    AQS_ERR_NO_MEMORY = 6         /* device allocation failed */
} aqs_status;

typedef enum aqs_precision {
    AQS_C64 = 0,  /* complex64: the reference's c32 state (include/quantum.h:107-111) */
    AQS_C128 = 1  /* complex128: FP64 state, used to measure the complex64 error */
} aqs_precision;

/* Maximum qubits per state.  The reference caps at 30
 * (include/quantum.h:122); indices here are 64-bit, so the cap is memory. */
#define AQS_MAX_QUBITS 34

/* ---- gate operations ---------------------------------------------------- */

typedef enum aqs_op_kind {
    /* dense 2x2 matrix m on targets[0]: H, RotX, RotY, U3, Y, CY, CH, CRot*  */
    AQS_OP_MAT1 = 1,
    /* diagonal diag(m[0..1], m[6..7]) on targets[0]: RotZ, CRotZ           */
    AQS_OP_DIAG1 = 2,
    /* flip targets[0]: X, CX, CCNot, Or (with AQS_CTRL_ANY)               */
    AQS_OP_X = 3,
    /* exchange targets[0] and targets[1]: Swap, CSwap                      */
    AQS_OP_SWAP = 4,
    /* dense 2^k x 2^k matrix u, k = ntargets (1..5): fused blocks,
     * Gate / ControlGate.  targets[0] is the most significant bit of the
     * row/column index of u.                                               */
    AQS_OP_MATK = 5,
    /* multiply by the phase m[0..1] where every target and control bit is 1:
     * Z, Phase, CZ, CPhase                                                  */
    AQS_OP_PHASE = 6,
    /* only in fused schedules (aqs_schedule_get): a diagonal table pass that
     * applies many deferred phase terms in one sweep; targets lists its
     * anchor qubits.  Not accepted by aqs_apply.                           */
    AQS_OP_DIAGT = 7,
    /* only in fused schedules: an in-place permutation of index bits that
     * restores the qubit order after swaps were turned into relabellings;
     * targets lists (up to 5 of) the qubits it moves.                      */
    AQS_OP_PERMB = 8,
    /* only in tiled schedules (aqs_fuse_tiled): one sweep of the state in
     * LDS-resident tiles that applies a run of gates; ntargets = 0.        */
    AQS_OP_TILE = 9
} aqs_op_kind;

/* Control semantics: default = all control bits must be 1.  AQS_CTRL_ANY =
 * at least one control bit is 1 (the reference's Or gate,
 * src/quantum.cpp:1614-1653). */
#define AQS_CTRL_ANY 1

typedef struct aqs_op {
    int32_t kind;        /* aqs_op_kind */
    int32_t ntargets;    /* 1 (MAT1/DIAG1/X/PHASE), 2 (SWAP), 1..5 (MATK); PHASE may list up to 5 */
    int32_t targets[5];  /* qubit indices */
    int32_t ncontrols;   /* 0..3 */
    int32_t controls[3];
    int32_t flags;       /* AQS_CTRL_ANY */
    double m[8];         /* 2x2 complex, row-major, interleaved */
    const double* u;     /* MATK: 2^k x 2^k complex, row-major, interleaved */
    uint64_t more_controls; /* further controls as a qubit mask (bit q = qubit q), all of which must be 1,
                             * for gates with more than 3 controls (the reference's NControl_Gate and nested
                             * ControlGates, src/quantum_gates.cpp:80-114); 0 with AQS_CTRL_ANY */
} aqs_op;

typedef struct aqs_state_s* aqs_state;
typedef struct aqs_schedule_s* aqs_schedule;

/* ---- library ------------------------------------------------------------ */

/* Last error message of the calling thread ("" when none). */
const char* aqs_last_error(void);

/* Select the device (replaces aqs::initialize's af::setDevice,
 * src/quantum.cpp:69-86).  Idempotent. */
aqs_status aqs_init(int device);

/* Device name, compute-unit count and HBM bytes. */
aqs_status aqs_device_info(char* name, size_t name_len, int* cu_count, size_t* hbm_bytes);

/* ---- state lifetime (QSimulator ctors / copy, src/quantum.cpp:212-259) -- */

aqs_status aqs_state_create(int nqubits, int precision, aqs_state* out);
aqs_status aqs_state_destroy(aqs_state s);
/* dst <- src (same qubit count; precisions may differ: converts). */
aqs_status aqs_state_copy(aqs_state dst, aqs_state src);
aqs_status aqs_state_qubits(aqs_state s, int* nqubits);
aqs_status aqs_state_precision(aqs_state s, int* precision);
/* Raw device pointer to the 2^n amplitudes (interleaved re, im). */
aqs_status aqs_state_device_ptr(aqs_state s, void** ptr);
/* Use an external stream (hipStream_t) for this state's work; NULL restores
 * the library's stream. */
aqs_status aqs_state_set_stream(aqs_state s, void* hip_stream);
aqs_status aqs_state_sync(aqs_state s);

/* |index> (the |0..0> / |1..1> fast paths, src/quantum.cpp:219-224). */
aqs_status aqs_state_set_basis(aqs_state s, uint64_t index);
/* Kronecker product of per-qubit 2-vectors, qubit 0 first
 * (generate_statevector, src/quantum.cpp:261-275).  qubit_amps: n x 2
 * complex, interleaved; evaluated in FP64. */
aqs_status aqs_state_product(aqs_state s, const double* qubit_amps);
/* Seeded Gaussian state (re, im ~ N(0,1), counter-based), normalised in FP64. */
aqs_status aqs_state_gaussian(aqs_state s, uint64_t seed);
/* Host -> device in the state's precision; normalize != 0 divides by the
 * FP64 norm (the af::array ctor, src/quantum.cpp:242-259); a null vector is
 * AQS_ERR_INVALID_ARGUMENT. */
aqs_status aqs_state_upload(aqs_state s, const void* host, int normalize);
/* Device -> host in the state's precision (statevector().host()). */
aqs_status aqs_state_download(aqs_state s, void* host);

/* ---- gate application (QSimulator::simulate, src/quantum.cpp:277-291) --- */

/* Apply ops in order, one HIP kernel pass each; on a state of at most 12
 * qubits (complex64) or 11 (complex128), runs of two or more non-MATK ops
 * are applied by one LDS-resident workgroup, up to 64 (40) ops per launch. */
aqs_status aqs_apply(aqs_state s, const aqs_op* ops, size_t count);

/* Launch-shape autotuner for the one-qubit pass at this state's size: times
 * every variant for every target qubit with the identity matrix (the state
 * is unchanged) and keeps the fastest per (qubit count, target) for later
 * launches on this device.  variant_by_qubit (n ints, may be NULL) receives
 * the choice per target qubit. */
aqs_status aqs_autotune(aqs_state s, int reps, int* variant_by_qubit);

/* Pins the one-qubit launch shape (0..8) for this state's qubit count and
 * one target qubit on the current device, as aqs_autotune would store it.
 * Shapes: 0 (1 unit per thread, 32-bit index, 256 threads), 1 (2, 64-bit,
 * 256), 2 (4, 64-bit, 256), 3 (2, 32-bit, 256), 4 (1, 32-bit, 64),
 * 5 (2, 32-bit, 64), 6 (lane exchange: uncontrolled gates on target bits
 * 1..6 of 7..32-qubit states, else shape 4), 7 and 8 (shape 4 with
 * consecutive workgroups spread over 2 and 4 parts of the state); 32-bit
 * shapes fall back to 1 above 33 qubits. */
aqs_status aqs_tune_set(aqs_state s, int qubit, int variant);

/* ---- compile(): gate fusion (replaces QCircuit::compile, src/quantum.cpp:199-210) */

/* Fuse ops into blocks of at most max_qubits (1..5) qubits.  Each block is
 * a dense unitary built in FP64; runs of diagonal ops become phase passes. */
aqs_status aqs_fuse(const aqs_op* ops, size_t count, int nqubits, int max_qubits,
                    aqs_schedule* out);
/* Cache-blocked compile: runs of one-qubit, controlled (X, CX, CH, CRX..,
 * Toffoli, Or, CSwap) and diagonal ops become tile passes, each of which
 * sweeps the state once in LDS-resident tiles of 2^tile_qubits amplitudes
 * (tile_qubits 10..13, 0 = 13) and applies every gate of the run whose
 * non-diagonal qubits fall in the tile; gates only move past gates they
 * commute with.  Swaps become relabellings undone at the end; other ops
 * (dense Gate / ControlGate blocks) stay plain passes.  States of fewer than
 * 10 qubits get the aqs_fuse(.., 5, ..) schedule. */
aqs_status aqs_fuse_tiled(const aqs_op* ops, size_t count, int nqubits, int tile_qubits, aqs_schedule* out);
/* Modelled cost of running the schedule, in full-state sweeps: every pass
 * counts 1 (a plain pass of a sparse gate its touched fraction), plus the
 * in-tile gate work of tile passes and the MFMA work of dense blocks in
 * sweep-time units.  Used to choose between aqs_fuse and aqs_fuse_tiled. */
aqs_status aqs_schedule_cost(aqs_schedule sch, double* sweeps);
/* compile()'s default: the aqs_fuse(.., 5, ..) schedule, or from 20 qubits
 * up the aqs_fuse_tiled(.., 0, ..) one when its modelled cost is lower. */
aqs_status aqs_fuse_auto(const aqs_op* ops, size_t count, int nqubits, aqs_schedule* out);
aqs_status aqs_schedule_destroy(aqs_schedule sch);
/* Number of fused passes and the ops each came from. */
aqs_status aqs_schedule_info(aqs_schedule sch, size_t* passes, size_t* source_ops);
/* The i-th fused pass as an op (u stays owned by the schedule). */
aqs_status aqs_schedule_get(aqs_schedule sch, size_t i, aqs_op* op);
aqs_status aqs_apply_schedule(aqs_state s, aqs_schedule sch);

/* ---- measurement (src/quantum.cpp:293-531) ------------------------------ */

/* sum |a|^2 in FP64. */
aqs_status aqs_norm2(aqs_state s, double* out);
/* Divide by the FP64 norm. */
aqs_status aqs_normalize(aqs_state s);
/* P(qubit = 1) in FP64 (qubit_probability_true, src/quantum.cpp:372-391). */
aqs_status aqs_prob1(aqs_state s, int qubit, double* out);
/* P(q = 1) for every qubit q, one pass. out: n doubles. */
aqs_status aqs_prob1_all(aqs_state s, double* out);
/* <sum_q Z_q> in FP64. */
aqs_status aqs_expect_sum_z(aqs_state s, double* out);
/* |a_i|^2 as float for all i (probabilities(), src/quantum.cpp:404-414). out: host, 2^n floats. */
aqs_status aqs_probabilities(aqs_state s, float* out);
/* One amplitude (state(i), include/quantum.h:645-648). re_im: 2 doubles. */
aqs_status aqs_amplitude(aqs_state s, uint64_t index, double* re_im);
/* For each uniform u[k] in [0,1), the first index whose inclusive cumulative
 * probability exceeds u[k] (peek_measure_all / profile_measure_all rule,
 * src/quantum.cpp:344-359,467-501).  |a|^2 is formed in FP64 and summed
 * exactly in 64-bit fixed point (quantum 2^-60), so the result does not
 * depend on reduction order.  u at or above the total probability yields
 * the last index of non-zero probability. */
aqs_status aqs_sample(aqs_state s, const double* u, uint32_t shots, uint64_t* out_index);
/* Measure one qubit with uniform u (outcome = u < P(1)) and collapse:
 * the other half is zeroed and the kept half divided by sqrt(P)
 * (measure(q), src/quantum.cpp:308-342). */
aqs_status aqs_measure_qubit(aqs_state s, int qubit, double u, int* outcome, double* p1);
/* Collapse to |index> (measure_all, src/quantum.cpp:361-370). */
aqs_status aqs_collapse_to(aqs_state s, uint64_t index);
/* <a|b> in FP64 for two states of equal qubit count (any precisions). */
aqs_status aqs_inner(aqs_state a, aqs_state b, double* re_im);

/* ---- sampling without allocation (no reference counterpart) --------------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same rule
 * and bit-identical indices as aqs_sample, on caller-owned device buffers:
 * amps = 2^nqubits amplitudes (complex64 or complex128 as `precision` says),
 * u = `shots` uniforms (double), out_index = `shots` indices (uint64), all
 * device pointers; stream = a hipStream_t (NULL: the library's stream).  The
 * work is enqueued and the call returns without waiting.  It calls no
 * hipMalloc / hipFree: the sampling tables live in one fixed device array per
 * device (24 MiB + 256 B, held until the process ends), which limits nqubits
 * to AQS_NOALLOC_MAX_QUBITS (larger: AQS_ERR_INVALID_ARGUMENT).  Every call
 * on a device shares that array, so calls that overlap on one device (other
 * streams, other threads) race on it and must be serialised by the caller. */
#define AQS_NOALLOC_MAX_QUBITS 30
aqs_status aqs_sample_noalloc(const void* amps, int nqubits, int precision, const double* u, uint32_t shots,
                              uint64_t* out_index, void* stream);

/* ---- batched small states (no reference counterpart: the reference runs
 * one state per QSimulator) ----------------------------------------------- */

typedef struct aqs_batch_s* aqs_batch;

/* count states; state i has nqubits[i] (1..20) qubits and the ops
 * ops[op_offsets[i] .. op_offsets[i+1]) (MAT1, DIAG1, PHASE, X, SWAP).
 * Every run starts each state from |0..0>. */
aqs_status aqs_batch_create(int count, const int* nqubits, const aqs_op* ops, const size_t* op_offsets,
                            aqs_batch* out);
/* Times the (workgroup size, LDS tile) candidates and keeps the fastest. */
aqs_status aqs_batch_autotune(aqs_batch b, int* block, int* lds_qubits, double* best_ms);
/* Pins the workgroup size (256, 512 or 1024) and the LDS tile (12..14
 * qubits).  States of more qubits than the tile run as phases of LDS tiles. */
aqs_status aqs_batch_configure(aqs_batch b, int block, int lds_qubits);
/* One persistent launch (replayed from a hipGraph) over all states;
 * expect_sum_z (host, count doubles, may be NULL) receives <sum_q Z_q> of
 * every final state in FP64; ms the device time of the launch. */
aqs_status aqs_batch_run(aqs_batch b, double* expect_sum_z, double* ms);
/* Final complex64 amplitudes of state `index` (2^n values). */
aqs_status aqs_batch_state(aqs_batch b, int index, void* host);
aqs_status aqs_batch_info(aqs_batch b, size_t* passes, size_t* source_ops, int* block, int* lds_qubits, int* grid);
aqs_status aqs_batch_destroy(aqs_batch b);

/* ---- request server (no reference counterpart) ---------------------------
 * Continuous batching of circuits of mixed size: W worker threads, each with
 * its own HIP stream, run admitted requests concurrently; a request is
 * admitted when a worker is free and the resident device states plus its
 * own fit the HBM budget; idle states stay resident for reuse. */

typedef struct aqs_server_s* aqs_server;

typedef struct aqs_request_result {
    uint64_t id;
    int32_t status;      /* aqs_status of the request */
    int32_t nqubits;
    uint32_t shots;
    uint32_t reserved;
    uint64_t passes;     /* fused device passes it ran */
    double t_submit;     /* steady-clock seconds */
    double t_admit;
    double t_done;
    double norm2;        /* FP64 norm^2 of the final state */
} aqs_request_result;

typedef struct aqs_server_stats_t {
    uint64_t admitted, completed, rejected, evictions, pending;
    int32_t running, max_concurrent;
    uint64_t resident_bytes, peak_resident_bytes, budget_bytes;
    uint64_t cancelled;  /* requests that ended through aqs_server_cancel */
    uint64_t groups;     /* coalesced launches (aqs_server_set_grouping) */
    uint64_t grouped;    /* requests that ran in them */
} aqs_server_stats_t;

/* fuse_qubits: 1..5 = aqs_fuse blocks of that width, 0 = aqs_fuse_auto. */
aqs_status aqs_server_create(size_t hbm_budget_bytes, int workers, int fuse_qubits, aqs_server* out);
/* A request waiting longer than this blocks later requests from jumping
 * ahead of it (default 0.25 s). */
aqs_status aqs_server_set_starvation_limit(aqs_server s, double seconds);
/* Coalescing of small requests (default on: 16 qubits, 64 requests).  A
 * worker that takes a request of at most max_qubits qubits (0..16; 0 turns
 * coalescing off) and no MATK op also takes the later queued requests of that
 * kind, up to max_requests (1..256) and a 96 MiB device arena, and runs them
 * as one batched launch (every state in LDS tiles, no host fusion), then
 * samples each state as aqs_sample does.  The arena counts against the HBM
 * budget while resident; coalescing is off when it does not fit. */
aqs_status aqs_server_set_grouping(aqs_server s, int max_qubits, int max_requests);
/* Queue a circuit (ops and uniforms are copied).  Fails with
 * AQS_ERR_INVALID_ARGUMENT when the request can never fit the budget. */
aqs_status aqs_server_submit(aqs_server s, uint64_t id, int nqubits, const aqs_op* ops, size_t nops,
                             const double* uniforms, uint32_t shots, int keep_state);
/* Wait for request id (timeout_ms < 0: no limit); indices receives `shots`
 * sampled basis indices, state_host the final complex64 state when the
 * request was submitted with keep_state.  A result is handed over once: a
 * second wait on the same id (or one that loses the race to another waiter)
 * returns AQS_ERR_OUT_OF_RANGE. */
aqs_status aqs_server_wait(aqs_server s, uint64_t id, int timeout_ms, aqs_request_result* res, uint64_t* indices,
                           void* state_host);
/* Cancel request id: a queued request is dropped, a running one stops at its
 * next stage (after compile, after the gate passes); either then completes
 * with AQS_ERR_RUNTIME "request cancelled", to be collected by
 * aqs_server_wait.  *cancelled = 0 when it had already finished. */
aqs_status aqs_server_cancel(aqs_server s, uint64_t id, int* cancelled);
aqs_status aqs_server_stats(aqs_server s, aqs_server_stats_t* out);
/* Finishes queued work, then frees every resident state. */
aqs_status aqs_server_destroy(aqs_server s);

/* ---- timing ------------------------------------------------------------- */

/* Per-kind kernel timing with HIP events on the state's stream.  While
 * enabled every gate kernel launch is bracketed by an event pair. */
aqs_status aqs_profile_enable(aqs_state s, int enable);
/* Slots (arrays of AQS_PROFILE_SLOTS): 0 state preparation and measurement,
 * 1..6 the AQS_OP_* kind of a plain pass (MATK on the scalar kernel), 7 MATK
 * on MFMA, 8 diagonal table passes, 9 bit-permutation passes, 10 tile
 * passes, 11 runs of ops on a small state held in LDS (aqs_apply).  Total ms
 * and launch count since enable. */
#define AQS_PROFILE_SLOTS 12
aqs_status aqs_profile_read(aqs_state s, double* ms_by_kind, uint64_t* launches_by_kind);
/* Event pair on the state's stream around arbitrary work. */
aqs_status aqs_timer_start(aqs_state s);
aqs_status aqs_timer_stop(aqs_state s, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* AQS_HIP_H */
