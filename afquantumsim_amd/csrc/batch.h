// Batched states: the per-state program the persistent batch kernel runs
// (batch.hip), shared with the host emulation of the CPU tests.
#ifndef AQS_BATCH_H
#define AQS_BATCH_H

#include <cstdint>
#include <vector>

#include "internal.h"

namespace aqs_hip {

// One gate of a batched state's program.  In a register segment t0 is a
// register slot (0..3) and creg / preg are masks over slots; in a plain
// segment t0, t1 are tile-local positions and creg = preg = 0.  cmask and
// pmask are tile-local positions outside the segment's register bits;
// gmask holds the controls (for a PHASE: the required bits) outside the tile
// in state-index positions, resolved once per tile.
// An uncontrolled one-qubit gate factored as a rotation after a phase: the
// target's 1-half times z = (m[4], m[5]), then the rotation as three shears
// p -= t q, q += s p, p -= t q with t = m[0], s = m[1].  The gate's diagonal
// on the left is carried by the host into the qubit's next gate.
constexpr int32_t kBatchOpMatP = 64;

struct BatchOp {
    int32_t kind;    // AQS_OP_MAT1 / DIAG1 / PHASE / X / SWAP, kBatchOpMatP
    int32_t t0, t1;  // targets (t1 for SWAP, plain segments only)
    int32_t any;     // AQS_CTRL_ANY: at least one control bit must be set
    uint32_t cmask;  // control bits inside the tile, outside the register bits
    uint32_t pmask;  // PHASE: required bits inside the tile, outside the register bits
    uint32_t gmask;  // control / PHASE bits outside the tile
    uint16_t creg;   // control register slots
    uint16_t preg;   // PHASE: required register slots
    float m[8];
};

// Register segments: every thread holds 2^kBatchRegBits amplitudes of a
// tile, the combinations of 4 tile-local bits (the slots), and runs the
// segment's gates on them in registers between one LDS load and one store.
constexpr int kBatchRegBits = 4;

// A run of gates whose changed bits lie within 4 tile-local positions
// (rmask, ascending = slots 0..3), or, with rmask = 0, gates applied one LDS
// pass each.
struct BatchSeg {
    uint32_t op_begin, op_count;
    uint32_t rmask;
    uint32_t pad;
};

// One sweep over a state: every tile of 2^m amplitudes that agree on the
// state-index bits `gbits` is loaded into LDS, runs segments [seg_begin,
// seg_begin + seg_count) and is stored back.  The tile's local index lists
// the other bits in ascending order.  A state of at most m qubits is one
// phase with gbits = 0.
struct BatchPhase {
    uint32_t seg_begin, seg_count;
    uint32_t gbits;
    uint32_t pad;
};

// A state's ops in state-index positions: runs of uncontrolled one-qubit
// gates on a qubit are multiplied into one 2x2 (FP64 products, rounded once)
// until a multi-qubit gate touches that qubit.  MATK ops are rejected.
std::vector<BatchOp> batch_absorb(const aqs_op* ops, size_t count, int n);

// Cuts a state's absorbed ops into phases of 2^m-amplitude tiles (n > m;
// otherwise one phase).  Each phase keeps bits 0..4 in the tile (runs of at
// least 32 amplitudes per load), adds the qubits the pending ops need in
// order of first use, and takes every op whose non-diagonal bits are in the
// tile and that commutes with each op it passes over.  A phase's ops are
// then cut, in order, into register segments (swaps as three X gates) and,
// for tiles of fewer than 4 bits or any-of controlled gates, plain segments.
// Ops, segments and phases are appended to the vectors (indices absolute).
void batch_phases(const std::vector<BatchOp>& flat, int n, int m, std::vector<BatchOp>& ops,
                  std::vector<BatchSeg>& segs, std::vector<BatchPhase>& phases);

// ---- request groups: small circuits of the request server run together ----

// Largest state of a group member, and the launch shape of a group.
constexpr int kGroupMaxQubits = 16;
constexpr int kGroupLdsQubits = 13;
constexpr int kGroupBlock = 512;

// One member of a group: its circuit from |0..0>, its uniforms (one per
// shot) and where its results go (host memory).
struct GroupItem {
    int n = 0;                     // 1..kGroupMaxQubits
    const aqs_op* ops = nullptr;   // MAT1, DIAG1, PHASE, X, SWAP
    size_t nops = 0;
    const double* u = nullptr;
    uint32_t shots = 0;
    uint64_t* indices = nullptr;   // shots entries
    double* norm2 = nullptr;       // FP64 norm^2 of the final state
    float* state = nullptr;        // complex64 interleaved (2^n pairs), or null
    size_t passes = 0;             // out: sweeps (phases) of its program
};

// Upper bound of the arena bytes one member takes (state, program, sampling
// tables, uniforms, indices and its share of the arena's fixed part).
size_t group_item_bytes(int n, size_t nops, uint32_t shots);

// Runs the members' circuits as one batched launch (k_batch, tiles of
// 2^kGroupLdsQubits amplitudes) on `stream` inside the device arena, then
// samples every final state with the exact fixed-point CDF of aqs_sample and
// reduces its norm; returns when the results are in host memory.  The
// members' group_item_bytes must sum to at most arena_bytes.
void run_group(GroupItem* items, int count, void* arena, size_t arena_bytes, hipStream_t stream);

}  // namespace aqs_hip

#endif
