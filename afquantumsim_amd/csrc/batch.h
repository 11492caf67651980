// Batched states: the per-state program the persistent batch kernel runs
// (batch.hip), shared with the host emulation of the CPU tests.
#ifndef AQS_BATCH_H
#define AQS_BATCH_H

#include <cstdint>
#include <vector>

#include "internal.h"

namespace aqs_hip {

// One pass of the batch kernel over an LDS-resident tile of a state.
// Positions t0, t1, cmask and pmask are tile-local (they equal the
// state-index positions when the whole state is one tile).  gmask holds the
// controls (for a PHASE: the required bits) that lie outside the tile, in
// state-index positions; they are resolved once per tile.
struct BatchOp {
    int32_t kind;    // AQS_OP_MAT1 / DIAG1 / PHASE / X / SWAP
    int32_t t0, t1;  // target positions (t1 for SWAP)
    int32_t any;     // AQS_CTRL_ANY: at least one control bit must be set
    uint32_t cmask;  // control bits inside the tile
    uint32_t pmask;  // PHASE: bits inside the tile that must be set
    uint32_t gmask;  // control / PHASE bits outside the tile
    uint32_t pad;
    float m[8];
};

// One sweep over a state: every tile of 2^m amplitudes that agree on the
// state-index bits `gbits` is loaded into LDS, runs ops [op_begin,
// op_begin + op_count) and is stored back.  The tile's local index lists the
// other bits in ascending order.  A state of at most m qubits is one phase
// with gbits = 0.
struct BatchPhase {
    uint32_t op_begin, op_count;
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
// tile and that commutes with each op it passes over.  Ops are appended to
// `out` (tile-local positions), phases to `phases` (op_begin relative to
// the start of `out`).
void batch_phases(const std::vector<BatchOp>& flat, int n, int m, std::vector<BatchOp>& out,
                  std::vector<BatchPhase>& phases);

}  // namespace aqs_hip

#endif
