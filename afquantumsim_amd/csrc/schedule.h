// Compiled schedules: the passes aqs_fuse / aqs_fuse_tiled produce and
// aqs_apply_schedule runs (shared by fusion.cpp, tiling.cpp, tile.hip).
#ifndef AQS_SCHEDULE_H
#define AQS_SCHEDULE_H

#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "internal.h"

namespace aqs_hip {

using cd = std::complex<double>;

struct DiagPass {
    std::vector<int> required;            // amplitude bit positions
    std::vector<int> anchor;              // amplitude bit positions (<= 5)
    std::vector<std::vector<int>> chunks; // amplitude bit positions (<= 6 x 5)
    std::vector<cd> table;                // [2^|anchor|][chunks][32]
};

// Tile pass: the state is swept in tiles of 2^m amplitudes that share every
// index bit outside `local` (m ascending bit positions, bits 0..4 always
// among them).  A tile is held in LDS; each segment gives every thread 16 of
// its amplitudes (4 register slots over local positions `reg`, the thread id
// spread over the other local positions `thr`, ascending) and runs a list of
// ops on them in registers.
constexpr int kTileRegBits = 4;
constexpr int kTileMaxThreadBits = 9;
constexpr int kTileMaxLocal = kTileRegBits + kTileMaxThreadBits;

enum TileOpKind : int32_t {
    TILE_MAT = 0,     // 2x2 on register slot `slot`
    TILE_XCH = 1,     // exchange across register slot `slot`
    TILE_PHASES = 2,  // a run of diagonal gates as phase terms [term_begin, +term_count)
    TILE_MATR = 3,    // real 2x2 on register slot `slot` (half the multiplies of TILE_MAT)
    TILE_MATP = 4,    // uncontrolled: the slot's 1-half times z, then a rotation (5/8 of TILE_MAT)
};

// Kernel dispatch code of an op (one switch per op, no divergent branch
// around the register array):
//   form 0  no control
//   form 1  controls on register slots only (one uniform test per pair)
//   form 2  controls on thread-held positions or outside the tile, or "any"
//           controls (per-thread predicate, results selected)
// MAT: 0 + 4*form + slot, XCH: 12 + 4*form + slot, PHASES: 24 direct, 25 zeta,
// MATR: 26 + 4*form + slot, MATP: 38 + slot (form 0 only), MATP set:
// 42 + slot mask (two or more slots) on the first of popcount(mask)
// consecutive MATP ops whose slots are the mask's bits in ascending order --
// uncontrolled gates on distinct qubits, dispatched once.
constexpr int kTileCodePhases = 24;
constexpr int kTileCodeMatR = 26;
constexpr int kTileCodeMatP = 38;
constexpr int kTileCodeMatPSet = 42;
constexpr int kTileCodeCount = 58;
inline int tile_code(int kind, int form, int slot, bool zeta) {
    if (kind == TILE_PHASES) return kTileCodePhases + (zeta ? 1 : 0);
    if (kind == TILE_MATP) return form == 0 ? kTileCodeMatP + slot : -1;
    if (kind == TILE_MATR) return kTileCodeMatR + 4 * form + slot;
    return (kind == TILE_XCH ? 12 : 0) + 4 * form + slot;
}
constexpr bool tile_code_is_phases(int code) { return code == kTileCodePhases || code == kTileCodePhases + 1; }
// form and slot of a MAT / XCH / MATR / MATP code
constexpr int tile_code_form(int code) {
    return code >= kTileCodeMatP ? 0 : code >= kTileCodeMatR ? (code - kTileCodeMatR) / 4 : (code % 12) / 4;
}
constexpr int tile_code_slot(int code) {
    return code >= kTileCodeMatPSet ? __builtin_ctz((unsigned)(code - kTileCodeMatPSet))
           : code >= kTileCodeMatP  ? code - kTileCodeMatP
                                    : (code >= kTileCodeMatR ? code - kTileCodeMatR : code) % 4;
}
constexpr bool tile_code_is_matp(int code) { return code >= kTileCodeMatP; }

// 64 bytes, read with scalar loads.  Complex64 kernels read the FP32 matrix
// here; complex128 kernels read TilePass::m128[8 * op].
struct TileOpDev {
    int32_t code;     // tile_code()
    uint32_t any;     // controls: at least one set instead of all set
    uint32_t c_reg;   // control register slots (mask over the 4 slots)
    uint32_t c_thr;   // control local positions held by the thread (mask over local positions)
    uint64_t c_glb;   // control index bits outside the tile
    uint32_t term_begin, term_count;
    float mf[8];      // MAT: row-major 2x2 in FP32 (complex); MATR: mf[0..3] = real row-major 2x2;
                      // MATP: mf[0] = t, mf[1] = s of the rotation's shears, mf[4..5] = z
};
static_assert(sizeof(TileOpDev) == 64, "compact tile op");

// Phase term: multiply by z where every bit of the three masks is set
// (register slots, thread-held local positions, index bits outside the
// tile).  Terms of one PHASES op are sorted by register mask; the first term
// of each equal-mask group carries the group length, the number of folded
// terms (the group's last ones), and real = 1 when every z of the group is
// real (no folded term): the group's product then scales amplitudes by a
// real factor.  z in FP32 here, in FP64 in TilePass::z128[2 * term].  A
// folded term (fold >= 0) takes its z from the per-tile value of fold
// `fold` instead.
struct TileTermDev {
    uint16_t m_reg;
    uint16_t group;
    uint32_t m_thr;
    uint64_t m_glb;
    float zf[2];
    int32_t fold;
    uint16_t real;
    uint16_t folded;
};
static_assert(sizeof(TileTermDev) == 32, "compact tile term");
static_assert(offsetof(TileTermDev, m_thr) == 4 && offsetof(TileTermDev, m_glb) == 8 &&
                  offsetof(TileTermDev, zf) == 16 && offsetof(TileTermDev, fold) == 24,
              "the kernel reads a term as two 16-byte words");

// Fold: the product of the phase terms that pair one local bit with bits
// outside the tile, evaluated once per tile (it only depends on the tile's
// outer bits): value = product over [begin, begin + count) of fold terms of
// (z where the tile has every bit of m_glb).  The run's folded term then
// applies the value to the amplitudes with that local bit set.
struct TileFoldDev {
    uint32_t begin, count;
};
struct TileFoldTermDev {
    uint64_t m_glb;
    float zf[2];
};
static_assert(sizeof(TileFoldTermDev) == 16 && offsetof(TileFoldTermDev, zf) == 8,
              "the kernel reads a fold term as one 16-byte word");
constexpr int kTileMaxFolds = 256;

struct TileSegDev {
    int32_t reg[kTileRegBits];
    int32_t thr[kTileMaxThreadBits + 1];
    uint32_t op_begin, op_count;
};

struct TilePass {
    int m = 0;                     // local bits = 4 + log2(threads)
    std::vector<int> local;        // ascending index-bit positions
    std::vector<TileSegDev> segs;
    std::vector<TileOpDev> ops;
    std::vector<double> m128;  // [ops][8]
    std::vector<TileTermDev> terms;
    std::vector<double> z128;  // [terms][2]
    std::vector<TileFoldDev> folds;
    std::vector<TileFoldTermDev> fold_terms;
    std::vector<double> fold_z128;  // [fold_terms][2]
};

// Device copies of one tile pass's program.
struct TileDevProgram {
    const void* segs;
    const void* ops;
    const void* terms;
    const void* m128;  // complex128 states only
    const void* z128;
    const void* folds;
    const void* fold_terms;
    const void* fold_z128;
};

// Runs one tile pass.
void apply_tile_pass(aqs_state s, const TilePass& pass, const TileDevProgram& prog);

}  // namespace aqs_hip

struct aqs_schedule_s {
    int n = 0;
    size_t source_ops = 0;
    std::vector<aqs_op> passes;
    std::vector<std::vector<double>> mats;  // owned FP64 matrices (MATK passes)
    std::vector<int> diag_index;            // -1 or index into diags
    std::vector<aqs_hip::DiagPass> diags;
    std::vector<std::pair<int, int>> perm_index;  // (pass, index into perms)
    std::vector<std::vector<int>> perms;          // bit involutions (AQS_OP_PERMB passes)
    std::vector<int> tile_index;                  // -1 or index into tiles
    std::vector<aqs_hip::TilePass> tiles;
    // device copies, built at first use for a given precision
    void* dev = nullptr;
    int dev_precision = -1;
    std::vector<size_t> dev_off;
    std::vector<std::array<size_t, 8>> dev_tile;  // tile passes: TileDevProgram fields
    std::vector<char> dev_mfma;

    // appends a pass with no side data
    void push_plain(const aqs_op& op, std::vector<double> mat = {}) {
        passes.push_back(op);
        mats.push_back(std::move(mat));
        diag_index.push_back(-1);
        tile_index.push_back(-1);
    }
};

namespace aqs_hip {

// Logical qubit q sits at physical qubit L[q]: appends the (at most two)
// bit-involution passes that move every amplitude to its logical index.
void emit_restore_order(aqs_schedule_s* sch, const std::vector<int>& L);

// Modelled run time of a schedule in full-state sweeps (tiling.cpp).
double schedule_cost(const aqs_schedule_s* sch);

}  // namespace aqs_hip

#endif
