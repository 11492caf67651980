// Test-only host emulation of compiled schedules, used by
// tests/test_tiling.py to check the tile scheduler on the CPU.
//
// tile_emulate() builds a schedule with aqs_fuse_tiled (host code only) and
// runs it on an FP64 host state, reading the tile passes exactly as the
// device kernel (afquantumsim_amd/csrc/tile.hip) does: per tile, per
// segment, per thread the 16 register amplitudes, with the op encodings
// (register slots, thread-fixed positions, global bits) interpreted as the
// kernel interprets them.  Plain passes and bit permutations are applied
// directly.  batch_emulate() runs the program of one batched state
// (afquantumsim_amd/csrc/batch.hip) phase by phase and tile by tile.
// block_emulate() does the same for the dense-block schedule of
// aqs_fuse, diagonal tables included (afquantumsim_amd/csrc/diag.hip:
// factor = product over chunks of table[anchor value][chunk][chunk value]
// on amplitudes whose required bits are all set).  The results are compared
// with the oracle by the Python tests.
#include <complex>
#include <cstring>
#include <vector>

#include "batch.h"
#include "schedule.h"

using namespace aqs_hip;
using cplx = std::complex<double>;

namespace {

// A batched rotation op (kBatchOpMatP, no controls) on the pair (x, y):
// y *= z, then the shears x -= t y, y += su x, x -= t y (m[0] = t,
// m[1] = su, m[4..5] = z).
void batch_rotation(cplx& x, cplx& y, const float* m) {
    y *= cplx(m[4], m[5]);
    x -= (double)m[0] * y;
    y += (double)m[1] * x;
    x -= (double)m[0] * y;
}

void plain_apply(std::vector<cplx>& v, int n, const aqs_op& op) {
    const uint64_t N = 1ull << n;
    auto bit = [&](int q) { return 1ull << (n - 1 - q); };
    uint64_t cmask = 0;
    for (int j = 0; j < op.ncontrols; ++j) cmask |= bit(op.controls[j]);
    for (int q = 0; q < n; ++q)
        if ((op.more_controls >> q) & 1ull) cmask |= bit(q);
    const bool any = (op.flags & AQS_CTRL_ANY) != 0;
    auto ok = [&](uint64_t i) { return !cmask || (any ? (i & cmask) != 0 : (i & cmask) == cmask); };
    const cplx m00(op.m[0], op.m[1]), m01(op.m[2], op.m[3]), m10(op.m[4], op.m[5]), m11(op.m[6], op.m[7]);
    switch (op.kind) {
        case AQS_OP_MAT1:
        case AQS_OP_DIAG1:
        case AQS_OP_X: {
            const uint64_t t = bit(op.targets[0]);
            for (uint64_t i = 0; i < N; ++i) {
                if ((i & t) || !ok(i)) continue;
                const cplx a = v[i], b = v[i | t];
                if (op.kind == AQS_OP_MAT1) {
                    v[i] = m00 * a + m01 * b;
                    v[i | t] = m10 * a + m11 * b;
                } else if (op.kind == AQS_OP_DIAG1) {
                    v[i] = m00 * a;
                    v[i | t] = m11 * b;
                } else {
                    v[i] = b;
                    v[i | t] = a;
                }
            }
            break;
        }
        case AQS_OP_SWAP: {
            const uint64_t a = bit(op.targets[0]), b = bit(op.targets[1]);
            for (uint64_t i = 0; i < N; ++i)
                if ((i & a) && !(i & b) && ok(i)) std::swap(v[i], v[i ^ a ^ b]);
            break;
        }
        case AQS_OP_PHASE: {
            uint64_t mask = cmask;
            for (int j = 0; j < op.ntargets; ++j) mask |= bit(op.targets[j]);
            for (uint64_t i = 0; i < N; ++i)
                if ((i & mask) == mask) v[i] *= m00;
            break;
        }
        case AQS_OP_MATK: {
            const int k = op.ntargets;
            const uint64_t K = 1ull << k;
            std::vector<uint64_t> off(K);
            uint64_t tmask = 0;
            for (uint64_t r = 0; r < K; ++r)
                for (int j = 0; j < k; ++j)
                    if ((r >> (k - 1 - j)) & 1) off[r] |= bit(op.targets[j]);
            for (int j = 0; j < k; ++j) tmask |= bit(op.targets[j]);
            std::vector<cplx> x(K);
            for (uint64_t i = 0; i < N; ++i) {
                if ((i & tmask) || !ok(i)) continue;
                for (uint64_t c = 0; c < K; ++c) x[c] = v[i | off[c]];
                for (uint64_t r = 0; r < K; ++r) {
                    cplx acc = 0;
                    for (uint64_t c = 0; c < K; ++c) acc += cplx(op.u[2 * (r * K + c)], op.u[2 * (r * K + c) + 1]) * x[c];
                    v[i | off[r]] = acc;
                }
            }
            break;
        }
        default:
            break;
    }
}

void diag_table_apply(std::vector<cplx>& v, const DiagPass& d) {
    const size_t nc = d.chunks.size();
    for (uint64_t i = 0; i < v.size(); ++i) {
        bool req = true;
        for (int b : d.required) req = req && ((i >> b) & 1);
        if (!req) continue;
        uint32_t g = 0;
        for (size_t j = 0; j < d.anchor.size(); ++j) g |= (uint32_t)((i >> d.anchor[j]) & 1) << j;
        cplx f = 1.0;
        for (size_t c = 0; c < nc; ++c) {
            uint32_t x = 0;
            for (size_t j = 0; j < d.chunks[c].size(); ++j) x |= (uint32_t)((i >> d.chunks[c][j]) & 1) << j;
            f *= d.table[(g * nc + c) * 32 + x];
        }
        v[i] *= f;
    }
}

void permute_bits(std::vector<cplx>& v, int n, const std::vector<int>& beta) {
    std::vector<cplx> out(v.size());
    for (uint64_t i = 0; i < v.size(); ++i) {
        uint64_t j = 0;
        for (int b = 0; b < n; ++b)
            if ((i >> b) & 1) j |= 1ull << beta[b];
        out[j] = v[i];
    }
    v.swap(out);
}

bool ctrl_ok(int j, uint32_t creg, bool any, bool g_any, bool t_all, bool t_any) {
    if (any) return g_any || t_any || ((uint32_t)j & creg) != 0u;
    return t_all && (((uint32_t)j & creg) == creg);
}

void tile_apply(std::vector<cplx>& v, int n, const TilePass& tp) {
    const int m = tp.m, W = m - kTileRegBits;
    const uint64_t tiles = (1ull << n) >> m;
    std::vector<cplx> tile((size_t)1 << m);
    for (uint64_t t = 0; t < tiles; ++t) {
        uint64_t base = t;
        for (int p = 0; p < m; ++p) {
            const int q = tp.local[p];
            base = ((base >> q) << (q + 1)) | (base & ((1ull << q) - 1ull));
        }
        auto gidx = [&](uint32_t l) {
            uint64_t g = base;
            for (int p = 0; p < m; ++p)
                if ((l >> p) & 1) g |= 1ull << tp.local[p];
            return g;
        };
        for (uint32_t l = 0; l < tile.size(); ++l) tile[l] = v[gidx(l)];
        // folded terms: one value per tile from the outer bits
        std::vector<cplx> fold_val(tp.folds.size(), cplx(1));
        for (size_t f = 0; f < tp.folds.size(); ++f)
            for (uint32_t e = tp.folds[f].begin; e < tp.folds[f].begin + tp.folds[f].count; ++e)
                if ((base & tp.fold_terms[e].m_glb) == tp.fold_terms[e].m_glb)
                    fold_val[f] *= cplx(tp.fold_z128[2 * e], tp.fold_z128[2 * e + 1]);
        for (const auto& S : tp.segs) {
            for (uint32_t tid = 0; tid < (1u << W); ++tid) {
                uint32_t lthr = 0;
                for (int b = 0; b < W; ++b) lthr |= ((tid >> b) & 1u) << S.thr[b];
                auto addr = [&](int j) {
                    uint32_t l = lthr;
                    for (int r = 0; r < kTileRegBits; ++r)
                        if ((j >> r) & 1) l |= 1u << S.reg[r];
                    return l;
                };
                cplx x[16];
                for (int j = 0; j < 16; ++j) x[j] = tile[addr(j)];
                for (uint32_t o = 0; o < S.op_count; ++o) {
                    const TileOpDev& op = tp.ops[S.op_begin + o];
                    if (tile_code_is_phases(op.code)) {
                        // every term multiplies the amplitudes whose register
                        // index holds its register mask (both device forms
                        // compute this product)
                        for (uint32_t k = 0; k < op.term_count; ++k) {
                            const size_t e = op.term_begin + k;
                            const TileTermDev& t = tp.terms[e];
                            if ((base & t.m_glb) != t.m_glb || (lthr & t.m_thr) != t.m_thr) continue;
                            const cplx z = t.fold >= 0 ? fold_val[t.fold] : cplx(tp.z128[2 * e], tp.z128[2 * e + 1]);
                            for (int j = 0; j < 16; ++j)
                                if (((uint32_t)j & t.m_reg) == t.m_reg) x[j] *= z;
                        }
                        continue;
                    }
                    const bool any = op.any != 0;
                    const bool g_all = (base & op.c_glb) == op.c_glb, g_any = (base & op.c_glb) != 0;
                    if (!any && !g_all) continue;
                    const bool t_all = (lthr & op.c_thr) == op.c_thr, t_any = (lthr & op.c_thr) != 0;
                    // code: MAT 4 * form + slot, XCH 12 + ..., MATR 26 + ..., MATP 38 + slot,
                    // MATP set 42 + mask (read as the MATP of its first slot; the
                    // set's other ops follow as MATP ops);
                    // the form only selects how the kernel evaluates the
                    // same controls
                    const int slot = tile_code_slot(op.code);
                    const bool mat = op.code < 12;
                    const bool real = op.code >= kTileCodeMatR && op.code < kTileCodeMatP;
                    const bool rotp = op.code >= kTileCodeMatP;
                    const int form = tile_code_form(op.code);
                    if (form == 0 && (op.c_reg || op.c_thr || op.c_glb || any)) throw 1;
                    if (form == 1 && (op.c_thr || op.c_glb || any || !op.c_reg)) throw 1;
                    for (int j = 0; j < 16; ++j) {
                        if ((j >> slot) & 1) continue;
                        if (!ctrl_ok(j, op.c_reg, any, g_any, t_all, t_any)) continue;
                        const int k = j | (1 << slot);
                        if (mat) {
                            const double* m = &tp.m128[8 * (size_t)(S.op_begin + o)];
                            const cplx p = x[j], q = x[k];
                            x[j] = cplx(m[0], m[1]) * p + cplx(m[2], m[3]) * q;
                            x[k] = cplx(m[4], m[5]) * p + cplx(m[6], m[7]) * q;
                        } else if (real) {
                            const double* m = &tp.m128[8 * (size_t)(S.op_begin + o)];
                            const cplx p = x[j], q = x[k];
                            x[j] = m[0] * p + m[1] * q;
                            x[k] = m[2] * p + m[3] * q;
                        } else if (rotp) {
                            // phase, then the rotation as three shears
                            const double* m = &tp.m128[8 * (size_t)(S.op_begin + o)];
                            cplx p = x[j], q = x[k] * cplx(m[4], m[5]);
                            p -= m[0] * q;
                            q += m[1] * p;
                            p -= m[0] * q;
                            x[j] = p;
                            x[k] = q;
                        } else {
                            std::swap(x[j], x[k]);
                        }
                    }
                }
                for (int j = 0; j < 16; ++j) tile[addr(j)] = x[j];
            }
        }
        for (uint32_t l = 0; l < tile.size(); ++l) v[gidx(l)] = tile[l];
    }
}

// One op of a batched state's plain segment on a tile held as `tile`
// (tile-local positions), read as the batch kernel reads it.
void batch_op_apply(std::vector<cplx>& tile, const BatchOp& op) {
    const uint32_t T = (uint32_t)tile.size();
    auto ok = [&](uint32_t i) { return !op.cmask || (op.any ? (i & op.cmask) != 0 : (i & op.cmask) == op.cmask); };
    const cplx m00(op.m[0], op.m[1]), m01(op.m[2], op.m[3]), m10(op.m[4], op.m[5]), m11(op.m[6], op.m[7]);
    const uint32_t t = 1u << op.t0;
    for (uint32_t i = 0; i < T; ++i) {
        switch (op.kind) {
            case AQS_OP_MAT1:
                if (!(i & t) && ok(i)) {
                    const cplx a = tile[i], b = tile[i | t];
                    tile[i] = m00 * a + m01 * b;
                    tile[i | t] = m10 * a + m11 * b;
                }
                break;
            case AQS_OP_X:
                if (!(i & t) && ok(i)) std::swap(tile[i], tile[i | t]);
                break;
            case kBatchOpMatP:
                if (op.cmask) throw 1;
                if (!(i & t)) batch_rotation(tile[i], tile[i | t], op.m);
                break;
            case AQS_OP_DIAG1:
                if (ok(i)) tile[i] *= (i & t) ? m11 : m00;
                break;
            case AQS_OP_PHASE:
                if ((i & op.pmask) == op.pmask) tile[i] *= m00;
                break;
            case AQS_OP_SWAP: {
                const uint32_t a = 1u << op.t0, b = 1u << op.t1;
                if ((i & a) && !(i & b) && ok(i)) std::swap(tile[i], tile[i ^ a ^ b]);
                break;
            }
            default:
                throw 1;
        }
    }
}

// One op of a register segment on the 16 amplitudes a thread holds (register
// k = the combination k of the segment's 4 slots; `base` = tile index of
// register 0), read as the batch kernel reads it: controls outside the
// registers test `base`, slot controls test k (all-of; register segments
// hold no any-of controlled gates, so one there is an error).
void batch_reg_apply(cplx (&a)[16], const BatchOp& e, uint32_t base) {
    const cplx m00(e.m[0], e.m[1]), m01(e.m[2], e.m[3]), m10(e.m[4], e.m[5]), m11(e.m[6], e.m[7]);
    if (e.kind == AQS_OP_PHASE) {
        if ((base & e.pmask) != e.pmask) return;
        for (uint32_t k = 0; k < 16; ++k)
            if ((k & e.preg) == e.preg) a[k] *= m00;
        return;
    }
    if (e.any) throw 1;
    if ((base & e.cmask) != e.cmask) return;
    if (e.t0 < 0 || e.t0 > 3) throw 1;
    const uint32_t t = 1u << e.t0;
    for (uint32_t k = 0; k < 16; ++k) {
        if ((k & e.creg) != e.creg) continue;
        switch (e.kind) {
            case AQS_OP_MAT1:
                if (!(k & t)) {
                    const cplx x = a[k], y = a[k | t];
                    a[k] = m00 * x + m01 * y;
                    a[k | t] = m10 * x + m11 * y;
                }
                break;
            case AQS_OP_X:
                if (!(k & t)) std::swap(a[k], a[k | t]);
                break;
            case kBatchOpMatP:
                if (e.creg || e.cmask) throw 1;
                if (!(k & t)) batch_rotation(a[k], a[k | t], e.m);
                break;
            case AQS_OP_DIAG1:
                a[k] *= (k & t) ? m11 : m00;
                break;
            default:
                throw 1;
        }
    }
}

// Controls outside the tile, given the tile's bits g (as the kernel's
// resolve): false when the op does nothing on this tile.
bool batch_resolve(BatchOp& e, uint32_t g) {
    if (!e.gmask) return true;
    if (e.kind == AQS_OP_PHASE || !e.any) return (g & e.gmask) == e.gmask;
    if (g & e.gmask) {
        e.any = 0;
        e.cmask = 0;
        e.creg = 0;
        return true;
    }
    return (e.cmask | e.creg) != 0;
}

}  // namespace

extern "C" {

// Builds the tiled schedule of ops and applies it to `state` (2^n complex
// FP64, interleaved) on the host.  Returns the aqs_status of the build;
// counts[0..3] = passes, tile passes, segments, folded term groups.
int tile_emulate(const aqs_op* ops, size_t count, int n, int tile_qubits, double* state, size_t* counts) {
    aqs_schedule sch = nullptr;
    const int st = aqs_fuse_tiled(ops, count, n, tile_qubits, &sch);
    if (st != AQS_OK) return st;
    std::vector<cplx> v((size_t)1 << n);
    std::memcpy(v.data(), state, v.size() * sizeof(cplx));
    size_t tiles = 0, segs = 0, folds = 0;
    for (size_t p = 0; p < sch->passes.size(); ++p) {
        const aqs_op& op = sch->passes[p];
        if (sch->tile_index[p] >= 0) {
            const TilePass& tp = sch->tiles[sch->tile_index[p]];
            try {
                tile_apply(v, n, tp);
            } catch (int) {
                aqs_schedule_destroy(sch);
                return -2;  // dispatch form inconsistent with the controls
            }
            ++tiles;
            segs += tp.segs.size();
            folds += tp.folds.size();
        } else if (sch->diag_index[p] >= 0) {
            aqs_schedule_destroy(sch);
            return -1;  // diagonal tables are not emulated here
        } else if (op.kind == AQS_OP_PERMB) {
            for (auto& pi : sch->perm_index)
                if (pi.first == (int)p) permute_bits(v, n, sch->perms[pi.second]);
        } else {
            plain_apply(v, n, op);
        }
    }
    std::memcpy(state, v.data(), v.size() * sizeof(cplx));
    counts[0] = sch->passes.size();
    counts[1] = tiles;
    counts[2] = segs;
    counts[3] = folds;
    aqs_schedule_destroy(sch);
    return AQS_OK;
}

// Builds the dense-block schedule of ops (aqs_fuse, blocks of at most kmax
// qubits) and applies it to `state` on the host.  counts[0..3] = passes,
// dense blocks, diagonal tables, bit permutations.
int block_emulate(const aqs_op* ops, size_t count, int n, int kmax, double* state, size_t* counts) {
    aqs_schedule sch = nullptr;
    const int st = aqs_fuse(ops, count, n, kmax, &sch);
    if (st != AQS_OK) return st;
    std::vector<cplx> v((size_t)1 << n);
    std::memcpy(v.data(), state, v.size() * sizeof(cplx));
    size_t blocks = 0, tables = 0, perms = 0;
    for (size_t p = 0; p < sch->passes.size(); ++p) {
        const aqs_op& op = sch->passes[p];
        if (sch->diag_index[p] >= 0) {
            diag_table_apply(v, sch->diags[sch->diag_index[p]]);
            ++tables;
        } else if (op.kind == AQS_OP_PERMB) {
            for (auto& pi : sch->perm_index)
                if (pi.first == (int)p) permute_bits(v, n, sch->perms[pi.second]);
            ++perms;
        } else {
            blocks += op.kind == AQS_OP_MATK;
            plain_apply(v, n, op);
        }
    }
    std::memcpy(state, v.data(), v.size() * sizeof(cplx));
    counts[0] = sch->passes.size();
    counts[1] = blocks;
    counts[2] = tables;
    counts[3] = perms;
    aqs_schedule_destroy(sch);
    return AQS_OK;
}

// Lowers the ops of one batched state for LDS tiles of 2^m amplitudes
// (batch_absorb + batch_phases) and applies the program to `state` (2^n
// complex FP64) on the host.  Each phase is read as the batch kernel reads
// it: per tile t, local index j -> state index with zeros inserted at the
// phase's outside bits (ascending) and t's bits placed there; controls
// outside the tile resolved once per tile (all-of: every bit set, else the
// op is skipped; any-of: one set bit makes the op unconditional, none and no
// local controls skips it).  A plain segment applies its ops over the tile
// one by one; a register segment gives each group of 16 amplitudes (the
// combinations of its 4 slot bits) every op in turn.  A register-segment
// op with any-of controls is an error (the kernel does not read them there).  The ops' float
// matrices are applied in FP64.  counts[0..3] = ops, phases, tile loads,
// register segments.
int batch_emulate(const aqs_op* ops, size_t count, int n, int m, double* state, size_t* counts) {
    std::vector<BatchOp> prog;
    std::vector<BatchSeg> segs;
    std::vector<BatchPhase> phases;
    try {
        batch_phases(batch_absorb(ops, count, n), n, m, prog, segs, phases);
    } catch (...) {
        return -1;
    }
    std::vector<cplx> v((size_t)1 << n);
    std::memcpy(v.data(), state, v.size() * sizeof(cplx));
    const int local = n < m ? n : m;
    std::vector<cplx> tile((size_t)1 << local);
    size_t loads = 0, reg_segs = 0;
    for (const BatchPhase& ph : phases) {
        if (__builtin_popcount(ph.gbits) != n - local || (n < 32 && (ph.gbits >> n))) return -3;
        auto state_index = [&](uint32_t j, uint32_t t) {
            uint32_t i = 0;
            int jb = 0, tb = 0;
            for (int b = 0; b < n; ++b) {
                const bool outside = (ph.gbits >> b) & 1u;
                const uint32_t bit = outside ? (t >> tb++) & 1u : (j >> jb++) & 1u;
                i |= bit << b;
            }
            return i;
        };
        for (uint32_t s = ph.seg_begin; s < ph.seg_begin + ph.seg_count; ++s) reg_segs += segs[s].rmask != 0;
        for (uint32_t t = 0; t < (1u << (n - local)); ++t) {
            const uint32_t g = state_index(0, t);
            for (uint32_t j = 0; j < tile.size(); ++j) tile[j] = v[state_index(j, t)];
            for (uint32_t s = ph.seg_begin; s < ph.seg_begin + ph.seg_count; ++s) {
                const BatchSeg& sg = segs[s];
                try {
                    if (!sg.rmask) {
                        for (uint32_t o = 0; o < sg.op_count; ++o) {
                            BatchOp e = prog[sg.op_begin + o];
                            if (batch_resolve(e, g)) batch_op_apply(tile, e);
                        }
                        continue;
                    }
                    if (__builtin_popcount(sg.rmask) != 4 || (sg.rmask >> local)) return -4;
                    int pos[4];
                    for (int q = 0, r = 0; r < 32; ++r)
                        if ((sg.rmask >> r) & 1u) pos[q++] = r;
                    for (uint32_t grp = 0; grp < (tile.size() >> 4); ++grp) {
                        uint32_t base = 0;  // grp's bits at the non-slot positions, ascending
                        for (int r = 0, gb = 0; r < local; ++r)
                            if (!((sg.rmask >> r) & 1u)) base |= ((grp >> gb++) & 1u) << r;
                        cplx a[16];
                        auto at = [&](uint32_t k) {
                            uint32_t i = base;
                            for (int q = 0; q < 4; ++q)
                                if ((k >> q) & 1u) i |= 1u << pos[q];
                            return i;
                        };
                        for (uint32_t k = 0; k < 16; ++k) a[k] = tile[at(k)];
                        for (uint32_t o = 0; o < sg.op_count; ++o) {
                            BatchOp e = prog[sg.op_begin + o];
                            if (batch_resolve(e, g)) batch_reg_apply(a, e, base);
                        }
                        for (uint32_t k = 0; k < 16; ++k) tile[at(k)] = a[k];
                    }
                } catch (int) {
                    return -2;
                }
            }
            for (uint32_t j = 0; j < tile.size(); ++j) v[state_index(j, t)] = tile[j];
            ++loads;
        }
    }
    std::memcpy(state, v.data(), v.size() * sizeof(cplx));
    counts[0] = prog.size();
    counts[1] = phases.size();
    counts[2] = loads;
    counts[3] = reg_segs;
    return AQS_OK;
}

// The shape of one batched state's program, without running it:
// out[0..5] = ops, phases, segments, register segments, register-segment
// ops, tile loads.
int batch_program_shape(const aqs_op* ops, size_t count, int n, int m, size_t* out) {
    std::vector<BatchOp> prog;
    std::vector<BatchSeg> segs;
    std::vector<BatchPhase> phases;
    try {
        batch_phases(batch_absorb(ops, count, n), n, m, prog, segs, phases);
    } catch (...) {
        return -1;
    }
    size_t reg = 0, reg_ops = 0;
    for (const BatchSeg& s : segs)
        if (s.rmask) {
            ++reg;
            reg_ops += s.op_count;
        }
    out[0] = prog.size();
    out[1] = phases.size();
    out[2] = segs.size();
    out[3] = reg;
    out[4] = reg_ops;
    out[5] = phases.size() << (n > m ? n - m : 0);
    return AQS_OK;
}

}  // extern "C"
