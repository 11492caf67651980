// Cache-blocked compile(): packs gates into tile passes (tile.hip).
//
// The reference's QCircuit::compile (src/quantum.cpp:199-210) multiplies
// every gate into one 2^n x 2^n matrix and simulate() then makes one dense
// product; the per-gate path makes one state sweep per gate.  A tile pass
// makes one sweep for a whole run of gates: the state is cut into tiles of
// 2^m amplitudes that share every index bit outside a chosen local set, and
// every gate whose non-diagonal qubits are local is applied while the tile
// sits in LDS.  Diagonal gates and controls may sit on any qubit (outside
// the tile their bits are fixed per tile).
//
// Pass construction walks the remaining gates in program order.  A gate joins
// the pass when (a) its non-diagonal qubits are local or can still be made
// local (the local set grows up to m bits, bits 0..4 always in it), and
// (b) it commutes with every gate passed over so far: per index bit the walk
// records whether a skipped gate used it non-diagonally (then no later gate
// may use it) or only diagonally -- as a control or a diagonal target --
// (then later gates may use it diagonally too).  Two gates that share only
// bits both use diagonally commute, so the pass applies its gates in program
// order and every skipped gate still sees the same operator before it.
//
// Inside a pass, segments give each thread 4 register slots; a segment runs
// the pass's gates, in order and under the same commutation rule, whose
// non-diagonal qubit is in its slots, and the next segment's slots are the
// first 4 distinct targets of the gates that are ready next.
//
// Uncontrolled swaps become relabellings undone by bit permutations at the
// end (as in aqs_fuse); a controlled swap is CX(b->a) CCX(c, a->b) CX(b->a);
// dense Gate / ControlGate blocks stay plain passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"
#include "rotation.h"
#include "schedule.h"

using namespace aqs_hip;

namespace {

constexpr int kTileAutoMinQubits = 20;

// A phase term: multiply by z where every listed index bit is 1.
struct Term {
    std::vector<int> bits;
    cd z;
};

struct GOp {
    aqs_op op{};           // plain passes: the op itself
    int kind = -1;         // TILE_MAT / TILE_XCH / TILE_PHASES (diagonal gate), -1 = plain pass
    int tbit = -1;         // MAT / XCH target bit
    std::vector<int> nd;   // bits used non-diagonally
    std::vector<int> dg;   // bits used diagonally (controls, diagonal targets)
    std::vector<int> ctrl; // control bits
    bool any = false;
    double m[8] = {};
    std::vector<Term> terms;  // diagonal gates
};

class Tiler {
   public:
    Tiler(aqs_schedule_s* sch, int n, int m) : sch_(sch), n_(n), m_(m) {}

    void run(const aqs_op* ops_in, size_t count) {
        std::vector<int> L(n_);
        for (int q = 0; q < n_; ++q) L[q] = q;
        pend0_.assign(n_, cd(1));
        pend1_.assign(n_, cd(1));
        for (size_t i = 0; i < count; ++i) {
            aqs_op op = ops_in[i];
            if (op.kind == AQS_OP_SWAP && op.ncontrols == 0 && !(op.flags & AQS_CTRL_ANY)) {
                std::swap(L[op.targets[0]], L[op.targets[1]]);
                continue;
            }
            for (int j = 0; j < op.ntargets; ++j) op.targets[j] = L[op.targets[j]];
            for (int j = 0; j < op.ncontrols; ++j) op.controls[j] = L[op.controls[j]];
            uint64_t more = 0;
            for (int q = 0; q < n_; ++q)
                if ((op.more_controls >> q) & 1ull) more |= 1ull << L[q];
            op.more_controls = more;
            lower(op);
        }
        for (int b = 0; b < n_; ++b) flush_pending(b);
        std::vector<char> done(g_.size(), 0);
        size_t first = 0;
        while (true) {
            while (first < g_.size() && done[first]) ++first;
            if (first == g_.size()) break;
            if (g_[first].kind < 0) {
                plain(g_[first].op);
                done[first] = 1;
                continue;
            }
            build_pass(first, done);
        }
        emit_restore_order(sch_, L);
    }

   private:
    aqs_schedule_s* sch_;
    int n_, m_;
    std::vector<GOp> g_;
    // Per index bit, a diagonal diag(p0, p1) not yet applied: phases of
    // one-qubit gates on the bit, carried forward (a diagonal commutes with
    // every gate that uses the bit only diagonally) into the next rotation
    // on the bit, or emitted as phase terms before the next gate that uses
    // the bit otherwise.
    std::vector<cd> pend0_, pend1_;

    void flush_pending(int b) {
        const cd p0 = pend0_[b], p1 = pend1_[b];
        pend0_[b] = pend1_[b] = cd(1);
        if (p0 == cd(1) && p1 == cd(1)) return;
        GOp ph;
        ph.kind = TILE_PHASES;
        ph.dg = {b};
        if (p0 != cd(1)) ph.terms.push_back({{}, p0});
        if (p1 != p0) ph.terms.push_back({{b}, p1 / p0});
        g_.push_back(std::move(ph));
    }

    // Every gate enters the list here, after the pending diagonals of the
    // bits it uses non-diagonally.
    void push(GOp&& o) {
        for (int b : o.nd) flush_pending(b);
        g_.push_back(std::move(o));
    }

    int bit(int q) const { return n_ - 1 - q; }

    void lower(const aqs_op& op) {
        GOp o;
        o.op = op;
        o.any = (op.flags & AQS_CTRL_ANY) != 0;
        for (int j = 0; j < op.ncontrols; ++j) o.ctrl.push_back(bit(op.controls[j]));
        for (int q = 0; q < n_; ++q)
            if ((op.more_controls >> q) & 1ull) o.ctrl.push_back(bit(q));
        switch (op.kind) {
            case AQS_OP_MAT1:
            case AQS_OP_X:
                o.kind = op.kind == AQS_OP_MAT1 ? TILE_MAT : TILE_XCH;
                o.tbit = bit(op.targets[0]);
                o.nd = {o.tbit};
                o.dg = o.ctrl;
                std::memcpy(o.m, op.m, sizeof(o.m));
                if (op.kind == AQS_OP_MAT1 && !o.any) {
                    if (o.ctrl.empty() && lower_rot(o)) return;
                    if (lower_real(o)) return;
                }
                if (op.kind == AQS_OP_X && o.ctrl.empty() && !o.any) {
                    // X diag(p0, p1) = diag(p1, p0) X
                    std::swap(pend0_[o.tbit], pend1_[o.tbit]);
                    g_.push_back(std::move(o));
                    return;
                }
                break;
            case AQS_OP_DIAG1: {
                if (o.any) break;
                // diag(d0, d1) on bit t under controls C = [C] d0 * [C, t] d1 / d0
                o.kind = TILE_PHASES;
                const int t = bit(op.targets[0]);
                o.dg = o.ctrl;
                o.dg.push_back(t);
                const cd d0(op.m[0], op.m[1]), d1(op.m[6], op.m[7]);
                if (o.ctrl.empty() && d0 != cd(0) && d1 != cd(0)) {
                    pend0_[t] *= d0;
                    pend1_[t] *= d1;
                    return;
                }
                if (d0 == cd(0)) break;  // not unitary: keep the plain pass
                o.terms.push_back({o.ctrl, d0});
                o.terms.push_back({o.dg, d1 / d0});
                break;
            }
            case AQS_OP_PHASE:
                if (o.any) break;
                if (o.ctrl.empty() && op.ntargets == 1 && cd(op.m[0], op.m[1]) != cd(0)) {
                    pend1_[bit(op.targets[0])] *= cd(op.m[0], op.m[1]);
                    return;
                }
                o.kind = TILE_PHASES;
                o.dg = o.ctrl;
                for (int j = 0; j < op.ntargets; ++j) o.dg.push_back(bit(op.targets[j]));
                o.terms.push_back({o.dg, cd(op.m[0], op.m[1])});
                break;
            case AQS_OP_SWAP:
                if (!o.any) {
                    // controlled swap: CX(b->a) CCX(c.., a->b) CX(b->a)
                    const int a = op.targets[0], b = op.targets[1];
                    aqs_op x{};
                    x.kind = AQS_OP_X;
                    x.ntargets = 1;
                    x.targets[0] = a;
                    x.ncontrols = 1;
                    x.controls[0] = b;
                    lower(x);
                    aqs_op t = x;
                    t.targets[0] = b;
                    t.ncontrols = 0;
                    for (int j = 0; j < op.ncontrols; ++j) t.controls[t.ncontrols++] = op.controls[j];
                    t.more_controls = op.more_controls;
                    lower_wide_x(t, a);
                    lower(x);
                    return;
                }
                break;
            default:
                break;
        }
        if (o.kind == TILE_PHASES && o.terms.empty()) o.kind = -1;
        if (o.kind < 0) {
            o.nd.clear();
            o.dg.clear();
            for (int j = 0; j < op.ntargets; ++j) o.nd.push_back(bit(op.targets[j]));
            for (int c : o.ctrl) o.nd.push_back(c);
        }
        push(std::move(o));
    }

    // An uncontrolled one-qubit gate as U = diag(g0, g1) R diag(1, z) with
    // R = [[c, -s], [s, c]] real (c = |u00|, s = |u10|).  The bit's pending
    // diagonal diag(p0, p1) is applied first, so the op run here is
    // R diag(1, z p1 / p0) (TILE_MATP: 10 real multiply-adds per amplitude
    // pair against 16 for a complex 2x2) and p0 diag(g0, g1) becomes the pending
    // diagonal.  A qubit's phases between rotations then cost nothing.
    // Real matrices with nothing pending stay TILE_MATR.  The form holds
    // when u00 u11 / (u01 u10) is real and negative (every unitary; FP32
    // rounding of a complex one can break it): a matrix it does not
    // reproduce to 1e-12 stays TILE_MAT.
    bool lower_rot(const GOp& o) {
        const double* m = o.op.m;
        const int b = o.tbit;
        const cd u00(m[0], m[1]), u01(m[2], m[3]), u10(m[4], m[5]), u11(m[6], m[7]);
        const cd p0 = pend0_[b], p1 = pend1_[b];
        const bool real = m[1] == 0 && m[3] == 0 && m[5] == 0 && m[7] == 0;
        if (real && p0 == cd(1) && p1 == cd(1)) return false;
        if (u01 == cd(0) && u10 == cd(0)) {
            if (u00 == cd(0) || u11 == cd(0)) return false;
            pend0_[b] = p0 * u00;
            pend1_[b] = p1 * u11;
            return true;
        }
        RotationForm f;
        if (!factor_rotation(u00, u01, u10, u11, f)) return false;
        const cd ze = f.z * p1 / p0;
        GOp r = o;
        r.kind = ze == cd(1) ? TILE_MATR : TILE_MATP;
        // MATR: the real 2x2; MATP: the shears of the unit rotation, then z;
        // the rotation's norm joins the pending diagonal
        double t, su, nr;
        rotation_shears(f.c, f.s, t, su, nr);
        const double mr[8] = {f.c, -f.s, f.s, f.c, 0, 0, 0, 0};
        const double mp[8] = {t, su, 0, 0, ze.real(), ze.imag(), 0, 0};
        std::memcpy(r.m, r.kind == TILE_MATR ? mr : mp, sizeof(r.m));
        const double scale = r.kind == TILE_MATR ? 1.0 : nr;
        g_.push_back(std::move(r));
        pend0_[b] = p0 * f.g0 * scale;
        pend1_[b] = p0 * f.g1 * scale;
        return true;
    }

    // A one-qubit gate with a real matrix (H, RotY), controlled or not, as a
    // real 2x2: half the multiplies of a complex one.
    bool lower_real(const GOp& o) {
        const double* m = o.op.m;
        if (m[1] != 0 || m[3] != 0 || m[5] != 0 || m[7] != 0) return false;
        GOp r = o;
        r.kind = TILE_MATR;
        const double re[8] = {m[0], m[2], m[4], m[6], 0, 0, 0, 0};
        std::memcpy(r.m, re, sizeof(r.m));
        push(std::move(r));
        return true;
    }

    // X on `t.targets[0]` controlled by t's controls plus qubit `extra` (the
    // aqs_op form holds at most 3 controls; the tile form has no limit).
    void lower_wide_x(const aqs_op& t, int extra) {
        GOp o;
        o.op = t;
        o.kind = TILE_XCH;
        o.tbit = bit(t.targets[0]);
        o.nd = {o.tbit};
        for (int j = 0; j < t.ncontrols; ++j) o.ctrl.push_back(bit(t.controls[j]));
        for (int q = 0; q < n_; ++q)
            if ((t.more_controls >> q) & 1ull) o.ctrl.push_back(bit(q));
        o.ctrl.push_back(bit(extra));
        o.dg = o.ctrl;
        push(std::move(o));
    }

    void plain(const aqs_op& op) {
        std::vector<double> mat;
        if (op.kind == AQS_OP_MATK) {
            const size_t K = (size_t)1 << op.ntargets;
            mat.assign(op.u, op.u + 2 * K * K);
        }
        sch_->push_plain(op, std::move(mat));
    }

    // Per-bit commutation state of the gates passed over: 0 untouched,
    // 1 used only diagonally, 2 used non-diagonally.
    static bool blocked(const GOp& o, const std::vector<char>& st) {
        for (int b : o.nd)
            if (st[b] != 0) return true;
        for (int b : o.dg)
            if (st[b] == 2) return true;
        return false;
    }
    static void pass_over(const GOp& o, std::vector<char>& st) {
        for (int b : o.nd) st[b] = 2;
        for (int b : o.dg) st[b] = std::max<char>(st[b], 1);
    }

    void build_pass(size_t first, std::vector<char>& done) {
        std::vector<char> local(n_, 0), st(n_, 0);
        int nlocal = 0;
        for (int b = 0; b < std::min(5, n_); ++b) local[b] = 1, ++nlocal;
        std::vector<size_t> members;
        int saturated = 0;  // bits at state 2
        for (size_t i = first; i < g_.size() && saturated < n_; ++i) {
            if (done[i]) continue;
            const GOp& o = g_[i];
            bool ok = o.kind >= 0 && !blocked(o, st);
            int need = 0;
            if (ok)
                for (int b : o.nd)
                    if (!local[b]) ++need;
            if (ok && nlocal + need > m_) ok = false;
            if (ok) {
                for (int b : o.nd)
                    if (!local[b]) local[b] = 1, ++nlocal;
                members.push_back(i);
            } else {
                for (int b : o.nd)
                    if (st[b] != 2) ++saturated;
                pass_over(o, st);
            }
        }
        // fill the local set with the lowest remaining bits (longer HBM runs)
        for (int b = 0; b < n_ && nlocal < m_; ++b)
            if (!local[b]) local[b] = 1, ++nlocal;
        TilePass tp;
        tp.m = m_;
        std::vector<int> lp(n_, -1);
        for (int b = 0; b < n_; ++b)
            if (local[b]) {
                lp[b] = (int)tp.local.size();
                tp.local.push_back(b);
            }
        segments(members, lp, tp);
        check_pass(tp);
        for (size_t i : members) done[i] = 1;
        if (tp.segs.empty()) return;  // the gates cancelled exactly
        aqs_op op{};
        op.kind = AQS_OP_TILE;
        sch_->push_plain(op);
        sch_->tile_index.back() = (int)sch_->tiles.size();
        sch_->tiles.push_back(std::move(tp));
    }

    void segments(const std::vector<size_t>& members, const std::vector<int>& lp, TilePass& tp) {
        std::vector<char> mdone(members.size(), 0);
        size_t left = members.size();
        while (left) {
            // slots: the first 4 distinct targets among gates that are ready
            std::vector<int> reg;
            std::vector<char> st(n_, 0);
            for (size_t k = 0; k < members.size() && reg.size() < (size_t)kTileRegBits; ++k) {
                if (mdone[k]) continue;
                const GOp& o = g_[members[k]];
                if (blocked(o, st)) {
                    pass_over(o, st);
                    continue;
                }
                if (o.kind == TILE_PHASES) continue;
                const int p = lp[o.tbit];
                if (std::find(reg.begin(), reg.end(), p) == reg.end()) reg.push_back(p);
            }
            for (int p = 0; p < m_ && reg.size() < (size_t)kTileRegBits; ++p)
                if (std::find(reg.begin(), reg.end(), p) == reg.end()) reg.push_back(p);
            // run every gate that is ready and fits the slots
            TileSegDev sd{};
            for (int r = 0; r < kTileRegBits; ++r) sd.reg[r] = reg[r];
            int t = 0;
            for (int p = 0; p < m_; ++p)
                if (std::find(reg.begin(), reg.end(), p) == reg.end()) sd.thr[t++] = p;
            sd.op_begin = (uint32_t)tp.ops.size();
            std::fill(st.begin(), st.end(), 0);
            std::vector<const GOp*> run;  // consecutive diagonal gates, merged into one op
            const size_t left_before = left;
            for (size_t k = 0; k < members.size(); ++k) {
                if (mdone[k]) continue;
                const GOp& o = g_[members[k]];
                const bool fits = o.kind == TILE_PHASES || std::find(reg.begin(), reg.end(), lp[o.tbit]) != reg.end();
                if (!fits || blocked(o, st)) {
                    pass_over(o, st);
                    continue;
                }
                if (o.kind == TILE_PHASES) {
                    run.push_back(&o);
                } else {
                    flush_phases(run, lp, reg, tp);
                    push_op(o, lp, reg, tp);
                }
                mdone[k] = 1;
                --left;
            }
            flush_phases(run, lp, reg, tp);
            group_rotations(tp, sd.op_begin);
            sd.op_count = (uint32_t)tp.ops.size() - sd.op_begin;
            if (left == left_before) fail(AQS_ERR_RUNTIME, "tile schedule made no progress");
            // (phase terms that multiply to exactly 1 leave nothing to run)
            if (sd.op_count) tp.segs.push_back(sd);
        }
    }

    // Every position the kernel turns into an LDS address or shift stays in
    // range: slots < 4, local positions < m and distinct, index bits < n.
    void check_pass(const TilePass& tp) const {
        const int m = tp.m, w = m - kTileRegBits;
        bool ok = m >= 10 && m <= kTileMaxLocal && (int)tp.local.size() == m;
        for (const auto& s : tp.segs) {
            uint32_t seen = 0;
            for (int r = 0; r < kTileRegBits; ++r) {
                ok = ok && s.reg[r] >= 0 && s.reg[r] < m && !(seen >> s.reg[r] & 1u);
                if (ok) seen |= 1u << s.reg[r];
            }
            for (int b = 0; b < w; ++b) {
                ok = ok && s.thr[b] >= 0 && s.thr[b] < m && !(seen >> s.thr[b] & 1u);
                if (ok) seen |= 1u << s.thr[b];
            }
            ok = ok && (uint64_t)s.op_begin + s.op_count <= tp.ops.size();
            // a MATP set: the following ops of the segment are MATP ops on
            // the mask's other slots, ascending
            for (uint32_t o = s.op_begin; ok && o < s.op_begin + s.op_count; ++o) {
                const int code = tp.ops[o].code;
                if (code < kTileCodeMatPSet || code >= kTileCodeCount) continue;
                const uint32_t mask = (uint32_t)(code - kTileCodeMatPSet);
                const uint32_t pc = (uint32_t)__builtin_popcount(mask);
                ok = pc >= 2 && o + pc <= s.op_begin + s.op_count;
                uint32_t rest = mask & (mask - 1);
                for (uint32_t k = 1; ok && k < pc; ++k, rest &= rest - 1)
                    ok = tp.ops[o + k].code == kTileCodeMatP + __builtin_ctz(rest);
            }
        }
        for (const auto& o : tp.ops) {
            ok = ok && o.code >= 0 && o.code < kTileCodeCount && o.c_reg < (1u << kTileRegBits) &&
                 o.c_thr < (1u << m) && (o.c_glb >> n_) == 0;
            if (!tile_code_is_phases(o.code)) {
                const int form = tile_code_form(o.code);
                ok = ok && (form == 2 || (!o.any && o.c_thr == 0 && o.c_glb == 0 && (form == 1) == (o.c_reg != 0)));
            }
            if (tile_code_is_phases(o.code)) {
                ok = ok && o.term_count > 0 && (uint64_t)o.term_begin + o.term_count <= tp.terms.size();
                // groups tile the op's terms exactly; a real group holds
                // real, unfolded terms only
                for (uint32_t k = 0; ok && k < o.term_count;) {
                    const TileTermDev& first = tp.terms[o.term_begin + k];
                    const uint32_t g = first.group;
                    ok = g > 0 && k + g <= o.term_count && first.real <= 1 && first.folded <= g;
                    for (uint32_t e = k; ok && e < k + g; ++e) {
                        const size_t t = o.term_begin + e;
                        ok = (tp.terms[t].fold >= 0) == (e >= k + g - first.folded);
                        if (first.real) ok = ok && tp.terms[t].zf[1] == 0.0f && tp.z128[2 * t + 1] == 0.0;
                    }
                    k += g ? g : 1;
                }
            }
        }
        ok = ok && tp.m128.size() == 8 * tp.ops.size() && tp.z128.size() == 2 * tp.terms.size();
        for (const auto& t : tp.terms)
            ok = ok && t.m_reg < (1u << kTileRegBits) && t.m_thr < (1u << m) && (t.m_glb >> n_) == 0 &&
                 t.fold >= -1 && t.fold < (int32_t)tp.folds.size() && (t.fold < 0 || t.m_glb == 0);
        ok = ok && tp.folds.size() <= (size_t)kTileMaxFolds && tp.fold_z128.size() == 2 * tp.fold_terms.size();
        for (const auto& f : tp.folds) ok = ok && (uint64_t)f.begin + f.count <= tp.fold_terms.size();
        for (const auto& f : tp.fold_terms) ok = ok && (f.m_glb >> n_) == 0;
        if (!ok) fail(AQS_ERR_RUNTIME, "tile schedule malformed");
    }

    // Consecutive MATP ops on distinct slots commute (uncontrolled gates on
    // distinct qubits): each maximal run of them is put in slot order and
    // marked on its first op as one MATP set, which the kernel dispatches
    // once.
    static void group_rotations(TilePass& tp, size_t begin) {
        const size_t end = tp.ops.size();
        for (size_t i = begin; i < end;) {
            uint32_t mask = 0;
            size_t j = i;
            while (j < end && tp.ops[j].code >= kTileCodeMatP && tp.ops[j].code < kTileCodeMatPSet &&
                   !((mask >> (tp.ops[j].code - kTileCodeMatP)) & 1u)) {
                mask |= 1u << (tp.ops[j].code - kTileCodeMatP);
                ++j;
            }
            if (j - i < 2) {
                i = std::max(j, i + 1);
                continue;
            }
            std::vector<TileOpDev> ops(tp.ops.begin() + i, tp.ops.begin() + j);
            std::vector<double> m(tp.m128.begin() + 8 * i, tp.m128.begin() + 8 * j);
            std::vector<size_t> order(j - i);
            for (size_t k = 0; k < order.size(); ++k) order[k] = k;
            std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ops[a].code < ops[b].code; });
            for (size_t k = 0; k < order.size(); ++k) {
                tp.ops[i + k] = ops[order[k]];
                std::copy(m.begin() + 8 * order[k], m.begin() + 8 * order[k] + 8, tp.m128.begin() + 8 * (i + k));
            }
            tp.ops[i].code = kTileCodeMatPSet + (int)mask;
            i = j;
        }
    }

    static int slot_of(const std::vector<int>& reg, int p) {
        return (int)(std::find(reg.begin(), reg.end(), p) - reg.begin());
    }

    // Split index bits into register slots, thread-held positions and bits
    // outside the tile for the segment layout `reg`.
    void masks(const std::vector<int>& bits, const std::vector<int>& lp, const std::vector<int>& reg, uint32_t& m_reg,
               uint32_t& m_thr, uint64_t& m_glb) const {
        m_reg = m_thr = 0;
        m_glb = 0;
        for (int b : bits) {
            if (lp[b] < 0)
                m_glb |= 1ull << b;
            else if (slot_of(reg, lp[b]) < kTileRegBits)
                m_reg |= 1u << slot_of(reg, lp[b]);
            else
                m_thr |= 1u << lp[b];
        }
    }

    void push_op(const GOp& o, const std::vector<int>& lp, const std::vector<int>& reg, TilePass& tp) const {
        TileOpDev d{};
        d.any = o.any ? 1 : 0;
        masks(o.ctrl, lp, reg, d.c_reg, d.c_thr, d.c_glb);
        // controls outside the tile go through the per-thread select form,
        // so the kernel's op loop has no path that leaves the registers as
        // they are (such a path made the compiler copy all 16 amplitudes
        // before every op)
        const int form = (d.any || d.c_thr || d.c_glb) ? 2 : d.c_reg ? 1 : 0;
        d.code = tile_code(o.kind, form, slot_of(reg, lp[o.tbit]), false);
        for (int i = 0; i < 8; ++i) d.mf[i] = (float)o.m[i];
        tp.ops.push_back(d);
        tp.m128.insert(tp.m128.end(), o.m, o.m + 8);
    }

    // One PHASES op for a run of diagonal gates: their terms, with equal
    // masks multiplied together (FP64) and factors of exactly 1 dropped,
    // sorted by register mask into groups.  Per thread the kernel forms each
    // group's product; the direct form multiplies it into the 16 / 2^|mask|
    // amplitudes of the group, the zeta form into one of 16 subset factors
    // (then 32 + 16 products); the cheaper is chosen.
    void flush_phases(std::vector<const GOp*>& run, const std::vector<int>& lp, const std::vector<int>& reg,
                      TilePass& tp) const {
        if (run.empty()) return;
        struct M {
            uint32_t reg, thr;
            uint64_t glb;
            cd z;
        };
        std::vector<M> merged;
        for (const GOp* o : run)
            for (const Term& t : o->terms) {
                M d{0, 0, 0, t.z};
                masks(t.bits, lp, reg, d.reg, d.thr, d.glb);
                size_t i = 0;
                while (i < merged.size() && !(merged[i].reg == d.reg && merged[i].thr == d.thr && merged[i].glb == d.glb))
                    ++i;
                if (i == merged.size())
                    merged.push_back(d);
                else
                    merged[i].z *= t.z;
            }
        run.clear();
        merged.erase(std::remove_if(merged.begin(), merged.end(), [](const M& m) { return m.z == cd(1); }),
                     merged.end());
        if (merged.empty()) return;
        // Terms pairing one local bit with bits outside the tile: per local
        // bit, one folded term whose value the kernel forms once per tile.
        std::vector<M> kept;
        std::vector<std::pair<uint32_t, uint32_t>> fold_keys;  // (reg, thr) of each new fold
        std::vector<std::vector<M>> fold_members;
        for (const M& t : merged) {
            const int local = __builtin_popcount(t.reg) + __builtin_popcount(t.thr);
            if (local != 1 || t.glb == 0) {
                kept.push_back(t);
                continue;
            }
            size_t f = 0;
            while (f < fold_keys.size() && fold_keys[f] != std::make_pair(t.reg, t.thr)) ++f;
            if (f == fold_keys.size()) {
                if (tp.folds.size() + fold_keys.size() >= (size_t)kTileMaxFolds) {
                    kept.push_back(t);
                    continue;
                }
                fold_keys.push_back({t.reg, t.thr});
                fold_members.emplace_back();
            }
            fold_members[f].push_back(t);
        }
        std::vector<int32_t> fold_of(kept.size(), -1);
        for (size_t f = 0; f < fold_keys.size(); ++f) {
            if (fold_members[f].size() < 2) {  // a lone term gains nothing from folding
                kept.push_back(fold_members[f][0]);
                fold_of.push_back(-1);
                continue;
            }
            TileFoldDev fd{(uint32_t)tp.fold_terms.size(), (uint32_t)fold_members[f].size()};
            for (const M& t : fold_members[f]) {
                tp.fold_terms.push_back(TileFoldTermDev{t.glb, {(float)t.z.real(), (float)t.z.imag()}});
                tp.fold_z128.push_back(t.z.real());
                tp.fold_z128.push_back(t.z.imag());
            }
            kept.push_back(M{fold_keys[f].first, fold_keys[f].second, 0, cd(1)});
            fold_of.push_back((int32_t)tp.folds.size());
            tp.folds.push_back(fd);
        }
        std::vector<size_t> order(kept.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        // by register mask; in a group the folded terms last (the kernel
        // reads them from LDS in a loop of their own)
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            return kept[a].reg != kept[b].reg ? kept[a].reg < kept[b].reg : fold_of[a] < 0 && fold_of[b] >= 0;
        });
        std::vector<M> sorted_terms;
        std::vector<int32_t> sorted_fold;
        for (size_t i : order) {
            sorted_terms.push_back(kept[i]);
            sorted_fold.push_back(fold_of[i]);
        }
        merged.swap(sorted_terms);
        TileOpDev op{};
        op.term_begin = (uint32_t)tp.terms.size();
        op.term_count = (uint32_t)merged.size();
        double direct = 0, groups = 0;
        size_t gend = 0;  // end of the current group
        for (size_t i = 0; i < merged.size(); ++i) {
            TileTermDev d{};
            d.m_reg = (uint16_t)merged[i].reg;
            d.m_thr = merged[i].thr;
            d.m_glb = merged[i].glb;
            d.zf[0] = (float)merged[i].z.real();
            d.zf[1] = (float)merged[i].z.imag();
            d.fold = sorted_fold[i];
            if (i == gend) {
                size_t e = i;
                bool real = true;
                uint32_t folded = 0;
                while (e < merged.size() && merged[e].reg == merged[i].reg && e - i < 0xFFFF) {
                    real = real && sorted_fold[e] < 0 && merged[e].z.imag() == 0.0;
                    folded += sorted_fold[e] >= 0 ? 1u : 0u;
                    ++e;
                }
                d.group = (uint16_t)(e - i);
                d.real = real ? 1 : 0;
                d.folded = (uint16_t)folded;
                gend = e;
                direct += 16.0 / (double)(1u << __builtin_popcount(merged[i].reg));
                groups += 1;
            }
            tp.terms.push_back(d);
            tp.z128.push_back(merged[i].z.real());
            tp.z128.push_back(merged[i].z.imag());
        }
        op.code = tile_code(TILE_PHASES, 0, 0, 48.0 + groups < direct);
        tp.ops.push_back(op);
        tp.m128.insert(tp.m128.end(), 8, 0.0);
    }
};

// Modelled cost of one pass in full-state sweeps (read + write of every
// amplitude at the streaming rate).  Tile passes are fitted to MI355X
// timings of brickwork (depth 8, 20), random-layer (depth 10, 40) and QFT
// circuits at 26..28 qubits (mean error 1.4%, at most 4%): a pass costs
// 1.03 sweeps plus its in-register work -- a rotation with its phase (MATP)
// 0.053 of a sweep, a real 2x2 0.05, an exchange 0.066, a phase term 0.02
// and each product a phase run applies per thread 0.0045; a complex 2x2,
// absent from those circuits, is priced at 0.085 (16 multiply-adds per
// pair against a rotation's 10).  Other passes (fitted earlier at 22..28
// qubits, mean error 9%): an MFMA block 1.17-1.2, a
// diagonal table 1.25 (0.05 when it rides in a neighbouring MFMA block), a
// bit permutation 1.8, a plain one-qubit pass 1, a sparse pass its touched
// fraction.
double pass_cost(const aqs_schedule_s* sch, size_t p) {
    const aqs_op& op = sch->passes[p];
    if (sch->tile_index[p] >= 0) {
        const TilePass& t = sch->tiles[sch->tile_index[p]];
        double c = 1.03;
        for (const auto& o : t.ops) {
            if (o.code < 12) {
                c += 0.085;
            } else if (o.code >= kTileCodeMatP) {
                c += 0.053;
            } else if (o.code >= kTileCodeMatR) {
                c += 0.05;
            } else if (!tile_code_is_phases(o.code)) {
                c += 0.066;
            } else {
                const bool zeta = o.code == kTileCodePhases + 1;
                double products = zeta ? 48.0 : 0.0;
                for (uint32_t k = 0; k < o.term_count; ++k) {
                    const TileTermDev& tt = t.terms[o.term_begin + k];
                    if (tt.group) products += zeta ? 1.0 : 16.0 / (double)(1u << __builtin_popcount(tt.m_reg));
                }
                c += 0.02 * o.term_count + 0.0045 * products;
            }
        }
        return c;
    }
    if (sch->diag_index[p] >= 0) return 1.25;
    if (op.more_controls) return 0.4;  // a small fraction of the state, found by index tests
    switch (op.kind) {
        case AQS_OP_PERMB:
            return 1.8;
        case AQS_OP_MATK:
            if (op.ncontrols == 0 && op.ntargets == 5) return 1.2;
            if (op.ncontrols == 0 && op.ntargets == 4) return 1.17;
            return 1.0 + 0.05 * (double)((size_t)1 << op.ntargets);
        case AQS_OP_PHASE: {
            const int bits = op.ntargets + op.ncontrols;
            return bits >= 2 ? 0.4 : 0.6;
        }
        case AQS_OP_SWAP:
            return 0.6;
        default:
            return op.ncontrols ? 0.6 : 1.0;
    }
}

// Index bits of the tile an MFMA block sweeps (its targets and the five
// lowest other bits), or empty when the op is not an MFMA block.
std::vector<int> mfma_tile_bits(const aqs_schedule_s* sch, size_t p) {
    if (p >= sch->passes.size()) return {};
    const aqs_op& op = sch->passes[p];
    const int n = sch->n;
    if (op.kind != AQS_OP_MATK || op.ncontrols || op.more_controls || (op.ntargets != 4 && op.ntargets != 5) ||
        n < op.ntargets + 5)
        return {};
    std::vector<int> T;
    for (int j = 0; j < op.ntargets; ++j) T.push_back(n - 1 - op.targets[j]);
    for (int b = 0, added = 0; b < n && added < 5; ++b)
        if (std::find(T.begin(), T.end(), b) == T.end()) {
            T.push_back(b);
            ++added;
        }
    return T;
}

// A diagonal table next to an MFMA block whose tile holds its anchor runs
// inside that block's sweep (apply_matk_mfma epilogue / prologue).
bool diag_rides_block(const aqs_schedule_s* sch, size_t p) {
    const auto& anchor = sch->diags[sch->diag_index[p]].anchor;
    for (size_t q : {p - 1, p + 1}) {
        if (p == 0 && q == p - 1) continue;
        const std::vector<int> T = mfma_tile_bits(sch, q);
        if (T.empty()) continue;
        bool inside = true;
        for (int b : anchor) inside = inside && std::find(T.begin(), T.end(), b) != T.end();
        if (inside) return true;
    }
    return false;
}

}  // namespace

double aqs_hip::schedule_cost(const aqs_schedule_s* sch) {
    double c = 0;
    for (size_t p = 0; p < sch->passes.size(); ++p)
        c += (sch->diag_index[p] >= 0 && diag_rides_block(sch, p)) ? 0.05 : pass_cost(sch, p);
    return c;
}

extern "C" {

aqs_status aqs_fuse_tiled(const aqs_op* ops, size_t count, int nqubits, int tile_qubits, aqs_schedule* out) {
    if (nqubits < 10) return aqs_fuse(ops, count, nqubits, 5, out);
    return guard([&] {
        if (!out) fail(AQS_ERR_INVALID_ARGUMENT, "null output schedule");
        *out = nullptr;
        if (count && !ops) fail(AQS_ERR_INVALID_ARGUMENT, "null op array");
        if (nqubits > AQS_MAX_QUBITS) fail(AQS_ERR_INVALID_ARGUMENT, "bad qubit count");
        if (tile_qubits == 0) tile_qubits = kTileMaxLocal;
        if (tile_qubits < 10 || tile_qubits > kTileMaxLocal)
            fail(AQS_ERR_INVALID_ARGUMENT, "tile_qubits must be 10..13 (0 = 13)");
        for (size_t i = 0; i < count; ++i) validate_op(ops[i], nqubits);
        auto* sch = new aqs_schedule_s();
        sch->n = nqubits;
        sch->source_ops = count;
        try {
            Tiler t(sch, nqubits, std::min(tile_qubits, nqubits));
            t.run(ops, count);
            for (size_t p = 0; p < sch->passes.size(); ++p)
                if (sch->passes[p].kind == AQS_OP_MATK) sch->passes[p].u = sch->mats[p].data();
        } catch (...) {
            delete sch;
            throw;
        }
        *out = sch;
    });
}

aqs_status aqs_schedule_cost(aqs_schedule sch, double* sweeps) {
    return guard([&] {
        if (!sch) fail(AQS_ERR_INVALID_ARGUMENT, "null schedule");
        if (sweeps) *sweeps = schedule_cost(sch);
    });
}

aqs_status aqs_fuse_auto(const aqs_op* ops, size_t count, int nqubits, aqs_schedule* out) {
    if (!out) {
        set_error("null output schedule");
        return AQS_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    aqs_schedule blocks = nullptr, tiled = nullptr;
    // below 2^20 amplitudes the state stays in the caches and launches of
    // the block schedule are cheaper than in-tile work (measured: at 19
    // qubits blocks win, at 20-21 tiles by 5-40% on brickwork and random
    // layers)
    if (nqubits < kTileAutoMinQubits) return aqs_fuse(ops, count, nqubits, 5, out);
    // both schedules, the tile one on a thread of its own (its error
    // message, kept per thread, is carried back)
    aqs_status tst = AQS_OK;
    std::string terr;
    std::thread worker;
    bool threaded = false;
    auto fuse_tiles = [&] {
        tst = aqs_fuse_tiled(ops, count, nqubits, 0, &tiled);
        if (tst != AQS_OK) terr = aqs_last_error();
    };
    try {
        worker = std::thread(fuse_tiles);
        threaded = true;
    } catch (...) {
        // no thread to be had: both run here
    }
    aqs_status st = aqs_fuse(ops, count, nqubits, 5, &blocks);
    if (threaded) worker.join();
    if (st != AQS_OK) {
        if (tiled) aqs_schedule_destroy(tiled);
        return st;
    }
    if (!threaded) fuse_tiles();
    if (tst != AQS_OK) {
        aqs_schedule_destroy(blocks);
        set_error(terr);
        return tst;
    }
    const bool keep_tiled = schedule_cost(tiled) < schedule_cost(blocks);
    aqs_schedule_destroy(keep_tiled ? blocks : tiled);
    *out = keep_tiled ? tiled : blocks;
    return AQS_OK;
}

}  // extern "C"
