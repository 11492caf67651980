// compile(): commutation-aware gate fusion (host side, FP64).
//
// The reference's QCircuit::compile (src/quantum.cpp:199-210) multiplies
// every gate into one 2^n x 2^n matrix.  This pass turns the gate list into
// device passes of three kinds:
//
//   dense blocks   <= max_qubits (<= 5) qubits, an FP64 unitary applied with
//                  MFMA (4, 5 qubits) or the small-block kernel;
//   diagonal table one sweep applying many deferred phase terms (diag.hip);
//   plain ops      permutations that fit no open block, ops wider than a block.
//
// Rules, in program order:
//   - open dense blocks have disjoint qubit sets, so they commute and grow in
//     any order; an op that meets open blocks merges with them when the union
//     fits, else those blocks are emitted first;
//   - a run of consecutive diagonal ops (they commute) is reordered: ops that
//     add no new qubit to the open blocks first, then by how soon a
//     non-diagonal op next touches the qubits they would add;
//   - a diagonal op that does not fit is deferred as phase terms instead of
//     closing blocks; a later non-diagonal op touching a deferred term's
//     qubits first emits the open blocks that meet the terms, then the terms
//     (every gate added to those blocks after the term commutes with it: it is
//     diagonal or acts on other qubits);
//   - a Swap that fits no open block becomes a plain pass (it moves half the
//     state; a dense block would move all of it).
// aqs_fuse also fuses a second program order in which every diagonal op is
// hoisted to just after the last non-diagonal op sharing one of its qubits
// (diagonal ops commute with each other and with ops on other qubits), and
// keeps the schedule of lower modelled cost.  Ladders written "phases, then
// the Hadamard" (inverse_fourier_transform) become "Hadamard, then its
// phases", whose deferred terms flush once per block instead of per qubit.
// Each block's unitary is accumulated in FP64 and rounded to FP32 once, at
// upload for a complex64 state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "internal.h"
#include "schedule.h"

using namespace aqs_hip;

namespace {

// Apply an op to a local state vector of k qubits (local qubit j = bit k-1-j).
void host_apply(std::vector<cd>& v, int k, const aqs_op& op) {
    const uint64_t N = 1ull << k;
    auto bit = [&](int q) { return 1ull << (k - 1 - q); };
    uint64_t cmask = 0;
    for (int j = 0; j < op.ncontrols; ++j) cmask |= bit(op.controls[j]);
    const bool any = (op.flags & AQS_CTRL_ANY) != 0;
    auto ctrl_ok = [&](uint64_t i) {
        if (!cmask) return true;
        return any ? (i & cmask) != 0 : (i & cmask) == cmask;
    };
    const cd m00(op.m[0], op.m[1]), m01(op.m[2], op.m[3]), m10(op.m[4], op.m[5]), m11(op.m[6], op.m[7]);
    switch (op.kind) {
        case AQS_OP_MAT1:
        case AQS_OP_DIAG1:
        case AQS_OP_X: {
            const uint64_t t = bit(op.targets[0]);
            for (uint64_t i = 0; i < N; ++i) {
                if ((i & t) || !ctrl_ok(i)) continue;
                cd a = v[i], b = v[i | t];
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
                if ((i & a) && !(i & b) && ctrl_ok(i)) std::swap(v[i], v[i ^ a ^ b]);
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
            const int kk = op.ntargets;
            const uint64_t K = 1ull << kk;
            std::vector<uint64_t> off(K);
            uint64_t tmask = 0;
            for (uint64_t r = 0; r < K; ++r) {
                uint64_t o = 0;
                for (int j = 0; j < kk; ++j)
                    if ((r >> (kk - 1 - j)) & 1) o |= bit(op.targets[j]);
                off[r] = o;
            }
            for (int j = 0; j < kk; ++j) tmask |= bit(op.targets[j]);
            std::vector<cd> x(K);
            for (uint64_t i = 0; i < N; ++i) {
                if ((i & tmask) || !ctrl_ok(i)) continue;
                for (uint64_t c = 0; c < K; ++c) x[c] = v[i | off[c]];
                for (uint64_t r = 0; r < K; ++r) {
                    cd acc = 0;
                    for (uint64_t c = 0; c < K; ++c) acc += cd(op.u[2 * (r * K + c)], op.u[2 * (r * K + c) + 1]) * x[c];
                    v[i | off[r]] = acc;
                }
            }
            break;
        }
        default:
            break;
    }
}

// Apply an op to every column of a row-major 2^k x 2^k matrix M (M <- G M):
// the op's pairs and groups are rows, each updated along its whole length.
void apply_rows(std::vector<cd>& M, int k, const aqs_op& op) {
    const uint64_t N = 1ull << k;
    auto bit = [&](int q) { return 1ull << (k - 1 - q); };
    uint64_t cmask = 0;
    for (int j = 0; j < op.ncontrols; ++j) cmask |= bit(op.controls[j]);
    const bool any = (op.flags & AQS_CTRL_ANY) != 0;
    auto ctrl_ok = [&](uint64_t i) {
        if (!cmask) return true;
        return any ? (i & cmask) != 0 : (i & cmask) == cmask;
    };
    auto row = [&](uint64_t i) { return M.data() + i * N; };
    const cd m00(op.m[0], op.m[1]), m01(op.m[2], op.m[3]), m10(op.m[4], op.m[5]), m11(op.m[6], op.m[7]);
    switch (op.kind) {
        case AQS_OP_MAT1:
        case AQS_OP_DIAG1:
        case AQS_OP_X: {
            const uint64_t t = bit(op.targets[0]);
            for (uint64_t i = 0; i < N; ++i) {
                if ((i & t) || !ctrl_ok(i)) continue;
                cd* a = row(i);
                cd* b = row(i | t);
                if (op.kind == AQS_OP_MAT1) {
                    for (uint64_t c = 0; c < N; ++c) {
                        const cd x = a[c], y = b[c];
                        a[c] = m00 * x + m01 * y;
                        b[c] = m10 * x + m11 * y;
                    }
                } else if (op.kind == AQS_OP_DIAG1) {
                    for (uint64_t c = 0; c < N; ++c) {
                        a[c] *= m00;
                        b[c] *= m11;
                    }
                } else {
                    std::swap_ranges(a, a + N, b);
                }
            }
            break;
        }
        case AQS_OP_SWAP: {
            const uint64_t a = bit(op.targets[0]), b = bit(op.targets[1]);
            for (uint64_t i = 0; i < N; ++i)
                if ((i & a) && !(i & b) && ctrl_ok(i)) std::swap_ranges(row(i), row(i) + N, row(i ^ a ^ b));
            break;
        }
        case AQS_OP_PHASE: {
            uint64_t mask = cmask;
            for (int j = 0; j < op.ntargets; ++j) mask |= bit(op.targets[j]);
            for (uint64_t i = 0; i < N; ++i)
                if ((i & mask) == mask)
                    for (uint64_t c = 0; c < N; ++c) row(i)[c] *= m00;
            break;
        }
        case AQS_OP_MATK: {
            const int kk = op.ntargets;
            const uint64_t K = 1ull << kk;
            std::vector<uint64_t> off(K);
            uint64_t tmask = 0;
            for (uint64_t r = 0; r < K; ++r) {
                uint64_t o = 0;
                for (int j = 0; j < kk; ++j)
                    if ((r >> (kk - 1 - j)) & 1) o |= bit(op.targets[j]);
                off[r] = o;
            }
            for (int j = 0; j < kk; ++j) tmask |= bit(op.targets[j]);
            std::vector<cd> x(K * N);
            for (uint64_t i = 0; i < N; ++i) {
                if ((i & tmask) || !ctrl_ok(i)) continue;
                for (uint64_t c = 0; c < K; ++c) std::copy(row(i | off[c]), row(i | off[c]) + N, x.data() + c * N);
                for (uint64_t r = 0; r < K; ++r) {
                    cd* out = row(i | off[r]);
                    std::fill(out, out + N, cd(0));
                    for (uint64_t c = 0; c < K; ++c) {
                        const cd u(op.u[2 * (r * K + c)], op.u[2 * (r * K + c) + 1]);
                        if (u == cd(0)) continue;
                        const cd* xc = x.data() + c * N;
                        for (uint64_t j = 0; j < N; ++j) out[j] += u * xc[j];
                    }
                }
            }
            break;
        }
        default:
            break;
    }
}

std::vector<int> op_qubits(const aqs_op& op) {
    std::vector<int> q;
    for (int j = 0; j < op.ntargets; ++j) q.push_back(op.targets[j]);
    for (int j = 0; j < op.ncontrols; ++j)
        if (std::find(q.begin(), q.end(), op.controls[j]) == q.end()) q.push_back(op.controls[j]);
    for (int b = 0; b < 64; ++b)
        if ((op.more_controls >> b) & 1ull) q.push_back(b);
    return q;
}

// A qubit mask with every qubit q moved to L[q].
uint64_t relabel_mask(uint64_t mask, const std::vector<int>& L) {
    uint64_t r = 0;
    for (size_t q = 0; q < L.size(); ++q)
        if ((mask >> q) & 1ull) r |= 1ull << L[q];
    return r;
}

bool is_diagonal(const aqs_op& op) {
    if (op.flags & AQS_CTRL_ANY) return false;
    return op.kind == AQS_OP_PHASE || op.kind == AQS_OP_DIAG1;
}

// A swap touches half of the amplitudes as a plain pass but all of them as a
// dense block, so it only joins blocks it can merge into.
bool is_permutation(const aqs_op& op) { return op.kind == AQS_OP_SWAP && !(op.flags & AQS_CTRL_ANY); }

struct Block {
    std::vector<int> qubits;  // local qubit j = bit k-1-j of the block index
    std::vector<cd> U;        // row-major 2^k x 2^k
};

int local_index(const std::vector<int>& qs, int q) {
    for (size_t j = 0; j < qs.size(); ++j)
        if (qs[j] == q) return (int)j;
    return -1;
}

// op with its qubits renumbered to their positions in qs.
aqs_op localize(const aqs_op& op, const std::vector<int>& qs) {
    aqs_op loc = op;
    for (int j = 0; j < op.ntargets; ++j) loc.targets[j] = local_index(qs, op.targets[j]);
    for (int j = 0; j < op.ncontrols; ++j) loc.controls[j] = local_index(qs, op.controls[j]);
    return loc;
}

std::vector<cd> op_matrix(const aqs_op& op, const std::vector<int>& qs) {
    const int k = (int)qs.size();
    const uint64_t K = 1ull << k;
    const aqs_op loc = localize(op, qs);
    std::vector<cd> M(K * K);
    std::vector<cd> e(K);
    for (uint64_t c = 0; c < K; ++c) {
        std::fill(e.begin(), e.end(), cd(0));
        e[c] = 1.0;
        host_apply(e, k, loc);
        for (uint64_t r = 0; r < K; ++r) M[r * K + c] = e[r];
    }
    return M;
}

bool diagonal_matrix(const std::vector<cd>& U, uint64_t K) {
    for (uint64_t i = 0; i < K; ++i)
        for (uint64_t j = 0; j < K; ++j)
            if (i != j && U[i * K + j] != cd(0)) return false;
    return true;
}

// A deferred phase term: factor `phase` on amplitudes whose qubits in `q` are all 1.
struct Term {
    std::vector<int> q;
    cd phase;
};

class Fuser {
   public:
    // keep_partial: when an op does not fit the union of the open blocks it
    // touches, close only the blocks it must (keep_for) instead of all of
    // them, so that small blocks keep growing rather than leaving as
    // fragments.
    Fuser(aqs_schedule_s* sch, int n, int kmax, bool keep_partial = false)
        : sch_(sch), n_(n), kmax_(kmax), keep_partial_(keep_partial) {}

    void run(const aqs_op* ops_in, size_t count, bool hoist) {
        // Uncontrolled swaps become relabellings: L[q] is the physical qubit
        // holding logical qubit q; every other op is rewritten through L.
        std::vector<int> L(n_);
        for (int q = 0; q < n_; ++q) L[q] = q;
        std::vector<aqs_op> ops;
        ops.reserve(count);
        for (size_t i = 0; i < count; ++i) {
            aqs_op op = ops_in[i];
            if (op.kind == AQS_OP_SWAP && op.ncontrols == 0) {
                std::swap(L[op.targets[0]], L[op.targets[1]]);
                continue;
            }
            for (int j = 0; j < op.ntargets; ++j) op.targets[j] = L[op.targets[j]];
            for (int j = 0; j < op.ncontrols; ++j) op.controls[j] = L[op.controls[j]];
            op.more_controls = relabel_mask(op.more_controls, L);
            ops.push_back(op);
        }
        count = ops.size();
        if (hoist) ops = hoist_diagonals(ops);
        // non-diagonal uses per qubit, for the look-ahead of diagonal runs
        std::vector<std::vector<size_t>> uses(n_);
        for (size_t i = 0; i < count; ++i)
            if (!is_diagonal(ops[i]))
                for (int q : op_qubits(ops[i])) uses[q].push_back(i);
        size_t i = 0;
        while (i < count) {
            if (!is_diagonal(ops[i])) {
                nondiag(ops[i]);
                ++i;
                continue;
            }
            size_t e = i;
            while (e < count && is_diagonal(ops[e])) ++e;
            std::vector<size_t> run;
            for (size_t j = i; j < e; ++j) run.push_back(j);
            // qubits already in open blocks at the start of the run
            std::vector<char> inblock(n_, 0);
            for (auto& b : open_)
                for (int q : b.qubits) inblock[q] = 1;
            auto next_use = [&](int q) {
                auto it = std::lower_bound(uses[q].begin(), uses[q].end(), e);
                return it == uses[q].end() ? (size_t)-1 : *it;
            };
            struct Key {
                int fresh;
                size_t soon;
                size_t idx;
            };
            std::vector<Key> keys;
            for (size_t j : run) {
                int fresh = 0;
                size_t soon = (size_t)-1;
                for (int q : op_qubits(ops[j]))
                    if (!inblock[q]) {
                        ++fresh;
                        soon = std::min(soon, next_use(q));
                    }
                keys.push_back({fresh, soon, j});
            }
            std::stable_sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
                if (a.fresh != b.fresh) return a.fresh < b.fresh;
                return a.soon < b.soon;
            });
            for (auto& k : keys) diag(ops[k.idx]);
            i = e;
        }
        flush_terms();
        while (!open_.empty()) close_block(0);
        restore_order(L);
    }

   private:
    aqs_schedule_s* sch_;
    int n_, kmax_;
    bool keep_partial_;
    std::vector<Block> open_;
    std::vector<Term> terms_;

    // Every diagonal op moves to just after the last non-diagonal op that
    // shares a qubit with it (or to the front); non-diagonal ops keep their
    // order, and so do diagonal ops that land on the same spot.
    std::vector<aqs_op> hoist_diagonals(const std::vector<aqs_op>& ops) const {
        std::vector<long> last(n_, -1);       // last non-diagonal op on each qubit
        std::vector<std::vector<size_t>> after(ops.size() + 1);  // slot 0 = front
        for (size_t i = 0; i < ops.size(); ++i) {
            if (!is_diagonal(ops[i])) {
                for (int q : op_qubits(ops[i])) last[q] = (long)i;
                continue;
            }
            long a = -1;
            for (int q : op_qubits(ops[i])) a = std::max(a, last[q]);
            after[a + 1].push_back(i);
        }
        std::vector<aqs_op> out;
        out.reserve(ops.size());
        for (size_t j : after[0]) out.push_back(ops[j]);
        for (size_t i = 0; i < ops.size(); ++i) {
            if (is_diagonal(ops[i])) continue;
            out.push_back(ops[i]);
            for (size_t j : after[i + 1]) out.push_back(ops[j]);
        }
        return out;
    }

    std::vector<size_t> touching(const std::vector<int>& S) {
        std::vector<size_t> hit;
        for (size_t b = 0; b < open_.size(); ++b)
            for (int q : S)
                if (local_index(open_[b].qubits, q) >= 0) {
                    hit.push_back(b);
                    break;
                }
        return hit;
    }

    std::vector<int> union_with(const std::vector<int>& S, const std::vector<size_t>& hit) {
        std::vector<int> U = S;
        for (size_t b : hit)
            for (int q : open_[b].qubits)
                if (std::find(U.begin(), U.end(), q) == U.end()) U.push_back(q);
        std::sort(U.begin(), U.end());
        return U;
    }

    // The blocks `hit` and then op, as one block on U.  The open blocks act
    // on disjoint qubits, so together they are the Kronecker product of their
    // unitaries (identity on U's other qubits): entry (r, c) is the product
    // of each block's entry at r's and c's bits on its qubits.  op then
    // updates the rows, O(K^2) per 1-2 qubit gate, instead of dense K^3
    // products.
    void merge(const aqs_op& op, const std::vector<size_t>& hit, const std::vector<int>& U) {
        const int k = (int)U.size();
        const uint64_t K = 1ull << k;
        const aqs_op loc = localize(op, U);
        std::vector<cd> M;
        if (hit.size() == 1 && open_[hit[0]].qubits == U) {
            M = std::move(open_[hit[0]].U);
        } else {
            // sub-index of each U-local index on each block's qubits
            std::vector<std::vector<uint32_t>> sub(hit.size(), std::vector<uint32_t>(K));
            uint64_t covered = 0;
            for (size_t h = 0; h < hit.size(); ++h) {
                const Block& ob = open_[hit[h]];
                const int kb = (int)ob.qubits.size();
                for (int j = 0; j < kb; ++j) covered |= 1ull << (k - 1 - local_index(U, ob.qubits[j]));
                for (uint64_t r = 0; r < K; ++r) {
                    uint32_t v = 0;
                    for (int j = 0; j < kb; ++j)
                        if ((r >> (k - 1 - local_index(U, ob.qubits[j]))) & 1ull) v |= 1u << (kb - 1 - j);
                    sub[h][r] = v;
                }
            }
            M.assign(K * K, cd(0));
            for (uint64_t r = 0; r < K; ++r)
                for (uint64_t c = 0; c < K; ++c) {
                    if ((r ^ c) & ~covered) continue;  // identity on qubits no block holds
                    cd v = 1.0;
                    for (size_t h = 0; h < hit.size(); ++h) {
                        const Block& ob = open_[hit[h]];
                        v *= ob.U[((uint64_t)sub[h][r] << ob.qubits.size()) + sub[h][c]];
                    }
                    M[r * K + c] = v;
                }
        }
        apply_rows(M, k, loc);
        for (size_t j = hit.size(); j-- > 0;) open_.erase(open_.begin() + hit[j]);
        Block nb;
        nb.qubits = U;
        nb.U = std::move(M);
        open_.push_back(std::move(nb));
    }

    void close_block(size_t idx) {
        Block b = std::move(open_[idx]);
        open_.erase(open_.begin() + idx);
        const int k = (int)b.qubits.size();
        const uint64_t K = 1ull << k;
        aqs_op op{};
        if (k == 1) {
            op.ntargets = 1;
            op.targets[0] = b.qubits[0];
            op.kind = diagonal_matrix(b.U, 2) ? AQS_OP_DIAG1 : AQS_OP_MAT1;
            for (int i = 0; i < 4; ++i) {
                op.m[2 * i] = b.U[i].real();
                op.m[2 * i + 1] = b.U[i].imag();
            }
            push(op, {});
            return;
        }
        op.kind = AQS_OP_MATK;
        op.ntargets = k;
        for (int j = 0; j < k; ++j) op.targets[j] = b.qubits[j];
        std::vector<double> u(2 * K * K);
        for (uint64_t i = 0; i < K * K; ++i) {
            u[2 * i] = b.U[i].real();
            u[2 * i + 1] = b.U[i].imag();
        }
        push(op, std::move(u));
    }

    void push(const aqs_op& op, std::vector<double> mat) {
        sch_->push_plain(op, std::move(mat));
    }

    // Closes the open blocks touching S; returns the qubits of the block
    // emitted last (empty when none was).
    std::vector<int> close_touching(const std::vector<int>& S) {
        auto hit = touching(S);
        std::vector<int> last;
        for (size_t j = hit.size(); j-- > 0;) {
            last = open_[hit[j]].qubits;
            close_block(hit[j]);
        }
        return last;
    }

    // Which of the open blocks `hit` (touching S) to keep open so that op
    // can join them: the subset whose union with S is largest within kmax
    // (ties: more blocks kept).  Open blocks act on disjoint qubits, so
    // closing some of them first only reorders commuting gates.
    std::vector<size_t> keep_for(const std::vector<int>& S, const std::vector<size_t>& hit) {
        size_t best = 0;
        int best_size = -1, best_count = -1;
        for (size_t mask = 0; mask < ((size_t)1 << hit.size()); ++mask) {
            std::vector<size_t> sub;
            for (size_t j = 0; j < hit.size(); ++j)
                if ((mask >> j) & 1) sub.push_back(hit[j]);
            const int size = (int)union_with(S, sub).size();
            const int count = __builtin_popcountll(mask);
            if (size > kmax_) continue;
            if (size > best_size || (size == best_size && count > best_count)) {
                best = mask;
                best_size = size;
                best_count = count;
            }
        }
        std::vector<size_t> keep;
        for (size_t j = 0; j < hit.size(); ++j)
            if ((best >> j) & 1) keep.push_back(hit[j]);
        return keep;
    }

    // Close the blocks of `hit` outside `keep` and merge op into the rest
    // (keep not empty).
    void merge_keeping(const aqs_op& op, const std::vector<int>& S, const std::vector<size_t>& hit,
                       const std::vector<size_t>& keep) {
        for (size_t j = hit.size(); j-- > 0;)
            if (std::find(keep.begin(), keep.end(), hit[j]) == keep.end()) close_block(hit[j]);
        auto kept = touching(S);
        merge(op, kept, union_with(S, kept));
    }

    void nondiag(const aqs_op& op) {
        std::vector<int> S = op_qubits(op);
        // deferred phase terms on these qubits go first
        bool need = false;
        for (auto& t : terms_)
            for (int q : t.q)
                if (std::find(S.begin(), S.end(), q) != S.end()) need = true;
        if (need) flush_terms();
        auto hit = touching(S);
        if ((int)S.size() > kmax_ || op.more_controls) {
            close_touching(S);
            raw(op);
            return;
        }
        std::vector<int> U = union_with(S, hit);
        if ((int)U.size() <= kmax_ && (!hit.empty() || !is_permutation(op))) {
            merge(op, hit, U);
            return;
        }
        if (keep_partial_ && !hit.empty()) {
            const auto keep = keep_for(S, hit);
            if (!keep.empty()) {
                merge_keeping(op, S, hit, keep);
                return;
            }
        }
        close_touching(S);
        if (is_permutation(op)) {
            raw(op);
            return;
        }
        std::sort(S.begin(), S.end());
        Block nb;
        nb.qubits = S;
        nb.U = op_matrix(op, S);
        open_.push_back(std::move(nb));
    }

    void diag(const aqs_op& op) {
        std::vector<int> S = op_qubits(op);
        if (op.more_controls) {  // many-controlled: its own pass
            close_touching(S);
            raw(op);
            return;
        }
        auto hit = touching(S);
        std::vector<int> U = union_with(S, hit);
        if ((int)U.size() <= kmax_) {
            merge(op, hit, U);
            return;
        }
        // Defer only ops that would pull a qubit no open block holds yet into
        // a full block (the long controlled-phase ladders); an op between
        // qubits that are all in blocks closes them, as a non-diagonal op does.
        bool fresh = false;
        for (int q : S) {
            bool held = false;
            for (auto& b : open_)
                if (local_index(b.qubits, q) >= 0) held = true;
            if (!held) fresh = true;
        }
        if (!fresh || (int)S.size() > 5) {
            if (keep_partial_ && (int)S.size() <= kmax_) {
                const auto keep = keep_for(S, hit);
                if (!keep.empty()) {
                    merge_keeping(op, S, hit, keep);
                    return;
                }
            }
            close_touching(S);
            if ((int)S.size() > kmax_) {
                raw(op);
                return;
            }
            std::sort(S.begin(), S.end());
            Block nb;
            nb.qubits = S;
            nb.U = op_matrix(op, S);
            open_.push_back(std::move(nb));
            return;
        }
        // deferred as phase terms (DIAG1 = global d0 and d1/d0 on the target)
        std::vector<int> ctl;
        for (int j = 0; j < op.ncontrols; ++j) ctl.push_back(op.controls[j]);
        if (op.kind == AQS_OP_PHASE) {
            std::vector<int> q = ctl;
            for (int j = 0; j < op.ntargets; ++j) q.push_back(op.targets[j]);
            terms_.push_back({q, cd(op.m[0], op.m[1])});
        } else {
            const cd d0(op.m[0], op.m[1]), d1(op.m[6], op.m[7]);
            terms_.push_back({ctl, d0});
            std::vector<int> q = ctl;
            q.push_back(op.targets[0]);
            terms_.push_back({q, d1 / d0});
        }
    }

    void raw(const aqs_op& op) {
        std::vector<double> u;
        if (op.kind == AQS_OP_MATK) {
            const size_t K = (size_t)1 << op.ntargets;
            u.assign(op.u, op.u + 2 * K * K);
        }
        push(op, std::move(u));
    }

    // Emit the deferred terms: close the open blocks they meet, then table passes.
    void flush_terms() {
        if (terms_.empty()) return;
        std::vector<int> Q;
        for (auto& t : terms_)
            for (int q : t.q)
                if (std::find(Q.begin(), Q.end(), q) == Q.end()) Q.push_back(q);
        const std::vector<int> last = close_touching(Q);
        cd global = 1.0;
        std::vector<Term> left;
        for (auto& t : terms_) {
            if (t.q.empty())
                global *= t.phase;
            else
                left.push_back(t);
        }
        terms_.clear();
        while (!left.empty() || global != cd(1.0)) {
            // anchor group: the (up to) 5 qubits in the most terms; among
            // equals, those of the block emitted right before the table, so
            // that the table can ride that block's store (the epilogue form
            // of the MFMA block holds fewer registers than the prologue form
            // of the next one)
            std::map<int, int> deg;
            for (auto& t : left)
                for (int q : t.q) deg[q]++;
            std::vector<std::pair<int, int>> byd(deg.begin(), deg.end());
            auto in_last = [&](int q) { return std::find(last.begin(), last.end(), q) != last.end(); };
            std::stable_sort(byd.begin(), byd.end(), [&](auto& a, auto& b) {
                if (a.second != b.second) return a.second > b.second;
                return in_last(a.first) && !in_last(b.first);
            });
            std::vector<int> G;
            for (auto& p : byd)
                if ((int)G.size() < 5) G.push_back(p.first);
            std::vector<Term> covered, rest;
            std::vector<int> outside;
            for (auto& t : left) {
                std::vector<int> o;
                for (int q : t.q)
                    if (std::find(G.begin(), G.end(), q) == G.end()) o.push_back(q);
                bool ok = o.size() <= 1;
                if (ok && o.size() == 1 && std::find(outside.begin(), outside.end(), o[0]) == outside.end()) {
                    if (outside.size() >= 30)
                        ok = false;
                    else
                        outside.push_back(o[0]);
                }
                (ok ? covered : rest).push_back(t);
            }
            if (covered.empty() && global == cd(1.0)) {
                // cannot factorise: emit each remaining term as its own phase pass
                for (auto& t : rest) {
                    aqs_op op{};
                    op.kind = AQS_OP_PHASE;
                    op.ntargets = std::min<int>(5, (int)t.q.size());
                    for (int j = 0; j < op.ntargets; ++j) op.targets[j] = t.q[j];
                    for (size_t j = op.ntargets; j < t.q.size(); ++j) op.controls[op.ncontrols++] = t.q[j];
                    op.m[0] = t.phase.real();
                    op.m[1] = t.phase.imag();
                    push(op, {});
                }
                break;
            }
            emit_table(G, outside, covered, global);
            global = 1.0;
            left = std::move(rest);
        }
    }

    int pos(int q) const { return n_ - 1 - q; }

    void restore_order(const std::vector<int>& L) { emit_restore_order(sch_, L); }

    void emit_table(std::vector<int> G, std::vector<int> outside, const std::vector<Term>& terms, cd global) {
        // drop anchor qubits no covered term uses
        std::vector<int> used;
        for (int g : G)
            for (auto& t : terms)
                if (std::find(t.q.begin(), t.q.end(), g) != t.q.end()) {
                    used.push_back(g);
                    break;
                }
        G = used;
        // ascending bit positions let the kernel read the groups with one shift
        std::sort(G.begin(), G.end(), [&](int a, int b) { return pos(a) < pos(b); });
        // required qubits: in every term
        std::vector<int> R;
        if (global == cd(1.0) && !terms.empty())
            for (int q : terms[0].q) {
                bool all = true;
                for (auto& t : terms)
                    if (std::find(t.q.begin(), t.q.end(), q) == t.q.end()) all = false;
                if (all) R.push_back(q);
            }
        std::sort(outside.begin(), outside.end(), [&](int a, int b) { return pos(a) < pos(b); });
        DiagPass d;
        for (int q : R) d.required.push_back(pos(q));
        for (int q : G) d.anchor.push_back(pos(q));
        std::vector<std::vector<int>> chunk_q;
        for (size_t j = 0; j < outside.size(); j += 5) {
            std::vector<int> c(outside.begin() + j, outside.begin() + std::min(outside.size(), j + 5));
            chunk_q.push_back(c);
        }
        if (chunk_q.empty()) chunk_q.push_back({});
        for (auto& c : chunk_q) {
            std::vector<int> p;
            for (int q : c) p.push_back(pos(q));
            d.chunks.push_back(p);
        }
        const size_t ng = (size_t)1 << G.size(), nc = chunk_q.size();
        d.table.assign(ng * nc * 32, cd(1.0));
        for (size_t g = 0; g < ng; ++g)
            for (size_t v = 0; v < 32; ++v) d.table[(g * nc + 0) * 32 + v] *= global;
        for (auto& t : terms) {
            uint32_t gmask = 0;
            int oq = -1;
            for (int q : t.q) {
                int gi = local_index(G, q);
                if (gi >= 0)
                    gmask |= 1u << gi;
                else
                    oq = q;
            }
            size_t c = 0;
            uint32_t vbit = 0;
            if (oq >= 0) {
                for (size_t ci = 0; ci < nc; ++ci) {
                    int b = local_index(chunk_q[ci], oq);
                    if (b >= 0) {
                        c = ci;
                        vbit = 1u << b;
                    }
                }
            }
            for (uint32_t g = 0; g < ng; ++g) {
                if ((g & gmask) != gmask) continue;
                for (uint32_t v = 0; v < 32; ++v)
                    if ((v & vbit) == vbit) d.table[(g * nc + c) * 32 + v] *= t.phase;
            }
        }
        aqs_op op{};
        op.kind = AQS_OP_DIAGT;
        op.ntargets = std::max(1, (int)G.size());
        for (size_t j = 0; j < G.size(); ++j) op.targets[j] = G[j];
        if (G.empty()) op.targets[0] = outside.empty() ? 0 : outside[0];
        sch_->push_plain(op);
        sch_->diag_index.back() = (int)sch_->diags.size();
        sch_->diags.push_back(std::move(d));
    }
};

}  // namespace

namespace aqs_hip {

// Logical qubit q sits at physical qubit L[q]: move index bit pos(L[q]) to
// pos(q).  beta = tau2 o tau1 with both involutions (per cycle
// x_0 -> x_1 -> ..: tau1(x_i) = x_{-i}, tau2(x_i) = x_{1-i}).
void emit_restore_order(aqs_schedule_s* sch, const std::vector<int>& L) {
    const int n = sch->n;
    auto pos = [n](int q) { return n - 1 - q; };
    std::vector<int> beta(n);
    bool identity = true;
    for (int q = 0; q < n; ++q) {
        beta[pos(L[q])] = pos(q);
        if (L[q] != q) identity = false;
    }
    if (identity) return;
    std::vector<int> t1(n), t2(n);
    std::vector<char> seen(n, 0);
    for (int s = 0; s < n; ++s) {
        if (seen[s]) continue;
        std::vector<int> cyc;
        for (int x = s; !seen[x]; x = beta[x]) {
            seen[x] = 1;
            cyc.push_back(x);
        }
        const int k = (int)cyc.size();
        for (int j = 0; j < k; ++j) {
            t1[cyc[j]] = cyc[((-j) % k + k) % k];
            t2[cyc[j]] = cyc[((1 - j) % k + k) % k];
        }
    }
    for (const auto* t : {&t1, &t2}) {
        bool id = true;
        for (int b = 0; b < n; ++b)
            if ((*t)[b] != b) id = false;
        if (id) continue;
        aqs_op op{};
        op.kind = AQS_OP_PERMB;
        op.ntargets = 0;
        for (int b = 0; b < n && op.ntargets < 5; ++b)
            if ((*t)[b] != b) op.targets[op.ntargets++] = n - 1 - b;
        sch->push_plain(op);
        sch->perm_index.push_back({(int)sch->passes.size() - 1, (int)sch->perms.size()});
        sch->perms.push_back(*t);
    }
}

}  // namespace aqs_hip

namespace {

// The bit involution of pass p when p is a bit-permutation pass made of
// exactly one involution, else null.
const std::vector<int>* single_involution(const aqs_schedule_s* sch, size_t p) {
    if (p >= sch->passes.size() || sch->passes[p].kind != AQS_OP_PERMB) return nullptr;
    const std::vector<int>* beta = nullptr;
    int count = 0;
    for (const auto& pi : sch->perm_index)
        if (pi.first == (int)p) {
            beta = &sch->perms[pi.second];
            ++count;
        }
    return count == 1 ? beta : nullptr;
}

}  // namespace

extern "C" {

// Op lists at least this long fuse their variants on parallel threads.
constexpr size_t kParallelFuseOps = 256;

aqs_status aqs_fuse(const aqs_op* ops, size_t count, int nqubits, int max_qubits, aqs_schedule* out) {
    return guard([&] {
        if (!out) fail(AQS_ERR_INVALID_ARGUMENT, "null output schedule");
        *out = nullptr;
        if (count && !ops) fail(AQS_ERR_INVALID_ARGUMENT, "null op array");
        if (nqubits < 1 || nqubits > AQS_MAX_QUBITS) fail(AQS_ERR_INVALID_ARGUMENT, "bad qubit count");
        if (max_qubits < 1 || max_qubits > 5) fail(AQS_ERR_INVALID_ARGUMENT, "max_qubits must be 1..5");
        for (size_t i = 0; i < count; ++i) validate_op(ops[i], nqubits);
        // Four independent fusions (closing policy x program order), built on
        // their own threads for long op lists; the cheapest by the cost model
        // wins, the first of equals in variant order.
        constexpr int kVariants = 4;
        std::unique_ptr<aqs_schedule_s> sch[kVariants];
        std::exception_ptr err[kVariants];
        auto build = [&](int variant) {
            try {
                auto s = std::make_unique<aqs_schedule_s>();
                s->n = nqubits;
                s->source_ops = count;
                Fuser f(s.get(), nqubits, max_qubits, variant >> 1);
                f.run(ops, count, variant & 1);
                for (size_t p = 0; p < s->passes.size(); ++p)
                    if (s->passes[p].kind == AQS_OP_MATK) s->passes[p].u = s->mats[p].data();
                sch[variant] = std::move(s);
            } catch (...) {
                err[variant] = std::current_exception();
            }
        };
        std::vector<std::thread> threads;
        int started = 1;  // variants 1 .. started-1 run on their own threads
        if (count >= kParallelFuseOps) {
            try {
                for (int v = 1; v < kVariants; ++v) {
                    threads.emplace_back(build, v);
                    started = v + 1;
                }
            } catch (...) {
                // no thread to be had: the rest run here
            }
        }
        build(0);
        for (int v = started; v < kVariants; ++v) build(v);
        for (auto& t : threads) t.join();
        for (int v = 0; v < kVariants; ++v)
            if (err[v]) std::rethrow_exception(err[v]);
        int best = 0;
        double best_cost = schedule_cost(sch[0].get());
        for (int v = 1; v < kVariants; ++v) {
            const double c = schedule_cost(sch[v].get());
            if (c < best_cost) {
                best = v;
                best_cost = c;
            }
        }
        *out = sch[best].release();
    });
}

aqs_status aqs_schedule_destroy(aqs_schedule sch) {
    return guard([&] {
        if (!sch) return;
        if (sch->dev) (void)hipFree(sch->dev);
        delete sch;
    });
}

aqs_status aqs_schedule_info(aqs_schedule sch, size_t* passes, size_t* source_ops) {
    return guard([&] {
        if (!sch) fail(AQS_ERR_INVALID_ARGUMENT, "null schedule");
        if (passes) *passes = sch->passes.size();
        if (source_ops) *source_ops = sch->source_ops;
    });
}

aqs_status aqs_schedule_get(aqs_schedule sch, size_t i, aqs_op* op) {
    return guard([&] {
        if (!sch) fail(AQS_ERR_INVALID_ARGUMENT, "null schedule");
        if (i >= sch->passes.size()) fail(AQS_ERR_OUT_OF_RANGE, "pass index out of range");
        if (op) *op = sch->passes[i];
    });
}

aqs_status aqs_apply_schedule(aqs_state s, aqs_schedule sch) {
    return guard([&] {
        if (!s || !s->amps) fail(AQS_ERR_INVALID_ARGUMENT, "null state handle");
        if (!sch) fail(AQS_ERR_INVALID_ARGUMENT, "null schedule");
        if (sch->n != s->n) fail(AQS_ERR_INVALID_ARGUMENT, "schedule and state qubit counts differ");
        if (sch->dev_precision != s->precision) {
            // Device copies of the pass data, in the layout each kernel reads:
            // MFMA-packed FP32 (complex64, k = 4, 5), plain FP32 / FP64
            // matrices, and diagonal tables as float2 / double2.
            if (sch->dev) {
                AQS_HIP(hipDeviceSynchronize());
                AQS_HIP(hipFree(sch->dev));
                sch->dev = nullptr;
            }
            std::vector<char> host;
            sch->dev_off.assign(sch->passes.size(), 0);
            sch->dev_tile.assign(sch->passes.size(), {});
            sch->dev_mfma.assign(sch->passes.size(), 0);
            auto append = [&](const void* p, size_t bytes) {
                size_t at = (host.size() + 255) & ~(size_t)255;
                host.resize(at + bytes);
                if (bytes) std::memcpy(host.data() + at, p, bytes);
                return at;
            };
            for (size_t p = 0; p < sch->passes.size(); ++p) {
                const aqs_op& op = sch->passes[p];
                if (sch->tile_index[p] >= 0) {
                    const auto& t = sch->tiles[sch->tile_index[p]];
                    auto& o = sch->dev_tile[p];
                    o[0] = append(t.segs.data(), t.segs.size() * sizeof(TileSegDev));
                    o[1] = append(t.ops.data(), t.ops.size() * sizeof(TileOpDev));
                    o[2] = append(t.terms.data(), t.terms.size() * sizeof(TileTermDev));
                    o[5] = append(t.folds.data(), t.folds.size() * sizeof(TileFoldDev));
                    o[6] = append(t.fold_terms.data(), t.fold_terms.size() * sizeof(TileFoldTermDev));
                    if (s->precision == AQS_C128) {
                        o[3] = append(t.m128.data(), t.m128.size() * sizeof(double));
                        o[4] = append(t.z128.data(), t.z128.size() * sizeof(double));
                        o[7] = append(t.fold_z128.data(), t.fold_z128.size() * sizeof(double));
                    }
                    continue;
                }
                if (sch->diag_index[p] >= 0) {
                    const auto& d = sch->diags[sch->diag_index[p]];
                    if (s->precision == AQS_C64) {
                        std::vector<float> t(2 * d.table.size());
                        for (size_t i = 0; i < d.table.size(); ++i) {
                            t[2 * i] = (float)d.table[i].real();
                            t[2 * i + 1] = (float)d.table[i].imag();
                        }
                        sch->dev_off[p] = append(t.data(), t.size() * sizeof(float));
                    } else {
                        sch->dev_off[p] = append(d.table.data(), d.table.size() * sizeof(cd));
                    }
                    continue;
                }
                if (op.kind != AQS_OP_MATK) continue;
                const size_t K = (size_t)1 << op.ntargets;
                if (s->precision == AQS_C64) {
                    std::vector<float> v;
                    if ((op.ntargets == 4 || op.ntargets == 5) && op.ncontrols == 0) {
                        v = pack_mfma_operand(op.ntargets, op.u);
                        sch->dev_mfma[p] = 1;
                    } else {
                        v.resize(2 * K * K);
                        for (size_t i = 0; i < v.size(); ++i) v[i] = (float)op.u[i];
                    }
                    sch->dev_off[p] = append(v.data(), v.size() * sizeof(float));
                } else {
                    sch->dev_off[p] = append(op.u, 2 * K * K * sizeof(double));
                }
            }
            if (!host.empty()) {
                AQS_HIP(hipMalloc(&sch->dev, host.size()));
                AQS_HIP(hipMemcpy(sch->dev, host.data(), host.size(), hipMemcpyHostToDevice));
            }
            sch->dev_precision = s->precision;
        }
        for (size_t p = 0; p < sch->passes.size(); ++p) {
            const aqs_op& op = sch->passes[p];
            const void* d = (const char*)sch->dev + sch->dev_off[p];
            if (sch->tile_index[p] >= 0) {
                const auto& o = sch->dev_tile[p];
                const char* base = (const char*)sch->dev;
                apply_tile_pass(s, sch->tiles[sch->tile_index[p]],
                                TileDevProgram{base + o[0], base + o[1], base + o[2], base + o[3], base + o[4],
                                               base + o[5], base + o[6], base + o[7]});
                continue;
            }
            if (sch->diag_index[p] >= 0) {
                const auto& dp = sch->diags[sch->diag_index[p]];
                // or in the prologue of the next MFMA block when its anchor
                // lies in that block's tile (with a restore permutation after
                // the block folded into its store when the tile allows)
                if (p + 1 < sch->passes.size() && sch->dev_mfma[p + 1] && sch->passes[p + 1].kind == AQS_OP_MATK &&
                    sch->tile_index[p + 1] < 0 && sch->diag_index[p + 1] < 0) {
                    DiagEpilogueDesc epi;
                    epi.required = dp.required;
                    epi.anchor = dp.anchor;
                    epi.chunks = dp.chunks;
                    epi.dev_tab = d;
                    const float* apack = (const float*)((const char*)sch->dev + sch->dev_off[p + 1]);
                    const std::vector<int>* beta = single_involution(sch, p + 2);
                    if (beta && apply_matk_mfma(s, sch->passes[p + 1], apack, &epi, true, beta)) {
                        p += 2;
                        continue;
                    }
                    if (apply_matk_mfma(s, sch->passes[p + 1], apack, &epi, true)) {
                        ++p;
                        continue;
                    }
                }
                apply_diag_table(s, dp.required, dp.anchor, dp.chunks, d);
                continue;
            }
            if (op.kind == AQS_OP_PERMB) {
                for (auto& pi : sch->perm_index)
                    if (pi.first == (int)p) apply_bit_involution(s, sch->perms[pi.second]);
                continue;
            }
            if (op.kind != AQS_OP_MATK) {
                apply_op(s, op);
                continue;
            }
            if (sch->dev_mfma[p]) {
                // a diagonal table pass right after the block rides in its
                // epilogue when its anchor lies in the block's tile
                if (p + 1 < sch->passes.size() && sch->diag_index[p + 1] >= 0) {
                    const auto& dp = sch->diags[sch->diag_index[p + 1]];
                    DiagEpilogueDesc epi;
                    epi.required = dp.required;
                    epi.anchor = dp.anchor;
                    epi.chunks = dp.chunks;
                    epi.dev_tab = (const char*)sch->dev + sch->dev_off[p + 1];
                    if (apply_matk_mfma(s, op, (const float*)d, &epi)) {
                        ++p;
                        continue;
                    }
                }
                // a restore permutation right after the block: folded into
                // its store when it maps the block's tile onto itself
                if (const std::vector<int>* beta = single_involution(sch, p + 1)) {
                    if (apply_matk_mfma(s, op, (const float*)d, nullptr, false, beta)) {
                        ++p;
                        continue;
                    }
                }
                if (apply_matk_mfma(s, op, (const float*)d)) continue;
                apply_op(s, op);  // MFMA shape not usable at this size: plain matrix path
                continue;
            }
            validate_op(op, s->n);
            apply_matk_device(s, op, d);
        }
    });
}

}  // extern "C"
