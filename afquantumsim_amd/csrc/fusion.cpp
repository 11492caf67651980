// compile(): gate fusion into dense k-qubit blocks (host side, FP64).
//
// The reference's QCircuit::compile (src/quantum.cpp:199-210) multiplies
// every gate into one 2^n x 2^n matrix.  This pass instead groups the gate
// list into blocks of at most max_qubits (<= 5) qubits:
//
//   - open blocks have disjoint qubit sets, so they commute and can grow in
//     any order;
//   - an op whose qubits meet open blocks is merged with them when the union
//     fits in max_qubits, otherwise those blocks are closed (emitted) first
//     and the op starts a new block;
//   - an op wider than max_qubits is emitted unfused.
//
// Each block's unitary is accumulated in FP64 (op matrices are built by
// applying the op to the block's basis vectors) and rounded to FP32 once,
// when the schedule is uploaded for a complex64 state.  One-qubit blocks are
// emitted as MAT1 / DIAG1 passes, wider blocks as MATK passes (MFMA for
// 4 and 5 qubits on complex64 states).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

#include "internal.h"

using namespace aqs_hip;
using cd = std::complex<double>;

struct aqs_schedule_s {
    int n = 0;
    size_t source_ops = 0;
    std::vector<aqs_op> passes;
    std::vector<std::vector<double>> mats;  // owned FP64 matrices (MATK passes)
    // device copies, built at first use for a given precision
    void* dev = nullptr;
    int dev_precision = -1;
    std::vector<size_t> dev_off;   // byte offset of each pass's matrix
    std::vector<char> dev_mfma;    // pass uses the MFMA packed layout
};

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

std::vector<int> op_qubits(const aqs_op& op) {
    std::vector<int> q;
    for (int j = 0; j < op.ntargets; ++j) q.push_back(op.targets[j]);
    for (int j = 0; j < op.ncontrols; ++j)
        if (std::find(q.begin(), q.end(), op.controls[j]) == q.end()) q.push_back(op.controls[j]);
    return q;
}

struct Block {
    std::vector<int> qubits;  // local qubit j = bit k-1-j of the block index
    std::vector<cd> U;        // row-major 2^k x 2^k
    size_t sources = 0;
};

int local_index(const std::vector<int>& qs, int q) {
    for (size_t j = 0; j < qs.size(); ++j)
        if (qs[j] == q) return (int)j;
    return -1;
}

// The op's matrix over the qubit list `qs` (all op qubits are in qs).
std::vector<cd> op_matrix(const aqs_op& op, const std::vector<int>& qs) {
    const int k = (int)qs.size();
    const uint64_t K = 1ull << k;
    aqs_op loc = op;
    for (int j = 0; j < op.ntargets; ++j) loc.targets[j] = local_index(qs, op.targets[j]);
    for (int j = 0; j < op.ncontrols; ++j) loc.controls[j] = local_index(qs, op.controls[j]);
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

// Re-express a block's matrix over a larger qubit list (block qubits subset).
std::vector<cd> embed(const Block& b, const std::vector<int>& qs) {
    aqs_op op{};
    op.kind = AQS_OP_MATK;
    op.ntargets = (int)b.qubits.size();
    for (int j = 0; j < op.ntargets; ++j) op.targets[j] = b.qubits[j];
    std::vector<double> u(2 * b.U.size());
    for (size_t i = 0; i < b.U.size(); ++i) {
        u[2 * i] = b.U[i].real();
        u[2 * i + 1] = b.U[i].imag();
    }
    op.u = u.data();
    return op_matrix(op, qs);
}

std::vector<cd> matmul(const std::vector<cd>& A, const std::vector<cd>& B, uint64_t K) {
    std::vector<cd> C(K * K);
    for (uint64_t i = 0; i < K; ++i)
        for (uint64_t l = 0; l < K; ++l) {
            const cd a = A[i * K + l];
            if (a == cd(0)) continue;
            for (uint64_t j = 0; j < K; ++j) C[i * K + j] += a * B[l * K + j];
        }
    return C;
}

bool is_diagonal(const std::vector<cd>& U, uint64_t K) {
    for (uint64_t i = 0; i < K; ++i)
        for (uint64_t j = 0; j < K; ++j)
            if (i != j && U[i * K + j] != cd(0)) return false;
    return true;
}

void emit_block(aqs_schedule_s* sch, Block&& b) {
    const int k = (int)b.qubits.size();
    const uint64_t K = 1ull << k;
    aqs_op op{};
    if (k == 1) {
        op.ntargets = 1;
        op.targets[0] = b.qubits[0];
        op.kind = is_diagonal(b.U, 2) ? AQS_OP_DIAG1 : AQS_OP_MAT1;
        for (int i = 0; i < 4; ++i) {
            op.m[2 * i] = b.U[i].real();
            op.m[2 * i + 1] = b.U[i].imag();
        }
        sch->passes.push_back(op);
        sch->mats.emplace_back();
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
    sch->passes.push_back(op);
    sch->mats.push_back(std::move(u));
}

}  // namespace

extern "C" {

aqs_status aqs_fuse(const aqs_op* ops, size_t count, int nqubits, int max_qubits, aqs_schedule* out) {
    return guard([&] {
        if (!out) fail(AQS_ERR_INVALID_ARGUMENT, "null output schedule");
        *out = nullptr;
        if (count && !ops) fail(AQS_ERR_INVALID_ARGUMENT, "null op array");
        if (nqubits < 1 || nqubits > AQS_MAX_QUBITS) fail(AQS_ERR_INVALID_ARGUMENT, "bad qubit count");
        if (max_qubits < 1 || max_qubits > 5) fail(AQS_ERR_INVALID_ARGUMENT, "max_qubits must be 1..5");
        auto* sch = new aqs_schedule_s();
        sch->n = nqubits;
        sch->source_ops = count;
        try {
            std::vector<Block> open;
            auto close_block = [&](size_t idx) {
                emit_block(sch, std::move(open[idx]));
                open.erase(open.begin() + idx);
            };
            for (size_t i = 0; i < count; ++i) {
                const aqs_op& op = ops[i];
                validate_op(op, nqubits);
                std::vector<int> S = op_qubits(op);
                std::vector<size_t> hit;
                for (size_t b = 0; b < open.size(); ++b)
                    for (int q : S)
                        if (local_index(open[b].qubits, q) >= 0) {
                            hit.push_back(b);
                            break;
                        }
                std::vector<int> U = S;
                for (size_t b : hit)
                    for (int q : open[b].qubits)
                        if (std::find(U.begin(), U.end(), q) == U.end()) U.push_back(q);
                if ((int)S.size() > max_qubits) {
                    // too wide to fuse: close what it touches, emit as is
                    for (size_t j = hit.size(); j-- > 0;) close_block(hit[j]);
                    aqs_op cp = op;
                    std::vector<double> u;
                    if (op.kind == AQS_OP_MATK) {
                        const size_t K = (size_t)1 << op.ntargets;
                        u.assign(op.u, op.u + 2 * K * K);
                    }
                    sch->passes.push_back(cp);
                    sch->mats.push_back(std::move(u));
                    continue;
                }
                if ((int)U.size() <= max_qubits) {
                    // merge op and every touched block into one block over U
                    std::sort(U.begin(), U.end());
                    const uint64_t K = 1ull << U.size();
                    std::vector<cd> M(K * K);
                    for (uint64_t d = 0; d < K; ++d) M[d * K + d] = 1.0;
                    size_t src = 1;
                    for (size_t b : hit) {
                        M = matmul(embed(open[b], U), M, K);
                        src += open[b].sources;
                    }
                    M = matmul(op_matrix(op, U), M, K);
                    for (size_t j = hit.size(); j-- > 0;) open.erase(open.begin() + hit[j]);
                    Block nb;
                    nb.qubits = U;
                    nb.U = std::move(M);
                    nb.sources = src;
                    open.push_back(std::move(nb));
                } else {
                    for (size_t j = hit.size(); j-- > 0;) close_block(hit[j]);
                    std::sort(S.begin(), S.end());
                    Block nb;
                    nb.qubits = S;
                    nb.U = op_matrix(op, S);
                    nb.sources = 1;
                    open.push_back(std::move(nb));
                }
            }
            while (!open.empty()) close_block(0);
            // MATK passes point at their owned matrices
            for (size_t p = 0; p < sch->passes.size(); ++p)
                if (sch->passes[p].kind == AQS_OP_MATK) sch->passes[p].u = sch->mats[p].data();
        } catch (...) {
            delete sch;
            throw;
        }
        *out = sch;
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
            // Build the device copy of every MATK matrix in the layout its
            // kernel reads: MFMA-packed FP32 (complex64, k = 4, 5), plain FP32
            // (complex64, k <= 3) or plain FP64 (complex128).
            if (sch->dev) {
                AQS_HIP(hipDeviceSynchronize());
                AQS_HIP(hipFree(sch->dev));
                sch->dev = nullptr;
            }
            std::vector<char> host;
            sch->dev_off.assign(sch->passes.size(), 0);
            sch->dev_mfma.assign(sch->passes.size(), 0);
            for (size_t p = 0; p < sch->passes.size(); ++p) {
                const aqs_op& op = sch->passes[p];
                if (op.kind != AQS_OP_MATK) continue;
                const size_t K = (size_t)1 << op.ntargets;
                size_t at = (host.size() + 255) & ~(size_t)255;
                host.resize(at);
                sch->dev_off[p] = at;
                if (s->precision == AQS_C64) {
                    std::vector<float> v;
                    if ((op.ntargets == 4 || op.ntargets == 5) && op.ncontrols == 0) {
                        v = pack_mfma_operand(op.ntargets, op.u);
                        sch->dev_mfma[p] = 1;
                    } else {
                        v.resize(2 * K * K);
                        for (size_t i = 0; i < v.size(); ++i) v[i] = (float)op.u[i];
                    }
                    host.resize(at + v.size() * sizeof(float));
                    std::memcpy(host.data() + at, v.data(), v.size() * sizeof(float));
                } else {
                    host.resize(at + 2 * K * K * sizeof(double));
                    std::memcpy(host.data() + at, op.u, 2 * K * K * sizeof(double));
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
            if (op.kind != AQS_OP_MATK) {
                apply_op(s, op);
                continue;
            }
            const void* d = (const char*)sch->dev + sch->dev_off[p];
            if (sch->dev_mfma[p] && apply_matk_mfma(s, op, (const float*)d)) continue;
            if (sch->dev_mfma[p]) {
                // MFMA shape not usable for this state size: plain matrix path
                aqs_op cp = op;
                apply_op(s, cp);
                continue;
            }
            validate_op(op, s->n);
            apply_matk_device(s, op, d);
        }
    });
}

}  // extern "C"
