// Batched simulation of many small, independent states in one persistent
// launch (variational ansatz workload: thousands of 10-16 qubit circuits,
// each followed by <sum Z> in FP64).
//
// The reference simulates one state at a time with a 2^n x 2^n operator per
// gate (src/quantum.cpp:277-291).  Here every workgroup of a persistent grid
// takes the next state from an atomic queue (states sorted by descending
// cost), builds |0..0> in LDS when the state fits the LDS tile (2^lds_qubits
// amplitudes) or in its HBM slot otherwise (then it stays L2-resident), runs
// all of that state's gate passes with a workgroup barrier between passes,
// reduces <sum_q Z_q> in FP64 and writes the final amplitudes back to HBM.
// The launch (queue reset + persistent kernel) is captured once into a
// hipGraph and replayed.  An autotuner times (workgroup size, LDS tile)
// candidates and keeps the fastest.
//
// Per-state host preprocessing absorbs runs of uncontrolled one-qubit gates
// on a qubit into one 2x2 matrix (FP64 products) until a multi-qubit gate
// touches that qubit, so an RY-RZ layer costs one pass per qubit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <numeric>
#include <vector>

#include "internal.h"

using namespace aqs_hip;

namespace {

struct BOp {
    int32_t kind;    // AQS_OP_MAT1 / DIAG1 / PHASE / X / SWAP
    int32_t t0, t1;  // target bit positions (t1 for SWAP)
    int32_t any;     // AQS_CTRL_ANY
    uint32_t cmask;  // control bits (state index space)
    uint32_t pmask;  // PHASE: all bits that must be set
    float m[8];
};

struct Desc {
    uint64_t amp_offset;  // into the HBM pool (amplitudes)
    uint32_t op_begin, op_count;
    int32_t n;
    int32_t index;        // position in the caller's order
};

__device__ __forceinline__ uint32_t ins0(uint32_t p, int t) { return ((p >> t) << (t + 1)) | (p & ((1u << t) - 1u)); }

__device__ __forceinline__ bool ctrl_ok(uint32_t i, uint32_t cmask, int any) {
    if (!cmask) return true;
    return any ? (i & cmask) != 0 : (i & cmask) == cmask;
}

template <int BLOCK>
__device__ void apply_block(float2* sv, uint32_t N, const BOp& op) {
    const int tid = threadIdx.x;
    switch (op.kind) {
        case AQS_OP_MAT1: {
            for (uint32_t p = tid; p < N / 2; p += BLOCK) {
                const uint32_t i0 = ins0(p, op.t0), i1 = i0 | (1u << op.t0);
                if (!ctrl_ok(i0, op.cmask, op.any)) continue;
                const float2 x = sv[i0], y = sv[i1];
                float2 a, b;
                a.x = op.m[0] * x.x - op.m[1] * x.y + op.m[2] * y.x - op.m[3] * y.y;
                a.y = op.m[0] * x.y + op.m[1] * x.x + op.m[2] * y.y + op.m[3] * y.x;
                b.x = op.m[4] * x.x - op.m[5] * x.y + op.m[6] * y.x - op.m[7] * y.y;
                b.y = op.m[4] * x.y + op.m[5] * x.x + op.m[6] * y.y + op.m[7] * y.x;
                sv[i0] = a;
                sv[i1] = b;
            }
            break;
        }
        case AQS_OP_DIAG1: {
            for (uint32_t i = tid; i < N; i += BLOCK) {
                if (!ctrl_ok(i, op.cmask, op.any)) continue;
                const bool one = (i >> op.t0) & 1u;
                const float dr = one ? op.m[6] : op.m[0], di = one ? op.m[7] : op.m[1];
                const float2 x = sv[i];
                sv[i] = make_float2(dr * x.x - di * x.y, dr * x.y + di * x.x);
            }
            break;
        }
        case AQS_OP_PHASE: {
            for (uint32_t i = tid; i < N; i += BLOCK) {
                if ((i & op.pmask) != op.pmask) continue;
                const float2 x = sv[i];
                sv[i] = make_float2(op.m[0] * x.x - op.m[1] * x.y, op.m[0] * x.y + op.m[1] * x.x);
            }
            break;
        }
        case AQS_OP_X: {
            for (uint32_t p = tid; p < N / 2; p += BLOCK) {
                const uint32_t i0 = ins0(p, op.t0), i1 = i0 | (1u << op.t0);
                if (!ctrl_ok(i0, op.cmask, op.any)) continue;
                const float2 x = sv[i0];
                sv[i0] = sv[i1];
                sv[i1] = x;
            }
            break;
        }
        case AQS_OP_SWAP: {
            const uint32_t ma = 1u << op.t0, mb = 1u << op.t1;
            for (uint32_t i = tid; i < N; i += BLOCK) {
                if (!(i & ma) || (i & mb) || !ctrl_ok(i, op.cmask, op.any)) continue;
                const uint32_t j = i ^ ma ^ mb;
                const float2 x = sv[i];
                sv[i] = sv[j];
                sv[j] = x;
            }
            break;
        }
        default:
            break;
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch(const Desc* __restrict__ desc, int count, const BOp* __restrict__ ops,
                                                  float2* __restrict__ pool, double* __restrict__ out,
                                                  unsigned* __restrict__ queue, int lds_qubits) {
    extern __shared__ float2 lds[];
    __shared__ int s_idx;
    __shared__ double s_red[BLOCK / 64];
    const int tid = threadIdx.x;
    while (true) {
        if (tid == 0) s_idx = (int)atomicAdd(queue, 1u);
        __syncthreads();
        const int k = s_idx;
        __syncthreads();
        if (k >= count) break;  // every wave of the workgroup leaves together
        const Desc d = desc[k];
        const uint32_t N = 1u << d.n;
        float2* hbm = pool + d.amp_offset;
        float2* sv = (d.n <= lds_qubits) ? lds : hbm;
        for (uint32_t i = tid; i < N; i += BLOCK) sv[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
        __syncthreads();
        for (uint32_t o = 0; o < d.op_count; ++o) {
            apply_block<BLOCK>(sv, N, ops[d.op_begin + o]);
            __syncthreads();
        }
        // <sum Z> in FP64, and the final amplitudes to HBM
        double acc = 0.0;
        for (uint32_t i = tid; i < N; i += BLOCK) {
            const float2 v = sv[i];
            const double p = (double)v.x * v.x + (double)v.y * v.y;
            acc += p * (double)(d.n - 2 * (int)__popc(i));
            if (sv != hbm) hbm[i] = v;
        }
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((tid & 63) == 0) s_red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < BLOCK / 64; ++w) s += s_red[w];
            out[d.index] = s;
        }
        __syncthreads();
    }
}

__global__ void k_reset_queue(unsigned* q) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *q = 0u;
}

using cd = std::complex<double>;

struct M2 {
    cd a[4];
};

M2 mat_of(const aqs_op& op) {
    M2 m;
    if (op.kind == AQS_OP_DIAG1) {
        m.a[0] = cd(op.m[0], op.m[1]);
        m.a[1] = 0;
        m.a[2] = 0;
        m.a[3] = cd(op.m[6], op.m[7]);
    } else {
        for (int i = 0; i < 4; ++i) m.a[i] = cd(op.m[2 * i], op.m[2 * i + 1]);
    }
    return m;
}

M2 mul(const M2& x, const M2& y) {  // x * y
    M2 r;
    r.a[0] = x.a[0] * y.a[0] + x.a[1] * y.a[2];
    r.a[1] = x.a[0] * y.a[1] + x.a[1] * y.a[3];
    r.a[2] = x.a[2] * y.a[0] + x.a[3] * y.a[2];
    r.a[3] = x.a[2] * y.a[1] + x.a[3] * y.a[3];
    return r;
}

}  // namespace

struct aqs_batch_s {
    int count = 0;
    std::vector<int> nq;
    std::vector<Desc> desc;  // sorted by descending cost
    std::vector<BOp> ops;
    uint64_t total_amps = 0;
    size_t source_ops = 0;
    void* d_desc = nullptr;
    void* d_ops = nullptr;
    float2* d_pool = nullptr;
    double* d_out = nullptr;
    unsigned* d_queue = nullptr;
    hipStream_t stream = nullptr;
    int block = 512;
    int lds_qubits = 13;
    int grid = 0;
    bool tuned = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
};

namespace {

void destroy_graph(aqs_batch b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void launch_kernel(aqs_batch b, int block, int lds_qubits, int grid) {
    const size_t lds_bytes = sizeof(float2) << lds_qubits;
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, b->stream, b->d_queue);
    switch (block) {
        case 256:
            hipLaunchKernelGGL(k_batch<256>, dim3(grid), dim3(256), lds_bytes, b->stream, (const Desc*)b->d_desc, b->count,
                               (const BOp*)b->d_ops, b->d_pool, b->d_out, b->d_queue, lds_qubits);
            break;
        case 512:
            hipLaunchKernelGGL(k_batch<512>, dim3(grid), dim3(512), lds_bytes, b->stream, (const Desc*)b->d_desc, b->count,
                               (const BOp*)b->d_ops, b->d_pool, b->d_out, b->d_queue, lds_qubits);
            break;
        default:
            hipLaunchKernelGGL(k_batch<1024>, dim3(grid), dim3(1024), lds_bytes, b->stream, (const Desc*)b->d_desc,
                               b->count, (const BOp*)b->d_ops, b->d_pool, b->d_out, b->d_queue, lds_qubits);
            break;
    }
    AQS_HIP(hipGetLastError());
}

int grid_for(int block, int lds_qubits) {
    int dev = current_device();
    hipDeviceProp_t prop;
    AQS_HIP(hipGetDeviceProperties(&prop, dev));
    const size_t lds_bytes = sizeof(float2) << lds_qubits;
    const size_t lds_per_cu = 160 * 1024;
    int per_cu_lds = (int)std::max<size_t>(1, lds_per_cu / (lds_bytes + 1024));
    int per_cu_threads = std::max(1, 2048 / block);
    int per_cu = std::min(per_cu_lds, per_cu_threads);
    return prop.multiProcessorCount * per_cu;
}

void set_lds_limit() {
    const int bytes = 160 * 1024 - 4096;
    AQS_HIP(hipFuncSetAttribute((const void*)k_batch<256>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    AQS_HIP(hipFuncSetAttribute((const void*)k_batch<512>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    AQS_HIP(hipFuncSetAttribute((const void*)k_batch<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
}

void build_graph(aqs_batch b) {
    destroy_graph(b);
    AQS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    try {
        launch_kernel(b, b->block, b->lds_qubits, b->grid);
    } catch (...) {
        hipGraph_t g;
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    AQS_HIP(hipStreamEndCapture(b->stream, &b->graph));
    AQS_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

double time_config(aqs_batch b, int block, int lds_qubits) {
    const int grid = grid_for(block, lds_qubits);
    launch_kernel(b, block, lds_qubits, grid);  // warm
    AQS_HIP(hipEventRecord(b->e0, b->stream));
    launch_kernel(b, block, lds_qubits, grid);
    AQS_HIP(hipEventRecord(b->e1, b->stream));
    AQS_HIP(hipEventSynchronize(b->e1));
    float ms = 0.f;
    AQS_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
    return ms;
}

aqs_batch check_batch(aqs_batch b) {
    if (!b) fail(AQS_ERR_INVALID_ARGUMENT, "null batch handle");
    return b;
}

}  // namespace

extern "C" {

aqs_status aqs_batch_create(int count, const int* nqubits, const aqs_op* ops, const size_t* op_offsets,
                            aqs_batch* out) {
    return guard([&] {
        if (!out) fail(AQS_ERR_INVALID_ARGUMENT, "null output handle");
        *out = nullptr;
        if (count < 1 || !nqubits || !op_offsets) fail(AQS_ERR_INVALID_ARGUMENT, "empty batch");
        auto* b = new aqs_batch_s();
        try {
            b->count = count;
            b->nq.assign(nqubits, nqubits + count);
            b->source_ops = op_offsets[count];
            std::vector<uint64_t> cost(count);
            std::vector<std::vector<BOp>> per(count);
            for (int s = 0; s < count; ++s) {
                const int n = nqubits[s];
                if (n < 1 || n > 20) fail(AQS_ERR_INVALID_ARGUMENT, "batched states take 1..20 qubits");
                std::vector<M2> pending(n);
                std::vector<char> has(n, 0);
                auto flush = [&](int q) {
                    if (!has[q]) return;
                    BOp o{};
                    o.kind = AQS_OP_MAT1;
                    o.t0 = n - 1 - q;
                    for (int i = 0; i < 4; ++i) {
                        o.m[2 * i] = (float)pending[q].a[i].real();
                        o.m[2 * i + 1] = (float)pending[q].a[i].imag();
                    }
                    per[s].push_back(o);
                    has[q] = 0;
                };
                for (size_t i = op_offsets[s]; i < op_offsets[s + 1]; ++i) {
                    const aqs_op& op = ops[i];
                    validate_op(op, n);
                    if (op.kind == AQS_OP_MATK) fail(AQS_ERR_INVALID_ARGUMENT, "batched states take no MATK ops");
                    const bool one_q =
                        (op.kind == AQS_OP_MAT1 || op.kind == AQS_OP_DIAG1) && op.ncontrols == 0 && !op.more_controls;
                    if (one_q) {
                        const int q = op.targets[0];
                        M2 m = mat_of(op);
                        if (has[q]) {
                            pending[q] = mul(m, pending[q]);
                        } else {
                            pending[q] = m;
                            has[q] = 1;
                        }
                        continue;
                    }
                    for (int j = 0; j < op.ntargets; ++j) flush(op.targets[j]);
                    for (int j = 0; j < op.ncontrols; ++j) flush(op.controls[j]);
                    for (int q = 0; q < n; ++q)
                        if ((op.more_controls >> q) & 1ull) flush(q);
                    BOp o{};
                    o.kind = op.kind;
                    o.t0 = n - 1 - op.targets[0];
                    o.t1 = op.ntargets > 1 ? n - 1 - op.targets[1] : 0;
                    o.any = (op.flags & AQS_CTRL_ANY) ? 1 : 0;
                    for (int j = 0; j < op.ncontrols; ++j) o.cmask |= 1u << (n - 1 - op.controls[j]);
                    for (int q = 0; q < n; ++q)
                        if ((op.more_controls >> q) & 1ull) o.cmask |= 1u << (n - 1 - q);
                    if (op.kind == AQS_OP_PHASE) {
                        o.pmask = o.cmask;
                        for (int j = 0; j < op.ntargets; ++j) o.pmask |= 1u << (n - 1 - op.targets[j]);
                    }
                    for (int j = 0; j < 8; ++j) o.m[j] = (float)op.m[j];
                    per[s].push_back(o);
                }
                for (int q = 0; q < n; ++q) flush(q);
                cost[s] = (uint64_t)per[s].size() << n;
            }
            std::vector<int> order(count);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
            std::vector<uint64_t> offset(count);
            uint64_t amps = 0;
            for (int s = 0; s < count; ++s) {
                offset[s] = amps;
                amps += 1ull << nqubits[s];
            }
            b->total_amps = amps;
            for (int k = 0; k < count; ++k) {
                const int s = order[k];
                Desc d{};
                d.amp_offset = offset[s];
                d.op_begin = (uint32_t)b->ops.size();
                d.op_count = (uint32_t)per[s].size();
                d.n = nqubits[s];
                d.index = s;
                b->desc.push_back(d);
                b->ops.insert(b->ops.end(), per[s].begin(), per[s].end());
            }
            current_device();
            AQS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
            AQS_HIP(hipEventCreate(&b->e0));
            AQS_HIP(hipEventCreate(&b->e1));
            AQS_HIP(hipMalloc(&b->d_desc, sizeof(Desc) * b->desc.size()));
            AQS_HIP(hipMalloc(&b->d_ops, sizeof(BOp) * std::max<size_t>(1, b->ops.size())));
            AQS_HIP(hipMalloc(&b->d_pool, sizeof(float2) * amps));
            AQS_HIP(hipMalloc(&b->d_out, sizeof(double) * count));
            AQS_HIP(hipMalloc(&b->d_queue, 256));
            AQS_HIP(hipMemcpy(b->d_desc, b->desc.data(), sizeof(Desc) * b->desc.size(), hipMemcpyHostToDevice));
            if (!b->ops.empty())
                AQS_HIP(hipMemcpy(b->d_ops, b->ops.data(), sizeof(BOp) * b->ops.size(), hipMemcpyHostToDevice));
            set_lds_limit();
            b->grid = grid_for(b->block, b->lds_qubits);
        } catch (...) {
            aqs_batch_destroy(b);
            throw;
        }
        *out = b;
    });
}

aqs_status aqs_batch_autotune(aqs_batch b, int* block, int* lds_qubits, double* best_ms) {
    return guard([&] {
        check_batch(b);
        double best = 1e30;
        int bb = b->block, bl = b->lds_qubits;
        for (int blk : {256, 512, 1024})
            for (int lq : {12, 13, 14}) {
                double ms = time_config(b, blk, lq);
                if (ms < best) {
                    best = ms;
                    bb = blk;
                    bl = lq;
                }
            }
        b->block = bb;
        b->lds_qubits = bl;
        b->grid = grid_for(bb, bl);
        b->tuned = true;
        destroy_graph(b);
        if (block) *block = bb;
        if (lds_qubits) *lds_qubits = bl;
        if (best_ms) *best_ms = best;
    });
}

aqs_status aqs_batch_run(aqs_batch b, double* expect_sum_z, double* ms) {
    return guard([&] {
        check_batch(b);
        if (!b->exec) build_graph(b);
        AQS_HIP(hipEventRecord(b->e0, b->stream));
        AQS_HIP(hipGraphLaunch(b->exec, b->stream));
        AQS_HIP(hipEventRecord(b->e1, b->stream));
        if (expect_sum_z)
            AQS_HIP(hipMemcpyAsync(expect_sum_z, b->d_out, sizeof(double) * b->count, hipMemcpyDeviceToHost, b->stream));
        AQS_HIP(hipStreamSynchronize(b->stream));
        float t = 0.f;
        AQS_HIP(hipEventElapsedTime(&t, b->e0, b->e1));
        b->last_ms = t;
        if (ms) *ms = t;
    });
}

aqs_status aqs_batch_state(aqs_batch b, int index, void* host) {
    return guard([&] {
        check_batch(b);
        if (index < 0 || index >= b->count) fail(AQS_ERR_OUT_OF_RANGE, "batch index out of range");
        if (!host) fail(AQS_ERR_INVALID_ARGUMENT, "null host buffer");
        uint64_t off = 0;
        for (int s = 0; s < index; ++s) off += 1ull << b->nq[s];
        AQS_HIP(hipMemcpyAsync(host, b->d_pool + off, sizeof(float2) << b->nq[index], hipMemcpyDeviceToHost, b->stream));
        AQS_HIP(hipStreamSynchronize(b->stream));
    });
}

aqs_status aqs_batch_info(aqs_batch b, size_t* passes, size_t* source_ops, int* block, int* lds_qubits, int* grid) {
    return guard([&] {
        check_batch(b);
        if (passes) *passes = b->ops.size();
        if (source_ops) *source_ops = b->source_ops;
        if (block) *block = b->block;
        if (lds_qubits) *lds_qubits = b->lds_qubits;
        if (grid) *grid = b->grid;
    });
}

aqs_status aqs_batch_destroy(aqs_batch b) {
    return guard([&] {
        if (!b) return;
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        destroy_graph(b);
        if (b->d_desc) (void)hipFree(b->d_desc);
        if (b->d_ops) (void)hipFree(b->d_ops);
        if (b->d_pool) (void)hipFree(b->d_pool);
        if (b->d_out) (void)hipFree(b->d_out);
        if (b->d_queue) (void)hipFree(b->d_queue);
        if (b->e0) (void)hipEventDestroy(b->e0);
        if (b->e1) (void)hipEventDestroy(b->e1);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
    });
}

}  // extern "C"
