// Batched simulation of many small, independent states in one persistent
// launch (variational ansatz workload: thousands of 10-16 qubit circuits,
// each followed by <sum Z> in FP64).
//
// The reference simulates one state at a time with a 2^n x 2^n operator per
// gate (src/quantum.cpp:277-291).  Here every workgroup of a persistent grid
// takes the next state from an atomic queue (states sorted by descending
// cost).  A state that fits the LDS tile (2^lds_qubits amplitudes) is built
// as |0..0> in LDS, runs all its gate passes there with a workgroup barrier
// between passes, and is written to HBM once.  A larger state runs as a few
// phases (batch_phases): per phase every tile of 2^lds_qubits amplitudes is
// loaded into LDS, runs all of the phase's gates and is stored, so the state
// crosses the memory hierarchy once per phase instead of once per gate.
// <sum_q Z_q> is reduced in FP64 from the final tiles.  The launch (queue
// reset + persistent kernel) is captured once into a hipGraph and replayed.
// An autotuner times (workgroup size, LDS tile) candidates and keeps the
// fastest.
//
// Per-state host preprocessing absorbs runs of uncontrolled one-qubit gates
// on a qubit into one 2x2 matrix (FP64 products) until a multi-qubit gate
// touches that qubit, so an RY-RZ layer costs one pass per qubit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#include "batch.h"
#include "rotation.h"

using namespace aqs_hip;

namespace {

struct Desc {
    uint64_t amp_offset;  // into the HBM pool (amplitudes)
    uint32_t phase_begin, phase_count;
    int32_t n;
    int32_t index;  // position in the caller's order
};

__device__ __forceinline__ uint32_t ins0(uint32_t p, int t) { return ((p >> t) << (t + 1)) | (p & ((1u << t) - 1u)); }

// LDS slot of tile index i: one padding slot per 32 amplitudes, so threads
// whose indices differ by 2^k (k >= 4) fall on different banks.  For i and
// j with disjoint bits, lds_slot(i | j) = lds_slot(i) + lds_slot(j).
__device__ __forceinline__ uint32_t lds_slot(uint32_t i) { return i + (i >> 5); }

__device__ __forceinline__ bool ctrl_ok(uint32_t i, uint32_t cmask, int any) {
    if (!cmask) return true;
    return any ? (i & cmask) != 0 : (i & cmask) == cmask;
}

// Tile-local index j -> state index: zeros inserted at the (ascending) bit
// positions of gbits, then the tile's own bits g or'ed in.
__device__ __forceinline__ uint32_t spread(uint32_t j, uint32_t gbits) {
    for (uint32_t g = gbits; g; g &= g - 1u) j = ins0(j, __ffs(g) - 1);
    return j;
}

// Tile number t -> its bits at the positions of gbits.
__device__ __forceinline__ uint32_t deposit(uint32_t t, uint32_t gbits) {
    uint32_t r = 0;
    for (uint32_t g = gbits; g; g &= g - 1u, t >>= 1)
        if (t & 1u) r |= g & (0u - g);
    return r;
}

// Controls outside the tile, given the tile's bits g: false when the op does
// nothing on this tile; an any-of control met outside makes it unconditional.
__device__ __forceinline__ bool resolve(BatchOp& e, uint32_t g) {
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

// One gate as one pass over the tile in LDS (plain segments).
template <int BLOCK>
__device__ void apply_block(float2* lds, uint32_t N, const BatchOp& op) {
    const int tid = threadIdx.x;
    switch (op.kind) {
        case AQS_OP_MAT1: {
            for (uint32_t p = tid; p < N / 2; p += BLOCK) {
                const uint32_t i0 = ins0(p, op.t0), i1 = i0 | (1u << op.t0);
                if (!ctrl_ok(i0, op.cmask, op.any)) continue;
                const float2 x = lds[lds_slot(i0)], y = lds[lds_slot(i1)];
                float2 a, b;
                a.x = op.m[0] * x.x - op.m[1] * x.y + op.m[2] * y.x - op.m[3] * y.y;
                a.y = op.m[0] * x.y + op.m[1] * x.x + op.m[2] * y.y + op.m[3] * y.x;
                b.x = op.m[4] * x.x - op.m[5] * x.y + op.m[6] * y.x - op.m[7] * y.y;
                b.y = op.m[4] * x.y + op.m[5] * x.x + op.m[6] * y.y + op.m[7] * y.x;
                lds[lds_slot(i0)] = a;
                lds[lds_slot(i1)] = b;
            }
            break;
        }
        case kBatchOpMatP: {
            for (uint32_t p = tid; p < N / 2; p += BLOCK) {
                const uint32_t i0 = ins0(p, op.t0), i1 = i0 | (1u << op.t0);
                const float2 x = lds[lds_slot(i0)], y = lds[lds_slot(i1)];
                float2 a = x, b = make_float2(op.m[4] * y.x - op.m[5] * y.y, op.m[4] * y.y + op.m[5] * y.x);
                a.x -= op.m[0] * b.x;
                a.y -= op.m[0] * b.y;
                b.x += op.m[1] * a.x;
                b.y += op.m[1] * a.y;
                a.x -= op.m[0] * b.x;
                a.y -= op.m[0] * b.y;
                lds[lds_slot(i0)] = a;
                lds[lds_slot(i1)] = b;
            }
            break;
        }
        case AQS_OP_DIAG1: {
            for (uint32_t i = tid; i < N; i += BLOCK) {
                if (!ctrl_ok(i, op.cmask, op.any)) continue;
                const bool one = (i >> op.t0) & 1u;
                const float dr = one ? op.m[6] : op.m[0], di = one ? op.m[7] : op.m[1];
                const float2 x = lds[lds_slot(i)];
                lds[lds_slot(i)] = make_float2(dr * x.x - di * x.y, dr * x.y + di * x.x);
            }
            break;
        }
        case AQS_OP_PHASE: {
            for (uint32_t i = tid; i < N; i += BLOCK) {
                if ((i & op.pmask) != op.pmask) continue;
                const float2 x = lds[lds_slot(i)];
                lds[lds_slot(i)] = make_float2(op.m[0] * x.x - op.m[1] * x.y, op.m[0] * x.y + op.m[1] * x.x);
            }
            break;
        }
        case AQS_OP_X: {
            for (uint32_t p = tid; p < N / 2; p += BLOCK) {
                const uint32_t i0 = ins0(p, op.t0), i1 = i0 | (1u << op.t0);
                if (!ctrl_ok(i0, op.cmask, op.any)) continue;
                const float2 x = lds[lds_slot(i0)];
                lds[lds_slot(i0)] = lds[lds_slot(i1)];
                lds[lds_slot(i1)] = x;
            }
            break;
        }
        case AQS_OP_SWAP: {
            const uint32_t ma = 1u << op.t0, mb = 1u << op.t1;
            for (uint32_t i = tid; i < N; i += BLOCK) {
                if (!(i & ma) || (i & mb) || !ctrl_ok(i, op.cmask, op.any)) continue;
                const uint32_t j = i ^ ma ^ mb;
                const float2 x = lds[lds_slot(i)];
                lds[lds_slot(i)] = lds[lds_slot(j)];
                lds[lds_slot(j)] = x;
            }
            break;
        }
        default:
            break;
    }
}

constexpr int kRegs = 1 << kBatchRegBits;

// Register ops.  Every variant assigns all 16 amplitudes in straight-line
// code: a register whose controls fail gets identity coefficients (or keeps
// its operand in a select) instead of a branch.  Partial updates under
// branches make the compiler copy the whole array at every join, which
// costs more than the selects.  Slot controls `creg` are uniform over the
// workgroup; a lane's controls outside the registers (`p`) only select.

// 2x2 [[m0 m1] [m2 m3]] (complex, interleaved in m) on slot S, no slot
// controls.
template <int S>
__device__ __forceinline__ void reg_mat_u(float2 (&a)[kRegs], const float (&m)[8]) {
#pragma unroll
    for (int k = 0; k < kRegs; ++k) {
        if ((k >> S) & 1) continue;
        const float2 x = a[k], y = a[k | (1 << S)];
        a[k] = make_float2(m[0] * x.x - m[1] * x.y + m[2] * y.x - m[3] * y.y,
                           m[0] * x.y + m[1] * x.x + m[2] * y.y + m[3] * y.x);
        a[k | (1 << S)] = make_float2(m[4] * x.x - m[5] * x.y + m[6] * y.x - m[7] * y.y,
                                      m[4] * x.y + m[5] * x.x + m[6] * y.y + m[7] * y.x);
    }
}

// The same for registers whose slots `creg` are all set (identity on the
// others).
template <int S>
__device__ __forceinline__ void reg_mat_c(float2 (&a)[kRegs], const float (&m)[8], uint32_t creg) {
#pragma unroll
    for (int k = 0; k < kRegs; ++k) {
        if ((k >> S) & 1) continue;
        const bool ok = ((uint32_t)k & creg) == creg;
        const float c0 = ok ? m[0] : 1.f, c1 = ok ? m[1] : 0.f, c2 = ok ? m[2] : 0.f, c3 = ok ? m[3] : 0.f;
        const float c4 = ok ? m[4] : 0.f, c5 = ok ? m[5] : 0.f, c6 = ok ? m[6] : 1.f, c7 = ok ? m[7] : 0.f;
        const float2 x = a[k], y = a[k | (1 << S)];
        a[k] = make_float2(c0 * x.x - c1 * x.y + c2 * y.x - c3 * y.y, c0 * x.y + c1 * x.x + c2 * y.y + c3 * y.x);
        a[k | (1 << S)] =
            make_float2(c4 * x.x - c5 * x.y + c6 * y.x - c7 * y.y, c4 * x.y + c5 * x.x + c6 * y.y + c7 * y.x);
    }
}

// A rotation after a phase (kBatchOpMatP) on slot S: no controls.
template <int S>
__device__ __forceinline__ void reg_matp(float2 (&a)[kRegs], const float (&m)[8]) {
#pragma unroll
    for (int k = 0; k < kRegs; ++k) {
        if ((k >> S) & 1) continue;
        float2 x = a[k];
        const float2 y = a[k | (1 << S)];
        float2 w = make_float2(m[4] * y.x - m[5] * y.y, m[4] * y.y + m[5] * y.x);
        x.x -= m[0] * w.x;
        x.y -= m[0] * w.y;
        w.x += m[1] * x.x;
        w.y += m[1] * x.y;
        x.x -= m[0] * w.x;
        x.y -= m[0] * w.y;
        a[k] = x;
        a[k | (1 << S)] = w;
    }
}

// Exchange across slot S on registers whose slots `creg` are all set, on
// lanes with p.
template <int S>
__device__ __forceinline__ void reg_x(float2 (&a)[kRegs], uint32_t creg, bool p) {
#pragma unroll
    for (int k = 0; k < kRegs; ++k) {
        if ((k >> S) & 1) continue;
        const bool ok = p && ((uint32_t)k & creg) == creg;
        const float2 x = a[k], y = a[k | (1 << S)];
        a[k] = make_float2(ok ? y.x : x.x, ok ? y.y : x.y);
        a[k | (1 << S)] = make_float2(ok ? x.x : y.x, ok ? x.y : y.y);
    }
}

// One gate on a thread's registers; `base` is the tile index of register 0
// (zero at the register bits).  (Register segments hold no any-of
// controlled gates.)
__device__ __forceinline__ void apply_regs(float2 (&a)[kRegs], const BatchOp& e, uint32_t base) {
    const uint32_t outside = e.cmask | e.pmask;
    const bool p = (base & outside) == outside;
    if (e.kind == AQS_OP_X) {
        switch (e.t0) {
            case 0: reg_x<0>(a, e.creg, p); break;
            case 1: reg_x<1>(a, e.creg, p); break;
            case 2: reg_x<2>(a, e.creg, p); break;
            default: reg_x<3>(a, e.creg, p); break;
        }
        return;
    }
    if (e.kind == AQS_OP_PHASE) {
        // one case per register mask: each multiplies its fixed set of
        // registers (16 / 2^|mask| of them); a real factor (CZ, Z) scales
        const float zr = p ? e.m[0] : 1.f, zi = p ? e.m[1] : 0.f;
        if (e.m[1] == 0.f) {
            switch (e.preg) {
#define AQS_PREG_REAL(M)                                               \
    case M:                                                            \
        _Pragma("unroll") for (int k = 0; k < kRegs; ++k) if ((k & (M)) == (M)) { \
            a[k].x *= zr;                                              \
            a[k].y *= zr;                                              \
        }                                                              \
        break;
                AQS_PREG_REAL(0) AQS_PREG_REAL(1) AQS_PREG_REAL(2) AQS_PREG_REAL(3)
                AQS_PREG_REAL(4) AQS_PREG_REAL(5) AQS_PREG_REAL(6) AQS_PREG_REAL(7)
                AQS_PREG_REAL(8) AQS_PREG_REAL(9) AQS_PREG_REAL(10) AQS_PREG_REAL(11)
                AQS_PREG_REAL(12) AQS_PREG_REAL(13) AQS_PREG_REAL(14)
#undef AQS_PREG_REAL
                default:  // mask 15
                    a[15].x *= zr;
                    a[15].y *= zr;
                    break;
            }
            return;
        }
        switch (e.preg) {
#define AQS_PREG_CPLX(M)                                               \
    case M:                                                            \
        _Pragma("unroll") for (int k = 0; k < kRegs; ++k) if ((k & (M)) == (M)) { \
            const float2 x = a[k];                                     \
            a[k] = make_float2(zr * x.x - zi * x.y, zr * x.y + zi * x.x); \
        }                                                              \
        break;
            AQS_PREG_CPLX(0) AQS_PREG_CPLX(1) AQS_PREG_CPLX(2) AQS_PREG_CPLX(3)
            AQS_PREG_CPLX(4) AQS_PREG_CPLX(5) AQS_PREG_CPLX(6) AQS_PREG_CPLX(7)
            AQS_PREG_CPLX(8) AQS_PREG_CPLX(9) AQS_PREG_CPLX(10) AQS_PREG_CPLX(11)
            AQS_PREG_CPLX(12) AQS_PREG_CPLX(13) AQS_PREG_CPLX(14)
#undef AQS_PREG_CPLX
            default: {  // mask 15
                const float2 x = a[15];
                a[15] = make_float2(zr * x.x - zi * x.y, zr * x.y + zi * x.x);
                break;
            }
        }
        return;
    }
    if (e.kind == kBatchOpMatP) {
        const float m[8] = {e.m[0], e.m[1], 0.f, 0.f, e.m[4], e.m[5], 0.f, 0.f};
        switch (e.t0) {
            case 0: reg_matp<0>(a, m); break;
            case 1: reg_matp<1>(a, m); break;
            case 2: reg_matp<2>(a, m); break;
            default: reg_matp<3>(a, m); break;
        }
        return;
    }
    // MAT1, or DIAG1 as a 2x2 with zero off-diagonal entries
    const bool off = p && e.kind != AQS_OP_DIAG1;
    const float m[8] = {p ? e.m[0] : 1.f, p ? e.m[1] : 0.f, off ? e.m[2] : 0.f, off ? e.m[3] : 0.f,
                        off ? e.m[4] : 0.f, off ? e.m[5] : 0.f, p ? e.m[6] : 1.f, p ? e.m[7] : 0.f};
    if (e.creg == 0) {
        switch (e.t0) {
            case 0: reg_mat_u<0>(a, m); break;
            case 1: reg_mat_u<1>(a, m); break;
            case 2: reg_mat_u<2>(a, m); break;
            default: reg_mat_u<3>(a, m); break;
        }
    } else {
        switch (e.t0) {
            case 0: reg_mat_c<0>(a, m, e.creg); break;
            case 1: reg_mat_c<1>(a, m, e.creg); break;
            case 2: reg_mat_c<2>(a, m, e.creg); break;
            default: reg_mat_c<3>(a, m, e.creg); break;
        }
    }
}

// A register segment over a tile of T amplitudes in LDS: every thread takes
// groups of 16 amplitudes (the register bits' combinations), runs the
// segment's gates on them and stores them back.
template <int BLOCK>
__device__ void run_reg_segment(float2* lds, uint32_t T, const BatchSeg& sg, const BatchOp* __restrict__ ops,
                                uint32_t g) {
    uint32_t pos[kBatchRegBits];
    uint32_t r = sg.rmask;
#pragma unroll
    for (int s = 0; s < kBatchRegBits; ++s) {
        pos[s] = __ffs(r) - 1;
        r &= r - 1u;
    }
    uint32_t off[kRegs];
#pragma unroll
    for (int k = 0; k < kRegs; ++k) {
        uint32_t o = 0;
#pragma unroll
        for (int s = 0; s < kBatchRegBits; ++s)
            if ((k >> s) & 1) o |= 1u << pos[s];
        off[k] = lds_slot(o);
    }
    for (uint32_t grp = threadIdx.x; grp < (T >> kBatchRegBits); grp += BLOCK) {
        uint32_t base = grp;
#pragma unroll
        for (int s = 0; s < kBatchRegBits; ++s) base = ins0(base, pos[s]);
        const uint32_t pb = lds_slot(base);
        float2 a[kRegs];
#pragma unroll
        for (int k = 0; k < kRegs; ++k) a[k] = lds[pb + off[k]];
        for (uint32_t o = 0; o < sg.op_count; ++o) {
            BatchOp e = ops[sg.op_begin + o];
            if (!resolve(e, g)) continue;
            apply_regs(a, e, base);
        }
#pragma unroll
        for (int k = 0; k < kRegs; ++k) lds[pb + off[k]] = a[k];
    }
    __syncthreads();
}

template <int BLOCK>
__device__ double block_sum(double acc, double* s_red) {
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < BLOCK / 64; ++w) s += s_red[w];
    return s;
}

// Persistent: each workgroup takes states from the queue until none is left.
// A state of n qubits is swept in tiles of 2^min(n, lds_qubits) amplitudes;
// its first phase starts from |0..0> in LDS, its last phase reduces
// <sum Z> while storing.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_batch(const Desc* __restrict__ desc, int count,
                                                  const BatchPhase* __restrict__ phases,
                                                  const BatchSeg* __restrict__ segs, const BatchOp* __restrict__ ops,
                                                  float2* __restrict__ pool, double* __restrict__ out,
                                                  unsigned* __restrict__ queue, int lds_qubits) {
    extern __shared__ float2 lds[];
    __shared__ int s_idx;
    __shared__ double s_red[BLOCK / 64];
    const int tid = threadIdx.x;
    while (true) {
        if (tid == 0) s_idx = (int)atomicAdd(queue, 1u);
        __syncthreads();
        const int k = __builtin_amdgcn_readfirstlane(s_idx);
        __syncthreads();
        if (k >= count) break;  // every wave of the workgroup leaves together
        const Desc d = desc[k];
        float2* hbm = pool + d.amp_offset;
        const int m = d.n < lds_qubits ? d.n : lds_qubits;
        const uint32_t T = 1u << m, tiles = 1u << (d.n - m);
        double acc = 0.0;
        for (uint32_t p = 0; p < d.phase_count; ++p) {
            const BatchPhase ph = phases[d.phase_begin + p];
            const bool last = p + 1 == d.phase_count;
            for (uint32_t t = 0; t < tiles; ++t) {
                const uint32_t g = deposit(t, ph.gbits);
                if (p == 0) {
                    for (uint32_t j = tid; j < T; j += BLOCK)
                        lds[lds_slot(j)] = make_float2(t == 0 && j == 0 ? 1.f : 0.f, 0.f);
                } else {
                    for (uint32_t j = tid; j < T; j += BLOCK) lds[lds_slot(j)] = hbm[spread(j, ph.gbits) | g];
                }
                __syncthreads();
                for (uint32_t s = 0; s < ph.seg_count; ++s) {
                    const BatchSeg sg = segs[ph.seg_begin + s];
                    if (sg.rmask) {
                        run_reg_segment<BLOCK>(lds, T, sg, ops, g);
                        continue;
                    }
                    for (uint32_t o = 0; o < sg.op_count; ++o) {
                        BatchOp e = ops[sg.op_begin + o];
                        if (!resolve(e, g)) continue;  // uniform over the workgroup
                        apply_block<BLOCK>(lds, T, e);
                        __syncthreads();
                    }
                }
                for (uint32_t j = tid; j < T; j += BLOCK) {
                    const float2 v = lds[lds_slot(j)];
                    const uint32_t i = spread(j, ph.gbits) | g;
                    if (last) acc += ((double)v.x * v.x + (double)v.y * v.y) * (double)(d.n - 2 * (int)__popc(i));
                    hbm[i] = v;
                }
                __syncthreads();  // the next tile's load reuses the LDS
            }
        }
        const double s = block_sum<BLOCK>(acc, s_red);
        if (tid == 0) out[d.index] = s;
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

// Bits an op must find inside the tile, bits it changes (non-diagonal
// use) and bits it only reads or multiplies on (diagonal use).
struct Uses {
    uint32_t need, nd, d;
};

Uses uses_of(const BatchOp& o) {
    const uint32_t t0 = 1u << o.t0;
    switch (o.kind) {
        case AQS_OP_MAT1:
        case AQS_OP_X:
        case kBatchOpMatP:
            return {t0, t0, o.cmask};
        case AQS_OP_SWAP:
            return {t0 | (1u << o.t1), t0 | (1u << o.t1), o.cmask};
        case AQS_OP_DIAG1:
            return {t0, 0u, t0 | o.cmask};
        default:  // PHASE
            return {0u, 0u, o.pmask};
    }
}

// Positions of the bits of `mask` inside the tile whose bits are `local`.
uint32_t compress(uint32_t mask, uint32_t local) {
    uint32_t r = 0;
    int k = 0;
    for (int b = 0; b < 32; ++b) {
        if (!((local >> b) & 1u)) continue;
        if ((mask >> b) & 1u) r |= 1u << k;
        ++k;
    }
    return r;
}

int local_pos(int b, uint32_t local) { return __builtin_popcount(local & ((1u << b) - 1u)); }

// Appends a phase's ops (tile-local positions, tile of `bits` bits) as
// segments, in order: each register segment takes ops while the bits they
// change fit in 4 positions (padded with the lowest free ones).  Gates move
// to slot form; a swap of a and b under controls C becomes CX(a -> b),
// (C and b) -> X on a, CX(a -> b).  Tiles of fewer than 4 bits and any-of
// controlled gates go to plain segments.
void cut_segments(const std::vector<BatchOp>& in, int bits, std::vector<BatchOp>& ops, std::vector<BatchSeg>& segs) {
    auto plain = [&](const BatchOp& o) { return bits < kBatchRegBits || o.any; };
    size_t i = 0;
    while (i < in.size()) {
        BatchSeg sg{(uint32_t)ops.size(), 0u, 0u, 0u};
        if (plain(in[i])) {
            for (; i < in.size() && plain(in[i]); ++i) {
                ops.push_back(in[i]);
                ++sg.op_count;
            }
            segs.push_back(sg);
            continue;
        }
        uint32_t R = 0;
        size_t j = i;
        for (; j < in.size() && !plain(in[j]); ++j) {
            const uint32_t need = uses_of(in[j]).need;
            if (__builtin_popcount(R | need) > kBatchRegBits) break;
            R |= need;
        }
        for (int b = 0; __builtin_popcount(R) < kBatchRegBits; ++b) R |= 1u << b;
        sg.rmask = R;
        auto slot = [&](int pos) { return __builtin_popcount(R & ((1u << pos) - 1u)); };
        auto slots = [&](uint32_t mask) { return (uint16_t)compress(mask & R, R); };
        for (; i < j; ++i) {
            BatchOp o = in[i];
            if (o.kind == AQS_OP_PHASE) {
                o.preg = slots(o.pmask);
                o.pmask &= ~R;
                o.cmask = 0;
                o.creg = 0;
            } else if (o.kind == AQS_OP_SWAP) {
                BatchOp cx{};
                cx.kind = AQS_OP_X;
                cx.t0 = slot(o.t1);
                cx.creg = (uint16_t)(1u << slot(o.t0));
                BatchOp mid = o;
                mid.kind = AQS_OP_X;
                mid.t0 = slot(o.t0);
                mid.t1 = 0;
                mid.creg = (uint16_t)(slots(o.cmask) | (1u << slot(o.t1)));
                mid.cmask = o.cmask & ~R;
                ops.push_back(cx);
                ops.push_back(mid);
                ops.push_back(cx);
                sg.op_count += 3;
                continue;
            } else {
                o.t0 = slot(o.t0);
                o.creg = slots(o.cmask);
                o.cmask &= ~R;
            }
            ops.push_back(o);
            ++sg.op_count;
        }
        segs.push_back(sg);
    }
}

}  // namespace

namespace aqs_hip {

std::vector<BatchOp> batch_absorb(const aqs_op* ops, size_t count, int n) {
    std::vector<BatchOp> out;
    std::vector<M2> pending(n);
    std::vector<char> has(n, 0);
    // the qubit's last rotation op (kBatchOpMatP) and its 2x2, to merge a
    // pending diagonal back into it at the end instead of a separate op
    std::vector<long> last_rot(n, -1);
    std::vector<M2> rot_mat(n);
    auto emit_mat = [&](int q, const M2& m) {
        BatchOp o{};
        o.kind = AQS_OP_MAT1;
        o.t0 = n - 1 - q;
        for (int i = 0; i < 4; ++i) {
            o.m[2 * i] = (float)m.a[i].real();
            o.m[2 * i + 1] = (float)m.a[i].imag();
        }
        return o;
    };
    // the whole pending 2x2, before a gate that uses the qubit otherwise
    auto flush = [&](int q) {
        last_rot[q] = -1;
        if (!has[q]) return;
        out.push_back(emit_mat(q, pending[q]));
        has[q] = 0;
    };
    // before a gate that uses the qubit only diagonally (a control, a phase):
    // the pending 2x2 U = diag(g0, g1) R diag(1, z) runs as one rotation op
    // and diag(g0, g1) stays pending (a diagonal commutes with that gate);
    // 10 multiply-adds per amplitude pair instead of 16
    auto flush_rotation = [&](int q) {
        if (!has[q]) return;
        const M2& u = pending[q];
        if (u.a[1] == cd(0) && u.a[2] == cd(0)) return;  // diagonal: stays pending
        RotationForm f;
        if (!factor_rotation(u.a[0], u.a[1], u.a[2], u.a[3], f)) {
            flush(q);
            return;
        }
        double t, su, nr;
        rotation_shears(f.c, f.s, t, su, nr);
        BatchOp o{};
        o.kind = kBatchOpMatP;
        o.t0 = n - 1 - q;
        o.m[0] = (float)t;
        o.m[1] = (float)su;
        o.m[4] = (float)f.z.real();
        o.m[5] = (float)f.z.imag();
        last_rot[q] = (long)out.size();
        out.push_back(o);
        // the op's 2x2: R diag(1, z) with R the unit rotation
        const double cu = f.c / nr;
        rot_mat[q].a[0] = cu;
        rot_mat[q].a[1] = -su * f.z;
        rot_mat[q].a[2] = su;
        rot_mat[q].a[3] = cu * f.z;
        pending[q].a[0] = f.g0 * nr;
        pending[q].a[1] = 0;
        pending[q].a[2] = 0;
        pending[q].a[3] = f.g1 * nr;
    };
    for (size_t i = 0; i < count; ++i) {
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
        const bool diagonal = op.kind == AQS_OP_PHASE || op.kind == AQS_OP_DIAG1;
        for (int j = 0; j < op.ntargets; ++j) {
            if (diagonal)
                flush_rotation(op.targets[j]);
            else
                flush(op.targets[j]);
        }
        for (int j = 0; j < op.ncontrols; ++j) flush_rotation(op.controls[j]);
        for (int q = 0; q < n; ++q)
            if ((op.more_controls >> q) & 1ull) flush_rotation(q);
        BatchOp o{};
        o.kind = op.kind;
        o.t0 = n - 1 - op.targets[0];
        o.t1 = op.ntargets > 1 ? n - 1 - op.targets[1] : 0;
        o.any = (op.flags & AQS_CTRL_ANY) && op.ncontrols > 0 ? 1 : 0;
        for (int j = 0; j < op.ncontrols; ++j) o.cmask |= 1u << (n - 1 - op.controls[j]);
        for (int q = 0; q < n; ++q)
            if ((op.more_controls >> q) & 1ull) o.cmask |= 1u << (n - 1 - q);
        if (op.kind == AQS_OP_PHASE) {
            o.pmask = o.cmask;
            for (int j = 0; j < op.ntargets; ++j) o.pmask |= 1u << (n - 1 - op.targets[j]);
        }
        for (int j = 0; j < 8; ++j) o.m[j] = (float)op.m[j];
        out.push_back(o);
    }
    for (int q = 0; q < n; ++q) {
        if (has[q] && last_rot[q] >= 0 && pending[q].a[1] == cd(0) && pending[q].a[2] == cd(0)) {
            // every op on the qubit since its last rotation is diagonal on
            // it (the pending product was diagonal at each of them and is
            // at the end), so they commute: the rotation takes the pending
            // diagonal and runs as the whole 2x2
            out[(size_t)last_rot[q]] = emit_mat(q, mul(pending[q], rot_mat[q]));
            has[q] = 0;
        }
        flush(q);
    }
    return out;
}

void batch_phases(const std::vector<BatchOp>& flat, int n, int m, std::vector<BatchOp>& ops,
                  std::vector<BatchSeg>& segs, std::vector<BatchPhase>& phases) {
    if (n <= m) {
        const uint32_t s0 = (uint32_t)segs.size();
        cut_segments(flat, n, ops, segs);
        phases.push_back({s0, (uint32_t)segs.size() - s0, 0u, 0u});
        return;
    }
    const uint32_t all = (n >= 32) ? ~0u : (1u << n) - 1u;
    std::vector<Uses> use(flat.size());
    for (size_t i = 0; i < flat.size(); ++i) use[i] = uses_of(flat[i]);
    std::vector<uint32_t> rem(flat.size()), next;
    std::iota(rem.begin(), rem.end(), 0u);
    std::vector<BatchOp> taken;
    while (!rem.empty()) {
        uint32_t local = (1u << std::min(5, m)) - 1u;
        for (uint32_t i : rem) {
            for (uint32_t b = use[i].need; b && __builtin_popcount(local) < m; b &= b - 1u) local |= b & (0u - b);
            if (__builtin_popcount(local) == m) break;
        }
        for (int b = 0; __builtin_popcount(local) < m; ++b) local |= 1u << b;
        uint32_t blocked_nd = 0, blocked_d = 0;
        next.clear();
        taken.clear();
        for (uint32_t i : rem) {
            const Uses& u = use[i];
            const bool ok = !(u.need & ~local) && !(u.nd & (blocked_nd | blocked_d)) && !(u.d & blocked_nd);
            if (!ok) {
                next.push_back(i);
                blocked_nd |= u.nd;
                blocked_d |= u.d;
                continue;
            }
            BatchOp o = flat[i];
            o.t0 = local_pos(o.t0, local);
            if (o.kind == AQS_OP_SWAP) o.t1 = local_pos(o.t1, local);
            if (o.kind == AQS_OP_PHASE) {
                o.gmask = o.pmask & ~local;
                o.pmask = compress(o.pmask, local);
                o.t0 = 0;
            } else {
                o.gmask = o.cmask & ~local;
            }
            o.cmask = compress(o.cmask, local);
            taken.push_back(o);
        }
        const uint32_t s0 = (uint32_t)segs.size();
        cut_segments(taken, m, ops, segs);
        phases.push_back({s0, (uint32_t)segs.size() - s0, all & ~local, 0u});
        rem.swap(next);
    }
}

}  // namespace aqs_hip

struct aqs_batch_s {
    // The program for one LDS tile size.
    struct Prog {
        std::vector<Desc> desc;  // sorted by descending cost
        std::vector<BatchOp> ops;
        std::vector<BatchSeg> segs;
        std::vector<BatchPhase> phases;
        void* d_desc = nullptr;
        void* d_ops = nullptr;
        void* d_segs = nullptr;
        void* d_phases = nullptr;
    };
    int count = 0;
    std::vector<int> nq;
    std::vector<std::vector<BatchOp>> flat;  // per state, absorbed, state-index positions
    std::vector<int> order;                  // states by descending cost
    std::vector<uint64_t> offset;            // amplitude offset of each state in the pool
    std::map<int, Prog> progs;               // by LDS tile qubits
    uint64_t total_amps = 0;
    size_t source_ops = 0;
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

void free_prog(aqs_batch_s::Prog& p) {
    if (p.d_desc) (void)hipFree(p.d_desc);
    if (p.d_ops) (void)hipFree(p.d_ops);
    if (p.d_segs) (void)hipFree(p.d_segs);
    if (p.d_phases) (void)hipFree(p.d_phases);
    p.d_desc = p.d_ops = p.d_segs = p.d_phases = nullptr;
}

// The program for LDS tiles of 2^m amplitudes, built and uploaded on first use.
aqs_batch_s::Prog& prog_for(aqs_batch b, int m) {
    auto it = b->progs.find(m);
    if (it != b->progs.end()) return it->second;
    aqs_batch_s::Prog& p = b->progs[m];
    try {
        for (int k = 0; k < b->count; ++k) {
            const int s = b->order[k];
            Desc d{};
            d.amp_offset = b->offset[s];
            d.phase_begin = (uint32_t)p.phases.size();
            batch_phases(b->flat[s], b->nq[s], m, p.ops, p.segs, p.phases);
            d.phase_count = (uint32_t)(p.phases.size() - d.phase_begin);
            d.n = b->nq[s];
            d.index = s;
            p.desc.push_back(d);
        }
        if (p.ops.size() >= (1ull << 32) || p.segs.size() >= (1ull << 32))
            fail(AQS_ERR_INVALID_ARGUMENT, "batch program exceeds 2^32 ops");
        AQS_HIP(hipMalloc(&p.d_desc, sizeof(Desc) * p.desc.size()));
        AQS_HIP(hipMalloc(&p.d_ops, sizeof(BatchOp) * std::max<size_t>(1, p.ops.size())));
        AQS_HIP(hipMalloc(&p.d_segs, sizeof(BatchSeg) * std::max<size_t>(1, p.segs.size())));
        AQS_HIP(hipMalloc(&p.d_phases, sizeof(BatchPhase) * p.phases.size()));
        AQS_HIP(hipMemcpy(p.d_desc, p.desc.data(), sizeof(Desc) * p.desc.size(), hipMemcpyHostToDevice));
        if (!p.ops.empty())
            AQS_HIP(hipMemcpy(p.d_ops, p.ops.data(), sizeof(BatchOp) * p.ops.size(), hipMemcpyHostToDevice));
        if (!p.segs.empty())
            AQS_HIP(hipMemcpy(p.d_segs, p.segs.data(), sizeof(BatchSeg) * p.segs.size(), hipMemcpyHostToDevice));
        AQS_HIP(hipMemcpy(p.d_phases, p.phases.data(), sizeof(BatchPhase) * p.phases.size(), hipMemcpyHostToDevice));
    } catch (...) {
        free_prog(p);
        b->progs.erase(m);
        throw;
    }
    return p;
}

void destroy_graph(aqs_batch b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Dynamic LDS of a tile of 2^lds_qubits amplitudes with one padding slot
// per 32 (lds_slot).
size_t lds_bytes_for(int lds_qubits) {
    const size_t T = 1ull << lds_qubits;
    return sizeof(float2) * (T + (T >> 5));
}

void launch_kernel(aqs_batch b, int block, int lds_qubits, int grid) {
    aqs_batch_s::Prog& p = prog_for(b, lds_qubits);
    const size_t lds_bytes = lds_bytes_for(lds_qubits);
    const auto* desc = (const Desc*)p.d_desc;
    const auto* phases = (const BatchPhase*)p.d_phases;
    const auto* segs = (const BatchSeg*)p.d_segs;
    const auto* ops = (const BatchOp*)p.d_ops;
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, b->stream, b->d_queue);
    switch (block) {
        case 256:
            hipLaunchKernelGGL(k_batch<256>, dim3(grid), dim3(256), lds_bytes, b->stream, desc, b->count, phases, segs, ops,
                               b->d_pool, b->d_out, b->d_queue, lds_qubits);
            break;
        case 512:
            hipLaunchKernelGGL(k_batch<512>, dim3(grid), dim3(512), lds_bytes, b->stream, desc, b->count, phases, segs, ops,
                               b->d_pool, b->d_out, b->d_queue, lds_qubits);
            break;
        default:
            hipLaunchKernelGGL(k_batch<1024>, dim3(grid), dim3(1024), lds_bytes, b->stream, desc, b->count, phases, segs, ops,
                               b->d_pool, b->d_out, b->d_queue, lds_qubits);
            break;
    }
    AQS_HIP(hipGetLastError());
}

int grid_for(int block, int lds_qubits) {
    int dev = current_device();
    hipDeviceProp_t prop;
    AQS_HIP(hipGetDeviceProperties(&prop, dev));
    const size_t lds_bytes = lds_bytes_for(lds_qubits);
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
    prog_for(b, b->lds_qubits);  // allocates and uploads: not allowed while capturing
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

// Median of three timed replays of the configuration's captured graph (as
// aqs_batch_run launches it) after one warm-up replay.  Leaves the batch
// configured as (block, lds_qubits).
double time_config(aqs_batch b, int block, int lds_qubits) {
    b->block = block;
    b->lds_qubits = lds_qubits;
    b->grid = grid_for(block, lds_qubits);
    build_graph(b);
    AQS_HIP(hipGraphLaunch(b->exec, b->stream));
    float t[3];
    for (float& ms : t) {
        AQS_HIP(hipEventRecord(b->e0, b->stream));
        AQS_HIP(hipGraphLaunch(b->exec, b->stream));
        AQS_HIP(hipEventRecord(b->e1, b->stream));
        AQS_HIP(hipEventSynchronize(b->e1));
        AQS_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
    }
    std::sort(t, t + 3);
    return t[1];
}

aqs_batch check_batch(aqs_batch b) {
    if (!b) fail(AQS_ERR_INVALID_ARGUMENT, "null batch handle");
    return b;
}

}  // namespace

// ---- request groups --------------------------------------------------------

namespace aqs_hip {

namespace {

constexpr size_t kArenaAlign = 256;
constexpr int kGroupRegions = 10;  // aligned regions of the arena layout

size_t align_up(size_t x) { return (x + kArenaAlign - 1) / kArenaAlign * kArenaAlign; }

}  // namespace

// A member's program holds at most nops absorbed ops, each cut into at most
// three (a swap becomes three X gates), at most one segment per op and at
// most one phase per absorbed op plus one.
size_t group_item_bytes(int n, size_t nops, uint32_t shots) {
    const size_t prog = sizeof(Desc) + (nops + 1) * sizeof(BatchPhase) + 3 * nops * sizeof(BatchSeg) +
                        3 * nops * sizeof(BatchOp);
    const size_t io = sizeof(double) * 2 + 2 * sizeof(uint64_t) * (size_t)shots;  // expectation, norm, u, indices
    return prog + io + align_up(sizeof(float2) << n) + align_up(sample_norm_scratch_bytes(n)) +
           kGroupRegions * kArenaAlign;
}

void run_group(GroupItem* items, int count, void* arena, size_t arena_bytes, hipStream_t stream) {
    if (count < 1) return;
    // programs, biggest state first (the persistent grid takes states in order)
    std::vector<int> order(count);
    std::vector<uint64_t> cost(count);
    std::vector<std::vector<BatchOp>> flat(count);
    for (int k = 0; k < count; ++k) {
        const GroupItem& it = items[k];
        if (it.n < 1 || it.n > kGroupMaxQubits) fail(AQS_ERR_INVALID_ARGUMENT, "group member qubits out of range");
        flat[k] = batch_absorb(it.ops, it.nops, it.n);
        cost[k] = (uint64_t)(flat[k].size() + 1) << it.n;
    }
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    std::vector<Desc> desc;
    std::vector<BatchOp> ops;
    std::vector<BatchSeg> segs;
    std::vector<BatchPhase> phases;
    size_t shots_total = 0;
    std::vector<size_t> shot_off(count);
    for (int k = 0; k < count; ++k) {
        shot_off[k] = shots_total;
        shots_total += items[k].shots;
    }
    // arena layout: queue | <sum Z> | upload (desc, phases, segs, ops, u) |
    // download (norms, indices) | states | sampling scratch
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += align_up(bytes);
        return at;
    };
    for (int k : order) {
        Desc d{};
        d.phase_begin = (uint32_t)phases.size();
        batch_phases(flat[k], items[k].n, kGroupLdsQubits, ops, segs, phases);
        d.phase_count = (uint32_t)(phases.size() - d.phase_begin);
        d.n = items[k].n;
        d.index = k;
        items[k].passes = d.phase_count;
        desc.push_back(d);
    }
    const size_t o_queue = take(sizeof(unsigned));
    const size_t o_zout = take(sizeof(double) * count);
    const size_t o_desc = take(sizeof(Desc) * desc.size());
    const size_t o_phases = take(sizeof(BatchPhase) * phases.size());
    const size_t o_segs = take(sizeof(BatchSeg) * std::max<size_t>(1, segs.size()));
    const size_t o_ops = take(sizeof(BatchOp) * std::max<size_t>(1, ops.size()));
    const size_t o_u = take(sizeof(double) * std::max<size_t>(1, shots_total));
    const size_t o_norm = take(sizeof(double) * count);
    const size_t o_idx = take(sizeof(uint64_t) * std::max<size_t>(1, shots_total));
    std::vector<size_t> o_state(count), o_scr(count);
    for (int k = 0; k < count; ++k) o_state[k] = take(sizeof(float2) << items[k].n);
    for (int k = 0; k < count; ++k) o_scr[k] = take(sample_norm_scratch_bytes(items[k].n));
    if (off > arena_bytes) fail(AQS_ERR_RUNTIME, "request group exceeds its device arena");
    char* base = (char*)arena;
    // amplitude offsets from the first state (the kernel's pool)
    for (Desc& d : desc) d.amp_offset = (o_state[d.index] - o_state[0]) / sizeof(float2);

    // one upload: desc .. u are contiguous
    std::vector<char> up(o_norm - o_desc);
    std::memcpy(up.data(), desc.data(), sizeof(Desc) * desc.size());
    std::memcpy(up.data() + (o_phases - o_desc), phases.data(), sizeof(BatchPhase) * phases.size());
    if (!segs.empty()) std::memcpy(up.data() + (o_segs - o_desc), segs.data(), sizeof(BatchSeg) * segs.size());
    if (!ops.empty()) std::memcpy(up.data() + (o_ops - o_desc), ops.data(), sizeof(BatchOp) * ops.size());
    for (int k = 0; k < count; ++k)
        if (items[k].shots)
            std::memcpy(up.data() + (o_u - o_desc) + sizeof(double) * shot_off[k], items[k].u,
                        sizeof(double) * items[k].shots);
    AQS_HIP(hipMemcpyAsync(base + o_desc, up.data(), up.size(), hipMemcpyHostToDevice, stream));

    set_lds_limit();
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, stream, (unsigned*)(base + o_queue));
    const unsigned grid = (unsigned)std::min(grid_for(kGroupBlock, kGroupLdsQubits), count);
    hipLaunchKernelGGL(k_batch<kGroupBlock>, dim3(grid), dim3(kGroupBlock), lds_bytes_for(kGroupLdsQubits), stream,
                       (const Desc*)(base + o_desc), count, (const BatchPhase*)(base + o_phases),
                       (const BatchSeg*)(base + o_segs), (const BatchOp*)(base + o_ops),
                       (float2*)(base + o_state[0]), (double*)(base + o_zout), (unsigned*)(base + o_queue),
                       kGroupLdsQubits);
    AQS_HIP(hipGetLastError());
    for (int k = 0; k < count; ++k)
        queue_sample_norm(base + o_state[k], items[k].n, base + o_scr[k],
                          (const double*)(base + o_u) + shot_off[k], items[k].shots,
                          (uint64_t*)(base + o_idx) + shot_off[k], (double*)(base + o_norm) + k, stream);
    std::vector<char> down(o_state[0] - o_norm);
    AQS_HIP(hipMemcpyAsync(down.data(), base + o_norm, down.size(), hipMemcpyDeviceToHost, stream));
    for (int k = 0; k < count; ++k)
        if (items[k].state)
            AQS_HIP(hipMemcpyAsync(items[k].state, base + o_state[k], sizeof(float2) << items[k].n,
                                   hipMemcpyDeviceToHost, stream));
    AQS_HIP(hipStreamSynchronize(stream));
    for (int k = 0; k < count; ++k) {
        if (items[k].norm2) std::memcpy(items[k].norm2, down.data() + sizeof(double) * k, sizeof(double));
        if (items[k].shots)
            std::memcpy(items[k].indices, down.data() + (o_idx - o_norm) + sizeof(uint64_t) * shot_off[k],
                        sizeof(uint64_t) * items[k].shots);
    }
}

}  // namespace aqs_hip

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
            b->flat.resize(count);
            for (int s = 0; s < count; ++s) {
                const int n = nqubits[s];
                if (n < 1 || n > 20) fail(AQS_ERR_INVALID_ARGUMENT, "batched states take 1..20 qubits");
                b->flat[s] = batch_absorb(ops + op_offsets[s], op_offsets[s + 1] - op_offsets[s], n);
                cost[s] = (uint64_t)b->flat[s].size() << n;
            }
            b->order.resize(count);
            std::iota(b->order.begin(), b->order.end(), 0);
            std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
            b->offset.resize(count);
            uint64_t amps = 0;
            for (int s = 0; s < count; ++s) {
                b->offset[s] = amps;
                amps += 1ull << nqubits[s];
            }
            b->total_amps = amps;
            current_device();
            AQS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
            AQS_HIP(hipEventCreate(&b->e0));
            AQS_HIP(hipEventCreate(&b->e1));
            AQS_HIP(hipMalloc(&b->d_pool, sizeof(float2) * amps));
            AQS_HIP(hipMalloc(&b->d_out, sizeof(double) * count));
            AQS_HIP(hipMalloc(&b->d_queue, 256));
            set_lds_limit();
            prog_for(b, b->lds_qubits);
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
        for (auto it = b->progs.begin(); it != b->progs.end();) {
            if (it->first == bl) {
                ++it;
                continue;
            }
            free_prog(it->second);
            it = b->progs.erase(it);
        }
        if (block) *block = bb;
        if (lds_qubits) *lds_qubits = bl;
        if (best_ms) *best_ms = best;
    });
}

aqs_status aqs_batch_configure(aqs_batch b, int block, int lds_qubits) {
    return guard([&] {
        check_batch(b);
        if (block != 256 && block != 512 && block != 1024)
            fail(AQS_ERR_INVALID_ARGUMENT, "batch workgroup size must be 256, 512 or 1024");
        if (lds_qubits < 12 || lds_qubits > 14) fail(AQS_ERR_INVALID_ARGUMENT, "batch LDS tile must be 12..14 qubits");
        prog_for(b, lds_qubits);
        b->block = block;
        b->lds_qubits = lds_qubits;
        b->grid = grid_for(block, lds_qubits);
        destroy_graph(b);
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
        if (passes) *passes = prog_for(b, b->lds_qubits).ops.size();
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
        for (auto& kv : b->progs) free_prog(kv.second);
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
