// Dense k-qubit blocks (k = 4, 5) on CDNA4 matrix cores, complex64 state.
//
// A fused block is a 2^k x 2^k complex unitary U applied to every group of
// 2^k amplitudes that differ only in the k target bits.  With the groups as
// columns, the pass is the complex GEMM  Y = U X  (X: 2^k x 2^(n-k)).  It is
// done as one real GEMM on the stacked operands
//
//     [Yr]   [Ur  -Ui] [Xr]
//     [Yi] = [Ui   Ur] [Xi]        (2K x 2K real times 2K x columns)
//
// with v_mfma_f32_32x32x2_f32 (FP32 products, FP32 accumulation).  One
// wavefront owns a tile of 32 columns: the stacked U (64 x 64 real for
// k = 5, 32 x 32 for k = 4) lives in VGPRs for the whole kernel, every lane
// loads the 16 (k = 5) or 8 (k = 4) complex amplitudes its B operands need
// and writes the complex results straight from the accumulators, so no LDS
// round trip is needed.  The 32 columns of a tile are consecutive in the
// non-target bits; when the targets are not the lowest bits each row of a
// tile is one 256-byte run of memory.
//
// Operand layouts of v_mfma_f32_32x32x2_f32 (wave64):
//   A (32 x 2):  lane l holds A[l % 32][l / 32]
//   B (2 x 32):  lane l holds B[l / 32][l % 32]
//   D (32 x 32): lane l, register r holds D[8*(r/4) + 4*(l/32) + r%4][l % 32]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.h"

using namespace aqs_hip;

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kWaves = 4;  // wavefronts per workgroup
constexpr int kBlock = 64 * kWaves;

struct MfmaArgs {
    float2* sv;
    uint64_t tiles;     // number of 32-column tiles
    int nfixed;
    int fixed[5];       // sorted target bit positions
    uint32_t off[32];   // amplitude offset of sub-index r
    const float* apack; // packed stacked operand
};

__device__ __forceinline__ uint64_t insert_zeros5(uint64_t t, const MfmaArgs& a) {
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        if (j >= a.nfixed) break;
        const uint64_t p = (uint64_t)a.fixed[j];
        t = ((t >> p) << (p + 1)) | (t & ((1ull << p) - 1ull));
    }
    return t;
}

// k = 5: K = 32 complex rows, stacked real 64 x 64.
__global__ __launch_bounds__(kBlock) void k_matk32_mfma(MfmaArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 31;
    const int half = lane >> 5;

    // A operand: [rt][s] for rt = 0,1 (real / imag output rows), s = 0..31.
    float A0[32], A1[32];
#pragma unroll
    for (int s = 0; s < 32; ++s) {
        A0[s] = a.apack[(0 * 32 + s) * 64 + lane];
        A1[s] = a.apack[(1 * 32 + s) * 64 + lane];
    }
    uint32_t loff[16], soff[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) loff[s] = a.off[2 * s + half];
#pragma unroll
    for (int r = 0; r < 16; ++r) soff[r] = a.off[8 * (r / 4) + 4 * half + (r % 4)];

    const uint64_t wstride = (uint64_t)gridDim.x * kWaves;
    for (uint64_t tile = (uint64_t)blockIdx.x * kWaves + wave; tile < a.tiles; tile += wstride) {
        const uint64_t base = insert_zeros5(tile * 32 + col, a);
        float2 x[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) x[s] = ld_nt(&a.sv[base + loff[s]]);
        floatx16 acc0 = {0}, acc1 = {0};
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(A0[s], x[s].x, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A1[s], x[s].x, acc1, 0, 0, 0);
        }
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(A0[16 + s], x[s].y, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A1[16 + s], x[s].y, acc1, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r)
            st_nt(make_float2(acc0[r], acc1[r]), &a.sv[base + soff[r]]);
    }
}

// k = 4: K = 16 complex rows, stacked real 32 x 32; two column tiles per
// iteration keep 16 loads in flight per lane.
__global__ __launch_bounds__(kBlock) void k_matk16_mfma(MfmaArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 31;
    const int half = lane >> 5;

    float A[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) A[s] = a.apack[s * 64 + lane];
    uint32_t loff[8], soff[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) loff[s] = a.off[2 * s + half];
#pragma unroll
    for (int r = 0; r < 8; ++r) soff[r] = a.off[8 * (r / 4) + 4 * half + (r % 4)];

    const uint64_t pairs = (a.tiles + 1) / 2;
    const uint64_t wstride = (uint64_t)gridDim.x * kWaves;
    for (uint64_t p = (uint64_t)blockIdx.x * kWaves + wave; p < pairs; p += wstride) {
        const uint64_t t0 = 2 * p, t1 = 2 * p + 1;
        const bool has1 = t1 < a.tiles;
        const uint64_t b0 = insert_zeros5(t0 * 32 + col, a);
        const uint64_t b1 = has1 ? insert_zeros5(t1 * 32 + col, a) : b0;
        float2 x0[8], x1[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) x0[s] = ld_nt(&a.sv[b0 + loff[s]]);
        if (has1) {
#pragma unroll
            for (int s = 0; s < 8; ++s) x1[s] = ld_nt(&a.sv[b1 + loff[s]]);
        }
        floatx16 acc0 = {0}, acc1 = {0};
#pragma unroll
        for (int s = 0; s < 8; ++s) acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(A[s], x0[s].x, acc0, 0, 0, 0);
#pragma unroll
        for (int s = 0; s < 8; ++s) acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(A[8 + s], x0[s].y, acc0, 0, 0, 0);
        if (has1) {
#pragma unroll
            for (int s = 0; s < 8; ++s) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A[s], x1[s].x, acc1, 0, 0, 0);
#pragma unroll
            for (int s = 0; s < 8; ++s)
                acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A[8 + s], x1[s].y, acc1, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 8; ++r)
            st_nt(make_float2(acc0[r], acc0[r + 8]), &a.sv[b0 + soff[r]]);
        if (has1) {
#pragma unroll
            for (int r = 0; r < 8; ++r)
                st_nt(make_float2(acc1[r], acc1[r + 8]), &a.sv[b1 + soff[r]]);
        }
    }
}

}  // namespace

namespace aqs_hip {

// Stacked real operand in per-lane order.
//   k = 5: apack[(rt*32 + s)*64 + l] = S[32*rt + l%32][2s + l/32]   (S: 64 x 64)
//   k = 4: apack[s*64 + l]           = S[l%32][2s + l/32]           (S: 32 x 32)
// with S[R][C] = Re/Im blocks of U as in the header comment.
std::vector<float> pack_mfma_operand(int k, const double* u) {
    const int K = 1 << k;
    auto S = [&](int R, int C) -> double {
        const int r = R % K, c = C % K;
        const double ur = u[2 * (r * K + c)], ui = u[2 * (r * K + c) + 1];
        const bool rim = R >= K, cim = C >= K;
        if (!rim && !cim) return ur;
        if (!rim && cim) return -ui;
        if (rim && !cim) return ui;
        return ur;
    };
    std::vector<float> out;
    if (k == 5) {
        out.resize(2 * 32 * 64);
        for (int rt = 0; rt < 2; ++rt)
            for (int s = 0; s < 32; ++s)
                for (int l = 0; l < 64; ++l) out[(rt * 32 + s) * 64 + l] = (float)S(32 * rt + l % 32, 2 * s + l / 32);
    } else if (k == 4) {
        out.resize(16 * 64);
        for (int s = 0; s < 16; ++s)
            for (int l = 0; l < 64; ++l) out[s * 64 + l] = (float)S(l % 32, 2 * s + l / 32);
    }
    return out;
}

bool apply_matk_mfma(aqs_state s, const aqs_op& op, const float* dev_apack) {
    const int k = op.ntargets;
    if (s->precision != AQS_C64 || op.ncontrols != 0 || (k != 4 && k != 5)) return false;
    if (s->n > 32 || s->n < k + 5) return false;  // 32-bit offsets; at least one 32-column tile
    MfmaArgs a{};
    a.sv = (float2*)s->amps;
    std::vector<int> pos;
    for (int j = 0; j < k; ++j) pos.push_back(qpos(s->n, op.targets[j]));
    std::sort(pos.begin(), pos.end());
    a.nfixed = k;
    for (int j = 0; j < k; ++j) a.fixed[j] = pos[j];
    const int K = 1 << k;
    for (int r = 0; r < K; ++r) {
        uint32_t o = 0;
        for (int j = 0; j < k; ++j)
            if ((r >> (k - 1 - j)) & 1) o |= (uint32_t)qbit(s->n, op.targets[j]);
        a.off[r] = o;
    }
    a.tiles = (s->N >> k) / 32;
    a.apack = dev_apack;
    const uint64_t work = (k == 5) ? a.tiles : (a.tiles + 1) / 2;
    uint64_t blocks = (work + kWaves - 1) / kWaves;
    blocks = std::min<uint64_t>(std::max<uint64_t>(blocks, 1), 8192);
    ProfScope ps(s, kSlotMfma);
    if (k == 5)
        hipLaunchKernelGGL(k_matk32_mfma, dim3((unsigned)blocks), dim3(kBlock), 0, s->stream, a);
    else
        hipLaunchKernelGGL(k_matk16_mfma, dim3((unsigned)blocks), dim3(kBlock), 0, s->stream, a);
    AQS_HIP(hipGetLastError());
    return true;
}

}  // namespace aqs_hip
