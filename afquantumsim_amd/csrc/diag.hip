// Diagonal table pass: many commuting phase gates applied in one sweep.
//
// compile() defers controlled-phase terms that do not fit the open dense
// blocks (e.g. the long CPhase ladders of the QFT, src/quantum_algo.cpp:
// 103-114) and emits them here.  A pass is described by
//   - an anchor group G of <= 5 bit positions (g = the G-bits of the index),
//   - up to 6 chunks of <= 5 bit positions each (v_c = the chunk-c bits),
//   - a table tab[g][c][v] of complex factors,
// and multiplies amplitude i by  prod_c tab[g(i)][c][v_c(i)].  Any product
// of phase terms that each touch at most one qubit outside G factorises
// this way.  Amplitudes whose `required` bits are not all set get factor 1
// and are not touched.  The table lives in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.h"

using namespace aqs_hip;

namespace {

constexpr int kBlock = 256;

struct DiagTArgs {
    void* sv;
    uint64_t units;       // amplitudes enumerated (required bits inserted)
    int nreq;
    int req[5];           // sorted required bit positions (set to 1)
    uint64_t reqmask;
    int ng;
    int gpos[5];
    int nc;
    int cpos[6][5];
    int clen[6];
    // bit groups whose positions are consecutive ascending are read with
    // one shift and mask: start bit, or -1 when the group is scattered
    int gstart;
    int cstart[6];
    const void* tab;      // [2^ng][nc][32] complex (float2 or double2)
};

__device__ __forceinline__ uint32_t group_bits(uint64_t i, int start, const int* pos, int len);

__device__ __forceinline__ uint32_t pext_list(uint64_t i, const int* pos, int len) {
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        if (b >= len) break;
        v |= (uint32_t)((i >> pos[b]) & 1ull) << b;
    }
    return v;
}

__device__ __forceinline__ uint32_t group_bits(uint64_t i, int start, const int* pos, int len) {
    if (start >= 0) return (uint32_t)(i >> start) & ((1u << len) - 1u);
    return pext_list(i, pos, len);
}

template <typename T2, typename R>
__global__ __launch_bounds__(kBlock) void k_diagt(DiagTArgs a) {
    extern __shared__ unsigned char smem[];
    T2* tab = reinterpret_cast<T2*>(smem);
    const int entries = (1 << a.ng) * a.nc * 32;
    const T2* gt = reinterpret_cast<const T2*>(a.tab);
    for (int e = threadIdx.x; e < entries; e += kBlock) tab[e] = gt[e];
    __syncthreads();
    T2* sv = reinterpret_cast<T2*>(a.sv);
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < a.units; u += stride) {
        uint64_t i = u;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            if (j >= a.nreq) break;
            const int p = a.req[j];
            i = ((i >> p) << (p + 1)) | (i & ((1ull << p) - 1ull));
        }
        i |= a.reqmask;
        const uint32_t g = group_bits(i, a.gstart, a.gpos, a.ng);
        R pr = 1, pi = 0;
        for (int c = 0; c < a.nc; ++c) {
            const T2 f = tab[(g * a.nc + c) * 32 + group_bits(i, a.cstart[c], a.cpos[c], a.clen[c])];
            const R nr = pr * f.x - pi * f.y;
            pi = pr * f.y + pi * f.x;
            pr = nr;
        }
        T2 x = sv[i];
        T2 y;
        y.x = pr * x.x - pi * x.y;
        y.y = pr * x.y + pi * x.x;
        sv[i] = y;
    }
}

// Complex64 variant: each lane takes U float4s (two amplitudes each, indices
// 2f and 2f+1) with all loads issued before the table work.  Used when no
// required bit is bit 0, so both halves of a float4 are enumerated together.
constexpr int kDiagU = 4;

__global__ __launch_bounds__(kBlock) void k_diagt_v4(DiagTArgs a) {
    extern __shared__ unsigned char smem[];
    float2* tab = reinterpret_cast<float2*>(smem);
    const int entries = (1 << a.ng) * a.nc * 32;
    const float2* gt = reinterpret_cast<const float2*>(a.tab);
    for (int e = threadIdx.x; e < entries; e += kBlock) tab[e] = gt[e];
    __syncthreads();
    float4* sv = reinterpret_cast<float4*>(a.sv);
    const uint64_t units = a.units >> 1;  // float4s enumerated
    const uint64_t stride = (uint64_t)gridDim.x * kBlock * kDiagU;
    for (uint64_t base = (uint64_t)blockIdx.x * kBlock * kDiagU + threadIdx.x; base < units; base += stride) {
        float4 v[kDiagU];
        uint64_t f[kDiagU];
#pragma unroll
        for (int j = 0; j < kDiagU; ++j) {
            uint64_t u = base + (uint64_t)j * kBlock;
            // required bits (>= 1) in float4-index space, then back to amplitudes
            for (int r = 0; r < a.nreq; ++r) {
                const int p = a.req[r] - 1;
                u = ((u >> p) << (p + 1)) | (u & ((1ull << p) - 1ull));
            }
            f[j] = u | (a.reqmask >> 1);
            if (base + (uint64_t)j * kBlock < units) v[j] = ld_nt(&sv[f[j]]);
        }
#pragma unroll
        for (int j = 0; j < kDiagU; ++j) {
            if (base + (uint64_t)j * kBlock >= units) continue;
            float4 x = v[j], y;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint64_t i = 2 * f[j] + h;
                const uint32_t g = group_bits(i, a.gstart, a.gpos, a.ng);
                float pr = 1.f, pi = 0.f;
                for (int c = 0; c < a.nc; ++c) {
                    const float2 t = tab[(g * a.nc + c) * 32 + group_bits(i, a.cstart[c], a.cpos[c], a.clen[c])];
                    const float nr = pr * t.x - pi * t.y;
                    pi = pr * t.y + pi * t.x;
                    pr = nr;
                }
                const float xr = h ? x.z : x.x, xi = h ? x.w : x.y;
                const float yr = pr * xr - pi * xi, yi = pr * xi + pi * xr;
                if (h) {
                    y.z = yr;
                    y.w = yi;
                } else {
                    y.x = yr;
                    y.y = yi;
                }
            }
            st_nt(y, &sv[f[j]]);
        }
    }
}

}  // namespace

namespace aqs_hip {

// desc: required bits, anchor bits, chunks (all amplitude bit positions);
// dev_tab: the table on the device in the state's precision.
void apply_diag_table(aqs_state s, const std::vector<int>& required, const std::vector<int>& anchor,
                      const std::vector<std::vector<int>>& chunks, const void* dev_tab) {
    if (anchor.size() > 5 || chunks.size() > 6 || required.size() > 5)
        fail(AQS_ERR_INVALID_ARGUMENT, "diagonal table pass too wide");
    DiagTArgs a{};
    a.sv = s->amps;
    std::vector<int> req = required;
    std::sort(req.begin(), req.end());
    a.nreq = (int)req.size();
    for (size_t j = 0; j < req.size(); ++j) {
        a.req[j] = req[j];
        a.reqmask |= 1ull << req[j];
    }
    a.units = s->N >> req.size();
    auto consecutive = [](const std::vector<int>& p) {
        for (size_t j = 1; j < p.size(); ++j)
            if (p[j] != p[0] + (int)j) return -1;
        return p.empty() ? 0 : p[0];
    };
    a.ng = (int)anchor.size();
    for (size_t j = 0; j < anchor.size(); ++j) a.gpos[j] = anchor[j];
    a.gstart = consecutive(anchor);
    a.nc = (int)chunks.size();
    for (size_t c = 0; c < chunks.size(); ++c) {
        if (chunks[c].size() > 5) fail(AQS_ERR_INVALID_ARGUMENT, "diagonal chunk too wide");
        a.clen[c] = (int)chunks[c].size();
        for (size_t j = 0; j < chunks[c].size(); ++j) a.cpos[c][j] = chunks[c][j];
        a.cstart[c] = consecutive(chunks[c]);
    }
    a.tab = dev_tab;
    const size_t entries = ((size_t)1 << a.ng) * std::max(1, a.nc) * 32;
    uint64_t blocks = (a.units + kBlock - 1) / kBlock;
    blocks = std::min<uint64_t>(std::max<uint64_t>(blocks, 1), 8192);
    ProfScope ps(s, kSlotDiagTable);
    const bool bit0_required = !req.empty() && req[0] == 0;
    if (s->precision == AQS_C64 && !bit0_required && s->n >= 2) {
        const size_t lds = entries * sizeof(float2);
        uint64_t b4 = (a.units / 2 + kBlock * kDiagU - 1) / (kBlock * kDiagU);
        b4 = std::min<uint64_t>(std::max<uint64_t>(b4, 1), 16384);
        hipLaunchKernelGGL(k_diagt_v4, dim3((unsigned)b4), dim3(kBlock), lds, s->stream, a);
    } else if (s->precision == AQS_C64) {
        const size_t lds = entries * sizeof(float2);
        hipLaunchKernelGGL((k_diagt<float2, float>), dim3((unsigned)blocks), dim3(kBlock), lds, s->stream, a);
    } else {
        const size_t lds = entries * sizeof(double2);
        AQS_HIP(hipFuncSetAttribute((const void*)k_diagt<double2, double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)std::max<size_t>(lds, 65536)));
        hipLaunchKernelGGL((k_diagt<double2, double>), dim3((unsigned)blocks), dim3(kBlock), lds, s->stream, a);
    }
    AQS_HIP(hipGetLastError());
}

}  // namespace aqs_hip
