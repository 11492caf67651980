// Gate-application kernels for gfx950.
//
// The reference applies every gate as a 2^n x 2^n operator: a CSR SpMV for
// the permutation / diagonal gates (e.g. X src/quantum.cpp:540-562, CZ
// :1115-1152) and a dense Kronecker-product GEMV for H / RotX / RotY / RotZ
// (:671-831).  Here each gate is one in-place pass over the amplitudes it
// can change:
//
//   MAT1   2x2 butterfly on (i0, i1 = i0 | target bit)        f = 1 (1/2 per control)
//   DIAG1  per-amplitude phase chosen by the target bit        f = 1
//   PHASE  phase on amplitudes whose target+control bits are 1 f = 1/2^(#bits)
//   X      exchange of the two halves of each pair             f = 1 (Or: 3/4)
//   SWAP   exchange of |..1..0..> and |..0..1..>               f = 1/2
//   MATK   dense 2^k x 2^k on k target bits (k <= 3 here; k = 4, 5 complex64
//          go to the MFMA kernel in mfma.hip)
//
// Complex64 states move two amplitudes per 16-byte access (float4) whenever
// neither the target nor a control is bit 0, so every wavefront touches
// 64 x 16 B contiguous runs; bit-0 targets pair the two halves of one float4
// in registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <exception>

#include "internal.h"

using namespace aqs_hip;

namespace {

constexpr int kBlock = 256;

// One-wave blocks with one unit per lane are the fastest streaming shape;
// they are used while the grid stays within 2^31 threads (and, for 32-bit
// index variants, the float4 index space within 2^32), else 256-thread
// blocks of several units with 64-bit indices.
inline bool one_wave_shape(uint64_t units) { return units <= (1ull << 31); }

// Sorted bit positions that are held fixed while enumerating work units.
struct Fixed {
    int n;
    int pos[8];
};

__device__ __forceinline__ uint64_t insert_zeros(uint64_t t, const Fixed& f) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j >= f.n) break;
        const uint64_t p = (uint64_t)f.pos[j];
        t = ((t >> p) << (p + 1)) | (t & ((1ull << p) - 1ull));
    }
    return t;
}

inline unsigned blocks_for(uint64_t units, int per_thread, int block = kBlock) {
    uint64_t threads = (units + per_thread - 1) / per_thread;
    uint64_t b = (threads + block - 1) / block;
    return (unsigned)std::max<uint64_t>(b, 1);
}

template <typename R>
struct C2 {
    R re, im;
};

template <typename R>
__device__ __forceinline__ void cmul_add(R ar, R ai, R br, R bi, R& accr, R& acci) {
    accr += ar * br - ai * bi;
    acci += ar * bi + ai * br;
}

// ---------------------------------------------------------------------------
// MAT1, complex64, float4 path (target bit >= 1, control bits >= 1).
// Work unit u -> float4 index f (2 amplitudes) with bit (t-1) = 0 and the
// control bits (c-1) = 1; partner float4 at f + 2^(t-1).
struct Mat1V4Args {
    float4* sv;
    uint64_t units;
    Fixed fixed;      // positions in float4-index space
    uint64_t cmask;   // control mask in float4-index space
    uint64_t stride;  // 2^(t-1)
    float m[8];
};

template <typename IT>
__device__ __forceinline__ IT insert_zeros_t(IT t, const Fixed& f) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j >= f.n) break;
        const int p = f.pos[j];
        t = ((t >> p) << (p + 1)) | (t & ((((IT)1) << p) - (IT)1));
    }
    return t;
}

// U independent units per thread, BS threads per block; IT = index type
// (32-bit when the float4 index space fits).  The variant is chosen per
// (qubits, target bit) by aqs_autotune.  One-wave blocks with one unit per
// lane stream fastest for most partner distances (6.5-6.8 TB/s at 28 qubits
// for partners 2^7..2^19 float4 apart, against 5.6-5.9 with 256-thread
// blocks of 4 units); 256-thread blocks win for near and far partners.
template <int U, typename IT, int BS>
__global__ __launch_bounds__(BS) void k_mat1_v4(Mat1V4Args a) {
    const IT units = (IT)a.units, stride = (IT)a.stride, cmask = (IT)a.cmask;
    const IT base = (IT)blockIdx.x * (IT)(BS * U) + (IT)threadIdx.x;
    float4 lo[U], hi[U];
    IT idx[U];
    bool ok[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const IT u = base + (IT)j * BS;
        ok[j] = u < units;
        idx[j] = insert_zeros_t<IT>(u, a.fixed) | cmask;
        if (ok[j]) {
            lo[j] = ld_nt(&a.sv[idx[j]]);
            hi[j] = ld_nt(&a.sv[idx[j] + stride]);
        }
    }
    const float m0r = a.m[0], m0i = a.m[1], m1r = a.m[2], m1i = a.m[3];
    const float m2r = a.m[4], m2i = a.m[5], m3r = a.m[6], m3i = a.m[7];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (!ok[j]) continue;
        float4 x = lo[j], y = hi[j], nx, ny;
        // amplitude pair (x.xy, y.xy)
        nx.x = m0r * x.x - m0i * x.y + m1r * y.x - m1i * y.y;
        nx.y = m0r * x.y + m0i * x.x + m1r * y.y + m1i * y.x;
        ny.x = m2r * x.x - m2i * x.y + m3r * y.x - m3i * y.y;
        ny.y = m2r * x.y + m2i * x.x + m3r * y.y + m3i * y.x;
        // amplitude pair (x.zw, y.zw)
        nx.z = m0r * x.z - m0i * x.w + m1r * y.z - m1i * y.w;
        nx.w = m0r * x.w + m0i * x.z + m1r * y.w + m1i * y.z;
        ny.z = m2r * x.z - m2i * x.w + m3r * y.z - m3i * y.w;
        ny.w = m2r * x.w + m2i * x.z + m3r * y.w + m3i * y.z;
        st_nt(nx, &a.sv[idx[j]]);
        st_nt(ny, &a.sv[idx[j] + stride]);
    }
}

// Uncontrolled MAT1 whose partner float4 is 1..32 float4s away (target
// bits 1..6): every wave streams 64 consecutive float4s and takes each
// partner from lane (lane ^ stride), so loads and stores are plain 1 KiB
// runs instead of interleaved pairs.
template <typename IT>
__global__ __launch_bounds__(64) void k_mat1_lx(Mat1V4Args a) {
    const IT f = (IT)blockIdx.x * (IT)64 + (IT)threadIdx.x;
    const int s = (int)a.stride;
    const float4 x = ld_nt(&a.sv[f]);
    float4 y;
    y.x = __shfl_xor(x.x, s);
    y.y = __shfl_xor(x.y, s);
    y.z = __shfl_xor(x.z, s);
    y.w = __shfl_xor(x.w, s);
    // lower member: m00 x + m01 y; upper member: m10 y + m11 x
    const bool upper = (threadIdx.x & (unsigned)s) != 0u;
    const float ar = upper ? a.m[6] : a.m[0], ai = upper ? a.m[7] : a.m[1];
    const float br = upper ? a.m[4] : a.m[2], bi = upper ? a.m[5] : a.m[3];
    float4 n;
    n.x = ar * x.x - ai * x.y + br * y.x - bi * y.y;
    n.y = ar * x.y + ai * x.x + br * y.y + bi * y.x;
    n.z = ar * x.z - ai * x.w + br * y.z - bi * y.w;
    n.w = ar * x.w + ai * x.z + br * y.w + bi * y.z;
    st_nt(n, &a.sv[f]);
}

// Variant 4's shape (one unit per lane, one-wave workgroups) with
// consecutive workgroups taking their units from F equal parts of the unit
// range in turn, so that the pass advances on F fronts (2F streams with the
// partners) instead of one.
template <int F>
__global__ __launch_bounds__(64) void k_mat1_fronts(Mat1V4Args a) {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    const uint32_t c = (b % F) * (G / F) + b / F;  // a bijection of 0..G-1 (F divides G)
    const uint32_t u = c * 64u + threadIdx.x;
    if (u >= (uint32_t)a.units) return;
    const uint32_t idx = insert_zeros_t<uint32_t>(u, a.fixed) | (uint32_t)a.cmask;
    const uint32_t stride = (uint32_t)a.stride;
    const float4 x = ld_nt(&a.sv[idx]), y = ld_nt(&a.sv[idx + stride]);
    const float m0r = a.m[0], m0i = a.m[1], m1r = a.m[2], m1i = a.m[3];
    const float m2r = a.m[4], m2i = a.m[5], m3r = a.m[6], m3i = a.m[7];
    float4 nx, ny;
    nx.x = m0r * x.x - m0i * x.y + m1r * y.x - m1i * y.y;
    nx.y = m0r * x.y + m0i * x.x + m1r * y.y + m1i * y.x;
    ny.x = m2r * x.x - m2i * x.y + m3r * y.x - m3i * y.y;
    ny.y = m2r * x.y + m2i * x.x + m3r * y.y + m3i * y.x;
    nx.z = m0r * x.z - m0i * x.w + m1r * y.z - m1i * y.w;
    nx.w = m0r * x.w + m0i * x.z + m1r * y.w + m1i * y.z;
    ny.z = m2r * x.z - m2i * x.w + m3r * y.z - m3i * y.w;
    ny.w = m2r * x.w + m2i * x.z + m3r * y.w + m3i * y.z;
    st_nt(nx, &a.sv[idx]);
    st_nt(ny, &a.sv[idx + stride]);
}

// Launch-shape variants of the float4 MAT1 pass: (units per thread, index
// type, block size) = 0 (1, 32-bit, 256), 1 (2, 64-bit, 256), 2 (4, 64-bit,
// 256), 3 (2, 32-bit, 256), 4 (1, 32-bit, 64), 5 (2, 32-bit, 64); 6 = the
// lane-exchange pass (uncontrolled, partner 1..32 float4s away, 2^6..2^31
// float4s; otherwise variant 4 or 1); 7, 8 = variant 4 on 2 and 4 fronts
// (when the fronts divide the grid; otherwise variant 4).
constexpr int kMat1Variants = 9;

void launch_mat1_v4(const Mat1V4Args& a, int variant, uint64_t n_float4, hipStream_t st) {
    const bool fits32 = n_float4 <= (1ull << 32);
    if (variant == 6 && !(a.cmask == 0 && a.fixed.n == 1 && a.stride <= 32 && n_float4 >= 64 &&
                          n_float4 <= (1ull << 31) && 2 * a.units == n_float4))
        variant = 4;
    if ((variant == 7 || variant == 8) && blocks_for(a.units, 1, 64) % (variant == 7 ? 2u : 4u) != 0) variant = 4;
    if (!fits32 && variant != 1 && variant != 2) variant = 1;
    switch (variant) {
        case 0:
            hipLaunchKernelGGL((k_mat1_v4<1, uint32_t, 256>), dim3(blocks_for(a.units, 1, 256)), dim3(256), 0, st, a);
            break;
        case 1:
            hipLaunchKernelGGL((k_mat1_v4<2, uint64_t, 256>), dim3(blocks_for(a.units, 2, 256)), dim3(256), 0, st, a);
            break;
        case 2:
            hipLaunchKernelGGL((k_mat1_v4<4, uint64_t, 256>), dim3(blocks_for(a.units, 4, 256)), dim3(256), 0, st, a);
            break;
        case 3:
            hipLaunchKernelGGL((k_mat1_v4<2, uint32_t, 256>), dim3(blocks_for(a.units, 2, 256)), dim3(256), 0, st, a);
            break;
        case 4:
            hipLaunchKernelGGL((k_mat1_v4<1, uint32_t, 64>), dim3(blocks_for(a.units, 1, 64)), dim3(64), 0, st, a);
            break;
        case 6:
            hipLaunchKernelGGL((k_mat1_lx<uint32_t>), dim3((unsigned)(n_float4 / 64)), dim3(64), 0, st, a);
            break;
        case 7:
            hipLaunchKernelGGL((k_mat1_fronts<2>), dim3(blocks_for(a.units, 1, 64)), dim3(64), 0, st, a);
            break;
        case 8:
            hipLaunchKernelGGL((k_mat1_fronts<4>), dim3(blocks_for(a.units, 1, 64)), dim3(64), 0, st, a);
            break;
        default:
            hipLaunchKernelGGL((k_mat1_v4<2, uint32_t, 64>), dim3(blocks_for(a.units, 2, 64)), dim3(64), 0, st, a);
            break;
    }
}

// Tuned variant per (device, qubits, target bit); -1 = not tuned (variant 0).
std::mutex g_tune_mutex;
int g_tune[8][AQS_MAX_QUBITS + 1][AQS_MAX_QUBITS];
bool g_tune_init = false;

int tuned_variant(int n, int t) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (!g_tune_init) return 0;
    const int d = current_device() & 7;
    const int v = g_tune[d][n][t];
    return v < 0 ? 0 : v;
}

// MAT1, complex64, target bit 0 (controls >= 1): the pair is one float4.
template <int U, int BS>
__global__ __launch_bounds__(BS) void k_mat1_b0(Mat1V4Args a) {
    const uint64_t base = (uint64_t)blockIdx.x * (BS * U) + threadIdx.x;
    float4 v[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t u = base + (uint64_t)j * BS;
        idx[j] = (u < a.units) ? (insert_zeros(u, a.fixed) | a.cmask) : ~0ull;
        if (idx[j] != ~0ull) v[j] = ld_nt(&a.sv[idx[j]]);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (idx[j] == ~0ull) continue;
        float4 x = v[j], n;
        n.x = a.m[0] * x.x - a.m[1] * x.y + a.m[2] * x.z - a.m[3] * x.w;
        n.y = a.m[0] * x.y + a.m[1] * x.x + a.m[2] * x.w + a.m[3] * x.z;
        n.z = a.m[4] * x.x - a.m[5] * x.y + a.m[6] * x.z - a.m[7] * x.w;
        n.w = a.m[4] * x.y + a.m[5] * x.x + a.m[6] * x.w + a.m[7] * x.z;
        st_nt(n, &a.sv[idx[j]]);
    }
}

// ---------------------------------------------------------------------------
// Generic per-amplitude-pair kernels (any bit positions, both precisions).
template <typename T2, typename R>
struct PairArgs {
    T2* sv;
    uint64_t units;
    Fixed fixed;      // amplitude-index space
    uint64_t cmask;   // bits forced to 1
    uint64_t tmask;   // partner offset (target bit, or pa|pb for SWAP)
    uint64_t anymask; // AQS_CTRL_ANY: at least one of these bits must be 1 (0 = unused)
    uint64_t reqmask; // further controls: all of these bits must be 1 (0 = unused)
    R m[8];
};

// op: 0 = MAT1, 1 = X (exchange), 2 = SWAP (exchange), 3 = DIAG1 (per-bit phase)
template <typename T2, typename R, int OP>
__global__ __launch_bounds__(kBlock) void k_pair(PairArgs<T2, R> a) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < a.units; u += stride) {
        const uint64_t i0 = insert_zeros(u, a.fixed) | a.cmask;
        if (a.anymask && !(i0 & a.anymask)) continue;
        if ((i0 & a.reqmask) != a.reqmask) continue;
        const uint64_t i1 = i0 ^ a.tmask;
        T2 x = a.sv[i0], y = a.sv[i1];
        if (OP == 0) {
            T2 nx, ny;
            nx.x = a.m[0] * x.x - a.m[1] * x.y + a.m[2] * y.x - a.m[3] * y.y;
            nx.y = a.m[0] * x.y + a.m[1] * x.x + a.m[2] * y.y + a.m[3] * y.x;
            ny.x = a.m[4] * x.x - a.m[5] * x.y + a.m[6] * y.x - a.m[7] * y.y;
            ny.y = a.m[4] * x.y + a.m[5] * x.x + a.m[6] * y.y + a.m[7] * y.x;
            a.sv[i0] = nx;
            a.sv[i1] = ny;
        } else if (OP == 1 || OP == 2) {
            a.sv[i0] = y;
            a.sv[i1] = x;
        } else {
            T2 nx, ny;
            nx.x = a.m[0] * x.x - a.m[1] * x.y;
            nx.y = a.m[0] * x.y + a.m[1] * x.x;
            ny.x = a.m[6] * y.x - a.m[7] * y.y;
            ny.y = a.m[6] * y.y + a.m[7] * y.x;
            a.sv[i0] = nx;
            a.sv[i1] = ny;
        }
    }
}

// Exchange kernels, complex64 float4 path (all fixed bits >= 1).
struct XchgV4Args {
    float4* sv;
    uint64_t units;
    Fixed fixed;
    uint64_t setmask;  // bits forced to 1 (controls, SWAP's bit a)
    uint64_t partner;  // xor to the partner float4
};

template <int U, int BS>
__global__ __launch_bounds__(BS) void k_xchg_v4(XchgV4Args a) {
    const uint64_t base = (uint64_t)blockIdx.x * (BS * U) + threadIdx.x;
    float4 lo[U], hi[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t u = base + (uint64_t)j * BS;
        idx[j] = (u < a.units) ? (insert_zeros(u, a.fixed) | a.setmask) : ~0ull;
        if (idx[j] != ~0ull) {
            lo[j] = ld_nt(&a.sv[idx[j]]);
            hi[j] = ld_nt(&a.sv[idx[j] ^ a.partner]);
        }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (idx[j] == ~0ull) continue;
        st_nt(hi[j], &a.sv[idx[j]]);
        st_nt(lo[j], &a.sv[idx[j] ^ a.partner]);
    }
}

void launch_xchg_v4(const XchgV4Args& a, hipStream_t st) {
    if (one_wave_shape(a.units))
        hipLaunchKernelGGL((k_xchg_v4<1, 64>), dim3(blocks_for(a.units, 1, 64)), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL((k_xchg_v4<4, 256>), dim3(blocks_for(a.units, 4, 256)), dim3(256), 0, st, a);
}

// X on bit 0 (complex64): swap the halves of each float4 (controls >= 1).
__global__ __launch_bounds__(kBlock) void k_x_b0(XchgV4Args a) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < a.units; u += stride) {
        const uint64_t f = insert_zeros(u, a.fixed) | a.setmask;
        float4 v = a.sv[f];
        a.sv[f] = make_float4(v.z, v.w, v.x, v.y);
    }
}

// Phase on amplitudes with all fixed bits set.
template <typename T2, typename R>
struct PhaseArgs {
    T2* sv;
    uint64_t units;
    Fixed fixed;
    uint64_t setmask;
    uint64_t reqmask;  // generic kernel: further controls, all must be 1 (0 = unused)
    R pr, pi;
};

template <typename T2, typename R>
__global__ __launch_bounds__(kBlock) void k_phase(PhaseArgs<T2, R> a) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < a.units; u += stride) {
        const uint64_t i = insert_zeros(u, a.fixed) | a.setmask;
        if ((i & a.reqmask) != a.reqmask) continue;
        T2 x = a.sv[i], y;
        y.x = a.pr * x.x - a.pi * x.y;
        y.y = a.pr * x.y + a.pi * x.x;
        a.sv[i] = y;
    }
}

// U float4s per lane, BS threads per block; IT = index type (32-bit when
// the float4 index space fits).  One-wave blocks with one float4 per lane
// are the fastest streaming shape.
template <int U, typename IT, int BS>
__global__ __launch_bounds__(BS) void k_phase_v4(PhaseArgs<float4, float> a) {
    const IT units = (IT)a.units, setmask = (IT)a.setmask;
    const IT base = (IT)blockIdx.x * (IT)(BS * U) + (IT)threadIdx.x;
    float4 v[U];
    IT idx[U];
    bool ok[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const IT u = base + (IT)j * BS;
        ok[j] = u < units;
        idx[j] = insert_zeros_t<IT>(u, a.fixed) | setmask;
        if (ok[j]) v[j] = ld_nt(&a.sv[idx[j]]);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (!ok[j]) continue;
        float4 x = v[j], y;
        y.x = a.pr * x.x - a.pi * x.y;
        y.y = a.pr * x.y + a.pi * x.x;
        y.z = a.pr * x.z - a.pi * x.w;
        y.w = a.pr * x.w + a.pi * x.z;
        st_nt(y, &a.sv[idx[j]]);
    }
}

// Phase whose required bits include some of bits 0..3: the amplitudes it
// changes do not fill whole 128-byte lines, so it sweeps every line that
// holds one (bits >= 4 fixed by insertion) and tests the low required bits
// per amplitude.  With bit 3 required, the 64-byte half line without it is
// neither read nor written (CZ on bits 2,3 at 28 qubits: 645 -> 486 us; on
// bits 3,4: 344 -> 291 us).  Skipping 32- or 16-byte pieces instead was
// slower (960 us for bits 0,2): HBM moves 64-byte granules.  For the same
// reason every loaded float4 is stored back, changed or not: storing only
// the changed ones (partial granules) took 685 against 486 us on bits 2,3
// and 972 against 638 us on bits 0,1.
struct PhaseLineArgs {
    float4* sv;
    uint64_t units;
    Fixed fixed;       // required bits >= 4, as float4-index bits
    uint64_t setmask;  // the same bits set (float4 index)
    uint32_t lowmask;  // required bits 0..3 (amplitude index)
    float pr, pi;
};

template <int U, typename IT, int BS>
__global__ __launch_bounds__(BS) void k_phase_lines(PhaseLineArgs a) {
    const IT units = (IT)a.units, setmask = (IT)a.setmask;
    const IT base = (IT)blockIdx.x * (IT)(BS * U) + (IT)threadIdx.x;
    float4 v[U];
    IT idx[U];
    bool need[U];
    const uint32_t gmask = a.lowmask & 8u;  // amplitude bit 3 selects a 64-byte half line
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const IT u = base + (IT)j * BS;
        idx[j] = insert_zeros_t<IT>(u, a.fixed) | setmask;
        const uint32_t i0 = 2u * (uint32_t)(idx[j] & 7u);
        need[j] = u < units && ((i0 | 1u) & gmask) == gmask;
        if (need[j]) v[j] = ld_nt(&a.sv[idx[j]]);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (!need[j]) continue;
        const uint32_t i0 = 2u * (uint32_t)(idx[j] & 7u);
        const bool c0 = (i0 & a.lowmask) == a.lowmask, c1 = ((i0 | 1u) & a.lowmask) == a.lowmask;
        float4 x = v[j], y = x;
        if (c0) {
            y.x = a.pr * x.x - a.pi * x.y;
            y.y = a.pr * x.y + a.pi * x.x;
        }
        if (c1) {
            y.z = a.pr * x.z - a.pi * x.w;
            y.w = a.pr * x.w + a.pi * x.z;
        }
        st_nt(y, &a.sv[idx[j]]);
    }
}

// DIAG1 complex64 float4 path (controls >= 1; any target bit).
struct Diag1V4Args {
    float4* sv;
    uint64_t units;
    Fixed fixed;
    uint64_t cmask;
    int tbit;  // target bit in amplitude space
    float d0r, d0i, d1r, d1i;
};

template <int U, int BS>
__global__ __launch_bounds__(BS) void k_diag1_v4(Diag1V4Args a) {
    const uint64_t base = (uint64_t)blockIdx.x * (BS * U) + threadIdx.x;
    float4 v[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t u = base + (uint64_t)j * BS;
        idx[j] = (u < a.units) ? (insert_zeros(u, a.fixed) | a.cmask) : ~0ull;
        if (idx[j] != ~0ull) v[j] = ld_nt(&a.sv[idx[j]]);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (idx[j] == ~0ull) continue;
        float ar, ai, br, bi;
        if (a.tbit == 0) {
            ar = a.d0r; ai = a.d0i; br = a.d1r; bi = a.d1i;
        } else {
            const bool one = (idx[j] >> (a.tbit - 1)) & 1ull;
            ar = one ? a.d1r : a.d0r;
            ai = one ? a.d1i : a.d0i;
            br = ar;
            bi = ai;
        }
        float4 x = v[j], y;
        y.x = ar * x.x - ai * x.y;
        y.y = ar * x.y + ai * x.x;
        y.z = br * x.z - bi * x.w;
        y.w = br * x.w + bi * x.z;
        st_nt(y, &a.sv[idx[j]]);
    }
}

// ---------------------------------------------------------------------------
// MATK scalar kernel: one thread per group of K = 2^k amplitudes.
template <typename T2, typename R, int K>
struct MatkArgs {
    T2* sv;
    uint64_t units;
    Fixed fixed;
    uint64_t cmask;
    uint64_t reqmask;  // further controls, all must be 1 (0 = unused)
    uint64_t off[K];   // amplitude offset of sub-index r
    const R* u;        // K x K complex, row-major, interleaved
};

template <typename T2, typename R, int K>
__global__ __launch_bounds__(kBlock) void k_matk(MatkArgs<T2, R, K> a) {
    __shared__ R su[2 * K * K];
    for (int i = threadIdx.x; i < 2 * K * K; i += kBlock) su[i] = a.u[i];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < a.units; t += stride) {
        const uint64_t base = insert_zeros(t, a.fixed) | a.cmask;
        if ((base & a.reqmask) != a.reqmask) continue;
        R xr[K], xi[K];
#pragma unroll
        for (int r = 0; r < K; ++r) {
            T2 v = a.sv[base + a.off[r]];
            xr[r] = v.x;
            xi[r] = v.y;
        }
#pragma unroll
        for (int r = 0; r < K; ++r) {
            R accr = 0, acci = 0;
#pragma unroll
            for (int c = 0; c < K; ++c) cmul_add(su[2 * (r * K + c)], su[2 * (r * K + c) + 1], xr[c], xi[c], accr, acci);
            T2 o;
            o.x = accr;
            o.y = acci;
            a.sv[base + a.off[r]] = o;
        }
    }
}

// ---------------------------------------------------------------------------
// Host side.

Fixed make_fixed(std::vector<int> pos) {
    std::sort(pos.begin(), pos.end());
    Fixed f{};
    f.n = (int)pos.size();
    for (int i = 0; i < f.n; ++i) f.pos[i] = pos[i];
    return f;
}

unsigned grid_stride_blocks(uint64_t units) {
    uint64_t b = (units + kBlock - 1) / kBlock;
    if (b > 65536) b = 65536;
    return (unsigned)std::max<uint64_t>(b, 1);
}

template <typename T2, typename R>
void launch_pair(aqs_state s, int op_code, const aqs_op& op, std::vector<int> fixed_pos, uint64_t cmask,
                 uint64_t tmask, uint64_t anymask, uint64_t units, uint64_t reqmask = 0) {
    PairArgs<T2, R> a{};
    a.sv = (T2*)s->amps;
    a.units = units;
    a.fixed = make_fixed(fixed_pos);
    a.cmask = cmask;
    a.tmask = tmask;
    a.anymask = anymask;
    a.reqmask = reqmask;
    for (int i = 0; i < 8; ++i) a.m[i] = (R)op.m[i];
    dim3 g(grid_stride_blocks(units)), b(kBlock);
    switch (op_code) {
        case 0: hipLaunchKernelGGL((k_pair<T2, R, 0>), g, b, 0, s->stream, a); break;
        case 1: hipLaunchKernelGGL((k_pair<T2, R, 1>), g, b, 0, s->stream, a); break;
        case 2: hipLaunchKernelGGL((k_pair<T2, R, 2>), g, b, 0, s->stream, a); break;
        default: hipLaunchKernelGGL((k_pair<T2, R, 3>), g, b, 0, s->stream, a); break;
    }
}

// Amplitude-index bits of an op's further controls (more_controls).
uint64_t more_control_bits(int n, const aqs_op& op) {
    uint64_t r = 0;
    for (int q = 0; q < n; ++q)
        if ((op.more_controls >> q) & 1ull) r |= qbit(n, q);
    return r;
}

template <typename T2, typename R, int K>
void launch_matk(aqs_state s, const aqs_op& op, const void* dev_u) {
    const int n = s->n, k = op.ntargets;
    MatkArgs<T2, R, K> a{};
    a.sv = (T2*)s->amps;
    a.reqmask = more_control_bits(n, op);
    std::vector<int> fp;
    for (int j = 0; j < k; ++j) fp.push_back(qpos(n, op.targets[j]));
    uint64_t cmask = 0;
    for (int j = 0; j < op.ncontrols; ++j) {
        fp.push_back(qpos(n, op.controls[j]));
        cmask |= qbit(n, op.controls[j]);
    }
    a.fixed = make_fixed(fp);
    a.cmask = cmask;
    a.units = s->N >> (k + op.ncontrols);
    for (int r = 0; r < K; ++r) {
        uint64_t o = 0;
        for (int j = 0; j < k; ++j)
            if ((r >> (k - 1 - j)) & 1) o |= qbit(n, op.targets[j]);
        a.off[r] = o;
    }
    a.u = (const R*)dev_u;
    hipLaunchKernelGGL((k_matk<T2, R, K>), dim3(grid_stride_blocks(a.units)), dim3(kBlock), 0, s->stream, a);
}

template <typename T2, typename R>
void dispatch_matk(aqs_state s, const aqs_op& op, const void* dev_u) {
    switch (op.ntargets) {
        case 1: launch_matk<T2, R, 2>(s, op, dev_u); break;
        case 2: launch_matk<T2, R, 4>(s, op, dev_u); break;
        case 3: launch_matk<T2, R, 8>(s, op, dev_u); break;
        case 4: launch_matk<T2, R, 16>(s, op, dev_u); break;
        default: launch_matk<T2, R, 32>(s, op, dev_u); break;
    }
}

bool all_ge1(const std::vector<int>& p) {
    for (int x : p)
        if (x < 1) return false;
    return true;
}

void apply_c64_fast_or_generic(aqs_state s, const aqs_op& op);

}  // namespace

namespace aqs_hip {

void validate_op(const aqs_op& op, int n) {
    auto chk_q = [&](int q) {
        if (q < 0 || q >= n) fail(AQS_ERR_OUT_OF_RANGE, "Cannot add gate at the given qubit position");
    };
    if (op.ntargets < 1 || op.ntargets > 5) fail(AQS_ERR_INVALID_ARGUMENT, "op target count must be 1..5");
    if (op.ncontrols < 0 || op.ncontrols > 3) fail(AQS_ERR_INVALID_ARGUMENT, "op control count must be 0..3");
    switch (op.kind) {
        case AQS_OP_MAT1:
        case AQS_OP_DIAG1:
        case AQS_OP_X:
            if (op.ntargets != 1) fail(AQS_ERR_INVALID_ARGUMENT, "one-qubit op needs exactly one target");
            break;
        case AQS_OP_SWAP:
            if (op.ntargets != 2) fail(AQS_ERR_INVALID_ARGUMENT, "swap needs two targets");
            break;
        case AQS_OP_MATK:
            if (!op.u) fail(AQS_ERR_INVALID_ARGUMENT, "MATK op without matrix");
            break;
        case AQS_OP_PHASE:
            break;
        default:
            fail(AQS_ERR_INVALID_ARGUMENT, "unknown op kind " + std::to_string(op.kind));
    }
    uint64_t seen = 0;
    for (int j = 0; j < op.ntargets; ++j) {
        chk_q(op.targets[j]);
        uint64_t b = 1ull << op.targets[j];
        if (seen & b) {
            if (op.kind == AQS_OP_SWAP) fail(AQS_ERR_INVALID_ARGUMENT, "Cannot use the swap gate on the same target qubits");
            fail(AQS_ERR_INVALID_ARGUMENT, "repeated target qubit");
        }
        seen |= b;
    }
    uint64_t cseen = 0;
    for (int j = 0; j < op.ncontrols; ++j) {
        chk_q(op.controls[j]);
        uint64_t b = 1ull << op.controls[j];
        if (seen & b) fail(AQS_ERR_INVALID_ARGUMENT, "Control qubit cannot be the same as the target qubit");
        if (cseen & b) fail(AQS_ERR_INVALID_ARGUMENT, "repeated control qubit");
        cseen |= b;
    }
    if (op.more_controls) {
        if (n < 64 && (op.more_controls >> n)) fail(AQS_ERR_OUT_OF_RANGE, "Cannot add gate at the given qubit position");
        if (op.more_controls & seen) fail(AQS_ERR_INVALID_ARGUMENT, "Control qubit cannot be the same as the target qubit");
        if (op.more_controls & cseen) fail(AQS_ERR_INVALID_ARGUMENT, "repeated control qubit");
        if (op.flags & AQS_CTRL_ANY) fail(AQS_ERR_INVALID_ARGUMENT, "further controls cannot take AQS_CTRL_ANY");
    }
}

}  // namespace aqs_hip

namespace {

// Collects fixed positions / masks for an op (controls de-duplicated: the
// reference's CCNot accepts control A == control B, src/quantum.cpp:1524-1540).
struct OpBits {
    std::vector<int> cpos;
    uint64_t cmask = 0;
};

OpBits control_bits(int n, const aqs_op& op) {
    OpBits b;
    for (int j = 0; j < op.ncontrols; ++j) {
        uint64_t m = qbit(n, op.controls[j]);
        if (b.cmask & m) continue;
        b.cmask |= m;
        b.cpos.push_back(qpos(n, op.controls[j]));
    }
    return b;
}

void apply_c64_fast_or_generic(aqs_state s, const aqs_op& op) {
    const int n = s->n;
    OpBits cb = control_bits(n, op);
    const bool any = (op.flags & AQS_CTRL_ANY) && op.ncontrols > 0;
    float mf[8];
    for (int i = 0; i < 8; ++i) mf[i] = (float)op.m[i];

    switch (op.kind) {
        case AQS_OP_MAT1: {
            ProfScope ps(s, AQS_OP_MAT1);
            const int t = qpos(n, op.targets[0]);
            if (!any && all_ge1(cb.cpos) && n >= 2) {
                Mat1V4Args a{};
                a.sv = (float4*)s->amps;
                std::vector<int> fp;
                for (int p : cb.cpos) fp.push_back(p - 1);
                a.cmask = cb.cmask >> 1;
                for (int i = 0; i < 8; ++i) a.m[i] = mf[i];
                if (t >= 1) {
                    fp.push_back(t - 1);
                    a.stride = 1ull << (t - 1);
                    a.units = s->N >> (2 + cb.cpos.size());
                    a.fixed = make_fixed(fp);
                    launch_mat1_v4(a, tuned_variant(n, t), s->N / 2, s->stream);
                } else {
                    a.units = s->N >> (1 + cb.cpos.size());
                    a.fixed = make_fixed(fp);
                    if (one_wave_shape(a.units))
                        hipLaunchKernelGGL((k_mat1_b0<1, 64>), dim3(blocks_for(a.units, 1, 64)), dim3(64), 0,
                                           s->stream, a);
                    else
                        hipLaunchKernelGGL((k_mat1_b0<4, 256>), dim3(blocks_for(a.units, 4, 256)), dim3(256), 0,
                                           s->stream, a);
                }
            } else {
                std::vector<int> fp = any ? std::vector<int>{} : cb.cpos;
                fp.push_back(t);
                uint64_t units = s->N >> fp.size();
                launch_pair<float2, float>(s, 0, op, fp, any ? 0 : cb.cmask, 1ull << t, any ? cb.cmask : 0, units);
            }
            break;
        }
        case AQS_OP_DIAG1: {
            ProfScope ps(s, AQS_OP_DIAG1);
            const int t = qpos(n, op.targets[0]);
            if (!any && all_ge1(cb.cpos) && n >= 2) {
                Diag1V4Args a{};
                a.sv = (float4*)s->amps;
                std::vector<int> fp;
                for (int p : cb.cpos) fp.push_back(p - 1);
                a.fixed = make_fixed(fp);
                a.cmask = cb.cmask >> 1;
                a.units = s->N >> (1 + cb.cpos.size());
                a.tbit = t;
                a.d0r = mf[0]; a.d0i = mf[1]; a.d1r = mf[6]; a.d1i = mf[7];
                if (one_wave_shape(a.units))
                    hipLaunchKernelGGL((k_diag1_v4<1, 64>), dim3(blocks_for(a.units, 1, 64)), dim3(64), 0, s->stream, a);
                else
                    hipLaunchKernelGGL((k_diag1_v4<4, 256>), dim3(blocks_for(a.units, 4, 256)), dim3(256), 0,
                                       s->stream, a);
            } else {
                std::vector<int> fp = any ? std::vector<int>{} : cb.cpos;
                fp.push_back(t);
                launch_pair<float2, float>(s, 3, op, fp, any ? 0 : cb.cmask, 1ull << t, any ? cb.cmask : 0,
                                           s->N >> fp.size());
            }
            break;
        }
        case AQS_OP_PHASE: {
            ProfScope ps(s, AQS_OP_PHASE);
            std::vector<int> fp = cb.cpos;
            uint64_t mask = cb.cmask;
            for (int j = 0; j < op.ntargets; ++j) {
                uint64_t m = qbit(n, op.targets[j]);
                if (mask & m) continue;
                mask |= m;
                fp.push_back(qpos(n, op.targets[j]));
            }
            uint32_t lowmask = 0;
            std::vector<int> high;
            for (int p : fp) {
                if (p < 4)
                    lowmask |= 1u << p;
                else
                    high.push_back(p - 1);
            }
            if (lowmask && n >= 4) {
                PhaseLineArgs a{};
                a.sv = (float4*)s->amps;
                a.fixed = make_fixed(high);
                a.setmask = (mask >> 1) & ~(uint64_t)7;
                a.lowmask = lowmask;
                a.units = s->N >> (1 + high.size());
                a.pr = mf[0];
                a.pi = mf[1];
                if (s->N / 2 <= (1ull << 32) && one_wave_shape(a.units))
                    hipLaunchKernelGGL((k_phase_lines<1, uint32_t, 64>), dim3(blocks_for(a.units, 1, 64)), dim3(64),
                                       0, s->stream, a);
                else
                    hipLaunchKernelGGL((k_phase_lines<4, uint64_t, 256>), dim3(blocks_for(a.units, 4, 256)),
                                       dim3(256), 0, s->stream, a);
            } else if (all_ge1(fp) && n >= 2) {
                PhaseArgs<float4, float> a{};
                a.sv = (float4*)s->amps;
                std::vector<int> f4;
                for (int p : fp) f4.push_back(p - 1);
                a.fixed = make_fixed(f4);
                a.setmask = mask >> 1;
                a.units = s->N >> (1 + fp.size());
                a.pr = mf[0];
                a.pi = mf[1];
                if (s->N / 2 <= (1ull << 32) && one_wave_shape(a.units))
                    hipLaunchKernelGGL((k_phase_v4<1, uint32_t, 64>), dim3(blocks_for(a.units, 1, 64)), dim3(64), 0,
                                       s->stream, a);
                else
                    hipLaunchKernelGGL((k_phase_v4<4, uint64_t, 256>), dim3(blocks_for(a.units, 4, 256)), dim3(256),
                                       0, s->stream, a);
            } else {
                PhaseArgs<float2, float> a{};
                a.sv = (float2*)s->amps;
                a.fixed = make_fixed(fp);
                a.setmask = mask;
                a.units = s->N >> fp.size();
                a.pr = mf[0];
                a.pi = mf[1];
                hipLaunchKernelGGL((k_phase<float2, float>), dim3(grid_stride_blocks(a.units)), dim3(kBlock), 0,
                                   s->stream, a);
            }
            break;
        }
        case AQS_OP_X: {
            ProfScope ps(s, AQS_OP_X);
            const int t = qpos(n, op.targets[0]);
            if (!any && all_ge1(cb.cpos) && n >= 2) {
                XchgV4Args a{};
                a.sv = (float4*)s->amps;
                std::vector<int> fp;
                for (int p : cb.cpos) fp.push_back(p - 1);
                a.setmask = cb.cmask >> 1;
                if (t >= 1) {
                    fp.push_back(t - 1);
                    a.fixed = make_fixed(fp);
                    a.partner = 1ull << (t - 1);
                    a.units = s->N >> (2 + cb.cpos.size());
                    launch_xchg_v4(a, s->stream);
                } else {
                    a.fixed = make_fixed(fp);
                    a.units = s->N >> (1 + cb.cpos.size());
                    hipLaunchKernelGGL(k_x_b0, dim3(grid_stride_blocks(a.units)), dim3(kBlock), 0, s->stream, a);
                }
            } else {
                std::vector<int> fp = any ? std::vector<int>{} : cb.cpos;
                fp.push_back(t);
                launch_pair<float2, float>(s, 1, op, fp, any ? 0 : cb.cmask, 1ull << t, any ? cb.cmask : 0,
                                           s->N >> fp.size());
            }
            break;
        }
        case AQS_OP_SWAP: {
            ProfScope ps(s, AQS_OP_SWAP);
            const int pa = qpos(n, op.targets[0]), pb = qpos(n, op.targets[1]);
            std::vector<int> fp = any ? std::vector<int>{} : cb.cpos;
            fp.push_back(pa);
            fp.push_back(pb);
            const uint64_t set = (any ? 0 : cb.cmask) | (1ull << pa);
            const uint64_t partner = (1ull << pa) | (1ull << pb);
            if (!any && all_ge1(fp) && n >= 3) {
                XchgV4Args a{};
                a.sv = (float4*)s->amps;
                std::vector<int> f4;
                for (int p : fp) f4.push_back(p - 1);
                a.fixed = make_fixed(f4);
                a.setmask = set >> 1;
                a.partner = partner >> 1;
                a.units = s->N >> (1 + fp.size());
                launch_xchg_v4(a, s->stream);
            } else {
                launch_pair<float2, float>(s, 2, op, fp, set, partner, any ? cb.cmask : 0, s->N >> fp.size());
            }
            break;
        }
        default:
            break;
    }
    AQS_HIP(hipGetLastError());
}

void apply_c128(aqs_state s, const aqs_op& op) {
    const int n = s->n;
    OpBits cb = control_bits(n, op);
    const bool any = (op.flags & AQS_CTRL_ANY) && op.ncontrols > 0;
    ProfScope ps(s, op.kind);
    switch (op.kind) {
        case AQS_OP_MAT1:
        case AQS_OP_DIAG1:
        case AQS_OP_X: {
            const int t = qpos(n, op.targets[0]);
            std::vector<int> fp = any ? std::vector<int>{} : cb.cpos;
            fp.push_back(t);
            int code = op.kind == AQS_OP_MAT1 ? 0 : (op.kind == AQS_OP_X ? 1 : 3);
            launch_pair<double2, double>(s, code, op, fp, any ? 0 : cb.cmask, 1ull << t, any ? cb.cmask : 0,
                                         s->N >> fp.size());
            break;
        }
        case AQS_OP_SWAP: {
            const int pa = qpos(n, op.targets[0]), pb = qpos(n, op.targets[1]);
            std::vector<int> fp = any ? std::vector<int>{} : cb.cpos;
            fp.push_back(pa);
            fp.push_back(pb);
            launch_pair<double2, double>(s, 2, op, fp, (any ? 0 : cb.cmask) | (1ull << pa), (1ull << pa) | (1ull << pb),
                                         any ? cb.cmask : 0, s->N >> fp.size());
            break;
        }
        case AQS_OP_PHASE: {
            std::vector<int> fp = cb.cpos;
            uint64_t mask = cb.cmask;
            for (int j = 0; j < op.ntargets; ++j) {
                uint64_t m = qbit(n, op.targets[j]);
                if (mask & m) continue;
                mask |= m;
                fp.push_back(qpos(n, op.targets[j]));
            }
            PhaseArgs<double2, double> a{};
            a.sv = (double2*)s->amps;
            a.fixed = make_fixed(fp);
            a.setmask = mask;
            a.units = s->N >> fp.size();
            a.pr = op.m[0];
            a.pi = op.m[1];
            hipLaunchKernelGGL((k_phase<double2, double>), dim3(grid_stride_blocks(a.units)), dim3(kBlock), 0,
                               s->stream, a);
            break;
        }
        default:
            break;
    }
    AQS_HIP(hipGetLastError());
}

// Ops with further controls (more_controls): the generic kernels over the
// listed controls and targets, skipping indices that lack a further control.
template <typename T2, typename R>
void apply_masked_t(aqs_state s, const aqs_op& op) {
    const int n = s->n;
    OpBits cb = control_bits(n, op);
    const uint64_t req = more_control_bits(n, op);
    ProfScope ps(s, op.kind);
    switch (op.kind) {
        case AQS_OP_MAT1:
        case AQS_OP_DIAG1:
        case AQS_OP_X: {
            const int t = qpos(n, op.targets[0]);
            std::vector<int> fp = cb.cpos;
            fp.push_back(t);
            const int code = op.kind == AQS_OP_MAT1 ? 0 : (op.kind == AQS_OP_X ? 1 : 3);
            launch_pair<T2, R>(s, code, op, fp, cb.cmask, 1ull << t, 0, s->N >> fp.size(), req);
            break;
        }
        case AQS_OP_SWAP: {
            const int pa = qpos(n, op.targets[0]), pb = qpos(n, op.targets[1]);
            std::vector<int> fp = cb.cpos;
            fp.push_back(pa);
            fp.push_back(pb);
            launch_pair<T2, R>(s, 2, op, fp, cb.cmask | (1ull << pa), (1ull << pa) | (1ull << pb), 0,
                               s->N >> fp.size(), req);
            break;
        }
        case AQS_OP_PHASE: {
            std::vector<int> fp = cb.cpos;
            uint64_t mask = cb.cmask;
            for (int j = 0; j < op.ntargets; ++j) {
                const uint64_t m = qbit(n, op.targets[j]);
                if (mask & m) continue;
                mask |= m;
                fp.push_back(qpos(n, op.targets[j]));
            }
            PhaseArgs<T2, R> a{};
            a.sv = (T2*)s->amps;
            a.fixed = make_fixed(fp);
            a.setmask = mask;
            a.reqmask = req;
            a.units = s->N >> fp.size();
            a.pr = (R)op.m[0];
            a.pi = (R)op.m[1];
            hipLaunchKernelGGL((k_phase<T2, R>), dim3(grid_stride_blocks(a.units)), dim3(kBlock), 0, s->stream, a);
            break;
        }
        default:
            break;
    }
    AQS_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Small states (at most 12 qubits in complex64, 11 in complex128): one
// workgroup loads the whole state into LDS, applies a run of up to CH ops
// with a barrier between ops, and stores it back.  A 10-qubit Grover search
// emits about 1560 ops; one launch per op made it launch-bound (about 3 us
// per op), one launch per CH ops is not.
constexpr int kSmallQubitsC64 = 12;
constexpr int kSmallQubitsC128 = 11;
constexpr int kSmallBlock = 512;

// One op in amplitude-index bit masks (n <= 12).  Units are the N/2
// indices with a zero inserted at bit `lo`; i0 = that index, i1 = i0 ^ flip.
//   MAT1 / X / DIAG1: lo = flip = target; the pair (i0, i1) is updated.
//   SWAP: lo = bit b, `allmask` also holds bit a; (i0, i1) are exchanged.
//   PHASE: lo = flip = one required bit; i1 is multiplied when every bit of
//   `allmask` (targets, controls, further controls) is set in it.
// Controls: every bit of `allmask` set in i0 (PHASE: in i1) and, when
// `anymask` is nonzero (AQS_CTRL_ANY), at least one of its bits.
template <typename R>
struct SmallOp {
    uint32_t code;  // 0 MAT1, 1 X, 2 SWAP, 3 DIAG1, 4 PHASE
    uint32_t lo, flip, allmask, anymask, pad;
    R m[8];
};

template <typename R, int CH>
struct SmallRunArgs {
    void* sv;
    uint32_t n, count;
    SmallOp<R> ops[CH];
};

// 64 complex64 ops (3.6 KiB) or 40 complex128 ops (3.5 KiB) per launch keep
// the arguments under 4 KiB.
template <typename R>
constexpr int small_chunk() { return sizeof(R) == 4 ? 64 : 40; }

template <typename T2, typename R, int CH>
__global__ __launch_bounds__(kSmallBlock) void k_small_run(SmallRunArgs<R, CH> a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    T2* S = reinterpret_cast<T2*>(smem);
    T2* sv = reinterpret_cast<T2*>(a.sv);
    const uint32_t N = 1u << a.n, half = N >> 1;
    for (uint32_t i = threadIdx.x; i < N; i += kSmallBlock) S[i] = sv[i];
    __syncthreads();
    for (uint32_t k = 0; k < a.count; ++k) {
        const SmallOp<R>& op = a.ops[k];
        const uint32_t lo = op.lo, low = (1u << lo) - 1u;
        for (uint32_t u = threadIdx.x; u < half; u += kSmallBlock) {
            const uint32_t i0 = ((u & ~low) << 1) | (u & low);
            const uint32_t i1 = i0 ^ op.flip;
            const uint32_t c = op.code == 4 ? i1 : i0;
            if ((c & op.allmask) != op.allmask) continue;
            if (op.anymask && !(c & op.anymask)) continue;
            if (op.code == 4) {
                const T2 y = S[i1];
                T2 ny;
                ny.x = op.m[0] * y.x - op.m[1] * y.y;
                ny.y = op.m[0] * y.y + op.m[1] * y.x;
                S[i1] = ny;
                continue;
            }
            const T2 x = S[i0], y = S[i1];
            if (op.code == 1 || op.code == 2) {
                S[i0] = y;
                S[i1] = x;
            } else if (op.code == 0) {
                T2 nx, ny;
                nx.x = op.m[0] * x.x - op.m[1] * x.y + op.m[2] * y.x - op.m[3] * y.y;
                nx.y = op.m[0] * x.y + op.m[1] * x.x + op.m[2] * y.y + op.m[3] * y.x;
                ny.x = op.m[4] * x.x - op.m[5] * x.y + op.m[6] * y.x - op.m[7] * y.y;
                ny.y = op.m[4] * x.y + op.m[5] * x.x + op.m[6] * y.y + op.m[7] * y.x;
                S[i0] = nx;
                S[i1] = ny;
            } else {
                T2 nx, ny;
                nx.x = op.m[0] * x.x - op.m[1] * x.y;
                nx.y = op.m[0] * x.y + op.m[1] * x.x;
                ny.x = op.m[6] * y.x - op.m[7] * y.y;
                ny.y = op.m[6] * y.y + op.m[7] * y.x;
                S[i0] = nx;
                S[i1] = ny;
            }
        }
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < N; i += kSmallBlock) sv[i] = S[i];
}

// An op (validated, not MATK) in the small kernel's form.
template <typename R>
SmallOp<R> small_op(int n, const aqs_op& op) {
    SmallOp<R> o{};
    // PHASE takes its controls as required bits whatever the flag, as the
    // per-op and batch kernels do
    const bool any = (op.flags & AQS_CTRL_ANY) && op.ncontrols > 0 && op.kind != AQS_OP_PHASE;
    uint32_t ctrl = 0;
    for (int j = 0; j < op.ncontrols; ++j) ctrl |= (uint32_t)qbit(n, op.controls[j]);
    const uint32_t req = (uint32_t)more_control_bits(n, op);
    o.allmask = (any ? 0u : ctrl) | req;
    o.anymask = any ? ctrl : 0u;
    for (int i = 0; i < 8; ++i) o.m[i] = (R)op.m[i];
    switch (op.kind) {
        case AQS_OP_MAT1:
        case AQS_OP_X:
        case AQS_OP_DIAG1: {
            const int t = qpos(n, op.targets[0]);
            o.code = op.kind == AQS_OP_MAT1 ? 0u : (op.kind == AQS_OP_X ? 1u : 3u);
            o.lo = (uint32_t)t;
            o.flip = 1u << t;
            break;
        }
        case AQS_OP_SWAP: {
            const int pa = qpos(n, op.targets[0]), pb = qpos(n, op.targets[1]);
            o.code = 2;
            o.lo = (uint32_t)pb;
            o.flip = (1u << pa) | (1u << pb);
            o.allmask |= 1u << pa;
            break;
        }
        default: {  // AQS_OP_PHASE
            const int p = qpos(n, op.targets[0]);
            o.code = 4;
            o.lo = (uint32_t)p;
            o.flip = 1u << p;
            for (int j = 0; j < op.ntargets; ++j) o.allmask |= (uint32_t)qbit(n, op.targets[j]);
            break;
        }
    }
    return o;
}

template <typename T2, typename R>
void launch_small_run(aqs_state s, const aqs_op* ops, size_t count) {
    constexpr int CH = small_chunk<R>();
    const size_t lds = s->N * sizeof(T2);
    for (size_t b = 0; b < count; b += CH) {
        SmallRunArgs<R, CH> a{};
        a.sv = s->amps;
        a.n = (uint32_t)s->n;
        a.count = (uint32_t)std::min<size_t>(CH, count - b);
        for (uint32_t k = 0; k < a.count; ++k) a.ops[k] = small_op<R>(s->n, ops[b + k]);
        ProfScope ps(s, kSlotSmall);
        hipLaunchKernelGGL((k_small_run<T2, R, CH>), dim3(1), dim3(kSmallBlock), lds, s->stream, a);
        AQS_HIP(hipGetLastError());
    }
}

}  // namespace

namespace aqs_hip {

void apply_op(aqs_state s, const aqs_op& op) {
    validate_op(op, s->n);
    if (op.kind == AQS_OP_MATK) {
        const int k = op.ntargets;
        const size_t K = (size_t)1 << k;
        if (op.flags & AQS_CTRL_ANY) fail(AQS_ERR_INVALID_ARGUMENT, "MATK does not take AQS_CTRL_ANY");
        if (s->precision == AQS_C64) {
            if (k >= 4 && op.ncontrols == 0 && !op.more_controls) {
                std::vector<float> pk = pack_mfma_operand(k, op.u);
                ensure_staging(s, pk.size() * sizeof(float));
                AQS_HIP(hipMemcpyAsync(s->staging, pk.data(), pk.size() * sizeof(float), hipMemcpyHostToDevice,
                                       s->stream));
                if (apply_matk_mfma(s, op, (const float*)s->staging)) return;
            }
            std::vector<float> uf(2 * K * K);
            for (size_t i = 0; i < uf.size(); ++i) uf[i] = (float)op.u[i];
            ensure_staging(s, uf.size() * sizeof(float));
            AQS_HIP(hipMemcpyAsync(s->staging, uf.data(), uf.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
            ProfScope ps(s, AQS_OP_MATK);
            dispatch_matk<float2, float>(s, op, s->staging);
        } else {
            ensure_staging(s, 2 * K * K * sizeof(double));
            AQS_HIP(hipMemcpyAsync(s->staging, op.u, 2 * K * K * sizeof(double), hipMemcpyHostToDevice, s->stream));
            ProfScope ps(s, AQS_OP_MATK);
            dispatch_matk<double2, double>(s, op, s->staging);
        }
        AQS_HIP(hipGetLastError());
        return;
    }
    if (op.more_controls) {
        if (s->precision == AQS_C64)
            apply_masked_t<float2, float>(s, op);
        else
            apply_masked_t<double2, double>(s, op);
        return;
    }
    if (s->precision == AQS_C64)
        apply_c64_fast_or_generic(s, op);
    else
        apply_c128(s, op);
}

// Apply a MATK whose matrix is already on the device (FP32 or FP64 per the
// state's precision, plain row-major layout).
void apply_matk_device(aqs_state s, const aqs_op& op, const void* dev_u) {
    ProfScope ps(s, AQS_OP_MATK);
    if (s->precision == AQS_C64)
        dispatch_matk<float2, float>(s, op, dev_u);
    else
        dispatch_matk<double2, double>(s, op, dev_u);
    AQS_HIP(hipGetLastError());
}

}  // namespace aqs_hip

extern "C" aqs_status aqs_apply(aqs_state s, const aqs_op* ops, size_t count) {
    return guard([&] {
        if (!s || !s->amps) fail(AQS_ERR_INVALID_ARGUMENT, "null state handle");
        if (count && !ops) fail(AQS_ERR_INVALID_ARGUMENT, "null op array");
        const bool small = s->n <= (s->precision == AQS_C64 ? kSmallQubitsC64 : kSmallQubitsC128);
        for (size_t i = 0; i < count;) {
            if (small && ops[i].kind != AQS_OP_MATK) {
                // a run of two or more ops goes through the LDS-resident kernel
                // (ops before an invalid one are applied, as on the per-op path)
                size_t e = i;
                std::exception_ptr err;
                while (e < count && ops[e].kind != AQS_OP_MATK) {
                    try {
                        validate_op(ops[e], s->n);
                    } catch (...) {
                        err = std::current_exception();
                        break;
                    }
                    ++e;
                }
                if (e - i >= 2) {
                    if (s->precision == AQS_C64)
                        launch_small_run<float2, float>(s, ops + i, e - i);
                    else
                        launch_small_run<double2, double>(s, ops + i, e - i);
                    if (err) std::rethrow_exception(err);
                    i = e;
                    continue;
                }
            }
            // A staged MATK matrix must not be overwritten while a previous
            // kernel may still read it.
            if (ops[i].kind == AQS_OP_MATK && i > 0) AQS_HIP(hipStreamSynchronize(s->stream));
            apply_op(s, ops[i]);
            ++i;
        }
    });
}

// Launch-shape autotuner for the one-qubit pass: for every target bit of the
// state's size, times each variant with the identity matrix (the state is
// left unchanged) and keeps the fastest for that (qubits, target bit).
extern "C" aqs_status aqs_autotune(aqs_state s, int reps, int* variant_by_target) {
    return guard([&] {
        if (!s || !s->amps) fail(AQS_ERR_INVALID_ARGUMENT, "null state handle");
        if (s->precision != AQS_C64) fail(AQS_ERR_INVALID_ARGUMENT, "autotune runs on a complex64 state");
        const int n = s->n;
        if (reps < 1) reps = 3;
        {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            if (!g_tune_init) {
                for (auto& d : g_tune)
                    for (auto& q : d)
                        for (auto& t : q) t = -1;
                g_tune_init = true;
            }
        }
        hipEvent_t e0, e1;
        AQS_HIP(hipEventCreate(&e0));
        AQS_HIP(hipEventCreate(&e1));
        const int d = current_device() & 7;
        for (int t = 0; t < n; ++t) {
            int best = 0;
            if (t >= 1 && n >= 2) {
                Mat1V4Args a{};
                a.sv = (float4*)s->amps;
                a.fixed = make_fixed({t - 1});
                a.stride = 1ull << (t - 1);
                a.units = s->N >> 2;
                a.m[0] = 1.f;
                a.m[6] = 1.f;
                float best_ms = 1e30f;
                for (int v = 0; v < kMat1Variants; ++v) {
                    launch_mat1_v4(a, v, s->N / 2, s->stream);
                    AQS_HIP(hipEventRecord(e0, s->stream));
                    for (int r = 0; r < reps; ++r) launch_mat1_v4(a, v, s->N / 2, s->stream);
                    AQS_HIP(hipEventRecord(e1, s->stream));
                    AQS_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    AQS_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = v;
                    }
                }
                AQS_HIP(hipGetLastError());
            }
            {
                std::lock_guard<std::mutex> lock(g_tune_mutex);
                g_tune[d][n][t] = best;
            }
            // qubit q has target bit n-1-q
            if (variant_by_target) variant_by_target[n - 1 - t] = best;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

// Pins the one-qubit launch shape for this state's size and one target
// qubit on the current device (the choice aqs_autotune would store).
extern "C" aqs_status aqs_tune_set(aqs_state s, int qubit, int variant) {
    return guard([&] {
        if (!s || !s->amps) fail(AQS_ERR_INVALID_ARGUMENT, "null state handle");
        if (qubit < 0 || qubit >= s->n) fail(AQS_ERR_OUT_OF_RANGE, "qubit index out of range");
        if (variant < 0 || variant >= kMat1Variants) fail(AQS_ERR_INVALID_ARGUMENT, "no such launch-shape variant");
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        if (!g_tune_init) {
            for (auto& d : g_tune)
                for (auto& q : d)
                    for (auto& t : q) t = -1;
            g_tune_init = true;
        }
        g_tune[current_device() & 7][s->n][s->n - 1 - qubit] = variant;
    });
}
