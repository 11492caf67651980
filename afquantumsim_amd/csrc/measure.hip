// Measurement: FP64 reductions, exact fixed-point sampling, collapse.
//
// The reference forms |a|^2 in FP32 and reduces / scans it with ArrayFire
// (af::sum, af::accum: src/quantum.cpp:349-356,372-391,467-501); sampling
// there tiles the CDF against every uniform, O(shots * 2^n).  Here:
//   - norms, P(q=1) and <sum Z> are FP64 block reductions with a fixed grid,
//     so results are run-to-run identical;
//   - sampling quantises |a|^2 (formed in FP64) to 2^-60 fixed point and
//     builds the cumulative sum exactly in uint64: per-1024-amplitude block
//     sums, one exact scan over the blocks, then per shot a binary search
//     over block prefixes and a wavefront scan inside the chosen block.
//     Exact integer sums make the sampled index independent of summation
//     order, so a CPU restatement reproduces it bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

using namespace aqs_hip;

namespace {

constexpr int kBlock = 256;
static_assert(kBlock == 256, "the block reductions combine four waves");
constexpr int kRedGrid = 1024;       // fixed reduction grid (deterministic order)
constexpr int kSampleBlock = 1024;   // amplitudes per sampling block
constexpr double kFixScale = 1152921504606846976.0;  // 2^60

// Device scratch of aqs_sample_noalloc: the sampling tables of a state of up
// to AQS_NOALLOC_MAX_QUBITS qubits (3 words per 1024 amplitudes, 3 per 2^20
// amplitudes and two), 24 MiB + 24 KiB + 256 B, held by every device the
// library runs on until the process ends.  Host code never uses the array's name as an address: it uses the
// per-device pointer from hipGetSymbolAddress (noalloc_scratch_ptr).
constexpr size_t kNoallocScratchBytes =
    (size_t)3 * ((1ull << AQS_NOALLOC_MAX_QUBITS) / kSampleBlock) * 8 +
    (size_t)3 * ((1ull << AQS_NOALLOC_MAX_QUBITS) / kSampleBlock / 1024) * 8 + 256;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
void* noalloc_scratch_ptr() {
    int dev = 0;
    AQS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(AQS_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        AQS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(AQS_ERR_DEVICE, "hipGetSymbolAddress returned no address for the sampling scratch");
        g_noalloc_ptr[dev] = p;
    }
    return g_noalloc_ptr[dev];
}

template <typename T2>
__device__ __forceinline__ double prob_of(const T2& v) {
    const double re = (double)v.x, im = (double)v.y;
    return fma(re, re, im * im);
}

__device__ __forceinline__ uint64_t quantise(double p) {
    return (uint64_t)rint(p * kFixScale);
}

template <typename T>
__device__ T block_reduce_sum(T v, T* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    T r = sh[0];
    __syncthreads();
    return r;
}

// Sums over the 64 lanes of a wave (a fixed shuffle tree, so the order is
// fixed); lane 0 holds the result.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (uint64_t)__shfl_xor((unsigned long long)v, o);
    return v;
}

__device__ __forceinline__ int64_t wave_max(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (int64_t)__shfl_xor((long long)v, o));
    return v;
}

// The term of amplitude i with probability p in reduction MODE.
template <int MODE>
__device__ __forceinline__ double reduce_term(uint64_t i, double p, uint64_t mask, int n) {
    if (MODE == 0) return p;
    if (MODE == 1) return (i & mask) ? p : 0.0;
    return p * (double)(n - 2 * (int)__popcll(i));
}

// MODE 0: sum |a|^2; 1: sum |a|^2 over (i & mask) != 0; 2: sum |a|^2 (n - 2 popcount(i)).
// 16-byte loads (two complex64 or one complex128 amplitude), four in
// flight per thread, four accumulators; the grid is fixed by N, so the
// summation order is too.  The block's sum: waves by shuffles, then one
// barrier.  (One 8-byte load per step and an LDS tree with a barrier per
// level read a 30-qubit complex64 state at 3.9 TB/s.)
template <typename T2, int MODE>
__global__ __launch_bounds__(kBlock) void k_reduce(const T2* __restrict__ a, uint64_t N, uint64_t mask, int n,
                                                   double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64];
    constexpr int A = sizeof(T2) == 8 ? 2 : 1;  // amplitudes per 16-byte load
    constexpr int U = 4;
    const float4* av = reinterpret_cast<const float4*>(a);
    const uint64_t NV = N / A;
    double acc[U] = {0.0, 0.0, 0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * kBlock * U;
    for (uint64_t b0 = (uint64_t)blockIdx.x * kBlock * U + threadIdx.x; b0 < NV; b0 += stride) {
        float4 v[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const uint64_t k = b0 + (uint64_t)j * kBlock;
            if (k < NV) v[j] = ld_nt(&av[k]);
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const uint64_t k = b0 + (uint64_t)j * kBlock;
            if (k >= NV) continue;
            if (A == 2) {
                acc[j] += reduce_term<MODE>(2 * k, prob_of(make_float2(v[j].x, v[j].y)), mask, n);
                acc[j] += reduce_term<MODE>(2 * k + 1, prob_of(make_float2(v[j].z, v[j].w)), mask, n);
            } else {
                double2 d;
                __builtin_memcpy(&d, &v[j], sizeof(d));
                acc[j] += reduce_term<MODE>(k, prob_of(d), mask, n);
            }
        }
    }
    const double r = wave_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// <a|b> = sum conj(a_i) b_i
template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_inner(const TA* __restrict__ a, const TB* __restrict__ b, uint64_t N,
                                                  double* __restrict__ partial) {
    __shared__ double sh[kBlock];
    double re = 0.0, im = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
        const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
        re += ar * br + ai * bi;
        im += ar * bi - ai * br;
    }
    double r0 = block_reduce_sum(re, sh);
    double r1 = block_reduce_sum(im, sh);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = r0;
        partial[2 * blockIdx.x + 1] = r1;
    }
}

// Sum `count` partial vectors of width w in a fixed order (one block).
__global__ __launch_bounds__(kBlock) void k_reduce_final(const double* __restrict__ partial, int count, int w,
                                                         double* __restrict__ out) {
    __shared__ double sh[kBlock];
    for (int c = 0; c < w; ++c) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < count; i += kBlock) acc += partial[i * w + c];
        double r = block_reduce_sum(acc, sh);
        if (threadIdx.x == 0) out[c] = r;
    }
}

// The amplitudes of 16-byte unit k of a state (two complex64 amplitudes,
// 2k and 2k + 1, or one complex128 amplitude k) times f where keep(index)
// holds, exactly 0 where it does not.
template <typename T2, typename R, typename Keep>
__device__ __forceinline__ float4 scale_unit(float4 v, uint64_t k, R f, Keep keep) {
    if constexpr (sizeof(T2) == 8) {
        const bool k0 = keep(2 * k), k1 = keep(2 * k + 1);
        return make_float4(k0 ? v.x * f : 0.f, k0 ? v.y * f : 0.f, k1 ? v.z * f : 0.f, k1 ? v.w * f : 0.f);
    } else {
        double2 d;
        __builtin_memcpy(&d, &v, sizeof(d));
        const bool k0 = keep(k);
        d.x = k0 ? d.x * f : 0.0;
        d.y = k0 ? d.y * f : 0.0;
        __builtin_memcpy(&v, &d, sizeof(d));
        return v;
    }
}

// Every amplitude through scale_unit: one 16-byte unit per lane, one-wave
// workgroups (the shape of the streaming gate passes).  (One 8-byte
// amplitude per step in a grid-stride loop moved a 30-qubit complex64 state
// at 4.5 TB/s, four 16-byte units per step at 4.8.)
constexpr int kUnitBlock = 64;

template <typename T2, typename R, typename Keep>
__device__ __forceinline__ void scale_pass(T2* a, uint64_t N, R f, Keep keep) {
    constexpr int A = sizeof(T2) == 8 ? 2 : 1;
    float4* av = reinterpret_cast<float4*>(a);
    const uint64_t k = (uint64_t)blockIdx.x * kUnitBlock + threadIdx.x;
    if (k < N / A) st_nt(scale_unit<T2>(ld_nt(&av[k]), k, f, keep), &av[k]);
}

// Workgroups of scale_pass over a state of N amplitudes.
unsigned unit_grid(uint64_t N, int precision) {
    const uint64_t units = precision == AQS_C64 ? N / 2 : N;
    return (unsigned)std::max<uint64_t>((units + kUnitBlock - 1) / kUnitBlock, 1);
}

template <typename T2, typename R>
__global__ __launch_bounds__(kUnitBlock) void k_scale(T2* __restrict__ a, uint64_t N, R f) {
    scale_pass(a, N, f, [](uint64_t) { return true; });
}

// Keep amplitudes whose bit matches `keep_one`, scaled by f; zero the rest.
template <typename T2, typename R>
__global__ __launch_bounds__(kUnitBlock) void k_collapse_qubit(T2* __restrict__ a, uint64_t N, uint64_t mask, int keep_one,
                                                           R f) {
    const bool want = keep_one != 0;
    scale_pass(a, N, f, [=](uint64_t i) { return ((i & mask) != 0) == want; });
}

template <typename T2>
__global__ __launch_bounds__(kBlock) void k_probs(const T2* __restrict__ a, uint64_t N, float* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += stride) {
        const T2 v = a[i];
        out[i] = (float)(v.x * v.x + v.y * v.y);
    }
}

// ---- sampling -------------------------------------------------------------

// sums[b] = exact sum of quantised |a|^2 over block b; lastnz[b] = last index
// with non-zero quantised probability in block b (-1 when none).  Each
// thread reads 16 bytes per load: complex64 amplitude pairs, complex128
// single amplitudes.
// One wavefront per sampling block (kSampleBlock amplitudes, 8 or 16
// 16-byte loads per lane), four per workgroup: shuffles reduce the block's
// integer sum and last non-zero index, with no LDS and no barrier (exact in
// any order).  (A 256-thread workgroup per block with LDS trees read a
// 30-qubit complex64 state at 4.9 TB/s.)
template <typename T2>
__global__ __launch_bounds__(kBlock) void k_block_sums(const T2* __restrict__ a, uint64_t N, uint64_t nb,
                                                       uint64_t* __restrict__ sums, int64_t* __restrict__ lastnz) {
    constexpr int A = sizeof(T2) == 8 ? 2 : 1;       // amplitudes per 16-byte unit
    constexpr int UPB = kSampleBlock / A;            // units per block
    constexpr int UPL = UPB / 64;                    // units per lane
    const uint64_t b = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (b >= nb) return;
    const int lane = threadIdx.x & 63;
    const float4* av = reinterpret_cast<const float4*>(a);
    const uint64_t NV = N / A;
    float4 v[UPL];
#pragma unroll
    for (int j = 0; j < UPL; ++j) {
        const uint64_t k = b * UPB + (uint64_t)j * 64 + lane;
        v[j] = k < NV ? ld_nt(&av[k]) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    uint64_t acc = 0;
    int64_t last = -1;
#pragma unroll
    for (int j = 0; j < UPL; ++j) {
        const uint64_t k = b * UPB + (uint64_t)j * 64 + lane;
        if (A == 2) {
            const uint64_t q0 = quantise(prob_of(make_float2(v[j].x, v[j].y)));
            const uint64_t q1 = quantise(prob_of(make_float2(v[j].z, v[j].w)));
            acc += q0 + q1;
            if (q0) last = max(last, (int64_t)(2 * k));
            if (q1) last = max(last, (int64_t)(2 * k + 1));
        } else {
            double2 d;
            __builtin_memcpy(&d, &v[j], sizeof(d));
            const uint64_t q = quantise(prob_of(d));
            acc += q;
            if (q) last = max(last, (int64_t)k);
        }
    }
    acc = wave_sum(acc);
    last = wave_max(last);
    if (lane == 0) {
        sums[b] = acc;
        lastnz[b] = last;
    }
}

// The exact inclusive scan of the block sums in three steps, so that no
// single workgroup walks all of them: each workgroup scans a chunk of
// kScanChunk blocks (k_scan_chunks), one workgroup scans the chunk totals
// (k_scan_top, which also forms the total and the last non-zero index), and
// every block adds its chunk's offset (k_scan_fixup).
constexpr int kScanChunk = 1024;

__global__ __launch_bounds__(kScanChunk) void k_scan_chunks(const uint64_t* __restrict__ sums,
                                                            const int64_t* __restrict__ lastnz, uint64_t nb,
                                                            uint64_t* __restrict__ incl, uint64_t* __restrict__ ctot,
                                                            int64_t* __restrict__ clast) {
    __shared__ uint64_t sh[kScanChunk];
    __shared__ int64_t shl[kScanChunk];
    const int t = threadIdx.x;
    const uint64_t b = (uint64_t)blockIdx.x * kScanChunk + t;
    sh[t] = b < nb ? sums[b] : 0;
    shl[t] = b < nb ? lastnz[b] : -1;
    __syncthreads();
    for (int off = 1; off < kScanChunk; off <<= 1) {
        const uint64_t v = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    if (b < nb) incl[b] = sh[t];
    if (t == kScanChunk - 1) ctot[blockIdx.x] = sh[t];
    for (int st = kScanChunk / 2; st > 0; st >>= 1) {
        if (t < st) shl[t] = max(shl[t], shl[t + st]);
        __syncthreads();
    }
    if (t == 0) clast[blockIdx.x] = shl[0];
}

// Exclusive prefix of the nc chunk totals (one workgroup of 1024 threads,
// each over a contiguous range), the total and the last non-zero index.
__global__ __launch_bounds__(1024) void k_scan_top(const uint64_t* __restrict__ ctot, const int64_t* __restrict__ clast,
                                                   uint64_t nc, uint64_t* __restrict__ cexcl, uint64_t* __restrict__ misc) {
    __shared__ uint64_t sh[1024];
    __shared__ int64_t shl[1024];
    const int t = threadIdx.x;
    const uint64_t per = (nc + 1023) / 1024;
    const uint64_t lo = (uint64_t)t * per, hi = min(nc, lo + per);
    uint64_t acc = 0;
    int64_t last = -1;
    for (uint64_t i = lo; i < hi; ++i) {
        acc += ctot[i];
        last = max(last, clast[i]);
    }
    sh[t] = acc;
    shl[t] = last;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint64_t v = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    uint64_t run = t > 0 ? sh[t - 1] : 0;
    for (uint64_t i = lo; i < hi; ++i) {
        cexcl[i] = run;
        run += ctot[i];
    }
    for (int st = 512; st > 0; st >>= 1) {
        if (t < st) shl[t] = max(shl[t], shl[t + st]);
        __syncthreads();
    }
    if (t == 0) {
        misc[0] = sh[1023];          // total
        misc[1] = (uint64_t)shl[0];  // last non-zero index (as int64 bits)
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_fixup(uint64_t* __restrict__ incl, const uint64_t* __restrict__ cexcl,
                                                       uint64_t nb) {
    const uint64_t b = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (b < nb) incl[b] += cexcl[b / kScanChunk];
}

// One wavefront per shot.
template <typename T2>
__global__ __launch_bounds__(kBlock) void k_sample(const T2* __restrict__ a, uint64_t N, const uint64_t* __restrict__ sums,
                                                   const uint64_t* __restrict__ incl, const uint64_t* __restrict__ misc,
                                                   uint64_t nb, const double* __restrict__ u, uint32_t shots,
                                                   uint64_t* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const uint32_t shot = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (shot >= shots) return;
    const uint64_t thr = (uint64_t)floor(u[shot] * kFixScale);
    const uint64_t total = misc[0];
    if (thr >= total) {
        // at or above the total: the last index of non-zero probability
        // (0 for a null state)
        if (lane == 0) out[shot] = ((int64_t)misc[1] < 0) ? 0 : misc[1];
        return;
    }
    // first block with incl[b] > thr
    uint64_t lo = 0, hi = nb;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (incl[mid] > thr) hi = mid;
        else lo = mid + 1;
    }
    const uint64_t b = lo;
    const uint64_t before = incl[b] - sums[b];
    constexpr int per = kSampleBlock / 64;  // 16 amplitudes per lane
    const uint64_t i0 = b * kSampleBlock + (uint64_t)lane * per;
    uint64_t q[per];
    uint64_t lsum = 0;
#pragma unroll
    for (int j = 0; j < per; ++j) {
        const uint64_t i = i0 + j;
        q[j] = (i < N) ? quantise(prob_of(a[i])) : 0;
        lsum += q[j];
    }
    // inclusive wavefront scan of lane sums
    uint64_t incl_l = lsum;
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t v = __shfl_up(incl_l, off, 64);
        if (lane >= off) incl_l += v;
    }
    const uint64_t excl_l = incl_l - lsum;
    // the lane whose range crosses thr
    const bool crosses = (before + excl_l <= thr) && (before + incl_l > thr);
    const unsigned long long ballot = __ballot(crosses);
    const int owner = ballot ? (__ffsll((long long)ballot) - 1) : 63;
    if (lane == owner) {
        uint64_t run = before + excl_l;
        uint64_t idx = i0 + per - 1;
#pragma unroll
        for (int j = 0; j < per; ++j) {
            run += q[j];
            if (run > thr) {
                idx = i0 + j;
                break;
            }
        }
        out[shot] = idx;
    }
}

unsigned red_grid(uint64_t N) {
    uint64_t b = (N + kBlock - 1) / kBlock;
    return (unsigned)std::min<uint64_t>(std::max<uint64_t>(b, 1), kRedGrid);
}

// Tables of the sampling pass over nb blocks, carved from one scratch area:
// per-block sums, inclusive prefixes, last non-zero indices, then the total
// and the global last non-zero index.
struct SampleTables {
    uint64_t* sums;
    uint64_t* incl;
    int64_t* lastnz;
    uint64_t* misc;
    uint64_t* ctot;   // per chunk of kScanChunk blocks: total,
    int64_t* clast;   // last non-zero index,
    uint64_t* cexcl;  // exclusive prefix
};

constexpr uint64_t scan_chunks(uint64_t nb) { return (nb + kScanChunk - 1) / kScanChunk; }
constexpr size_t sample_table_bytes(uint64_t nb) {
    return (size_t)(3 * nb + 2 + 3 * scan_chunks(nb)) * sizeof(uint64_t);
}

SampleTables sample_tables(void* base, uint64_t nb) {
    SampleTables t;
    t.sums = (uint64_t*)base;
    t.incl = t.sums + nb;
    t.lastnz = (int64_t*)(t.incl + nb);
    t.misc = (uint64_t*)(t.lastnz + nb);
    t.ctot = t.misc + 2;
    t.clast = (int64_t*)(t.ctot + scan_chunks(nb));
    t.cexcl = (uint64_t*)(t.clast + scan_chunks(nb));
    return t;
}

// Block sums, the exact scan and one wavefront per shot, on `stream`.
void launch_sampling(const void* amps, int precision, uint64_t N, const SampleTables& t, const double* du,
                     uint32_t shots, uint64_t* dout, hipStream_t stream) {
    const uint64_t nb = (N + kSampleBlock - 1) / kSampleBlock;
    const unsigned gb = (unsigned)((nb + kBlock / 64 - 1) / (kBlock / 64));
    if (precision == AQS_C64)
        hipLaunchKernelGGL(k_block_sums<float2>, dim3(gb), dim3(kBlock), 0, stream, (const float2*)amps, N, nb, t.sums,
                           t.lastnz);
    else
        hipLaunchKernelGGL(k_block_sums<double2>, dim3(gb), dim3(kBlock), 0, stream, (const double2*)amps, N, nb,
                           t.sums, t.lastnz);
    const uint64_t nc = scan_chunks(nb);
    hipLaunchKernelGGL(k_scan_chunks, dim3((unsigned)nc), dim3(kScanChunk), 0, stream, (const uint64_t*)t.sums,
                       (const int64_t*)t.lastnz, nb, t.incl, t.ctot, t.clast);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, (const uint64_t*)t.ctot, (const int64_t*)t.clast, nc,
                       t.cexcl, t.misc);
    hipLaunchKernelGGL(k_scan_fixup, dim3((unsigned)((nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, t.incl,
                       (const uint64_t*)t.cexcl, nb);
    const unsigned g = (shots + (kBlock / 64) - 1) / (kBlock / 64);
    if (precision == AQS_C64)
        hipLaunchKernelGGL(k_sample<float2>, dim3(g), dim3(kBlock), 0, stream, (const float2*)amps, N,
                           (const uint64_t*)t.sums, (const uint64_t*)t.incl, (const uint64_t*)t.misc, nb, du, shots,
                           dout);
    else
        hipLaunchKernelGGL(k_sample<double2>, dim3(g), dim3(kBlock), 0, stream, (const double2*)amps, N,
                           (const uint64_t*)t.sums, (const uint64_t*)t.incl, (const uint64_t*)t.misc, nb, du, shots,
                           dout);
    AQS_HIP(hipGetLastError());
}

double* partial_buf(aqs_state s) { return (double*)s->scratch; }
double* result_buf(aqs_state s) { return (double*)s->scratch + 2 * kRedGrid; }

void read_results(aqs_state s, double* out, int w) {
    ensure_pinned(s, sizeof(double) * w);
    AQS_HIP(hipMemcpyAsync(s->pinned, result_buf(s), sizeof(double) * w, hipMemcpyDeviceToHost, s->stream));
    AQS_HIP(hipStreamSynchronize(s->stream));
    std::copy((double*)s->pinned, (double*)s->pinned + w, out);
}

template <int MODE>
double reduce_mode(aqs_state s, uint64_t mask) {
    const unsigned g = red_grid(s->N);
    {
        ProfScope ps(s, 0);
        if (s->precision == AQS_C64)
            hipLaunchKernelGGL((k_reduce<float2, MODE>), dim3(g), dim3(kBlock), 0, s->stream, (const float2*)s->amps, s->N,
                               mask, s->n, partial_buf(s));
        else
            hipLaunchKernelGGL((k_reduce<double2, MODE>), dim3(g), dim3(kBlock), 0, s->stream, (const double2*)s->amps,
                               s->N, mask, s->n, partial_buf(s));
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, s->stream, (const double*)partial_buf(s), (int)g, 1,
                           result_buf(s));
        AQS_HIP(hipGetLastError());
    }
    double r;
    read_results(s, &r, 1);
    return r;
}

aqs_state check(aqs_state s) {
    if (!s || !s->amps) fail(AQS_ERR_INVALID_ARGUMENT, "null state handle");
    return s;
}

void check_qubit(aqs_state s, int q, const char* msg) {
    if (q < 0 || q >= s->n) fail(AQS_ERR_OUT_OF_RANGE, msg);
}

}  // namespace

namespace aqs_hip {

double reduce_norm2(aqs_state s) { return reduce_mode<0>(s, 0); }

size_t sample_norm_scratch_bytes(int n) {
    const uint64_t N = 1ull << n;
    const uint64_t nb = (N + kSampleBlock - 1) / kSampleBlock;
    return sample_table_bytes(nb) + sizeof(double) * red_grid(N);
}

void queue_sample_norm(const void* amps, int n, void* scratch, const double* du, uint32_t shots, uint64_t* dout,
                       double* dnorm, hipStream_t stream) {
    const uint64_t N = 1ull << n;
    const uint64_t nb = (N + kSampleBlock - 1) / kSampleBlock;
    double* partial = (double*)((char*)scratch + sample_table_bytes(nb));
    const unsigned g = red_grid(N);
    hipLaunchKernelGGL((k_reduce<float2, 0>), dim3(g), dim3(kBlock), 0, stream, (const float2*)amps, N, 0ull, n,
                       partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, stream, (const double*)partial, (int)g, 1, dnorm);
    AQS_HIP(hipGetLastError());
    if (shots) launch_sampling(amps, AQS_C64, N, sample_tables(scratch, nb), du, shots, dout, stream);
}

void scale_state(aqs_state s, double factor) {
    const unsigned g = unit_grid(s->N, s->precision);
    ProfScope ps(s, 0);
    if (s->precision == AQS_C64)
        hipLaunchKernelGGL((k_scale<float2, float>), dim3(g), dim3(kUnitBlock), 0, s->stream, (float2*)s->amps, s->N,
                           (float)factor);
    else
        hipLaunchKernelGGL((k_scale<double2, double>), dim3(g), dim3(kUnitBlock), 0, s->stream, (double2*)s->amps,
                           s->N, factor);
    AQS_HIP(hipGetLastError());
}

}  // namespace aqs_hip

extern "C" {

aqs_status aqs_norm2(aqs_state s, double* out) {
    return guard([&] {
        check(s);
        double r = reduce_norm2(s);
        if (out) *out = r;
    });
}

aqs_status aqs_normalize(aqs_state s) {
    return guard([&] {
        check(s);
        double n2 = reduce_norm2(s);
        if (!(n2 > 0.0)) fail(AQS_ERR_INVALID_ARGUMENT, "Cannot normalize a null state");
        scale_state(s, 1.0 / std::sqrt(n2));
    });
}

aqs_status aqs_prob1(aqs_state s, int qubit, double* out) {
    return guard([&] {
        check(s);
        check_qubit(s, qubit, "Cannot obtain probability of the given qubit");
        double r = reduce_mode<1>(s, qbit(s->n, qubit));
        if (out) *out = r;
    });
}

aqs_status aqs_prob1_all(aqs_state s, double* out) {
    return guard([&] {
        check(s);
        for (int q = 0; q < s->n; ++q) {
            double r = reduce_mode<1>(s, qbit(s->n, q));
            if (out) out[q] = r;
        }
    });
}

aqs_status aqs_expect_sum_z(aqs_state s, double* out) {
    return guard([&] {
        check(s);
        double r = reduce_mode<2>(s, 0);
        if (out) *out = r;
    });
}

aqs_status aqs_probabilities(aqs_state s, float* out) {
    return guard([&] {
        check(s);
        if (!out) fail(AQS_ERR_INVALID_ARGUMENT, "null output");
        // chunks of at most 2^24 floats through the state's staging buffer:
        // no allocation of N floats per call, and no hipFree stalling the device
        const uint64_t chunk = std::min<uint64_t>(s->N, 1ull << 24);
        ensure_staging(s, chunk * sizeof(float));
        float* d = (float*)s->staging;
        for (uint64_t off = 0; off < s->N; off += chunk) {
            const uint64_t len = std::min<uint64_t>(chunk, s->N - off);
            const unsigned g = red_grid(len) * 4;
            if (s->precision == AQS_C64)
                hipLaunchKernelGGL(k_probs<float2>, dim3(g), dim3(kBlock), 0, s->stream, (const float2*)s->amps + off,
                                   len, d);
            else
                hipLaunchKernelGGL(k_probs<double2>, dim3(g), dim3(kBlock), 0, s->stream, (const double2*)s->amps + off,
                                   len, d);
            AQS_HIP(hipGetLastError());
            AQS_HIP(hipMemcpyAsync(out + off, d, len * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        }
        AQS_HIP(hipStreamSynchronize(s->stream));
    });
}

aqs_status aqs_amplitude(aqs_state s, uint64_t index, double* re_im) {
    return guard([&] {
        check(s);
        if (index >= s->N) fail(AQS_ERR_OUT_OF_RANGE, "amplitude index out of range");
        ensure_pinned(s, 16);
        AQS_HIP(hipMemcpyAsync(s->pinned, (const char*)s->amps + index * s->amp_bytes(), s->amp_bytes(),
                               hipMemcpyDeviceToHost, s->stream));
        AQS_HIP(hipStreamSynchronize(s->stream));
        if (re_im) {
            if (s->precision == AQS_C64) {
                re_im[0] = ((float*)s->pinned)[0];
                re_im[1] = ((float*)s->pinned)[1];
            } else {
                re_im[0] = ((double*)s->pinned)[0];
                re_im[1] = ((double*)s->pinned)[1];
            }
        }
    });
}

aqs_status aqs_sample(aqs_state s, const double* u, uint32_t shots, uint64_t* out_index) {
    return guard([&] {
        check(s);
        if (shots == 0) return;
        if (!u || !out_index) fail(AQS_ERR_INVALID_ARGUMENT, "null uniforms or output");
        const uint64_t nb = (s->N + kSampleBlock - 1) / kSampleBlock;
        if (sample_table_bytes(nb) > s->scratch_bytes) fail(AQS_ERR_RUNTIME, "sampling scratch too small");
        // uniforms in, indices out through the staging buffer
        const size_t ubytes = sizeof(double) * shots, obytes = sizeof(uint64_t) * shots;
        ensure_staging(s, ubytes + obytes);
        AQS_HIP(hipMemcpyAsync(s->staging, u, ubytes, hipMemcpyHostToDevice, s->stream));
        double* du = (double*)s->staging;
        uint64_t* dout = (uint64_t*)((char*)s->staging + ubytes);
        {
            ProfScope ps(s, 0);
            launch_sampling(s->amps, s->precision, s->N, sample_tables(s->scratch, nb), du, shots, dout, s->stream);
        }
        AQS_HIP(hipMemcpyAsync(out_index, dout, obytes, hipMemcpyDeviceToHost, s->stream));
        AQS_HIP(hipStreamSynchronize(s->stream));
    });
}

aqs_status aqs_measure_qubit(aqs_state s, int qubit, double u, int* outcome, double* p1_out) {
    return guard([&] {
        check(s);
        check_qubit(s, qubit, "Cannot measure the state of the given qubit");
        const uint64_t mask = qbit(s->n, qubit);
        const double p1 = reduce_mode<1>(s, mask);
        const int one = u < p1 ? 1 : 0;
        const double p = one ? p1 : 1.0 - p1;
        if (!(p > 0.0)) fail(AQS_ERR_RUNTIME, "measured outcome has zero probability");
        const double f = 1.0 / std::sqrt(p);
        const unsigned g = unit_grid(s->N, s->precision);
        ProfScope ps(s, 0);
        if (s->precision == AQS_C64)
            hipLaunchKernelGGL((k_collapse_qubit<float2, float>), dim3(g), dim3(kUnitBlock), 0, s->stream,
                               (float2*)s->amps, s->N, mask, one, (float)f);
        else
            hipLaunchKernelGGL((k_collapse_qubit<double2, double>), dim3(g), dim3(kUnitBlock), 0, s->stream,
                               (double2*)s->amps, s->N, mask, one, f);
        AQS_HIP(hipGetLastError());
        if (outcome) *outcome = one;
        if (p1_out) *p1_out = p1;
    });
}

aqs_status aqs_collapse_to(aqs_state s, uint64_t index) { return aqs_state_set_basis(s, index); }

aqs_status aqs_sample_noalloc(const void* amps, int nqubits, int precision, const double* u, uint32_t shots,
                              uint64_t* out_index, void* stream) {
    return guard([&] {
        if (nqubits < 1 || nqubits > AQS_NOALLOC_MAX_QUBITS)
            fail(AQS_ERR_INVALID_ARGUMENT, "aqs_sample_noalloc: nqubits must be 1.." +
                                               std::to_string(AQS_NOALLOC_MAX_QUBITS) + " (fixed device scratch)");
        if (precision != AQS_C64 && precision != AQS_C128) fail(AQS_ERR_INVALID_ARGUMENT, "unknown precision");
        if (shots == 0) return;
        if (!amps || !u || !out_index) fail(AQS_ERR_INVALID_ARGUMENT, "null state, uniforms or output");
        const uint64_t N = 1ull << nqubits;
        const uint64_t nb = (N + kSampleBlock - 1) / kSampleBlock;
        static_assert(sample_table_bytes(((1ull << AQS_NOALLOC_MAX_QUBITS) + kSampleBlock - 1) / kSampleBlock) <=
                          kNoallocScratchBytes,
                      "scratch holds the tables of the largest state");
        void* base = noalloc_scratch_ptr();
        launch_sampling(amps, precision, N, sample_tables(base, nb), u, shots, out_index,
                        stream ? (hipStream_t)stream : default_stream());
    });
}

aqs_status aqs_inner(aqs_state a, aqs_state b, double* re_im) {
    return guard([&] {
        check(a);
        check(b);
        if (a->n != b->n) fail(AQS_ERR_INVALID_ARGUMENT, "inner product of states with different qubit counts");
        if (a->stream != b->stream) AQS_HIP(hipStreamSynchronize(b->stream));
        const unsigned g = red_grid(a->N);
        double* part = partial_buf(a);
        if (a->precision == AQS_C64 && b->precision == AQS_C64)
            hipLaunchKernelGGL((k_inner<float2, float2>), dim3(g), dim3(kBlock), 0, a->stream, (const float2*)a->amps,
                               (const float2*)b->amps, a->N, part);
        else if (a->precision == AQS_C64)
            hipLaunchKernelGGL((k_inner<float2, double2>), dim3(g), dim3(kBlock), 0, a->stream, (const float2*)a->amps,
                               (const double2*)b->amps, a->N, part);
        else if (b->precision == AQS_C64)
            hipLaunchKernelGGL((k_inner<double2, float2>), dim3(g), dim3(kBlock), 0, a->stream, (const double2*)a->amps,
                               (const float2*)b->amps, a->N, part);
        else
            hipLaunchKernelGGL((k_inner<double2, double2>), dim3(g), dim3(kBlock), 0, a->stream,
                               (const double2*)a->amps, (const double2*)b->amps, a->N, part);
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, a->stream, (const double*)part, (int)g, 2,
                           result_buf(a));
        AQS_HIP(hipGetLastError());
        double r[2];
        read_results(a, r, 2);
        if (re_im) {
            re_im[0] = r[0];
            re_im[1] = r[1];
        }
    });
}

}  // extern "C"
