// In-place index-bit permutation pass (an involution of bit positions).
//
// compile() turns uncontrolled Swap gates into qubit relabellings and, at
// the end of a schedule, restores the qubit order with at most two of these
// passes (every permutation is a product of two involutions).  The pass
// moves amplitude i to P(i), where P sends index bit b to bit beta(b) and
// beta is an involution, so P pairs amplitudes and works in place.
//
// Tiling: the tile bits T are bits 0..4, their partners and the pairs added
// below (10-11 bits), so tiles are runs of >= 32 amplitudes and T is closed
// under beta; the other bits R are permuted among themselves.  A workgroup
// takes an outer index o over R and its partner o' = beta_R(o) (o < o';
// o == o' is a tile permuted within itself), stages both tiles in LDS and
// writes each one, permuted by beta_T, to the other tile's place with
// coalesced stores.
//
// Small tiles starve the pass (a 5-bit tile, as when beta only moves middle
// bits, ran at 0.1 TB/s at 30 qubits; 8 bits at 0.9 TB/s), so T is grown to
// at least 10 bits with further outer bits and their partners, lowest first:
// 10-11 bits run at 2.8-3.1 TB/s on the same permutations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.h"

using namespace aqs_hip;

namespace {

constexpr int kBlock = 256;

struct PermArgs {
    void* sv;
    int nt, nr;
    int tpos[12];     // tile bit positions (ascending; tpos[j] = j for j < 5)
    int tmap[12];     // tile bit j -> tile bit tmap[j]
    int rpos[34];     // outer bit positions (ascending)
    int rmap[34];     // outer bit j -> outer bit rmap[j]
    uint64_t outer;   // 2^nr
};

__device__ __forceinline__ uint64_t deposit(uint64_t v, const int* pos, int len) {
    uint64_t r = 0;
    for (int j = 0; j < len; ++j) r |= ((v >> j) & 1ull) << pos[j];
    return r;
}

__device__ __forceinline__ uint64_t permute_bits(uint64_t v, const int* map, int len) {
    uint64_t r = 0;
    for (int j = 0; j < len; ++j) r |= ((v >> j) & 1ull) << map[j];
    return r;
}

// Tile-local tables, built once per workgroup: index offsets of tile bits
// 5.. and the permuted position of every tile element.
constexpr int kMaxTileBits = 11;
constexpr int kMaxTile = 1 << kMaxTileBits;
// Elements per thread and tile: complex64 moves pairs (2^11 / 2 / 256 = 4),
// complex128 single amplitudes of tiles of at most 2^10 (4).
constexpr int kUnroll = kMaxTile / 2 / kBlock;
static_assert(kUnroll * kBlock * 2 >= kMaxTile, "a thread's share of a complex64 tile fits the unroll");
static_assert(kUnroll * kBlock >= (1 << 10), "a thread's share of a complex128 tile fits the unroll");

// LDS slot of tile element p: one element of padding per 32.  Permuted
// writes (a bit reversal sends consecutive elements 32 apart) then spread
// over the banks instead of hitting one; the contiguous reads stay
// contiguous within each run of 32.
__device__ __forceinline__ int lds_slot(int p) { return p + (p >> 5); }

template <typename T2>
__global__ __launch_bounds__(kBlock) void k_bitperm(PermArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int tile = 1 << a.nt;
    T2* A = reinterpret_cast<T2*>(smem);
    T2* B = A + lds_slot(tile);
    __shared__ uint64_t hi_off[kMaxTile >> 5];
    __shared__ uint16_t pmap[kMaxTile];
    // tile bits 0..min(nt, 5)-1 are index bits 0..4 (all of them when nt < 5)
    const int hi = a.nt > 5 ? a.nt - 5 : 0;
    for (int v = threadIdx.x; v < (1 << hi); v += kBlock) {
        uint64_t o = 0;
        for (int j = 0; j < hi; ++j)
            if ((v >> j) & 1) o |= 1ull << a.tpos[5 + j];
        hi_off[v] = o;
    }
    for (int t = threadIdx.x; t < tile; t += kBlock)
        pmap[t] = (uint16_t)lds_slot((int)permute_bits((uint64_t)t, a.tmap, a.nt));
    __syncthreads();
    T2* sv = reinterpret_cast<T2*>(a.sv);
    // Whether a workgroup skips o depends on o's low bits, so with a plain
    // grid stride (low bits fixed per workgroup) some workgroups would move
    // nearly every pair and others none.  Workgroup b takes o = m G + (b ^
    // (m mod G)) in step m (G = gridDim.x, a power of two dividing outer):
    // every o once, each workgroup seeing every low-bit pattern.
    const uint64_t G = gridDim.x, steps = a.outer / G;
    for (uint64_t m = 0; m < steps; ++m) {
        const uint64_t o = m * G + ((uint64_t)blockIdx.x ^ (m & (G - 1)));
        const uint64_t o2 = permute_bits(o, a.rmap, a.nr);
        if (o2 < o) continue;  // the partner tile's workgroup does this pair
        const uint64_t b1 = deposit(o, a.rpos, a.nr), b2 = deposit(o2, a.rpos, a.nr);
        const bool pair = o2 != o;
        // Every load of the pair is issued before the first LDS write: a
        // thread moves up to kUnroll elements of each tile, so the whole tile
        // pair is in flight at once instead of one load round trip per element.
        if constexpr (sizeof(T2) == 8) {
            // complex64: amplitude pairs (16 bytes) along the contiguous bits
            float4* sv4 = reinterpret_cast<float4*>(a.sv);
            const int units = tile / 2;
            float4 xa[kUnroll], xb[kUnroll];
#pragma unroll
            for (int k = 0; k < kUnroll; ++k) {
                const int f = threadIdx.x + k * kBlock;
                if (f >= units) break;
                const int t = 2 * f;
                const uint64_t off = (uint64_t)(t & 31) | hi_off[t >> 5];
                xa[k] = ld_nt(&sv4[(b1 | off) >> 1]);
                if (pair) xb[k] = ld_nt(&sv4[(b2 | off) >> 1]);
            }
#pragma unroll
            for (int k = 0; k < kUnroll; ++k) {
                const int f = threadIdx.x + k * kBlock;
                if (f >= units) break;
                const int t = 2 * f;
                A[pmap[t]] = make_float2(xa[k].x, xa[k].y);
                A[pmap[t + 1]] = make_float2(xa[k].z, xa[k].w);
                if (pair) {
                    B[pmap[t]] = make_float2(xb[k].x, xb[k].y);
                    B[pmap[t + 1]] = make_float2(xb[k].z, xb[k].w);
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kUnroll; ++k) {
                const int f = threadIdx.x + k * kBlock;
                if (f >= units) break;
                const int t = 2 * f;
                const uint64_t off = (uint64_t)(t & 31) | hi_off[t >> 5];
                const int l = lds_slot(t);  // t even: t + 1 shares its run
                const float2 p = A[l], q = A[l + 1];
                st_nt(make_float4(p.x, p.y, q.x, q.y), &sv4[((pair ? b2 : b1) | off) >> 1]);
                if (pair) {
                    const float2 u = B[l], v = B[l + 1];
                    st_nt(make_float4(u.x, u.y, v.x, v.y), &sv4[(b1 | off) >> 1]);
                }
            }
        } else {
            T2 xa[kUnroll], xb[kUnroll];
#pragma unroll
            for (int k = 0; k < kUnroll; ++k) {
                const int t = threadIdx.x + k * kBlock;
                if (t >= tile) break;
                const uint64_t off = (uint64_t)(t & 31) | hi_off[t >> 5];
                xa[k] = sv[b1 | off];
                if (pair) xb[k] = sv[b2 | off];
            }
#pragma unroll
            for (int k = 0; k < kUnroll; ++k) {
                const int t = threadIdx.x + k * kBlock;
                if (t >= tile) break;
                A[pmap[t]] = xa[k];
                if (pair) B[pmap[t]] = xb[k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kUnroll; ++k) {
                const int t = threadIdx.x + k * kBlock;
                if (t >= tile) break;
                const uint64_t off = (uint64_t)(t & 31) | hi_off[t >> 5];
                sv[(pair ? b2 : b1) | off] = A[lds_slot(t)];
                if (pair) sv[b1 | off] = B[lds_slot(t)];
            }
        }
        __syncthreads();
    }
}

}  // namespace

namespace aqs_hip {

// beta: involution of bit positions 0..n-1 (beta[beta[b]] == b).
void apply_bit_involution(aqs_state s, const std::vector<int>& beta) {
    const int n = s->n;
    if ((int)beta.size() != n) fail(AQS_ERR_INVALID_ARGUMENT, "permutation size mismatch");
    for (int b = 0; b < n; ++b)
        if (beta[b] < 0 || beta[b] >= n || beta[beta[b]] != b) fail(AQS_ERR_INVALID_ARGUMENT, "not an involution");
    std::vector<char> in_t(n, 0);
    int nt = 0;
    for (int b = 0; b < std::min(n, 5); ++b) {
        nt += !in_t[b];
        in_t[b] = 1;
        nt += !in_t[beta[b]];
        in_t[beta[b]] = 1;
    }
    // grow the tile to >= 10 bits (LDS: two tiles of 2^11 complex64 or
    // 2^10 complex128 amplitudes at most)
    const int cap = s->precision == AQS_C64 ? kMaxTileBits : 10;
    for (int b = 5; b < n && nt < 10; ++b) {
        if (in_t[b]) continue;
        const int add = beta[b] == b ? 1 : 2;
        if (nt + add > cap) continue;
        in_t[b] = in_t[beta[b]] = 1;
        nt += add;
    }
    std::vector<int> T, R;
    for (int b = 0; b < n; ++b) (in_t[b] ? T : R).push_back(b);
    PermArgs a{};
    a.sv = s->amps;
    a.nt = (int)T.size();
    a.nr = (int)R.size();
    for (int j = 0; j < a.nt; ++j) {
        a.tpos[j] = T[j];
        a.tmap[j] = (int)(std::find(T.begin(), T.end(), beta[T[j]]) - T.begin());
    }
    for (int j = 0; j < a.nr; ++j) {
        a.rpos[j] = R[j];
        a.rmap[j] = (int)(std::find(R.begin(), R.end(), beta[R[j]]) - R.begin());
    }

    a.outer = 1ull << a.nr;
    if (a.nt < 1 || a.nt > cap) fail(AQS_ERR_INVALID_ARGUMENT, "bit permutation tile out of range");
    const uint64_t blocks = std::min<uint64_t>(a.outer, 16384);  // powers of two: blocks divides outer
    const size_t slots = ((size_t)1 << a.nt) + (((size_t)1 << a.nt) >> 5);
    const size_t lds = 2 * slots * s->amp_bytes();
    ProfScope ps(s, kSlotPermute);
    if (s->precision == AQS_C64)
        hipLaunchKernelGGL(k_bitperm<float2>, dim3((unsigned)blocks), dim3(kBlock), lds, s->stream, a);
    else
        hipLaunchKernelGGL(k_bitperm<double2>, dim3((unsigned)blocks), dim3(kBlock), lds, s->stream, a);
    AQS_HIP(hipGetLastError());
}

}  // namespace aqs_hip
