// In-place index-bit permutation pass (an involution of bit positions).
//
// compile() turns uncontrolled Swap gates into qubit relabellings and, at
// the end of a schedule, restores the qubit order with at most two of these
// passes (every permutation is a product of two involutions).  The pass
// moves amplitude i to P(i), where P sends index bit b to bit beta(b) and
// beta is an involution, so P pairs amplitudes and works in place.
//
// Tiling: the tile bits T are bits 0..4 and their partners (<= 10 bits), so
// tiles are runs of >= 32 amplitudes and T is closed under beta; the other
// bits R are permuted among themselves.  A workgroup takes an outer index o
// over R and its partner o' = beta_R(o) (o < o'; o == o' is a tile permuted
// within itself), stages both tiles in LDS and writes each one, permuted by
// beta_T, to the other tile's place with coalesced stores.
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
    int tpos[10];     // tile bit positions (ascending; tpos[j] = j for j < 5)
    int tmap[10];     // tile bit j -> tile bit tmap[j]
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

template <typename T2>
__global__ __launch_bounds__(kBlock) void k_bitperm(PermArgs a) {
    extern __shared__ unsigned char smem[];
    const int tile = 1 << a.nt;
    T2* A = reinterpret_cast<T2*>(smem);
    T2* B = A + tile;
    T2* sv = reinterpret_cast<T2*>(a.sv);
    for (uint64_t o = blockIdx.x; o < a.outer; o += gridDim.x) {
        const uint64_t o2 = permute_bits(o, a.rmap, a.nr);
        if (o2 < o) continue;  // the partner tile's workgroup does this pair
        const uint64_t b1 = deposit(o, a.rpos, a.nr), b2 = deposit(o2, a.rpos, a.nr);
        for (int t = threadIdx.x; t < tile; t += kBlock) {
            const uint64_t off = deposit((uint64_t)t, a.tpos, a.nt);
            const int pt = (int)permute_bits((uint64_t)t, a.tmap, a.nt);
            A[pt] = sv[b1 | off];
            if (o2 != o) B[pt] = sv[b2 | off];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < tile; t += kBlock) {
            const uint64_t off = deposit((uint64_t)t, a.tpos, a.nt);
            if (o2 != o) {
                sv[b2 | off] = A[t];
                sv[b1 | off] = B[t];
            } else {
                sv[b1 | off] = A[t];
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
    for (int b = 0; b < std::min(n, 5); ++b) {
        in_t[b] = 1;
        in_t[beta[b]] = 1;
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
    const uint64_t blocks = std::min<uint64_t>(a.outer, 16384);
    const size_t lds = 2 * ((size_t)1 << a.nt) * s->amp_bytes();
    ProfScope ps(s, kSlotPermute);
    if (s->precision == AQS_C64)
        hipLaunchKernelGGL(k_bitperm<float2>, dim3((unsigned)blocks), dim3(kBlock), lds, s->stream, a);
    else
        hipLaunchKernelGGL(k_bitperm<double2>, dim3((unsigned)blocks), dim3(kBlock), lds, s->stream, a);
    AQS_HIP(hipGetLastError());
}

}  // namespace aqs_hip
