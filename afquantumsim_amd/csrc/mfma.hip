// Dense k-qubit blocks (k = 4, 5) on CDNA4 matrix cores, complex64 state.
//
// A fused block is a 2^k x 2^k complex unitary U applied to every group of
// 2^k amplitudes that differ only in the k target bits; with the groups as
// columns the pass is the complex GEMM  Y = U X.
//
// Tiling.  A tile is the 2^(k+5) amplitudes spanned by the k target bits and
// the 5 lowest non-target bits (32 columns).  Bits 0..4 of the amplitude
// index always belong to that set, so a tile is a union of contiguous runs
// of at least 32 amplitudes (256 B) whatever the targets are: every HBM
// access below is a coalesced 16-byte-per-lane run.  One wavefront owns a
// tile at a time: it streams the tile into a private LDS region (padded by
// 16 B every 256 B), computes from LDS, writes the result back to LDS and
// streams it out.  The next tile's global loads are issued before the
// matrix work of the current one.
//
// Arithmetic (FP32 MFMA, FP32 accumulation), v_mfma_f32_32x32x2_f32:
//   k = 5: three real products (Gauss): T1 = Ur Xr, T2 = Ui Xi,
//          T3 = (Ur + Ui)(Xr + Xi);  Yr = T1 - T2, Yi = T3 - T1 - T2
//          -> 48 MFMAs per 32 columns instead of 64.
//   k = 4: the stacked [Ur; Ui] (32 x 16) times Xr and times Xi
//          -> 16 MFMAs per 32 columns.
//
// Diagonal epilogue.  A diagonal table pass (diag.hip) that follows a block,
// whose anchor bits lie in the block's tile and which has at most one chunk
// touching the tile, is applied while the tile is written out: per tile the
// chunks outside the tile give one factor per anchor value (formed by the
// first lanes into LDS), and the chunk inside the tile, if any, is looked up
// per amplitude in a copy of its table slice held in LDS for the whole
// launch.  The table pass then costs no sweep of its own.  The same table
// can instead precede the block (prologue): the factors then scale the tile
// as it is read in.
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

// Row length of the LDS copy of a table chunk's slice (32 entries + 1 pad).
constexpr int kTinRow = 33;

// Diagonal table applied in the epilogue (amplitude bit positions).
struct MfmaEpilogue {
    const float2* tab;    // [2^ng][nc][32]
    uint64_t reqmask;     // bits that must be 1 for a factor other than 1
    int ng, gstart;       // anchor bits (gstart >= 0: consecutive from gstart)
    int gpos[5];
    int nc;
    uint32_t outside;     // chunks whose bits all lie outside the tile
    int inside;           // the chunk touching the tile, or -1
    int cstart[6];
    int clen[6];
    int cpos[6][5];
};

// Bit involution applied while the block's tiles are written out: a restore
// pass right after the block whose permutation maps the tile bits onto
// themselves moves every amplitude i to P(i) (bit b to bit beta(b)).  Tile
// t (outer bits o) then lands in tile beta(o), its local index permuted.
struct MfmaPerm {
    int nr;          // outer (non-tile) bits, ascending positions
    int rmap[40];    // outer bit j -> outer bit rmap[j]
    int tmap[10];    // tile bit j -> tile bit tmap[j]
};

struct MfmaArgs {
    float4* sv4;
    uint64_t tiles;
    int ntile_bits;       // k + 5
    int tpos[10];         // sorted amplitude bit positions of the tile
    uint32_t gj[8];       // float4 offset of the j-th 64-lane chunk: pdep(128 j) / 2
    uint32_t glane[64];   // float4 offset of lane l inside a chunk: pdep(2 l) / 2
    uint32_t tr[32];      // tile-index bits of target sub-index r
    uint32_t tc[32];      // tile-index bits of column c
    uint32_t ptr[32];     // permuted blocks: tr under the tile-local permutation
    const float* apack;
    MfmaEpilogue epi;
    MfmaPerm perm;
};

__device__ __forceinline__ uint32_t bits_of(uint64_t i, int start, const int* pos, int len) {
    if (start >= 0) return (uint32_t)(i >> start) & ((1u << len) - 1u);
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
        if (b >= len) break;
        v |= (uint32_t)((i >> pos[b]) & 1ull) << b;
    }
    return v;
}

// Factor of amplitude i: the per-tile part pg[g] times the inside chunk's
// entry tin[g][v] (tin: that chunk's table slice in LDS, rows of kTinRow
// entries: the pad entry puts lanes that read one v under different anchor
// values g on different banks).
__device__ __forceinline__ float2 epi_factor(const MfmaEpilogue& e, const float2* pg, const float2* tin, uint64_t i) {
    if ((i & e.reqmask) != e.reqmask) return make_float2(1.f, 0.f);
    const uint32_t g = bits_of(i, e.gstart, e.gpos, e.ng);
    float2 f = pg[g];
    if (e.inside >= 0) {
        const int c = e.inside;
        const float2 t = tin[g * kTinRow + bits_of(i, e.cstart[c], e.cpos[c], e.clen[c])];
        f = make_float2(f.x * t.x - f.y * t.y, f.x * t.y + f.y * t.x);
    }
    return f;
}

__device__ __forceinline__ float2 cmul2(float2 f, float xr, float xi) {
    return make_float2(f.x * xr - f.y * xi, f.x * xi + f.y * xr);
}

__device__ __forceinline__ uint32_t padded(uint32_t t) { return t + ((t >> 5) << 1); }

// The parts of an amplitude's table lookup that depend only on its offset in
// the tile (the anchor lies in the tile): anchor value g (bits 0-4), the
// inside chunk's bits that lie in the tile (bits 5-9), and whether the
// required bits that lie in the tile are set (bit 10).
__device__ __forceinline__ uint32_t epi_lane_entry(const MfmaEpilogue& e, uint64_t off, uint64_t tilemask) {
    const uint64_t rin = e.reqmask & tilemask;
    uint32_t v = bits_of(off, e.gstart, e.gpos, e.ng);
    if (e.inside >= 0) {
        const int c = e.inside;
        v |= bits_of(off, e.cstart[c], e.cpos[c], e.clen[c]) << 5;
    }
    if ((off & rin) == rin) v |= 1u << 10;
    return v;
}

// Factor of an amplitude from its lane entry and the tile's part: the
// required bits outside the tile set (req_out) and the inside chunk's bits
// outside the tile (vout).
__device__ __forceinline__ float2 epi_lane_factor(const MfmaEpilogue& e, const float2* pg, const float2* tin,
                                                  uint32_t ent, bool req_out, uint32_t vout) {
    if (!req_out || !(ent & (1u << 10))) return make_float2(1.f, 0.f);
    const uint32_t g = ent & 31u;
    float2 f = pg[g];
    if (e.inside >= 0) {
        const float2 t = tin[g * kTinRow + (((ent >> 5) & 31u) | vout)];
        f = make_float2(f.x * t.x - f.y * t.y, f.x * t.y + f.y * t.x);
    }
    return f;
}

__device__ __forceinline__ uint64_t tile_base(uint64_t t, const MfmaArgs& a) {
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        if (j >= a.ntile_bits) break;
        const uint64_t p = (uint64_t)a.tpos[j];
        t = ((t >> p) << (p + 1)) | (t & ((1ull << p) - 1ull));
    }
    return t;  // amplitude index of the tile's first element (even)
}

// Per-tile factors of the diagonal table's chunks outside the tile (one per
// anchor value, first lanes) into pg.
__device__ __forceinline__ void outside_factors(const MfmaArgs& a, float2* pg, uint64_t base, int lane) {
    if (lane >= (1 << a.epi.ng)) return;
    float2 f = make_float2(1.f, 0.f);
    for (int c = 0; c < a.epi.nc; ++c) {
        if (!((a.epi.outside >> c) & 1u)) continue;
        const float2 v = a.epi.tab[(lane * a.epi.nc + c) * 32 +
                                   bits_of(base, a.epi.cstart[c], a.epi.cpos[c], a.epi.clen[c])];
        f = make_float2(f.x * v.x - f.y * v.y, f.x * v.y + f.y * v.x);
    }
    pg[lane] = f;
}

// The same with the loads issued together (the chunk loop unrolled): the
// tile waits for one table read instead of one per chunk.  9 more VGPRs;
// the permutation-folding kernel keeps the loop above (its prologue form
// would drop to one wave per SIMD).
__device__ __forceinline__ void outside_factors_together(const MfmaArgs& a, float2* pg, uint64_t base, int lane) {
    if (lane >= (1 << a.epi.ng)) return;
    float2 v[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        v[c] = make_float2(1.f, 0.f);
        if (c < a.epi.nc && ((a.epi.outside >> c) & 1u))
            v[c] = a.epi.tab[(lane * a.epi.nc + c) * 32 + bits_of(base, a.epi.cstart[c], a.epi.cpos[c], a.epi.clen[c])];
    }
    float2 f = v[0];
#pragma unroll
    for (int c = 1; c < 6; ++c) f = make_float2(f.x * v[c].x - f.y * v[c].y, f.x * v[c].y + f.y * v[c].x);
    pg[lane] = f;
}

// Diagonal table applied to a tile's amplitudes as they are read in.
template <int CHUNKS>
__device__ __forceinline__ void prologue(const MfmaArgs& a, const float2* pg, const float2* tin, float4 (&pf)[CHUNKS],
                                         uint64_t b4, uint32_t glane) {
#pragma unroll
    for (int j = 0; j < CHUNKS; ++j) {
        const uint64_t i0 = (b4 + a.gj[j] + glane) << 1;
        const float2 f0 = epi_factor(a.epi, pg, tin, i0), f1 = epi_factor(a.epi, pg, tin, i0 | 1ull);
        const float2 y0 = cmul2(f0, pf[j].x, pf[j].y), y1 = cmul2(f1, pf[j].z, pf[j].w);
        pf[j] = make_float4(y0.x, y0.y, y1.x, y1.y);
    }
}

// The block's product on the tile held in LDS (memory order), result back
// into LDS.  Ends before the final barrier.
template <int K, int NA, typename Args>
__device__ __forceinline__ void tile_product(float2* lds, const float (&A)[NA], const Args& a, uint32_t rd_lane,
                                             uint32_t wr_lane, const uint32_t* wtr) {
    if (K == 32) {
        floatx16 t1 = {0}, t2 = {0}, t3 = {0};
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const float2 x = lds[padded(a.tr[2 * s] | rd_lane)];
            t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(A[s], x.x, t1, 0, 0, 0);
            t2 = __builtin_amdgcn_mfma_f32_32x32x2f32(A[16 + s], x.y, t2, 0, 0, 0);
            t3 = __builtin_amdgcn_mfma_f32_32x32x2f32(A[32 + s], x.x + x.y, t3, 0, 0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t tau = wtr[8 * (r / 4) + (r % 4)] | wr_lane;
            lds[padded(tau)] = make_float2(t1[r] - t2[r], t3[r] - t1[r] - t2[r]);
        }
    } else {
        floatx16 p = {0}, q = {0};
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const float2 x = lds[padded(a.tr[2 * s] | rd_lane)];
            p = __builtin_amdgcn_mfma_f32_32x32x2f32(A[s], x.x, p, 0, 0, 0);
            q = __builtin_amdgcn_mfma_f32_32x32x2f32(A[s], x.y, q, 0, 0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const uint32_t tau = wtr[8 * (r / 4) + (r % 4)] | wr_lane;
            lds[padded(tau)] = make_float2(p[r] - q[r + 8], p[r + 8] + q[r]);
        }
    }
}

// DIAG: 0 none, 1 diagonal table after the block (epilogue), 2 before it (prologue).
template <int K, int DIAG>
__global__ __launch_bounds__(64) void k_matk_mfma(MfmaArgs a) {
    constexpr bool EPI = DIAG != 0;
    constexpr int TILE = 32 * K;            // amplitudes per tile
    constexpr int CHUNKS = TILE / 128;      // 64 lanes x float4 = 128 amplitudes
    constexpr int LDS_F2 = TILE + TILE / 16;
    __shared__ __attribute__((aligned(16))) float2 lds[LDS_F2];
    __shared__ float2 pg[EPI ? 32 : 1];     // epilogue: per-tile factor per anchor value
    __shared__ float2 tin[EPI ? 32 * kTinRow : 1];  // epilogue: the inside chunk's table slice [g][v]
    if (EPI && a.epi.inside >= 0) {
        for (int e = threadIdx.x; e < (32 << a.epi.ng); e += 64)
            tin[(e >> 5) * kTinRow + (e & 31)] = a.epi.tab[((e >> 5) * a.epi.nc + a.epi.inside) * 32 + (e & 31)];
    }
    const int lane = threadIdx.x;
    const int col = lane & 31;
    const int half = lane >> 5;

    // A operands for the whole kernel
    constexpr int NA = (K == 32) ? 48 : 8;
    float A[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) A[i] = a.apack[i * 64 + lane];

    const uint32_t glane = a.glane[lane];
    const uint32_t rd_lane = a.tr[half] | a.tc[col];        // B rows 2s + half
    const uint32_t wr_lane = a.tr[4 * half] | a.tc[col];    // D rows 8(r/4) + 4 half + r%4

    // table lookups of this lane's amplitudes: the in-tile parts, once
    uint64_t tilemask = 0;
    uint32_t ent[EPI ? CHUNKS : 1];
    if (EPI) {
        for (int j = 0; j < a.ntile_bits; ++j) tilemask |= 1ull << a.tpos[j];
#pragma unroll
        for (int j = 0; j < CHUNKS; ++j) {
            const uint64_t off = (uint64_t)(a.gj[j] + glane) << 1;
            ent[j] = epi_lane_entry(a.epi, off, tilemask) | (epi_lane_entry(a.epi, off | 1ull, tilemask) << 16);
        }
    }
    const uint64_t rout = a.epi.reqmask & ~tilemask;

    uint64_t t = blockIdx.x;
    const uint64_t stride = gridDim.x;
    float4 pf[CHUNKS];
    if (t < a.tiles) {
        const uint64_t b4 = tile_base(t, a) >> 1;
#pragma unroll
        for (int j = 0; j < CHUNKS; ++j) pf[j] = ld_nt(&a.sv4[b4 + a.gj[j] + glane]);
    }
    while (t < a.tiles) {
        const uint64_t b4 = tile_base(t, a) >> 1;
        bool req_out = true;
        uint32_t vout = 0;
        if (EPI) {
            outside_factors_together(a, pg, b4 << 1, lane);
            req_out = ((b4 << 1) & rout) == rout;
            if (a.epi.inside >= 0) {
                const int c = a.epi.inside;
                vout = bits_of(b4 << 1, a.epi.cstart[c], a.epi.cpos[c], a.epi.clen[c]);
            }
        }
        if (DIAG == 2) {
            __syncthreads();  // pg of this tile before the prologue reads it
#pragma unroll
            for (int j = 0; j < CHUNKS; ++j) {
                const float2 f0 = epi_lane_factor(a.epi, pg, tin, ent[j] & 0xffffu, req_out, vout);
                const float2 f1 = epi_lane_factor(a.epi, pg, tin, ent[j] >> 16, req_out, vout);
                const float2 y0 = cmul2(f0, pf[j].x, pf[j].y), y1 = cmul2(f1, pf[j].z, pf[j].w);
                pf[j] = make_float4(y0.x, y0.y, y1.x, y1.y);
            }
        }
        // tile -> LDS (memory order)
#pragma unroll
        for (int j = 0; j < CHUNKS; ++j) {
            const uint32_t tau = 2u * (uint32_t)(j * 64 + lane);
            *reinterpret_cast<float4*>(&lds[padded(tau)]) = pf[j];
        }
        __syncthreads();
        // prefetch the next tile, at a raised wave priority so that its loads
        // issue ahead of other waves' LDS and matrix work (measured: 1.5%
        // faster k = 5 passes at 28 qubits)
        const uint64_t tn = t + stride;
        if (tn < a.tiles) {
            const uint64_t n4 = tile_base(tn, a) >> 1;
            __builtin_amdgcn_s_setprio(2);
#pragma unroll
            for (int j = 0; j < CHUNKS; ++j) pf[j] = ld_nt(&a.sv4[n4 + a.gj[j] + glane]);
            __builtin_amdgcn_s_setprio(0);
        }
        tile_product<K>(lds, A, a, rd_lane, wr_lane, a.tr);
        __syncthreads();
        // LDS -> tile (memory order)
#pragma unroll
        for (int j = 0; j < CHUNKS; ++j) {
            const uint32_t tau = 2u * (uint32_t)(j * 64 + lane);
            float4 v = *reinterpret_cast<const float4*>(&lds[padded(tau)]);
            if (DIAG == 1) {
                const float2 f0 = epi_lane_factor(a.epi, pg, tin, ent[j] & 0xffffu, req_out, vout);
                const float2 f1 = epi_lane_factor(a.epi, pg, tin, ent[j] >> 16, req_out, vout);
                const float2 y0 = cmul2(f0, v.x, v.y), y1 = cmul2(f1, v.z, v.w);
                v = make_float4(y0.x, y0.y, y1.x, y1.y);
            }
            st_nt(v, &a.sv4[b4 + a.gj[j] + glane]);
        }
        __syncthreads();
        t = tn;
    }
}

__device__ __forceinline__ uint64_t permute_index(uint64_t v, const int* map, int len) {
    uint64_t r = 0;
    for (int j = 0; j < len; ++j) r |= ((v >> j) & 1ull) << map[j];
    return r;
}

// The block followed by a bit involution (a.perm), DIAG 0 or 2 (prologue).
// A wave takes an outer index t and its image t2 = beta(t) (skipping t2 < t)
// and handles both tiles with one LDS buffer: tile A goes to LDS while tile
// B's loads are in flight into registers, A's product is stored into B's
// place (B is already read: every lane stores exactly the addresses it
// loaded), then B's product into A's place.  The product's results are
// written to LDS at their places under the tile-local permutation, and the
// store reads LDS in memory order, so HBM writes stay coalesced runs.
//
// Which t a wave skips depends on t's low bits, so a plain grid stride (the
// low bits fixed per wave) would leave some waves with nearly every pair and
// others with none.  Wave b instead takes t = m G + (b ^ (m mod G)) in step
// m (G = gridDim.x, a power of two): every t once, each wave seeing every
// low-bit pattern.  The next pair's first tile is loaded while the current
// pair's second tile is multiplied.
__device__ __forceinline__ uint64_t perm_step_tile(uint64_t m, uint64_t G, uint64_t b) {
    return m * G + (b ^ (m & (G - 1)));
}

template <int K, int DIAG>
__global__ __launch_bounds__(64) void k_matk_mfma_perm(MfmaArgs a) {
    static_assert(DIAG == 0 || DIAG == 2, "the permuted store takes no epilogue");
    constexpr bool EPI = DIAG != 0;
    constexpr int TILE = 32 * K;
    constexpr int CHUNKS = TILE / 128;
    constexpr int LDS_F2 = TILE + TILE / 16;
    __shared__ __attribute__((aligned(16))) float2 lds[LDS_F2];
    __shared__ float2 pg[EPI ? 32 : 1];
    __shared__ float2 tin[EPI ? 32 * kTinRow : 1];
    const int lane = threadIdx.x;
    if (EPI && a.epi.inside >= 0) {
        for (int e = lane; e < (32 << a.epi.ng); e += 64)
            tin[(e >> 5) * kTinRow + (e & 31)] = a.epi.tab[((e >> 5) * a.epi.nc + a.epi.inside) * 32 + (e & 31)];
    }
    const int col = lane & 31;
    const int half = lane >> 5;
    constexpr int NA = (K == 32) ? 48 : 8;
    float A[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) A[i] = a.apack[i * 64 + lane];
    const uint32_t glane = a.glane[lane];
    const uint32_t rd_lane = a.tr[half] | a.tc[col];
    // results are written to LDS at their permuted places (the permutation
    // acts bitwise, so it maps the row part and the lane part separately),
    // and the store then reads LDS in memory order
    const uint32_t wr_lane =
        (uint32_t)permute_index((uint64_t)(a.tr[4 * half] | a.tc[col]), a.perm.tmap, a.ntile_bits);
    __syncthreads();

    const uint64_t G = gridDim.x, b = blockIdx.x, steps = a.tiles / G;
    // the first step m >= from whose tile leads its pair (t <= beta(t))
    auto next_lead = [&](uint64_t from) {
        uint64_t m = from;
        for (; m < steps; ++m) {
            const uint64_t t = perm_step_tile(m, G, b);
            if (permute_index(t, a.perm.rmap, a.perm.nr) >= t) break;
        }
        return m;
    };
    uint64_t m = next_lead(0);
    float4 pf[CHUNKS];
    if (m < steps) {
        const uint64_t b4 = tile_base(perm_step_tile(m, G, b), a) >> 1;
#pragma unroll
        for (int j = 0; j < CHUNKS; ++j) pf[j] = ld_nt(&a.sv4[b4 + a.gj[j] + glane]);
    }
    while (m < steps) {
        const uint64_t t = perm_step_tile(m, G, b);
        const uint64_t t2 = permute_index(t, a.perm.rmap, a.perm.nr);
        const int ntiles = t2 == t ? 1 : 2;
        const uint64_t b4a = tile_base(t, a) >> 1, b4b = tile_base(t2, a) >> 1;
        const uint64_t mn = next_lead(m + 1);
        for (int s = 0; s < ntiles; ++s) {
            const uint64_t src = s ? b4b : b4a;
            if (EPI) {
                outside_factors(a, pg, src << 1, lane);
                __syncthreads();
                prologue(a, pg, tin, pf, src, glane);
            }
#pragma unroll
            for (int j = 0; j < CHUNKS; ++j) {
                const uint32_t tau = 2u * (uint32_t)(j * 64 + lane);
                *reinterpret_cast<float4*>(&lds[padded(tau)]) = pf[j];
            }
            __syncthreads();
            if (s + 1 < ntiles) {
                // the partner tile, read before this tile's result lands on it
#pragma unroll
                for (int j = 0; j < CHUNKS; ++j) pf[j] = ld_nt(&a.sv4[b4b + a.gj[j] + glane]);
            } else if (mn < steps) {
                // the next pair's first tile: no other wave touches it
                const uint64_t n4 = tile_base(perm_step_tile(mn, G, b), a) >> 1;
#pragma unroll
                for (int j = 0; j < CHUNKS; ++j) pf[j] = ld_nt(&a.sv4[n4 + a.gj[j] + glane]);
            }
            tile_product<K>(lds, A, a, rd_lane, wr_lane, a.ptr);
            __syncthreads();
            const uint64_t dst = (ntiles - 1 - s) ? b4b : b4a;
#pragma unroll
            for (int j = 0; j < CHUNKS; ++j) {
                const uint32_t tau = 2u * (uint32_t)(j * 64 + lane);
                st_nt(*reinterpret_cast<const float4*>(&lds[padded(tau)]), &a.sv4[dst + a.gj[j] + glane]);
            }
            __syncthreads();
        }
        m = mn;
    }
}

// Wide tiles for a plain 5-qubit block (no table, no permutation): the tile
// is the 5 targets and the 7 lowest other bits, 4096 amplitudes, so that it
// is read and written as runs of at least 128 contiguous amplitudes (1 KB)
// instead of 32.  A workgroup of four waves moves the tile through LDS, each
// lane 8 float4 each way, and wave w multiplies column group w (the 32
// columns whose two highest free bits equal w).  Streaming probes of this
// shape (one tile per workgroup, far target bits) moved 28 qubits in
// 0.69-0.72 ms against 0.85 ms for the one-wave 1024-amplitude tile loop.
constexpr int kWideBits = 12;     // 5 targets + 7 free bits
constexpr int kWideThreads = 256; // four waves
constexpr int kWideUnits = 8;     // float4 per lane each way

struct MfmaWideArgs {
    float4* sv4;
    uint64_t tiles;
    int tpos[kWideBits];          // sorted amplitude bit positions of the tile
    uint32_t glane[kWideThreads]; // float4 offset of thread l's first element: pdep(2 l) / 2
    uint32_t gj[kWideUnits];      // float4 offset of element group u: pdep(512 u) / 2
    uint32_t tr[32];              // tile-index bits of target sub-index r
    uint32_t tc[128];             // tile-index bits of column c
    const float* apack;
};

__device__ __forceinline__ uint64_t wide_tile_base(uint64_t t, const MfmaWideArgs& a) {
#pragma unroll
    for (int j = 0; j < kWideBits; ++j) {
        const uint64_t p = (uint64_t)a.tpos[j];
        t = ((t >> p) << (p + 1)) | (t & ((1ull << p) - 1ull));
    }
    return t;
}

__global__ __launch_bounds__(kWideThreads) void k_matk_mfma_wide(MfmaWideArgs a) {
    constexpr int TILE = 1 << kWideBits;
    __shared__ __attribute__((aligned(16))) float2 lds[TILE + TILE / 16];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int col = lane & 31, half = lane >> 5;
    float A[48];
#pragma unroll
    for (int i = 0; i < 48; ++i) A[i] = a.apack[i * 64 + lane];
    const uint32_t rd_lane = a.tr[half] | a.tc[32 * w + col];
    const uint32_t wr_lane = a.tr[4 * half] | a.tc[32 * w + col];
    const uint64_t b4 = wide_tile_base(blockIdx.x, a) >> 1;
    const uint32_t gl = a.glane[tid];
    float4 v[kWideUnits];
#pragma unroll
    for (int u = 0; u < kWideUnits; ++u) v[u] = ld_nt(&a.sv4[b4 + a.gj[u] + gl]);
#pragma unroll
    for (int u = 0; u < kWideUnits; ++u) {
        const uint32_t tau = 2u * (uint32_t)(tid + u * kWideThreads);
        *reinterpret_cast<float4*>(&lds[padded(tau)]) = v[u];
    }
    __syncthreads();
    tile_product<32>(lds, A, a, rd_lane, wr_lane, a.tr);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kWideUnits; ++u) {
        const uint32_t tau = 2u * (uint32_t)(tid + u * kWideThreads);
        st_nt(*reinterpret_cast<const float4*>(&lds[padded(tau)]), &a.sv4[b4 + a.gj[u] + gl]);
    }
}

// pdep of `v` into the bit positions listed in `pos` (ascending).
uint64_t deposit(uint64_t v, const std::vector<int>& pos) {
    uint64_t r = 0;
    for (size_t j = 0; j < pos.size(); ++j)
        if ((v >> j) & 1ull) r |= 1ull << pos[j];
    return r;
}

}  // namespace

namespace aqs_hip {

// Per-lane operand order.
//   k = 5: [Ur | Ui | Ur + Ui] x 16 steps: apack[(m*16 + s)*64 + l] = M_m[l % 32][2s + l/32]
//   k = 4: apack[s*64 + l] = [Ur; Ui][l % 32][2s + l/32], s < 8
std::vector<float> pack_mfma_operand(int k, const double* u) {
    const int K = 1 << k;
    auto ur = [&](int r, int c) { return u[2 * (r * K + c)]; };
    auto ui = [&](int r, int c) { return u[2 * (r * K + c) + 1]; };
    std::vector<float> out;
    if (k == 5) {
        out.resize(3 * 16 * 64);
        for (int m = 0; m < 3; ++m)
            for (int s = 0; s < 16; ++s)
                for (int l = 0; l < 64; ++l) {
                    const int r = l % 32, c = 2 * s + l / 32;
                    const double v = m == 0 ? ur(r, c) : (m == 1 ? ui(r, c) : ur(r, c) + ui(r, c));
                    out[(m * 16 + s) * 64 + l] = (float)v;
                }
    } else if (k == 4) {
        out.resize(8 * 64);
        for (int s = 0; s < 8; ++s)
            for (int l = 0; l < 64; ++l) {
                const int R = l % 32, c = 2 * s + l / 32;
                out[s * 64 + l] = (float)(R < 16 ? ur(R, c) : ui(R - 16, c));
            }
    }
    return out;
}

namespace {

// Plain 5-qubit blocks on states of at least kWideMinQubits qubits take the
// wide tiles.  Below that a launch has too few tiles of 4096 amplitudes for
// the chip (brickwork depth 20, block schedule: 0.21 against 0.18-0.20 ms at
// 14 qubits, equal at 18, 0.38 against 0.40 ms at 20, 51.5 against 57.4 ms
// at 28).
constexpr bool kWideBlocks = true;
constexpr int kWideMinQubits = 18;

// A plain k = 5 block on wide tiles (k_matk_mfma_wide); false when it does
// not apply.
bool apply_matk_mfma_wide(aqs_state s, const aqs_op& op, const float* dev_apack) {
    const int n = s->n;
    if (!kWideBlocks || op.ntargets != 5 || n < kWideMinQubits) return false;
    std::vector<int> tbits;
    for (int j = 0; j < 5; ++j) tbits.push_back(qpos(n, op.targets[j]));
    std::vector<int> freeb;
    for (int b = 0; b < n && (int)freeb.size() < kWideBits - 5; ++b)
        if (std::find(tbits.begin(), tbits.end(), b) == tbits.end()) freeb.push_back(b);
    std::vector<int> T = tbits;
    T.insert(T.end(), freeb.begin(), freeb.end());
    std::sort(T.begin(), T.end());
    if (T.size() != (size_t)kWideBits || T[0] != 0) return false;
    auto tidx = [&](int bit) { return (int)(std::find(T.begin(), T.end(), bit) - T.begin()); };
    MfmaWideArgs a{};
    a.sv4 = (float4*)s->amps;
    a.tiles = s->N >> kWideBits;
    for (int j = 0; j < kWideBits; ++j) a.tpos[j] = T[j];
    for (int l = 0; l < kWideThreads; ++l) a.glane[l] = (uint32_t)(deposit((uint64_t)(2 * l), T) >> 1);
    for (int u = 0; u < kWideUnits; ++u) a.gj[u] = (uint32_t)(deposit((uint64_t)(2 * kWideThreads * u), T) >> 1);
    for (int r = 0; r < 32; ++r) {
        uint32_t v = 0;
        for (int j = 0; j < 5; ++j)
            if ((r >> (4 - j)) & 1) v |= 1u << tidx(tbits[j]);
        a.tr[r] = v;
    }
    for (int c = 0; c < 128; ++c) {
        uint32_t v = 0;
        for (int j = 0; j < kWideBits - 5; ++j)
            if ((c >> j) & 1) v |= 1u << tidx(freeb[j]);
        a.tc[c] = v;
    }
    a.apack = dev_apack;
    ProfScope ps(s, kSlotMfma);
    hipLaunchKernelGGL(k_matk_mfma_wide, dim3((unsigned)a.tiles), dim3(kWideThreads), 0, s->stream, a);
    AQS_HIP(hipGetLastError());
    return true;
}

}  // namespace

bool apply_matk_mfma(aqs_state s, const aqs_op& op, const float* dev_apack, const DiagEpilogueDesc* epi,
                     bool before, const std::vector<int>* beta) {
    const int k = op.ntargets;
    if (s->precision != AQS_C64 || op.ncontrols != 0 || op.more_controls || (k != 4 && k != 5)) return false;
    if (s->n < k + 5 || s->n > 36) return false;
    if (!epi && !beta && apply_matk_mfma_wide(s, op, dev_apack)) return true;
    const int n = s->n;
    // target bit positions (op order: targets[0] = most significant sub-index bit)
    std::vector<int> tbits;
    for (int j = 0; j < k; ++j) tbits.push_back(qpos(n, op.targets[j]));
    std::vector<int> sorted_t = tbits;
    std::sort(sorted_t.begin(), sorted_t.end());
    std::vector<int> freeb;
    for (int b = 0; b < n && (int)freeb.size() < 5; ++b)
        if (std::find(tbits.begin(), tbits.end(), b) == tbits.end()) freeb.push_back(b);
    std::vector<int> T = sorted_t;
    T.insert(T.end(), freeb.begin(), freeb.end());
    std::sort(T.begin(), T.end());
    // the tile covers bits 0..4, so float4 pairs and 256-byte runs hold
    if (T[0] != 0 || T.size() != (size_t)(k + 5)) return false;
    auto tidx = [&](int bit) { return (int)(std::find(T.begin(), T.end(), bit) - T.begin()); };

    MfmaArgs a{};
    a.sv4 = (float4*)s->amps;
    a.tiles = s->N >> (k + 5);
    a.ntile_bits = k + 5;
    for (int j = 0; j < k + 5; ++j) a.tpos[j] = T[j];
    const int chunks = (32 << k) / 128;
    for (int j = 0; j < chunks; ++j) a.gj[j] = (uint32_t)(deposit((uint64_t)(128 * j), T) >> 1);
    for (int l = 0; l < 64; ++l) a.glane[l] = (uint32_t)(deposit((uint64_t)(2 * l), T) >> 1);
    const int K = 1 << k;
    for (int r = 0; r < 32; ++r) {
        uint32_t v = 0;
        if (r < K)
            for (int j = 0; j < k; ++j)
                if ((r >> (k - 1 - j)) & 1) v |= 1u << tidx(tbits[j]);
        a.tr[r] = v;
    }
    for (int c = 0; c < 32; ++c) {
        uint32_t v = 0;
        for (int j = 0; j < 5; ++j)
            if ((c >> j) & 1) v |= 1u << tidx(freeb[j]);
        a.tc[c] = v;
    }
    a.apack = dev_apack;
    if (beta) {
        if ((int)beta->size() != n || (epi && !before)) return false;
        std::vector<int> R;
        for (int b = 0; b < n; ++b)
            if (std::find(T.begin(), T.end(), b) == T.end()) R.push_back(b);
        for (int b : T)
            if (std::find(T.begin(), T.end(), (*beta)[b]) == T.end()) return false;
        MfmaPerm& pm = a.perm;
        pm.nr = (int)R.size();
        for (size_t j = 0; j < T.size(); ++j) pm.tmap[j] = tidx((*beta)[T[j]]);
        for (int r = 0; r < 32; ++r) {
            uint32_t v = 0;
            for (size_t j = 0; j < T.size(); ++j)
                if ((a.tr[r] >> j) & 1u) v |= 1u << pm.tmap[j];
            a.ptr[r] = v;
        }
        for (size_t j = 0; j < R.size(); ++j)
            pm.rmap[j] = (int)(std::find(R.begin(), R.end(), (*beta)[R[j]]) - R.begin());
    }
    if (epi) {
        // the anchor must lie in the tile (it indexes the per-tile factors);
        // a chunk is either wholly outside the tile or looked up per amplitude
        auto in_tile = [&](int b) { return std::find(T.begin(), T.end(), b) != T.end(); };
        if (epi->anchor.size() > 5 || epi->chunks.size() > 6 || !epi->dev_tab) return false;
        for (int b : epi->anchor)
            if (!in_tile(b)) return false;
        MfmaEpilogue& e = a.epi;
        e.tab = reinterpret_cast<const float2*>(epi->dev_tab);
        for (int b : epi->required) e.reqmask |= 1ull << b;
        auto consecutive = [](const std::vector<int>& p) {
            for (size_t j = 1; j < p.size(); ++j)
                if (p[j] != p[0] + (int)j) return -1;
            return p.empty() ? 0 : p[0];
        };
        e.ng = (int)epi->anchor.size();
        for (int j = 0; j < e.ng; ++j) e.gpos[j] = epi->anchor[j];
        e.gstart = consecutive(epi->anchor);
        e.nc = (int)epi->chunks.size();
        for (int c = 0; c < e.nc; ++c) {
            const auto& ch = epi->chunks[c];
            if (ch.size() > 5) return false;
            e.clen[c] = (int)ch.size();
            for (size_t j = 0; j < ch.size(); ++j) e.cpos[c][j] = ch[j];
            e.cstart[c] = consecutive(ch);
            bool outside = true;
            for (int b : ch) outside = outside && !in_tile(b);
            if (outside) e.outside |= 1u << c;
        }
        e.inside = -1;
        for (int c = 0; c < e.nc; ++c) {
            if ((e.outside >> c) & 1u) continue;
            if (e.inside >= 0) return false;  // two chunks touch the tile: separate pass
            e.inside = c;
        }
    }
    const int cus = device_cu_count();
    uint64_t blocks = std::min<uint64_t>(a.tiles, (uint64_t)cus * 16);
    if (beta) {
        // the permuted store's wave-to-tile map needs a power-of-two grid
        // (tiles is one too, so the grid divides it)
        while (blocks & (blocks - 1)) blocks &= blocks - 1;
    }
    ProfScope ps(s, kSlotMfma);
    const int mode = !epi ? 0 : before ? 2 : 1;
    const dim3 grid((unsigned)blocks), block(64);
    if (beta) {
        if (k == 5) {
            if (mode == 0) hipLaunchKernelGGL((k_matk_mfma_perm<32, 0>), grid, block, 0, s->stream, a);
            if (mode == 2) hipLaunchKernelGGL((k_matk_mfma_perm<32, 2>), grid, block, 0, s->stream, a);
        } else {
            if (mode == 0) hipLaunchKernelGGL((k_matk_mfma_perm<16, 0>), grid, block, 0, s->stream, a);
            if (mode == 2) hipLaunchKernelGGL((k_matk_mfma_perm<16, 2>), grid, block, 0, s->stream, a);
        }
        AQS_HIP(hipGetLastError());
        return true;
    }
    if (k == 5) {
        if (mode == 0) hipLaunchKernelGGL((k_matk_mfma<32, 0>), grid, block, 0, s->stream, a);
        if (mode == 1) hipLaunchKernelGGL((k_matk_mfma<32, 1>), grid, block, 0, s->stream, a);
        if (mode == 2) hipLaunchKernelGGL((k_matk_mfma<32, 2>), grid, block, 0, s->stream, a);
    } else {
        if (mode == 0) hipLaunchKernelGGL((k_matk_mfma<16, 0>), grid, block, 0, s->stream, a);
        if (mode == 1) hipLaunchKernelGGL((k_matk_mfma<16, 1>), grid, block, 0, s->stream, a);
        if (mode == 2) hipLaunchKernelGGL((k_matk_mfma<16, 2>), grid, block, 0, s->stream, a);
    }
    AQS_HIP(hipGetLastError());
    return true;
}

}  // namespace aqs_hip
