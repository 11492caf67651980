// Tile pass: a run of one-qubit, controlled and diagonal gates applied in
// one HBM sweep (cache blocking).
//
// The state is cut into tiles of 2^m amplitudes that agree on every index
// bit outside the pass's local set (m ascending bit positions, bits 0..4
// always among them, so a tile is read and written as runs of >= 32
// contiguous amplitudes).  A workgroup of 2^(m-4) threads takes a tile into
// LDS and runs the pass's segments: in each segment every thread holds 16
// amplitudes of the tile in registers -- 4 local positions are its register
// slots, the thread id is spread over the other m-4 -- and applies the
// segment's ops to them:
//   MAT     2x2 matrix across one register slot (with controls anywhere);
//   MATR    real 2x2 across one register slot (H, RotY, controlled or not);
//   MATP    uncontrolled one-qubit gate as a phase on the slot's 1-half and
//           a real 2x2 (the gate's other phases are carried by the host
//           into the qubit's next gate); consecutive MATP ops on distinct
//           slots are dispatched once, as a set;
//   XCH     exchange across one register slot (X, CX, Toffoli, Or);
//   PHASES  a run of diagonal gates (Z, S, T, Phase, RZ, CZ, CPhase and
//           controlled forms) as phase terms "multiply by z where these bits
//           are all 1"; long runs accumulate one factor per subset of the
//           register slots per thread and multiply each amplitude once.
// Between segments the tile goes back to LDS and is re-read in the next
// segment's layout.  The host side (tiling.cpp) chooses the local sets, the
// segments and the op order; every gate of the reference's per-gate loop
// (src/quantum.cpp:277-291) that it places here is applied exactly once, in
// an order that only exchanges commuting gates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "internal.h"
#include "schedule.h"

using namespace aqs_hip;

namespace {

// The program arrays are separate __restrict__ kernel parameters: the
// compiler can then prove that the state writes never alias them and reads
// them with scalar loads (one per workgroup, off the vector memory path).
struct TileArgs {
    uint64_t tiles;
    int nseg;
    int nfold;
    int local[kTileMaxLocal];
};

// Two padding amplitudes after every 32 keep LDS accesses of strided layouts
// spread over the banks and keep amplitude pairs 16-byte aligned.
__device__ __forceinline__ uint32_t lds_pad(uint32_t l) { return l + ((l >> 5) << 1); }
constexpr uint32_t lds_len(uint32_t tile) { return tile + (tile >> 4); }

template <typename T2>
struct Cplx;
template <>
struct Cplx<float2> {
    using R = float;
};
template <>
struct Cplx<double2> {
    using R = double;
};

template <typename T2, typename R>
__device__ __forceinline__ T2 cmul(R fr, R fi, T2 x) {
    T2 y;
    y.x = fr * x.x - fi * x.y;
    y.y = fr * x.y + fi * x.x;
    return y;
}

template <typename T2, typename R>
__device__ __forceinline__ void butterfly(T2& p, T2& q, const R (&m)[8]) {
    T2 u, v;
    u.x = m[0] * p.x - m[1] * p.y + m[2] * q.x - m[3] * q.y;
    u.y = m[0] * p.y + m[1] * p.x + m[2] * q.y + m[3] * q.x;
    v.x = m[4] * p.x - m[5] * p.y + m[6] * q.x - m[7] * q.y;
    v.y = m[4] * p.y + m[5] * p.x + m[6] * q.y + m[7] * q.x;
    p = u;
    q = v;
}

// Real 2x2 (m[0..3] row-major) on complex amplitudes: 8 multiplies per pair
// instead of 16, for one-qubit gates whose matrix is real (H, RotY).
template <typename T2, typename R>
__device__ __forceinline__ void butterfly_real(T2& p, T2& q, const R (&m)[8]) {
    T2 u, v;
    u.x = m[0] * p.x + m[1] * q.x;
    u.y = m[0] * p.y + m[1] * q.y;
    v.x = m[2] * p.x + m[3] * q.x;
    v.y = m[2] * p.y + m[3] * q.y;
    p = u;
    q = v;
}

// q times z (m[4], m[5]), then the rotation [[c, -s], [s, c]] (c, s >= 0)
// as three shears, t = m[0] = s / (1 + c), s = m[1]: p -= t q, q += s p,
// p -= t q.  A one-qubit gate factored as a rotation between phases in 10
// multiply-adds per pair (a complex 2x2 takes 16).
template <typename T2, typename R>
__device__ __forceinline__ void butterfly_rotp(T2& p, T2& q, const R (&m)[8]) {
    T2 w;
    w.x = m[4] * q.x - m[5] * q.y;
    w.y = m[4] * q.y + m[5] * q.x;
    p.x -= m[0] * w.x;
    p.y -= m[0] * w.y;
    w.x += m[1] * p.x;
    w.y += m[1] * p.y;
    p.x -= m[0] * w.x;
    p.y -= m[0] * w.y;
    q = w;
}

// Pair operations across register slot SL (OP: 0 complex 2x2, 1 exchange,
// 2 real 2x2, 3 phase then real 2x2).  Form 0: every pair.  Form 1:
// pairs whose register index holds the control slots `creg` (uniform test,
// the partner has the same control bits since a control is never the
// target).  Form 2: per-thread predicate `ok(j)`, results selected.
template <int SL, int OP, int FORM, typename T2, typename R, typename OK>
__device__ __forceinline__ void pairs(T2 (&x)[16], const R (&m)[8], uint32_t creg, OK ok) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (j & (1 << SL)) continue;
        const int k = j | (1 << SL);
        if (FORM == 1 && ((uint32_t)j & creg) != creg) continue;
        T2 p = x[j], q = x[k];
        if (OP == 1) {
            const T2 t = p;
            p = q;
            q = t;
        } else if (OP == 2) {
            butterfly_real(p, q, m);
        } else if (OP == 3) {
            butterfly_rotp(p, q, m);
        } else {
            butterfly(p, q, m);
        }
        if (FORM == 2) {
            const bool c = ok(j);
            x[j] = c ? p : x[j];
            x[k] = c ? q : x[k];
        } else {
            x[j] = p;
            x[k] = q;
        }
    }
}

// One slot of a MATP set starting at op `at`: the op for slot SL is the
// k-th of the set, k = number of the mask's slots below SL.
template <int SL, int M, typename T2, typename R, typename OK>
__device__ __forceinline__ void matp_set_slot(T2 (&x)[16], const TileOpDev* __restrict__ ops,
                                              const double* __restrict__ m128, uint32_t at, OK ok) {
    if constexpr (((M >> SL) & 1) != 0) {
        constexpr int k = __builtin_popcount(M & ((1 << SL) - 1));
        R m[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (sizeof(R) == 4)
                m[i] = ops[at + k].mf[i];
            else
                m[i] = m128[8 * (size_t)(at + k) + i];
        }
        pairs<SL, 3, 0>(x, m, 0u, ok);
    }
}

// A MATP set: one-qubit gates on the register slots of mask M, in slot order.
template <int M, typename T2, typename R, typename OK>
__device__ __forceinline__ void matp_set(T2 (&x)[16], const TileOpDev* __restrict__ ops,
                                         const double* __restrict__ m128, uint32_t at, OK ok) {
    matp_set_slot<0, M, T2, R>(x, ops, m128, at, ok);
    matp_set_slot<1, M, T2, R>(x, ops, m128, at, ok);
    matp_set_slot<2, M, T2, R>(x, ops, m128, at, ok);
    matp_set_slot<3, M, T2, R>(x, ops, m128, at, ok);
}

template <typename R>
__device__ __forceinline__ void cmul_acc(R& ar, R& ai, R zr, R zi) {
    const R t = ar * zr - ai * zi;
    ai = ar * zi + ai * zr;
    ar = t;
}

// A run of phase terms, grouped by register mask.  Per thread each group's
// product P is formed (thread and tile bits tested per term); the direct
// form multiplies P into the amplitudes whose register index holds the
// mask, the zeta form into the subset factor f[mask], then F[j] = product
// of f[s] over the subsets s of j and one multiplication per amplitude.
// Product over one group of terms (same register mask) for this thread;
// folded terms read their per-tile value from LDS.
template <typename T2, typename R>
__device__ __forceinline__ void group_product(const TileTermDev* __restrict__ tt, const double* __restrict__ z128,
                                              const T2* fold_val, uint32_t k, uint32_t g, uint32_t lthr,
                                              uint64_t base, R& pr, R& pi) {
    pr = R(1);
    pi = R(0);
    // plain terms (z from the program, scalar loads), then the folded ones
    // (z from LDS); one loop with either source made the compiler select
    // between the two address spaces and wait on a flat load per term
    // each term is read whole (two 16-byte scalar loads, one wait) and
    // tested without short-circuit (a divergent branch per term otherwise)
    const uint32_t mid = k + g - tt[k].folded;
    const uint4* w = reinterpret_cast<const uint4*>(tt);
#pragma unroll 2
    for (uint32_t e = k; e < mid; ++e) {
        const uint4 a = w[2 * e], b = w[2 * e + 1];
        const uint64_t m_glb = (uint64_t)a.z | ((uint64_t)a.w << 32);
        const bool c = ((lthr & a.y) == a.y) & ((base & m_glb) == m_glb);
        R zr, zi;
        if constexpr (sizeof(R) == 4) {
            zr = __uint_as_float(b.x);
            zi = __uint_as_float(b.y);
        } else {
            zr = z128[2 * e];
            zi = z128[2 * e + 1];
        }
        cmul_acc(pr, pi, c ? zr : R(1), c ? zi : R(0));
    }
#pragma unroll 2
    for (uint32_t e = mid; e < k + g; ++e) {
        const uint4 a = w[2 * e], b = w[2 * e + 1];
        const bool c = (lthr & a.y) == a.y;
        const T2 v = fold_val[b.z];
        cmul_acc(pr, pi, c ? v.x : R(1), c ? v.y : R(0));
    }
}

template <typename T2, typename R>
__device__ __forceinline__ void phases_direct(T2 (&x)[16], const TileTermDev* __restrict__ tt,
                                           const double* __restrict__ z128, const T2* fold_val, uint32_t count,
                                           uint32_t lthr, uint64_t base) {
    for (uint32_t k = 0; k < count;) {
        const uint32_t g = tt[k].group, mr = tt[k].m_reg;
        const bool real = tt[k].real != 0;
        R pr, pi;
        group_product<T2, R>(tt, z128, fold_val, k, g, lthr, base, pr, pi);
        k += g;
        // one case per register mask: each multiplies a fixed set of
        // registers in place (a runtime test per register made the compiler
        // carry extra copies of the register array through the op loop);
        // real products (CZ, Z, controlled Z) scale: 2 multiplies per
        // amplitude instead of 4
        if (real) {
            switch (mr) {
#define AQS_REAL_CASE(M)                                                      \
    case M:                                                                   \
        _Pragma("unroll") for (int j = 0; j < 16; ++j) if ((j & (M)) == (M)) { \
            x[j].x *= pr;                                                     \
            x[j].y *= pr;                                                     \
        }                                                                     \
        break;
                AQS_REAL_CASE(0) AQS_REAL_CASE(1) AQS_REAL_CASE(2) AQS_REAL_CASE(3)
                AQS_REAL_CASE(4) AQS_REAL_CASE(5) AQS_REAL_CASE(6) AQS_REAL_CASE(7)
                AQS_REAL_CASE(8) AQS_REAL_CASE(9) AQS_REAL_CASE(10) AQS_REAL_CASE(11)
                AQS_REAL_CASE(12) AQS_REAL_CASE(13) AQS_REAL_CASE(14)
                default:  // mask 15
                    x[15].x *= pr;
                    x[15].y *= pr;
                    break;
#undef AQS_REAL_CASE
            }
            continue;
        }
        switch (mr) {
#define AQS_MASK_CASE(M)                                                      \
    case M:                                                                   \
        _Pragma("unroll") for (int j = 0; j < 16; ++j) if ((j & (M)) == (M)) \
            x[j] = cmul<T2, R>(pr, pi, x[j]);                                 \
        break;
            AQS_MASK_CASE(0) AQS_MASK_CASE(1) AQS_MASK_CASE(2) AQS_MASK_CASE(3)
            AQS_MASK_CASE(4) AQS_MASK_CASE(5) AQS_MASK_CASE(6) AQS_MASK_CASE(7)
            AQS_MASK_CASE(8) AQS_MASK_CASE(9) AQS_MASK_CASE(10) AQS_MASK_CASE(11)
            AQS_MASK_CASE(12) AQS_MASK_CASE(13) AQS_MASK_CASE(14)
            default:  // mask 15
                x[15] = cmul<T2, R>(pr, pi, x[15]);
                break;
#undef AQS_MASK_CASE
        }
    }
}

template <typename T2, typename R>
__device__ __forceinline__ void phases_zeta(T2 (&x)[16], const TileTermDev* __restrict__ tt,
                                         const double* __restrict__ z128, const T2* fold_val, uint32_t count,
                                         uint32_t lthr, uint64_t base) {
    R fr[16], fi[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        fr[j] = R(1);
        fi[j] = R(0);
    }
    for (uint32_t k = 0; k < count;) {
        const uint32_t g = tt[k].group, mr = tt[k].m_reg;
        R pr, pi;
        group_product<T2, R>(tt, z128, fold_val, k, g, lthr, base, pr, pi);
        k += g;
        switch (mr) {
#define AQS_TERM_CASE(S) \
    case S: cmul_acc(fr[S], fi[S], pr, pi); break;
            AQS_TERM_CASE(0) AQS_TERM_CASE(1) AQS_TERM_CASE(2) AQS_TERM_CASE(3)
            AQS_TERM_CASE(4) AQS_TERM_CASE(5) AQS_TERM_CASE(6) AQS_TERM_CASE(7)
            AQS_TERM_CASE(8) AQS_TERM_CASE(9) AQS_TERM_CASE(10) AQS_TERM_CASE(11)
            AQS_TERM_CASE(12) AQS_TERM_CASE(13) AQS_TERM_CASE(14) default: cmul_acc(fr[15], fi[15], pr, pi); break;
#undef AQS_TERM_CASE
        }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j & (1 << b)) cmul_acc(fr[j], fi[j], fr[j ^ (1 << b)], fi[j ^ (1 << b)]);
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = cmul<T2, R>(fr[j], fi[j], x[j]);
}

// HBM <-> LDS for one tile.  Complex64 moves amplitude pairs (16 bytes),
// complex128 single amplitudes.
template <typename T2, int M, int T>
__device__ __forceinline__ void load_tile(const T2* __restrict__ sv, uint64_t base, const uint64_t* hi_tab, T2* lds,
                                          int tid) {
    if constexpr (sizeof(T2) == 8) {
        constexpr int U = (1 << M) / 2 / T;
        float4 v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const uint32_t l = 2u * (uint32_t)(tid + k * T);
            const uint64_t g = base | (l & 31u) | hi_tab[l >> 5];
            v[k] = ld_nt(reinterpret_cast<const float4*>(sv) + (g >> 1));
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const uint32_t l = 2u * (uint32_t)(tid + k * T);
            *reinterpret_cast<float4*>(&lds[lds_pad(l)]) = v[k];
        }
    } else {
        constexpr int U = (1 << M) / T;
        double2 v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const uint32_t l = (uint32_t)(tid + k * T);
            v[k] = reinterpret_cast<const double2*>(sv)[base | (l & 31u) | hi_tab[l >> 5]];
        }
#pragma unroll
        for (int k = 0; k < U; ++k) lds[lds_pad((uint32_t)(tid + k * T))] = v[k];
    }
}

template <typename T2, int M, int T>
__device__ __forceinline__ void store_tile(T2* __restrict__ sv, uint64_t base, const uint64_t* hi_tab, const T2* lds,
                                           int tid) {
    if constexpr (sizeof(T2) == 8) {
        constexpr int U = (1 << M) / 2 / T;
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const uint32_t l = 2u * (uint32_t)(tid + k * T);
            const uint64_t g = base | (l & 31u) | hi_tab[l >> 5];
            st_nt(*reinterpret_cast<const float4*>(&lds[lds_pad(l)]), reinterpret_cast<float4*>(sv) + (g >> 1));
        }
    } else {
        constexpr int U = (1 << M) / T;
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const uint32_t l = (uint32_t)(tid + k * T);
            reinterpret_cast<double2*>(sv)[base | (l & 31u) | hi_tab[l >> 5]] = lds[lds_pad(l)];
        }
    }
}

template <typename T2, int W>
__global__ __launch_bounds__(1 << W, (sizeof(T2) == 8 ? 4 : 2)) void k_tile(
    T2* __restrict__ sv, const TileSegDev* __restrict__ segs, const TileOpDev* __restrict__ ops,
    const TileTermDev* __restrict__ terms, const double* __restrict__ m128, const double* __restrict__ z128,
    const TileFoldDev* __restrict__ folds, const TileFoldTermDev* __restrict__ fold_terms,
    const double* __restrict__ fold_z128, TileArgs a) {
    using R = typename Cplx<T2>::R;
    constexpr int T = 1 << W;
    constexpr int M = W + kTileRegBits;
    constexpr int HI = 1 << (M - 5);
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    uint64_t* hi_tab = reinterpret_cast<uint64_t*>(smem);
    T2* lds = reinterpret_cast<T2*>(smem + HI * sizeof(uint64_t));
    T2* fold_val = lds + lds_len(1u << M);
    const int tid = threadIdx.x;
    // index offsets of the local positions 5..M-1
    for (int v = tid; v < HI; v += T) {
        uint64_t o = 0;
        for (int b = 0; b < M - 5; ++b)
            if ((v >> b) & 1) o |= 1ull << a.local[5 + b];
        hi_tab[v] = o;
    }
    __syncthreads();
    for (uint64_t t = blockIdx.x; t < a.tiles; t += gridDim.x) {
        // tile index -> amplitude index with zeros at the local positions
        uint64_t base = t;
        for (int p = 0; p < M; ++p) {
            const int q = a.local[p];
            base = ((base >> q) << (q + 1)) | (base & ((1ull << q) - 1ull));
        }
        load_tile<T2, M, T>(sv, base, hi_tab, lds, tid);
        // per-tile values of the folded phase terms (outer bits only)
        for (int f = tid; f < a.nfold; f += T) {
            R pr = R(1), pi = R(0);
            const TileFoldDev fd = folds[f];
            // whole 16-byte fold terms, four loads in flight
            const uint4* ftw = reinterpret_cast<const uint4*>(fold_terms);
#pragma unroll 4
            for (uint32_t e = fd.begin; e < fd.begin + fd.count; ++e) {
                const uint4 a = ftw[e];
                const uint64_t m = (uint64_t)a.x | ((uint64_t)a.y << 32);
                const bool c = (base & m) == m;
                R zr, zi;
                if constexpr (sizeof(R) == 4) {
                    zr = __uint_as_float(a.z);
                    zi = __uint_as_float(a.w);
                } else {
                    zr = fold_z128[2 * e];
                    zi = fold_z128[2 * e + 1];
                }
                cmul_acc(pr, pi, c ? zr : R(1), c ? zi : R(0));
            }
            T2 v;
            v.x = pr;
            v.y = pi;
            fold_val[f] = v;
        }
        __syncthreads();
        for (int sg = 0; sg < a.nseg; ++sg) {
            const TileSegDev& S = segs[sg];
            uint32_t lthr = 0;
#pragma unroll
            for (int b = 0; b < W; ++b) lthr |= ((uint32_t)(tid >> b) & 1u) << S.thr[b];
            // disjoint bits: pad(lthr | u) = pad(lthr) + pad(u), and pad(u)
            // is the same for the whole workgroup
            const uint32_t r0 = 1u << S.reg[0], r1 = 1u << S.reg[1], r2 = 1u << S.reg[2], r3 = 1u << S.reg[3];
            const uint32_t pthr = lds_pad(lthr);
            auto addr = [&](int j) {
                return pthr + lds_pad(((j & 1) ? r0 : 0u) | ((j & 2) ? r1 : 0u) | ((j & 4) ? r2 : 0u) |
                                      ((j & 8) ? r3 : 0u));
            };
            T2 x[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) x[j] = lds[addr(j)];
            const uint32_t ob = S.op_begin, oc = S.op_count;
            for (uint32_t o = 0; o < oc; ++o) {
                const TileOpDev& op = ops[ob + o];
                const int code = op.code;
                if (tile_code_is_phases(code)) {
                    if (code == kTileCodePhases)
                        phases_direct<T2, R>(x, terms + op.term_begin, z128 + 2 * (size_t)op.term_begin, fold_val,
                                             op.term_count, lthr, base);
                    else
                        phases_zeta<T2, R>(x, terms + op.term_begin, z128 + 2 * (size_t)op.term_begin, fold_val,
                                           op.term_count, lthr, base);
                    continue;
                }
                const bool any = op.any != 0;
                const uint32_t creg = op.c_reg, cthr = op.c_thr;
                // forms 0 and 1 have no control outside the tile; form 2
                // selects its results per thread (all-set or any-set
                // controls on thread-held and outer bits)
                const bool g_all = (base & op.c_glb) == op.c_glb, g_any = (base & op.c_glb) != 0;
                const bool t_all = (lthr & cthr) == cthr && g_all, t_any = (lthr & cthr) != 0u;
                auto ok = [&](int j) {
                    if (any) return g_any || t_any || ((uint32_t)j & creg) != 0u;
                    return t_all && ((uint32_t)j & creg) == creg;
                };
                R m[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if constexpr (sizeof(R) == 4)
                        m[i] = op.mf[i];
                    else
                        m[i] = m128[8 * (size_t)(ob + o) + i];
                }
                switch (code) {
#define AQS_PAIRS(C, SL, OP, FORM) \
    case C: pairs<SL, OP, FORM>(x, m, creg, ok); break;
#define AQS_PAIRS4(C, OP, FORM) \
    AQS_PAIRS(C, 0, OP, FORM) AQS_PAIRS(C + 1, 1, OP, FORM) AQS_PAIRS(C + 2, 2, OP, FORM) AQS_PAIRS(C + 3, 3, OP, FORM)
                    AQS_PAIRS4(0, 0, 0)
                    AQS_PAIRS4(4, 0, 1)
                    AQS_PAIRS4(8, 0, 2)
                    AQS_PAIRS4(12, 1, 0)
                    AQS_PAIRS4(16, 1, 1)
                    AQS_PAIRS4(20, 1, 2)
                    AQS_PAIRS4(kTileCodeMatR, 2, 0)
                    AQS_PAIRS4(kTileCodeMatR + 4, 2, 1)
                    AQS_PAIRS4(kTileCodeMatR + 8, 2, 2)
                    AQS_PAIRS4(kTileCodeMatP, 3, 0)
#undef AQS_PAIRS4
#undef AQS_PAIRS
#define AQS_MATPS(M)                                                   \
    case kTileCodeMatPSet + M:                                         \
        matp_set<M, T2, R>(x, ops, m128, ob + o, ok);                  \
        o += __builtin_popcount(M) - 1;                                \
        break;
                    AQS_MATPS(3) AQS_MATPS(5) AQS_MATPS(6) AQS_MATPS(7) AQS_MATPS(9) AQS_MATPS(10)
                    AQS_MATPS(11) AQS_MATPS(12) AQS_MATPS(13) AQS_MATPS(14) AQS_MATPS(15)
#undef AQS_MATPS
                    default: break;
                }
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) lds[addr(j)] = x[j];
            __syncthreads();
        }
        store_tile<T2, M, T>(sv, base, hi_tab, lds, tid);
        __syncthreads();
    }
}

template <typename T2, int W>
void launch(aqs_state s, const TileArgs& a, const TileDevProgram& p, int per_cu) {
    constexpr int M = W + kTileRegBits;
    constexpr size_t lds_fixed = (size_t)(1 << (M - 5)) * sizeof(uint64_t) + lds_len(1u << M) * sizeof(T2);
    const size_t lds = lds_fixed + (size_t)a.nfold * sizeof(T2);
    constexpr size_t lds_max = lds_fixed + (size_t)kTileMaxFolds * sizeof(T2);
    if (lds_max > 65536) {
        // once per kernel instantiation: raising the limit is not free per launch
        static std::once_flag once;
        static hipError_t rc = hipSuccess;
        std::call_once(once, [&] {
            rc = hipFuncSetAttribute((const void*)k_tile<T2, W>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds_max);
        });
        AQS_HIP(rc);
    }
    const uint64_t blocks = std::min<uint64_t>(a.tiles, (uint64_t)device_cu_count() * per_cu);
    hipLaunchKernelGGL((k_tile<T2, W>), dim3((unsigned)blocks), dim3(1 << W), lds, s->stream,
                       reinterpret_cast<T2*>(s->amps), reinterpret_cast<const TileSegDev*>(p.segs),
                       reinterpret_cast<const TileOpDev*>(p.ops), reinterpret_cast<const TileTermDev*>(p.terms),
                       reinterpret_cast<const double*>(p.m128), reinterpret_cast<const double*>(p.z128),
                       reinterpret_cast<const TileFoldDev*>(p.folds),
                       reinterpret_cast<const TileFoldTermDev*>(p.fold_terms),
                       reinterpret_cast<const double*>(p.fold_z128), a);
    AQS_HIP(hipGetLastError());
}

}  // namespace

namespace aqs_hip {

void apply_tile_pass(aqs_state s, const TilePass& pass, const TileDevProgram& prog) {
    const int m = pass.m;
    if (m < 10 || m > kTileMaxLocal || (int)pass.local.size() != m || m > s->n)
        fail(AQS_ERR_INVALID_ARGUMENT, "tile pass does not fit this state");
    for (int p = 0; p < m; ++p)
        if (pass.local[p] < 0 || pass.local[p] >= s->n || (p < 5 && pass.local[p] != p) ||
            (p > 0 && pass.local[p] <= pass.local[p - 1]))
            fail(AQS_ERR_INVALID_ARGUMENT, "tile pass local bits malformed");
    TileArgs a{};
    a.tiles = s->N >> m;
    a.nseg = (int)pass.segs.size();
    a.nfold = (int)pass.folds.size();
    if (a.nfold > kTileMaxFolds) fail(AQS_ERR_INVALID_ARGUMENT, "tile pass has too many folded terms");
    for (int p = 0; p < m; ++p) a.local[p] = pass.local[p];
    ProfScope ps(s, kSlotTile);
    const bool c64 = s->precision == AQS_C64;
    switch (m) {
        case 10:
            c64 ? launch<float2, 6>(s, a, prog, 16)
                : launch<double2, 6>(s, a, prog, 8);
            break;
        case 11:
            c64 ? launch<float2, 7>(s, a, prog, 8)
                : launch<double2, 7>(s, a, prog, 4);
            break;
        case 12:
            c64 ? launch<float2, 8>(s, a, prog, 4)
                : launch<double2, 8>(s, a, prog, 2);
            break;
        default:
            c64 ? launch<float2, 9>(s, a, prog, 2)
                : launch<double2, 9>(s, a, prog, 1);
            break;
    }
}

}  // namespace aqs_hip
