/*
 * cpu_sv.c -- plain complex64 CPU state-vector passes: the CPU baseline that
 * bench.py times beside the GPU (its cpu_baseline leg), built into the same
 * test-infrastructure library as the FP64 oracle.
 *
 * This is what a straightforward CPU simulator does, not the reference's
 * method (which builds a 2^n x 2^n operator per gate, src/quantum.cpp:676-696):
 *   - the state is 2^n interleaved FP32 pairs, updated in place, one pass per
 *     gate, no copy and no allocation per gate;
 *   - a one-qubit gate (or a controlled one) updates amplitude pairs with its
 *     FP32 2x2 matrix (the oracle's matrices, rounded once); diagonal gates
 *     touch only the amplitudes they change; permutations move amplitudes;
 *   - every pass splits its index range over the OpenMP threads (static
 *     schedule), so each thread streams its own contiguous share;
 *   - sampling forms the inclusive FP64 CDF once, then binary-searches per shot
 *     (the reference's rule: first index whose CDF exceeds u).
 * Gates: X Y Z H Phase RotX RotY RotZ U3 (one qubit), CX CY CZ CH CPhase
 * CRotX CRotY CRotZ (one control), Swap, CCX.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "aqs_oracle.h"

static inline uint64_t bitpos(int n, int q) { return (uint64_t)(n - 1 - q); }

/* i with a zero inserted at bit position p. */
static inline uint64_t insert0(uint64_t i, uint64_t p) {
    const uint64_t low = i & ((1ull << p) - 1);
    return ((i >> p) << (p + 1)) | low;
}

static int is_diagonal(int kind) {
    return kind == ORC_Z || kind == ORC_PHASE || kind == ORC_ROTZ || kind == ORC_CZ || kind == ORC_CPHASE ||
           kind == ORC_CROTZ;
}

static int is_controlled(int kind) {
    return kind == ORC_CX || kind == ORC_CY || kind == ORC_CZ || kind == ORC_CH || kind == ORC_CPHASE ||
           kind == ORC_CROTX || kind == ORC_CROTY || kind == ORC_CROTZ;
}

/* 2x2 on bit position t over the indices whose bit c is set (c < 0: all). */
static void pass_2x2(float* sv, int n, uint64_t t, int c, const float m[8], int threads) {
    const uint64_t half = 1ull << (n - 1);
    const uint64_t st = 1ull << t;
    const uint64_t cm = c >= 0 ? 1ull << (uint64_t)c : 0;
    const float a_r = m[0], a_i = m[1], b_r = m[2], b_i = m[3], c_r = m[4], c_i = m[5], d_r = m[6], d_i = m[7];
#pragma omp parallel for schedule(static) num_threads(threads)
    for (uint64_t k = 0; k < half; ++k) {
        const uint64_t i0 = insert0(k, t);
        if ((i0 & cm) != cm) continue;
        const uint64_t i1 = i0 | st;
        const float x_r = sv[2 * i0], x_i = sv[2 * i0 + 1], y_r = sv[2 * i1], y_i = sv[2 * i1 + 1];
        sv[2 * i0] = a_r * x_r - a_i * x_i + b_r * y_r - b_i * y_i;
        sv[2 * i0 + 1] = a_r * x_i + a_i * x_r + b_r * y_i + b_i * y_r;
        sv[2 * i1] = c_r * x_r - c_i * x_i + d_r * y_r - d_i * y_i;
        sv[2 * i1 + 1] = c_r * x_i + c_i * x_r + d_r * y_i + d_i * y_r;
    }
}

/* Multiply the amplitudes with bit t equal to `one` (and bit c set, c >= 0)
 * by z: only those amplitudes are read and written. */
static void pass_phase(float* sv, int n, uint64_t t, int one, int c, float z_r, float z_i, int threads) {
    if (c < 0) {
        const uint64_t half = 1ull << (n - 1);
#pragma omp parallel for schedule(static) num_threads(threads)
        for (uint64_t k = 0; k < half; ++k) {
            const uint64_t i = insert0(k, t) | (one ? 1ull << t : 0);
            const float x_r = sv[2 * i], x_i = sv[2 * i + 1];
            sv[2 * i] = z_r * x_r - z_i * x_i;
            sv[2 * i + 1] = z_r * x_i + z_i * x_r;
        }
        return;
    }
    const uint64_t lo = t < (uint64_t)c ? t : (uint64_t)c, hi = t < (uint64_t)c ? (uint64_t)c : t;
    const uint64_t quarter = 1ull << (n - 2);
    const uint64_t set = (1ull << (uint64_t)c) | (one ? 1ull << t : 0);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (uint64_t k = 0; k < quarter; ++k) {
        const uint64_t i = insert0(insert0(k, lo), hi) | set;
        const float x_r = sv[2 * i], x_i = sv[2 * i + 1];
        sv[2 * i] = z_r * x_r - z_i * x_i;
        sv[2 * i + 1] = z_r * x_i + z_i * x_r;
    }
}

/* Exchange amplitude pairs differing in bits a and b (a set, b clear <-> a
 * clear, b set), over indices with bit c set when c >= 0; c2 likewise. */
static void pass_swap(float* sv, int n, uint64_t a, uint64_t b, int c, int c2, int threads) {
    const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
    const uint64_t quarter = 1ull << (n - 2);
    const uint64_t cm = (c >= 0 ? 1ull << (uint64_t)c : 0) | (c2 >= 0 ? 1ull << (uint64_t)c2 : 0);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (uint64_t k = 0; k < quarter; ++k) {
        const uint64_t base = insert0(insert0(k, lo), hi);
        if ((base & cm) != cm) continue;
        const uint64_t i = base | (1ull << a), j = base | (1ull << b);
        const float r = sv[2 * i], m = sv[2 * i + 1];
        sv[2 * i] = sv[2 * j];
        sv[2 * i + 1] = sv[2 * j + 1];
        sv[2 * j] = r;
        sv[2 * j + 1] = m;
    }
}

/* Flip bit t over indices with every bit of mask cm set. */
static void pass_flip(float* sv, int n, uint64_t t, uint64_t cm, int threads) {
    const uint64_t half = 1ull << (n - 1);
    const uint64_t st = 1ull << t;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (uint64_t k = 0; k < half; ++k) {
        const uint64_t i0 = insert0(k, t);
        if ((i0 & cm) != cm) continue;
        const uint64_t i1 = i0 | st;
        const float r = sv[2 * i0], m = sv[2 * i0 + 1];
        sv[2 * i0] = sv[2 * i1];
        sv[2 * i0 + 1] = sv[2 * i1 + 1];
        sv[2 * i1] = r;
        sv[2 * i1 + 1] = m;
    }
}

int csv_apply(float* sv, int n, const orc_gate* g, int threads) {
    const int kind = g->kind;
    if (kind == ORC_BARRIER) return 1;
    if (kind == ORC_SWAP) {
        pass_swap(sv, n, bitpos(n, g->q0), bitpos(n, g->q1), -1, -1, threads);
        return 1;
    }
    if (kind == ORC_CSWAP) {
        pass_swap(sv, n, bitpos(n, g->q1), bitpos(n, g->q2), (int)bitpos(n, g->q0), -1, threads);
        return 1;
    }
    if (kind == ORC_CCX) {
        pass_flip(sv, n, bitpos(n, g->q2), (1ull << bitpos(n, g->q0)) | (1ull << bitpos(n, g->q1)), threads);
        return 1;
    }
    const int ctrl = is_controlled(kind) ? (int)bitpos(n, g->q0) : -1;
    const uint64_t t = bitpos(n, is_controlled(kind) ? g->q1 : g->q0);
    if (kind == ORC_X || kind == ORC_CX) {
        pass_flip(sv, n, t, ctrl >= 0 ? 1ull << (uint64_t)ctrl : 0, threads);
        return 1;
    }
    double md[8];
    if (!orc_matrix_1q(g, md)) return 0;
    float m[8];
    for (int k = 0; k < 8; ++k) m[k] = (float)md[k];
    if (is_diagonal(kind)) {
        if (m[0] != 1.0f || m[1] != 0.0f) pass_phase(sv, n, t, 0, ctrl, m[0], m[1], threads);
        pass_phase(sv, n, t, 1, ctrl, m[6], m[7], threads);
        return 1;
    }
    pass_2x2(sv, n, t, ctrl, m, threads);
    return 1;
}

int csv_run(float* sv, int n, const orc_gate* gates, int count, int threads) {
    for (int i = 0; i < count; ++i)
        if (!csv_apply(sv, n, &gates[i], threads)) return i;
    return count;
}

/* Inclusive FP64 CDF in `cdf` (caller's 2^n doubles), then per shot the first
 * index whose CDF exceeds u (the last index when none does). */
void csv_sample(const float* sv, int n, double* cdf, const double* u, int shots, uint32_t* out, int threads) {
    const uint64_t N = 1ull << n;
    int nt = threads > 0 ? threads : 1;
    /* per-thread partial sums, then each thread scans its share with its offset */
    double* part = (double*)calloc((size_t)nt + 1, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int t = 0, T = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        T = omp_get_num_threads();
#endif
        const uint64_t lo = N * (uint64_t)t / (uint64_t)T, hi = N * (uint64_t)(t + 1) / (uint64_t)T;
        double acc = 0.0;
        for (uint64_t i = lo; i < hi; ++i) {
            const double re = sv[2 * i], im = sv[2 * i + 1];
            acc += re * re + im * im;
            cdf[i] = acc;
        }
        part[t + 1] = acc;
#pragma omp barrier
#pragma omp single
        for (int k = 1; k <= T; ++k) part[k] += part[k - 1];
        const double off = part[t];
        for (uint64_t i = lo; i < hi; ++i) cdf[i] += off;
    }
    free(part);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int s = 0; s < shots; ++s) {
        uint64_t a = 0, b = N;
        while (a < b) {
            const uint64_t mid = a + (b - a) / 2;
            if (cdf[mid] > u[s]) b = mid;
            else a = mid + 1;
        }
        out[s] = (uint32_t)(a < N ? a : N - 1);
    }
}
