/*
 * aqs_oracle.c -- FP64 CPU restatement of the afQuantumSim state-vector path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg.  Nothing in the product links it.
 *
 * Each gate follows the operator the reference builds in src/quantum.cpp:
 * sparse gates as "row i of the CSR matrix" (out[i] = val(i) * in[col(i)]),
 * dense gates as the Kronecker product I(2^q) (x) U (x) I(2^(n-q-1)), and
 * circuit gates as the embedding of a 2^k x 2^k block at gen_index positions.
 * The arithmetic is FP64 and exact in the constants: H uses 1/sqrt(2) where
 * the reference stores the float 0.70710678118f, and the trigonometry of the
 * float angles is evaluated in double.
 */
#include "aqs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_SQRT1_2 0.70710678118654752440

typedef struct {
    double re, im;
} cplx;

static inline cplx c_make(double re, double im) {
    cplx c = {re, im};
    return c;
}
static inline cplx c_mul(cplx a, cplx b) {
    return c_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx c_add(cplx a, cplx b) { return c_make(a.re + b.re, a.im + b.im); }
static inline cplx c_ld(const double* v, uint64_t i) { return c_make(v[2 * i], v[2 * i + 1]); }
static inline void c_st(double* v, uint64_t i, cplx c) {
    v[2 * i] = c.re;
    v[2 * i + 1] = c.im;
}

/* Bit position of reference qubit q (qubit 0 = MSB), include/quantum.h:102-106. */
static inline uint64_t qmask(int n, int q) { return 1ull << (uint64_t)(n - 1 - q); }

static int g_threads = 1;

/* ---------------------------------------------------------------------------
 * Gate matrices (row-major 2x2, interleaved)
 * ------------------------------------------------------------------------- */

/* H: src/quantum.cpp:40-49 (h_mat, transposed and symmetric) with 1/sqrt(2) exact. */
/* RotX: src/quantum.cpp:683-688  [[c, -is], [-is, c]], c = cos(a/2), s = sin(a/2). */
/* RotY: src/quantum.cpp:730-735  column-major {c,-s,s,c} then .T() -> [[c, -s], [s, c]]. */
/* RotZ: src/quantum.cpp:777-782  diag(c - is, c + is). */
/* Phase: src/quantum.cpp:861-866 diag(1, cos a + i sin a). */
/* U3 (no reference counterpart): Phase(phi) * RotY(theta) * Phase(lambda). */
int orc_matrix_1q(const orc_gate* g, double m[8]) {
    double c, s;
    memset(m, 0, 8 * sizeof(double));
    switch (g->kind) {
        case ORC_X:
        case ORC_CX:
            m[2] = 1.0; m[4] = 1.0;
            return 1;
        case ORC_Y:
        case ORC_CY:
            m[3] = -1.0; m[5] = 1.0;
            return 1;
        case ORC_Z:
        case ORC_CZ:
            m[0] = 1.0; m[6] = -1.0;
            return 1;
        case ORC_H:
        case ORC_CH:
            m[0] = ORC_SQRT1_2; m[2] = ORC_SQRT1_2;
            m[4] = ORC_SQRT1_2; m[6] = -ORC_SQRT1_2;
            return 1;
        case ORC_PHASE:
        case ORC_CPHASE:
            m[0] = 1.0;
            m[6] = cos((double)g->angle);
            m[7] = sin((double)g->angle);
            return 1;
        case ORC_ROTX:
        case ORC_CROTX:
            c = cos((double)g->angle / 2.0); s = sin((double)g->angle / 2.0);
            m[0] = c; m[3] = -s; m[5] = -s; m[6] = c;
            return 1;
        case ORC_ROTY:
        case ORC_CROTY:
            c = cos((double)g->angle / 2.0); s = sin((double)g->angle / 2.0);
            m[0] = c; m[2] = -s; m[4] = s; m[6] = c;
            return 1;
        case ORC_ROTZ:
        case ORC_CROTZ:
            c = cos((double)g->angle / 2.0); s = sin((double)g->angle / 2.0);
            m[0] = c; m[1] = -s; m[6] = c; m[7] = s;
            return 1;
        case ORC_U3: {
            double th = (double)g->angle, ph = (double)g->phi, la = (double)g->lam;
            c = cos(th / 2.0); s = sin(th / 2.0);
            m[0] = c;
            m[2] = -cos(la) * s; m[3] = -sin(la) * s;
            m[4] = cos(ph) * s;  m[5] = sin(ph) * s;
            m[6] = cos(ph + la) * c; m[7] = sin(ph + la) * c;
            return 1;
        }
        default:
            return 0;
    }
}

/* ---------------------------------------------------------------------------
 * Row formulation of the sparse (one non-zero per row) gates.
 * Returns the column index of row i and writes the value.
 * ------------------------------------------------------------------------- */
static uint64_t swap_bits(uint64_t i, uint64_t pa, uint64_t pb) {
    /* src/quantum.cpp:929-937: temp = ((i >> pA) ^ (i >> pB)) & 1; i ^ (temp<<pA | temp<<pB) */
    uint64_t t = ((i >> pa) ^ (i >> pb)) & 1ull;
    return i ^ ((t << pa) | (t << pb));
}

static uint64_t sparse_row(int n, const orc_gate* g, uint64_t i, cplx* val) {
    *val = c_make(1.0, 0.0);
    switch (g->kind) {
        case ORC_X: { /* src/quantum.cpp:540-562: col = i ^ mask */
            return i ^ qmask(n, g->q0);
        }
        case ORC_Y: { /* src/quantum.cpp:580-609: col = i ^ mask; val = -i (bit 0), +i (bit 1) */
            uint64_t m = qmask(n, g->q0);
            *val = (i & m) ? c_make(0.0, 1.0) : c_make(0.0, -1.0);
            return i ^ m;
        }
        case ORC_Z: { /* src/quantum.cpp:627-653: val = 1 (bit 0), -1 (bit 1) */
            if (i & qmask(n, g->q0)) *val = c_make(-1.0, 0.0);
            return i;
        }
        case ORC_PHASE: { /* src/quantum.cpp:849-877: val = e^{ia} where bit set */
            if (i & qmask(n, g->q0))
                *val = c_make(cos((double)g->angle), sin((double)g->angle));
            return i;
        }
        case ORC_SWAP: { /* src/quantum.cpp:912-952 */
            return swap_bits(i, (uint64_t)(n - 1 - g->q0), (uint64_t)(n - 1 - g->q1));
        }
        case ORC_CX: { /* src/quantum.cpp:979-1014: flip target where control set */
            uint64_t cm = qmask(n, g->q0), tm = qmask(n, g->q1);
            return (i & cm) ? (i ^ tm) : i;
        }
        case ORC_CY: { /* src/quantum.cpp:1041-1088 */
            uint64_t cm = qmask(n, g->q0), tm = qmask(n, g->q1), mask = cm | tm;
            uint64_t op = i & mask;
            if (op == mask) *val = c_make(0.0, 1.0);
            else if (op == cm) *val = c_make(0.0, -1.0);
            return (i & cm) ? (i ^ tm) : i;
        }
        case ORC_CZ: { /* src/quantum.cpp:1115-1152: -1 where both bits set */
            uint64_t mask = qmask(n, g->q0) | qmask(n, g->q1);
            if ((i & mask) == mask) *val = c_make(-1.0, 0.0);
            return i;
        }
        case ORC_CPHASE: { /* src/quantum.cpp:1179-1218: e^{ia} where both bits set */
            uint64_t mask = qmask(n, g->q0) | qmask(n, g->q1);
            if ((i & mask) == mask)
                *val = c_make(cos((double)g->angle), sin((double)g->angle));
            return i;
        }
        case ORC_CSWAP: { /* src/quantum.cpp:1265-1313: swap A,B bits where control set */
            uint64_t cm = qmask(n, g->q0);
            if (!(i & cm)) return i;
            return swap_bits(i, (uint64_t)(n - 1 - g->q1), (uint64_t)(n - 1 - g->q2));
        }
        case ORC_CCX: { /* src/quantum.cpp:1542-1580: flip target where both controls set
                           (A == B is accepted and acts as one control, :1524-1540) */
            uint64_t cm = qmask(n, g->q0) | qmask(n, g->q1), tm = qmask(n, g->q2);
            return ((i & cm) == cm) ? (i ^ tm) : i;
        }
        case ORC_OR: { /* src/quantum.cpp:1614-1653: flip target where either control set */
            uint64_t cm = qmask(n, g->q0) | qmask(n, g->q1), tm = qmask(n, g->q2);
            return (i & cm) ? (i ^ tm) : i;
        }
        default:
            return i;
    }
}

static int is_sparse(int kind) {
    switch (kind) {
        case ORC_X: case ORC_Y: case ORC_Z: case ORC_PHASE: case ORC_SWAP:
        case ORC_CX: case ORC_CY: case ORC_CZ: case ORC_CPHASE: case ORC_CSWAP:
        case ORC_CCX: case ORC_OR:
            return 1;
        default:
            return 0;
    }
}

/* Dense single-target kinds (with the control qubit, if any). */
static int dense_1q(const orc_gate* g, int* target, int* control) {
    switch (g->kind) {
        case ORC_H: case ORC_ROTX: case ORC_ROTY: case ORC_ROTZ: case ORC_U3:
            *target = g->q0; *control = -1; return 1;
        /* CH: src/quantum.cpp:1342-1400 (two non-zeros per row: controlled H).
         * CRotX/Y/Z: src/quantum.cpp:1424-1510, ControlGate(Rot*::gate(a), c, t). */
        case ORC_CH: case ORC_CROTX: case ORC_CROTY: case ORC_CROTZ:
            *target = g->q1; *control = g->q0; return 1;
        default:
            return 0;
    }
}

/* One output amplitude of a gate, out[i], read from in[]. */
static cplx gate_row(int n, const orc_gate* g, const double* in, uint64_t i, const double m[8]) {
    int t, c;
    if (g->kind == ORC_BARRIER) return c_ld(in, i);
    if (is_sparse(g->kind)) {
        cplx v;
        uint64_t col = sparse_row(n, g, i, &v);
        return c_mul(v, c_ld(in, col));
    }
    if (dense_1q(g, &t, &c)) {
        if (c >= 0 && !(i & qmask(n, c))) return c_ld(in, i);
        uint64_t tm = qmask(n, t);
        int r = (i & tm) ? 1 : 0;
        cplx a0 = c_ld(in, i & ~tm), a1 = c_ld(in, i | tm);
        cplx u0 = c_make(m[4 * r], m[4 * r + 1]), u1 = c_make(m[4 * r + 2], m[4 * r + 3]);
        return c_add(c_mul(u0, a0), c_mul(u1, a1));
    }
    if (g->kind == ORC_CIRCUIT || g->kind == ORC_CONTROL_CIRCUIT) {
        /* Gate (src/quantum.cpp:1752-1809) / ControlGate (:1888-1950): the
         * sub-unitary acts on qubits [begin, begin+k), sub-qubit 0 = qubit begin. */
        int k = g->k;
        int begin = (g->kind == ORC_CIRCUIT) ? g->q0 : g->q1;
        if (g->kind == ORC_CONTROL_CIRCUIT && !(i & qmask(n, g->q0))) return c_ld(in, i);
        int pos = n - begin - k;
        uint64_t field = ((1ull << k) - 1ull) << pos;
        uint64_t r = (i & field) >> pos;
        uint64_t base = i & ~field;
        uint64_t dim = 1ull << k;
        cplx acc = c_make(0.0, 0.0);
        for (uint64_t col = 0; col < dim; ++col) {
            cplx u = c_make(g->u[2 * (r * dim + col)], g->u[2 * (r * dim + col) + 1]);
            acc = c_add(acc, c_mul(u, c_ld(in, base | (col << pos))));
        }
        return acc;
    }
    return c_ld(in, i);
}

static void apply_oop(const double* in, double* out, int n, const orc_gate* g, int threads) {
    double m[8];
    orc_matrix_1q(g, m);
    const int64_t N = (int64_t)1 << n;
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (int64_t i = 0; i < N; ++i) c_st(out, (uint64_t)i, gate_row(n, g, in, (uint64_t)i, m));
}

void orc_apply(double* sv, int n, const orc_gate* g) {
    size_t bytes = sizeof(double) * 2 * ((size_t)1 << n);
    double* tmp = (double*)malloc(bytes);
    if (!tmp) return;
    memcpy(tmp, sv, bytes);
    apply_oop(tmp, sv, n, g, 1);
    free(tmp);
}

static void run_impl(double* sv, int n, const orc_gate* gates, int count, int threads) {
    size_t bytes = sizeof(double) * 2 * ((size_t)1 << n);
    double* tmp = (double*)malloc(bytes);
    if (!tmp) return;
    double* a = sv;
    double* b = tmp;
    for (int i = 0; i < count; ++i) {
        if (gates[i].kind == ORC_BARRIER) continue;
        apply_oop(a, b, n, &gates[i], threads);
        double* t = a;
        a = b;
        b = t;
    }
    if (a != sv) memcpy(sv, a, bytes);
    free(tmp);
}

void orc_run(double* sv, int n, const orc_gate* gates, int count) {
    run_impl(sv, n, gates, count, g_threads);
}

void orc_run_threads(double* sv, int n, const orc_gate* gates, int count, int threads) {
    run_impl(sv, n, gates, count, threads < 1 ? 1 : threads);
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------------------
 * Reference method: explicit 2^n x 2^n operators.
 * ------------------------------------------------------------------------- */

/* kron(A (ra x ca), B (rb x cb)) -> out (ra*rb x ca*cb), row-major complex,
 * the tensor_product of src/utils.cpp:169-195 (A is the more significant factor). */
static void kron(const double* A, int ra, int ca, const double* B, int rb, int cb, double* out) {
    int oc = ca * cb;
    for (int i = 0; i < ra; ++i)
        for (int j = 0; j < ca; ++j) {
            cplx a = c_ld(A, (uint64_t)i * ca + j);
            for (int k = 0; k < rb; ++k)
                for (int l = 0; l < cb; ++l) {
                    cplx b = c_ld(B, (uint64_t)k * cb + l);
                    c_st(out, (uint64_t)(i * rb + k) * oc + (j * cb + l), c_mul(a, b));
                }
        }
}

static double* identity(int dim) {
    double* I = (double*)calloc((size_t)dim * dim * 2, sizeof(double));
    for (int i = 0; i < dim; ++i) I[2 * ((size_t)i * dim + i)] = 1.0;
    return I;
}

void orc_dense_operator(int n, const orc_gate* g, double* op) {
    const int N = 1 << n;
    int t, c;
    memset(op, 0, sizeof(double) * 2 * (size_t)N * N);
    if (dense_1q(g, &t, &c) && c < 0) {
        /* I(2^t) (x) U (x) I(2^(n-t-1)), src/quantum.cpp:676-696 */
        double m[8];
        orc_matrix_1q(g, m);
        int L = 1 << t, R = 1 << (n - t - 1);
        double* IL = identity(L);
        double* IR = identity(R);
        double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)(2 * L) * (2 * L));
        kron(IL, L, L, m, 2, 2, tmp);
        kron(tmp, 2 * L, 2 * L, IR, R, R, op);
        free(IL);
        free(IR);
        free(tmp);
        return;
    }
    /* Everything else: column j of the operator is the gate applied to e_j. */
    double* e = (double*)calloc((size_t)N * 2, sizeof(double));
    double* col = (double*)malloc(sizeof(double) * 2 * (size_t)N);
    double m[8];
    orc_matrix_1q(g, m);
    for (int j = 0; j < N; ++j) {
        memset(e, 0, sizeof(double) * 2 * (size_t)N);
        e[2 * j] = 1.0;
        for (int i = 0; i < N; ++i) c_st(col, (uint64_t)i, gate_row(n, g, e, (uint64_t)i, m));
        for (int i = 0; i < N; ++i) {
            op[2 * ((size_t)i * N + j)] = col[2 * i];
            op[2 * ((size_t)i * N + j) + 1] = col[2 * i + 1];
        }
    }
    free(e);
    free(col);
}

void orc_gemv(int n, const double* op, const double* in, double* out) {
    const uint64_t N = 1ull << n;
    for (uint64_t i = 0; i < N; ++i) {
        cplx acc = c_make(0.0, 0.0);
        for (uint64_t j = 0; j < N; ++j) acc = c_add(acc, c_mul(c_ld(op, i * N + j), c_ld(in, j)));
        c_st(out, i, acc);
    }
}

/* ---------------------------------------------------------------------------
 * State preparation
 * ------------------------------------------------------------------------- */

/* |0...0> -> sv[0] = 1, |1...1> -> sv[N-1] = 1 (src/quantum.cpp:219-224). */
void orc_basis_state(double* sv, int n, uint64_t index) {
    memset(sv, 0, sizeof(double) * 2 * ((size_t)1 << n));
    sv[2 * index] = 1.0;
}

/* generate_statevector: kron(q0, q1, ..., q_{n-1}), src/quantum.cpp:261-275. */
void orc_product_state(double* sv, int n, const double* qamps) {
    const uint64_t N = 1ull << n;
    for (uint64_t i = 0; i < N; ++i) {
        cplx acc = c_make(1.0, 0.0);
        for (int q = 0; q < n; ++q) {
            int b = (i & qmask(n, q)) ? 1 : 0;
            acc = c_mul(acc, c_make(qamps[4 * q + 2 * b], qamps[4 * q + 2 * b + 1]));
        }
        c_st(sv, i, acc);
    }
}

static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* Uniform in (0, 1) from 53 random bits of a counter-based hash. */
static double unit_uniform(uint64_t seed, uint64_t ctr) {
    uint64_t h = splitmix64(seed ^ splitmix64(ctr));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* Box-Muller on the counter pair (2i, 2i+1). */
double orc_gaussian_component(uint64_t seed, uint64_t index, int imag_part) {
    double u1 = unit_uniform(seed, 2 * index);
    double u2 = unit_uniform(seed, 2 * index + 1);
    double r = sqrt(-2.0 * log(u1));
    double a = 6.283185307179586476925 * u2;
    return imag_part ? r * sin(a) : r * cos(a);
}

void orc_gaussian_state(double* sv, int n, uint64_t seed) {
    const uint64_t N = 1ull << n;
    for (uint64_t i = 0; i < N; ++i) {
        sv[2 * i] = orc_gaussian_component(seed, i, 0);
        sv[2 * i + 1] = orc_gaussian_component(seed, i, 1);
    }
    double nrm = sqrt(orc_norm2(sv, n));
    for (uint64_t i = 0; i < 2 * N; ++i) sv[i] /= nrm;
}

/* ---------------------------------------------------------------------------
 * Measurement
 * ------------------------------------------------------------------------- */

/* |a|^2 = real(s * conj(s)), src/quantum.cpp:349,404-414. */
void orc_probabilities(const double* sv, int n, double* p) {
    const uint64_t N = 1ull << n;
    for (uint64_t i = 0; i < N; ++i) p[i] = fma(sv[2 * i], sv[2 * i], sv[2 * i + 1] * sv[2 * i + 1]);
}

double orc_norm2(const double* sv, int n) {
    const uint64_t N = 1ull << n;
    long double acc = 0.0L;
    for (uint64_t i = 0; i < N; ++i) acc += (long double)sv[2 * i] * sv[2 * i] + (long double)sv[2 * i + 1] * sv[2 * i + 1];
    return (double)acc;
}

/* qubit_probability_true: sum of |a|^2 over indices with bit (n-1-q) set,
 * src/quantum.cpp:372-391. */
double orc_prob1(const double* sv, int n, int qubit) {
    const uint64_t N = 1ull << n;
    const uint64_t m = qmask(n, qubit);
    long double acc = 0.0L;
    for (uint64_t i = 0; i < N; ++i)
        if (i & m) acc += (long double)sv[2 * i] * sv[2 * i] + (long double)sv[2 * i + 1] * sv[2 * i + 1];
    return (double)acc;
}

/* <sum_i Z_i> = sum_q (P0(q) - P1(q)) = sum_q (norm2 - 2 P1(q)). */
double orc_expect_sum_z(const double* sv, int n) {
    const uint64_t N = 1ull << n;
    long double acc = 0.0L;
    for (uint64_t i = 0; i < N; ++i) {
        long double p = (long double)sv[2 * i] * sv[2 * i] + (long double)sv[2 * i + 1] * sv[2 * i + 1];
        int ones = __builtin_popcountll(i);
        acc += p * (long double)(n - 2 * ones);
    }
    return (double)acc;
}

#define ORC_FIX_SCALE 1152921504606846976.0 /* 2^60 */

static uint64_t quantise(double re, double im) {
    double p = fma(re, re, im * im);
    return (uint64_t)nearbyint(p * ORC_FIX_SCALE);
}

static void sample_fixed(const uint64_t* q, uint64_t N, const double* u, int shots, uint32_t* out) {
    /* inclusive cumulative sums, exact */
    uint64_t* cdf = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t acc = 0, last_nz = 0;
    for (uint64_t i = 0; i < N; ++i) {
        acc += q[i];
        cdf[i] = acc;
        if (q[i]) last_nz = i;
    }
    for (int s = 0; s < shots; ++s) {
        uint64_t thr = (uint64_t)floor(u[s] * ORC_FIX_SCALE);
        /* first i with cdf[i] > thr (binary search on the non-decreasing cdf) */
        uint64_t lo = 0, hi = N;
        while (lo < hi) {
            uint64_t mid = lo + (hi - lo) / 2;
            if (cdf[mid] > thr) hi = mid;
            else lo = mid + 1;
        }
        out[s] = (uint32_t)(lo < N ? lo : last_nz);
    }
    free(cdf);
}

void orc_sample_c64(const float* amps, int n, const double* u, int shots, uint32_t* out) {
    const uint64_t N = 1ull << n;
    uint64_t* q = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (uint64_t i = 0; i < N; ++i) q[i] = quantise((double)amps[2 * i], (double)amps[2 * i + 1]);
    sample_fixed(q, N, u, shots, out);
    free(q);
}

void orc_sample_c128(const double* sv, int n, const double* u, int shots, uint32_t* out) {
    const uint64_t N = 1ull << n;
    uint64_t* q = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (uint64_t i = 0; i < N; ++i) q[i] = quantise(sv[2 * i], sv[2 * i + 1]);
    sample_fixed(q, N, u, shots, out);
    free(q);
}

void orc_sample_fp64_cdf(const double* sv, int n, const double* u, int shots, uint32_t* out) {
    const uint64_t N = 1ull << n;
    double* cdf = (double*)malloc(sizeof(double) * N);
    double acc = 0.0;
    uint64_t last_nz = 0;
    for (uint64_t i = 0; i < N; ++i) {
        double p = fma(sv[2 * i], sv[2 * i], sv[2 * i + 1] * sv[2 * i + 1]);
        acc += p;
        cdf[i] = acc;
        if (p > 0.0) last_nz = i;
    }
    for (int s = 0; s < shots; ++s) {
        uint64_t i = 0;
        while (i < N && !(cdf[i] > u[s])) ++i;
        out[s] = (uint32_t)(i < N ? i : last_nz);
    }
    free(cdf);
}

/* ---------------------------------------------------------------------------
 * Timing restatement of the reference's method (CPU baseline only).
 *
 * simulate() on an uncompiled circuit (src/quantum.cpp:277-291) first
 * multiplies the state by the circuit's N x N operator, the identity while
 * nothing is compiled (:283), then applies every gate's operator.  Gates the
 * reference holds as dense N x N operators -- H, RotX, RotY, RotZ built as
 * I (x) U (x) I (src/utils.cpp:169-195, src/quantum.cpp:676-831), the
 * controlled rotations and Gate / ControlGate (src/quantum.cpp:1424-1510,
 * 1705-1960) -- cost a dense build plus an N x N product; the CSR gates (X, Y,
 * Z, Phase, Swap, CX, CY, CZ, CH, CPhase, CSwap, CCNot, Or) an O(N) row pass.
 * U3, which the reference lacks, is its decomposition Phase(phi) RotY(theta)
 * Phase(lambda).  FP64, rows split over OpenMP threads.
 * ------------------------------------------------------------------------- */

static int reference_dense_kind(int kind) {
    switch (kind) {
        case ORC_H: case ORC_ROTX: case ORC_ROTY: case ORC_ROTZ:
        case ORC_CROTX: case ORC_CROTY: case ORC_CROTZ:
        case ORC_CIRCUIT: case ORC_CONTROL_CIRCUIT:
            return 1;
        default:
            return 0;
    }
}

static void gemv_threads(int n, const double* op, double* sv, double* tmp, int threads) {
    const int64_t N = (int64_t)1 << n;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double re = 0.0, im = 0.0;
        const double* row = op + 2 * (size_t)i * (size_t)N;
        for (int64_t j = 0; j < N; ++j) {
            re += row[2 * j] * sv[2 * j] - row[2 * j + 1] * sv[2 * j + 1];
            im += row[2 * j] * sv[2 * j + 1] + row[2 * j + 1] * sv[2 * j];
        }
        tmp[2 * i] = re;
        tmp[2 * i + 1] = im;
    }
    memcpy(sv, tmp, sizeof(double) * 2 * (size_t)N);
}

static void reference_gate(double* sv, int n, const orc_gate* g, double* op, double* tmp, int threads) {
    if (reference_dense_kind(g->kind)) {
        orc_dense_operator(n, g, op);
        gemv_threads(n, op, sv, tmp, threads);
    } else {
        orc_apply(sv, n, g);
    }
}

int orc_refmethod_run(double* sv, int n, const orc_gate* gates, int count, int threads) {
    const size_t N = (size_t)1 << n;
    double* op = (double*)malloc(sizeof(double) * 2 * N * N);
    double* tmp = (double*)malloc(sizeof(double) * 2 * N);
    if (!op || !tmp) {
        free(op);
        free(tmp);
        return -1;
    }
    /* the uncompiled circuit's operator: the identity (src/quantum.cpp:159-186) */
    memset(op, 0, sizeof(double) * 2 * N * N);
    for (size_t i = 0; i < N; ++i) op[2 * (i * N + i)] = 1.0;
    gemv_threads(n, op, sv, tmp, threads);
    for (int k = 0; k < count; ++k) {
        const orc_gate* g = &gates[k];
        if (g->kind == ORC_U3) {
            orc_gate p = *g, r = *g;
            p.kind = ORC_PHASE;
            p.angle = g->lam;
            reference_gate(sv, n, &p, op, tmp, threads);
            r.kind = ORC_ROTY;
            r.angle = g->angle;
            reference_gate(sv, n, &r, op, tmp, threads);
            p.angle = g->phi;
            reference_gate(sv, n, &p, op, tmp, threads);
        } else {
            reference_gate(sv, n, g, op, tmp, threads);
        }
    }
    free(op);
    free(tmp);
    return 0;
}
