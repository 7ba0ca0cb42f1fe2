/*
 * CPU oracle for the FDTD hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a restatement of the reference CPU backend's algorithm
 * (src/DeepGPR/lib/deepgpr_cpu.c) written as a checker for the HIP
 * implementation. It is never linked into the product: only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * It is compiled twice:
 *   -DORACLE_REAL=float  -> oracle_f32_*   (op-for-op float arithmetic; must be
 *                                            bitwise equal to the reference CPU)
 *   -DORACLE_REAL=double -> oracle_f64_*   (the FP64 error oracle)
 *
 * The ABI of oracle_<p>_forward / oracle_<p>_backward follows the reference
 * ABI-6 prototypes (src/DeepGPR/lib/deepgpr.h:89-152) with ORACLE_REAL in place
 * of float and one extra trailing argument, the FDTD order, instead of the
 * process-global set_fdtd_order (deepgpr_cpu.c:276-279).
 *
 * Pinning (what checks each part): every function is compared bitwise with
 * the reference compiled from source (oracle/Makefile, target "ref"), against
 * the golden vectors in tests/golden/, and through the adjoint dot-product,
 * Taylor and closed-form tests in tests/test_oracle*.py. The INT8 history
 * codec has no CPU counterpart in the reference (it is CUDA-only,
 * src/DeepGPR/lib/deepgpr.cu:2885-3030); it is pinned by the reference's
 * decoder known-answer test (tests/test_numerics.py:55-70), by the layout
 * formula (src/DeepGPR/solver/storage.py:377-414) and by round-trip bounds.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORACLE_REAL
#define ORACLE_REAL float
#endif
typedef ORACLE_REAL real;

#define ORACLE_CAT2(a, b) a##b
#define ORACLE_CAT(a, b) ORACLE_CAT2(a, b)
#ifndef ORACLE_PREFIX
#define ORACLE_PREFIX oracle_f32_
#endif
#define API(name) ORACLE_CAT(ORACLE_PREFIX, name)
#define EXPORT __attribute__((visibility("default")))

/* CODATA 2018 constants (deepgpr.h:21-24). The float build uses the float
 * literals directly so no double rounding can differ from the reference. */
#ifdef ORACLE_DOUBLE
#define EPS0 8.8541878128e-12
#define MU0 1.25663706212e-06
#else
#define EPS0 8.8541878128e-12f
#define MU0 1.25663706212e-06f
#endif
#define PEC_SIGMA ((real)100.0)

/* storage_type bits (src/DeepGPR/config/constants.py:125-200). */
#define KIND_MASK 0x0f
#define KIND_F32 0
#define KIND_F16 1
#define KIND_BF16 2
#define KIND_INT8 3
#define FLAG_RHS_FROM_E (1 << 4)
#define FLAG_PHYSICAL (1 << 5)
#define FLAG_NO_HISTORY (1 << 26)

static __thread char g_error[512];

static void set_error(const char* msg) { snprintf(g_error, sizeof(g_error), "%s", msg); }

EXPORT const char* API(last_error)(void) { return g_error; }

/* ------------------------------------------------------------------------ */
/* Half / bfloat16 / int8 codecs                                            */
/* ------------------------------------------------------------------------ */

/* IEEE binary32 -> binary16, round to nearest even; NaN -> quiet NaN with the
 * input sign (the convention of deepgpr_cpu.c:100-141). */
static uint16_t f32_to_f16(float value)
{
    uint32_t u;
    memcpy(&u, &value, 4);
    uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t mag = u & 0x7fffffffu;
    if (mag > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);
    if (mag >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 -> inf */
    int exp32 = (int)(mag >> 23);
    uint32_t frac = mag & 0x7fffffu;
    int e16 = exp32 - 112; /* biased binary16 exponent */
    uint32_t keep, drop, half;
    int shift;
    if (e16 <= 0) {
        /* binary16 subnormal (or zero): value / 2^-24 rounded to an integer. */
        if (e16 < -10) return (uint16_t)sign;
        frac |= 0x800000u;
        shift = 14 - e16;
    } else {
        shift = 13;
    }
    keep = frac >> shift;
    drop = frac & ((1u << shift) - 1u);
    half = 1u << (shift - 1);
    if (drop > half || (drop == half && (keep & 1u))) keep += 1u;
    if (e16 <= 0) return (uint16_t)(sign | keep); /* carry into 0x400 is the smallest normal */
    /* keep holds 10 mantissa bits; a carry to 0x400 bumps the exponent. */
    return (uint16_t)(sign + (((uint32_t)e16 << 10) + keep));
}

static float f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t m = (uint32_t)h & 0x3ffu;
    uint32_t bits;
    if (e == 0u) {
        if (m == 0u) {
            bits = sign;
        } else {
            /* subnormal: m * 2^-24, normalised */
            int shift = 0;
            while ((m & 0x400u) == 0u) {
                m <<= 1;
                ++shift;
            }
            m &= 0x3ffu;
            bits = sign | ((uint32_t)(113 - shift) << 23) | (m << 13);
        }
    } else if (e == 0x1fu) {
        bits = sign | 0x7f800000u | (m << 13);
    } else {
        bits = sign | ((e + 112u) << 23) | (m << 13);
    }
    float out;
    memcpy(&out, &bits, 4);
    return out;
}

static uint16_t f32_to_bf16(float value)
{
    uint32_t u;
    memcpy(&u, &value, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    uint32_t lsb = (u >> 16) & 1u;
    return (uint16_t)((u + 0x7fffu + lsb) >> 16);
}

static float bf16_to_f32(uint16_t b)
{
    uint32_t u = (uint32_t)b << 16;
    float out;
    memcpy(&out, &u, 4);
    return out;
}

EXPORT void API(codec_roundtrip)(const float* input, void* encoded, float* decoded,
                                 long long count, int kind)
{
    for (long long n = 0; n < count; ++n) {
        if (kind == KIND_F16) {
            uint16_t h = f32_to_f16(input[n]);
            ((uint16_t*)encoded)[n] = h;
            decoded[n] = f16_to_f32(h);
        } else if (kind == KIND_BF16) {
            uint16_t h = f32_to_bf16(input[n]);
            ((uint16_t*)encoded)[n] = h;
            decoded[n] = bf16_to_f32(h);
        } else {
            ((float*)encoded)[n] = input[n];
            decoded[n] = input[n];
        }
    }
}

static size_t kind_bytes(int kind)
{
    if (kind == KIND_F32) return sizeof(real);
    if (kind == KIND_INT8) return 1;
    return 2;
}

static void store_value(void* base, long long n, real v, int kind)
{
    if (kind == KIND_F16) ((uint16_t*)base)[n] = f32_to_f16((float)v);
    else if (kind == KIND_BF16) ((uint16_t*)base)[n] = f32_to_bf16((float)v);
    else ((real*)base)[n] = v;
}

static real load_value(const void* base, long long n, int kind)
{
    if (kind == KIND_F16) return (real)f16_to_f32(((const uint16_t*)base)[n]);
    if (kind == KIND_BF16) return (real)bf16_to_f32(((const uint16_t*)base)[n]);
    return ((const real*)base)[n];
}

/* ------------------------------------------------------------------------ */
/* Grid description                                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    int N[3];             /* NX, NY, NZ (field sizes incl. Yee halo) */
    long long st[3];      /* strides of x, y, z */
    long long cells;      /* NX*NY*NZ */
    int nshot;
    int order;
    int pml[6];
    real d[3];            /* dx, dy, dz */
    const real* coef[2][6];   /* [0=E,1=H][face] -> (4, p) RA, RB, RE, RF */
    real* psi[2][6][2];       /* [E/H][face][phi1/phi2] */
} Grid;

/* Target component of phi1/phi2 on a face of the given axis, and the
 * differentiated source component (src/DeepGPR/lib/deepgpr_cpu.c:722-1100).
 * The sign is the sign of that term in the curl. */
static const int E_TARGET[3][2] = {{1, 2}, {0, 2}, {0, 1}};
static const int E_SOURCE[3][2] = {{2, 1}, {2, 0}, {1, 0}};
static const int E_SIGN[3][2] = {{-1, 1}, {1, -1}, {-1, 1}};
static const int H_TARGET[3][2] = {{1, 2}, {0, 2}, {0, 1}};
static const int H_SOURCE[3][2] = {{2, 1}, {2, 0}, {1, 0}};
static const int H_SIGN[3][2] = {{1, -1}, {-1, 1}, {1, -1}};

/* Shape of one phi array: the face axis holds the thickness (+1 for E); for
 * E the target axis holds N-1, for H the source axis holds N-1
 * (src/DeepGPR/solver/pml.py:310-348). */
static void psi_dims(const Grid* g, int field, int face, int which, int dims[3])
{
    int axis = face / 2, p = g->pml[face];
    for (int a = 0; a < 3; ++a) dims[a] = g->N[a];
    if (field == 0) {
        dims[E_TARGET[axis][which]] -= 1;
        dims[axis] = p + 1;
    } else {
        dims[H_SOURCE[axis][which]] -= 1;
        dims[axis] = p;
    }
}

/* Index of the cell along the face axis inside the face profile, or -1 when
 * the cell is outside the face region. */
static int face_q(const Grid* g, int field, int face, int c)
{
    int axis = face / 2, p = g->pml[face], n = g->N[axis];
    if (p <= 0) return -1;
    if (face % 2 == 0) {
        if (field == 0) return (c >= 1 && c <= p) ? p - c : -1;
        return (c >= 0 && c <= p - 1) ? p - 1 - c : -1;
    }
    if (c >= n - 1 - p && c <= n - 2) {
        if (field == 0 && c <= 0) return -1;
        return c - (n - 1 - p);
    }
    return -1;
}

static long long psi_index(const int dims[3], int s, const int c[3], int axis, int q)
{
    int cc[3] = {c[0], c[1], c[2]};
    cc[axis] = q;
    return (((long long)s * dims[0] + cc[0]) * dims[1] + cc[1]) * dims[2] + cc[2];
}

/* ------------------------------------------------------------------------ */
/* Staggered stencils (deepgpr_cpu.c:287-511)                                */
/* ------------------------------------------------------------------------ */

static int radius_of(int order) { return order >= 8 ? 4 : (order >= 4 ? 2 : 1); }

static real stencil_c(int radius, int r)
{
    switch (radius) {
    case 4:
        return r == 1 ? (real)1225 / (real)1024 : r == 2 ? (real)-245 / (real)3072
             : r == 3 ? (real)49 / (real)5120 : (real)-5 / (real)7168;
    case 3:
        return r == 1 ? (real)75 / (real)64 : r == 2 ? (real)-25 / (real)384 : (real)3 / (real)640;
    case 2:
        return r == 1 ? (real)9 / (real)8 : (real)-1 / (real)24;
    default:
        return (real)1;
    }
}

/* Largest radius <= want whose backward stencil at c fits in [0, n). */
static int fit_bwd(long long c, long long n, int want)
{
    int r = want;
    while (r > 1 && (c - r < 0 || c + r - 1 >= n)) --r;
    return r;
}

static int fit_fwd(long long c, long long n, int want)
{
    int r = want;
    while (r > 1 && (c - r + 1 < 0 || c + r >= n)) --r;
    return r;
}

/* sum_r c_r (f[c + r - 1] - f[c - r]) */
static real dbwd(const real* f, long long id, long long st, long long c, long long n, int order)
{
    if (order <= 2) return f[id] - f[id - st];
    int R = fit_bwd(c, n, radius_of(order));
    real acc = 0;
    for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (f[id + (r - 1) * st] - f[id - r * st]);
    return acc;
}

/* sum_r c_r (f[c + r] - f[c - r + 1]) */
static real dfwd(const real* f, long long id, long long st, long long c, long long n, int order)
{
    if (order <= 2) return f[id + st] - f[id];
    int R = fit_fwd(c, n, radius_of(order));
    real acc = 0;
    for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (f[id + r * st] - f[id - (r - 1) * st]);
    return acc;
}

/* Transpose of dbwd evaluated at target t: each contributing source point s
 * uses its own clamped radius. */
static real tbwd(const real* w, long long t, long long st, long long c, long long n, int order)
{
    if (order <= 2) {
        real acc = w[t];
        if (c + 1 < n) acc -= w[t + st];
        return acc;
    }
    int R = radius_of(order);
    real acc = 0;
    if (c >= 2 * R - 1 && c + 2 * R - 1 < n) {
        for (int r = 1; r <= R; ++r) {
            real k = stencil_c(R, r);
            acc += k * w[t - (r - 1) * st];
            acc -= k * w[t + r * st];
        }
        return acc;
    }
    for (int r = 1; r <= R; ++r) {
        long long sp = c - (r - 1);
        if (sp >= 0) {
            int rs = fit_bwd(sp, n, R);
            if (r <= rs) acc += stencil_c(rs, r) * w[t - (r - 1) * st];
        }
        long long sm = c + r;
        if (sm < n) {
            int rs = fit_bwd(sm, n, R);
            if (r <= rs) acc -= stencil_c(rs, r) * w[t + r * st];
        }
    }
    return acc;
}

static real tfwd(const real* w, long long t, long long st, long long c, long long n, int order)
{
    if (order <= 2) {
        real acc = 0;
        if (c >= 1) acc += w[t - st];
        acc -= w[t];
        return acc;
    }
    int R = radius_of(order);
    real acc = 0;
    if (c >= 2 * R - 1 && c + 2 * R <= n) {
        for (int r = 1; r <= R; ++r) {
            real k = stencil_c(R, r);
            acc += k * w[t - r * st];
            acc -= k * w[t + (r - 1) * st];
        }
        return acc;
    }
    for (int r = 1; r <= R; ++r) {
        long long sp = c - r;
        if (sp >= 0) {
            int rs = fit_fwd(sp, n, R);
            if (r <= rs) acc += stencil_c(rs, r) * w[t - r * st];
        }
        long long sm = c + (r - 1);
        if (sm < n) {
            int rs = fit_fwd(sm, n, R);
            if (r <= rs) acc -= stencil_c(rs, r) * w[t + (r - 1) * st];
        }
    }
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Update masks (deepgpr_cpu.c:697-699, 913-915, 1538-1551)                  */
/* ------------------------------------------------------------------------ */

static int e_updated(const Grid* g, int comp, const int c[3])
{
    int a1 = (comp + 1) % 3, a2 = (comp + 2) % 3;
    /* A component is frozen when both transverse axes are degenerate. */
    if (g->N[a1] - 1 == 1 && g->N[a2] - 1 == 1) return 0;
    if (c[comp] >= g->N[comp] - 1) return 0;
    if (c[a1] <= 0 || c[a1] >= g->N[a1] - 1) return 0;
    if (c[a2] <= 0 || c[a2] >= g->N[a2] - 1) return 0;
    return 1;
}

static int h_updated(const Grid* g, int comp, const int c[3])
{
    int a1 = (comp + 1) % 3, a2 = (comp + 2) % 3;
    if (g->N[comp] - 1 == 1) return 0;
    if (c[comp] <= 0 || c[comp] >= g->N[comp] - 1) return 0;
    if (c[a1] >= g->N[a1] - 1 || c[a2] >= g->N[a2] - 1) return 0;
    return 1;
}

/* ratio factor of a curl coefficient on axis a: c1 * dx / d_a unless d_a == dx */
static real axis_coeff(const Grid* g, real c1, int a)
{
    if (a == 0 || g->d[a] == g->d[0]) return c1;
    return c1 * g->d[0] / g->d[a];
}

/* ------------------------------------------------------------------------ */
/* Update coefficients (deepgpr_cpu.c:547-584)                               */
/* ------------------------------------------------------------------------ */

static void build_coefficients(const Grid* g, const real* eps, const real* sig, const real* mu,
                               real* ca, real* ccurl, real* cb,
                               real* hh, real* hcurl, real* hrhs, real dt)
{
    long long total = g->cells;
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < total; ++n) {
        long long i = n / g->st[0], j = (n / g->st[1]) % g->N[1], k = n % g->N[2];
        if (!(i < g->N[0] - 1 && j < g->N[1] - 1 && k < g->N[2] - 1)) continue;
        real ha = MU0 * mu[n] / dt;
        hh[n] = (real)1;
        hcurl[n] = ((real)1 / g->d[0]) / ha;
        hrhs[n] = (real)1 / ha;
        if (sig[n] > PEC_SIGMA) {
            ca[n] = 0;
            ccurl[n] = 0;
            cb[n] = 0;
        } else {
            real et = EPS0 * eps[n] / dt;
            real st = (real)0.5 * sig[n];
            real ea = et + st, eb = et - st;
            ca[n] = eb / ea;
            ccurl[n] = ((real)1 / g->d[0]) / ea;
            cb[n] = (real)1 / ea;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Forward half steps                                                       */
/* ------------------------------------------------------------------------ */

static void cell_coords(const Grid* g, long long n, int* s, int c[3])
{
    *s = (int)(n / g->cells);
    long long r = n % g->cells;
    c[0] = (int)(r / g->st[0]);
    c[1] = (int)((r / g->st[1]) % g->N[1]);
    c[2] = (int)(r % g->N[2]);
}

/* CPML corrections of one cell in face order x0, xm, y0, ym, z0, zm. */
static void cpml_cell(const Grid* g, int field, int s, const int c[3], long long id, real upd,
                      real* const tgt[3], const real* const src[3])
{
    for (int face = 0; face < 6; ++face) {
        int axis = face / 2;
        int q = face_q(g, field, face, c[axis]);
        if (q < 0) continue;
        const real* R = g->coef[field][face];
        int p = g->pml[face];
        real ra1 = R[q] - (real)1, rb = R[p + q], re = R[2 * p + q], rf = R[3 * p + q];
        for (int which = 0; which < 2; ++which) {
            int dims[3];
            psi_dims(g, field, face, which, dims);
            int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
            if (c[a1] >= dims[a1] || c[a2] >= dims[a2]) continue;
            int t = field == 0 ? E_TARGET[axis][which] : H_TARGET[axis][which];
            int sc = field == 0 ? E_SOURCE[axis][which] : H_SOURCE[axis][which];
            int sign = field == 0 ? E_SIGN[axis][which] : H_SIGN[axis][which];
            real der = field == 0
                ? dbwd(src[sc], id, g->st[axis], c[axis], g->N[axis], g->order) / g->d[axis]
                : dfwd(src[sc], id, g->st[axis], c[axis], g->N[axis], g->order) / g->d[axis];
            real* psi = g->psi[field][face][which];
            long long pi = psi_index(dims, s, c, axis, q);
            real phi = psi[pi];
            real corr = upd * (ra1 * der + rb * phi);
            if (sign < 0) tgt[t][id] -= corr;
            else tgt[t][id] += corr;
            psi[pi] = re * phi - rf * der;
        }
    }
}

static int has_cpml(const Grid* g)
{
    for (int f = 0; f < 6; ++f)
        if (g->pml[f]) return 1;
    return 0;
}

/* H^(n+1/2): base curl update then CPML (deepgpr_cpu.c:886-1100). The two are
 * fused per cell: both read only E, and H is written only at its own cell. */
static void h_step(const Grid* g, const real* hh, const real* hcurl, const real* hrhs,
                   real* const E[3], real* const H[3])
{
    long long total = (long long)g->nshot * g->cells;
    int cpml = has_cpml(g);
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < total; ++n) {
        int s, c[3];
        cell_coords(g, n, &s, c);
        long long m = n % g->cells;
        real u0 = hh[m], u1 = hcurl[m];
        real u[3] = {u1, axis_coeff(g, u1, 1), axis_coeff(g, u1, 2)};
        /* H_t = u0 H_t - u_a2 D+_a2(E_a1') ... written per component */
        if (h_updated(g, 0, c)) {
            real dEz = dfwd(E[2], n, g->st[1], c[1], g->N[1], g->order);
            real dEy = dfwd(E[1], n, g->st[2], c[2], g->N[2], g->order);
            H[0][n] = u0 * H[0][n] - u[1] * dEz + u[2] * dEy;
        }
        if (h_updated(g, 1, c)) {
            real dEx = dfwd(E[0], n, g->st[2], c[2], g->N[2], g->order);
            real dEz = dfwd(E[2], n, g->st[0], c[0], g->N[0], g->order);
            H[1][n] = u0 * H[1][n] - u[2] * dEx + u[0] * dEz;
        }
        if (h_updated(g, 2, c)) {
            real dEy = dfwd(E[1], n, g->st[0], c[0], g->N[0], g->order);
            real dEx = dfwd(E[0], n, g->st[1], c[1], g->N[1], g->order);
            H[2][n] = u0 * H[2][n] - u[0] * dEy + u[1] * dEx;
        }
        if (cpml) cpml_cell(g, 1, s, c, n, hrhs[m], H, (const real* const*)E);
    }
}

static void e_step(const Grid* g, const real* ca, const real* ccurl, const real* cb,
                   real* const E[3], real* const H[3])
{
    long long total = (long long)g->nshot * g->cells;
    int cpml = has_cpml(g);
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < total; ++n) {
        int s, c[3];
        cell_coords(g, n, &s, c);
        long long m = n % g->cells;
        real u0 = ca[m], u1 = ccurl[m];
        real u[3] = {u1, axis_coeff(g, u1, 1), axis_coeff(g, u1, 2)};
        if (e_updated(g, 0, c)) {
            real dHz = dbwd(H[2], n, g->st[1], c[1], g->N[1], g->order);
            real dHy = dbwd(H[1], n, g->st[2], c[2], g->N[2], g->order);
            E[0][n] = u0 * E[0][n] + u[1] * dHz - u[2] * dHy;
        }
        if (e_updated(g, 1, c)) {
            real dHx = dbwd(H[0], n, g->st[2], c[2], g->N[2], g->order);
            real dHz = dbwd(H[2], n, g->st[0], c[0], g->N[0], g->order);
            E[1][n] = u0 * E[1][n] + u[2] * dHx - u[0] * dHz;
        }
        if (e_updated(g, 2, c)) {
            real dHy = dbwd(H[1], n, g->st[0], c[0], g->N[0], g->order);
            real dHx = dbwd(H[0], n, g->st[1], c[1], g->N[1], g->order);
            E[2][n] = u0 * E[2][n] + u[0] * dHy - u[1] * dHx;
        }
        if (cpml) cpml_cell(g, 0, s, c, n, cb[m], E, (const real* const*)H);
    }
}

/* ------------------------------------------------------------------------ */
/* Sources and receivers (deepgpr_cpu.c:600-668, 1479-1534)                  */
/* ------------------------------------------------------------------------ */

static long long loc_index(const Grid* g, const int* loc)
{
    return (long long)loc[0] * g->st[0] + (long long)loc[1] * g->st[1] + loc[2];
}

static void inject(const Grid* g, int it, int nt, int nsrc, const int* sloc, const real* wave,
                   int comp, real* const E[3], const real* cb)
{
    if (comp < 0 || comp > 2) return;
    real dl = g->d[comp];
    for (int s = 0; s < g->nshot; ++s) {
        for (int k = 0; k < nsrc; ++k) {
            real scale = wave[(long long)k * nt + it] * dl / (g->d[0] * g->d[1] * g->d[2]);
            long long m = loc_index(g, sloc + ((long long)s * nsrc + k) * 3);
            E[comp][(long long)s * g->cells + m] -= cb[m] * scale;
        }
    }
}

static void sample(const Grid* g, int it, int nt, int nrx, const int* rloc, real* out,
                   int comp, real* const E[3])
{
    const real* f = E[comp < 0 || comp > 2 ? 2 : comp];
    for (int s = 0; s < g->nshot; ++s)
        for (int r = 0; r < nrx; ++r) {
            long long m = loc_index(g, rloc + ((long long)s * nrx + r) * 3);
            out[((long long)s * nt + it) * nrx + r] = f[(long long)s * g->cells + m];
        }
}

/* ------------------------------------------------------------------------ */
/* History frames                                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    int x0, y0, z0, nx, ny, nz;
    long long cells; /* nx*ny*nz */
} Box;

static void box_axis(int physical, int cells, int lo, int hi, int tile, int* o, int* e)
{
    if (!physical) {
        *o = 0;
        *e = cells;
        return;
    }
    int begin = (lo / tile) * tile;
    int end = ((cells - hi + tile - 1) / tile) * tile;
    if (end > cells) end = cells;
    *o = begin;
    *e = end - begin;
}

/* deepgpr_cpu.c:253-268; with INT8 tiles the physical box is widened to whole
 * tiles (deepgpr.cu:317-342). */
static Box make_box(const Grid* g, int physical, const int tile[3])
{
    Box b;
    box_axis(physical, g->N[0] - 1, g->pml[0], g->pml[1], tile[0], &b.x0, &b.nx);
    box_axis(physical, g->N[1] - 1, g->pml[2], g->pml[3], tile[1], &b.y0, &b.ny);
    box_axis(physical, g->N[2] - 1, g->pml[4], g->pml[5], tile[2], &b.z0, &b.nz);
    b.cells = (long long)b.nx * b.ny * b.nz;
    return b;
}

static long long box_field_index(const Grid* g, const Box* b, int s, long long h)
{
    long long i = h / ((long long)b->ny * b->nz) + b->x0;
    long long j = (h / b->nz) % b->ny + b->y0;
    long long k = h % b->nz + b->z0;
    return (long long)s * g->cells + i * g->st[0] + j * g->st[1] + k;
}

/* INT8 packed history (storage.py:377-414, deepgpr.cu:2885-3030). */
typedef struct {
    int bx, by, bz;
    long long nbx, nby, nbz, tiles; /* tiles per shot */
    long long q_aligned;            /* payload bytes rounded up to 4 */
} Int8Layout;

static Int8Layout int8_layout(const Box* b, int comps, int frames, int nshot, const int tile[3])
{
    Int8Layout L;
    L.bx = tile[0];
    L.by = tile[1];
    L.bz = tile[2];
    L.nbx = (b->nx + L.bx - 1) / L.bx;
    L.nby = (b->ny + L.by - 1) / L.by;
    L.nbz = (b->nz + L.bz - 1) / L.bz;
    L.tiles = L.nbx * L.nby * L.nbz;
    long long q = (long long)comps * frames * nshot * b->cells;
    L.q_aligned = (q + 3) & ~3LL;
    return L;
}

/* Quantise one (component, frame, shot) snapshot given as box-ordered values. */
static void int8_encode_snapshot(void* packed, const Int8Layout* L, const Box* b, int comp,
                                 int frames, int frame, int nshot, int s, const real* values)
{
    signed char* q = (signed char*)packed;
    float* scales = (float*)((unsigned char*)packed + L->q_aligned);
    long long snap = (long long)nshot * b->cells;
    long long qbase = ((long long)comp * frames + frame) * snap + (long long)s * b->cells;
    long long sbase = (((long long)comp * frames + frame) * nshot + s) * L->tiles;
    for (long long tx = 0; tx < L->nbx; ++tx)
        for (long long ty = 0; ty < L->nby; ++ty)
            for (long long tz = 0; tz < L->nbz; ++tz) {
                float peak = 0.0f;
                for (int a = 0; a < L->bx; ++a)
                    for (int bb = 0; bb < L->by; ++bb)
                        for (int cc = 0; cc < L->bz; ++cc) {
                            long long x = tx * L->bx + a, y = ty * L->by + bb, z = tz * L->bz + cc;
                            if (x >= b->nx || y >= b->ny || z >= b->nz) continue;
                            float v = (float)values[(x * b->ny + y) * b->nz + z];
                            float m = v < 0.0f ? -v : v;
                            if (m > peak) peak = m;
                        }
                float scale = peak >= 1.0e-30f ? peak / 127.0f : 1.0f;
                scales[sbase + (tx * L->nby + ty) * L->nbz + tz] = scale;
                for (int a = 0; a < L->bx; ++a)
                    for (int bb = 0; bb < L->by; ++bb)
                        for (int cc = 0; cc < L->bz; ++cc) {
                            long long x = tx * L->bx + a, y = ty * L->by + bb, z = tz * L->bz + cc;
                            if (x >= b->nx || y >= b->ny || z >= b->nz) continue;
                            long long h = (x * b->ny + y) * b->nz + z;
                            float r = rintf((float)values[h] / scale);
                            if (r > 127.0f) r = 127.0f;
                            if (r < -127.0f) r = -127.0f;
                            q[qbase + h] = (signed char)(int)r;
                        }
            }
}

static real int8_decode(const void* packed, const Int8Layout* L, const Box* b, int comp,
                        int frames, int frame, int nshot, int s, long long h)
{
    const signed char* q = (const signed char*)packed;
    const float* scales = (const float*)((const unsigned char*)packed + L->q_aligned);
    long long snap = (long long)nshot * b->cells;
    long long x = h / ((long long)b->ny * b->nz), y = (h / b->nz) % b->ny, z = h % b->nz;
    long long tile = ((x / L->bx) * L->nby + y / L->by) * L->nbz + z / L->bz;
    float scale = scales[(((long long)comp * frames + frame) * nshot + s) * L->tiles + tile];
    signed char v = q[((long long)comp * frames + frame) * snap + (long long)s * b->cells + h];
    return (real)(scale * (float)v);
}

/* ------------------------------------------------------------------------ */
/* Forward driver (deepgpr_cpu.c:1782-1934)                                  */
/* ------------------------------------------------------------------------ */

#define PHI_PARAMS(p) \
    real *p##x0e1, real *p##x0e2, real *p##x0h1, real *p##x0h2, \
    real *p##xme1, real *p##xme2, real *p##xmh1, real *p##xmh2, \
    real *p##y0e1, real *p##y0e2, real *p##y0h1, real *p##y0h2, \
    real *p##yme1, real *p##yme2, real *p##ymh1, real *p##ymh2, \
    real *p##z0e1, real *p##z0e2, real *p##z0h1, real *p##z0h2, \
    real *p##zme1, real *p##zme2, real *p##zmh1, real *p##zmh2

#define PHI_TABLE(p) { \
    {p##x0e1, p##x0e2, p##x0h1, p##x0h2}, {p##xme1, p##xme2, p##xmh1, p##xmh2}, \
    {p##y0e1, p##y0e2, p##y0h1, p##y0h2}, {p##yme1, p##yme2, p##ymh1, p##ymh2}, \
    {p##z0e1, p##z0e2, p##z0h1, p##z0h2}, {p##zme1, p##zme2, p##zmh1, p##zmh2}}

static void fill_grid(Grid* g, int NX, int NY, int NZ, int nshot, int order, const int pml[6],
                      real dx, real dy, real dz, const real* ce[6], const real* chh[6],
                      real* phi[6][4])
{
    g->N[0] = NX;
    g->N[1] = NY;
    g->N[2] = NZ;
    g->st[2] = 1;
    g->st[1] = NZ;
    g->st[0] = (long long)NY * NZ;
    g->cells = (long long)NX * NY * NZ;
    g->nshot = nshot;
    g->order = (order == 4 || order == 8) ? order : 2;
    g->d[0] = dx;
    g->d[1] = dy;
    g->d[2] = dz;
    for (int f = 0; f < 6; ++f) {
        g->pml[f] = pml[f];
        g->coef[0][f] = ce[f];
        g->coef[1][f] = chh[f];
        g->psi[0][f][0] = phi[f][0];
        g->psi[0][f][1] = phi[f][1];
        g->psi[1][f][0] = phi[f][2];
        g->psi[1][f][1] = phi[f][3];
    }
}

static void tile_of(int storage_type, int NZ, int tile[3])
{
    if ((storage_type & KIND_MASK) == KIND_INT8) {
        tile[0] = (storage_type >> 8) & 0x3f;
        tile[1] = (storage_type >> 14) & 0x3f;
        tile[2] = (storage_type >> 20) & 0x3f;
        if (tile[0] <= 0) tile[0] = NZ == 2 ? 8 : 4;
        if (tile[1] <= 0) tile[1] = NZ == 2 ? 8 : 4;
        if (tile[2] <= 0) tile[2] = NZ == 2 ? 1 : 4;
    } else {
        tile[0] = tile[1] = tile[2] = 1;
    }
}

EXPORT void API(forward)(
    const real* eps, const real* sig, const real* mu, void* E_saved, void* R_saved,
    real* Ex, real* Ey, real* Ez, real* Hx, real* Hy, real* Hz,
    real* ca, real* ccurl, real* cb, real* hh, real* hcurl, real* hrhs,
    PHI_PARAMS(),
    int pml0, int pml1, int pml2, int pml3, int pml4, int pml5,
    const real* x0e, const real* xme, const real* y0e, const real* yme, const real* z0e, const real* zme,
    const real* x0h, const real* xmh, const real* y0h, const real* ymh, const real* z0h, const real* zmh,
    real dt, int nt, int nshot, int nrx, real dx, real dy, real dz,
    const int* rloc, real* rdata, int rcomp,
    int NX, int NY, int NZ, int nsrc, const int* sloc, const real* wave, int scomp,
    int S, int mode, int storage_type, int save_model_history, int use_async_offload, int order)
{
    (void)use_async_offload;
    g_error[0] = 0;
    int pml[6] = {pml0, pml1, pml2, pml3, pml4, pml5};
    const real* ce[6] = {x0e, xme, y0e, yme, z0e, zme};
    const real* chh[6] = {x0h, xmh, y0h, ymh, z0h, zmh};
    real* phi[6][4] = PHI_TABLE();
    Grid g;
    fill_grid(&g, NX, NY, NZ, nshot, order, pml, dx, dy, dz, ce, chh, phi);
    real* E[3] = {Ex, Ey, Ez};
    real* H[3] = {Hx, Hy, Hz};

    int history = (storage_type & FLAG_NO_HISTORY) == 0;
    int kind = storage_type & KIND_MASK;
    int e_only = history && save_model_history && (storage_type & FLAG_RHS_FROM_E);
    int store_r = history && save_model_history && !e_only;
    int physical = (storage_type & FLAG_PHYSICAL) != 0;
    int tile[3];
    tile_of(storage_type, NZ, tile);
    Box box = make_box(&g, physical, tile);
    int comps = mode == 3 ? 3 : 1;
    int nt_saved = (nt + S - 1) / S;
    long long snap = (long long)nshot * box.cells;
    long long cstride = (long long)nt_saved * snap;

    if (history && save_model_history && has_cpml(&g) && physical) {
        set_error("oracle forward: material gradients with PML need the extended history");
        return;
    }
    if (e_only && S != 1) {
        set_error("oracle forward: E-only history requires sampling interval 1");
        return;
    }
    if (sizeof(real) == sizeof(double) && history && kind != KIND_F32) {
        set_error("oracle forward: the float64 oracle stores lossless histories only");
        return;
    }

    Int8Layout Lh = int8_layout(&box, comps, nt_saved, nshot, tile);
    Int8Layout Lf = int8_layout(&box, comps, 1, nshot, tile);
    real* exact = NULL;   /* exact E^n per component for R (lossy kinds) */
    real* scratch = NULL; /* one box-ordered snapshot for INT8 encoding */
    if (history) {
        exact = (real*)malloc(sizeof(real) * (size_t)(comps * snap + 1));
        scratch = (real*)malloc(sizeof(real) * (size_t)(box.cells + 1));
        if (!exact || !scratch) {
            free(exact);
            free(scratch);
            set_error("oracle forward: out of memory");
            return;
        }
    }

    build_coefficients(&g, eps, sig, mu, ca, ccurl, cb, hh, hcurl, hrhs, dt);

    for (int it = 0; it < nt; ++it) {
        int save = history && (it % S) == 0;
        int ts = it / S;
        if (save) {
            for (int ci = 0; ci < comps; ++ci) {
                int comp = mode == 3 ? ci : 2;
                for (int s = 0; s < nshot; ++s) {
                    for (long long h = 0; h < box.cells; ++h) {
                        real v = E[comp][box_field_index(&g, &box, s, h)];
                        exact[(long long)ci * snap + (long long)s * box.cells + h] = v;
                        if (kind != KIND_INT8)
                            store_value(E_saved, (long long)ci * cstride + (long long)ts * snap
                                                     + (long long)s * box.cells + h, v, kind);
                    }
                    if (kind == KIND_INT8)
                        int8_encode_snapshot(E_saved, &Lh, &box, ci, nt_saved, ts, nshot, s,
                                             exact + (long long)ci * snap + (long long)s * box.cells);
                }
            }
        }
        h_step(&g, hh, hcurl, hrhs, E, H);
        e_step(&g, ca, ccurl, cb, E, H);
        inject(&g, it, nt, nsrc, sloc, wave, scomp, E, cb);
        if (save && store_r) {
            for (int ci = 0; ci < comps; ++ci) {
                int comp = mode == 3 ? ci : 2;
                for (int s = 0; s < nshot; ++s) {
                    for (long long h = 0; h < box.cells; ++h) {
                        long long f = box_field_index(&g, &box, s, h);
                        long long m = f % g.cells;
                        int cc[3] = {(int)(m / g.st[0]), (int)((m / g.st[1]) % NY), (int)(m % NZ)};
                        real eold = exact[(long long)ci * snap + (long long)s * box.cells + h];
                        real a = e_updated(&g, comp, cc) ? ca[m] : (real)1;
                        real rhs = cb[m] != 0 ? (E[comp][f] - a * eold) / cb[m] : (real)0;
                        if (kind == KIND_INT8) scratch[h] = rhs;
                        else
                            store_value(R_saved, (long long)ci * cstride + (long long)ts * snap
                                                     + (long long)s * box.cells + h, rhs, kind);
                    }
                    if (kind == KIND_INT8)
                        int8_encode_snapshot(R_saved, &Lh, &box, ci, nt_saved, ts, nshot, s, scratch);
                }
            }
        }
        sample(&g, it, nt, nrx, rloc, rdata, rcomp, E);
    }

    if (e_only) {
        /* the final frame E^nt closes the last (E^n, E^(n+1)) pair */
        for (int ci = 0; ci < comps; ++ci) {
            int comp = mode == 3 ? ci : 2;
            for (int s = 0; s < nshot; ++s) {
                for (long long h = 0; h < box.cells; ++h) {
                    real v = E[comp][box_field_index(&g, &box, s, h)];
                    if (kind == KIND_INT8) scratch[h] = v;
                    else store_value(R_saved, (long long)ci * snap + (long long)s * box.cells + h, v, kind);
                }
                if (kind == KIND_INT8)
                    int8_encode_snapshot(R_saved, &Lf, &box, ci, 1, 0, nshot, s, scratch);
            }
        }
    }
    free(exact);
    free(scratch);
}

/* ------------------------------------------------------------------------ */
/* Adjoint                                                                  */
/* ------------------------------------------------------------------------ */

/* Channel of a transposed stencil: (source component s, axis a), s != a.
 * Channel order per target (ascending axis): see the gather below. */
static int chan(int s, int a)
{
    /* (2,0)=0 (2,1)=1 (1,0)=2 (1,2)=3 (0,1)=4 (0,2)=5 */
    static const int T[3][3] = {{-1, 4, 5}, {2, -1, 3}, {0, 1, -1}};
    return T[s][a];
}

/* Transpose of one half step. field 0: (E-CPML + E update)^T -> lambda_H
 * gathers backward-difference transposes; field 1: (H-CPML + H update)^T ->
 * lambda_E gathers forward-difference transposes (deepgpr_cpu.c:1112-1459). */
static void adjoint_half(const Grid* g, int field, const real* hist, const real* curl,
                         const real* upd, real* const L[3], real* const M[3], real* w)
{
    long long total = (long long)g->nshot * g->cells;
    int cpml = has_cpml(g);
    real* W[6];
    for (int ch = 0; ch < 6; ++ch) W[ch] = w + (long long)ch * total;

#pragma omp parallel for schedule(static)
    for (long long n = 0; n < total; ++n) {
        int s, c[3];
        cell_coords(g, n, &s, c);
        long long m = n % g->cells;
        real lam[3] = {L[0][n], L[1][n], L[2][n]};
        real acc[6] = {0, 0, 0, 0, 0, 0};
        if (cpml) {
            real u = upd[m];
            for (int face = 0; face < 6; ++face) {
                int axis = face / 2;
                int q = face_q(g, field, face, c[axis]);
                if (q < 0) continue;
                const real* R = g->coef[field][face];
                int p = g->pml[face];
                real ra1 = R[q] - (real)1, rb = R[p + q], re = R[2 * p + q], rf = R[3 * p + q];
                real inv = (real)1 / g->d[axis];
                for (int which = 0; which < 2; ++which) {
                    int dims[3];
                    psi_dims(g, field, face, which, dims);
                    int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
                    if (c[a1] >= dims[a1] || c[a2] >= dims[a2]) continue;
                    int t = field == 0 ? E_TARGET[axis][which] : H_TARGET[axis][which];
                    int sc = field == 0 ? E_SOURCE[axis][which] : H_SOURCE[axis][which];
                    real sign = (real)(field == 0 ? E_SIGN[axis][which] : H_SIGN[axis][which]);
                    real* psi = g->psi[field][face][which];
                    long long pi = psi_index(dims, s, c, axis, q);
                    real old = psi[pi];
                    acc[chan(sc, axis)] += (sign * u * ra1 * lam[t] - rf * old) * inv;
                    psi[pi] = sign * u * rb * lam[t] + re * old;
                }
            }
        }
        real k1 = curl[m];
        real k[3] = {k1, axis_coeff(g, k1, 1), axis_coeff(g, k1, 2)};
        for (int t = 0; t < 3; ++t) {
            int up = field == 0 ? e_updated(g, t, c) : h_updated(g, t, c);
            if (!up) continue;
            int a1 = (t + 1) % 3, a2 = (t + 2) % 3;
            /* curl_t = d_{a1} F_{a2} - d_{a2} F_{a1}; H uses the negated curl */
            if (field == 0) {
                acc[chan(a2, a1)] += k[a1] * lam[t];
                acc[chan(a1, a2)] += -k[a2] * lam[t];
            } else {
                acc[chan(a2, a1)] += -k[a1] * lam[t];
                acc[chan(a1, a2)] += k[a2] * lam[t];
            }
            L[t][n] = hist[m] * lam[t];
        }
        for (int ch = 0; ch < 6; ++ch) W[ch][n] = acc[ch];
    }

#pragma omp parallel for schedule(static)
    for (long long n = 0; n < total; ++n) {
        int s, c[3];
        cell_coords(g, n, &s, c);
        for (int t = 2; t >= 0; --t) {
            /* target order z, y, x with channels in ascending axis order */
            for (int a = 0; a < 3; ++a) {
                if (a == t) continue;
                const real* ww = W[chan(t, a)];
                real v = field == 0 ? tbwd(ww, n, g->st[a], c[a], g->N[a], g->order)
                                    : tfwd(ww, n, g->st[a], c[a], g->N[a], g->order);
                M[t][n] += v;
            }
        }
    }
}

/* deepgpr_cpu.c:1659-1745 */
static void material_gradient(const Grid* g, const real* const L[3], const void* eframe,
                              long long estride, const void* second, long long sstride, int kind,
                              const Int8Layout* Le, int eframes, int eframe_index,
                              const Int8Layout* Ls, int sframes, int sframe_index,
                              const real* ca, const real* cb, const real* sig, real* geps,
                              real* gsig, real dt, int want_eps, int want_sig, int weight, int mode,
                              int e_only, const Box* box)
{
    long long total = box->cells;
    int comps = mode == 3 ? 3 : 1;
#pragma omp parallel for schedule(static)
    for (long long h = 0; h < total; ++h) {
        long long ix = h / ((long long)box->ny * box->nz) + box->x0;
        long long iy = (h / box->nz) % box->ny + box->y0;
        long long iz = h % box->nz + box->z0;
        long long m = ix * g->st[0] + iy * g->st[1] + iz;
        long long out = (ix * (g->N[1] - 1) + iy) * (g->N[2] - 1) + iz;
        int c3[3] = {(int)ix, (int)iy, (int)iz};
        real ge = 0, gs = 0;
        for (int s = 0; s < g->nshot; ++s) {
            long long f = (long long)s * g->cells + m;
            long long base = (long long)s * box->cells + h;
            for (int ci = 0; ci < comps; ++ci) {
                int comp = mode == 3 ? ci : 2;
                int up = e_updated(g, comp, c3);
                real e_old, sec;
                if (kind == KIND_INT8) {
                    e_old = int8_decode(eframe, Le, box, ci, eframes, eframe_index, g->nshot, s, h);
                    sec = int8_decode(second, Ls, box, ci, sframes, sframe_index, g->nshot, s, h);
                } else {
                    e_old = load_value(eframe, (long long)ci * estride + base, kind);
                    sec = load_value(second, (long long)ci * sstride + base, kind);
                }
                real lam = L[comp][f];
                real cav = ca[m], cbv = cb[m];
                real rhs = sec;
                if (e_only) {
                    real a = up ? cav : (real)1;
                    rhs = cbv != 0 ? (sec - a * e_old) / cbv : (real)0;
                }
                if (e_old == 0 && rhs == 0) lam = 0;
                real gca = up ? lam * e_old * (real)weight : (real)0;
                real gcb = lam * rhs * (real)weight;
                if (sig[m] <= PEC_SIGMA) {
                    if (want_eps) {
                        real dca = EPS0 * ((real)1 - cav) * cbv / dt;
                        real dcb = -EPS0 * cbv * cbv / dt;
                        ge += gca * dca + gcb * dcb;
                    }
                    if (want_sig) {
                        real dca = (real)-0.5 * ((real)1 + cav) * cbv;
                        real dcb = (real)-0.5 * cbv * cbv;
                        gs += gca * dca + gcb * dcb;
                    }
                }
            }
        }
        if (want_eps) geps[out] += ge;
        if (want_sig) gsig[out] += gs;
    }
}

EXPORT void API(backward)(
    const real* eps, const real* sig, const real* mu, const void* E_saved, const void* R_saved,
    real* lex, real* ley, real* lez, real* lhx, real* lhy, real* lhz,
    real* ca, real* ccurl, real* cb, real* hh, real* hcurl, real* hrhs,
    PHI_PARAMS(l),
    int pml0, int pml1, int pml2, int pml3, int pml4, int pml5,
    const real* x0e, const real* xme, const real* y0e, const real* yme, const real* z0e, const real* zme,
    const real* x0h, const real* xmh, const real* y0h, const real* ymh, const real* z0h, const real* zmh,
    real dt, int nt, int nshot, int nrx, real dx, real dy, real dz,
    int NX, int NY, int NZ, int ndata, const int* rloc, const real* dgrad, int rcomp,
    int nsrc, const int* sloc, int scomp, real* gsrc, int want_src,
    real* geps, real* gsig, int want_eps, int want_sig,
    int S, int mode, int storage_type, int use_async_offload, int order)
{
    (void)nrx;
    (void)use_async_offload;
    g_error[0] = 0;
    int pml[6] = {pml0, pml1, pml2, pml3, pml4, pml5};
    const real* ce[6] = {x0e, xme, y0e, yme, z0e, zme};
    const real* chh[6] = {x0h, xmh, y0h, ymh, z0h, zmh};
    real* phi[6][4] = PHI_TABLE(l);
    Grid g;
    fill_grid(&g, NX, NY, NZ, nshot, order, pml, dx, dy, dz, ce, chh, phi);
    real* L[3] = {lex, ley, lez};
    real* M[3] = {lhx, lhy, lhz};
    int kind = storage_type & KIND_MASK;
    int want_mat = want_eps || want_sig;
    int e_only = want_mat && (storage_type & FLAG_RHS_FROM_E);
    int physical = (storage_type & FLAG_PHYSICAL) != 0;
    int tile[3];
    tile_of(storage_type, NZ, tile);
    Box box = make_box(&g, physical, tile);
    int comps = mode == 3 ? 3 : 1;
    int nt_saved = (nt + S - 1) / S;
    long long snap = (long long)nshot * box.cells;
    long long cstride = (long long)nt_saved * snap;
    size_t kb = kind_bytes(kind);

    if (want_mat && has_cpml(&g) && physical) {
        set_error("oracle backward: material gradients with PML need the extended history");
        return;
    }
    if (e_only && S != 1) {
        set_error("oracle backward: E-only history requires sampling interval 1");
        return;
    }
    long long total = (long long)nshot * g.cells;
    real* w = (real*)malloc(sizeof(real) * (size_t)(6 * total));
    if (!w) {
        set_error("oracle backward: out of memory");
        return;
    }
    Int8Layout Lh = int8_layout(&box, comps, nt_saved, nshot, tile);
    Int8Layout Lf = int8_layout(&box, comps, 1, nshot, tile);

    build_coefficients(&g, eps, sig, mu, ca, ccurl, cb, hh, hcurl, hrhs, dt);

    for (int it = nt - 1; it >= 0; --it) {
        /* receivers^T: shots independent, receivers in index order */
        if (rcomp >= 0 && rcomp <= 2)
            for (int s = 0; s < nshot; ++s)
                for (int r = 0; r < ndata; ++r) {
                    long long mm = loc_index(&g, rloc + ((long long)s * ndata + r) * 3);
                    L[rcomp][(long long)s * g.cells + mm] += dgrad[((long long)s * nt + it) * ndata + r];
                }
        if (want_src == 1 && scomp >= 0 && scomp <= 2) {
            real geo = g.d[scomp] / (g.d[0] * g.d[1] * g.d[2]);
            for (int k = 0; k < nsrc; ++k)
                for (int s = 0; s < nshot; ++s) {
                    long long mm = loc_index(&g, sloc + ((long long)s * nsrc + k) * 3);
                    gsrc[(long long)k * nt + it] += -cb[mm] * geo * L[scomp][(long long)s * g.cells + mm];
                }
        }
        if (want_mat && it % S == 0) {
            int wgt = S;
            int ts = it / S;
            if (it + wgt > nt) wgt = nt - it;
            const void* ef;
            const void* sf;
            long long sstride = cstride;
            const Int8Layout* Ls = &Lh;
            int sframes = nt_saved, sidx = ts;
            if (kind == KIND_INT8) {
                ef = E_saved;
                if (!e_only) {
                    sf = R_saved;
                } else if (ts + 1 < nt_saved) {
                    sf = E_saved;
                    sidx = ts + 1;
                } else {
                    sf = R_saved;
                    Ls = &Lf;
                    sframes = 1;
                    sidx = 0;
                }
            } else {
                ef = (const unsigned char*)E_saved + (size_t)ts * (size_t)snap * kb;
                if (!e_only) sf = (const unsigned char*)R_saved + (size_t)ts * (size_t)snap * kb;
                else if (ts + 1 < nt_saved) sf = (const unsigned char*)E_saved + (size_t)(ts + 1) * (size_t)snap * kb;
                else {
                    sf = R_saved;
                    sstride = snap;
                }
            }
            material_gradient(&g, (const real* const*)L, ef, cstride, sf, sstride, kind, &Lh,
                              nt_saved, ts, Ls, sframes, sidx, ca, cb, sig, geps, gsig, dt,
                              want_eps == 1, want_sig == 1, wgt, mode, e_only, &box);
        }
        adjoint_half(&g, 0, ca, ccurl, cb, L, M, w);
        adjoint_half(&g, 1, hh, hcurl, hrhs, M, L, w);
    }
    free(w);
}

/* Number of saved frames and INT8 packed size, for the harness. */
EXPORT long long API(int8_packed_bytes)(int comps, int frames, int nshot, int hx, int hy, int hz,
                                        int bx, int by, int bz)
{
    long long q = (long long)comps * frames * nshot * hx * hy * hz;
    long long tiles = (long long)((hx + bx - 1) / bx) * ((hy + by - 1) / by) * ((hz + bz - 1) / bz);
    return ((q + 3) & ~3LL) + 4LL * comps * frames * nshot * tiles;
}

/* 1D views of the stencil machinery, for the closed-form tests:
 * kind 0: backward difference at coordinate c, 1: forward difference,
 * 2: transpose of the backward difference gathered at c, 3: transpose of the
 * forward difference. */
EXPORT real API(stencil_1d)(const real* f, int n, int c, int order, int kind)
{
    int ord = (order == 4 || order == 8) ? order : 2;
    switch (kind) {
    case 0: return dbwd(f, c, 1, c, n, ord);
    case 1: return dfwd(f, c, 1, c, n, ord);
    case 2: return tbwd(f, c, 1, c, n, ord);
    default: return tfwd(f, c, 1, c, n, ord);
    }
}

EXPORT real API(stencil_coefficient)(int radius, int r) { return stencil_c(radius, r); }
