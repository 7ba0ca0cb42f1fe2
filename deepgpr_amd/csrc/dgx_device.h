// Device-side building blocks of the MI355X (gfx950) FDTD hot path.
//
// Every arithmetic expression below keeps the operation order of the
// reference CPU backend (src/DeepGPR/lib/deepgpr_cpu.c) so that, compiled with
// -ffp-contract=off, the HIP path reproduces it bit for bit in FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dgx_sweep.h"

namespace dgx {

constexpr float kEps0 = 8.8541878128e-12f;   // deepgpr.h:23
constexpr float kMu0 = 1.25663706212e-06f;   // deepgpr.h:24
constexpr float kPecSigma = 100.0f;          // deepgpr.h:32

// Storage kinds of the wavefield history (constants.py:129-132).
enum : int { kF32 = 0, kF16 = 1, kBF16 = 2, kInt8 = 3 };

// Grid of one solve: field sizes include the Yee halo (NX = nx + 1).
struct Grid {
    int N[3];          // NX, NY, NZ
    long long st0;     // NY * NZ
    int plane;         // NY * NZ (fits int for every supported grid)
    long long cells;   // NX * NY * NZ
    int nshot;
    float d[3];        // dx, dy, dz
    int pml[6];        // x0, xm, y0, ym, z0, zm
};

// CPML data of one field (E or H): 6 faces x (RA, RB, RE, RF) profiles and the
// two auxiliary arrays per face, or their adjoint weights.
struct Cpml {
    const float* coef[6];
    float* psi[6][2];
};

struct F3 {
    float* c[3];
};
struct CF3 {
    const float* c[3];
};

// Saved-history box (deepgpr_cpu.c:253-268; INT8 widens it, deepgpr.cu:317-342).
struct Box {
    int x0, y0, z0, nx, ny, nz;
    long long cells;
};

// ---------------------------------------------------------------------------
// Staggered stencils (deepgpr_cpu.c:287-407)
// ---------------------------------------------------------------------------

template <int ORDER>
struct Radius {
    static constexpr int value = ORDER == 8 ? 4 : (ORDER == 4 ? 2 : 1);
};

__device__ __forceinline__ float stencil_c(int R, int r)
{
    if (R >= 4) {
        return r == 1 ? 1225.0f / 1024.0f
             : r == 2 ? -245.0f / 3072.0f
             : r == 3 ? 49.0f / 5120.0f
                      : -5.0f / 7168.0f;
    }
    if (R == 3) return r == 1 ? 75.0f / 64.0f : r == 2 ? -25.0f / 384.0f : 3.0f / 640.0f;
    if (R == 2) return r == 1 ? 9.0f / 8.0f : -1.0f / 24.0f;
    return 1.0f;
}

__device__ __forceinline__ int fit_bwd(int c, int n, int want)
{
    int r = want;
    while (r > 1 && (c < r || c + r - 1 >= n)) --r;
    return r;
}

__device__ __forceinline__ int fit_fwd(int c, int n, int want)
{
    int r = want;
    while (r > 1 && (c - r + 1 < 0 || c + r >= n)) --r;
    return r;
}

// sum_r c_r (f[c + r - 1] - f[c - r]) with the radius clamped at the faces.
// IDX is int for LDS tiles (32-bit address math) and long long for global
// fields; it is deduced from the index alone (the stride is converted to it).
template <typename T>
struct NoDeduce {
    using type = T;
};

template <int ORDER, typename IDX = long long>
__device__ __forceinline__ float dbwd(const float* __restrict__ f, IDX id, typename NoDeduce<IDX>::type st, int c,
                                      int n)
{
    if constexpr (ORDER <= 2) {
        return f[id] - f[id - st];
    } else {
        constexpr int R = Radius<ORDER>::value;
        int rr = fit_bwd(c, n, R);
        float acc = 0.0f;
        if (rr == R) {
#pragma unroll
            for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (f[id + (r - 1) * st] - f[id - r * st]);
        } else {
            for (int r = 1; r <= rr; ++r) acc += stencil_c(rr, r) * (f[id + (r - 1) * st] - f[id - r * st]);
        }
        return acc;
    }
}

// sum_r c_r (f[c + r] - f[c - r + 1])
template <int ORDER, typename IDX = long long>
__device__ __forceinline__ float dfwd(const float* __restrict__ f, IDX id, typename NoDeduce<IDX>::type st, int c,
                                      int n)
{
    if constexpr (ORDER <= 2) {
        return f[id + st] - f[id];
    } else {
        constexpr int R = Radius<ORDER>::value;
        int rr = fit_fwd(c, n, R);
        float acc = 0.0f;
        if (rr == R) {
#pragma unroll
            for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (f[id + r * st] - f[id - (r - 1) * st]);
        } else {
            for (int r = 1; r <= rr; ++r) acc += stencil_c(rr, r) * (f[id + r * st] - f[id - (r - 1) * st]);
        }
        return acc;
    }
}

// x, with its value hidden from the optimiser (an empty asm on a vector
// register, no instruction): a load at base + opaque(c ? off : 0) stays one
// unconditional load instead of becoming a load of base plus a masked load
// of base + off, which would wait for the first before branching.
__device__ __forceinline__ int opaque(int x)
{
    asm("" : "+v"(x));
    return x;
}

__device__ __forceinline__ int fit_bwd_sel(int c, int n, int R)
{
    int r = c < n - c ? c : n - c;
    return r > R ? R : (r < 1 ? 1 : r);
}

__device__ __forceinline__ int fit_fwd_sel(int c, int n, int R)
{
    int r = c + 1 < n - 1 - c ? c + 1 : n - 1 - c;
    return r > R ? R : (r < 1 ? 1 : r);
}

// Branch-free forms of dfwd / dbwd for the general (CPML shell) path of the
// 3D sweeps, in two halves: the taps (loads only) and their combination
// (dfwd_comb serves both), so that a caller issues the loads of all its
// differences before any of their arithmetic. With live, each tap the
// fitted stencil uses is loaded from the address dfwd / dbwd load, and the
// other taps read the cell itself; without live every tap reads the cell
// itself (the caller does not use the result). Every candidate radius's
// sum is formed in the reference's order and the fitted one selected, so
// the same bits come out, while a wavefront waits on memory once instead of
// once per branch.
template <int ORDER>
struct FwdTaps {
    static constexpr int R = Radius<ORDER>::value;
    float hi[R], lo[R];
    int rr;
};

template <int ORDER>
__device__ __forceinline__ void dfwd_taps(FwdTaps<ORDER>& t, const float* __restrict__ f, long long id, long long st,
                                          int c, int n, bool live)
{
    constexpr int R = Radius<ORDER>::value;
    int rr = c + 1 < n - 1 - c ? c + 1 : n - 1 - c;
    t.rr = rr > R ? R : (rr < 1 ? 1 : rr);
#pragma unroll
    for (int r = 1; r <= R; ++r) {
        const bool use = live && (ORDER <= 2 || r <= t.rr);
        const int hi = opaque(use ? r : 0), lo = opaque(use ? r - 1 : 0);
        t.hi[r - 1] = f[id + hi * st];
        t.lo[r - 1] = f[id - lo * st];
    }
}

template <int ORDER>
__device__ __forceinline__ float dfwd_comb(const FwdTaps<ORDER>& t)
{
    if constexpr (ORDER <= 2) {
        return t.hi[0] - t.lo[0];
    } else {
        constexpr int R = Radius<ORDER>::value;
        float out = 0.0f;
#pragma unroll
        for (int q = 1; q <= R; ++q) {
            float acc = 0.0f;
#pragma unroll
            for (int r = 1; r <= q; ++r) acc += stencil_c(q, r) * (t.hi[r - 1] - t.lo[r - 1]);
            out = t.rr == q ? acc : out;
        }
        return out;
    }
}

// The taps of dbwd.
template <int ORDER>
__device__ __forceinline__ void dbwd_taps(FwdTaps<ORDER>& t, const float* __restrict__ f, long long id, long long st,
                                          int c, int n, bool live)
{
    constexpr int R = Radius<ORDER>::value;
    int rr = c < n - c ? c : n - c;
    t.rr = rr > R ? R : (rr < 1 ? 1 : rr);
#pragma unroll
    for (int r = 1; r <= R; ++r) {
        const bool use = live && (ORDER <= 2 || r <= t.rr);
        const int hi = opaque(use ? r - 1 : 0), lo = opaque(use ? r : 0);
        t.hi[r - 1] = f[id + hi * st];
        t.lo[r - 1] = f[id - lo * st];
    }
}

// The unclamped differences (the full-radius branch of dbwd / dfwd), for
// cells at least R away from every face.
template <int ORDER, typename IDX = long long>
__device__ __forceinline__ float dbwd_full(const float* __restrict__ f, IDX id, typename NoDeduce<IDX>::type st)
{
    if constexpr (ORDER <= 2) {
        return f[id] - f[id - st];
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
#pragma unroll
        for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (f[id + (r - 1) * st] - f[id - r * st]);
        return acc;
    }
}

template <int ORDER, typename IDX = long long>
__device__ __forceinline__ float dfwd_full(const float* __restrict__ f, IDX id, typename NoDeduce<IDX>::type st)
{
    if constexpr (ORDER <= 2) {
        return f[id + st] - f[id];
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
#pragma unroll
        for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (f[id + r * st] - f[id - (r - 1) * st]);
        return acc;
    }
}

// Transposes of the staggered differences (deepgpr_cpu.c:430-511). The
// _full forms are the branches taken at least 2R - 1 cells from both faces
// (c >= 2R - 1 and c <= n - 2R), where no stencil is clamped.
template <int ORDER, typename WF>
__device__ __forceinline__ float gather_bwd_full(WF wf)
{
    if constexpr (ORDER <= 2) {
        float acc = wf(0);
        acc -= wf(1);
        return acc;
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
#pragma unroll
        for (int r = 1; r <= R; ++r) {
            const float kc = stencil_c(R, r);
            acc += kc * wf(-(r - 1));
            acc -= kc * wf(r);
        }
        return acc;
    }
}

template <int ORDER, typename WF>
__device__ __forceinline__ float gather_fwd_full(WF wf)
{
    if constexpr (ORDER <= 2) {
        float acc = 0.0f;
        acc += wf(-1);
        acc -= wf(0);
        return acc;
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
#pragma unroll
        for (int r = 1; r <= R; ++r) {
            const float kc = stencil_c(R, r);
            acc += kc * wf(-r);
            acc -= kc * wf(r - 1);
        }
        return acc;
    }
}

template <int ORDER, typename WF>
__device__ __forceinline__ float gather_bwd_T(WF wf, int c, int n)
{
    if constexpr (ORDER <= 2) {
        float acc = wf(0);
        if (c + 1 < n) acc -= wf(1);
        return acc;
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
        if (c >= 2 * R - 1 && c + 2 * R - 1 < n) return gather_bwd_full<ORDER>(wf);
        for (int r = 1; r <= R; ++r) {
            const int sp = c - (r - 1);
            if (sp >= 0) {
                const int rs = fit_bwd(sp, n, R);
                if (r <= rs) acc += stencil_c(rs, r) * wf(-(r - 1));
            }
            const int sm = c + r;
            if (sm < n) {
                const int rs = fit_bwd(sm, n, R);
                if (r <= rs) acc -= stencil_c(rs, r) * wf(r);
            }
        }
        return acc;
    }
}

template <int ORDER, typename WF>
__device__ __forceinline__ float gather_fwd_T(WF wf, int c, int n)
{
    if constexpr (ORDER <= 2) {
        float acc = 0.0f;
        if (c >= 1) acc += wf(-1);
        acc -= wf(0);
        return acc;
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
        if (c >= 2 * R - 1 && c + 2 * R <= n) return gather_fwd_full<ORDER>(wf);
        for (int r = 1; r <= R; ++r) {
            const int sp = c - r;
            if (sp >= 0) {
                const int rs = fit_fwd(sp, n, R);
                if (r <= rs) acc += stencil_c(rs, r) * wf(-r);
            }
            const int sm = c + (r - 1);
            if (sm < n) {
                const int rs = fit_fwd(sm, n, R);
                if (r <= rs) acc -= stencil_c(rs, r) * wf(r - 1);
            }
        }
        return acc;
    }
}

// Branch-free forms of gather_bwd_T / gather_fwd_T for the general path of
// the 3D sweeps: every tap is loaded (a tap outside the transposed stencil
// reads offset 0), and each term is added or subtracted through a select in
// the reference's order, which keeps the bits, the sign of a zero sum
// included. Inside the full-width range every tap is used with the full
// radius, the arithmetic of gather_*_full.
template <int ORDER, typename WF>
__device__ __forceinline__ float gather_bwd_sel(WF wf, int c, int n)
{
    if constexpr (ORDER <= 2) {
        const bool ok = c + 1 < n;
        const float w1 = wf(opaque(ok ? 1 : 0));
        float acc = wf(0);
        return ok ? acc - w1 : acc;
    } else {
        constexpr int R = Radius<ORDER>::value;
        float wp[R], wm[R];
        bool okp[R], okm[R];
        int rsp[R], rsm[R];
#pragma unroll
        for (int r = 1; r <= R; ++r) {
            const int sp = c - (r - 1), sm = c + r;
            rsp[r - 1] = fit_bwd_sel(sp, n, R);
            rsm[r - 1] = fit_bwd_sel(sm, n, R);
            okp[r - 1] = sp >= 0 && r <= rsp[r - 1];
            okm[r - 1] = sm < n && r <= rsm[r - 1];
            wp[r - 1] = wf(opaque(okp[r - 1] ? -(r - 1) : 0));
            wm[r - 1] = wf(opaque(okm[r - 1] ? r : 0));
        }
        float acc = 0.0f;
#pragma unroll
        for (int r = 1; r <= R; ++r) {
            const float up = acc + stencil_c(rsp[r - 1], r) * wp[r - 1];
            acc = okp[r - 1] ? up : acc;
            const float dn = acc - stencil_c(rsm[r - 1], r) * wm[r - 1];
            acc = okm[r - 1] ? dn : acc;
        }
        return acc;
    }
}

template <int ORDER, typename WF>
__device__ __forceinline__ float gather_fwd_sel(WF wf, int c, int n)
{
    if constexpr (ORDER <= 2) {
        const bool ok = c >= 1;
        const float w = wf(opaque(ok ? -1 : 0));
        const float w0 = wf(0);
        float acc = 0.0f;
        acc = ok ? acc + w : acc;
        return acc - w0;
    } else {
        constexpr int R = Radius<ORDER>::value;
        float wp[R], wm[R];
        bool okp[R], okm[R];
        int rsp[R], rsm[R];
#pragma unroll
        for (int r = 1; r <= R; ++r) {
            const int sp = c - r, sm = c + (r - 1);
            rsp[r - 1] = fit_fwd_sel(sp, n, R);
            rsm[r - 1] = fit_fwd_sel(sm, n, R);
            okp[r - 1] = sp >= 0 && r <= rsp[r - 1];
            okm[r - 1] = sm < n && r <= rsm[r - 1];
            wp[r - 1] = wf(opaque(okp[r - 1] ? -r : 0));
            wm[r - 1] = wf(opaque(okm[r - 1] ? r - 1 : 0));
        }
        float acc = 0.0f;
#pragma unroll
        for (int r = 1; r <= R; ++r) {
            const float up = acc + stencil_c(rsp[r - 1], r) * wp[r - 1];
            acc = okp[r - 1] ? up : acc;
            const float dn = acc - stencil_c(rsm[r - 1], r) * wm[r - 1];
            acc = okm[r - 1] ? dn : acc;
        }
        return acc;
    }
}

// ---------------------------------------------------------------------------
// Update masks (deepgpr_cpu.c:697-699, 913-915)
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool e_updated(const Grid& g, int comp, const int c[3])
{
    int a1 = comp == 2 ? 0 : comp + 1;
    int a2 = comp == 0 ? 2 : comp - 1;
    if (g.N[a1] - 1 == 1 && g.N[a2] - 1 == 1) return false;
    if (c[comp] >= g.N[comp] - 1) return false;
    if (c[a1] <= 0 || c[a1] >= g.N[a1] - 1) return false;
    if (c[a2] <= 0 || c[a2] >= g.N[a2] - 1) return false;
    return true;
}

__device__ __forceinline__ bool h_updated(const Grid& g, int comp, const int c[3])
{
    int a1 = comp == 2 ? 0 : comp + 1;
    int a2 = comp == 0 ? 2 : comp - 1;
    if (g.N[comp] - 1 == 1) return false;
    if (c[comp] <= 0 || c[comp] >= g.N[comp] - 1) return false;
    if (c[a1] >= g.N[a1] - 1 || c[a2] >= g.N[a2] - 1) return false;
    return true;
}

// A cell at least R away from every face: no difference is clamped there and
// every E and H component is updated (an axis with N = 2 has no such cell).
__device__ __forceinline__ bool plain_cell(const Grid& g, int R, const int c[3])
{
    return c[0] >= R && c[0] <= g.N[0] - 1 - R && c[1] >= R && c[1] <= g.N[1] - 1 - R && c[2] >= R &&
           c[2] <= g.N[2] - 1 - R;
}

// Curl coefficient on axis a: c1 * dx / d_a unless d_a == dx (deepgpr_cpu.c:693-694).
__device__ __forceinline__ float axis_coeff(const Grid& g, float c1, int a)
{
    if (a == 0 || g.d[a] == g.d[0]) return c1;
    return c1 * g.d[0] / g.d[a];
}

// ---------------------------------------------------------------------------
// CPML face tables (deepgpr_cpu.c:722-1100; pml.py:310-348)
// ---------------------------------------------------------------------------
// For a face on axis a the two auxiliary arrays couple target component
// TARGET[a][w] with the derivative along a of source component SOURCE[a][w];
// SIGN is the sign of that term in the curl (E) or negated curl (H).

__device__ __forceinline__ int face_target(int axis, int w) { return axis == 0 ? 1 + w : (w == 0 ? 0 : (axis == 1 ? 2 : 1)); }
__device__ __forceinline__ int face_source(int axis, int w) { return axis == 0 ? 2 - w : (axis == 1 ? (w == 0 ? 2 : 0) : (w == 0 ? 1 : 0)); }
__device__ __forceinline__ float e_face_sign(int axis, int w) { return (axis == 1) ? (w == 0 ? 1.0f : -1.0f) : (w == 0 ? -1.0f : 1.0f); }
__device__ __forceinline__ float h_face_sign(int axis, int w) { return -e_face_sign(axis, w); }

// Shape of one auxiliary array: face axis holds p+1 (E) or p (H); the E
// target axis / H source axis holds N-1; the remaining axis holds N.
__device__ __forceinline__ void psi_dims(const Grid& g, bool electric, int face, int w, int dims[3])
{
    int axis = face >> 1, p = g.pml[face];
    dims[0] = g.N[0];
    dims[1] = g.N[1];
    dims[2] = g.N[2];
    if (electric) {
        dims[face_target(axis, w)] -= 1;
        dims[axis] = p + 1;
    } else {
        dims[face_source(axis, w)] -= 1;
        dims[axis] = p;
    }
}

// Whether cell c has an element in auxiliary array w of face f (inside the
// face region along the face axis, inside the array's shape across it).
__device__ __forceinline__ bool psi_active(const Grid& g, bool electric, int face, int w, const int c[3])
{
    int dims[3];
    psi_dims(g, electric, face, w, dims);
    const int axis = face >> 1;
    const int a1 = axis == 2 ? 0 : axis + 1, a2 = axis == 0 ? 2 : axis - 1;
    return c[a1] < dims[a1] && c[a2] < dims[a2];
}

// The six staggered differences of a curl, (component, axis) per slot:
// z/y, y/z (first component), x/z, z/x (second), y/x, x/y (third).
constexpr int kDiffComp[6] = {2, 1, 0, 2, 1, 0};
constexpr int kDiffAxis[6] = {1, 2, 2, 0, 0, 1};
__device__ __forceinline__ int diff_slot(int comp, int axis)
{
    return axis == 0 ? (comp == 2 ? 3 : 4) : (axis == 1 ? (comp == 2 ? 0 : 5) : (comp == 1 ? 1 : 2));
}

// Profile index of coordinate c on a face, or -1 outside the face region.
__device__ __forceinline__ int face_q(const Grid& g, bool electric, int face, int c)
{
    int axis = face >> 1, p = g.pml[face], n = g.N[axis];
    if (p <= 0) return -1;
    if ((face & 1) == 0) {
        if (electric) return (c >= 1 && c <= p) ? p - c : -1;
        return (c >= 0 && c <= p - 1) ? p - 1 - c : -1;
    }
    if (c >= n - 1 - p && c <= n - 2) {
        if (electric && c <= 0) return -1;
        return c - (n - 1 - p);
    }
    return -1;
}

__device__ __forceinline__ long long psi_index(const int dims[3], int s, const int c[3], int axis, int q)
{
    int c0 = axis == 0 ? q : c[0];
    int c1 = axis == 1 ? q : c[1];
    int c2 = axis == 2 ? q : c[2];
    return (((long long)s * dims[0] + c0) * dims[1] + c1) * dims[2] + c2;
}

__device__ __forceinline__ bool in_any_face(const Grid& g, bool electric, const int c[3])
{
#pragma unroll
    for (int f = 0; f < 6; ++f)
        if (face_q(g, electric, f, c[f >> 1]) >= 0) return true;
    return false;
}

// ---------------------------------------------------------------------------
// History codecs
// ---------------------------------------------------------------------------

// binary32 -> binary16, round to nearest even; NaN -> quiet NaN keeping the
// sign (the convention of deepgpr_cpu.c:100-141). Written from the IEEE 754
// definition; checked against the hardware conversion by the codec tests.
__device__ __forceinline__ uint16_t f32_to_f16_bits(float value)
{
    uint32_t u = __float_as_uint(value);
    uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t mag = u & 0x7fffffffu;
    if (mag > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);
    if (mag >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
    int e16 = (int)(mag >> 23) - 112;
    uint32_t frac = mag & 0x7fffffu;
    int shift = 13;
    if (e16 <= 0) {
        if (e16 < -10) return (uint16_t)sign;
        frac |= 0x800000u;
        shift = 14 - e16;
    }
    uint32_t keep = frac >> shift;
    uint32_t drop = frac & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1);
    if (drop > half || (drop == half && (keep & 1u))) keep += 1u;
    if (e16 <= 0) return (uint16_t)(sign | keep);
    return (uint16_t)(sign + (((uint32_t)e16 << 10) + keep));
}

__device__ __forceinline__ float f16_bits_to_f32(uint16_t h)
{
    // the hardware conversion is exact for every binary16 value
    _Float16 v;
    __builtin_memcpy(&v, &h, 2);
    return (float)v;
}

__device__ __forceinline__ uint16_t f32_to_bf16_bits(float value)
{
    uint32_t u = __float_as_uint(value);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf16_bits_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ void store_hist(void* base, long long idx, float v, int kind)
{
    if (kind == kF16) ((uint16_t*)base)[idx] = f32_to_f16_bits(v);
    else if (kind == kBF16) ((uint16_t*)base)[idx] = f32_to_bf16_bits(v);
    else ((float*)base)[idx] = v;
}

__device__ __forceinline__ float load_hist(const void* base, long long idx, int kind)
{
    if (kind == kF16) return f16_bits_to_f32(((const uint16_t*)base)[idx]);
    if (kind == kBF16) return bf16_bits_to_f32(((const uint16_t*)base)[idx]);
    return ((const float*)base)[idx];
}

// INT8 packed history: payload (comp, frame, shot, box) then FP32 scales
// (comp, frame, shot, tiles) at a 4-byte aligned offset (storage.py:377-414).
struct Int8View {
    const signed char* q;
    const float* scale;
    long long q_cstride;      // elements between components
    long long s_cstride;      // scales between components
    int bx, by, bz;
    int nby, nbz;
    long long tiles;          // tiles per shot
};

// Tile of box cell (x, y, z) (box-relative coordinates).
__device__ __forceinline__ long long int8_tile(const Int8View& v, int x, int y, int z)
{
    return ((long long)(x / v.bx) * v.nby + y / v.by) * v.nbz + z / v.bz;
}

// Value of box cell h, whose tile the caller has formed once (int8_tile).
__device__ __forceinline__ float int8_load_at(const Int8View& v, const Box& b, int comp, int shot, long long h,
                                              long long tile)
{
    float sc = v.scale[comp * v.s_cstride + (long long)shot * v.tiles + tile];
    return sc * (float)v.q[comp * v.q_cstride + (long long)shot * b.cells + h];
}

__device__ __forceinline__ float int8_load(const Int8View& v, const Box& b, int comp, int shot, long long h)
{
    int x = (int)(h / ((long long)b.ny * b.nz));
    int y = (int)((h / b.nz) % b.ny);
    int z = (int)(h % b.nz);
    return int8_load_at(v, b, comp, shot, h, int8_tile(v, x, y, z));
}

}  // namespace dgx
