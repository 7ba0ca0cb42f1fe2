// HIP kernels of the FDTD hot path for MI355X (gfx950).
//
// General path: one thread per Yee cell, one kernel per half step, CPML fused
// into the curl sweep, source injection and history encoding fused into the
// electric sweep. The adjoint is a deterministic gather: every target cell
// re-evaluates the transposed stencil weights of its neighbours, so no
// atomics are used anywhere and results do not depend on scheduling.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "dgx_device.h"
#include "dgx_internal.h"

namespace dgx {

// ---------------------------------------------------------------------------
// Update coefficients (deepgpr_cpu.c:547-584)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_coefficients(
    Grid g, const float* __restrict__ eps, const float* __restrict__ sig, const float* __restrict__ mu,
    float* ca, float* ccurl, float* cb, float* hh, float* hcurl, float* hrhs, float dt)
{
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.cells) return;
    int i = (int)(n / g.st0);
    int rem = (int)(n - (long long)i * g.st0);
    int j = rem / g.N[2];
    int k = rem - j * g.N[2];
    if (!(i < g.N[0] - 1 && j < g.N[1] - 1 && k < g.N[2] - 1)) return;
    float ha = kMu0 * mu[n] / dt;
    hh[n] = 1.0f;
    hcurl[n] = (1.0f / g.d[0]) / ha;
    hrhs[n] = 1.0f / ha;
    if (sig[n] > kPecSigma) {
        ca[n] = 0.0f;
        ccurl[n] = 0.0f;
        cb[n] = 0.0f;
    } else {
        float et = kEps0 * eps[n] / dt;
        float st = 0.5f * sig[n];
        float ea = et + st;
        float eb = et - st;
        ca[n] = eb / ea;
        ccurl[n] = (1.0f / g.d[0]) / ea;
        cb[n] = 1.0f / ea;
    }
}

// ---------------------------------------------------------------------------
// Forward magnetic half step: curl update + H-CPML (deepgpr_cpu.c:886-1100)
// ---------------------------------------------------------------------------
// A plain cell (every component updated, no clamped difference, no CPML
// term): the curl update alone. m: cell in the model, n: cell of the shot.
template <int ORDER>
__device__ __forceinline__ void fwd_h_plain(const Grid& g, const float* __restrict__ hcurl, const CF3& E,
                                            const F3& H, long long m, long long n)
{
    const long long sx = g.st0, sy = g.N[2];
    float h0 = H.c[0][n], h1 = H.c[1][n], h2 = H.c[2][n];
    const float u1 = hcurl[m];
    const float uy = axis_coeff(g, u1, 1), uz = axis_coeff(g, u1, 2);
    {
        const float dEz = dfwd_full<ORDER>(E.c[2], n, sy);
        const float dEy = dfwd_full<ORDER>(E.c[1], n, 1);
        h0 = h0 - uy * dEz + uz * dEy;
    }
    {
        const float dEx = dfwd_full<ORDER>(E.c[0], n, 1);
        const float dEz = dfwd_full<ORDER>(E.c[2], n, sx);
        h1 = h1 - uz * dEx + u1 * dEz;
    }
    {
        const float dEy = dfwd_full<ORDER>(E.c[1], n, sx);
        const float dEx = dfwd_full<ORDER>(E.c[0], n, sy);
        h2 = h2 - u1 * dEy + uy * dEx;
    }
    H.c[0][n] = h0;
    H.c[1][n] = h1;
    H.c[2][n] = h2;
}

// A shell cell of the magnetic half step: CPML terms, clamped differences
// and update masks as the cell needs them.
template <int ORDER>
__device__ __forceinline__ void fwd_h_cell(const Grid& g, const float* __restrict__ hcurl,
                                           const float* __restrict__ hrhs, const CF3& E, const F3& H, const Cpml& cp,
                                           int has_cpml, int i, int j, int k, int s)
{
    const int c[3] = {i, j, k};
    const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
    const long long n = (long long)s * g.cells + m;
    const long long st[3] = {g.st0, (long long)g.N[2], 1};
    const bool pml = has_cpml && in_any_face(g, false, c);
    if (!pml && plain_cell(g, Radius<ORDER>::value, c)) {
        fwd_h_plain<ORDER>(g, hcurl, E, H, m, n);
        return;
    }
    const bool up0 = h_updated(g, 0, c), up1 = h_updated(g, 1, c), up2 = h_updated(g, 2, c);
    if (!(up0 || up1 || up2 || pml)) return;

    float h[3] = {H.c[0][n], H.c[1][n], H.c[2][n]};
    const float u1 = hcurl[m];
    const float uy = axis_coeff(g, u1, 1), uz = axis_coeff(g, u1, 2);
    // Each staggered difference is evaluated once and shared by the curl
    // update and the CPML terms that use it (the same call either way).
    bool need[6] = {up0, up0, up1, up1, up2, up2};
    if (pml) {
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int axis = f >> 1;
            if (face_q(g, false, f, c[axis]) < 0) continue;
#pragma unroll
            for (int w = 0; w < 2; ++w)
                if (psi_active(g, false, f, w, c)) need[diff_slot(face_source(axis, w), axis)] = true;
        }
    }
    float dv[6];
    {
        // every tap of the six differences is loaded before any arithmetic
        // (a difference the cell does not need reads the cell itself and is
        // not used)
        FwdTaps<ORDER> tp[6];
#pragma unroll
        for (int d = 0; d < 6; ++d)
            dfwd_taps<ORDER>(tp[d], E.c[kDiffComp[d]], n, st[kDiffAxis[d]], c[kDiffAxis[d]], g.N[kDiffAxis[d]],
                             need[d]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int d = 0; d < 6; ++d) dv[d] = dfwd_comb<ORDER>(tp[d]);
    }
    // selects rather than branches, so that every load above is issued at once
    const float n0 = h[0] - uy * dv[0] + uz * dv[1];  // dEz/dy, dEy/dz
    const float n1 = h[1] - uz * dv[2] + u1 * dv[3];  // dEx/dz, dEz/dx
    const float n2 = h[2] - u1 * dv[4] + uy * dv[5];  // dEy/dx, dEx/dy
    h[0] = up0 ? n0 : h[0];
    h[1] = up1 ? n1 : h[1];
    h[2] = up2 ? n2 : h[2];
    if (pml) {
        const float upd = hrhs[m];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int axis = f >> 1;
            const int q = face_q(g, false, f, c[axis]);
            if (q < 0) continue;
            const float* R = cp.coef[f];
            const int pp = g.pml[f];
            const float ra1 = R[q] - 1.0f, rb = R[pp + q], re = R[2 * pp + q], rf = R[3 * pp + q];
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                if (!psi_active(g, false, f, w, c)) continue;
                int dims[3];
                psi_dims(g, false, f, w, dims);
                const int t = face_target(axis, w), sc = face_source(axis, w);
                const float der = dv[diff_slot(sc, axis)] / g.d[axis];
                float* psi = cp.psi[f][w];
                const long long pi = psi_index(dims, s, c, axis, q);
                const float phi = psi[pi];
                const float corr = upd * (ra1 * der + rb * phi);
                if (h_face_sign(axis, w) < 0.0f) h[t] -= corr;
                else h[t] += corr;
                psi[pi] = re * phi - rf * der;
            }
        }
    }
    if (up0 || pml) H.c[0][n] = h[0];
    if (up1 || pml) H.c[1][n] = h[1];
    if (up2 || pml) H.c[2][n] = h[2];
}

template <int ORDER>
__global__ __launch_bounds__(256, ORDER == 8 ? 5 : 8) void k_fwd_h(Grid g, const float* __restrict__ hcurl,
                                               const float* __restrict__ hrhs, CF3 E, F3 H, Cpml cp,
                                               int has_cpml, CellMap cm)
{
    const int s = blockIdx.y;
    int i, j, k;
    int idx;
    const int part = sweep_part(cm, blockIdx.x, idx);
    if (part < 0) return;
    if (part == 1) {
        // interior box: every cell plain
        if (!box_cell(cm, idx, threadIdx.x, i, j, k)) return;
        const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
        fwd_h_plain<ORDER>(g, hcurl, E, H, m, (long long)s * g.cells + m);
        return;
    }
    if (!shell_cell(cm, idx, threadIdx.x, i, j, k)) return;
    fwd_h_cell<ORDER>(g, hcurl, hrhs, E, H, cp, has_cpml, i, j, k, s);
}

// ---------------------------------------------------------------------------
// Forward electric half step: curl update + E-CPML + sources + history
// (deepgpr_cpu.c:670-883, 644-668, 1559-1633)
// ---------------------------------------------------------------------------
// The curl update of a plain cell: e holds E^n and receives E^(n+1).
template <int ORDER>
__device__ __forceinline__ void fwd_e_plain(const Grid& g, const CF3& Hf, long long n, float u0, float u1,
                                            float uy, float uz, float (&e)[3])
{
    const long long sx = g.st0, sy = g.N[2];
    {
        const float dHz = dbwd_full<ORDER>(Hf.c[2], n, sy);
        const float dHy = dbwd_full<ORDER>(Hf.c[1], n, 1);
        e[0] = u0 * e[0] + uy * dHz - uz * dHy;
    }
    {
        const float dHx = dbwd_full<ORDER>(Hf.c[0], n, 1);
        const float dHz = dbwd_full<ORDER>(Hf.c[2], n, sx);
        e[1] = u0 * e[1] + uz * dHx - u1 * dHz;
    }
    {
        const float dHy = dbwd_full<ORDER>(Hf.c[1], n, sx);
        const float dHx = dbwd_full<ORDER>(Hf.c[0], n, sy);
        e[2] = u0 * e[2] + u1 * dHy - uy * dHx;
    }
}

// After the curl (and CPML) update of a cell: the Hertzian dipole sources of
// its block in source order, the history (E^n before the update and the
// effective right-hand side R^n), then the stores of the changed components.
__device__ __forceinline__ void fwd_e_finish(const SourceBlocks& src, const HistWrite& hist, int sb, int se,
                                             bool in_box, int i, int j, int k, int s, long long m, long long n,
                                             float (&e)[3], const float (&eold)[3], const bool (&up)[3], float u0,
                                             float cbv, bool touched, const F3& E)
{
    for (int q = sb; q < se; ++q) {
        const int2 ent = src.entries[q];
        if ((long long)ent.x == m) {
            const float scale = src.wave[(long long)ent.y * src.nt + src.it] * src.dl / src.vol;
            e[src.comp] -= cbv * scale;
        }
    }
    if (in_box) {
        const long long h = ((long long)(i - hist.box.x0) * hist.box.ny + (j - hist.box.y0)) * hist.box.nz +
                            (k - hist.box.z0);
        const long long base = (long long)s * hist.box.cells + h;
        const int comps = hist.mode == 3 ? 3 : 1;
        for (int ci = 0; ci < comps; ++ci) {
            const int comp = hist.mode == 3 ? ci : 2;
            if (hist.save_e) store_hist(hist.e_base, ci * hist.e_cstride + base, eold[comp], hist.e_kind);
            if (hist.save_r) {
                const float a = up[comp] ? u0 : 1.0f;
                const float rhs = cbv != 0.0f ? (e[comp] - a * eold[comp]) / cbv : 0.0f;
                store_hist(hist.r_base, ci * hist.r_cstride + base, rhs, hist.r_kind);
            }
        }
    }
    if (up[0] || touched) E.c[0][n] = e[0];
    if (up[1] || touched) E.c[1][n] = e[1];
    if (up[2] || touched) E.c[2][n] = e[2];
}

__device__ __forceinline__ bool in_hist_box(const HistWrite& hist, int i, int j, int k)
{
    return hist.active && i >= hist.box.x0 && i < hist.box.x0 + hist.box.nx && j >= hist.box.y0 &&
           j < hist.box.y0 + hist.box.ny && k >= hist.box.z0 && k < hist.box.z0 + hist.box.nz;
}

// A shell cell of the electric half step; sb..se are its block's source
// entries.
template <int ORDER>
__device__ __forceinline__ void fwd_e_cell(const Grid& g, const float* __restrict__ ca,
                                           const float* __restrict__ ccurl, const float* __restrict__ cb,
                                           const CF3& Hf, const F3& E, const Cpml& cp, int has_cpml,
                                           const SourceBlocks& src, const HistWrite& hist, int sb, int se, int i,
                                           int j, int k, int s)
{
    const int c[3] = {i, j, k};
    const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
    const long long n = (long long)s * g.cells + m;
    const long long st[3] = {g.st0, (long long)g.N[2], 1};
    const bool up[3] = {e_updated(g, 0, c), e_updated(g, 1, c), e_updated(g, 2, c)};
    const bool pml = has_cpml && in_any_face(g, true, c);
    const bool in_box = in_hist_box(hist, i, j, k);
    if (!(up[0] || up[1] || up[2] || pml || se > sb || in_box)) return;

    float e[3] = {E.c[0][n], E.c[1][n], E.c[2][n]};
    const float eold[3] = {e[0], e[1], e[2]};
    const float u0 = ca[m], u1 = ccurl[m];
    const float uy = axis_coeff(g, u1, 1), uz = axis_coeff(g, u1, 2);
    const float cbv = cb[m];
    if (!pml && plain_cell(g, Radius<ORDER>::value, c)) {
        fwd_e_plain<ORDER>(g, Hf, n, u0, u1, uy, uz, e);
    } else {
        // each staggered difference once, shared by the curl update and the CPML terms
        bool need[6] = {up[0], up[0], up[1], up[1], up[2], up[2]};
        if (pml) {
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                const int axis = f >> 1;
                if (face_q(g, true, f, c[axis]) < 0) continue;
#pragma unroll
                for (int w = 0; w < 2; ++w)
                    if (psi_active(g, true, f, w, c)) need[diff_slot(face_source(axis, w), axis)] = true;
            }
        }
        float dv[6];
        {
            // every tap of the six differences is loaded before any arithmetic
            // (a difference the cell does not need reads the cell itself and
            // is not used)
            FwdTaps<ORDER> tp[6];
#pragma unroll
            for (int d = 0; d < 6; ++d)
                dbwd_taps<ORDER>(tp[d], Hf.c[kDiffComp[d]], n, st[kDiffAxis[d]], c[kDiffAxis[d]], g.N[kDiffAxis[d]],
                                 need[d]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int d = 0; d < 6; ++d) dv[d] = dfwd_comb<ORDER>(tp[d]);
        }
        // selects rather than branches, so that every load above is issued at once
        const float n0 = u0 * e[0] + uy * dv[0] - uz * dv[1];  // dHz/dy, dHy/dz
        const float n1 = u0 * e[1] + uz * dv[2] - u1 * dv[3];  // dHx/dz, dHz/dx
        const float n2 = u0 * e[2] + u1 * dv[4] - uy * dv[5];  // dHy/dx, dHx/dy
        e[0] = up[0] ? n0 : e[0];
        e[1] = up[1] ? n1 : e[1];
        e[2] = up[2] ? n2 : e[2];
        if (pml) {
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                const int axis = f >> 1;
                const int q = face_q(g, true, f, c[axis]);
                if (q < 0) continue;
                const float* R = cp.coef[f];
                const int pp = g.pml[f];
                const float ra1 = R[q] - 1.0f, rb = R[pp + q], re = R[2 * pp + q], rf = R[3 * pp + q];
#pragma unroll
                for (int w = 0; w < 2; ++w) {
                    if (!psi_active(g, true, f, w, c)) continue;
                    int dims[3];
                    psi_dims(g, true, f, w, dims);
                    const int t = face_target(axis, w), sc = face_source(axis, w);
                    const float der = dv[diff_slot(sc, axis)] / g.d[axis];
                    float* psi = cp.psi[f][w];
                    const long long pi = psi_index(dims, s, c, axis, q);
                    const float phi = psi[pi];
                    const float corr = cbv * (ra1 * der + rb * phi);
                    if (e_face_sign(axis, w) < 0.0f) e[t] -= corr;
                    else e[t] += corr;
                    psi[pi] = re * phi - rf * der;
                }
            }
        }
    }
    fwd_e_finish(src, hist, sb, se, in_box, i, j, k, s, m, n, e, eold, up, u0, cbv, pml || se > sb, E);
}

template <int ORDER>
__global__ __launch_bounds__(256, ORDER == 8 ? 5 : 8) void k_fwd_e(Grid g, const float* __restrict__ ca,
                                               const float* __restrict__ ccurl,
                                               const float* __restrict__ cb, CF3 Hf, F3 E, Cpml cp,
                                               int has_cpml, SourceBlocks src, HistWrite hist, CellMap cm)
{
    const int s = blockIdx.y;
    // block-local source entries (sorted by block, then source order)
    const int blin = blockIdx.x + gridDim.x * blockIdx.y;
    int i, j, k;
    int idx;
    const int part = sweep_part(cm, blockIdx.x, idx);
    if (part < 0) return;
    if (part == 1) {
        // interior box: every cell plain and every component updated
        if (!box_cell(cm, idx, threadIdx.x, i, j, k)) return;
        const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
        const long long n = (long long)s * g.cells + m;
        int sb = 0, se = 0;
        if (src.count > 0) {
            sb = src.block_off[blin];
            se = src.block_off[blin + 1];
        }
        const bool up[3] = {true, true, true};
        float e[3] = {E.c[0][n], E.c[1][n], E.c[2][n]};
        const float eold[3] = {e[0], e[1], e[2]};
        const float u0 = ca[m], u1 = ccurl[m];
        const float uy = axis_coeff(g, u1, 1), uz = axis_coeff(g, u1, 2);
        const float cbv = cb[m];
        fwd_e_plain<ORDER>(g, Hf, n, u0, u1, uy, uz, e);
        fwd_e_finish(src, hist, sb, se, in_hist_box(hist, i, j, k), i, j, k, s, m, n, e, eold, up, u0, cbv,
                     se > sb, E);
        return;
    }
    if (!shell_cell(cm, idx, threadIdx.x, i, j, k)) return;
    int sb = 0, se = 0;
    if (src.count > 0) {
        sb = src.block_off[blin];
        se = src.block_off[blin + 1];
    }
    fwd_e_cell<ORDER>(g, ca, ccurl, cb, Hf, E, cp, has_cpml, src, hist, sb, se, i, j, k, s);
}

// Receiver gather: rx[s, it, r] = E_comp at the receiver (deepgpr_cpu.c:600-623)
__global__ __launch_bounds__(256) void k_sample(int nshot, const float* __restrict__ f, const int* __restrict__ loc,
                                                float* __restrict__ out, int nrx, int nt, int it, long long sst,
                                                long long xst, long long yst)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nshot * nrx) return;
    int s = idx / nrx, r = idx - s * nrx;
    const int* L = loc + (long long)idx * 3;
    long long m = (long long)L[0] * xst + (long long)L[1] * yst + L[2];
    out[((long long)s * nt + it) * nrx + r] = f[(long long)s * sst + m];
}

// Copy E components of the history box into one frame (final E^nt frame).
__global__ __launch_bounds__(256) void k_save_frame(Grid g, CF3 E, HistWrite hist)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)g.nshot * hist.box.cells) return;
    int s = (int)(idx / hist.box.cells);
    long long h = idx - (long long)s * hist.box.cells;
    long long x = h / ((long long)hist.box.ny * hist.box.nz) + hist.box.x0;
    long long y = (h / hist.box.nz) % hist.box.ny + hist.box.y0;
    long long z = h % hist.box.nz + hist.box.z0;
    long long n = (long long)s * g.cells + x * g.st0 + y * g.N[2] + z;
    int comps = hist.mode == 3 ? 3 : 1;
    for (int ci = 0; ci < comps; ++ci) {
        int comp = hist.mode == 3 ? ci : 2;
        store_hist(hist.e_base, ci * hist.e_cstride + idx, E.c[comp][n], hist.e_kind);
    }
}

// One workgroup per INT8 tile: max |v| over the tile, scale = max/127 (1 when
// max < 1e-30), q = rint(v/scale) clamped to +-127 (deepgpr.cu:2885-3030).
// Values come from the field arrays (box-mapped) or from a box-ordered FP32
// frame (comp stride src_cstride).
__global__ __launch_bounds__(256) void k_quant_int8(Grid g, CF3 E, const float* __restrict__ src,
                                                    long long src_cstride, Int8Write w)
{
    __shared__ float red[256];
    const int tile = blockIdx.x, s = blockIdx.y, ci = blockIdx.z;
    const int comp = w.mode == 3 ? ci : 2;
    const int tz = tile % w.nbz, ty = (tile / w.nbz) % w.nby, tx = tile / (w.nbz * w.nby);
    const int local = threadIdx.x;
    const int lx = local / (w.by * w.bz), ly = (local / w.bz) % w.by, lz = local % w.bz;
    const int x = tx * w.bx + lx, y = ty * w.by + ly, z = tz * w.bz + lz;
    const bool valid = x < w.box.nx && y < w.box.ny && z < w.box.nz;
    const long long h = valid ? ((long long)x * w.box.ny + y) * w.box.nz + z : 0;
    float v = 0.0f;
    if (valid) {
        if (src) {
            v = src[ci * src_cstride + (long long)s * w.box.cells + h];
        } else {
            long long n = (long long)s * g.cells + (long long)(x + w.box.x0) * g.st0 +
                          (long long)(y + w.box.y0) * g.N[2] + (z + w.box.z0);
            v = E.c[comp][n];
        }
    }
    red[local] = v < 0.0f ? -v : v;
    __syncthreads();
    for (int stride = blockDim.x >> 1; stride > 0; stride >>= 1) {
        if (local < stride && red[local + stride] > red[local]) red[local] = red[local + stride];
        __syncthreads();
    }
    const float peak = red[0];
    const float scale = peak >= 1.0e-30f ? peak / 127.0f : 1.0f;
    const long long snap = (long long)g.nshot * w.box.cells;
    if (local == 0) w.scale[((long long)ci * w.frames + w.frame) * g.nshot * w.tiles + (long long)s * w.tiles + tile] = scale;
    if (valid) {
        float r = rintf(v / scale);
        r = fminf(fmaxf(r, -127.0f), 127.0f);
        w.q[((long long)ci * w.frames + w.frame) * snap + (long long)s * w.box.cells + h] = (signed char)(int)r;
    }
}

// The same for 64-cell tiles (the default 4x4x4 and 8x8x1 blocks), one
// wavefront per tile and four tiles per workgroup. The maximum runs through
// wavefront shuffles in the pairing of the tree above (lane i takes lane
// i + stride when that value is larger, stride 32 down to 1), so the scale is
// the same value, NaN inputs included, with no barrier.
__global__ __launch_bounds__(256) void k_quant_int8_w64(Grid g, CF3 E, const float* __restrict__ src,
                                                        long long src_cstride, Int8Write w)
{
    const int local = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (int)(threadIdx.x >> 6), s = blockIdx.y, ci = blockIdx.z;
    if (tile >= w.tiles) return;  // whole wavefronts
    const int comp = w.mode == 3 ? ci : 2;
    // tile and in-tile coordinates by multiply-and-shift divisions
    const unsigned trow = div_q((unsigned)tile, w.div_nbz), tx = div_q((unsigned)tile, w.div_tplane);
    const int tz = tile - (int)(trow * (unsigned)w.nbz), ty = (int)(trow - tx * (unsigned)w.nby);
    const unsigned lrow = div_q((unsigned)local, w.div_bz), lx = div_q((unsigned)local, w.div_bplane);
    const int lz = local - (int)(lrow * (unsigned)w.bz), ly = (int)(lrow - lx * (unsigned)w.by);
    const int x = tx * w.bx + lx, y = ty * w.by + ly, z = tz * w.bz + lz;
    const bool valid = x < w.box.nx && y < w.box.ny && z < w.box.nz;
    const long long h = valid ? ((long long)x * w.box.ny + y) * w.box.nz + z : 0;
    float v = 0.0f;
    if (valid) {
        if (src) {
            v = src[ci * src_cstride + (long long)s * w.box.cells + h];
        } else {
            long long n = (long long)s * g.cells + (long long)(x + w.box.x0) * g.st0 +
                          (long long)(y + w.box.y0) * g.N[2] + (z + w.box.z0);
            v = E.c[comp][n];
        }
    }
    float m = v < 0.0f ? -v : v;
#pragma unroll
    for (int stride = 32; stride > 0; stride >>= 1) {
        const float o = __shfl_down(m, stride, 64);
        if (local < stride && o > m) m = o;
    }
    const float peak = __shfl(m, 0, 64);
    const float scale = peak >= 1.0e-30f ? peak / 127.0f : 1.0f;
    const long long snap = (long long)g.nshot * w.box.cells;
    if (local == 0) w.scale[((long long)ci * w.frames + w.frame) * g.nshot * w.tiles + (long long)s * w.tiles + tile] = scale;
    if (valid) {
        float r = rintf(v / scale);
        r = fminf(fmaxf(r, -127.0f), 127.0f);
        w.q[((long long)ci * w.frames + w.frame) * snap + (long long)s * w.box.cells + h] = (signed char)(int)r;
    }
}

// ---------------------------------------------------------------------------
// Adjoint: receivers^T and source^T (deepgpr_cpu.c:1479-1534)
// ---------------------------------------------------------------------------

// One thread per (shot, receiver cell): coincident receivers add in index order.
__global__ __launch_bounds__(256) void k_adj_receivers(Grid g, ReceiverGroups rg, float* __restrict__ lam,
                                                       const float* __restrict__ dgrad, int nrx, int nt, int it)
{
    int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= rg.ngroups) return;
    int s = rg.shot[gi];
    long long n = (long long)s * g.cells + rg.cell[gi];
    float acc = lam[n];
    for (int q = rg.off[gi]; q < rg.off[gi + 1]; ++q)
        acc += dgrad[((long long)s * nt + it) * nrx + rg.member[q]];
    lam[n] = acc;
}

// One thread per source: shots summed in order.
__global__ __launch_bounds__(64) void k_adj_source(Grid g, const int* __restrict__ loc, const float* __restrict__ cb,
                                                   const float* __restrict__ lam, float* __restrict__ gsrc,
                                                   int nsrc, int nt, int it, float geo)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsrc) return;
    float acc = gsrc[(long long)k * nt + it];
    for (int s = 0; s < g.nshot; ++s) {
        const int* L = loc + ((long long)s * nsrc + k) * 3;
        long long m = (long long)L[0] * g.st0 + (long long)L[1] * g.N[2] + L[2];
        acc += -cb[m] * geo * lam[(long long)s * g.cells + m];
    }
    gsrc[(long long)k * nt + it] = acc;
}

// ---------------------------------------------------------------------------
// Material gradient fused with the history decode (deepgpr_cpu.c:1659-1745)
// ---------------------------------------------------------------------------
template <typename ACC>
__global__ __launch_bounds__(256) void k_material_gradient(Grid g, CF3 L, MatGrad mg, ACC* __restrict__ geps,
                                                           ACC* __restrict__ gsig)
{
    long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const Box b = mg.box;
    if (h >= b.cells) return;
    int ix, iy, iz;
    if (b.cells <= 0x7fffffffLL) {
        // 32-bit multiply-and-shift divisions instead of three 64-bit ones
        const unsigned q = (unsigned)h;
        const unsigned row = div_q(q, mg.div_z);
        const unsigned px = div_q(row, mg.div_y);
        iz = (int)(q - row * (unsigned)b.nz) + b.z0;
        iy = (int)(row - px * (unsigned)b.ny) + b.y0;
        ix = (int)px + b.x0;
    } else {
        ix = (int)(h / ((long long)b.ny * b.nz)) + b.x0;
        iy = (int)((h / b.nz) % b.ny) + b.y0;
        iz = (int)(h % b.nz) + b.z0;
    }
    const int c[3] = {ix, iy, iz};
    const long long m = (long long)ix * g.st0 + (long long)iy * g.N[2] + iz;
    const long long out = ((long long)ix * (g.N[1] - 1) + iy) * (g.N[2] - 1) + iz;
    const float cav = mg.ca[m], cbv = mg.cb[m];
    const bool active = mg.sig[m] <= kPecSigma;
    const int comps = mg.mode == 3 ? 3 : 1;
    const float wgt = (float)mg.weight;
    // the INT8 tiles of this cell, formed once for every shot and component
    long long tile_e = 0, tile_s = 0;
    if (mg.kind == kInt8) {
        tile_e = int8_tile(mg.e8, ix - b.x0, iy - b.y0, iz - b.z0);
        tile_s = int8_tile(mg.s8, ix - b.x0, iy - b.y0, iz - b.z0);
    }
    ACC ge = 0, gs = 0;
    for (int s = 0; s < g.nshot; ++s) {
        const long long f = (long long)s * g.cells + m;
        const long long base = (long long)s * b.cells + h;
        for (int ci = 0; ci < comps; ++ci) {
            const int comp = mg.mode == 3 ? ci : 2;
            const bool up = e_updated(g, comp, c);
            float e_old, sec;
            if (mg.kind == kInt8) {
                e_old = int8_load_at(mg.e8, b, ci, s, h, tile_e);
                sec = int8_load_at(mg.s8, b, ci, s, h, tile_s);
            } else {
                e_old = load_hist(mg.eframe, ci * mg.e_cstride + base, mg.kind);
                sec = load_hist(mg.sframe, ci * mg.s_cstride + base, mg.kind);
            }
            float lam = L.c[comp][f];
            float rhs = sec;
            if (mg.e_only) {
                const float a = up ? cav : 1.0f;
                rhs = cbv != 0.0f ? (sec - a * e_old) / cbv : 0.0f;
            }
            if (e_old == 0.0f && rhs == 0.0f) lam = 0.0f;
            const float gca = up ? lam * e_old * wgt : 0.0f;
            const float gcb = lam * rhs * wgt;
            if (active) {
                if (mg.want_eps) {
                    const float dca = kEps0 * (1.0f - cav) * cbv / mg.dt;
                    const float dcb = -kEps0 * cbv * cbv / mg.dt;
                    ge += (ACC)(gca * dca + gcb * dcb);
                }
                if (mg.want_sig) {
                    const float dca = -0.5f * (1.0f + cav) * cbv;
                    const float dcb = -0.5f * cbv * cbv;
                    gs += (ACC)(gca * dca + gcb * dcb);
                }
            }
        }
    }
    if (mg.want_eps) geps[out] += ge;
    if (mg.want_sig) gsig[out] += gs;
}

__global__ __launch_bounds__(256) void k_add_f64_to_f32(const double* __restrict__ a, float* __restrict__ out,
                                                        long long count)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < count) out[idx] += (float)a[idx];
}

// ---------------------------------------------------------------------------
// Adjoint CPML weights: (E-CPML)^T or (H-CPML)^T of one half step
// (deepgpr_cpu.c:509-527, 1153-1239, 1332-1410)
// ---------------------------------------------------------------------------
// Two-pass adjoint: every cell evaluates its six transposed-stencil channel
// weights once (CPML weight, which also advances that cell's lambda psi, then
// the curl term), and the gathers read them.
// Channel (u, a) is stored at W + chan(u, a) * nshot * cells.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int chan_of(int u, int a) { return 2 * u + (a > u ? a - 1 : a); }

// Channel weights of cell (c, shot s) from its lambda values lt3 (curl
// coefficient cu; upd is the CPML source coefficient array): one store per
// channel, and the cell's lambda psi advanced.
// Channel weights of a cell with every component updated and no CPML term:
// the curl terms alone.
template <bool ELECTRIC>
__device__ __forceinline__ void adj_weights_plain(const Grid& g, float cu, const float (&lt3)[3],
                                                  float* __restrict__ W, long long n)
{
    const long long cstride = (long long)g.nshot * g.cells;
#pragma unroll
    for (int u = 0; u < 3; ++u) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (a == u) continue;
            const int t = 3 - u - a;
            const float kk = axis_coeff(g, cu, a);
            const int t1 = t == 2 ? 0 : t + 1;
            const bool plus = (a == t1) == ELECTRIC;
            W[chan_of(u, a) * cstride + n] = 0.0f + (plus ? kk : -kk) * lt3[t];
        }
    }
}

template <bool ELECTRIC>
__device__ __forceinline__ void adj_weights_cell(const Grid& g, float cu, const float* __restrict__ upd,
                                                 const float (&lt3)[3], const Cpml& cp, float* __restrict__ W,
                                                 bool cpml, const int (&c)[3], long long m, long long n, int s)
{
    const long long cstride = (long long)g.nshot * g.cells;
    if (!cpml && plain_cell(g, 1, c)) {
        adj_weights_plain<ELECTRIC>(g, cu, lt3, W, n);
        return;
    }
#pragma unroll
    for (int u = 0; u < 3; ++u) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (a == u) continue;
            const int t = 3 - u - a;
            float acc = 0.0f;
            if (cpml) {
                const int w = face_source(a, 0) == u ? 0 : 1;
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    const int f = 2 * a + side;
                    const int qq = face_q(g, ELECTRIC, f, c[a]);
                    if (qq < 0) continue;
                    int dims[3];
                    psi_dims(g, ELECTRIC, f, w, dims);
                    const int a1 = a == 2 ? 0 : a + 1, a2 = a == 0 ? 2 : a - 1;
                    if (c[a1] >= dims[a1] || c[a2] >= dims[a2]) continue;
                    const float* R = cp.coef[f];
                    const int pp = g.pml[f];
                    const float ra1 = R[qq] - 1.0f, rb = R[pp + qq], re = R[2 * pp + qq], rf = R[3 * pp + qq];
                    const float sign = ELECTRIC ? e_face_sign(a, w) : h_face_sign(a, w);
                    const float uc = upd[m];
                    const float inv = 1.0f / g.d[a];
                    float* psi = cp.psi[f][w];
                    const long long pi = psi_index(dims, s, c, a, qq);
                    const float lt = lt3[face_target(a, w)];
                    const float old = psi[pi];
                    acc += (sign * uc * ra1 * lt - rf * old) * inv;
                    psi[pi] = sign * uc * rb * lt + re * old;
                }
            }
            const bool up = ELECTRIC ? e_updated(g, t, c) : h_updated(g, t, c);
            if (up) {
                const float kk = axis_coeff(g, cu, a);
                const int t1 = t == 2 ? 0 : t + 1;
                const bool plus = (a == t1) == ELECTRIC;
                acc += (plus ? kk : -kk) * lt3[t];
            }
            W[chan_of(u, a) * cstride + n] = acc;
        }
    }
}

template <bool ELECTRIC>
__global__ __launch_bounds__(256, 8) void k_adj_weights(Grid g, const float* __restrict__ curl,
                                                     const float* __restrict__ upd, CF3 lam, Cpml cp,
                                                     float* __restrict__ W, int has_cpml, CellMap cm)
{
    const int s = blockIdx.y;
    int i, j, k;
    int idx;
    const int part = sweep_part(cm, blockIdx.x, idx);
    if (part < 0) return;
    if (part == 1) {
        if (!box_cell(cm, idx, threadIdx.x, i, j, k)) return;
        const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
        const long long n = (long long)s * g.cells + m;
        const float lt3[3] = {lam.c[0][n], lam.c[1][n], lam.c[2][n]};
        adj_weights_plain<ELECTRIC>(g, curl[m], lt3, W, n);
        return;
    }
    if (!shell_cell(cm, idx, threadIdx.x, i, j, k)) return;
    const int c[3] = {i, j, k};
    const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
    const long long n = (long long)s * g.cells + m;
    const bool cpml = has_cpml != 0 && in_any_face(g, ELECTRIC, c);
    const float lt3[3] = {lam.c[0][n], lam.c[1][n], lam.c[2][n]};
    adj_weights_cell<ELECTRIC>(g, curl[m], upd, lt3, cp, W, cpml, c, m, n, s);
}

// lambda_H[u] += D-^T W_E over the two axes a != u. MODE -2: a cell of the
// interior box (every gather full-width); -1: any cell.
template <int ORDER, int MODE>
__device__ __forceinline__ float gather_e_comp(const Grid& g, const float* __restrict__ W, long long n,
                                               const int (&c)[3], int u, float acc)
{
    const long long st[3] = {g.st0, (long long)g.N[2], 1};
    const long long cstride = (long long)g.nshot * g.cells;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == u) continue;
        const float* Wc = W + chan_of(u, a) * cstride + n;
        auto wf = [&](int off) { return Wc[off * st[a]]; };
        if (MODE == -2) acc += gather_bwd_full<ORDER>(wf);
        else acc += gather_bwd_sel<ORDER>(wf, c[a], g.N[a]);
    }
    return acc;
}

// lambda_E[t] (already scaled by ca where updated) += D+^T W_H over a != t
// (MODE as for gather_e_comp).
template <int ORDER, int MODE>
__device__ __forceinline__ float gather_h_comp(const Grid& g, const float* __restrict__ W, long long n,
                                               const int (&c)[3], int t, float acc)
{
    const long long st[3] = {g.st0, (long long)g.N[2], 1};
    const long long cstride = (long long)g.nshot * g.cells;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == t) continue;
        const float* Wc = W + chan_of(t, a) * cstride + n;
        auto wf = [&](int off) { return Wc[off * st[a]]; };
        if (MODE == -2) acc += gather_fwd_full<ORDER>(wf);
        else acc += gather_fwd_sel<ORDER>(wf, c[a], g.N[a]);
    }
    return acc;
}

// lambda_H += D-^T W_E (the E-side weights); lambda_E is scaled later.
template <int ORDER>
__global__ __launch_bounds__(256, ORDER == 8 ? 5 : 8) void k_adj_gather_e(Grid g, const float* __restrict__ W, F3 lamH, CellMap cm)
{
    const int s = blockIdx.y;
    int i, j, k;
    int idx;
    const int part = sweep_part(cm, blockIdx.x, idx);
    if (part < 0) return;
    const bool boxed = part == 1;
    if (boxed ? !box_cell(cm, idx, threadIdx.x, i, j, k) : !shell_cell(cm, idx, threadIdx.x, i, j, k)) return;
    const int c[3] = {i, j, k};
    const long long n = (long long)s * g.cells + (long long)i * g.st0 + (long long)j * g.N[2] + k;
#pragma unroll
    for (int u = 2; u >= 0; --u) {
        const float acc = boxed ? gather_e_comp<ORDER, -2>(g, W, n, c, u, lamH.c[u][n])
                                : gather_e_comp<ORDER, -1>(g, W, n, c, u, lamH.c[u][n]);
        lamH.c[u][n] = acc;
    }
}

// lambda_E = ca lambda_E (updated components) + D+^T W_H.
template <int ORDER>
__global__ __launch_bounds__(256, ORDER == 8 ? 5 : 8) void k_adj_gather_h(Grid g, const float* __restrict__ ca,
                                                      const float* __restrict__ W, F3 lamE, CellMap cm)
{
    const int s = blockIdx.y;
    int i, j, k;
    int idx;
    const int part = sweep_part(cm, blockIdx.x, idx);
    if (part < 0) return;
    if (part == 1) {
        if (!box_cell(cm, idx, threadIdx.x, i, j, k)) return;
        const int c[3] = {i, j, k};
        const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
        const long long n = (long long)s * g.cells + m;
        const float cav = ca[m];
#pragma unroll
        for (int t = 2; t >= 0; --t) lamE.c[t][n] = gather_h_comp<ORDER, -2>(g, W, n, c, t, cav * lamE.c[t][n]);
        return;
    }
    if (!shell_cell(cm, idx, threadIdx.x, i, j, k)) return;
    const int c[3] = {i, j, k};
    const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
    const long long n = (long long)s * g.cells + m;
    const float cav = ca[m];
#pragma unroll
    for (int t = 2; t >= 0; --t) {
        float acc = lamE.c[t][n];
        if (e_updated(g, t, c)) acc = cav * acc;
        lamE.c[t][n] = gather_h_comp<ORDER, -1>(g, W, n, c, t, acc);
    }
}

// A shell cell of k_adj_gather_e_w.
template <int ORDER>
__device__ __forceinline__ void gather_e_w_cell(const Grid& g, const float* __restrict__ We, const F3& lamH,
                                                const float* __restrict__ hcurl, const float* __restrict__ hrhs,
                                                const Cpml& cph, float* __restrict__ Wh, int has_cpml,
                                                const int (&c)[3], long long m, long long n, int s)
{
    float lt3[3];
#pragma unroll
    for (int u = 2; u >= 0; --u) {
        const float acc = gather_e_comp<ORDER, -1>(g, We, n, c, u, lamH.c[u][n]);
        lamH.c[u][n] = acc;
        lt3[u] = acc;
    }
    const bool cpml = has_cpml != 0 && in_any_face(g, false, c);
    adj_weights_cell<false>(g, hcurl[m], hrhs, lt3, cph, Wh, cpml, c, m, n, s);
}

// The E-side gather followed by the H-side weights of the same step, so that
// the new lambda_H is used while it is in registers: lambda_H += D-^T W_E,
// then W_H of the new lambda_H. W_E and W_H are separate arrays (every cell
// reads W_E at its neighbours and writes only its own W_H channels).
template <int ORDER>
__global__ __launch_bounds__(256, ORDER == 8 ? 5 : 8) void k_adj_gather_e_w(Grid g, const float* __restrict__ We, F3 lamH, CellMap cm,
                                                        const float* __restrict__ hcurl,
                                                        const float* __restrict__ hrhs, Cpml cph,
                                                        float* __restrict__ Wh, int has_cpml)
{
    const int s = blockIdx.y;
    int i, j, k;
    float lt3[3];
    int idx;
    const int part = sweep_part(cm, blockIdx.x, idx);
    if (part < 0) return;
    if (part == 1) {
        if (!box_cell(cm, idx, threadIdx.x, i, j, k)) return;
        const int c[3] = {i, j, k};
        const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
        const long long n = (long long)s * g.cells + m;
#pragma unroll
        for (int u = 2; u >= 0; --u) {
            const float acc = gather_e_comp<ORDER, -2>(g, We, n, c, u, lamH.c[u][n]);
            lamH.c[u][n] = acc;
            lt3[u] = acc;
        }
        adj_weights_plain<false>(g, hcurl[m], lt3, Wh, n);
        return;
    }
    if (!shell_cell(cm, idx, threadIdx.x, i, j, k)) return;
    const int c[3] = {i, j, k};
    const long long m = (long long)i * g.st0 + (long long)j * g.N[2] + k;
    const long long n = (long long)s * g.cells + m;
    gather_e_w_cell<ORDER>(g, We, lamH, hcurl, hrhs, cph, Wh, has_cpml, c, m, n, s);
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

// The interior box runs its own plain-cell path (same bits as the general
// path); DEEPGPR_AMD_3D_BOX=off sends every cell through the general path.
// Read at every launch, so one process can run both.
bool box_sweeps()
{
    const char* v = std::getenv("DEEPGPR_AMD_3D_BOX");
    return !(v && std::strcmp(v, "off") == 0);
}

// The shell and box blocks are interleaved in the launch order, so the
// latency-bound CPML blocks share the CUs with the bandwidth-bound box
// blocks; DEEPGPR_AMD_3D_INTERLEAVE=off launches every shell block first
// (same bits). Read at every launch.
bool interleave_sweeps()
{
    const char* v = std::getenv("DEEPGPR_AMD_3D_INTERLEAVE");
    return !(v && std::strcmp(v, "off") == 0);
}

CellSweep cell_sweep(const Grid& g, int order, bool use_box)
{
    CellSweep cs{};
    cs.cm = make_cell_map(g.N, g.pml, order / 2, use_box && box_sweeps());
    if (interleave_sweeps()) interleave(cs.cm);
    cs.grid = dim3((unsigned)cs.cm.per_shot, (unsigned)g.nshot, 1u);
    cs.blocks = (long long)cs.cm.per_shot * g.nshot;
    return cs;
}

long long cell_block(const CellSweep& cs, const Grid& g, int i, int j, int k, int s)
{
    const long long b = cell_map_block(cs.cm, i, j, k);
    return b < 0 ? -1 : b + (long long)cs.cm.per_shot * s;
}

void launch_coefficients(const Grid& g, const float* eps, const float* sig, const float* mu, float* const co[6],
                         float dt, hipStream_t st)
{
    unsigned blocks = (unsigned)((g.cells + 255) / 256);
    hipLaunchKernelGGL(k_coefficients, dim3(blocks), dim3(256), 0, st, g, eps, sig, mu, co[0], co[1], co[2],
                       co[3], co[4], co[5], dt);
}

void launch_fwd_h(int order, const Grid& g, const StepArgs& a, hipStream_t st)
{
    const CellSweep cs = cell_sweep(g, order);
    {
        const dim3 grid = cs.grid;
        const CellMap cm = cs.cm;
        if (order == 8)
            hipLaunchKernelGGL(k_fwd_h<8>, grid, dim3(256), 0, st, g, a.hcurl, a.hrhs, a.Ec, a.H, a.cph, a.has_cpml,
                               cm);
        else if (order == 4)
            hipLaunchKernelGGL(k_fwd_h<4>, grid, dim3(256), 0, st, g, a.hcurl, a.hrhs, a.Ec, a.H, a.cph, a.has_cpml,
                               cm);
        else
            hipLaunchKernelGGL(k_fwd_h<2>, grid, dim3(256), 0, st, g, a.hcurl, a.hrhs, a.Ec, a.H, a.cph, a.has_cpml,
                               cm);
    }
}

void launch_fwd_e(int order, const Grid& g, const StepArgs& a, hipStream_t st)
{
    const CellSweep cs = cell_sweep(g, order);
    {
        const dim3 grid = cs.grid;
        const CellMap cm = cs.cm;
        if (order == 8)
            hipLaunchKernelGGL(k_fwd_e<8>, grid, dim3(256), 0, st, g, a.ca, a.ccurl, a.cb, a.Hc, a.E, a.cpe,
                               a.has_cpml, a.src, a.hist, cm);
        else if (order == 4)
            hipLaunchKernelGGL(k_fwd_e<4>, grid, dim3(256), 0, st, g, a.ca, a.ccurl, a.cb, a.Hc, a.E, a.cpe,
                               a.has_cpml, a.src, a.hist, cm);
        else
            hipLaunchKernelGGL(k_fwd_e<2>, grid, dim3(256), 0, st, g, a.ca, a.ccurl, a.cb, a.Hc, a.E, a.cpe,
                               a.has_cpml, a.src, a.hist, cm);
    }
}

void launch_sample(const Grid& g, const float* f, const int* loc, float* out, int nrx, int nt, int it,
                   hipStream_t st, long long shot_stride, long long x_stride, long long y_stride)
{
    int total = g.nshot * nrx;
    if (total <= 0) return;
    if (shot_stride < 0) {
        shot_stride = g.cells;
        x_stride = g.st0;
        y_stride = g.N[2];
    }
    hipLaunchKernelGGL(k_sample, dim3((total + 255) / 256), dim3(256), 0, st, g.nshot, f, loc, out, nrx, nt, it,
                       shot_stride, x_stride, y_stride);
}

void launch_save_frame(const Grid& g, CF3 E, const HistWrite& hw, hipStream_t st)
{
    long long total = (long long)g.nshot * hw.box.cells;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_save_frame, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, g, E, hw);
}

void launch_quant_int8(const Grid& g, CF3 E, const float* src, long long src_cstride, const Int8Write& w,
                       hipStream_t st)
{
    int comps = w.mode == 3 ? 3 : 1;
    if (w.bx * w.by * w.bz == 64) {
        dim3 grid((unsigned)((w.tiles + 3) / 4), (unsigned)g.nshot, (unsigned)comps);
        hipLaunchKernelGGL(k_quant_int8_w64, grid, dim3(256), 0, st, g, E, src, src_cstride, w);
        return;
    }
    dim3 grid((unsigned)w.tiles, (unsigned)g.nshot, (unsigned)comps);
    hipLaunchKernelGGL(k_quant_int8, grid, dim3(w.bx * w.by * w.bz), 0, st, g, E, src, src_cstride, w);
}

void launch_adj_receivers(const Grid& g, const ReceiverGroups& rg, float* lam, const float* dgrad, int nrx,
                          int nt, int it, hipStream_t st)
{
    if (rg.ngroups <= 0) return;
    hipLaunchKernelGGL(k_adj_receivers, dim3((rg.ngroups + 255) / 256), dim3(256), 0, st, g, rg, lam, dgrad,
                       nrx, nt, it);
}

void launch_adj_source(const Grid& g, const int* loc, const float* cb, const float* lam, float* gsrc, int nsrc,
                       int nt, int it, float geo, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_source, dim3((nsrc + 63) / 64), dim3(64), 0, st, g, loc, cb, lam, gsrc, nsrc, nt,
                       it, geo);
}

void launch_material_gradient(const Grid& g, CF3 L, const MatGrad& mg, void* geps, void* gsig, bool fp64,
                              hipStream_t st)
{
    unsigned blocks = (unsigned)((mg.box.cells + 255) / 256);
    if (fp64)
        hipLaunchKernelGGL(k_material_gradient<double>, dim3(blocks), dim3(256), 0, st, g, L, mg, (double*)geps,
                           (double*)gsig);
    else
        hipLaunchKernelGGL(k_material_gradient<float>, dim3(blocks), dim3(256), 0, st, g, L, mg, (float*)geps,
                           (float*)gsig);
}

void launch_add_f64(const double* a, float* out, long long count, hipStream_t st)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_add_f64_to_f32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, a, out, count);
}

void launch_adj_weights(int order, bool electric, const Grid& g, const AdjArgs& a, float* W, hipStream_t st)
{
    const CellSweep cs = cell_sweep(g, order);
    {
        if (electric)
            hipLaunchKernelGGL(k_adj_weights<true>, cs.grid, dim3(256), 0, st, g, a.ccurl, a.cb, a.lEc, a.cpe, W,
                               a.has_cpml, cs.cm);
        else
            hipLaunchKernelGGL(k_adj_weights<false>, cs.grid, dim3(256), 0, st, g, a.hcurl, a.hrhs, a.lHc, a.cph,
                               W, a.has_cpml, cs.cm);
    }
}

void launch_adj_gather_e(int order, const Grid& g, const AdjArgs& a, const float* W, hipStream_t st)
{
    const CellSweep cs = cell_sweep(g, order);
    {
        const dim3 grid = cs.grid;
        if (order == 8) hipLaunchKernelGGL(k_adj_gather_e<8>, grid, dim3(256), 0, st, g, W, a.lH, cs.cm);
        else if (order == 4) hipLaunchKernelGGL(k_adj_gather_e<4>, grid, dim3(256), 0, st, g, W, a.lH, cs.cm);
        else hipLaunchKernelGGL(k_adj_gather_e<2>, grid, dim3(256), 0, st, g, W, a.lH, cs.cm);
    }
}

void launch_adj_gather_h(int order, const Grid& g, const AdjArgs& a, const float* W, hipStream_t st)
{
    const CellSweep cs = cell_sweep(g, order);
    {
        const dim3 grid = cs.grid;
        if (order == 8) hipLaunchKernelGGL(k_adj_gather_h<8>, grid, dim3(256), 0, st, g, a.ca, W, a.lE, cs.cm);
        else if (order == 4)
            hipLaunchKernelGGL(k_adj_gather_h<4>, grid, dim3(256), 0, st, g, a.ca, W, a.lE, cs.cm);
        else hipLaunchKernelGGL(k_adj_gather_h<2>, grid, dim3(256), 0, st, g, a.ca, W, a.lE, cs.cm);
    }
}

void launch_adj_gather_e_w(int order, const Grid& g, const AdjArgs& a, const float* We, float* Wh, hipStream_t st)
{
    const CellSweep cs = cell_sweep(g, order);
    {
        const dim3 grid = cs.grid;
        const CellMap cm = cs.cm;
        if (order == 8)
            hipLaunchKernelGGL(k_adj_gather_e_w<8>, grid, dim3(256), 0, st, g, We, a.lH, cm, a.hcurl, a.hrhs, a.cph,
                               Wh, a.has_cpml);
        else if (order == 4)
            hipLaunchKernelGGL(k_adj_gather_e_w<4>, grid, dim3(256), 0, st, g, We, a.lH, cm, a.hcurl, a.hrhs, a.cph,
                               Wh, a.has_cpml);
        else
            hipLaunchKernelGGL(k_adj_gather_e_w<2>, grid, dim3(256), 0, st, g, We, a.lH, cm, a.hcurl, a.hrhs, a.cph,
                               Wh, a.has_cpml);
    }
}

}  // namespace dgx
