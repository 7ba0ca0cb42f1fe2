// One kernel per 2D-TM time step, forward and adjoint.
//
// A workgroup owns a kTmBI x kTmBJ tile of one shot. It stages Ez (or
// lambda_Ez) on the tile plus a (2R-1)-cell halo in LDS, recomputes the
// magnetic half step (or the transposed electric weights) on the R-cell halo
// the electric stencil needs, and writes only the cells it owns. Fields and
// the active CPML auxiliaries ping-pong between two buffers, so a halo read
// never races another workgroup's write. Every value uses the expression of
// the split kernels in dgx_tm2d.hip (the same operations in the same order),
// so both agree bit for bit; those in turn follow the general path.
#include "dgx_device.h"
#include "dgx_internal.h"

namespace dgx {

namespace {

constexpr int BI = kTmBI;
constexpr int BJ = kTmBJ;
constexpr int NTF = DGX_TM_FWD_THREADS;  // threads per workgroup, forward
constexpr int NTA = DGX_TM_ADJ_THREADS;  // threads per workgroup, adjoint

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

template <int ORDER>
struct Tile {
    static constexpr int R = Radius<ORDER>::value;
    static constexpr int HE = 2 * R - 1;       // halo of the staged Ez / lambda_Ez
    static constexpr int LR = BI + 2 * HE;     // staged rows
    static constexpr int LC = BJ + 2 * HE;     // staged columns
    static constexpr int HR = BI + 2 * R - 1;  // rows of the Hy strip: R above the tile, R-1 below
    static constexpr int HC = BJ + 2 * R - 1;  // columns of the Hx strip
    static_assert(HE == tm_halo(ORDER), "halo width");
};

// Same XCD-aware placement as the split kernels: XCD x takes the x-th
// contiguous chunk of tiles.
__device__ __forceinline__ int step_tile(int tiles)
{
    const int chunk = gridDim.x >> 3;
    const int t = (blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
    return t < tiles ? t : -1;
}

__device__ __forceinline__ int face_q2(const Tm2dGrid& g, bool electric, int face, int c)
{
    const int p = g.pml[face];
    const int n = face < 2 ? g.NX : g.NY;
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

// x faces (s, q, j), y faces (s, i, q)
__device__ __forceinline__ long long psi_at(const Tm2dGrid& g, bool electric, int face, int s, int i, int j, int q)
{
    const int depth = g.pml[face] + (electric ? 1 : 0);
    if (face < 2) return ((long long)s * depth + q) * g.NY + j;
    return ((long long)s * g.NX + i) * depth + q;
}

__device__ __forceinline__ float axis_k(const Tm2dGrid& g, float c1, int a)
{
    if (a == 0 || g.d[a] == g.d[0]) return c1;
    return c1 * g.d[0] / g.d[a];
}

__device__ __forceinline__ bool ez_up(const Tm2dGrid& g, int i, int j)
{
    return ((g.NX - 1) != 1 || (g.NY - 1) != 1) && i > 0 && i < g.NX - 1 && j > 0 && j < g.NY - 1;
}
__device__ __forceinline__ bool hx_up(const Tm2dGrid& g, int i, int j)
{
    return (g.NX - 1) != 1 && i > 0 && i < g.NX - 1 && j < g.NY - 1;
}
__device__ __forceinline__ bool hy_up(const Tm2dGrid& g, int i, int j)
{
    return (g.NY - 1) != 1 && i < g.NX - 1 && j > 0 && j < g.NY - 1;
}

struct Cpml4 {
    float ra1, rb, re, rf;
};
__device__ __forceinline__ Cpml4 cpml_at(const float* R, int p, int q)
{
    return {R[q] - 1.0f, R[p + q], R[2 * p + q], R[3 * p + q]};
}

}  // namespace

// ---------------------------------------------------------------------------
// Forward: H half step (tile + halo), E half step, sources, history, receivers
//
// Wavefront w handles tile rows w, w + NW, ... of every phase; lane = column.
// Row indices, row pointers and x-face checks are therefore wave-uniform
// (scalar); per-lane column data (the halo column a lane also covers, y-face
// indices) is computed once.
// ---------------------------------------------------------------------------
template <int ORDER>
__device__ __forceinline__ void tm_fwd_tile(const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                                            const SourceBlocks& src, const TmReceiversFwd& rx, const HistWrite& hist,
                                            int it, int tile, int s)
{
    using T = Tile<ORDER>;
    constexpr int R = T::R, HE = T::HE, LR = T::LR, LC = T::LC, HR = T::HR, HC = T::HC;
    constexpr int NW = NTF / 64;
    constexpr int KS = cdiv(LR, NW), KY = cdiv(HR, NW), KT = cdiv(BI, NW);
    __shared__ float sE[LR * LC];
    __shared__ float sHy[HR * BJ];
    __shared__ float sHx[BI * HC];
    __shared__ float sEn[BI * BJ];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int ti = tile / g.ntj, tj = tile - ti * g.ntj;
    const int i0 = ti * BI, j0 = tj * BJ;
    const long long sb0 = (long long)s * g.plane;
    const float* __restrict__ Ez = pp.Ez[0] + sb0;
    const float* __restrict__ Hx = pp.Hx[0] + sb0;
    const float* __restrict__ Hy = pp.Hy[0] + sb0;

    // this lane's columns: the tile column gj, one staged halo column hc
    // (left: 0..HE-1, right: HE+BJ..LC-1) and one Hx-strip halo column xc
    const int gj = j0 + lane;
    const bool col_ok = gj < g.NY;
    const int hc = lane < HE ? lane : (lane < 2 * HE ? BJ + lane : -1);
    const int hgj = j0 - HE + hc;
    const bool h_ok = hc >= 0 && hgj >= 0 && hgj < g.NY;
    const int xc = lane < R ? lane : (lane < 2 * R - 1 ? BJ + lane : -1);
    const int xgj = j0 - R + xc;
    const bool x_ok = xc >= 0 && xgj >= 0 && xgj < g.NY;
    const bool col_in = gj > 0 && gj < g.NY - 1;
    const bool xcol_in = xgj > 0 && xgj < g.NY - 1;
    int qe[2], qh[2], qx[2];  // y-face depth indices of gj (E, H) and xgj (H)
#pragma unroll
    for (int f2 = 0; f2 < 2; ++f2) {
        qe[f2] = face_q2(g, true, 2 + f2, gj);
        qh[f2] = face_q2(g, false, 2 + f2, gj);
        qx[f2] = face_q2(g, false, 2 + f2, xgj);
    }

    // Every load that does not depend on LDS is issued before the first barrier.
    float st_m[KS], st_h[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int r = w + NW * k;
        const int gi = i0 - HE + r;
        const bool row = r < LR && gi >= 0 && gi < g.NX;
        st_m[k] = (row && col_ok) ? Ez[gi * g.NY + gj] : 0.0f;
        st_h[k] = (row && h_ok) ? Ez[gi * g.NY + hgj] : 0.0f;
    }
    float y_h[KY], y_u[KY];  // Hy and hcurl of the Hy strip
#pragma unroll
    for (int k = 0; k < KY; ++k) {
        const int r = w + NW * k;
        const int gi = i0 - R + r;
        const bool ok = r < HR && gi >= 0 && gi < g.NX && col_ok;
        y_h[k] = ok ? Hy[gi * g.NY + gj] : 0.0f;
        y_u[k] = ok ? f.hcurl[gi * g.NY + gj] : 0.0f;
    }
    float x_h[KT], x_u[KT], xe_h[KT], xe_u[KT], t_ca[KT], t_cc[KT], t_cb[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int r = w + NW * k;
        const int gi = i0 + r;
        const bool row = r < BI && gi < g.NX;
        const bool ok = row && col_ok, okx = row && x_ok;
        const int p = gi * g.NY + gj, px = gi * g.NY + xgj;
        x_h[k] = ok ? Hx[p] : 0.0f;
        x_u[k] = ok ? f.hcurl[p] : 0.0f;
        t_ca[k] = ok ? f.ca[p] : 0.0f;
        t_cc[k] = ok ? f.ccurl[p] : 0.0f;
        t_cb[k] = ok ? f.cb[p] : 0.0f;
        xe_h[k] = okx ? Hx[px] : 0.0f;
        xe_u[k] = okx ? f.hcurl[px] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int r = w + NW * k;
        if (r < LR) {
            sE[r * LC + HE + lane] = st_m[k];
            if (hc >= 0) sE[r * LC + hc] = st_h[k];
        }
    }
    __syncthreads();

    // Hy on rows [i0-R, i0+BI+R-1) of the tile's columns (x faces: H phi1)
#pragma unroll
    for (int k = 0; k < KY; ++k) {
        const int r = w + NW * k;
        const int gi = i0 - R + r;
        if (r >= HR || gi < 0 || gi >= g.NX || !col_ok) continue;
        const int p = gi * g.NY + gj;
        const bool own = gi >= i0 && gi < i0 + BI;
        const int le = (r + R - 1) * LC + HE + lane;  // staged index of (gi, gj)
        float hy = y_h[k];
        const float u1 = y_u[k];
        if ((g.NY - 1) != 1 && gi < g.NX - 1 && col_in) {
            const float uz = axis_k(g, u1, 2);
            const float dEz = dfwd<ORDER, int>(sE, le, LC, gi, g.NX);
            hy = hy - uz * 0.0f + u1 * dEz;
        }
#pragma unroll
        for (int face = 0; face < 2; ++face) {
            const int q = face_q2(g, false, face, gi);
            if (q < 0) continue;
            const Cpml4 kc = cpml_at(f.coef_h[face], g.pml[face], q);
            const float der = dfwd<ORDER, int>(sE, le, LC, gi, g.NX) / g.d[0];
            const long long pi = psi_at(g, false, face, s, gi, gj, q);
            const float phi = pp.psi_h[0][face][pi];
            hy += f.hrhs[p] * (kc.ra1 * der + kc.rb * phi);
            if (own) pp.psi_h[1][face][pi] = kc.re * phi - kc.rf * der;
        }
        sHy[r * BJ + lane] = hy;
        if (own) pp.Hy[1][sb0 + p] = hy;
    }
    // Hx on the tile's rows: the tile columns, then the strip's halo columns (y faces: H phi1)
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int r = w + NW * k;
        const int gi = i0 + r;
        if (r >= BI || gi >= g.NX) continue;
        const bool row_up = (g.NX - 1) != 1 && gi > 0 && gi < g.NX - 1;
        if (col_ok) {
            const int p = gi * g.NY + gj;
            const int le = (r + HE) * LC + HE + lane;
            float hx = x_h[k];
            const float u1 = x_u[k];
            if (row_up && gj < g.NY - 1) {
                const float uy = axis_k(g, u1, 1), uz = axis_k(g, u1, 2);
                const float dEz = dfwd<ORDER, int>(sE, le, 1, gj, g.NY);
                hx = hx - uy * dEz + uz * 0.0f;
            }
#pragma unroll
            for (int f2 = 0; f2 < 2; ++f2) {
                const int q = qh[f2];
                if (q < 0) continue;
                const int face = 2 + f2;
                const Cpml4 kc = cpml_at(f.coef_h[face], g.pml[face], q);
                const float der = dfwd<ORDER, int>(sE, le, 1, gj, g.NY) / g.d[1];
                const long long pi = psi_at(g, false, face, s, gi, gj, q);
                const float phi = pp.psi_h[0][face][pi];
                hx -= f.hrhs[p] * (kc.ra1 * der + kc.rb * phi);
                pp.psi_h[1][face][pi] = kc.re * phi - kc.rf * der;
            }
            sHx[r * HC + R + lane] = hx;
            pp.Hx[1][sb0 + p] = hx;
        }
        if (x_ok) {
            const int p = gi * g.NY + xgj;
            const int le = (r + HE) * LC + (xc - R + HE);
            float hx = xe_h[k];
            const float u1 = xe_u[k];
            if (row_up && xgj < g.NY - 1) {
                const float uy = axis_k(g, u1, 1), uz = axis_k(g, u1, 2);
                const float dEz = dfwd<ORDER, int>(sE, le, 1, xgj, g.NY);
                hx = hx - uy * dEz + uz * 0.0f;
            }
#pragma unroll
            for (int f2 = 0; f2 < 2; ++f2) {
                const int q = qx[f2];
                if (q < 0) continue;
                const int face = 2 + f2;
                const Cpml4 kc = cpml_at(f.coef_h[face], g.pml[face], q);
                const float der = dfwd<ORDER, int>(sE, le, 1, xgj, g.NY) / g.d[1];
                const float phi = pp.psi_h[0][face][psi_at(g, false, face, s, gi, xgj, q)];
                hx -= f.hrhs[p] * (kc.ra1 * der + kc.rb * phi);
            }
            sHx[r * HC + xc] = hx;
        }
    }
    __syncthreads();

    // Ez on the tile, sources, history
    int s_beg = 0, s_end = 0;
    if (src.count > 0) {
        s_beg = src.block_off[tile + g.ftiles * s];
        s_end = src.block_off[tile + g.ftiles * s + 1];
    }
    const bool ez_deg = (g.NX - 1) != 1 || (g.NY - 1) != 1;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int r = w + NW * k;
        const int gi = i0 + r;
        if (r >= BI || gi >= g.NX || !col_ok) continue;
        const int p = gi * g.NY + gj;
        const int e = r * BJ + lane;
        const float eold = sE[(r + HE) * LC + HE + lane];
        const int ly = (r + R) * BJ + lane;  // sHy index of (gi, gj)
        const int lx = r * HC + R + lane;    // sHx index of (gi, gj)
        const bool up = ez_deg && gi > 0 && gi < g.NX - 1 && col_in;
        const float u0 = t_ca[k];
        float ez = eold;
        if (up) {
            const float u1 = t_cc[k];
            const float uy = axis_k(g, u1, 1);
            const float dHy = dbwd<ORDER, int>(sHy, ly, BJ, gi, g.NX);
            const float dHx = dbwd<ORDER, int>(sHx, lx, 1, gj, g.NY);
            ez = u0 * ez + u1 * dHy - uy * dHx;
        }
        const float cbv = t_cb[k];
#pragma unroll
        for (int face = 0; face < 4; ++face) {
            const int q = face < 2 ? face_q2(g, true, face, gi) : qe[face - 2];
            if (q < 0) continue;
            const Cpml4 kc = cpml_at(f.coef_e[face], g.pml[face], q);
            const float der = face < 2 ? dbwd<ORDER, int>(sHy, ly, BJ, gi, g.NX) / g.d[0]
                                       : dbwd<ORDER, int>(sHx, lx, 1, gj, g.NY) / g.d[1];
            const long long pi = psi_at(g, true, face, s, gi, gj, q);
            const float phi = pp.psi_e[0][face][pi];
            const float corr = cbv * (kc.ra1 * der + kc.rb * phi);
            if (face < 2) ez += corr;
            else ez -= corr;
            pp.psi_e[1][face][pi] = kc.re * phi - kc.rf * der;
        }
        for (int q = s_beg; q < s_end; ++q) {
            const int2 ent = src.entries[q];
            if (ent.x == e) {
                const float scale = src.wave[(long long)ent.y * src.nt + it] * src.dl / src.vol;
                ez -= cbv * scale;
            }
        }
        if (hist.active && gi >= hist.box.x0 && gi < hist.box.x0 + hist.box.nx && gj >= hist.box.y0 &&
            gj < hist.box.y0 + hist.box.ny) {
            const long long h = (long long)(gi - hist.box.x0) * hist.box.ny + (gj - hist.box.y0);
            const long long base = (long long)s * hist.box.cells + h;
            const int comps = hist.mode == 3 ? 3 : 1;
            for (int ci = 0; ci < comps; ++ci) {
                const bool is_ez = ci == comps - 1;
                // Ex and Ey are identically zero on this path (and never updated: a = 1)
                const float eo = is_ez ? eold : 0.0f;
                const float en = is_ez ? ez : 0.0f;
                const float a = (is_ez && up) ? u0 : 1.0f;
                if (hist.save_e) store_hist(hist.e_base, ci * hist.e_cstride + base, eo, hist.e_kind);
                if (hist.save_r) {
                    const float rhs = cbv != 0.0f ? (en - a * eo) / cbv : 0.0f;
                    store_hist(hist.r_base, ci * hist.r_cstride + base, rhs, hist.r_kind);
                }
            }
        }
        pp.Ez[1][sb0 + p] = ez;
        sEn[e] = ez;
    }
    if (rx.count > 0) {
        __syncthreads();
        const int b0 = rx.off[tile + g.ftiles * s], b1 = rx.off[tile + g.ftiles * s + 1];
        for (int q = b0 + (int)threadIdx.x; q < b1; q += NTF) {
            const int2 ent = rx.ent[q];
            rx.out[((long long)s * rx.nt + it) * rx.nrx + ent.y] = sEn[ent.x];
        }
    }
}

template <int ORDER>
__global__ __launch_bounds__(NTF) void k_tm_step_fwd(Tm2dGrid g, Tm2dFields f, TmPingPong pp, SourceBlocks src,
                                                     TmReceiversFwd rx, HistWrite hist, int it)
{
    const int tile = step_tile(g.ftiles);
    if (tile < 0) return;
    tm_fwd_tile<ORDER>(g, f, pp, src, rx, hist, it, tile, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Adjoint: receivers^T, source and material terms, (E step)^T, (H step)^T
// ---------------------------------------------------------------------------
// sum over s of v[s * stride], in shot order; the loads of 8 shots are issued
// before their adds so a cell waits for memory once per 8 shots.
template <typename ACC>
__device__ __forceinline__ ACC sum_shots(const float* __restrict__ v, long long stride, int nshot)
{
    ACC acc = 0;
    int s = 0;
    for (; s + 8 <= nshot; s += 8) {
        float x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = v[(long long)(s + q) * stride];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc += (ACC)x[q];
    }
    for (; s < nshot; ++s) acc += (ACC)v[(long long)s * stride];
    return acc;
}

// Sum of the per-shot material terms of cell h of the history box, in shot
// order, added to the accumulators (one saved step).
template <typename ACC>
__device__ __forceinline__ void reduce_terms(const Tm2dGrid& g, const Box& b, const float* __restrict__ terms,
                                             long long h, int want_eps, int want_sig, ACC* __restrict__ geps,
                                             ACC* __restrict__ gsig)
{
    const int i = (int)(h / b.ny) + b.x0, j = (int)(h % b.ny) + b.y0;
    const long long out = (long long)i * (g.NY - 1) + j;
    const long long sstride = (long long)g.nshot * b.cells;
    if (want_eps) geps[out] += sum_shots<ACC>(terms + h, b.cells, g.nshot);
    if (want_sig) gsig[out] += sum_shots<ACC>(terms + sstride + h, b.cells, g.nshot);
}

template <int ORDER, typename ACC>
__device__ __forceinline__ void tm_adj_tile(const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                                            const SourceBlocks& src, const TmReceiversAdj& rx, const TmAdjGrad& ag,
                                            int it, int tile, int s)
{
    using T = Tile<ORDER>;
    constexpr int R = T::R, HE = T::HE, LR = T::LR, LC = T::LC, HR = T::HR, HC = T::HC;
    constexpr int NW = NTA / 64;
    constexpr int KS = cdiv(LR, NW), KY = cdiv(HR, NW), KT = cdiv(BI, NW);
    __shared__ float sL[LR * LC];   // lambda_Ez after receivers^T
    __shared__ float sA[LR * BJ];   // (Hy, x) weights of (E step)^T, then (Ez, x) weights of (H step)^T
    __shared__ float sB[BI * LC];   // (Hx, y) weights (row stride LC), then (Ez, y) weights (row stride HC)
    __shared__ float sHy[HR * BJ];  // new lambda_Hy on rows [i0-R, i0+BI+R-1)
    __shared__ float sHx[BI * HC];  // new lambda_Hx on columns [j0-R, j0+BJ+R-1)
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int ti = tile / g.ntj, tj = tile - ti * g.ntj;
    const int i0 = ti * BI, j0 = tj * BJ;
    const int lt = tile + g.ftiles * s;
    const long long sb0 = (long long)s * g.plane;
    const float* __restrict__ Lz = pp.Ez[0] + sb0;
    const float* __restrict__ Lx = pp.Hx[0] + sb0;
    const float* __restrict__ Ly = pp.Hy[0] + sb0;

    // this lane's columns: tile column gj, staged halo column hc, Hx-strip halo column xc
    const int gj = j0 + lane;
    const bool col_ok = gj < g.NY;
    const int hc = lane < HE ? lane : (lane < 2 * HE ? BJ + lane : -1);
    const int hgj = j0 - HE + hc;
    const bool h_ok = hc >= 0 && hgj >= 0 && hgj < g.NY;
    const int xc = lane < R ? lane : (lane < 2 * R - 1 ? BJ + lane : -1);
    const int xgj = j0 - R + xc;
    const bool x_ok = xc >= 0 && xgj >= 0 && xgj < g.NY;
    const bool ez_deg = (g.NX - 1) != 1 || (g.NY - 1) != 1;
    int qe[2], qeh[2], qh[2], qx[2];  // y-face depth indices: E at gj / hgj, H at gj / xgj
#pragma unroll
    for (int f2 = 0; f2 < 2; ++f2) {
        qe[f2] = face_q2(g, true, 2 + f2, gj);
        qeh[f2] = face_q2(g, true, 2 + f2, hgj);
        qh[f2] = face_q2(g, false, 2 + f2, gj);
        qx[f2] = face_q2(g, false, 2 + f2, xgj);
    }

    // Every load that does not depend on LDS is issued before the first barrier.
    float st_m[KS], st_h[KS], a_cc[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int r = w + NW * k;
        const int gi = i0 - HE + r;
        const bool row = r < LR && gi >= 0 && gi < g.NX;
        st_m[k] = (row && col_ok) ? Lz[gi * g.NY + gj] : 0.0f;
        st_h[k] = (row && h_ok) ? Lz[gi * g.NY + hgj] : 0.0f;
        a_cc[k] = (row && col_ok) ? f.ccurl[gi * g.NY + gj] : 0.0f;
    }
    float y_l[KY], y_u[KY];  // lambda_Hy and hcurl of the Hy strip
#pragma unroll
    for (int k = 0; k < KY; ++k) {
        const int r = w + NW * k;
        const int gi = i0 - R + r;
        const bool ok = r < HR && gi >= 0 && gi < g.NX && col_ok;
        y_l[k] = ok ? Ly[gi * g.NY + gj] : 0.0f;
        y_u[k] = ok ? f.hcurl[gi * g.NY + gj] : 0.0f;
    }
    const MatGrad& mg = ag.mg;
    const Box& hb = mg.box;
    float x_l[KT], x_u[KT], xe_l[KT], xe_u[KT], b_cm[KT], b_ch[KT], t_ca[KT], t_cb[KT];
    float m_e[KT], m_s[KT], m_sig[KT];  // history of a saved step and sigma (tile cells in the box)
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int r = w + NW * k;
        const int gi = i0 + r;
        const bool row = r < BI && gi < g.NX;
        const bool ok = row && col_ok, okx = row && x_ok, okh = row && h_ok;
        const int p = gi * g.NY + gj, px = gi * g.NY + xgj;
        x_l[k] = ok ? Lx[p] : 0.0f;
        x_u[k] = ok ? f.hcurl[p] : 0.0f;
        b_cm[k] = ok ? f.ccurl[p] : 0.0f;
        t_ca[k] = ok ? f.ca[p] : 0.0f;
        t_cb[k] = ok ? f.cb[p] : 0.0f;
        xe_l[k] = okx ? Lx[px] : 0.0f;
        xe_u[k] = okx ? f.hcurl[px] : 0.0f;
        b_ch[k] = okh ? f.ccurl[gi * g.NY + hgj] : 0.0f;
        m_e[k] = m_s[k] = m_sig[k] = 0.0f;
        if (ag.terms && ok && gi >= hb.x0 && gi < hb.x0 + hb.nx && gj >= hb.y0 && gj < hb.y0 + hb.ny) {
            const long long base = (long long)s * hb.cells + (long long)(gi - hb.x0) * hb.ny + (gj - hb.y0);
            m_e[k] = load_hist(mg.eframe, ag.ez_comp * mg.e_cstride + base, mg.kind);
            m_s[k] = load_hist(mg.sframe, ag.ez_comp * mg.s_cstride + base, mg.kind);
            m_sig[k] = mg.sig[(long long)gi * g.NY * 2 + (long long)gj * 2];  // full-layout material index
        }
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int r = w + NW * k;
        if (r < LR) {
            sL[r * LC + HE + lane] = st_m[k];
            if (hc >= 0) sL[r * LC + hc] = st_h[k];
        }
    }
    __syncthreads();
    if (rx.count > 0) {
        const int b0 = rx.off[lt], b1 = rx.off[lt + 1];
        for (int q = b0 + (int)threadIdx.x; q < b1; q += NTA) {
            const int4 grp = rx.grp[q];
            float acc = sL[grp.x];
            for (int m = grp.y; m < grp.y + grp.z; ++m)
                acc += rx.dgrad[((long long)s * rx.nt + it) * rx.nrx + rx.member[m]];
            sL[grp.x] = acc;
        }
        __syncthreads();
    }
    // source^T: one value per (step, shot, source), summed over shots later
    if (ag.src_vals && src.count > 0) {
        const int q0 = src.block_off[lt], q1 = src.block_off[lt + 1];
        for (int q = q0 + (int)threadIdx.x; q < q1; q += NTA) {
            const int2 ent = src.entries[q];
            const int r = ent.x / BJ, c = ent.x % BJ;
            const int p = (i0 + r) * g.NY + (j0 + c);
            ag.src_vals[((long long)it * g.nshot + s) * ag.nsrc + ent.y] =
                -f.cb[p] * ag.geo * sL[(r + HE) * LC + (c + HE)];
        }
    }
    // per-shot material-gradient terms of a saved step
    if (ag.terms) {
        const long long sstride = (long long)g.nshot * hb.cells;
        const float wgt = (float)mg.weight;
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const int r = w + NW * k;
            const int gi = i0 + r;
            if (r >= BI || gi >= g.NX || !col_ok) continue;
            if (gi < hb.x0 || gi >= hb.x0 + hb.nx || gj < hb.y0 || gj >= hb.y0 + hb.ny) continue;
            const long long base = (long long)s * hb.cells + (long long)(gi - hb.x0) * hb.ny + (gj - hb.y0);
            const float cav = t_ca[k], cbv = t_cb[k];
            const bool up = ez_deg && gi > 0 && gi < g.NX - 1 && gj > 0 && gj < g.NY - 1;
            const float e_old = m_e[k];
            const float sec = m_s[k];
            float lam = sL[(r + HE) * LC + HE + lane];
            float rhs = sec;
            if (mg.e_only) {
                const float a = up ? cav : 1.0f;
                rhs = cbv != 0.0f ? (sec - a * e_old) / cbv : 0.0f;
            }
            if (e_old == 0.0f && rhs == 0.0f) lam = 0.0f;
            const float gca = up ? lam * e_old * wgt : 0.0f;
            const float gcb = lam * rhs * wgt;
            const bool active = m_sig[k] <= kPecSigma;
            if (mg.want_eps) {
                float t = 0.0f;
                if (active) {
                    const float dca = kEps0 * (1.0f - cav) * cbv / mg.dt;
                    const float dcb = -kEps0 * cbv * cbv / mg.dt;
                    t = gca * dca + gcb * dcb;
                }
                ag.terms[base] = t;
            }
            if (mg.want_sig) {
                float t = 0.0f;
                if (active) {
                    const float dca = -0.5f * (1.0f + cav) * cbv;
                    const float dcb = -0.5f * cbv * cbv;
                    t = gca * dca + gcb * dcb;
                }
                ag.terms[sstride + base] = t;
            }
        }
    }

    // (E step)^T weights: x channel on the staged rows (owners advance lambda psi_E of the x faces)
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int r = w + NW * k;
        const int gi = i0 - HE + r;
        if (r >= LR || gi < 0 || gi >= g.NX || !col_ok) continue;
        const int p = gi * g.NY + gj;
        const float lz = sL[r * LC + HE + lane];
        const bool own = gi >= i0 && gi < i0 + BI;
        float acc = 0.0f;
#pragma unroll
        for (int face = 0; face < 2; ++face) {
            const int q = face_q2(g, true, face, gi);
            if (q < 0) continue;
            const Cpml4 kc = cpml_at(f.coef_e[face], g.pml[face], q);
            const long long pi = psi_at(g, true, face, s, gi, gj, q);
            const float old = pp.psi_e[0][face][pi];
            const float u = f.cb[p];
            const float inv = 1.0f / g.d[0];
            acc += (1.0f * u * kc.ra1 * lz - kc.rf * old) * inv;
            if (own) pp.psi_e[1][face][pi] = 1.0f * u * kc.rb * lz + kc.re * old;
        }
        if (ez_deg && gi > 0 && gi < g.NX - 1 && gj > 0 && gj < g.NY - 1) acc += axis_k(g, a_cc[k], 0) * lz;
        sA[r * BJ + lane] = acc;
    }
    // ... y channel on the tile rows: tile columns (owners advance lambda psi_E of the y faces), then halo columns
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int r = w + NW * k;
        const int gi = i0 + r;
        if (r >= BI || gi >= g.NX) continue;
        const bool row_in = ez_deg && gi > 0 && gi < g.NX - 1;
        if (col_ok) {
            const int p = gi * g.NY + gj;
            const float lz = sL[(r + HE) * LC + HE + lane];
            float acc = 0.0f;
#pragma unroll
            for (int f2 = 0; f2 < 2; ++f2) {
                const int q = qe[f2];
                if (q < 0) continue;
                const int face = 2 + f2;
                const Cpml4 kc = cpml_at(f.coef_e[face], g.pml[face], q);
                const long long pi = psi_at(g, true, face, s, gi, gj, q);
                const float old = pp.psi_e[0][face][pi];
                const float u = f.cb[p];
                const float inv = 1.0f / g.d[1];
                acc += (-1.0f * u * kc.ra1 * lz - kc.rf * old) * inv;
                pp.psi_e[1][face][pi] = -1.0f * u * kc.rb * lz + kc.re * old;
            }
            if (row_in && gj > 0 && gj < g.NY - 1) acc += -axis_k(g, b_cm[k], 1) * lz;
            sB[r * LC + HE + lane] = acc;
        }
        if (h_ok) {
            const int p = gi * g.NY + hgj;
            const float lz = sL[(r + HE) * LC + hc];
            float acc = 0.0f;
#pragma unroll
            for (int f2 = 0; f2 < 2; ++f2) {
                const int q = qeh[f2];
                if (q < 0) continue;
                const int face = 2 + f2;
                const Cpml4 kc = cpml_at(f.coef_e[face], g.pml[face], q);
                const float old = pp.psi_e[0][face][psi_at(g, true, face, s, gi, hgj, q)];
                const float u = f.cb[p];
                const float inv = 1.0f / g.d[1];
                acc += (-1.0f * u * kc.ra1 * lz - kc.rf * old) * inv;
            }
            if (row_in && hgj > 0 && hgj < g.NY - 1) acc += -axis_k(g, b_ch[k], 1) * lz;
            sB[r * LC + hc] = acc;
        }
    }
    __syncthreads();

    // lambda_Hy on rows [i0-R, i0+BI+R-1): the (Hy, x) channel, then the identically zero (Hy, z) channel
#pragma unroll
    for (int k = 0; k < KY; ++k) {
        const int r = w + NW * k;
        const int gi = i0 - R + r;
        if (r >= HR || gi < 0 || gi >= g.NX || !col_ok) continue;
        const int base = (r + R - 1) * BJ + lane;  // sA row of gi is gi - (i0 - HE)
        auto wx = [&](int off) { return sA[base + off * BJ]; };
        const float v = (y_l[k] + gather_bwd_T<ORDER>(wx, gi, g.NX)) + 0.0f;
        sHy[r * BJ + lane] = v;
        if (gi >= i0 && gi < i0 + BI) pp.Hy[1][sb0 + gi * g.NY + gj] = v;
    }
    // lambda_Hx on the tile rows: tile columns, then the strip's halo columns
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int r = w + NW * k;
        const int gi = i0 + r;
        if (r >= BI || gi >= g.NX) continue;
        if (col_ok) {
            const int base = r * LC + HE + lane;  // sB column of gj is gj - (j0 - HE)
            auto wy = [&](int off) { return sB[base + off]; };
            const float v = (x_l[k] + gather_bwd_T<ORDER>(wy, gj, g.NY)) + 0.0f;
            sHx[r * HC + R + lane] = v;
            pp.Hx[1][sb0 + gi * g.NY + gj] = v;
        }
        if (x_ok) {
            const int base = r * LC + (xc - R + HE);
            auto wy = [&](int off) { return sB[base + off]; };
            sHx[r * HC + xc] = (xe_l[k] + gather_bwd_T<ORDER>(wy, xgj, g.NY)) + 0.0f;
        }
    }
    __syncthreads();

    // (H step)^T weights: x channel into sA (rows [i0-R, ...)), y channel into sB (row stride HC)
#pragma unroll
    for (int k = 0; k < KY; ++k) {
        const int r = w + NW * k;
        const int gi = i0 - R + r;
        if (r >= HR || gi < 0 || gi >= g.NX || !col_ok) continue;
        const int p = gi * g.NY + gj;
        const float ly = sHy[r * BJ + lane];
        const bool own = gi >= i0 && gi < i0 + BI;
        float acc = 0.0f;
#pragma unroll
        for (int face = 0; face < 2; ++face) {
            const int q = face_q2(g, false, face, gi);
            if (q < 0) continue;
            const Cpml4 kc = cpml_at(f.coef_h[face], g.pml[face], q);
            const long long pi = psi_at(g, false, face, s, gi, gj, q);
            const float old = pp.psi_h[0][face][pi];
            const float u = f.hrhs[p];
            const float inv = 1.0f / g.d[0];
            acc += (1.0f * u * kc.ra1 * ly - kc.rf * old) * inv;
            if (own) pp.psi_h[1][face][pi] = 1.0f * u * kc.rb * ly + kc.re * old;
        }
        if ((g.NY - 1) != 1 && gi < g.NX - 1 && gj > 0 && gj < g.NY - 1) acc += axis_k(g, y_u[k], 0) * ly;
        sA[r * BJ + lane] = acc;
    }
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int r = w + NW * k;
        const int gi = i0 + r;
        if (r >= BI || gi >= g.NX) continue;
        const bool row_up = (g.NX - 1) != 1 && gi > 0 && gi < g.NX - 1;
        if (col_ok) {
            const int p = gi * g.NY + gj;
            const float lx = sHx[r * HC + R + lane];
            float acc = 0.0f;
#pragma unroll
            for (int f2 = 0; f2 < 2; ++f2) {
                const int q = qh[f2];
                if (q < 0) continue;
                const int face = 2 + f2;
                const Cpml4 kc = cpml_at(f.coef_h[face], g.pml[face], q);
                const long long pi = psi_at(g, false, face, s, gi, gj, q);
                const float old = pp.psi_h[0][face][pi];
                const float u = f.hrhs[p];
                const float inv = 1.0f / g.d[1];
                acc += (-1.0f * u * kc.ra1 * lx - kc.rf * old) * inv;
                pp.psi_h[1][face][pi] = -1.0f * u * kc.rb * lx + kc.re * old;
            }
            if (row_up && gj < g.NY - 1) acc += -axis_k(g, x_u[k], 1) * lx;
            sB[r * HC + R + lane] = acc;
        }
        if (x_ok) {
            const int p = gi * g.NY + xgj;
            const float lx = sHx[r * HC + xc];
            float acc = 0.0f;
#pragma unroll
            for (int f2 = 0; f2 < 2; ++f2) {
                const int q = qx[f2];
                if (q < 0) continue;
                const int face = 2 + f2;
                const Cpml4 kc = cpml_at(f.coef_h[face], g.pml[face], q);
                const float old = pp.psi_h[0][face][psi_at(g, false, face, s, gi, xgj, q)];
                const float u = f.hrhs[p];
                const float inv = 1.0f / g.d[1];
                acc += (-1.0f * u * kc.ra1 * lx - kc.rf * old) * inv;
            }
            if (row_up && xgj < g.NY - 1) acc += -axis_k(g, xe_u[k], 1) * lx;
            sB[r * HC + xc] = acc;
        }
    }
    __syncthreads();

    // lambda_Ez = (ca) lambda_Ez + D+^T (x weights) + D+^T (y weights)
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        const int r = w + NW * k;
        const int gi = i0 + r;
        if (r >= BI || gi >= g.NX || !col_ok) continue;
        float acc = sL[(r + HE) * LC + HE + lane];
        if (ez_deg && gi > 0 && gi < g.NX - 1 && gj > 0 && gj < g.NY - 1) acc = t_ca[k] * acc;
        const int bx = (r + R) * BJ + lane;  // sA row of gi is gi - (i0 - R)
        const int by = r * HC + R + lane;    // sB column of gj is gj - (j0 - R)
        auto vx = [&](int off) { return sA[bx + off * BJ]; };
        auto vy = [&](int off) { return sB[by + off]; };
        acc += gather_fwd_T<ORDER>(vx, gi, g.NX);
        acc += gather_fwd_T<ORDER>(vy, gj, g.NY);
        pp.Ez[1][sb0 + gi * g.NY + gj] = acc;
    }

    // Shot sums of the previous saved step: the workgroup of shot s takes the
    // s-th slice of its tile, so every cell is summed once.
    if (ag.pending) {
        const Box& b = ag.mg.box;
        const int chunk = (BI * BJ + g.nshot - 1) / g.nshot;
        const int e_end = min(BI * BJ, (s + 1) * chunk);
        for (int e = s * chunk + (int)threadIdx.x; e < e_end; e += NTA) {
            const int gi = i0 + e / BJ, gjj = j0 + e % BJ;
            if (gi >= g.NX || gjj >= g.NY || gi < b.x0 || gi >= b.x0 + b.nx || gjj < b.y0 || gjj >= b.y0 + b.ny)
                continue;
            reduce_terms<ACC>(g, b, ag.pending, (long long)(gi - b.x0) * b.ny + (gjj - b.y0), ag.mg.want_eps,
                              ag.mg.want_sig, (ACC*)ag.geps, (ACC*)ag.gsig);
        }
    }
}

template <int ORDER, typename ACC>
__global__ __launch_bounds__(NTA) void k_tm_step_adj(Tm2dGrid g, Tm2dFields f, TmPingPong pp, SourceBlocks src,
                                                     TmReceiversAdj rx, TmAdjGrad ag, int it)
{
    const int tile = step_tile(g.ftiles);
    if (tile < 0) return;
    tm_adj_tile<ORDER, ACC>(g, f, pp, src, rx, ag, it, tile, blockIdx.y);
}

// Material gradient of one saved step: per-shot terms summed in shot order.
template <typename ACC>
__global__ __launch_bounds__(256) void k_tm_grad_reduce(Tm2dGrid g, Box b, const float* __restrict__ terms,
                                                        int want_eps, int want_sig, ACC* __restrict__ geps,
                                                        ACC* __restrict__ gsig)
{
    const long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= b.cells) return;
    reduce_terms<ACC>(g, b, terms, h, want_eps, want_sig, geps, gsig);
}

// grad_source[k, it] += per-shot values, shots in order
__global__ __launch_bounds__(256) void k_tm_src_reduce(const float* __restrict__ vals, float* __restrict__ gsrc,
                                                       int nshot, int nsrc, int nt)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nsrc * nt) return;
    const int k = (int)(idx / nt), it = (int)(idx - (long long)k * nt);
    float acc = gsrc[idx];
    for (int s = 0; s < nshot; ++s) acc += vals[((long long)it * nshot + s) * nsrc + k];
    gsrc[idx] = acc;
}

// ---------------------------------------------------------------------------
// Batched steps: one launch advances every instance of the batch by one step.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int batch_instance(const int* __restrict__ prefix, int ninst, int b)
{
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int batch_group(int total)
{
    const int chunk = gridDim.x >> 3;
    const int b = (blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
    return b < total ? b : -1;
}

__device__ __forceinline__ TmPingPong batch_buffers(const TmPingPong& pp, int parity)
{
    if (!parity) return pp;
    TmPingPong q;
    q.Ez[0] = pp.Ez[1];
    q.Ez[1] = pp.Ez[0];
    q.Hx[0] = pp.Hx[1];
    q.Hx[1] = pp.Hx[0];
    q.Hy[0] = pp.Hy[1];
    q.Hy[1] = pp.Hy[0];
#pragma unroll
    for (int f4 = 0; f4 < 4; ++f4) {
        q.psi_e[0][f4] = pp.psi_e[1][f4];
        q.psi_e[1][f4] = pp.psi_e[0][f4];
        q.psi_h[0][f4] = pp.psi_h[1][f4];
        q.psi_h[1][f4] = pp.psi_h[0][f4];
    }
    return q;
}

__device__ __forceinline__ size_t hist_elem_bytes(int kind) { return kind == kF32 ? 4 : 2; }

template <int ORDER>
__global__ __launch_bounds__(NTF) void k_tm_batch_fwd(const TmBatchInst* __restrict__ insts,
                                                      const int* __restrict__ prefix, int ninst, int total,
                                                      int gstep)
{
    const int b = batch_group(total);
    if (b < 0) return;
    const int k = batch_instance(prefix, ninst, b);
    const TmBatchInst& I = insts[k];
    const Tm2dGrid g = I.g;
    const Tm2dFields f = I.f;
    const int local = b - prefix[k];
    const int s = local / g.ftiles, tile = local - s * g.ftiles;
    const int it = gstep - I.start;
    const TmPingPong pp = batch_buffers(I.pp, it & 1);
    HistWrite hw{};
    hw.mode = I.mode;
    hw.box = I.box;
    if (I.e_saved) {  // E-only history at every step (sampling interval 1)
        hw.active = 1;
        hw.save_e = 1;
        hw.e_kind = I.kind;
        hw.e_base = (unsigned char*)I.e_saved + (size_t)it * I.snap * hist_elem_bytes(I.kind);
        hw.e_cstride = I.cstride;
    }
    tm_fwd_tile<ORDER>(g, f, pp, I.src, I.rx, hw, it, tile, s);
}

template <int ORDER>
__global__ __launch_bounds__(NTA) void k_tm_batch_adj(const TmBatchInst* __restrict__ insts,
                                                      const int* __restrict__ prefix, int ninst, int total,
                                                      int gstep)
{
    const int b = batch_group(total);
    if (b < 0) return;
    const int k = batch_instance(prefix, ninst, b);
    const TmBatchInst& I = insts[k];
    const Tm2dGrid g = I.g;
    const Tm2dFields f = I.f;
    const int local = b - prefix[k];
    const int s = local / g.ftiles, tile = local - s * g.ftiles;
    const int done = gstep - I.start;
    const int it = I.nt - 1 - done;
    const TmPingPong pp = batch_buffers(I.pp, done & 1);
    const size_t eb = hist_elem_bytes(I.kind);
    TmAdjGrad ag{};
    ag.geps = I.geps;
    ag.gsig = I.gsig;
    ag.ez_comp = I.mode == 3 ? 2 : 0;
    MatGrad& mg = ag.mg;
    mg.sig = I.sig_full;
    mg.kind = I.kind;
    mg.e_only = 1;
    mg.mode = I.mode;
    mg.weight = 1;
    mg.want_eps = I.want_eps;
    mg.want_sig = I.want_sig;
    mg.dt = I.dt;
    mg.box = I.box;
    mg.eframe = (const unsigned char*)I.e_saved + (size_t)it * I.snap * eb;
    mg.e_cstride = I.cstride;
    if (it + 1 < I.nt_saved) {
        mg.sframe = (const unsigned char*)I.e_saved + (size_t)(it + 1) * I.snap * eb;
        mg.s_cstride = I.cstride;
    } else {
        mg.sframe = I.e_final;
        mg.s_cstride = I.snap;
    }
    // sampling interval 1: buffer (nt - 1 - it) & 1 takes this step's terms,
    // the other one holds the previous step's (summed by this launch)
    if (I.want_eps || I.want_sig) {
        ag.terms = I.terms + (size_t)(done & 1) * I.terms_len;
        if (done > 0) ag.pending = I.terms + (size_t)((done - 1) & 1) * I.terms_len;
    }
    SourceBlocks none{};
    tm_adj_tile<ORDER, double>(g, f, pp, none, I.rxa, ag, it, tile, s);
}

// The final E frame of an E-only history from the compact Ez (Ex, Ey are zero).
__global__ __launch_bounds__(256) void k_tm_final_frame(Tm2dGrid g, const float* __restrict__ ez, HistWrite hw)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)g.nshot * hw.box.cells) return;
    const int s = (int)(idx / hw.box.cells);
    const long long h = idx - (long long)s * hw.box.cells;
    const int i = (int)(h / hw.box.ny) + hw.box.x0, j = (int)(h % hw.box.ny) + hw.box.y0;
    const float v = ez[(long long)s * g.plane + (long long)i * g.NY + j];
    const int comps = hw.mode == 3 ? 3 : 1;
    for (int ci = 0; ci < comps; ++ci)
        store_hist(hw.e_base, ci * hw.e_cstride + idx, ci == comps - 1 ? v : 0.0f, hw.e_kind);
}

__global__ __launch_bounds__(256) void k_scale(const float* __restrict__ in, float* __restrict__ out, float a,
                                               long long count)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < count) out[idx] = a * in[idx];
}

void launch_tm_batch(int order, bool adjoint, const TmBatchInst* insts, const int* prefix, int ninst,
                     int total_groups, int gstep, hipStream_t st)
{
    if (ninst <= 0 || total_groups <= 0) return;
    const dim3 grid((unsigned)((total_groups + 7) / 8 * 8));
    if (adjoint) {
        if (order == 8)
            hipLaunchKernelGGL(k_tm_batch_adj<8>, grid, dim3(NTA), 0, st, insts, prefix, ninst, total_groups, gstep);
        else if (order == 4)
            hipLaunchKernelGGL(k_tm_batch_adj<4>, grid, dim3(NTA), 0, st, insts, prefix, ninst, total_groups, gstep);
        else
            hipLaunchKernelGGL(k_tm_batch_adj<2>, grid, dim3(NTA), 0, st, insts, prefix, ninst, total_groups, gstep);
    } else {
        if (order == 8)
            hipLaunchKernelGGL(k_tm_batch_fwd<8>, grid, dim3(NTF), 0, st, insts, prefix, ninst, total_groups, gstep);
        else if (order == 4)
            hipLaunchKernelGGL(k_tm_batch_fwd<4>, grid, dim3(NTF), 0, st, insts, prefix, ninst, total_groups, gstep);
        else
            hipLaunchKernelGGL(k_tm_batch_fwd<2>, grid, dim3(NTF), 0, st, insts, prefix, ninst, total_groups, gstep);
    }
}

void launch_tm_final_frame(const Tm2dGrid& g, const float* ez, const HistWrite& hw, hipStream_t st)
{
    const long long n = (long long)g.nshot * hw.box.cells;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_tm_final_frame, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, ez, hw);
}

void launch_scale(const float* in, float* out, float factor, long long count, hipStream_t st)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, in, out, factor, count);
}

// --- launchers -----------------------------------------------------------------

static dim3 step_grid(const Tm2dGrid& g) { return dim3((unsigned)((g.ftiles + 7) / 8 * 8), (unsigned)g.nshot); }

void launch_tm_step_fwd(int order, const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                        const SourceBlocks& src, const TmReceiversFwd& rx, const HistWrite& hw, int it,
                        hipStream_t st)
{
    const dim3 grid = step_grid(g);
    if (order == 8) hipLaunchKernelGGL(k_tm_step_fwd<8>, grid, dim3(NTF), 0, st, g, f, pp, src, rx, hw, it);
    else if (order == 4) hipLaunchKernelGGL(k_tm_step_fwd<4>, grid, dim3(NTF), 0, st, g, f, pp, src, rx, hw, it);
    else hipLaunchKernelGGL(k_tm_step_fwd<2>, grid, dim3(NTF), 0, st, g, f, pp, src, rx, hw, it);
}

template <typename ACC>
static void launch_adj_acc(int order, const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                           const SourceBlocks& src, const TmReceiversAdj& rx, const TmAdjGrad& ag, int it,
                           hipStream_t st)
{
    const dim3 grid = step_grid(g);
    if (order == 8) hipLaunchKernelGGL((k_tm_step_adj<8, ACC>), grid, dim3(NTA), 0, st, g, f, pp, src, rx, ag, it);
    else if (order == 4)
        hipLaunchKernelGGL((k_tm_step_adj<4, ACC>), grid, dim3(NTA), 0, st, g, f, pp, src, rx, ag, it);
    else hipLaunchKernelGGL((k_tm_step_adj<2, ACC>), grid, dim3(NTA), 0, st, g, f, pp, src, rx, ag, it);
}

void launch_tm_step_adj(int order, const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                        const SourceBlocks& src, const TmReceiversAdj& rx, const TmAdjGrad& ag, int it, bool fp64,
                        hipStream_t st)
{
    if (fp64) launch_adj_acc<double>(order, g, f, pp, src, rx, ag, it, st);
    else launch_adj_acc<float>(order, g, f, pp, src, rx, ag, it, st);
}

void launch_tm_grad_reduce(const Tm2dGrid& g, const Box& b, const float* terms, int want_eps, int want_sig,
                           void* geps, void* gsig, bool fp64, hipStream_t st)
{
    if (b.cells <= 0) return;
    const unsigned blocks = (unsigned)((b.cells + 255) / 256);
    if (fp64)
        hipLaunchKernelGGL(k_tm_grad_reduce<double>, dim3(blocks), dim3(256), 0, st, g, b, terms, want_eps,
                           want_sig, (double*)geps, (double*)gsig);
    else
        hipLaunchKernelGGL(k_tm_grad_reduce<float>, dim3(blocks), dim3(256), 0, st, g, b, terms, want_eps,
                           want_sig, (float*)geps, (float*)gsig);
}

void launch_tm_src_reduce(const float* vals, float* gsrc, int nshot, int nsrc, int nt, hipStream_t st)
{
    const long long n = (long long)nsrc * nt;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_tm_src_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vals, gsrc, nshot,
                       nsrc, nt);
}

}  // namespace dgx
