// 2D Ez-TM fast path (storage_type bit SOLVER_TM2D, nz = 1, Ez source, zero
// Ex/Ey/Hz states; src/DeepGPR/lib/deepgpr.h:77-83).
//
// Ez, Hx and Hy live in compact (nshot, NX, NY) buffers (y fastest) for the
// whole call; only the four CPML faces' active auxiliaries are touched. Every
// expression mirrors the general path (dgx_kernels.hip) with the inactive
// fields substituted by their exact value 0, including the "+ 0" terms, so
// both paths agree bit for bit.
#include "dgx_device.h"
#include "dgx_internal.h"

namespace dgx {

// Logical tile of this workgroup. Workgroups are placed on the 8 XCDs round
// robin; mapping XCD x to the x-th contiguous chunk of tiles keeps the rows a
// stencil touches in the same XCD's L2.
__device__ __forceinline__ int tm_tile(int tiles)
{
    const int chunk = gridDim.x >> 3;  // gridDim.x is padded to a multiple of 8
    const int t = (blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
    return t < tiles ? t : -1;
}

__device__ __forceinline__ int tm_face_q(const Tm2dGrid& g, bool electric, int face, int c)
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

// Index of the active auxiliary of a face: x faces (s, q, j), y faces (s, i, q).
__device__ __forceinline__ long long tm_psi_index(const Tm2dGrid& g, bool electric, int face, int s, int i, int j,
                                                  int q)
{
    const int depth = g.pml[face] + (electric ? 1 : 0);
    if (face < 2) return ((long long)s * depth + q) * g.NY + j;
    return ((long long)s * g.NX + i) * depth + q;
}

__device__ __forceinline__ float tm_axis_coeff(const Tm2dGrid& g, float c1, int a)
{
    if (a == 0 || g.d[a] == g.d[0]) return c1;
    return c1 * g.d[0] / g.d[a];
}

__device__ __forceinline__ bool tm_ez_updated(const Tm2dGrid& g, int i, int j)
{
    return ((g.NX - 1) != 1 || (g.NY - 1) != 1) && i > 0 && i < g.NX - 1 && j > 0 && j < g.NY - 1;
}
__device__ __forceinline__ bool tm_hx_updated(const Tm2dGrid& g, int i, int j)
{
    return (g.NX - 1) != 1 && i > 0 && i < g.NX - 1 && j < g.NY - 1;
}
__device__ __forceinline__ bool tm_hy_updated(const Tm2dGrid& g, int i, int j)
{
    return (g.NY - 1) != 1 && i < g.NX - 1 && j > 0 && j < g.NY - 1;
}

template <int ORDER>
__global__ __launch_bounds__(256) void k_tm_fwd_h(Tm2dGrid g, Tm2dFields f)
{
    const int tile = tm_tile(g.tiles);
    if (tile < 0) return;
    const int p = tile * 256 + threadIdx.x;
    if (p >= g.plane) return;
    const int s = blockIdx.y;
    const int i = p / g.NY, j = p - (p / g.NY) * g.NY;
    const long long n = (long long)s * g.plane + p;
    const bool upx = tm_hx_updated(g, i, j), upy = tm_hy_updated(g, i, j);
    const int qx[2] = {tm_face_q(g, false, 0, i), tm_face_q(g, false, 1, i)};
    const int qy[2] = {tm_face_q(g, false, 2, j), tm_face_q(g, false, 3, j)};
    const bool pml = qx[0] >= 0 || qx[1] >= 0 || qy[0] >= 0 || qy[1] >= 0;
    if (!(upx || upy || pml)) return;
    float hx = f.Hx[n], hy = f.Hy[n];
    const float u1 = f.hcurl[p];
    const float uy = tm_axis_coeff(g, u1, 1), uz = tm_axis_coeff(g, u1, 2);
    if (upx) {
        const float dEz = dfwd<ORDER>(f.Ez, n, 1, j, g.NY);
        hx = hx - uy * dEz + uz * 0.0f;
    }
    if (upy) {
        const float dEz = dfwd<ORDER>(f.Ez, n, g.NY, i, g.NX);
        hy = hy - uz * 0.0f + u1 * dEz;
    }
    if (pml) {
        const float upd = f.hrhs[p];
#pragma unroll
        for (int face = 0; face < 4; ++face) {
            const int q = face < 2 ? qx[face] : qy[face - 2];
            if (q < 0) continue;
            const float* R = f.coef_h[face];
            const int pp = g.pml[face];
            const float ra1 = R[q] - 1.0f, rb = R[pp + q], re = R[2 * pp + q], rf = R[3 * pp + q];
            float der;
            if (face < 2) der = dfwd<ORDER>(f.Ez, n, g.NY, i, g.NX) / g.d[0];
            else der = dfwd<ORDER>(f.Ez, n, 1, j, g.NY) / g.d[1];
            float* psi = f.psi_h[face];
            const long long pi = tm_psi_index(g, false, face, s, i, j, q);
            const float phi = psi[pi];
            const float corr = upd * (ra1 * der + rb * phi);
            if (face < 2) hy += corr;  // Hy += (x faces)
            else hx -= corr;           // Hx -= (y faces)
            psi[pi] = re * phi - rf * der;
        }
    }
    if (upx || pml) f.Hx[n] = hx;
    if (upy || pml) f.Hy[n] = hy;
}

template <int ORDER>
__global__ __launch_bounds__(256) void k_tm_fwd_e(Tm2dGrid g, Tm2dFields f, SourceBlocks src, HistWrite hist)
{
    const int tile = tm_tile(g.tiles);
    if (tile < 0) return;
    const int p = tile * 256 + threadIdx.x;
    if (p >= g.plane) return;
    const int s = blockIdx.y;
    const int i = p / g.NY, j = p - (p / g.NY) * g.NY;
    const long long n = (long long)s * g.plane + p;
    const bool up = tm_ez_updated(g, i, j);
    const int qx[2] = {tm_face_q(g, true, 0, i), tm_face_q(g, true, 1, i)};
    const int qy[2] = {tm_face_q(g, true, 2, j), tm_face_q(g, true, 3, j)};
    const bool pml = qx[0] >= 0 || qx[1] >= 0 || qy[0] >= 0 || qy[1] >= 0;
    int sb = 0, se = 0;
    if (src.count > 0) {
        sb = src.block_off[tile + g.tiles * s];
        se = src.block_off[tile + g.tiles * s + 1];
    }
    const bool in_box = hist.active && i >= hist.box.x0 && i < hist.box.x0 + hist.box.nx && j >= hist.box.y0 &&
                        j < hist.box.y0 + hist.box.ny;
    if (!(up || pml || se > sb || in_box)) return;
    float ez = f.Ez[n];
    const float eold = ez;
    const float u0 = f.ca[p], u1 = f.ccurl[p];
    const float uy = tm_axis_coeff(g, u1, 1);
    if (up) {
        const float dHy = dbwd<ORDER>(f.Hy, n, g.NY, i, g.NX);
        const float dHx = dbwd<ORDER>(f.Hx, n, 1, j, g.NY);
        ez = u0 * ez + u1 * dHy - uy * dHx;
    }
    const float cbv = f.cb[p];
    if (pml) {
#pragma unroll
        for (int face = 0; face < 4; ++face) {
            const int q = face < 2 ? qx[face] : qy[face - 2];
            if (q < 0) continue;
            const float* R = f.coef_e[face];
            const int pp = g.pml[face];
            const float ra1 = R[q] - 1.0f, rb = R[pp + q], re = R[2 * pp + q], rf = R[3 * pp + q];
            float der;
            if (face < 2) der = dbwd<ORDER>(f.Hy, n, g.NY, i, g.NX) / g.d[0];
            else der = dbwd<ORDER>(f.Hx, n, 1, j, g.NY) / g.d[1];
            float* psi = f.psi_e[face];
            const long long pi = tm_psi_index(g, true, face, s, i, j, q);
            const float phi = psi[pi];
            const float corr = cbv * (ra1 * der + rb * phi);
            if (face < 2) ez += corr;  // Ez += (x faces)
            else ez -= corr;           // Ez -= (y faces)
            psi[pi] = re * phi - rf * der;
        }
    }
    for (int q = sb; q < se; ++q) {
        const int2 ent = src.entries[q];
        if (ent.x == p) {
            const float scale = src.wave[(long long)ent.y * src.nt + src.it] * src.dl / src.vol;
            ez -= cbv * scale;
        }
    }
    if (in_box) {
        const long long h = (long long)(i - hist.box.x0) * hist.box.ny + (j - hist.box.y0);
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
    if (up || pml || se > sb) f.Ez[n] = ez;
}

// (nshot, NX, NY, 2) k = 0 plane <-> compact (nshot, NX, NY)
__global__ __launch_bounds__(256) void k_tm_pack(const float* __restrict__ full, float* __restrict__ compact,
                                                 long long count, int unpack)
{
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    if (unpack) ((float*)full)[2 * idx] = compact[idx];
    else compact[idx] = full[2 * idx];
}

// --- adjoint ----------------------------------------------------------------

template <bool ELECTRIC>
__global__ __launch_bounds__(256) void k_tm_adj_pml(Tm2dGrid g, Tm2dFields f, Tm2dWeights w)
{
    const int face = blockIdx.y, s = blockIdx.z;
    const int pp = g.pml[face];
    if (pp <= 0) return;
    const long long other = face < 2 ? g.NY : g.NX;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= other * pp) return;
    int i, j, c;
    if (face < 2) {  // x faces: (face row, j) with j fastest
        const int r = (int)(idx / g.NY);
        j = (int)(idx - (long long)r * g.NY);
        c = (face & 1) == 0 ? (ELECTRIC ? 1 + r : r) : g.NX - 1 - pp + r;
        i = c;
    } else {  // y faces: (i, face column) with the face coordinate fastest
        i = (int)(idx / pp);
        const int r = (int)(idx - (long long)i * pp);
        c = (face & 1) == 0 ? (ELECTRIC ? 1 + r : r) : g.NY - 1 - pp + r;
        j = c;
    }
    if (c < 0 || c >= (face < 2 ? g.NX : g.NY)) return;
    const int q = tm_face_q(g, ELECTRIC, face, c);
    if (q < 0) return;
    const int p = i * g.NY + j;
    const long long n = (long long)s * g.plane + p;
    const float* R = ELECTRIC ? f.coef_e[face] : f.coef_h[face];
    const float ra1 = R[q] - 1.0f, rb = R[pp + q], re = R[2 * pp + q], rf = R[3 * pp + q];
    const float u = ELECTRIC ? f.cb[p] : f.hrhs[p];
    const float inv = 1.0f / g.d[face < 2 ? 0 : 1];
    // E: x faces couple Ez (+), y faces Ez (-); H: x faces Hy (+), y faces Hx (-)
    const float sign = face < 2 ? 1.0f : -1.0f;
    const float lt = ELECTRIC ? f.Ez[n] : (face < 2 ? f.Hy[n] : f.Hx[n]);
    float* psi = ELECTRIC ? f.psi_e[face] : f.psi_h[face];
    float* wt = ELECTRIC ? w.we[face] : w.wh[face];
    const long long pi = tm_psi_index(g, ELECTRIC, face, s, i, j, q);
    const float old = psi[pi];
    wt[pi] = (sign * u * ra1 * lt - rf * old) * inv;
    psi[pi] = sign * u * rb * lt + re * old;
}

// weight of the (Hy, x) [axis 0] or (Hx, y) [axis 1] channel of (E update + E-CPML)^T
__device__ __forceinline__ float tm_e_channel(const Tm2dGrid& g, const Tm2dFields& f, const Tm2dWeights& w,
                                              bool cpml, int s, int axis, int i, int j)
{
    float acc = 0.0f;
    const int p = i * g.NY + j;
    if (cpml) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int face = 2 * axis + side;
            const int q = tm_face_q(g, true, face, axis == 0 ? i : j);
            if (q >= 0) acc += w.we[face][tm_psi_index(g, true, face, s, i, j, q)];
        }
    }
    if (tm_ez_updated(g, i, j)) {
        const float k = tm_axis_coeff(g, f.ccurl[p], axis);
        acc += (axis == 0 ? k : -k) * f.Ez[(long long)s * g.plane + p];
    }
    return acc;
}

// weight of the (Ez, x) [axis 0, from Hy] or (Ez, y) [axis 1, from Hx] channel of (H update + H-CPML)^T
__device__ __forceinline__ float tm_h_channel(const Tm2dGrid& g, const Tm2dFields& f, const Tm2dWeights& w,
                                              bool cpml, int s, int axis, int i, int j)
{
    float acc = 0.0f;
    const int p = i * g.NY + j;
    if (cpml) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int face = 2 * axis + side;
            const int q = tm_face_q(g, false, face, axis == 0 ? i : j);
            if (q >= 0) acc += w.wh[face][tm_psi_index(g, false, face, s, i, j, q)];
        }
    }
    const bool up = axis == 0 ? tm_hy_updated(g, i, j) : tm_hx_updated(g, i, j);
    if (up) {
        const float k = tm_axis_coeff(g, f.hcurl[p], axis);
        const float lt = axis == 0 ? f.Hy[(long long)s * g.plane + p] : f.Hx[(long long)s * g.plane + p];
        acc += (axis == 0 ? k : -k) * lt;
    }
    return acc;
}

template <int ORDER>
__global__ __launch_bounds__(256) void k_tm_adj_e(Tm2dGrid g, Tm2dFields f, Tm2dWeights w)
{
    const int tile = tm_tile(g.tiles);
    if (tile < 0) return;
    const int p = tile * 256 + threadIdx.x;
    if (p >= g.plane) return;
    const int s = blockIdx.y;
    const int i = p / g.NY, j = p - (p / g.NY) * g.NY;
    const long long n = (long long)s * g.plane + p;
    const bool cpml = g.has_cpml != 0;
    auto wy = [&](int off) { return tm_e_channel(g, f, w, cpml, s, 1, i, j + off); };
    auto wx = [&](int off) { return tm_e_channel(g, f, w, cpml, s, 0, i + off, j); };
    // lambda_Hx: (Hx, y) channel, then the identically zero (Hx, z) channel
    f.Hx[n] = (f.Hx[n] + gather_bwd_T<ORDER>(wy, j, g.NY)) + 0.0f;
    // lambda_Hy: (Hy, x) channel, then the zero (Hy, z) channel
    f.Hy[n] = (f.Hy[n] + gather_bwd_T<ORDER>(wx, i, g.NX)) + 0.0f;
}

template <int ORDER>
__global__ __launch_bounds__(256) void k_tm_adj_h(Tm2dGrid g, Tm2dFields f, Tm2dWeights w)
{
    const int tile = tm_tile(g.tiles);
    if (tile < 0) return;
    const int p = tile * 256 + threadIdx.x;
    if (p >= g.plane) return;
    const int s = blockIdx.y;
    const int i = p / g.NY, j = p - (p / g.NY) * g.NY;
    const long long n = (long long)s * g.plane + p;
    const bool cpml = g.has_cpml != 0;
    float acc = f.Ez[n];
    if (tm_ez_updated(g, i, j)) acc = f.ca[p] * acc;
    auto wx = [&](int off) { return tm_h_channel(g, f, w, cpml, s, 0, i + off, j); };
    auto wy = [&](int off) { return tm_h_channel(g, f, w, cpml, s, 1, i, j + off); };
    acc += gather_fwd_T<ORDER>(wx, i, g.NX);
    acc += gather_fwd_T<ORDER>(wy, j, g.NY);
    f.Ez[n] = acc;
}

template <typename ACC>
__global__ __launch_bounds__(256) void k_tm_material_gradient(Tm2dGrid g, Tm2dFields f, MatGrad mg, int ez_comp,
                                                              ACC* __restrict__ geps, ACC* __restrict__ gsig)
{
    long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const Box b = mg.box;
    if (h >= b.cells) return;
    const int i = (int)(h / b.ny) + b.x0;
    const int j = (int)(h % b.ny) + b.y0;
    const int p = i * g.NY + j;
    const long long m = (long long)i * g.NY * 2 + (long long)j * 2;  // full-layout material index
    const long long out = (long long)i * (g.NY - 1) + j;
    const float cav = mg.ca[m], cbv = mg.cb[m];
    const bool active = mg.sig[m] <= kPecSigma;
    const bool up = tm_ez_updated(g, i, j);
    const float wgt = (float)mg.weight;
    ACC ge = 0, gs = 0;
    for (int s = 0; s < g.nshot; ++s) {
        const long long base = (long long)s * b.cells + h;
        float e_old, sec;
        if (mg.kind == kInt8) {
            e_old = int8_load(mg.e8, b, ez_comp, s, h);
            sec = int8_load(mg.s8, b, ez_comp, s, h);
        } else {
            e_old = load_hist(mg.eframe, ez_comp * mg.e_cstride + base, mg.kind);
            sec = load_hist(mg.sframe, ez_comp * mg.s_cstride + base, mg.kind);
        }
        float lam = f.Ez[(long long)s * g.plane + p];
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
    if (mg.want_eps) geps[out] += ge;
    if (mg.want_sig) gsig[out] += gs;
}

// --- launch helpers -----------------------------------------------------------

static dim3 tm_grid(const Tm2dGrid& g)
{
    unsigned bx = (unsigned)((g.tiles + 7) / 8 * 8);
    return dim3(bx, (unsigned)g.nshot);
}

void launch_tm_fwd_h(int order, const Tm2dGrid& g, const Tm2dFields& f, hipStream_t st)
{
    dim3 grid = tm_grid(g);
    if (order == 8) hipLaunchKernelGGL(k_tm_fwd_h<8>, grid, dim3(256), 0, st, g, f);
    else if (order == 4) hipLaunchKernelGGL(k_tm_fwd_h<4>, grid, dim3(256), 0, st, g, f);
    else hipLaunchKernelGGL(k_tm_fwd_h<2>, grid, dim3(256), 0, st, g, f);
}

void launch_tm_fwd_e(int order, const Tm2dGrid& g, const Tm2dFields& f, const SourceBlocks& src,
                     const HistWrite& hw, hipStream_t st)
{
    dim3 grid = tm_grid(g);
    if (order == 8) hipLaunchKernelGGL(k_tm_fwd_e<8>, grid, dim3(256), 0, st, g, f, src, hw);
    else if (order == 4) hipLaunchKernelGGL(k_tm_fwd_e<4>, grid, dim3(256), 0, st, g, f, src, hw);
    else hipLaunchKernelGGL(k_tm_fwd_e<2>, grid, dim3(256), 0, st, g, f, src, hw);
}

void launch_tm_pack(const float* full, float* compact, long long count, bool unpack, hipStream_t st)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_tm_pack, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, full, compact, count,
                       unpack ? 1 : 0);
}

void launch_tm_adj_pml(bool electric, const Tm2dGrid& g, const Tm2dFields& f, const Tm2dWeights& w,
                       hipStream_t st)
{
    long long best = 0;
    for (int face = 0; face < 4; ++face) {
        long long c = (long long)g.pml[face] * (face < 2 ? g.NY : g.NX);
        if (c > best) best = c;
    }
    if (best <= 0) return;
    dim3 grid((unsigned)((best + 255) / 256), 4, (unsigned)g.nshot);
    if (electric) hipLaunchKernelGGL(k_tm_adj_pml<true>, grid, dim3(256), 0, st, g, f, w);
    else hipLaunchKernelGGL(k_tm_adj_pml<false>, grid, dim3(256), 0, st, g, f, w);
}

void launch_tm_adj_e(int order, const Tm2dGrid& g, const Tm2dFields& f, const Tm2dWeights& w, hipStream_t st)
{
    dim3 grid = tm_grid(g);
    if (order == 8) hipLaunchKernelGGL(k_tm_adj_e<8>, grid, dim3(256), 0, st, g, f, w);
    else if (order == 4) hipLaunchKernelGGL(k_tm_adj_e<4>, grid, dim3(256), 0, st, g, f, w);
    else hipLaunchKernelGGL(k_tm_adj_e<2>, grid, dim3(256), 0, st, g, f, w);
}

void launch_tm_adj_h(int order, const Tm2dGrid& g, const Tm2dFields& f, const Tm2dWeights& w, hipStream_t st)
{
    dim3 grid = tm_grid(g);
    if (order == 8) hipLaunchKernelGGL(k_tm_adj_h<8>, grid, dim3(256), 0, st, g, f, w);
    else if (order == 4) hipLaunchKernelGGL(k_tm_adj_h<4>, grid, dim3(256), 0, st, g, f, w);
    else hipLaunchKernelGGL(k_tm_adj_h<2>, grid, dim3(256), 0, st, g, f, w);
}

void launch_tm_material_gradient(const Tm2dGrid& g, const Tm2dFields& f, const MatGrad& mg, int ez_comp,
                                 void* geps, void* gsig, bool fp64, hipStream_t st)
{
    unsigned blocks = (unsigned)((mg.box.cells + 255) / 256);
    if (fp64)
        hipLaunchKernelGGL(k_tm_material_gradient<double>, dim3(blocks), dim3(256), 0, st, g, f, mg, ez_comp,
                           (double*)geps, (double*)gsig);
    else
        hipLaunchKernelGGL(k_tm_material_gradient<float>, dim3(blocks), dim3(256), 0, st, g, f, mg, ez_comp,
                           (float*)geps, (float*)gsig);
}

}  // namespace dgx
