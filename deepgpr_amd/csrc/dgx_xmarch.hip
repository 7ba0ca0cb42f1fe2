// x-marching cell sweeps of the 3D FDTD step and of its exact adjoint.
//
// A workgroup owns a (y, z) tile of one shot, one wavefront per tile row of
// 64 cells and one cell per thread, and walks a chunk of x-planes in order. Per plane it stages the values
// the y- and z-differences read (tile plus halo) in LDS, double-buffered so
// that one barrier per plane suffices; the x-differences read a register
// queue that holds the thread's own column over the 2R planes the stencil
// spans. Every value of a sweep therefore comes from HBM once (the halos
// mostly from L2), where a one-thread-per-cell sweep reads each y- and
// x-neighbour line again from other workgroups.
//
// Forward: the magnetic sweep reads E and advances H, the electric sweep
// reads H and advances E (CPML, sources and the history are fused into it),
// both in place: a sweep never writes what it reads from a neighbour.
//
// Adjoint: the channel weights of the transposed curl (deepgpr_cpu.c:1112-1459)
// are formed on the fly for the staged region instead of going through a
// scratch array: W(u, a) of a cell is its lambda of component 3-u-a times the
// curl coefficient, plus the CPML term that reads the cell's lambda psi. The
// E-side sweep gathers W_E into lambda_H; the H-side sweep gathers W_H into
// lambda_E. A sweep may only read the lambda psi of its neighbours, so each
// lambda psi is advanced by the sweep that reads it at no neighbour:
// lambda psi_E by the H-side sweep (from the lambda_E it holds before scaling
// it), lambda psi_H one step late by the next E-side sweep (from lambda_H
// before that sweep adds to it), and after the last step by a final pass.
// The arithmetic of every value, and the order of every sum, is that of the
// one-thread-per-cell kernels, so results are the same bits.
#include <algorithm>
#include <cstdlib>

#include "dgx_device.h"
#include "dgx_internal.h"

namespace dgx {

namespace {

constexpr int kXmTZ = 64;        // tile width along z: one wavefront per tile row

// Global-memory views of pointers read from the constant block: accesses
// through them are global loads and stores. Accessed through a plain pointer
// loaded from memory they would be flat accesses, which also count as LDS
// traffic, so every LDS wait in the plane loop would wait for the next
// plane's prefetch as well.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const float gcfloat;
__device__ __forceinline__ gcfloat* gconst(const float* p) { return (gcfloat*)p; }
__device__ __forceinline__ gfloat* gmut(float* p) { return (gfloat*)p; }

// Logical tile of this workgroup (XCD-contiguous order when remapped: block b
// runs on XCD b % 8, and XCD x takes tiles [x * per, (x + 1) * per)).
__device__ __forceinline__ bool xm_tile(const XmTiles& xt, int& tile, int& y0, int& z0)
{
    int b = (int)blockIdx.x;
    if (xt.remap) {
        const int per = (int)gridDim.x >> 3;
        b = (b & 7) * per + (b >> 3);
    }
    if (b >= xt.nty * xt.ntz) return false;
    const int ty = b / xt.ntz;
    tile = b;
    y0 = ty * xt.ty;
    z0 = (b - ty * xt.ntz) * kXmTZ;
    return true;
}

// Differences and transposed differences along x of one own cell from its
// register queue: q[B + off] holds plane i + off. Every index is a
// compile-time constant after unrolling, so the queue stays in registers.
template <int ORDER, int B, int QN>
__device__ __forceinline__ float qdfwd(const float (&q)[QN], int c, int n)
{
    if constexpr (ORDER <= 2) {
        return q[B + 1] - q[B];
    } else {
        constexpr int R = Radius<ORDER>::value;
        const int rr = fit_fwd(c, n, R);
        float acc = 0.0f;
        if (rr == R) {
#pragma unroll
            for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (q[B + r] - q[B - (r - 1)]);
        } else {
#pragma unroll
            for (int r = 1; r <= R; ++r)
                if (r <= rr) acc += stencil_c(rr, r) * (q[B + r] - q[B - (r - 1)]);
        }
        return acc;
    }
}

template <int ORDER, int B, int QN>
__device__ __forceinline__ float qdbwd(const float (&q)[QN], int c, int n)
{
    if constexpr (ORDER <= 2) {
        return q[B] - q[B - 1];
    } else {
        constexpr int R = Radius<ORDER>::value;
        const int rr = fit_bwd(c, n, R);
        float acc = 0.0f;
        if (rr == R) {
#pragma unroll
            for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (q[B + (r - 1)] - q[B - r]);
        } else {
#pragma unroll
            for (int r = 1; r <= R; ++r)
                if (r <= rr) acc += stencil_c(rr, r) * (q[B + (r - 1)] - q[B - r]);
        }
        return acc;
    }
}

template <int ORDER, int B, int QN>
__device__ __forceinline__ float qdfwd_full(const float (&q)[QN])
{
    if constexpr (ORDER <= 2) {
        return q[B + 1] - q[B];
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
#pragma unroll
        for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (q[B + r] - q[B - (r - 1)]);
        return acc;
    }
}

template <int ORDER, int B, int QN>
__device__ __forceinline__ float qdbwd_full(const float (&q)[QN])
{
    if constexpr (ORDER <= 2) {
        return q[B] - q[B - 1];
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
#pragma unroll
        for (int r = 1; r <= R; ++r) acc += stencil_c(R, r) * (q[B + (r - 1)] - q[B - r]);
        return acc;
    }
}

// gather_bwd_T / gather_fwd_T (dgx_device.h) with the weights in the queue.
template <int ORDER, int B, int QN>
__device__ __forceinline__ float qgather_bwd_T(const float (&q)[QN], int c, int n)
{
    if constexpr (ORDER <= 2) {
        float acc = q[B];
        if (c + 1 < n) acc -= q[B + 1];
        return acc;
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
        if (c >= 2 * R - 1 && c + 2 * R - 1 < n) {
#pragma unroll
            for (int r = 1; r <= R; ++r) {
                const float kc = stencil_c(R, r);
                acc += kc * q[B - (r - 1)];
                acc -= kc * q[B + r];
            }
            return acc;
        }
#pragma unroll
        for (int r = 1; r <= R; ++r) {
            const int sp = c - (r - 1);
            if (sp >= 0) {
                const int rs = fit_bwd(sp, n, R);
                if (r <= rs) acc += stencil_c(rs, r) * q[B - (r - 1)];
            }
            const int sm = c + r;
            if (sm < n) {
                const int rs = fit_bwd(sm, n, R);
                if (r <= rs) acc -= stencil_c(rs, r) * q[B + r];
            }
        }
        return acc;
    }
}

template <int ORDER, int B, int QN>
__device__ __forceinline__ float qgather_fwd_T(const float (&q)[QN], int c, int n)
{
    if constexpr (ORDER <= 2) {
        float acc = 0.0f;
        if (c >= 1) acc += q[B - 1];
        acc -= q[B];
        return acc;
    } else {
        constexpr int R = Radius<ORDER>::value;
        float acc = 0.0f;
        if (c >= 2 * R - 1 && c + 2 * R <= n) {
#pragma unroll
            for (int r = 1; r <= R; ++r) {
                const float kc = stencil_c(R, r);
                acc += kc * q[B - r];
                acc -= kc * q[B + (r - 1)];
            }
            return acc;
        }
#pragma unroll
        for (int r = 1; r <= R; ++r) {
            const int sp = c - r;
            if (sp >= 0) {
                const int rs = fit_fwd(sp, n, R);
                if (r <= rs) acc += stencil_c(rs, r) * q[B - r];
            }
            const int sm = c + (r - 1);
            if (sm < n) {
                const int rs = fit_fwd(sm, n, R);
                if (r <= rs) acc -= stencil_c(rs, r) * q[B + (r - 1)];
            }
        }
        return acc;
    }
}

// Halo cells of a staged plane: NHY cells in the LO rows below and HI rows
// above the tile (tile columns only), then NHZ cells in the LO columns left
// and HI columns right of the tile (tile rows only). Region coordinates put
// the tile at (LO, LO). Returns false past the halo.
template <int TY, int LO, int HI>
__device__ __forceinline__ bool halo_cell(int h, int y0, int z0, int& yy, int& zz, int& lidx, bool& yhalo)
{
    constexpr int RZ = kXmTZ + LO + HI;
    constexpr int NHY = (LO + HI) * kXmTZ;
    constexpr int NHZ = TY * (LO + HI);
    if (h < NHY) {
        const int r = h / kXmTZ, zc = h - r * kXmTZ;
        const int row = r < LO ? r : TY + r;  // region row
        yy = y0 - LO + row;
        zz = z0 + zc;
        lidx = row * RZ + LO + zc;
        yhalo = true;
        return true;
    }
    h -= NHY;
    if (h >= NHZ) return false;
    const int yr = h / (LO + HI), cz = h - yr * (LO + HI);
    const int col = cz < LO ? cz : kXmTZ + cz;  // region column
    yy = y0 + yr;
    zz = z0 - LO + col;
    lidx = (LO + yr) * RZ + col;
    yhalo = false;
    return true;
}

// v[i] for a runtime i without indexing a register array (select chains)
__device__ __forceinline__ float pick6(const float (&v)[6], int i)
{
    float r = v[0];
    r = i == 1 ? v[1] : r;
    r = i == 2 ? v[2] : r;
    r = i == 3 ? v[3] : r;
    r = i == 4 ? v[4] : r;
    r = i == 5 ? v[5] : r;
    return r;
}

// f[t] += corr (plus) or -= corr for a runtime component t
__device__ __forceinline__ void add_comp(float (&f)[3], int t, float corr, bool plus)
{
    const float v0 = plus ? f[0] + corr : f[0] - corr;
    const float v1 = plus ? f[1] + corr : f[1] - corr;
    const float v2 = plus ? f[2] + corr : f[2] - corr;
    f[0] = t == 0 ? v0 : f[0];
    f[1] = t == 1 ? v1 : f[1];
    f[2] = t == 2 ? v2 : f[2];
}

// The CPML terms below read the solve's grid, CPML profiles and auxiliaries
// through one pointer, and each sweep calls every one of them from a single
// place in a loop that is not unrolled, which keeps its register use low.

// The CPML part of channel weight W(u, a) of cell c (deepgpr_cpu.c:1153-1239
// for the E side, 1332-1410 for the H side), summed from 0 over the two faces
// of axis a: (sign uc (RA-1) lt - RF psi) / d. lt is the lambda of component
// t = 3 - u - a; the cell's lambda psi is read, not advanced.
template <bool ELECTRIC>
__device__ __forceinline__ float chan_weight_cpml(const XmConst* __restrict__ kc, int u, int a, float lt, int c0,
                                               int c1, int c2, long long m, int s)
{
    const Grid& g = kc->g;
    const Cpml& cp = ELECTRIC ? kc->cpe : kc->cph;
    gcfloat* upd = gconst(ELECTRIC ? kc->cb : kc->hrhs);
    const int c[3] = {c0, c1, c2};
    const int w = face_source(a, 0) == u ? 0 : 1;
    const int ca_ = a == 0 ? c0 : (a == 1 ? c1 : c2);
    float acc = 0.0f;
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
        const int f = 2 * a + side;
        const int qq = face_q(g, ELECTRIC, f, ca_);
        if (qq < 0) continue;
        if (!psi_active(g, ELECTRIC, f, w, c)) continue;
        int dims[3];
        psi_dims(g, ELECTRIC, f, w, dims);
        gcfloat* Rc = gconst(cp.coef[f]);
        const int pp = g.pml[f];
        const float ra1 = Rc[qq] - 1.0f, rf = Rc[3 * pp + qq];
        const float sign = ELECTRIC ? e_face_sign(a, w) : h_face_sign(a, w);
        const float uc = upd[m];
        const float inv = 1.0f / g.d[a];
        const float old = gconst(cp.psi[f][w])[psi_index(dims, s, c, a, qq)];
        acc += (sign * uc * ra1 * lt - rf * old) * inv;
    }
    return acc;
}

// Channel weight W(u, a) of cell c. cpml: the cell lies in a face region of
// this side.
template <bool ELECTRIC>
__device__ __forceinline__ float chan_weight(const XmConst* __restrict__ kc, const Grid& g, int u, int a, float lt,
                                             float cu, bool cpml, const int (&c)[3], long long m, int s)
{
    const int t = 3 - u - a;
    const float kk = axis_coeff(g, cu, a);
    const int t1 = t == 2 ? 0 : t + 1;
    const bool plus = (a == t1) == ELECTRIC;
    if (!cpml && plain_cell(g, 1, c)) return 0.0f + (plus ? kk : -kk) * lt;
    float acc = 0.0f;
    if (cpml) acc = chan_weight_cpml<ELECTRIC>(kc, u, a, lt, c[0], c[1], c[2], m, s);
    const bool up = ELECTRIC ? e_updated(g, t, c) : h_updated(g, t, c);
    if (up) acc += (plus ? kk : -kk) * lt;
    return acc;
}

// Advance the lambda psi of cell c: psi <- sign uc RB lambda_target + RE psi
// (the psi stores of deepgpr_cpu.c:1153-1239 / 1332-1410).
template <bool ELECTRIC>
__device__ __forceinline__ void advance_lambda_psi(const XmConst* __restrict__ kc, float l0, float l1, float l2,
                                                long long m, int c0, int c1, int c2, int s)
{
    const Grid& g = kc->g;
    const Cpml& cp = ELECTRIC ? kc->cpe : kc->cph;
    const float uc = gconst(ELECTRIC ? kc->cb : kc->hrhs)[m];
    const int c[3] = {c0, c1, c2};
#pragma unroll 1
    for (int f = 0; f < 6; ++f) {
        const int axis = f >> 1;
        const int qq = face_q(g, ELECTRIC, f, axis == 0 ? c0 : (axis == 1 ? c1 : c2));
        if (qq < 0) continue;
        gcfloat* Rc = gconst(cp.coef[f]);
        const int pp = g.pml[f];
        const float rb = Rc[pp + qq], re = Rc[2 * pp + qq];
#pragma unroll 1
        for (int w = 0; w < 2; ++w) {
            if (!psi_active(g, ELECTRIC, f, w, c)) continue;
            int dims[3];
            psi_dims(g, ELECTRIC, f, w, dims);
            const float sign = ELECTRIC ? e_face_sign(axis, w) : h_face_sign(axis, w);
            gfloat* psi = gmut(cp.psi[f][w]);
            const long long pi = psi_index(dims, s, c, axis, qq);
            const int tg = face_target(axis, w);
            const float lt = tg == 0 ? l0 : (tg == 1 ? l1 : l2);
            const float old = psi[pi];
            psi[pi] = sign * uc * rb * lt + re * old;
        }
    }
}

// CPML terms of a forward half step at cell c (deepgpr_cpu.c:722-883 for E,
// 939-1100 for H): for every face region holding c and each of its two
// auxiliaries, f[target] +-= upd ((RA-1) der + RB psi), psi <- RE psi - RF der,
// der the difference in dv divided by the spacing; faces in the order of the
// one-thread-per-cell kernels. Returns the three corrected components.
struct Vec3 {
    float v0, v1, v2;
};

template <bool ELECTRIC>
__device__ __forceinline__ Vec3 cpml_forward(const XmConst* __restrict__ kc, float f0, float f1, float f2, float d0,
                                          float d1, float d2, float d3, float d4, float d5, long long m, int c0,
                                          int c1, int c2, int s)
{
    const Grid& g = kc->g;
    const Cpml& cp = ELECTRIC ? kc->cpe : kc->cph;
    const float upd = gconst(ELECTRIC ? kc->cb : kc->hrhs)[m];
    const int c[3] = {c0, c1, c2};
    const float dv[6] = {d0, d1, d2, d3, d4, d5};
    float f[3] = {f0, f1, f2};
#pragma unroll 1
    for (int fc = 0; fc < 6; ++fc) {
        const int axis = fc >> 1;
        const int qq = face_q(g, ELECTRIC, fc, axis == 0 ? c0 : (axis == 1 ? c1 : c2));
        if (qq < 0) continue;
        gcfloat* Rc = gconst(cp.coef[fc]);
        const int pp = g.pml[fc];
        const float ra1 = Rc[qq] - 1.0f, rb = Rc[pp + qq], re = Rc[2 * pp + qq], rf = Rc[3 * pp + qq];
        const float da = g.d[axis];
#pragma unroll 1
        for (int w = 0; w < 2; ++w) {
            if (!psi_active(g, ELECTRIC, fc, w, c)) continue;
            int dims[3];
            psi_dims(g, ELECTRIC, fc, w, dims);
            const int tg = face_target(axis, w), sc = face_source(axis, w);
            const float der = pick6(dv, diff_slot(sc, axis)) / da;
            gfloat* psi = gmut(cp.psi[fc][w]);
            const long long pi = psi_index(dims, s, c, axis, qq);
            const float phi = psi[pi];
            const float corr = upd * (ra1 * der + rb * phi);
            const float sign = ELECTRIC ? e_face_sign(axis, w) : h_face_sign(axis, w);
            add_comp(f, tg, corr, !(sign < 0.0f));
            psi[pi] = re * phi - rf * der;
        }
    }
    return Vec3{f[0], f[1], f[2]};
}

// Staging geometry shared by the sweeps: own cell (y0 + wave, z0 + lane), its
// region index, and the halo cell of this thread (if any).
template <int TY, int LO, int HI>
struct XmCell {
    static constexpr int RZ = kXmTZ + LO + HI;
    static constexpr int NH = (LO + HI) * (kXmTZ + TY);
    int j, k, l, po;
    bool ok;
    int hl, hp, hyy, hzz;  // halo: region index (-1: none), plane offset, coordinates
    bool hv, hy;           // halo cell inside the grid; y halo (else z halo)
    __device__ __forceinline__ XmCell(const Grid& g, int y0, int z0)
    {
        const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
        j = y0 + wv;
        k = z0 + lane;
        ok = j < g.N[1] && k < g.N[2];
        l = (wv + LO) * RZ + LO + lane;
        po = j * g.N[2] + k;
        int yy = 0, zz = 0;
        bool yh = false;
        const bool have = (int)threadIdx.x < NH && halo_cell<TY, LO, HI>((int)threadIdx.x, y0, z0, yy, zz, hl, yh);
        hv = have && yy >= 0 && yy < g.N[1] && zz >= 0 && zz < g.N[2];
        hy = yh;
        hyy = yy;
        hzz = zz;
        hp = hv ? yy * g.N[2] + zz : 0;
        if (!have) hl = -1;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Forward magnetic sweep: H -= curl E (forward differences) + H-CPML
// (deepgpr_cpu.c:886-1100). Stages Ex (y and z halos), Ey (z), Ez (y); the
// queue holds Ey and Ez over planes i-R+1 .. i+R.
// ---------------------------------------------------------------------------
template <int ORDER, int TY>
__global__ __launch_bounds__(64 * TY) void k_xm_fwd_h(Grid g, XmTiles xt, const XmConst* __restrict__ kc, CF3 E,
                                                      F3 H)
{
    constexpr int R = Radius<ORDER>::value;
    constexpr int LO = R - 1, HI = R;
    using Cell = XmCell<TY, LO, HI>;
    constexpr int RZ = Cell::RZ, RC = (TY + LO + HI) * RZ;
    static_assert(Cell::NH <= 64 * TY, "one halo cell per thread");
    constexpr int QN = 2 * R, B = R - 1;
    __shared__ float sm[2][3][RC];

    int tile, y0, z0;
    if (!xm_tile(xt, tile, y0, z0)) return;
    const int NX = g.N[0], NY = g.N[1], NZ = g.N[2];
    const int xa = (int)blockIdx.y * xt.xc;
    const int xb = min(NX, xa + xt.xc);
    const int s = blockIdx.z;
    const long long sb = (long long)s * g.cells;
    const Cell cl(g, y0, z0);
    const int j = cl.j, k = cl.k, l = cl.l;
    gcfloat* hcurl = gconst(kc->hcurl);
    const bool has_cpml = g.pml[0] + g.pml[1] + g.pml[2] + g.pml[3] + g.pml[4] + g.pml[5] > 0;

    float q1[QN], q2[QN];  // Ey, Ez of the own column
#pragma unroll
    for (int p = 0; p < QN - 1; ++p) {
        const int x = xa - R + 1 + p;
        float v1 = 0.0f, v2 = 0.0f;
        if (cl.ok && x >= 0 && x < NX) {
            const long long n = sb + (long long)x * g.st0 + cl.po;
            v1 = E.c[1][n];
            v2 = E.c[2][n];
        }
        q1[p] = v1;
        q2[p] = v2;
    }
    // the next plane's values, loaded one plane ahead
    float nl1, nl2, nex, nh0, nh1, nh2, nhc, nha, nhb;
    auto load_plane = [&](int i) {
        const int xl = i + R;
        const long long pb = sb + (long long)i * g.st0;
        nl1 = 0.0f;
        nl2 = 0.0f;
        if (cl.ok && xl < NX) {
            const long long n = sb + (long long)xl * g.st0 + cl.po;
            nl1 = E.c[1][n];
            nl2 = E.c[2][n];
        }
        nex = cl.ok ? E.c[0][pb + cl.po] : 0.0f;
        nh0 = cl.ok ? H.c[0][pb + cl.po] : 0.0f;
        nh1 = cl.ok ? H.c[1][pb + cl.po] : 0.0f;
        nh2 = cl.ok ? H.c[2][pb + cl.po] : 0.0f;
        nhc = cl.ok ? hcurl[(long long)i * g.st0 + cl.po] : 0.0f;
        nha = cl.hv ? E.c[0][pb + cl.hp] : 0.0f;
        nhb = cl.hv ? (cl.hy ? E.c[2][pb + cl.hp] : E.c[1][pb + cl.hp]) : 0.0f;
    };
    load_plane(xa);
    for (int i = xa; i < xb; ++i) {
        q1[QN - 1] = nl1;
        q2[QN - 1] = nl2;
        const float ex = nex, hc = nhc, ha = nha, hb = nhb;
        float h0 = nh0, h1 = nh1, h2 = nh2;
        if (i + 1 < xb) load_plane(i + 1);
        const int bsel = (i - xa) & 1;
        float* s0 = sm[bsel][0];
        float* s1 = sm[bsel][1];
        float* s2 = sm[bsel][2];
        s0[l] = ex;
        s1[l] = q1[B];
        s2[l] = q2[B];
        if (cl.hl >= 0) {
            s0[cl.hl] = ha;
            if (cl.hy) s2[cl.hl] = hb;
            else s1[cl.hl] = hb;
        }
        __syncthreads();
        if (cl.ok) {
            const int c[3] = {i, j, k};
            const long long m = (long long)i * g.st0 + cl.po;
            const long long n = sb + m;
            const bool pml = has_cpml && in_any_face(g, false, c);
            if (!pml && plain_cell(g, R, c)) {
                const float u1 = hc;
                const float uy = axis_coeff(g, u1, 1), uz = axis_coeff(g, u1, 2);
                {
                    const float dEz = dfwd_full<ORDER, int>(s2, l, RZ);
                    const float dEy = dfwd_full<ORDER, int>(s1, l, 1);
                    h0 = h0 - uy * dEz + uz * dEy;
                }
                {
                    const float dEx = dfwd_full<ORDER, int>(s0, l, 1);
                    const float dEz = qdfwd_full<ORDER, B>(q2);
                    h1 = h1 - uz * dEx + u1 * dEz;
                }
                {
                    const float dEy = qdfwd_full<ORDER, B>(q1);
                    const float dEx = dfwd_full<ORDER, int>(s0, l, RZ);
                    h2 = h2 - u1 * dEy + uy * dEx;
                }
                H.c[0][n] = h0;
                H.c[1][n] = h1;
                H.c[2][n] = h2;
            } else {
                const bool up0 = h_updated(g, 0, c), up1 = h_updated(g, 1, c), up2 = h_updated(g, 2, c);
                if (up0 || up1 || up2 || pml) {
                    const float u1 = hc;
                    const float uy = axis_coeff(g, u1, 1), uz = axis_coeff(g, u1, 2);
                    // all six differences: every read stays in the staged
                    // region or the queue; those no term needs go unused
                    const float d0 = dfwd<ORDER, int>(s2, l, RZ, j, NY);  // dEz/dy
                    const float d1 = dfwd<ORDER, int>(s1, l, 1, k, NZ);   // dEy/dz
                    const float d2 = dfwd<ORDER, int>(s0, l, 1, k, NZ);   // dEx/dz
                    const float d3 = qdfwd<ORDER, B>(q2, i, NX);          // dEz/dx
                    const float d4 = qdfwd<ORDER, B>(q1, i, NX);          // dEy/dx
                    const float d5 = dfwd<ORDER, int>(s0, l, RZ, j, NY);  // dEx/dy
                    if (up0) h0 = h0 - uy * d0 + uz * d1;
                    if (up1) h1 = h1 - uz * d2 + u1 * d3;
                    if (up2) h2 = h2 - u1 * d4 + uy * d5;
                    if (pml) {
                        const Vec3 r = cpml_forward<false>(kc, h0, h1, h2, d0, d1, d2, d3, d4, d5, m, i, j, k, s);
                        h0 = r.v0;
                        h1 = r.v1;
                        h2 = r.v2;
                    }
                    if (up0 || pml) H.c[0][n] = h0;
                    if (up1 || pml) H.c[1][n] = h1;
                    if (up2 || pml) H.c[2][n] = h2;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < QN - 1; ++p) {
            q1[p] = q1[p + 1];
            q2[p] = q2[p + 1];
        }
    }
}

// ---------------------------------------------------------------------------
// Forward electric sweep: E = ca E + curl H (backward differences) + E-CPML,
// then the sources of the cell in source order, then the history E^n and R^n
// (deepgpr_cpu.c:644-883, 1559-1633). Stages Hx (y and z halos), Hy (z), Hz
// (y); the queue holds Hy and Hz over planes i-R .. i+R-1. Sources are listed
// per (shot, tile), in plane order and, within a plane, source order.
// ---------------------------------------------------------------------------
template <int ORDER, int TY>
__global__ __launch_bounds__(64 * TY) void k_xm_fwd_e(Grid g, XmTiles xt, const XmConst* __restrict__ kc, CF3 Hf,
                                                      F3 E, SourceBlocks src, HistWrite hist)
{
    constexpr int R = Radius<ORDER>::value;
    constexpr int LO = R, HI = R - 1;
    using Cell = XmCell<TY, LO, HI>;
    constexpr int RZ = Cell::RZ, RC = (TY + LO + HI) * RZ;
    static_assert(Cell::NH <= 64 * TY, "one halo cell per thread");
    constexpr int QN = 2 * R, B = R;
    __shared__ float sm[2][3][RC];

    int tile, y0, z0;
    if (!xm_tile(xt, tile, y0, z0)) return;
    const int NX = g.N[0], NY = g.N[1], NZ = g.N[2];
    const int xa = (int)blockIdx.y * xt.xc;
    const int xb = min(NX, xa + xt.xc);
    const int s = blockIdx.z;
    const long long sb = (long long)s * g.cells;
    const Cell cl(g, y0, z0);
    const int j = cl.j, k = cl.k, l = cl.l;
    gcfloat* ca = gconst(kc->ca);
    gcfloat* ccurl = gconst(kc->ccurl);
    gcfloat* cb = gconst(kc->cb);
    const bool has_cpml = g.pml[0] + g.pml[1] + g.pml[2] + g.pml[3] + g.pml[4] + g.pml[5] > 0;
    // this workgroup's source entries, plane by plane
    int sp = 0, se = 0;
    if (src.count > 0) {
        const int key = s * xt.nty * xt.ntz + tile;
        sp = src.block_off[key];
        se = src.block_off[key + 1];
        while (sp < se && (int)((long long)src.entries[sp].x / g.st0) < xa) ++sp;
    }

    float q1[QN], q2[QN];  // Hy, Hz of the own column
#pragma unroll
    for (int p = 0; p < QN - 1; ++p) {
        const int x = xa - R + p;
        float v1 = 0.0f, v2 = 0.0f;
        if (cl.ok && x >= 0 && x < NX) {
            const long long n = sb + (long long)x * g.st0 + cl.po;
            v1 = Hf.c[1][n];
            v2 = Hf.c[2][n];
        }
        q1[p] = v1;
        q2[p] = v2;
    }
    float nl1, nl2, nhx, ne0, ne1, ne2, nca, ncc, ncb, nha, nhb;
    auto load_plane = [&](int i) {
        const int xl = i + R - 1;
        const long long pb = sb + (long long)i * g.st0;
        const long long mb = (long long)i * g.st0;
        nl1 = 0.0f;
        nl2 = 0.0f;
        if (cl.ok && xl < NX) {
            const long long n = sb + (long long)xl * g.st0 + cl.po;
            nl1 = Hf.c[1][n];
            nl2 = Hf.c[2][n];
        }
        nhx = cl.ok ? Hf.c[0][pb + cl.po] : 0.0f;
        ne0 = cl.ok ? E.c[0][pb + cl.po] : 0.0f;
        ne1 = cl.ok ? E.c[1][pb + cl.po] : 0.0f;
        ne2 = cl.ok ? E.c[2][pb + cl.po] : 0.0f;
        nca = cl.ok ? ca[mb + cl.po] : 0.0f;
        ncc = cl.ok ? ccurl[mb + cl.po] : 0.0f;
        ncb = cl.ok ? cb[mb + cl.po] : 0.0f;
        nha = cl.hv ? Hf.c[0][pb + cl.hp] : 0.0f;
        nhb = cl.hv ? (cl.hy ? Hf.c[2][pb + cl.hp] : Hf.c[1][pb + cl.hp]) : 0.0f;
    };
    load_plane(xa);
    for (int i = xa; i < xb; ++i) {
        q1[QN - 1] = nl1;
        q2[QN - 1] = nl2;
        const float hx = nhx, u0 = nca, u1 = ncc, cbv = ncb, ha = nha, hb = nhb;
        float e[3] = {ne0, ne1, ne2};
        if (i + 1 < xb) load_plane(i + 1);
        // source entries of this plane: [sp, sq)
        int sq = sp;
        while (sq < se && (int)((long long)src.entries[sq].x / g.st0) == i) ++sq;
        const bool plane_src = sq > sp;
        const int bsel = (i - xa) & 1;
        float* s0 = sm[bsel][0];
        float* s1 = sm[bsel][1];
        float* s2 = sm[bsel][2];
        s0[l] = hx;
        s1[l] = q1[B];
        s2[l] = q2[B];
        if (cl.hl >= 0) {
            s0[cl.hl] = ha;
            if (cl.hy) s2[cl.hl] = hb;
            else s1[cl.hl] = hb;
        }
        __syncthreads();
        if (cl.ok) {
            const int c[3] = {i, j, k};
            const long long m = (long long)i * g.st0 + cl.po;
            const long long n = sb + m;
            const bool up[3] = {e_updated(g, 0, c), e_updated(g, 1, c), e_updated(g, 2, c)};
            const bool pml = has_cpml && in_any_face(g, true, c);
            const bool in_box = hist.active && i >= hist.box.x0 && i < hist.box.x0 + hist.box.nx &&
                                j >= hist.box.y0 && j < hist.box.y0 + hist.box.ny && k >= hist.box.z0 &&
                                k < hist.box.z0 + hist.box.nz;
            if (up[0] || up[1] || up[2] || pml || plane_src || in_box) {
                const float eold[3] = {e[0], e[1], e[2]};
                const float uy = axis_coeff(g, u1, 1), uz = axis_coeff(g, u1, 2);
                if (!pml && plain_cell(g, R, c)) {
                    {
                        const float dHz = dbwd_full<ORDER, int>(s2, l, RZ);
                        const float dHy = dbwd_full<ORDER, int>(s1, l, 1);
                        e[0] = u0 * e[0] + uy * dHz - uz * dHy;
                    }
                    {
                        const float dHx = dbwd_full<ORDER, int>(s0, l, 1);
                        const float dHz = qdbwd_full<ORDER, B>(q2);
                        e[1] = u0 * e[1] + uz * dHx - u1 * dHz;
                    }
                    {
                        const float dHy = qdbwd_full<ORDER, B>(q1);
                        const float dHx = dbwd_full<ORDER, int>(s0, l, RZ);
                        e[2] = u0 * e[2] + u1 * dHy - uy * dHx;
                    }
                } else {
                    const float d0 = dbwd<ORDER, int>(s2, l, RZ, j, NY);  // dHz/dy
                    const float d1 = dbwd<ORDER, int>(s1, l, 1, k, NZ);   // dHy/dz
                    const float d2 = dbwd<ORDER, int>(s0, l, 1, k, NZ);   // dHx/dz
                    const float d3 = qdbwd<ORDER, B>(q2, i, NX);          // dHz/dx
                    const float d4 = qdbwd<ORDER, B>(q1, i, NX);          // dHy/dx
                    const float d5 = dbwd<ORDER, int>(s0, l, RZ, j, NY);  // dHx/dy
                    if (up[0]) e[0] = u0 * e[0] + uy * d0 - uz * d1;
                    if (up[1]) e[1] = u0 * e[1] + uz * d2 - u1 * d3;
                    if (up[2]) e[2] = u0 * e[2] + u1 * d4 - uy * d5;
                    if (pml) {
                        const Vec3 r = cpml_forward<true>(kc, e[0], e[1], e[2], d0, d1, d2, d3, d4, d5, m, i, j, k, s);
                        e[0] = r.v0;
                        e[1] = r.v1;
                        e[2] = r.v2;
                    }
                }
                // Hertzian dipole sources of this plane, in source order
                for (int qs = sp; qs < sq; ++qs) {
                    const int2 ent = src.entries[qs];
                    if ((long long)ent.x == m) {
                        const float scale = src.wave[(long long)ent.y * src.nt + src.it] * src.dl / src.vol;
                        const int sc = src.comp;
                        const float v = (sc == 0 ? e[0] : (sc == 1 ? e[1] : e[2])) - cbv * scale;
                        e[0] = sc == 0 ? v : e[0];
                        e[1] = sc == 1 ? v : e[1];
                        e[2] = sc == 2 ? v : e[2];
                    }
                }
                if (in_box) {
                    const long long h = ((long long)(i - hist.box.x0) * hist.box.ny + (j - hist.box.y0)) *
                                            hist.box.nz + (k - hist.box.z0);
                    const long long base = (long long)s * hist.box.cells + h;
                    if (hist.mode == 3) {
#pragma unroll
                        for (int ci = 0; ci < 3; ++ci) {
                            if (hist.save_e) store_hist(hist.e_base, ci * hist.e_cstride + base, eold[ci], hist.e_kind);
                            if (hist.save_r) {
                                const float a = up[ci] ? u0 : 1.0f;
                                const float rhs = cbv != 0.0f ? (e[ci] - a * eold[ci]) / cbv : 0.0f;
                                store_hist(hist.r_base, ci * hist.r_cstride + base, rhs, hist.r_kind);
                            }
                        }
                    } else {
                        if (hist.save_e) store_hist(hist.e_base, base, eold[2], hist.e_kind);
                        if (hist.save_r) {
                            const float a = up[2] ? u0 : 1.0f;
                            const float rhs = cbv != 0.0f ? (e[2] - a * eold[2]) / cbv : 0.0f;
                            store_hist(hist.r_base, base, rhs, hist.r_kind);
                        }
                    }
                }
                const bool touched = pml || plane_src;
                if (up[0] || touched) E.c[0][n] = e[0];
                if (up[1] || touched) E.c[1][n] = e[1];
                if (up[2] || touched) E.c[2][n] = e[2];
            }
        }
        sp = sq;
#pragma unroll
        for (int p = 0; p < QN - 1; ++p) {
            q1[p] = q1[p + 1];
            q2[p] = q2[p + 1];
        }
    }
}

// ---------------------------------------------------------------------------
// Adjoint sweeps. ELECTRIC = true: the E-side sweep, lambda_H += D-^T W_E
// (gather_bwd_T, offsets -(R-1) .. R), W_E from lambda_E, ccurl, cb and
// lambda psi_E; first, with `defer`, the lambda psi_H update owed by the
// previous step. ELECTRIC = false: the H-side sweep, lambda_E = ca lambda_E
// (updated components) + D+^T W_H (gather_fwd_T, offsets -R .. R-1), W_H from
// lambda_H, hcurl, hrhs and lambda psi_H; first lambda psi_E is advanced from
// the lambda_E the cell holds. Stages W(0,1), W(2,1) (y halo) and W(0,2),
// W(1,2) (z halo); the queue holds W(1,0), W(2,0) of the own column.
// ---------------------------------------------------------------------------
template <int ORDER, int TY, bool ELECTRIC>
__global__ __launch_bounds__(64 * TY) void k_xm_adj(Grid g, XmTiles xt, const XmConst* __restrict__ kc, CF3 lw,
                                                    F3 lt, int defer)
{
    constexpr int R = Radius<ORDER>::value;
    constexpr int LO = ELECTRIC ? R - 1 : R, HI = ELECTRIC ? R : R - 1;
    using Cell = XmCell<TY, LO, HI>;
    constexpr int RZ = Cell::RZ, RC = (TY + LO + HI) * RZ;
    static_assert(Cell::NH <= 64 * TY, "one halo cell per thread");
    constexpr int QN = 2 * R, B = ELECTRIC ? R - 1 : R;
    constexpr int LEAD = ELECTRIC ? R : R - 1;  // lead plane of the queue, relative to i
    __shared__ float sm[2][4][RC];
    __shared__ float sx[2][64 * TY];

    int tile, y0, z0;
    if (!xm_tile(xt, tile, y0, z0)) return;
    const int NX = g.N[0], NY = g.N[1], NZ = g.N[2];
    const int xa = (int)blockIdx.y * xt.xc;
    const int xb = min(NX, xa + xt.xc);
    const int s = blockIdx.z;
    const long long sb = (long long)s * g.cells;
    const Cell cl(g, y0, z0);
    const int j = cl.j, k = cl.k, l = cl.l;
    // lw: the lambda the weights are formed from; lt: the lambda this sweep advances
    gcfloat* cw = gconst(ELECTRIC ? kc->ccurl : kc->hcurl);
    gcfloat* cav_src = gconst(kc->ca);
    const bool has_cpml = g.pml[0] + g.pml[1] + g.pml[2] + g.pml[3] + g.pml[4] + g.pml[5] > 0;

    // W of the x channels of the own cell at plane x: (1,0) from lambda_2, (2,0) from lambda_1
    auto x_weights = [&](int x, float l1, float l2, float cu, float& w1, float& w2) {
        w1 = 0.0f;
        w2 = 0.0f;
        if (!cl.ok || x < 0 || x >= NX) return;
        const int c[3] = {x, j, k};
        const long long m = (long long)x * g.st0 + cl.po;
        const bool cpml = has_cpml && in_any_face(g, ELECTRIC, c);
#pragma unroll 1
        for (int u = 1; u <= 2; ++u) {
            const float w = chan_weight<ELECTRIC>(kc, g, u, 0, u == 1 ? l2 : l1, cu, cpml, c, m, s);
            w1 = u == 1 ? w : w1;
            w2 = u == 2 ? w : w2;
        }
    };
    float q1[QN], q2[QN];  // W(1,0), W(2,0) of the own column
#pragma unroll
    for (int p = 0; p < QN - 1; ++p) {
        const int x = xa + LEAD - (QN - 1) + p;
        float l1 = 0.0f, l2 = 0.0f, cu = 0.0f;
        if (cl.ok && x >= 0 && x < NX) {
            const long long n = sb + (long long)x * g.st0 + cl.po;
            l1 = lw.c[1][n];
            l2 = lw.c[2][n];
            cu = cw[(long long)x * g.st0 + cl.po];
        }
        x_weights(x, l1, l2, cu, q1[p], q2[p]);
    }
    // raw values of the next plane: the lead plane's lambda_1, lambda_2 and
    // curl coefficient; the plane's lambda (3) and curl coefficient; the
    // lambda this sweep advances (and ca); the halo cell's two lambdas and
    // curl coefficient
    float nl1, nl2, nlc, nw0, nw1, nw2, nwc, nt0, nt1, nt2, nca, nha, nhb, nhc;
    auto load_plane = [&](int i) {
        const int xl = i + LEAD;
        const long long pb = sb + (long long)i * g.st0;
        const long long mb = (long long)i * g.st0;
        nl1 = 0.0f;
        nl2 = 0.0f;
        nlc = 0.0f;
        if (cl.ok && xl < NX) {
            const long long n = sb + (long long)xl * g.st0 + cl.po;
            nl1 = lw.c[1][n];
            nl2 = lw.c[2][n];
            nlc = cw[(long long)xl * g.st0 + cl.po];
        }
        nw0 = cl.ok ? lw.c[0][pb + cl.po] : 0.0f;
        nw1 = cl.ok ? lw.c[1][pb + cl.po] : 0.0f;
        nw2 = cl.ok ? lw.c[2][pb + cl.po] : 0.0f;
        nwc = cl.ok ? cw[mb + cl.po] : 0.0f;
        nt0 = cl.ok ? lt.c[0][pb + cl.po] : 0.0f;
        nt1 = cl.ok ? lt.c[1][pb + cl.po] : 0.0f;
        nt2 = cl.ok ? lt.c[2][pb + cl.po] : 0.0f;
        if constexpr (!ELECTRIC) nca = cl.ok ? cav_src[mb + cl.po] : 0.0f;
        // y halo: W(0,1) from lambda_2, W(2,1) from lambda_0; z halo: W(0,2) from lambda_1, W(1,2) from lambda_0
        nha = cl.hv ? (cl.hy ? lw.c[2][pb + cl.hp] : lw.c[1][pb + cl.hp]) : 0.0f;
        nhb = cl.hv ? lw.c[0][pb + cl.hp] : 0.0f;
        nhc = cl.hv ? cw[mb + cl.hp] : 0.0f;
    };
    load_plane(xa);
    for (int i = xa; i < xb; ++i) {
        const float ll1 = nl1, ll2 = nl2, llc = nlc, w0 = nw0, w1 = nw1, w2 = nw2, wc = nwc, ha = nha, hb = nhb,
                    hc = nhc;
        float t0 = nt0, t1 = nt1, t2 = nt2;
        float cav = 0.0f;
        if constexpr (!ELECTRIC) cav = nca;
        if (i + 1 < xb) load_plane(i + 1);
        // staged channels: 0 W(0,1), 1 W(2,1) (y halo); 2 W(0,2), 3 W(1,2) (z halo)
        const int bsel = (i - xa) & 1;
        const long long mb = (long long)i * g.st0;
        // The plane's weights, one item at a time: 0-1 the x channels of the
        // own column at the lead plane (through sx into the queue), 2-5 the y
        // and z channels of the own cell, 6-7 the two channels of the halo cell.
        const int tid = (int)threadIdx.x;
#pragma unroll 1
        for (int it = 0; it < 8; ++it) {
            const bool lead = it < 2, halo = it >= 6;
            if (halo && cl.hl < 0) continue;
            const bool valid = lead ? (cl.ok && i + LEAD < NX) : (halo ? cl.hv : cl.ok);
            int u, a, ch, idx;
            float lam, cu;
            if (lead) {
                u = it + 1;
                a = 0;
                lam = it == 0 ? ll2 : ll1;
                cu = llc;
                ch = -1;
                idx = it;
            } else if (!halo) {
                u = it == 3 ? 2 : (it == 5 ? 1 : 0);
                a = it < 4 ? 1 : 2;
                lam = it == 2 ? w2 : (it == 4 ? w1 : w0);
                cu = wc;
                ch = it - 2;
                idx = l;
            } else {
                a = cl.hy ? 1 : 2;
                u = it == 6 ? 0 : (cl.hy ? 2 : 1);
                lam = it == 6 ? ha : hb;
                cu = hc;
                ch = (cl.hy ? 0 : 2) + (it - 6);
                idx = cl.hl;
            }
            float wv = 0.0f;
            if (valid) {
                const int c[3] = {lead ? i + LEAD : i, halo ? cl.hyy : j, halo ? cl.hzz : k};
                const long long m = (long long)c[0] * g.st0 + (halo ? cl.hp : cl.po);
                const bool cpml = has_cpml && in_any_face(g, ELECTRIC, c);
                wv = chan_weight<ELECTRIC>(kc, g, u, a, lam, cu, cpml, c, m, s);
            }
            if (lead) sx[idx][tid] = wv;
            else sm[bsel][ch][idx] = wv;
        }
        q1[QN - 1] = sx[0][tid];
        q2[QN - 1] = sx[1][tid];
        __syncthreads();
        if (cl.ok) {
            const int c[3] = {i, j, k};
            const long long m = mb + cl.po;
            const long long n = sb + m;
            float tv[3] = {t0, t1, t2};
            if constexpr (ELECTRIC) {
                if (defer && has_cpml && in_any_face(g, false, c))
                    advance_lambda_psi<false>(kc, tv[0], tv[1], tv[2], m, i, j, k, s);
#pragma unroll
                for (int u = 2; u >= 0; --u) {
                    float acc = tv[u];
                    if (u != 0) acc += u == 1 ? qgather_bwd_T<ORDER, B>(q1, i, NX) : qgather_bwd_T<ORDER, B>(q2, i, NX);
                    if (u != 1) {
                        const int ch = u == 0 ? 0 : 1;
                        auto wf = [&](int off) { return sm[bsel][ch][l + off * RZ]; };
                        acc += gather_bwd_T<ORDER>(wf, j, NY);
                    }
                    if (u != 2) {
                        const int ch = u == 0 ? 2 : 3;
                        auto wf = [&](int off) { return sm[bsel][ch][l + off]; };
                        acc += gather_bwd_T<ORDER>(wf, k, NZ);
                    }
                    lt.c[u][n] = acc;
                }
            } else {
                if (has_cpml && in_any_face(g, true, c)) advance_lambda_psi<true>(kc, tv[0], tv[1], tv[2], m, i, j, k, s);
#pragma unroll
                for (int t = 2; t >= 0; --t) {
                    float acc = tv[t];
                    if (e_updated(g, t, c)) acc = cav * acc;
                    if (t != 0) acc += t == 1 ? qgather_fwd_T<ORDER, B>(q1, i, NX) : qgather_fwd_T<ORDER, B>(q2, i, NX);
                    if (t != 1) {
                        const int ch = t == 0 ? 0 : 1;
                        auto wf = [&](int off) { return sm[bsel][ch][l + off * RZ]; };
                        acc += gather_fwd_T<ORDER>(wf, j, NY);
                    }
                    if (t != 2) {
                        const int ch = t == 0 ? 2 : 3;
                        auto wf = [&](int off) { return sm[bsel][ch][l + off]; };
                        acc += gather_fwd_T<ORDER>(wf, k, NZ);
                    }
                    lt.c[t][n] = acc;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < QN - 1; ++p) {
            q1[p] = q1[p + 1];
            q2[p] = q2[p + 1];
        }
    }
}

// The lambda psi_H update owed by the last adjoint step (one thread per cell).
__global__ __launch_bounds__(256) void k_xm_psi_h_final(Grid g, const XmConst* __restrict__ kc, CF3 lH)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)g.nshot * g.cells) return;
    const int s = (int)(idx / g.cells);
    const long long m = idx - (long long)s * g.cells;
    const int i = (int)(m / g.st0);
    const int rem = (int)(m - (long long)i * g.st0);
    const int j = rem / g.N[2];
    const int c[3] = {i, j, rem - j * g.N[2]};
    if (!in_any_face(g, false, c)) return;
    advance_lambda_psi<false>(kc, lH.c[0][idx], lH.c[1][idx], lH.c[2][idx], m, c[0], c[1], c[2], s);
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

namespace {
int env_int(const char* name, int fallback)
{
    const char* v = std::getenv(name);
    if (!v || !*v) return fallback;
    char* end = nullptr;
    const long x = std::strtol(v, &end, 10);
    return (end && *end == '\0' && x > 0 && x < (1 << 20)) ? (int)x : fallback;
}
}  // namespace

int xm_tile_rows(int order)
{
    // order 8 needs 8 rows: its halo (7 rows and columns) must fit one cell per thread
    const int rows = env_int("DEEPGPR_AMD_XM_TILE_ROWS", 8) == 4 ? 4 : 8;
    return order == 8 ? 8 : rows;
}

int xm_chunk_planes() { return env_int("DEEPGPR_AMD_XM_PLANES", 16); }

int xm_remap() { return env_int("DEEPGPR_AMD_XM_REMAP", 2) == 1 ? 1 : 0; }

XmTiles xm_tiles(const Grid& g, int order)
{
    XmTiles xt{};
    xt.ty = xm_tile_rows(order);
    xt.nty = (g.N[1] + xt.ty - 1) / xt.ty;
    xt.ntz = (g.N[2] + kXmTZ - 1) / kXmTZ;
    xt.xc = xm_chunk_planes();
    xt.remap = xm_remap();
    return xt;
}

int xm_tile_of(const XmTiles& xt, int j, int k) { return (j / xt.ty) * xt.ntz + k / kXmTZ; }

namespace {

dim3 xm_grid(const Grid& g, const XmTiles& xt)
{
    unsigned tiles = (unsigned)(xt.nty * xt.ntz);
    if (xt.remap) tiles = (tiles + 7) / 8 * 8;
    return dim3(tiles, (unsigned)((g.N[0] + xt.xc - 1) / xt.xc), (unsigned)g.nshot);
}

template <int ORDER, int TY>
void xm_fwd_h_t(const Grid& g, const XmTiles& xt, const XmConst* kc, const StepArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL((k_xm_fwd_h<ORDER, TY>), xm_grid(g, xt), dim3(64 * TY), 0, st, g, xt, kc, a.Ec, a.H);
}

template <int ORDER, int TY>
void xm_fwd_e_t(const Grid& g, const XmTiles& xt, const XmConst* kc, const StepArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL((k_xm_fwd_e<ORDER, TY>), xm_grid(g, xt), dim3(64 * TY), 0, st, g, xt, kc, a.Hc, a.E, a.src,
                       a.hist);
}

template <int ORDER, int TY>
void xm_adj_t(bool electric, const Grid& g, const XmTiles& xt, const XmConst* kc, const AdjArgs& a, int defer,
              hipStream_t st)
{
    if (electric)
        hipLaunchKernelGGL((k_xm_adj<ORDER, TY, true>), xm_grid(g, xt), dim3(64 * TY), 0, st, g, xt, kc, a.lEc, a.lH,
                           defer);
    else
        hipLaunchKernelGGL((k_xm_adj<ORDER, TY, false>), xm_grid(g, xt), dim3(64 * TY), 0, st, g, xt, kc, a.lHc,
                           a.lE, defer);
}

#define DGX_XM_DISPATCH(FN, ...)                               \
    do {                                                        \
        const bool t4_ = xt.ty == 4;                            \
        if (order == 8) {                                       \
            FN<8, 8>(__VA_ARGS__);                              \
        } else if (order == 4) {                                \
            if (t4_) FN<4, 4>(__VA_ARGS__);                     \
            else FN<4, 8>(__VA_ARGS__);                         \
        } else {                                                \
            if (t4_) FN<2, 4>(__VA_ARGS__);                     \
            else FN<2, 8>(__VA_ARGS__);                         \
        }                                                       \
    } while (0)

}  // namespace

void launch_xm_fwd_h(int order, const Grid& g, const XmTiles& xt, const XmConst* kc, const StepArgs& a,
                     hipStream_t st)
{
    DGX_XM_DISPATCH(xm_fwd_h_t, g, xt, kc, a, st);
}

void launch_xm_fwd_e(int order, const Grid& g, const XmTiles& xt, const XmConst* kc, const StepArgs& a,
                     hipStream_t st)
{
    DGX_XM_DISPATCH(xm_fwd_e_t, g, xt, kc, a, st);
}

void launch_xm_adj(int order, bool electric, const Grid& g, const XmTiles& xt, const XmConst* kc,
                   const AdjArgs& a, bool defer, hipStream_t st)
{
    DGX_XM_DISPATCH(xm_adj_t, electric, g, xt, kc, a, defer ? 1 : 0, st);
}

void launch_xm_psi_h_final(const Grid& g, const XmConst* kc, const AdjArgs& a, hipStream_t st)
{
    const long long total = (long long)g.nshot * g.cells;
    if (total <= 0 || !a.has_cpml) return;
    hipLaunchKernelGGL(k_xm_psi_h_final, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, g, kc, a.lHc);
}

XmConst xm_const(const Grid& g, const Cpml& cpe, const Cpml& cph, const float* ca, const float* ccurl,
                 const float* cb, const float* hcurl, const float* hrhs)
{
    XmConst c{};
    c.g = g;
    c.cpe = cpe;
    c.cph = cph;
    c.ca = ca;
    c.ccurl = ccurl;
    c.cb = cb;
    c.hcurl = hcurl;
    c.hrhs = hrhs;
    return c;
}

}  // namespace dgx
