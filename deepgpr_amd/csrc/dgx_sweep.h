// Block -> cell mapping of the one-thread-per-cell 3D sweeps, shared by the
// kernels, the host launch code and a host-only test.
//
// A sweep advances every cell (i, j, k) of the field grid N0 x N1 x N2 of
// every shot once. The grid of a launch is (per_shot, nshot): block b of shot
// s is either
//  * a shell block (b < shell_end): up to three segments, each a set of
//    planes (two ranges), rows (two ranges) and z-runs (two runs), walked
//    plane by plane in chunks of 256 consecutive (row, z) cells; or
//  * a box block: the interior box, where every cell is plain (no CPML term,
//    no clamped difference, every component updated, full-width adjoint
//    gathers). Its 256-cell chunks are dealt to the 8 XCDs in bands: XCD x
//    (= b mod 8, the hardware's round-robin workgroup placement) takes chunks
//    [x * band, (x + 1) * band) of every plane, so a chunk's y- and
//    x-neighbours are mostly in the same L2.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DGX_HD __host__ __device__ __forceinline__
#else
#define DGX_HD inline
#endif

namespace dgx {

// Division of q < 2^31 by a constant d >= 1 as a multiply and a shift
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", 1994): with 2^(l-1) < d <= 2^l and
// m = floor(2^(31+l) / d) + 1 (< 2^32), floor(q * m / 2^(31+l)) = floor(q / d)
// for every q < 2^31, because 0 < m * d - 2^(31+l) <= d <= 2^l.
struct Div {
    unsigned m;
    unsigned shift;  // 31 + l
};

DGX_HD unsigned div_q(unsigned q, Div v)
{
    return (unsigned)(((unsigned long long)q * v.m) >> v.shift);
}

inline Div make_div(unsigned d)
{
    unsigned l = 0;
    while ((1ULL << l) < d) ++l;
    Div v;
    v.shift = 31 + l;
    v.m = (unsigned)((1ULL << v.shift) / d + 1ULL);
    return v;
}

struct SweepSeg {
    int p[2], np[2];  // planes [p[0], p[0] + np[0]) then [p[1], p[1] + np[1])
    int r[2], nr[2];  // rows
    int z[2], nz[2];  // z runs of every row
    int first;        // first block of the segment within a shot
    int bpp;          // blocks per plane: ceil(rows * (nz[0] + nz[1]) / 256)
    Div per_div;      // / (nz[0] + nz[1])
    Div bpp_div;      // / bpp
};

constexpr int kMaxSegs = 3;

struct CellMap {
    SweepSeg seg[kMaxSegs];
    int nseg;
    int shell_end;            // blocks [0, shell_end) are shell blocks (a multiple of 8)
    int bx, by, bz;           // interior box origin
    int bnx, bny, bnz;        // and extent (bnx == 0: no box)
    int chunks;               // 256-cell chunks of one box plane
    int band;                 // chunks per XCD band: ceil(chunks / 8)
    Div nz_div, band_div;     // / bnz, / band
    int box_blocks;           // 8 * band * bnx
    // Dispatch order: with ilv_s > 0 the launch interleaves periods of ilv_s
    // shell blocks and ilv_b box blocks (multiples of 8), so the latency-bound
    // CPML blocks run beside the bandwidth-bound box blocks instead of before
    // them; ilv_s == 0: every shell block, then every box block.
    int ilv_s, ilv_b;
    Div period_div;           // / (ilv_s + ilv_b)
    int per_shot;             // blocks per shot
};

// Which part block b of a shot belongs to: 0 shell, 1 box, -1 neither
// (padding); idx is its index within the part.
DGX_HD int sweep_part(const CellMap& cm, int b, int& idx)
{
    if (cm.ilv_s == 0) {
        if (b < cm.shell_end) {
            idx = b;
            return 0;
        }
        idx = b - cm.shell_end;
        return idx < cm.box_blocks ? 1 : -1;
    }
    const int P = cm.ilv_s + cm.ilv_b;
    const int g = (int)div_q((unsigned)b, cm.period_div);
    const int r = b - g * P;
    if (r < cm.ilv_s) {
        idx = g * cm.ilv_s + r;
        return idx < cm.shell_end ? 0 : -1;
    }
    idx = g * cm.ilv_b + (r - cm.ilv_s);
    return idx < cm.box_blocks ? 1 : -1;
}

// Box block L (0 <= L < box_blocks; L mod 8 is its XCD), thread t: its
// cell, or false.
DGX_HD bool box_cell(const CellMap& cm, int L, int t, int& i, int& j, int& k)
{
    const int x = L & 7;
    const unsigned slot = (unsigned)L >> 3;
    const int pl = (int)div_q(slot, cm.band_div);
    const int chunk = x * cm.band + (int)(slot - (unsigned)pl * (unsigned)cm.band);
    if (pl >= cm.bnx || chunk >= cm.chunks) return false;
    const unsigned q = (unsigned)chunk * 256u + (unsigned)t;
    if (q >= (unsigned)cm.bny * (unsigned)cm.bnz) return false;
    const int row = (int)div_q(q, cm.nz_div);
    i = cm.bx + pl;
    j = cm.by + row;
    k = cm.bz + (int)(q - (unsigned)row * (unsigned)cm.bnz);
    return true;
}

DGX_HD int seg_of(const CellMap& cm, int b)
{
    int q = 0;
    for (int t = 1; t < kMaxSegs; ++t)
        if (t < cm.nseg && b >= cm.seg[t].first) q = t;
    return q;
}

// Shell block b (0 <= b < shell_end), thread t: its cell, or false.
DGX_HD bool shell_cell(const CellMap& cm, int b, int t, int& i, int& j, int& k)
{
    // the block's segment by selects (an indexed read of the argument block
    // can make the compiler copy the table to scratch memory)
    const int si = seg_of(cm, b);
    const SweepSeg sg = si == 2 ? cm.seg[2] : (si == 1 ? cm.seg[1] : cm.seg[0]);
    const unsigned L = (unsigned)(b - sg.first);
    const int pl = (int)div_q(L, sg.bpp_div);
    if (pl >= sg.np[0] + sg.np[1]) return false;
    const unsigned q = (L - (unsigned)pl * (unsigned)sg.bpp) * 256u + (unsigned)t;
    const unsigned per = (unsigned)(sg.nz[0] + sg.nz[1]);
    if (q >= (unsigned)(sg.nr[0] + sg.nr[1]) * per) return false;
    const int row = (int)div_q(q, sg.per_div);
    const int rz = (int)(q - (unsigned)row * per);
    i = pl < sg.np[0] ? sg.p[0] + pl : sg.p[1] + (pl - sg.np[0]);
    j = row < sg.nr[0] ? sg.r[0] + row : sg.r[1] + (row - sg.nr[0]);
    k = rz < sg.nz[0] ? sg.z[0] + rz : sg.z[1] + (rz - sg.nz[0]);
    return true;
}

// Interior box of a grid for stencil radius R (order / 2) and CPML widths
// pml[6]: per axis, lo = max(2R - 1, pml_lo + 1 if pml_lo > 0) and
// hi = min(N - 2R, N - 2 - pml_hi if pml_hi > 0), inclusive. Cells there lie
// outside every E and H CPML face region (E: 1..p and N-1-p..N-2 on a face
// of width p, H: 0..p-1 and the same upper range), are at least R from
// every face (unclamped forward differences: R <= c <= N-1-R), have every E
// and H component updated (1 <= c <= N-2) and take the full-width branch of
// the adjoint gathers (2R-1 <= c <= N-2R).
inline void interior_box(const int N[3], const int pml[6], int R, int lo[3], int hi[3])
{
    for (int a = 0; a < 3; ++a) {
        const int plo = pml[2 * a], phi = pml[2 * a + 1];
        int l = 2 * R - 1, h = N[a] - 2 * R;
        if (plo > 0 && plo + 1 > l) l = plo + 1;
        if (phi > 0 && N[a] - 2 - phi < h) h = N[a] - 2 - phi;
        lo[a] = l;
        hi[a] = h;
    }
}

// The mapping of a sweep over grid N with radius R. use_box: take the
// interior box out of the shell when it is large enough to pay.
inline CellMap make_cell_map(const int N[3], const int pml[6], int R, bool use_box)
{
    CellMap cm{};
    int lo[3], hi[3];
    interior_box(N, pml, R, lo, hi);
    const long long bnx = hi[0] - lo[0] + 1, bny = hi[1] - lo[1] + 1, bnz = hi[2] - lo[2] + 1;
    const bool box = use_box && bnx >= 1 && bny >= 1 && bnz >= 32 && bnx * bny * bnz >= 65536 &&
                     bny * bnz < (1LL << 31);
    int nseg = 0;
    auto add = [&](int p0, int np0, int p1, int np1, int r0, int nr0, int r1, int nr1, int z0, int nz0, int z1,
                   int nz1) {
        if (np0 + np1 <= 0 || nr0 + nr1 <= 0 || nz0 + nz1 <= 0) return;
        SweepSeg& s = cm.seg[nseg++];
        s.p[0] = p0; s.np[0] = np0; s.p[1] = p1; s.np[1] = np1;
        s.r[0] = r0; s.nr[0] = nr0; s.r[1] = r1; s.nr[1] = nr1;
        s.z[0] = z0; s.nz[0] = nz0; s.z[1] = z1; s.nz[1] = nz1;
        const long long cells = (long long)(nr0 + nr1) * (nz0 + nz1);
        s.bpp = (int)((cells + 255) / 256);
        s.per_div = make_div((unsigned)(nz0 + nz1));
        s.bpp_div = make_div((unsigned)s.bpp);
    };
    if (box) {
        const int x0 = lo[0], x1 = hi[0] + 1, y0 = lo[1], y1 = hi[1] + 1, z0 = lo[2], z1 = hi[2] + 1;
        // z outside the box: every plane and row
        add(0, N[0], 0, 0, 0, N[1], 0, 0, 0, z0, z1, N[2] - z1);
        // planes outside the box, the box's z-run
        add(0, x0, x1, N[0] - x1, 0, N[1], 0, 0, z0, z1 - z0, 0, 0);
        // the box's planes, rows outside it, the box's z-run
        add(x0, x1 - x0, 0, 0, 0, y0, y1, N[1] - y1, z0, z1 - z0, 0, 0);
        cm.bx = x0; cm.by = y0; cm.bz = z0;
        cm.bnx = (int)bnx; cm.bny = (int)bny; cm.bnz = (int)bnz;
    } else {
        // no box: the z-interior of every row and the rest as two segments,
        // so a wavefront rarely mixes the interior and the CPML branches
        const int zl = lo[2], zh = hi[2];
        if (zh - zl + 1 >= 64) {
            add(0, N[0], 0, 0, 0, N[1], 0, 0, zl, zh - zl + 1, 0, 0);
            add(0, N[0], 0, 0, 0, N[1], 0, 0, 0, zl, zh + 1, N[2] - 1 - zh);
        } else {
            add(0, N[0], 0, 0, 0, N[1], 0, 0, 0, N[2], 0, 0);
        }
    }
    cm.nseg = nseg;
    int first = 0;
    for (int q = 0; q < nseg; ++q) {
        cm.seg[q].first = first;
        first += cm.seg[q].bpp * (cm.seg[q].np[0] + cm.seg[q].np[1]);
    }
    cm.shell_end = (first + 7) / 8 * 8;
    if (box) {
        cm.chunks = (int)((bny * bnz + 255) / 256);
        cm.band = (cm.chunks + 7) / 8;
        cm.nz_div = make_div((unsigned)bnz);
        cm.band_div = make_div((unsigned)cm.band);
        cm.box_blocks = 8 * cm.band * cm.bnx;
    }
    cm.per_shot = cm.shell_end + cm.box_blocks;
    return cm;
}

// Interleave the shell and box blocks of a map in periods of 8 shell blocks
// and the nearest multiple of 8 box blocks of the same share.
inline void interleave(CellMap& cm)
{
    if (cm.box_blocks == 0 || cm.shell_end == 0) return;
    cm.ilv_s = 8;
    long long b = (long long)cm.box_blocks * 8 / cm.shell_end;
    b = (b + 4) / 8 * 8;
    cm.ilv_b = (int)(b < 8 ? 8 : b);
    const long long gs = (cm.shell_end + cm.ilv_s - 1) / cm.ilv_s, gb = (cm.box_blocks + cm.ilv_b - 1) / cm.ilv_b;
    const long long groups = gs > gb ? gs : gb;
    const long long P = cm.ilv_s + cm.ilv_b;
    cm.per_shot = (int)(groups * P);
    cm.period_div = make_div((unsigned)P);
}

// Block (within a shot) that advances cell (i, j, k), or -1.
inline long long cell_map_block(const CellMap& cm, int i, int j, int k)
{
    if (cm.bnx > 0 && i >= cm.bx && i < cm.bx + cm.bnx && j >= cm.by && j < cm.by + cm.bny && k >= cm.bz &&
        k < cm.bz + cm.bnz) {
        const long long q = (long long)(j - cm.by) * cm.bnz + (k - cm.bz);
        const int chunk = (int)(q / 256);
        const int x = chunk / cm.band, cb = chunk % cm.band;
        const long long slot = (long long)(i - cm.bx) * cm.band + cb;
        const long long L = slot * 8 + x;
        if (cm.ilv_s == 0) return cm.shell_end + L;
        return (L / cm.ilv_b) * (cm.ilv_s + cm.ilv_b) + cm.ilv_s + L % cm.ilv_b;
    }
    for (int n = 0; n < cm.nseg; ++n) {
        const SweepSeg& s = cm.seg[n];
        auto index = [](int c, const int* at, const int* cnt) {
            if (c >= at[0] && c < at[0] + cnt[0]) return c - at[0];
            if (c >= at[1] && c < at[1] + cnt[1]) return cnt[0] + c - at[1];
            return -1;
        };
        const int pl = index(i, s.p, s.np), row = index(j, s.r, s.nr), rz = index(k, s.z, s.nz);
        if (pl < 0 || row < 0 || rz < 0) continue;
        const long long q = (long long)row * (s.nz[0] + s.nz[1]) + rz;
        const long long S = s.first + (long long)pl * s.bpp + q / 256;
        if (cm.ilv_s == 0) return S;
        return (S / cm.ilv_s) * (cm.ilv_s + cm.ilv_b) + S % cm.ilv_s;
    }
    return -1;
}

}  // namespace dgx
