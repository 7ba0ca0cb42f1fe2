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
#include <algorithm>
#include "dgx_device.h"
#include "dgx_internal.h"

namespace dgx {

namespace {

constexpr int BI = kTmBI;
constexpr int BJ = kTmBJ;
constexpr int NTF = DGX_TM_FWD_THREADS;  // threads per workgroup, forward
constexpr int NTA = DGX_TM_ADJ_THREADS;  // threads per workgroup, adjoint

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

template <int ORDER, int TBI = BI>
struct Tile {
    static constexpr int R = Radius<ORDER>::value;
    static constexpr int HE = 2 * R - 1;        // halo of the staged Ez / lambda_Ez
    static constexpr int LR = TBI + 2 * HE;     // staged rows
    static constexpr int LC = BJ + 2 * HE;      // staged columns
    static constexpr int HR = TBI + 2 * R - 1;  // rows of the Hy strip: R above the tile, R-1 below
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

// Element idx of an array within 4 GiB of its base: the byte offset is a
// 32-bit value, so a wave-uniform base is addressed as SGPR base + 32-bit
// VGPR offset instead of a 64-bit address per lane.
template <typename T>
__device__ __forceinline__ T& at32(T* base, unsigned idx)
{
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (size_t)(idx * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ const T& at32(const T* base, unsigned idx)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)(idx * (unsigned)sizeof(T)));
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
// The time step is it_off, plus *it_base when it_base is set (a replayed
// graph of several steps reads the step of its first node from the device).
// TBI: tile rows (g.bi).
template <int ORDER, int TBI>
__global__ __launch_bounds__(NTF) void k_tm_step_fwd(Tm2dGrid g, Tm2dFields f, TmPingPong pp, SourceBlocks src,
                                                     TmReceiversFwd rx, HistWrite hist, int it_off,
                                                     const int* __restrict__ it_base)
{
    constexpr int BI = TBI;
    const int tile = step_tile(g.ftiles);
    if (tile < 0) return;
    const int s = blockIdx.y;
    const int it = it_off + (it_base ? *it_base : 0);
#include "dgx_tm2d_tile_fwd.inc"
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

// Cell h of the history box: its output index in the (NX-1, NY-1) gradient.
__device__ __forceinline__ long long terms_out(const Tm2dGrid& g, const Box& b, long long h)
{
    const int i = (int)(h / b.ny) + b.x0, j = (int)(h % b.ny) + b.y0;
    return (long long)i * (g.NY - 1) + j;
}

// Sum of the per-shot material terms of cell h of the history box, in shot
// order, added to the accumulators (one saved step).
template <typename ACC>
__device__ __forceinline__ void reduce_terms(const Tm2dGrid& g, const Box& b, const float* __restrict__ terms,
                                             long long h, int want_eps, int want_sig, ACC* __restrict__ geps,
                                             ACC* __restrict__ gsig)
{
    const long long out = terms_out(g, b, h);
    const long long sstride = (long long)g.nshot * b.cells;
    if (want_eps) geps[out] += sum_shots<ACC>(terms + h, b.cells, g.nshot);
    if (want_sig) gsig[out] += sum_shots<ACC>(terms + sstride + h, b.cells, g.nshot);
}

template <int ORDER, typename ACC, int TBI>
__global__ __launch_bounds__(NTA) void k_tm_step_adj(Tm2dGrid g, Tm2dFields f, TmPingPong pp, SourceBlocks src,
                                                     TmReceiversAdj rx, TmAdjGrad ag, int it)
{
    constexpr int BI = TBI;
    const int tile = step_tile(g.ftiles);
    if (tile < 0) return;
    const int s = blockIdx.y;
#include "dgx_tm2d_tile_adj.inc"
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
                                                      int gstep, const int* __restrict__ gbase)
{
    if (gbase) gstep += *gbase;  // a replayed graph: the step is the base plus the node's offset
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
    const SourceBlocks& src = I.src;
    const TmReceiversFwd& rx = I.rx;
    const HistWrite& hist = hw;
    {
#include "dgx_tm2d_tile_fwd.inc"
    }
}

template <int ORDER>
__global__ __launch_bounds__(NTA) void k_tm_batch_adj(const TmBatchInst* __restrict__ insts,
                                                      const int* __restrict__ prefix, int ninst, int total,
                                                      int gstep, const int* __restrict__ gbase)
{
    if (gbase) gstep += *gbase;  // a replayed graph: the step is the base plus the node's offset
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
    const SourceBlocks src{};
    const TmReceiversAdj& rx = I.rxa;
    using ACC = double;
    {
#include "dgx_tm2d_tile_adj.inc"
    }
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

// Zeroes up to kMaxSpans float arrays in one launch: blockIdx.y picks the
// array, a grid-stride loop over x covers it.
__global__ __launch_bounds__(256) void k_clear_spans(ClearSpans sp)
{
    const int q = blockIdx.y;
    float* p = sp.p[q];
    const long long n = sp.n[q];
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x)
        p[idx] = 0.0f;
}

void launch_clear_spans(const ClearSpans& sp, hipStream_t st)
{
    if (sp.count <= 0) return;
    long long longest = 0;
    for (int q = 0; q < sp.count; ++q) longest = std::max(longest, sp.n[q]);
    const long long blocks = std::min<long long>((longest + 255) / 256, 1024);
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_clear_spans, dim3((unsigned)blocks, (unsigned)sp.count), dim3(256), 0, st, sp);
}

void launch_tm_batch(int order, bool adjoint, const TmBatchInst* insts, const int* prefix, int ninst,
                     int total_groups, int gstep, hipStream_t st, const int* gbase)
{
    if (ninst <= 0 || total_groups <= 0) return;
    const dim3 grid((unsigned)((total_groups + 7) / 8 * 8));
    if (adjoint) {
        if (order == 8)
            hipLaunchKernelGGL(k_tm_batch_adj<8>, grid, dim3(NTA), 0, st, insts, prefix, ninst, total_groups, gstep, gbase);
        else if (order == 4)
            hipLaunchKernelGGL(k_tm_batch_adj<4>, grid, dim3(NTA), 0, st, insts, prefix, ninst, total_groups, gstep, gbase);
        else
            hipLaunchKernelGGL(k_tm_batch_adj<2>, grid, dim3(NTA), 0, st, insts, prefix, ninst, total_groups, gstep, gbase);
    } else {
        if (order == 8)
            hipLaunchKernelGGL(k_tm_batch_fwd<8>, grid, dim3(NTF), 0, st, insts, prefix, ninst, total_groups, gstep, gbase);
        else if (order == 4)
            hipLaunchKernelGGL(k_tm_batch_fwd<4>, grid, dim3(NTF), 0, st, insts, prefix, ninst, total_groups, gstep, gbase);
        else
            hipLaunchKernelGGL(k_tm_batch_fwd<2>, grid, dim3(NTF), 0, st, insts, prefix, ninst, total_groups, gstep, gbase);
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

template <int TBI>
static void launch_fwd_rows(int order, const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                            const SourceBlocks& src, const TmReceiversFwd& rx, const HistWrite& hw, int it,
                            hipStream_t st, const int* it_base)
{
    const dim3 grid = step_grid(g);
    if (order == 8)
        hipLaunchKernelGGL((k_tm_step_fwd<8, TBI>), grid, dim3(NTF), 0, st, g, f, pp, src, rx, hw, it, it_base);
    else if (order == 4)
        hipLaunchKernelGGL((k_tm_step_fwd<4, TBI>), grid, dim3(NTF), 0, st, g, f, pp, src, rx, hw, it, it_base);
    else hipLaunchKernelGGL((k_tm_step_fwd<2, TBI>), grid, dim3(NTF), 0, st, g, f, pp, src, rx, hw, it, it_base);
}

void launch_tm_step_fwd(int order, const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                        const SourceBlocks& src, const TmReceiversFwd& rx, const HistWrite& hw, int it,
                        hipStream_t st, const int* it_base)
{
    if (g.bi == kTmBITiny) launch_fwd_rows<kTmBITiny>(order, g, f, pp, src, rx, hw, it, st, it_base);
    else if (g.bi == kTmBISmall) launch_fwd_rows<kTmBISmall>(order, g, f, pp, src, rx, hw, it, st, it_base);
    else launch_fwd_rows<kTmBI>(order, g, f, pp, src, rx, hw, it, st, it_base);
}

__global__ void k_step_counter(int* __restrict__ it, int value, int add)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *it = add ? *it + value : value;
}

void launch_step_counter(int* it, int value, bool add, hipStream_t st)
{
    hipLaunchKernelGGL(k_step_counter, dim3(1), dim3(64), 0, st, it, value, add ? 1 : 0);
}

template <typename ACC, int TBI>
static void launch_adj_acc(int order, const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                           const SourceBlocks& src, const TmReceiversAdj& rx, const TmAdjGrad& ag, int it,
                           hipStream_t st)
{
    const dim3 grid = step_grid(g);
    if (order == 8)
        hipLaunchKernelGGL((k_tm_step_adj<8, ACC, TBI>), grid, dim3(NTA), 0, st, g, f, pp, src, rx, ag, it);
    else if (order == 4)
        hipLaunchKernelGGL((k_tm_step_adj<4, ACC, TBI>), grid, dim3(NTA), 0, st, g, f, pp, src, rx, ag, it);
    else hipLaunchKernelGGL((k_tm_step_adj<2, ACC, TBI>), grid, dim3(NTA), 0, st, g, f, pp, src, rx, ag, it);
}

void launch_tm_step_adj(int order, const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                        const SourceBlocks& src, const TmReceiversAdj& rx, const TmAdjGrad& ag, int it, bool fp64,
                        hipStream_t st)
{
    if (g.bi == kTmBITiny) {
        if (fp64) launch_adj_acc<double, kTmBITiny>(order, g, f, pp, src, rx, ag, it, st);
        else launch_adj_acc<float, kTmBITiny>(order, g, f, pp, src, rx, ag, it, st);
    } else if (g.bi == kTmBISmall) {
        if (fp64) launch_adj_acc<double, kTmBISmall>(order, g, f, pp, src, rx, ag, it, st);
        else launch_adj_acc<float, kTmBISmall>(order, g, f, pp, src, rx, ag, it, st);
    } else {
        if (fp64) launch_adj_acc<double, kTmBI>(order, g, f, pp, src, rx, ag, it, st);
        else launch_adj_acc<float, kTmBI>(order, g, f, pp, src, rx, ag, it, st);
    }
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
