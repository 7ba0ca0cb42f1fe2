// Argument bundles shared by the kernels and the host drivers.
#pragma once

#include "dgx_device.h"

namespace dgx {

// Source entries of one forward sweep, grouped by the block that owns the
// source cell: block_off[b] .. block_off[b+1] index entries (cell, source).
struct SourceBlocks {
    const int* block_off;
    const int2* entries;
    const float* wave;  // (nsrc, nt)
    int count;
    int nt;
    int it;
    int comp;
    float dl;   // dipole length
    float vol;  // dx * dy * dz
};

// Where the electric sweep writes E^n and R^n for the current saved step.
struct HistWrite {
    int active;          // any history written this step
    int save_e, save_r;
    int e_kind, r_kind;  // kF32 / kF16 / kBF16 element format
    void* e_base;        // component 0 of this step's frame
    void* r_base;
    long long e_cstride; // elements between components
    long long r_cstride;
    int mode;            // 3: Ex, Ey, Ez; otherwise Ez only
    Box box;
};

struct Int8Write {
    signed char* q;
    float* scale;
    int frames, frame;
    int bx, by, bz;
    int nbx, nby, nbz;
    long long tiles;
    int mode;
    Box box;
    // divisions by nbz, nby * nbz, bz and by * bz (tile counts below 2^31)
    Div div_nbz, div_tplane, div_bz, div_bplane;
};

struct ReceiverGroups {
    const int* off;     // ngroups + 1
    const int* shot;
    const long long* cell;
    const int* member;  // receiver indices in index order
    int ngroups;
};


struct MatGrad {
    const float* ca;
    const float* cb;
    const float* sig;
    int kind;
    const void* eframe;  // component 0 of E^n (non-INT8)
    const void* sframe;  // component 0 of R^n or E^(n+1)
    long long e_cstride, s_cstride;
    Int8View e8, s8;
    int e_only;
    int mode;
    int weight;
    int want_eps, want_sig;
    float dt;
    Box box;
    Div div_z, div_y;  // division by box.nz and box.ny, for boxes of fewer than 2^31 cells
};

struct StepArgs {
    const float *ca, *ccurl, *cb, *hcurl, *hrhs;
    F3 E, H;
    CF3 Ec, Hc;
    Cpml cpe, cph;
    int has_cpml;
    SourceBlocks src;
    HistWrite hist;
};

struct AdjArgs {
    const float *ca, *ccurl, *cb, *hcurl, *hrhs;
    F3 lE, lH;
    CF3 lEc, lHc;
    Cpml cpe, cph;  // lambda psi arrays + profiles
    int has_cpml;
};

// 2D Ez-TM fast path: compact (nshot, NX, NY) state, y fastest.
struct Tm2dGrid {
    int NX, NY;
    int plane;   // NX * NY
    int tiles;   // ceil(plane / 256)
    int ftiles;  // tiles of the fused step kernels: nti * ntj
    int ntj;     // kTmBJ-wide tile columns
    int bi;      // tile rows of the fused step kernels: kTmBI, kTmBISmall or kTmBITiny
    int nshot;
    float d[3];
    int pml[4];  // x0, xm, y0, ym
    int has_cpml;
};

struct Tm2dFields {
    float *Ez, *Hx, *Hy;                           // or their adjoints
    const float *ca, *ccurl, *cb, *hcurl, *hrhs;   // compact k = 0 planes
    const float* coef_e[4];
    const float* coef_h[4];
    float* psi_e[4];  // E phi2 of x0, xm, y0, ym (Ez couplings)
    float* psi_h[4];  // H phi1 of x0, xm, y0, ym (Hy, Hx couplings)
};

struct Tm2dWeights {
    float* we[4];
    float* wh[4];
};

// Fused step kernels (dgx_tm2d_fused.hip): a workgroup owns kTmBI x kTmBJ
// cells of one shot; fields and the active auxiliaries ping-pong between
// index 0 (read) and index 1 (written) every time step.
#ifndef DGX_TM_TILE_ROWS
#define DGX_TM_TILE_ROWS 16
#endif
#ifndef DGX_TM_FWD_THREADS
#define DGX_TM_FWD_THREADS 512
#endif
#ifndef DGX_TM_ADJ_THREADS
#define DGX_TM_ADJ_THREADS 512
#endif
constexpr int kTmBI = DGX_TM_TILE_ROWS;
// Single solves whose kTmBI-row tiles (all shots) number fewer than
// kTmSmallTiles use kTmBISmall-row tiles: twice the workgroups, each with a
// shorter chain, for grids that would leave most of the GPU idle. When even
// kTmBITiny-row tiles number at most kTmTinyGroups (one generation of
// resident workgroups: 256 CUs x 3), those are used.
constexpr int kTmBISmall = 8;
constexpr int kTmSmallTiles = 1024;
constexpr int kTmBITiny = 4;
constexpr int kTmTinyGroups = 768;
constexpr int kTmBJ = 64;  // one wavefront per tile row
__host__ __device__ constexpr int tm_halo(int order) { return order == 8 ? 7 : (order == 4 ? 3 : 1); }

struct TmPingPong {
    float* Ez[2];
    float* Hx[2];
    float* Hy[2];
    float* psi_e[2][4];
    float* psi_h[2][4];
};

// Receivers of the forward sweep, grouped by owning (tile, shot):
// off[t] .. off[t+1] index ent = (cell index inside the tile, receiver).
struct TmReceiversFwd {
    const int* off;
    const int2* ent;
    float* out;  // (nshot, nt, nrx)
    int nrx, nt, count;
};

// Receiver groups of the adjoint sweep for every (tile, shot) whose staged
// region (tile plus halo) holds the cell: grp = (staged index, first member,
// member count, 0); members are receiver indices in index order.
struct TmReceiversAdj {
    const int* off;
    const int4* grp;
    const int* member;
    const float* dgrad;  // (nshot, nt, nrx)
    int nrx, nt, count;
};

// Per-shot gradient terms written by the fused adjoint and summed over shots
// in order afterwards (keeps the FP32 summation order of the unfused path).
struct TmAdjGrad {
    float* terms;          // [eps | sig] x (nshot, box cells) of this step; null when not a saved step
    const float* pending;  // terms of the previous saved step, summed over shots by this launch; or null
    void* geps;            // accumulators (float or double, as the kernel's ACC)
    void* gsig;
    float* src_vals;       // (nt, nshot, nsrc); null without a source gradient
    int nsrc;
    float geo;
    int ez_comp;
    MatGrad mg;
};

// One instance of a batched 2D TM launch (dgx_batch_*). Index 0 of the
// ping-pong buffers holds the state at the start of the current phase; the
// kernels derive the step, the buffer parity, the history frame and the
// material-term buffers from the global step counter.
struct TmBatchInst {
    Tm2dGrid g;
    Tm2dFields f;         // coefficient planes and CPML profiles (field pointers unused)
    TmPingPong pp;
    SourceBlocks src;     // forward sources
    TmReceiversFwd rx;    // forward receivers
    TmReceiversAdj rxa;   // adjoint receiver groups
    void* e_saved;        // E^n frames (E-only history)
    void* e_final;        // the final E frame
    long long snap;       // elements of one frame of one component
    long long cstride;    // elements between components of the history
    int kind, mode, nt_saved;
    Box box;
    float* terms;         // two buffers of terms_len floats
    long long terms_len;
    double* geps;
    double* gsig;
    const float* sig_full;
    float dt;
    int want_eps, want_sig;
    int nt;
    int phase;            // 0 forward, 1 adjoint
    int start;            // global step at which the phase started
};

// Batched step of every listed instance of one order and phase:
// prefix[k] .. prefix[k+1] are the workgroups of instance k. The global step
// is gstep, plus *gbase when gbase is set (the launches of a replayed graph).
void launch_tm_batch(int order, bool adjoint, const TmBatchInst* insts, const int* prefix, int ninst,
                     int total_groups, int gstep, hipStream_t st, const int* gbase = nullptr);
void launch_tm_final_frame(const Tm2dGrid& g, const float* ez, const HistWrite& hw, hipStream_t st);
void launch_scale(const float* in, float* out, float factor, long long count, hipStream_t st);
// Float arrays to zero in one launch (a fill kernel, not hipMemsetAsync).
constexpr int kMaxSpans = 16;
struct ClearSpans {
    float* p[kMaxSpans];
    long long n[kMaxSpans];
    int count;
    bool add(float* ptr, long long len)
    {
        if (len <= 0) return true;
        if (count >= kMaxSpans) return false;
        p[count] = ptr;
        n[count] = len;
        ++count;
        return true;
    }
};
void launch_clear_spans(const ClearSpans& sp, hipStream_t st);

void launch_tm_step_fwd(int order, const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                        const SourceBlocks& src, const TmReceiversFwd& rx, const HistWrite& hw, int it,
                        hipStream_t st, const int* it_base = nullptr);
// *it = value, or *it += value (one thread): the step counter of a replayed graph
void launch_step_counter(int* it, int value, bool add, hipStream_t st);
void launch_tm_step_adj(int order, const Tm2dGrid& g, const Tm2dFields& f, const TmPingPong& pp,
                        const SourceBlocks& src, const TmReceiversAdj& rx, const TmAdjGrad& ag, int it, bool fp64,
                        hipStream_t st);
void launch_tm_grad_reduce(const Tm2dGrid& g, const Box& b, const float* terms, int want_eps, int want_sig,
                           void* geps, void* gsig, bool fp64, hipStream_t st);
void launch_tm_src_reduce(const float* vals, float* gsrc, int nshot, int nsrc, int nt, hipStream_t st);

void launch_tm_fwd_h(int order, const Tm2dGrid& g, const Tm2dFields& f, hipStream_t st);
void launch_tm_fwd_e(int order, const Tm2dGrid& g, const Tm2dFields& f, const SourceBlocks& src,
                     const HistWrite& hw, hipStream_t st);
void launch_tm_pack(const float* full, float* compact, long long count, bool unpack, hipStream_t st);
void launch_tm_adj_pml(bool electric, const Tm2dGrid& g, const Tm2dFields& f, const Tm2dWeights& w,
                       hipStream_t st);
void launch_tm_adj_e(int order, const Tm2dGrid& g, const Tm2dFields& f, const Tm2dWeights& w, hipStream_t st);
void launch_tm_adj_h(int order, const Tm2dGrid& g, const Tm2dFields& f, const Tm2dWeights& w, hipStream_t st);
void launch_tm_material_gradient(const Tm2dGrid& g, const Tm2dFields& f, const MatGrad& mg, int ez_comp,
                                 void* geps, void* gsig, bool fp64, hipStream_t st);

// The launches of a one-thread-per-cell sweep (block -> cell map in dgx_sweep.h).
struct CellSweep {
    CellMap cm;
    dim3 grid;         // (blocks per shot, shots)
    long long blocks;  // workgroups of the launch
};
// use_box: advance the interior box with its own plain-cell path.
CellSweep cell_sweep(const Grid& g, int order, bool use_box = true);
// Workgroup (linear index in the launch) that advances cell (i, j, k) of shot s.
long long cell_block(const CellSweep& cs, const Grid& g, int i, int j, int k, int s);
void launch_coefficients(const Grid& g, const float* eps, const float* sig, const float* mu, float* const co[6],
                         float dt, hipStream_t st);
void launch_fwd_h(int order, const Grid& g, const StepArgs& a, hipStream_t st);
void launch_fwd_e(int order, const Grid& g, const StepArgs& a, hipStream_t st);
// Receiver gather from a field with strides (shot, x, y); z stride is 1.
void launch_sample(const Grid& g, const float* f, const int* loc, float* out, int nrx, int nt, int it,
                   hipStream_t st, long long shot_stride = -1, long long x_stride = -1, long long y_stride = -1);
void launch_save_frame(const Grid& g, CF3 E, const HistWrite& hw, hipStream_t st);
void launch_quant_int8(const Grid& g, CF3 E, const float* src, long long src_cstride, const Int8Write& w,
                       hipStream_t st);
void launch_adj_receivers(const Grid& g, const ReceiverGroups& rg, float* lam, const float* dgrad, int nrx,
                          int nt, int it, hipStream_t st);
void launch_adj_source(const Grid& g, const int* loc, const float* cb, const float* lam, float* gsrc, int nsrc,
                       int nt, int it, float geo, hipStream_t st);
void launch_material_gradient(const Grid& g, CF3 L, const MatGrad& mg, void* geps, void* gsig, bool fp64,
                              hipStream_t st);
void launch_add_f64(const double* a, float* out, long long count, hipStream_t st);
// two-pass adjoint: per-cell channel weights (6 x nshot x cells floats), then gathers
void launch_adj_weights(int order, bool electric, const Grid& g, const AdjArgs& a, float* W, hipStream_t st);
void launch_adj_gather_e(int order, const Grid& g, const AdjArgs& a, const float* W, hipStream_t st);
void launch_adj_gather_h(int order, const Grid& g, const AdjArgs& a, const float* W, hipStream_t st);
// the E-side gather fused with the H-side weights: W_E -> lambda_H -> W_H
void launch_adj_gather_e_w(int order, const Grid& g, const AdjArgs& a, const float* We, float* Wh, hipStream_t st);

// x-marching 3D sweeps (dgx_xmarch.hip): a workgroup owns a (y, z) tile of
// `ty` rows (one wavefront each) x 64 cells of one shot and walks `xc`
// x-planes; tiles are numbered (y tile) * ntz + (z tile), in XCD-contiguous
// launch order when `remap` is set.
struct XmTiles {
    int nty, ntz;
    int ty;
    int xc;
    int remap;
};
// What the sweeps' out-of-line CPML code reads, in device memory (one copy per
// solve): the grid, both CPML sets and the update coefficients.
struct XmConst {
    Grid g;
    Cpml cpe, cph;
    const float *ca, *ccurl, *cb, *hcurl, *hrhs;
};
XmConst xm_const(const Grid& g, const Cpml& cpe, const Cpml& cph, const float* ca, const float* ccurl,
                 const float* cb, const float* hcurl, const float* hrhs);
// Tile rows (4 or 8), planes per chunk and remap of the 3D sweeps;
// DEEPGPR_AMD_XM_TILE_ROWS, DEEPGPR_AMD_XM_PLANES and DEEPGPR_AMD_XM_REMAP
// override the defaults.
int xm_tile_rows(int order);
int xm_chunk_planes();
int xm_remap();
XmTiles xm_tiles(const Grid& g, int order);
int xm_tile_of(const XmTiles& xt, int j, int k);
void launch_xm_fwd_h(int order, const Grid& g, const XmTiles& xt, const XmConst* kc, const StepArgs& a,
                     hipStream_t st);
// sources: block_off indexed by shot * (nty * ntz) + tile, entries in plane
// order and, within a plane, source order
void launch_xm_fwd_e(int order, const Grid& g, const XmTiles& xt, const XmConst* kc, const StepArgs& a,
                     hipStream_t st);
// electric: lambda_H += D-^T W_E (with `defer`, first the lambda psi_H update
// owed by the previous step); otherwise lambda_E = ca lambda_E + D+^T W_H,
// advancing lambda psi_E
void launch_xm_adj(int order, bool electric, const Grid& g, const XmTiles& xt, const XmConst* kc,
                   const AdjArgs& a, bool defer, hipStream_t st);
// the lambda psi_H update owed by the last adjoint step
void launch_xm_psi_h_final(const Grid& g, const XmConst* kc, const AdjArgs& a, hipStream_t st);

// Kernel ids of the optional per-launch timing (dgx_set_kernel_timing).
enum KernelId : int {
    kKCoeff = 0, kKFwdH, kKFwdE, kKSample, kKQuant, kKSaveFrame, kKAdjRx, kKAdjSrc, kKMatGrad, kKAdjPmlE,
    kKAdjE, kKAdjPmlH, kKAdjH, kKTmStepFwd, kKTmStepAdj, kKTmGradReduce, kKTmSrcReduce, kKAdjWeightsE,
    kKAdjWeightsH, kKAdjEW, kKXmFwdH, kKXmFwdE, kKXmAdjE, kKXmAdjH, kKXmPsiH, kKCount
};

}  // namespace dgx
