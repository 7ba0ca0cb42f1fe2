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
    Cpml we, wh;    // CPML weight scratch shaped like the psi arrays
    int has_cpml;
};

// 2D Ez-TM fast path: compact (nshot, NX, NY) state, y fastest.
struct Tm2dGrid {
    int NX, NY;
    int plane;   // NX * NY
    int tiles;   // ceil(plane / 256)
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

dim3 cell_grid(const Grid& g);
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
void launch_adj_pml(bool electric, const Grid& g, const AdjArgs& a, hipStream_t st);
void launch_adj_e(int order, const Grid& g, const AdjArgs& a, hipStream_t st);
void launch_adj_h(int order, const Grid& g, const AdjArgs& a, hipStream_t st);

// Kernel ids of the optional per-launch timing (dgx_set_kernel_timing).
enum KernelId : int {
    kKCoeff = 0, kKFwdH, kKFwdE, kKSample, kKQuant, kKSaveFrame, kKAdjRx, kKAdjSrc, kKMatGrad, kKAdjPmlE,
    kKAdjE, kKAdjPmlH, kKAdjH, kKCount
};

}  // namespace dgx
