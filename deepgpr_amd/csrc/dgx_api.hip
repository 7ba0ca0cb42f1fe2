// Host drivers of the forward and adjoint time loops and the exported C ABI.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <mutex>
#include <utility>
#include <vector>

#include "deepgpr_amd.h"
#include "dgx_internal.h"
#include "dgx_retire.h"

using namespace dgx;

namespace {

thread_local char g_error[512];
int g_abi6_order = 2;  // set_fdtd_order(), read by the ABI-6 adapters only

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

// Scratch of dgx_forward_noalloc: a file-scope device array (each GPU the
// library runs on holds its own copy until the process ends). Host code uses
// the address hipGetSymbolAddress returns for the current device, never the
// array's name. Calls that overlap on one GPU share it and race.
constexpr size_t kNoAllocScratchBytes = 4u << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up once per device).
void* noalloc_scratch()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("dgx_forward_noalloc: no usable current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !ptr) {
            set_error("dgx_forward_noalloc: hipGetSymbolAddress of the scratch array failed: %s",
                      hipGetErrorString(rc));
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// Device allocation released at scope exit (after the stream has drained).
// Stream of the solve running on this thread. Scratch is allocated and freed
// stream-ordered on it, so concurrent solves on different streams never wait
// for each other (hipFree would synchronise the whole device).
thread_local hipStream_t g_call_stream = nullptr;

struct CallStream {
    hipStream_t prev;
    explicit CallStream(hipStream_t s) : prev(g_call_stream) { g_call_stream = s; }
    ~CallStream() { g_call_stream = prev; }
};

struct DevBuf {
    void* p = nullptr;
    hipStream_t st = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset()
    {
        if (p) (void)hipFreeAsync(p, st);
        p = nullptr;
    }
    bool alloc(size_t bytes)
    {
        if (bytes == 0) bytes = 16;
        st = g_call_stream;
        return hipMallocAsync(&p, bytes, st) == hipSuccess;
    }
    // The free follows the work queued on `s` instead: the stream that uses
    // the buffer last (a free on any other stream could return the memory
    // while kernels queued on `s` still read it).
    void free_after(hipStream_t s) { st = s; }
};

#define DGX_TRY(call, what)                                                         \
    do {                                                                            \
        hipError_t err_ = (call);                                                   \
        if (err_ != hipSuccess) {                                                   \
            set_error("%s: %s", (what), hipGetErrorString(err_));                   \
            return -1;                                                              \
        }                                                                           \
    } while (0)

// ---------------------------------------------------------------------------
// Optional per-launch timing: HIP events recorded on the launch stream around
// each launch of the hot kernels, summed per kernel after the call's final
// synchronisation. Off by default; enabled by dgx_set_kernel_timing(1).
// ---------------------------------------------------------------------------
std::mutex g_timing_mu;
int g_timing_on = 0;
double g_timing_ms[kKCount];
long long g_timing_count[kKCount];
const char* const kKernelNames[kKCount] = {
    "coefficients", "fwd_h", "fwd_e", "sample", "quant_int8", "save_frame", "adj_receivers",
    "adj_source", "material_gradient", "adj_pml_e", "adj_e", "adj_pml_h", "adj_h", "tm_step_fwd",
    "tm_step_adj", "tm_grad_reduce", "tm_src_reduce", "adj_weights_e", "adj_weights_h", "adj_e_w",
    "xm_fwd_h", "xm_fwd_e", "xm_adj_e", "xm_adj_h", "xm_psi_h"};

struct LaunchTimer {
    bool on;
    hipStream_t stream;
    struct Rec {
        int id;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    explicit LaunchTimer(hipStream_t s) : stream(s)
    {
        std::lock_guard<std::mutex> lock(g_timing_mu);
        on = g_timing_on != 0;
    }
    ~LaunchTimer()
    {
        for (Rec& r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
    int begin(int id)
    {
        if (!on) return -1;
        Rec r{id, nullptr, nullptr};
        if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
        (void)hipEventRecord(r.a, stream);
        recs.push_back(r);
        return (int)recs.size() - 1;
    }
    void end(int h)
    {
        if (h >= 0) (void)hipEventRecord(recs[h].b, stream);
    }
    // call after the stream has been synchronised
    void collect()
    {
        if (!on) return;
        std::lock_guard<std::mutex> lock(g_timing_mu);
        for (Rec& r : recs) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                g_timing_ms[r.id] += ms;
                g_timing_count[r.id] += 1;
            }
        }
    }
};

#define TIMED(timer, id, call)                 \
    do {                                       \
        int th_ = (timer).begin(id);           \
        call;                                  \
        (timer).end(th_);                      \
    } while (0)

constexpr int kFlagRhsFromE = 1 << 4;
constexpr int kFlagPhysical = 1 << 5;
constexpr int kFlagNoHistory = 1 << 26;

struct Plan {
    Grid g;
    int order;
    int kind;
    bool history;
    bool e_only;
    bool store_r;
    bool physical;
    int tile[3];
    Box box;
    int comps;
    int nt_saved;
    long long snap;     // elements of one frame of one component (all shots)
    long long cstride;  // nt_saved * snap
    bool has_cpml;
    Cpml cpe, cph;
};

void box_axis(bool physical, int cells, int lo, int hi, int tile, int* o, int* e)
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

void host_psi_dims(const Grid& g, bool electric, int face, int w, int dims[3])
{
    static const int target[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    static const int source[3][2] = {{2, 1}, {2, 0}, {1, 0}};
    int axis = face / 2, p = g.pml[face];
    for (int a = 0; a < 3; ++a) dims[a] = g.N[a];
    if (electric) {
        dims[target[axis][w]] -= 1;
        dims[axis] = p + 1;
    } else {
        dims[source[axis][w]] -= 1;
        dims[axis] = p;
    }
}

long long psi_elems(const Grid& g, bool electric, int face, int w)
{
    if (g.pml[face] <= 0) return 0;
    int d[3];
    host_psi_dims(g, electric, face, w, d);
    return (long long)g.nshot * d[0] * d[1] * d[2];
}

int make_plan(const dgx_solve* P, bool backward, Plan& pl)
{
    if (!P) {
        set_error("null solve descriptor");
        return -1;
    }
    if (P->abi_version != DGX_ABI_VERSION) {
        set_error("descriptor abi_version %d != %d", P->abi_version, DGX_ABI_VERSION);
        return -1;
    }
    Grid& g = pl.g;
    g.N[0] = P->nx_fields;
    g.N[1] = P->ny_fields;
    g.N[2] = P->nz_fields;
    if (g.N[0] < 2 || g.N[1] < 2 || g.N[2] < 2) {
        set_error("field sizes must be >= 2 (got %d, %d, %d)", g.N[0], g.N[1], g.N[2]);
        return -1;
    }
    if ((long long)g.N[1] * g.N[2] > (1LL << 30)) {
        set_error("y-z plane of %lld cells is too large", (long long)g.N[1] * g.N[2]);
        return -1;
    }
    if (P->nshot < 1 || P->nshot > 65535) {
        set_error("nshot must be in [1, 65535] (got %d)", P->nshot);
        return -1;
    }
    if (P->nt < 0 || P->sampling_interval < 1) {
        set_error("nt must be >= 0 and sampling_interval >= 1");
        return -1;
    }
    g.st0 = (long long)g.N[1] * g.N[2];
    g.plane = g.N[1] * g.N[2];
    g.cells = (long long)g.N[0] * g.st0;
    g.nshot = P->nshot;
    g.d[0] = P->dx;
    g.d[1] = P->dy;
    g.d[2] = P->dz;
    pl.has_cpml = false;
    for (int f = 0; f < 6; ++f) {
        g.pml[f] = P->pml[f];
        if (g.pml[f] < 0) {
            set_error("negative PML thickness on face %d", f);
            return -1;
        }
        if (g.pml[f]) pl.has_cpml = true;
    }
    for (int a = 0; a < 3; ++a) {
        if (g.pml[2 * a] + g.pml[2 * a + 1] > g.N[a] - 2) {
            set_error("PML thickness %d + %d leaves no physical cell on axis %d (field size %d)", g.pml[2 * a],
                      g.pml[2 * a + 1], a, g.N[a]);
            return -1;
        }
    }
    pl.order = (P->fdtd_order == 4 || P->fdtd_order == 8) ? P->fdtd_order : 2;
    int st = P->storage_type;
    pl.kind = st & 0x0f;
    if (pl.kind > kInt8) {
        set_error("unknown wavefield storage kind %d", pl.kind);
        return -1;
    }
    bool want_mat = P->eps_r_requires_grad || P->sigma_requires_grad;
    if (!backward) {
        pl.history = (st & kFlagNoHistory) == 0;
        pl.e_only = pl.history && P->save_model_history && (st & kFlagRhsFromE);
        pl.store_r = pl.history && P->save_model_history && !pl.e_only;
    } else {
        pl.history = want_mat;
        pl.e_only = want_mat && (st & kFlagRhsFromE);
        pl.store_r = !pl.e_only;
    }
    pl.physical = (st & kFlagPhysical) != 0;
    if (pl.kind == kInt8) {
        pl.tile[0] = (st >> 8) & 0x3f;
        pl.tile[1] = (st >> 14) & 0x3f;
        pl.tile[2] = (st >> 20) & 0x3f;
        for (int a = 0; a < 3; ++a)
            if (pl.tile[a] <= 0) {
                set_error("INT8 storage_type carries an empty block on axis %d", a);
                return -1;
            }
        int vol = pl.tile[0] * pl.tile[1] * pl.tile[2];
        if (vol > 256 || (vol & (vol - 1))) {
            set_error("INT8 block volume %d must be a power of two <= 256", vol);
            return -1;
        }
    } else {
        pl.tile[0] = pl.tile[1] = pl.tile[2] = 1;
    }
    Box& b = pl.box;
    box_axis(pl.physical, g.N[0] - 1, g.pml[0], g.pml[1], pl.tile[0], &b.x0, &b.nx);
    box_axis(pl.physical, g.N[1] - 1, g.pml[2], g.pml[3], pl.tile[1], &b.y0, &b.ny);
    box_axis(pl.physical, g.N[2] - 1, g.pml[4], g.pml[5], pl.tile[2], &b.z0, &b.nz);
    b.cells = (long long)b.nx * b.ny * b.nz;
    pl.comps = P->fwi_mode == 3 ? 3 : 1;
    pl.nt_saved = (P->nt + P->sampling_interval - 1) / P->sampling_interval;
    pl.snap = (long long)g.nshot * b.cells;
    pl.cstride = (long long)pl.nt_saved * pl.snap;

    const char* who = backward ? "backward" : "forward";
    bool needs_mat = backward ? want_mat : (pl.history && P->save_model_history);
    if (needs_mat && pl.has_cpml && pl.physical) {
        set_error("%s: material gradients with PML require extended wavefield history; physical history omits "
                  "PML material sensitivities", who);
        return -1;
    }
    if (pl.e_only && P->sampling_interval != 1) {
        set_error("%s: E-only wavefield history requires sampling interval 1", who);
        return -1;
    }
    for (int f = 0; f < 6; ++f) {
        pl.cpe.coef[f] = P->pml_coeff[f];
        pl.cph.coef[f] = P->pml_coeff[6 + f];
        pl.cpe.psi[f][0] = P->phi[4 * f + 0];
        pl.cpe.psi[f][1] = P->phi[4 * f + 1];
        pl.cph.psi[f][0] = P->phi[4 * f + 2];
        pl.cph.psi[f][1] = P->phi[4 * f + 3];
        if (g.pml[f] > 0 && (!pl.cpe.coef[f] || !pl.cph.coef[f] || !pl.cpe.psi[f][0] || !pl.cpe.psi[f][1] ||
                             !pl.cph.psi[f][0] || !pl.cph.psi[f][1])) {
            set_error("%s: CPML face %d has missing arrays", who, f);
            return -1;
        }
    }
    for (int c = 0; c < 3; ++c)
        if (!P->E[c] || !P->H[c]) {
            set_error("%s: null field pointer", who);
            return -1;
        }
    for (int c = 0; c < 6; ++c)
        if (!P->update_coeff[c]) {
            set_error("%s: null update-coefficient scratch", who);
            return -1;
        }
    return 0;
}

bool loc_ok(const Grid& g, const int* L)
{
    return L[0] >= 0 && L[0] < g.N[0] && L[1] >= 0 && L[1] < g.N[1] && L[2] >= 0 && L[2] < g.N[2];
}

size_t elem_bytes(int kind) { return kind == kF32 ? 4 : (kind == kInt8 ? 1 : 2); }

struct Int8Geom {
    int nbx, nby, nbz;
    long long tiles;
};

Int8Geom int8_geom(const Plan& pl)
{
    Int8Geom q;
    q.nbx = (pl.box.nx + pl.tile[0] - 1) / pl.tile[0];
    q.nby = (pl.box.ny + pl.tile[1] - 1) / pl.tile[1];
    q.nbz = (pl.box.nz + pl.tile[2] - 1) / pl.tile[2];
    q.tiles = (long long)q.nbx * q.nby * q.nbz;
    return q;
}

long long aligned4(long long v) { return (v + 3) & ~3LL; }

Int8Write int8_writer(const Plan& pl, void* packed, int frames, int frame)
{
    Int8Geom q = int8_geom(pl);
    Int8Write w;
    w.q = (signed char*)packed;
    w.scale = (float*)((unsigned char*)packed + aligned4((long long)pl.comps * frames * pl.snap));
    w.frames = frames;
    w.frame = frame;
    w.bx = pl.tile[0];
    w.by = pl.tile[1];
    w.bz = pl.tile[2];
    w.nbx = q.nbx;
    w.nby = q.nby;
    w.nbz = q.nbz;
    w.tiles = q.tiles;
    w.mode = pl.comps == 3 ? 3 : 2;
    w.box = pl.box;
    w.div_nbz = make_div((unsigned)q.nbz);
    w.div_tplane = make_div((unsigned)(q.nby * q.nbz));
    w.div_bz = make_div((unsigned)w.bz);
    w.div_bplane = make_div((unsigned)(w.by * w.bz));
    return w;
}

Int8View int8_view(const Plan& pl, const void* packed, int frames, int frame)
{
    Int8Geom q = int8_geom(pl);
    Int8View v;
    v.q = (const signed char*)packed + (long long)frame * pl.snap;
    v.q_cstride = (long long)frames * pl.snap;
    v.scale = (const float*)((const unsigned char*)packed + aligned4((long long)pl.comps * frames * pl.snap)) +
              (long long)frame * pl.g.nshot * q.tiles;
    v.s_cstride = (long long)frames * pl.g.nshot * q.tiles;
    v.bx = pl.tile[0];
    v.by = pl.tile[1];
    v.bz = pl.tile[2];
    v.nby = q.nby;
    v.nbz = q.nbz;
    v.tiles = q.tiles;
    return v;
}

constexpr int kFlagTm2d = 1 << 6;

// The 2D TM fast path serves storage_type's SOLVER_TM2D requests on nz = 1
// grids with an Ez source whose sources and receivers sit on the k = 0 layer
// (the caller guarantees zero Ex/Ey/Hz states, deepgpr.h:77-83); everything
// else runs the general path, which gives the same results.
bool tm2d_applies(const dgx_solve* P, const Plan& pl, const std::vector<int>& sloc, const std::vector<int>& rloc)
{
    if (!(P->storage_type & kFlagTm2d) || pl.g.N[2] != 2 || P->source_component != 2 || pl.kind == kInt8) return false;
    for (size_t q = 2; q < sloc.size(); q += 3)
        if (sloc[q] != 0) return false;
    for (size_t q = 2; q < rloc.size(); q += 3)
        if (rloc[q] != 0) return false;
    return true;
}

// Forward-only fused 2D TM solves replay their step loop as a graph of
// kGraphSteps step launches (DEEPGPR_AMD_TM2D_GRAPHS=0 turns this off).
constexpr int kGraphSteps = 32;  // even: a replay leaves the ping-pong parity unchanged

bool tm2d_graphs()
{
    static const bool on = [] {
        const char* v = std::getenv("DEEPGPR_AMD_TM2D_GRAPHS");
        return !(v && std::strcmp(v, "0") == 0);
    }();
    return on;
}

// Graphs are captured on a stream of the calling thread (capture is not
// possible on the null stream) and launched on the caller's stream.
hipStream_t capture_stream()
{
    thread_local hipStream_t streams[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (!streams[dev] && hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess)
        streams[dev] = nullptr;
    return streams[dev];
}

// An executable graph is destroyed once the event recorded after its last
// launch has completed.
struct RetiredGraph {
    hipGraphExec_t exec;
    hipGraph_t graph;
    hipEvent_t done;
};
std::mutex g_graph_mu;
std::vector<RetiredGraph> g_graphs;

void sweep_graphs()
{
    std::lock_guard<std::mutex> lock(g_graph_mu);
    for (size_t q = 0; q < g_graphs.size();) {
        if (hipEventQuery(g_graphs[q].done) != hipSuccess) {
            ++q;
            continue;
        }
        (void)hipGraphExecDestroy(g_graphs[q].exec);
        (void)hipGraphDestroy(g_graphs[q].graph);
        (void)hipEventDestroy(g_graphs[q].done);
        g_graphs.erase(g_graphs.begin() + (long)q);
    }
    (void)hipGetLastError();  // a pending event's "not ready" is no error of the caller's
}

// The split per-operator TM kernels stay selectable for comparison
// (DEEPGPR_AMD_TM2D_KERNELS=split); the fused step kernels are the default.
bool tm2d_fused()
{
    static const bool fused = [] {
        const char* v = std::getenv("DEEPGPR_AMD_TM2D_KERNELS");
        return !(v && std::strcmp(v, "split") == 0);
    }();
    return fused;
}

// The 3D adjoint runs the E-side gather fused with the H-side weights;
// DEEPGPR_AMD_3D_ADJOINT=split keeps the separate weight kernel (same bits).
bool adj3d_fused()
{
    static const bool fused = [] {
        const char* v = std::getenv("DEEPGPR_AMD_3D_ADJOINT");
        return !(v && std::strcmp(v, "split") == 0);
    }();
    return fused;
}

// 3D (and other non-TM) solves run the one-thread-per-cell sweeps;
// DEEPGPR_AMD_3D_SWEEP=xmarch selects the x-marching sweeps (same bits). Read
// at every solve, so one process can run both.
bool xmarch3d()
{
    const char* v = std::getenv("DEEPGPR_AMD_3D_SWEEP");
    return v && std::strcmp(v, "xmarch") == 0;
}

// Static rule for the tile rows of a single fused TM solve. Small solves
// leave most of the GPU idle and their time is the chain of one workgroup,
// which shorter tiles cut: kTmBITiny when its kTmBITiny-row tiles over all
// shots fit in one generation of resident workgroups (kTmTinyGroups), else
// kTmBISmall when the kTmBI-row tiles number fewer than kTmSmallTiles, else
// kTmBI.
int tm2d_tile_rows_rule(const Plan& pl)
{
    const long long cols = (pl.g.N[1] + kTmBJ - 1) / kTmBJ;
    const auto tiles = [&](int bi) { return (long long)((pl.g.N[0] + bi - 1) / bi) * cols * pl.g.nshot; };
    if (tiles(kTmBITiny) <= kTmTinyGroups) return kTmBITiny;
    return tiles(kTmBI) < kTmSmallTiles ? kTmBISmall : kTmBI;
}

Tm2dGrid tm_grid_of(const Plan& pl, int bi);

// Run-time tuner of the tile rows: the first solve of each shape (grid,
// shots, order, CPML) times the fused forward step kernel at 4, 8 and 16
// rows on buffers of its own (zeroed fields and CPML auxiliaries, the solve's
// coefficients, no sources, receivers or history: 2 warm-up and 6 timed
// steps each), and the fastest is kept for every later solve of that shape
// in the process, forward and adjoint. The tile rows change the work
// decomposition only, never the results. Shapes whose probe buffers would
// exceed kTuneMaxBytes, and every shape with DEEPGPR_AMD_TM2D_TUNE=0, use the
// static rule.
constexpr size_t kTuneMaxBytes = (size_t)512 << 20;
struct TuneKey {
    int n0, n1, nshot, order, cpml;
    bool operator==(const TuneKey& o) const
    {
        return n0 == o.n0 && n1 == o.n1 && nshot == o.nshot && order == o.order && cpml == o.cpml;
    }
};
std::mutex g_tune_mu;
std::vector<std::pair<TuneKey, int>> g_tuned;

int tm2d_probe(const Plan& pl, const dgx_solve* P, hipStream_t stream, double* us_out)
{
    const int cand[3] = {kTmBITiny, kTmBISmall, kTmBI};
    Tm2dGrid g0 = tm_grid_of(pl, kTmBI);
    const long long field = (long long)g0.nshot * g0.plane;
    long long psi_len = 0;
    long long len_e[4], len_h[4];
    for (int f4 = 0; f4 < 4; ++f4) {
        const long long across = (long long)g0.nshot * (f4 < 2 ? g0.NY : g0.NX);
        len_e[f4] = g0.pml[f4] > 0 ? across * (g0.pml[f4] + 1) : 0;
        len_h[f4] = g0.pml[f4] > 0 ? across * g0.pml[f4] : 0;
        psi_len += 2 * (len_e[f4] + len_h[f4]);
    }
    const size_t bytes = ((size_t)6 * field + 5 * (size_t)g0.plane + (size_t)psi_len) * sizeof(float);
    if (bytes > kTuneMaxBytes) return -1;
    DevBuf buf;
    if (!buf.alloc(bytes)) return -1;
    if (hipMemsetAsync(buf.p, 0, bytes, stream) != hipSuccess) return -1;
    float* b = (float*)buf.p;
    Tm2dFields f{};
    TmPingPong pp{};
    for (int q = 0; q < 2; ++q) {
        pp.Ez[q] = b + (3 * q) * field;
        pp.Hx[q] = b + (3 * q + 1) * field;
        pp.Hy[q] = b + (3 * q + 2) * field;
    }
    float* co = b + 6 * field;
    const float* full[5] = {P->update_coeff[0], P->update_coeff[1], P->update_coeff[2], P->update_coeff[4],
                            P->update_coeff[5]};
    for (int c = 0; c < 5; ++c) launch_tm_pack(full[c], co + (long long)c * g0.plane, g0.plane, false, stream);
    f.Ez = pp.Ez[0];
    f.Hx = pp.Hx[0];
    f.Hy = pp.Hy[0];
    f.ca = co;
    f.ccurl = co + g0.plane;
    f.cb = co + 2 * (long long)g0.plane;
    f.hcurl = co + 3 * (long long)g0.plane;
    f.hrhs = co + 4 * (long long)g0.plane;
    float* a = co + 5 * (long long)g0.plane;
    for (int f4 = 0; f4 < 4; ++f4) {
        f.coef_e[f4] = P->pml_coeff[f4];
        f.coef_h[f4] = P->pml_coeff[6 + f4];
        for (int q = 0; q < 2; ++q) {
            pp.psi_e[q][f4] = len_e[f4] ? a : nullptr;
            a += len_e[f4];
            pp.psi_h[q][f4] = len_h[f4] ? a : nullptr;
            a += len_h[f4];
        }
        f.psi_e[f4] = pp.psi_e[0][f4];
        f.psi_h[f4] = pp.psi_h[0][f4];
    }
    SourceBlocks src{};
    TmReceiversFwd rx{};
    HistWrite none{};
    none.mode = pl.comps == 3 ? 3 : 2;
    none.box = pl.box;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int best = -1;
    double best_us = 0.0;
    bool ok = hipEventCreate(&ev[0]) == hipSuccess && hipEventCreate(&ev[1]) == hipSuccess;
    for (int q = 0; q < 3 && ok; ++q) {
        const Tm2dGrid g = tm_grid_of(pl, cand[q]);
        TmPingPong cur = pp;
        auto step = [&](int it) {
            launch_tm_step_fwd(pl.order, g, f, cur, src, rx, none, it, stream);
            std::swap(cur.Ez[0], cur.Ez[1]);
            std::swap(cur.Hx[0], cur.Hx[1]);
            std::swap(cur.Hy[0], cur.Hy[1]);
            for (int f4 = 0; f4 < 4; ++f4) {
                std::swap(cur.psi_e[0][f4], cur.psi_e[1][f4]);
                std::swap(cur.psi_h[0][f4], cur.psi_h[1][f4]);
            }
        };
        step(0);
        step(1);
        ok = hipEventRecord(ev[0], stream) == hipSuccess;
        for (int it = 2; it < 8 && ok; ++it) step(it);
        ok = ok && hipEventRecord(ev[1], stream) == hipSuccess && hipEventSynchronize(ev[1]) == hipSuccess;
        float ms = 0.0f;
        ok = ok && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess;
        if (ok && (best < 0 || ms * 1e3 / 6 < best_us)) {
            best = cand[q];
            best_us = ms * 1e3 / 6;
        }
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (!ok || hipGetLastError() != hipSuccess) return -1;
    if (us_out) *us_out = best_us;
    return best;
}

// Tile rows of a single fused TM solve: DEEPGPR_AMD_TM2D_TILE_ROWS=4, 8 or 16
// forces one (read on every call, so one process can run all three); else
// the tuned choice for the solve's shape.
int tm2d_tile_rows(const Plan& pl, const dgx_solve* P, hipStream_t stream)
{
    const char* v = std::getenv("DEEPGPR_AMD_TM2D_TILE_ROWS");
    if (v && std::strcmp(v, "4") == 0) return kTmBITiny;
    if (v && std::strcmp(v, "8") == 0) return kTmBISmall;
    if (v && std::strcmp(v, "16") == 0) return kTmBI;
    const char* t = std::getenv("DEEPGPR_AMD_TM2D_TUNE");
    if ((t && std::strcmp(t, "0") == 0) || !P) return tm2d_tile_rows_rule(pl);
    const TuneKey key{pl.g.N[0], pl.g.N[1], pl.g.nshot, pl.order, pl.has_cpml ? 1 : 0};
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        for (const auto& kv : g_tuned)
            if (kv.first == key) return kv.second;
    }
    int bi = tm2d_probe(pl, P, stream, nullptr);
    if (bi < 0) {
        (void)hipGetLastError();
        bi = tm2d_tile_rows_rule(pl);
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned.push_back({key, bi});
    return bi;
}

Tm2dGrid tm_grid_of(const Plan& pl, int bi)
{
    Tm2dGrid g{};
    g.NX = pl.g.N[0];
    g.NY = pl.g.N[1];
    g.plane = g.NX * g.NY;
    g.tiles = (g.plane + 255) / 256;
    g.ntj = (g.NY + kTmBJ - 1) / kTmBJ;
    g.bi = bi;
    g.ftiles = ((g.NX + bi - 1) / bi) * g.ntj;
    g.nshot = pl.g.nshot;
    for (int a = 0; a < 3; ++a) g.d[a] = pl.g.d[a];
    for (int f4 = 0; f4 < 4; ++f4) g.pml[f4] = pl.g.pml[f4];
    g.has_cpml = pl.has_cpml;
    return g;
}

int tm_tile_of(const Tm2dGrid& g, int i, int j) { return (i / g.bi) * g.ntj + j / kTmBJ; }
int tm_local(const Tm2dGrid& g, int i, int j) { return (i % g.bi) * kTmBJ + j % kTmBJ; }

// Compact buffers of the 2D TM path: 3 fields (nshot * plane; 6 with the
// fused kernels' ping-pong copies) and 5 coefficient planes, plus the second
// copies of the 8 active CPML auxiliaries.
struct Tm2dState {
    Tm2dGrid g;
    Tm2dFields f;
    TmPingPong pp{};
    bool fused = false;
    float* home_e[4] = {};
    float* home_h[4] = {};
    long long psi_len_e[4] = {}, psi_len_h[4] = {};
    DevBuf buf, alt;
    int setup(const Plan& pl, const dgx_solve* P, hipStream_t stream, bool use_fused, int bi)
    {
        fused = use_fused;
        g = tm_grid_of(pl, bi);
        const long long field = (long long)g.nshot * g.plane;
        const int nf = fused ? 6 : 3;
        if (!buf.alloc(((size_t)nf * field + 5 * (size_t)g.plane) * sizeof(float))) return -1;
        float* b = (float*)buf.p;
        f.Ez = b;
        f.Hx = b + field;
        f.Hy = b + 2 * field;
        float* co = b + nf * field;
        const float* full[5] = {P->update_coeff[0], P->update_coeff[1], P->update_coeff[2], P->update_coeff[4],
                                P->update_coeff[5]};
        for (int c = 0; c < 5; ++c) launch_tm_pack(full[c], co + (long long)c * g.plane, g.plane, false, stream);
        f.ca = co;
        f.ccurl = co + g.plane;
        f.cb = co + 2 * (long long)g.plane;
        f.hcurl = co + 3 * (long long)g.plane;
        f.hrhs = co + 4 * (long long)g.plane;
        long long alt_len = 0;
        for (int f4 = 0; f4 < 4; ++f4) {
            f.coef_e[f4] = P->pml_coeff[f4];
            f.coef_h[f4] = P->pml_coeff[6 + f4];
            f.psi_e[f4] = P->phi[4 * f4 + 1];
            f.psi_h[f4] = P->phi[4 * f4 + 2];
            home_e[f4] = f.psi_e[f4];
            home_h[f4] = f.psi_h[f4];
            const long long across = (long long)g.nshot * (f4 < 2 ? g.NY : g.NX);
            psi_len_e[f4] = g.pml[f4] > 0 ? across * (g.pml[f4] + 1) : 0;
            psi_len_h[f4] = g.pml[f4] > 0 ? across * g.pml[f4] : 0;
            alt_len += psi_len_e[f4] + psi_len_h[f4];
        }
        launch_tm_pack(P->E[2], f.Ez, field, false, stream);
        launch_tm_pack(P->H[0], f.Hx, field, false, stream);
        launch_tm_pack(P->H[1], f.Hy, field, false, stream);
        if (!fused) return 0;
        pp.Ez[0] = f.Ez;
        pp.Hx[0] = f.Hx;
        pp.Hy[0] = f.Hy;
        pp.Ez[1] = b + 3 * field;
        pp.Hx[1] = b + 4 * field;
        pp.Hy[1] = b + 5 * field;
        if (alt_len > 0 && !alt.alloc((size_t)alt_len * sizeof(float))) return -1;
        float* a = (float*)alt.p;
        // The second copies start equal to the first: auxiliaries no cell
        // updates (the deepest electric layer) keep their value in both.
        for (int f4 = 0; f4 < 4; ++f4) {
            pp.psi_e[0][f4] = home_e[f4];
            pp.psi_h[0][f4] = home_h[f4];
            pp.psi_e[1][f4] = psi_len_e[f4] ? a : nullptr;
            if (psi_len_e[f4] &&
                hipMemcpyAsync(a, home_e[f4], psi_len_e[f4] * 4, hipMemcpyDeviceToDevice, stream) != hipSuccess)
                return -1;
            a += psi_len_e[f4];
            pp.psi_h[1][f4] = psi_len_h[f4] ? a : nullptr;
            if (psi_len_h[f4] &&
                hipMemcpyAsync(a, home_h[f4], psi_len_h[f4] * 4, hipMemcpyDeviceToDevice, stream) != hipSuccess)
                return -1;
            a += psi_len_h[f4];
        }
        return 0;
    }
    // after a fused step: what was written becomes what is read
    void flip()
    {
        std::swap(pp.Ez[0], pp.Ez[1]);
        std::swap(pp.Hx[0], pp.Hx[1]);
        std::swap(pp.Hy[0], pp.Hy[1]);
        for (int f4 = 0; f4 < 4; ++f4) {
            std::swap(pp.psi_e[0][f4], pp.psi_e[1][f4]);
            std::swap(pp.psi_h[0][f4], pp.psi_h[1][f4]);
        }
    }
    int unpack(const dgx_solve* P, hipStream_t stream)
    {
        const long long field = (long long)g.nshot * g.plane;
        const float* ez = fused ? pp.Ez[0] : f.Ez;
        const float* hx = fused ? pp.Hx[0] : f.Hx;
        const float* hy = fused ? pp.Hy[0] : f.Hy;
        launch_tm_pack(P->E[2], (float*)ez, field, true, stream);
        launch_tm_pack(P->H[0], (float*)hx, field, true, stream);
        launch_tm_pack(P->H[1], (float*)hy, field, true, stream);
        if (!fused) return 0;
        for (int f4 = 0; f4 < 4; ++f4) {
            if (psi_len_e[f4] && pp.psi_e[0][f4] != home_e[f4] &&
                hipMemcpyAsync(home_e[f4], pp.psi_e[0][f4], psi_len_e[f4] * 4, hipMemcpyDeviceToDevice, stream) !=
                    hipSuccess)
                return -1;
            if (psi_len_h[f4] && pp.psi_h[0][f4] != home_h[f4] &&
                hipMemcpyAsync(home_h[f4], pp.psi_h[0][f4], psi_len_h[f4] * 4, hipMemcpyDeviceToDevice, stream) !=
                    hipSuccess)
                return -1;
        }
        return 0;
    }
};

// Source entries of the fused TM kernels grouped by owning (tile, shot):
// (cell index inside the tile, source index), sources in index order.
void tm_source_lists(const Tm2dGrid& g, const std::vector<int>& sloc, int nsrc, std::vector<int>& off,
                     std::vector<int2>& ents)
{
    const long long nblk = (long long)g.ftiles * g.nshot;
    off.assign((size_t)nblk + 1, 0);
    ents.clear();
    std::vector<std::pair<long long, int2>> tmp;
    for (int s = 0; s < g.nshot; ++s)
        for (int k = 0; k < nsrc; ++k) {
            const int* L = &sloc[((size_t)s * nsrc + k) * 3];
            tmp.push_back({(long long)tm_tile_of(g, L[0], L[1]) + (long long)g.ftiles * s,
                           make_int2(tm_local(g, L[0], L[1]), k)});
        }
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const std::pair<long long, int2>& a, const std::pair<long long, int2>& b) {
                         return a.first < b.first;
                     });
    for (const auto& e : tmp) {
        off[(size_t)e.first + 1] += 1;
        ents.push_back(e.second);
    }
    for (size_t b = 1; b < off.size(); ++b) off[b] += off[b - 1];
}

// Receivers of the fused TM kernels grouped by owning (tile, shot): (cell
// index inside the tile, receiver index).
void tm_receiver_lists(const Tm2dGrid& g, const std::vector<int>& rloc, int nrx, std::vector<int>& off,
                       std::vector<int2>& ent)
{
    std::vector<std::pair<long long, int2>> tmp;
    for (int s = 0; s < g.nshot; ++s)
        for (int r = 0; r < nrx; ++r) {
            const int* L = &rloc[((size_t)s * nrx + r) * 3];
            tmp.push_back({(long long)tm_tile_of(g, L[0], L[1]) + (long long)g.ftiles * s,
                           make_int2(tm_local(g, L[0], L[1]), r)});
        }
    std::stable_sort(tmp.begin(), tmp.end(), [](const std::pair<long long, int2>& x,
                                                const std::pair<long long, int2>& y) { return x.first < y.first; });
    off.assign((size_t)g.ftiles * g.nshot + 1, 0);
    ent.clear();
    for (const auto& e : tmp) {
        off[(size_t)e.first + 1] += 1;
        ent.push_back(e.second);
    }
    for (size_t b = 1; b < off.size(); ++b) off[b] += off[b - 1];
}

// Receiver groups (one per shot and cell, members in receiver order; goff /
// gshot / gcell as built by the backward driver, cells in the compact plane)
// listed for every (tile, shot) whose staged region (tile plus halo) holds the
// cell: (staged index, first member, member count, 0).
void tm_receiver_groups(const Tm2dGrid& g, int order, const std::vector<int>& goff, const std::vector<int>& gshot,
                        const std::vector<long long>& gcell, std::vector<int>& off, std::vector<int4>& grp)
{
    const int he = tm_halo(order);
    const int lc = kTmBJ + 2 * he;
    const int bi = g.bi;
    const int nti = (g.NX + bi - 1) / bi;
    std::vector<std::pair<long long, int4>> tmp;
    for (size_t q = 0; q < gshot.size(); ++q) {
        const int i = (int)(gcell[q] / g.NY), j = (int)(gcell[q] % g.NY);
        for (int ti = std::max(0, (i - he) / bi - 1); ti <= std::min(nti - 1, (i + he) / bi); ++ti) {
            const int r = i - (ti * bi - he);
            if (r < 0 || r >= bi + 2 * he) continue;
            for (int tj = std::max(0, (j - he) / kTmBJ - 1); tj <= std::min(g.ntj - 1, (j + he) / kTmBJ); ++tj) {
                const int c = j - (tj * kTmBJ - he);
                if (c < 0 || c >= lc) continue;
                tmp.push_back({(long long)(ti * g.ntj + tj) + (long long)g.ftiles * gshot[q],
                               make_int4(r * lc + c, goff[q], goff[q + 1] - goff[q], 0)});
            }
        }
    }
    std::stable_sort(tmp.begin(), tmp.end(), [](const std::pair<long long, int4>& x,
                                                const std::pair<long long, int4>& y) { return x.first < y.first; });
    off.assign((size_t)g.ftiles * g.nshot + 1, 0);
    grp.clear();
    for (const auto& e : tmp) {
        off[(size_t)e.first + 1] += 1;
        grp.push_back(e.second);
    }
    for (size_t k = 1; k < off.size(); ++k) off[k] += off[k - 1];
}

// Receiver groups of one call: one per (shot, cell), members in receiver
// order; cell = linear index with strides (st0, st1, 1).
void receiver_groups(const std::vector<int>& rloc, int nshot, int nrx, long long st0, long long st1,
                     std::vector<int>& goff, std::vector<int>& gshot, std::vector<long long>& gcell,
                     std::vector<int>& gmem)
{
    goff.assign(1, 0);
    gshot.clear();
    gcell.clear();
    gmem.clear();
    for (int s = 0; s < nshot; ++s) {
        std::vector<std::pair<long long, int>> cells;
        for (int r = 0; r < nrx; ++r) {
            const int* L = &rloc[((size_t)s * nrx + r) * 3];
            cells.push_back({(long long)L[0] * st0 + (long long)L[1] * st1 + L[2], r});
        }
        std::stable_sort(cells.begin(), cells.end(),
                         [](const std::pair<long long, int>& x, const std::pair<long long, int>& y) {
                             return x.first < y.first;
                         });
        for (size_t q = 0; q < cells.size(); ++q) {
            if (q == 0 || cells[q].first != cells[q - 1].first) {
                if (q) goff.push_back((int)gmem.size());
                gshot.push_back(s);
                gcell.push_back(cells[q].first);
            }
            gmem.push_back(cells[q].second);
        }
        if (!cells.empty()) goff.push_back((int)gmem.size());
    }
}

// Device copy of a host-built list (nothing to do for an empty one).
template <typename T>
bool upload(DevBuf& d, const std::vector<T>& v, hipStream_t st)
{
    if (v.empty()) return true;
    return d.alloc(v.size() * sizeof(T)) &&
           hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st) == hipSuccess;
}

// With `fixed` (dgx_forward_noalloc), the solve allocates no device memory:
// it runs the one-thread-per-cell sweeps and places the electric sweep's
// source lists in `fixed` (fixed_bytes long) instead of allocating them.
int run_forward(const dgx_solve* P, hipStream_t stream, void* fixed = nullptr, size_t fixed_bytes = 0)
{
    CallStream call(stream);
    Plan pl;
    if (make_plan(P, false, pl)) return -1;
    const Grid& g = pl.g;
    const int nt = P->nt;
    const int S = P->sampling_interval;
    float* const* co = P->update_coeff;

    // --- sources: group entries by the block of the electric sweep ---------
    const CellSweep sweep = cell_sweep(g, pl.order);
    std::vector<int> hsloc((size_t)g.nshot * std::max(P->nsource, 0) * 3);
    if (P->nsource > 0) {
        if (!P->source_location || !P->source_waveform) {
            set_error("forward: null source arrays");
            return -1;
        }
        DGX_TRY(hipMemcpyAsync(hsloc.data(), P->source_location, hsloc.size() * sizeof(int),
                               hipMemcpyDeviceToHost, stream),
                "forward: copy source locations");
    }
    std::vector<int> hrloc((size_t)g.nshot * std::max(P->nreceiver, 0) * 3);
    if (P->nreceiver > 0) {
        if (!P->receiver_location || !P->receiver_data) {
            set_error("forward: null receiver arrays");
            return -1;
        }
        DGX_TRY(hipMemcpyAsync(hrloc.data(), P->receiver_location, hrloc.size() * sizeof(int),
                               hipMemcpyDeviceToHost, stream),
                "forward: copy receiver locations");
    }
    DGX_TRY(hipStreamSynchronize(stream), "forward: stream synchronize");
    for (size_t q = 0; q < hsloc.size(); q += 3)
        if (!loc_ok(g, &hsloc[q])) {
            set_error("forward: source location (%d, %d, %d) outside the field grid", hsloc[q], hsloc[q + 1],
                      hsloc[q + 2]);
            return -1;
        }
    for (size_t q = 0; q < hrloc.size(); q += 3)
        if (!loc_ok(g, &hrloc[q])) {
            set_error("forward: receiver location (%d, %d, %d) outside the field grid", hrloc[q], hrloc[q + 1],
                      hrloc[q + 2]);
            return -1;
        }

    const bool inject = P->nsource > 0 && P->source_component >= 0 && P->source_component <= 2;
    const bool tm = !fixed && tm2d_applies(P, pl, hsloc, hrloc);
    const bool fused = tm && tm2d_fused();
    const bool xm = !fixed && !tm && xmarch3d();
    const XmTiles xt = xm_tiles(g, pl.order);
    if (fixed && pl.kind == kInt8 && pl.history) {
        set_error("dgx_forward_noalloc: INT8 histories need a scratch frame; use dgx_forward");
        return -1;
    }
    const Tm2dGrid tg = tm_grid_of(pl, fused ? tm2d_tile_rows(pl, P, stream) : kTmBI);
    const int tm_tiles = (g.N[0] * g.N[1] + 255) / 256;
    long long nblocks = tm ? (long long)tm_tiles * g.nshot
                           : (xm ? (long long)xt.nty * xt.ntz * g.nshot : sweep.blocks);
    std::vector<int> boff((size_t)nblocks + 1, 0);
    std::vector<int2> ents;
    if (inject && fused) {
        tm_source_lists(tg, hsloc, P->nsource, boff, ents);
    } else if (inject) {
        struct E4 {
            long long blk;
            int order;
            int cell;
            int src;
        };
        std::vector<E4> tmp;
        for (int s = 0; s < g.nshot; ++s)
            for (int k = 0; k < P->nsource; ++k) {
                const int* L = &hsloc[((size_t)s * P->nsource + k) * 3];
                long long m = (long long)L[0] * g.st0 + (long long)L[1] * g.N[2] + L[2];
                // x-marching sweeps: one list per (shot, tile), in plane order
                long long blk = xm ? ((long long)s * xt.nty * xt.ntz + xm_tile_of(xt, L[1], L[2])) * g.N[0] + L[0]
                                   : cell_block(sweep, g, L[0], L[1], L[2], s);
                if (blk < 0 && !tm) {
                    set_error("forward: source cell outside the cell sweep");
                    return -1;
                }
                if (tm) {  // compact plane index and 2D tile
                    m = (long long)L[0] * g.N[1] + L[1];
                    blk = (long long)(m / 256) + (long long)tm_tiles * s;
                }
                tmp.push_back({blk, s * P->nsource + k, (int)m, k});
            }
        std::stable_sort(tmp.begin(), tmp.end(), [](const E4& a, const E4& b) { return a.blk < b.blk; });
        for (const E4& e : tmp) {
            const long long list = xm ? e.blk / g.N[0] : e.blk;  // x-marching: key (list, plane)
            boff[(size_t)list + 1] += 1;
            ents.push_back(make_int2(e.cell, e.src));
        }
        for (size_t b = 1; b < boff.size(); ++b) boff[b] += boff[b - 1];
    }
    DevBuf d_boff, d_ents;
    SourceBlocks src{};
    if (inject) {
        void* off_p = nullptr;
        void* ent_p = nullptr;
        const size_t off_bytes = boff.size() * sizeof(int), ent_bytes = ents.size() * sizeof(int2);
        if (fixed) {
            const size_t at = (off_bytes + 255) / 256 * 256;
            if (at + ent_bytes > fixed_bytes) {
                set_error("dgx_forward_noalloc: the source lists need %zu bytes, more than the %zu-byte scratch "
                          "(fewer shots per call, or dgx_forward)", at + ent_bytes, fixed_bytes);
                return -1;
            }
            off_p = fixed;
            ent_p = (unsigned char*)fixed + at;
        } else {
            if (!d_boff.alloc(off_bytes) || !d_ents.alloc(ent_bytes)) {
                set_error("forward: out of device memory for source lists");
                return -1;
            }
            off_p = d_boff.p;
            ent_p = d_ents.p;
        }
        DGX_TRY(hipMemcpyAsync(off_p, boff.data(), off_bytes, hipMemcpyHostToDevice, stream),
                "forward: upload source lists");
        DGX_TRY(hipMemcpyAsync(ent_p, ents.data(), ent_bytes, hipMemcpyHostToDevice, stream),
                "forward: upload source entries");
        src.block_off = (const int*)off_p;
        src.entries = (const int2*)ent_p;
        src.count = (int)ents.size();
        src.wave = P->source_waveform;
        src.nt = nt;
        src.comp = P->source_component;
        src.dl = g.d[P->source_component];
        src.vol = g.d[0] * g.d[1] * g.d[2];
    }

    // --- receivers of the fused 2D TM kernels: grouped by owning (tile, shot)
    const int rcomp = (P->receiver_component >= 0 && P->receiver_component <= 2) ? P->receiver_component : 2;
    const bool fused_rx = fused && P->nreceiver > 0 && rcomp == 2;
    DevBuf d_roff, d_rent;
    TmReceiversFwd trx{};
    if (fused_rx) {
        std::vector<int> roff;
        std::vector<int2> rent;
        tm_receiver_lists(tg, hrloc, P->nreceiver, roff, rent);
        if (!upload(d_roff, roff, stream) || !upload(d_rent, rent, stream)) {
            set_error("forward: out of device memory for receiver lists");
            return -1;
        }
        trx.off = (const int*)d_roff.p;
        trx.ent = (const int2*)d_rent.p;
        trx.out = P->receiver_data;
        trx.nrx = P->nreceiver;
        trx.nt = nt;
        trx.count = (int)rent.size();
    }

    // --- history buffers -----------------------------------------------------
    const bool int8 = pl.kind == kInt8;
    if (pl.history && (!P->E_saved || ((pl.store_r || pl.e_only) && !P->R_saved))) {
        set_error("forward: null history buffer");
        return -1;
    }
    DevBuf d_rscratch;  // FP32 R^n frame for INT8 encoding
    if (pl.history && pl.store_r && int8) {
        if (!d_rscratch.alloc((size_t)pl.comps * pl.snap * sizeof(float))) {
            set_error("forward: out of device memory for the INT8 R frame (%lld bytes)",
                      (long long)pl.comps * pl.snap * 4);
            return -1;
        }
    }

    LaunchTimer tmr(stream);
    TIMED(tmr, kKCoeff, launch_coefficients(g, P->eps_r_pad, P->sigma_pad, P->mu_r_pad, co, P->dt, stream));

    StepArgs a{};
    a.ca = co[0];
    a.ccurl = co[1];
    a.cb = co[2];
    a.hcurl = co[4];
    a.hrhs = co[5];
    for (int c = 0; c < 3; ++c) {
        a.E.c[c] = P->E[c];
        a.H.c[c] = P->H[c];
        a.Ec.c[c] = P->E[c];
        a.Hc.c[c] = P->H[c];
    }
    a.cpe = pl.cpe;
    a.cph = pl.cph;
    a.has_cpml = pl.has_cpml;
    a.src = src;
    const size_t eb = elem_bytes(pl.kind);
    DevBuf d_kc;  // grid, CPML and coefficients of the x-marching sweeps
    if (xm) {
        const XmConst kc = xm_const(g, pl.cpe, pl.cph, co[0], co[1], co[2], co[4], co[5]);
        if (!d_kc.alloc(sizeof(XmConst))) {
            set_error("forward: out of device memory for the sweep constants");
            return -1;
        }
        DGX_TRY(hipMemcpyAsync(d_kc.p, &kc, sizeof(XmConst), hipMemcpyHostToDevice, stream),
                "forward: upload sweep constants");
    }
    const XmConst* kcp = (const XmConst*)d_kc.p;
    Tm2dState tms;
    if (tm && tms.setup(pl, P, stream, fused, tg.bi)) {
        set_error("forward: out of device memory for the 2D TM state");
        return -1;
    }

    // Forward-only fused solves: kGraphSteps steps captured once (each node
    // reads its step from d_step), replayed nt / kGraphSteps times; the
    // remaining steps are launched one by one below.
    int it0 = 0;
    DevBuf d_step;
    const bool graphed = fused && !pl.history && !tmr.on && tm2d_graphs() && nt >= 2 * kGraphSteps &&
                         (P->nreceiver <= 0 || fused_rx);
    if (graphed) {
        // earlier launch errors are reported here: the fallback below clears
        // whatever a failed capture leaves behind
        DGX_TRY(hipGetLastError(), "forward: kernel launch");
        hipStream_t cs = capture_stream();
        if (cs && d_step.alloc(sizeof(int))) {
            int* step = (int*)d_step.p;
            HistWrite none{};
            none.mode = pl.comps == 3 ? 3 : 2;
            none.box = pl.box;
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                for (int k = 0; k < kGraphSteps; ++k) {
                    launch_tm_step_fwd(pl.order, tms.g, tms.f, tms.pp, a.src, trx, none, k, cs, step);
                    tms.flip();
                }
                launch_step_counter(step, kGraphSteps, true, cs);
                ok = hipStreamEndCapture(cs, &graph) == hipSuccess && graph;
            }
            ok = ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
            if (ok) {
                launch_step_counter(step, 0, false, stream);
                const int reps = nt / kGraphSteps;
                for (int r = 0; r < reps && ok; ++r) ok = hipGraphLaunch(exec, stream) == hipSuccess;
                hipEvent_t done = nullptr;
                if (ok && hipEventCreateWithFlags(&done, hipEventDisableTiming) == hipSuccess) {
                    if (hipEventRecord(done, stream) == hipSuccess) {
                        std::lock_guard<std::mutex> lock(g_graph_mu);
                        g_graphs.push_back({exec, graph, done});
                        exec = nullptr;
                        graph = nullptr;
                    } else {
                        (void)hipEventDestroy(done);
                    }
                }
                if (!ok) {
                    set_error("forward: graph launch failed");
                    return -1;
                }
                it0 = reps * kGraphSteps;
            }
            if (exec || graph) {  // not launched, or its completion event could not be recorded
                if (it0) (void)hipStreamSynchronize(stream);
                if (exec) (void)hipGraphExecDestroy(exec);
                if (graph) (void)hipGraphDestroy(graph);
            }
            (void)hipGetLastError();  // a failed capture falls back to launches one by one
        }
    }
    for (int it = it0; it < nt; ++it) {
        const bool save = pl.history && (it % S) == 0;
        const int ts = it / S;
        HistWrite hw{};
        hw.mode = pl.comps == 3 ? 3 : 2;
        hw.box = pl.box;
        if (save) {
            if (int8) {
                TIMED(tmr, kKQuant,
                      launch_quant_int8(g, a.Ec, nullptr, 0, int8_writer(pl, P->E_saved, pl.nt_saved, ts), stream));
            } else {
                hw.save_e = 1;
                hw.e_kind = pl.kind;
                hw.e_base = (unsigned char*)P->E_saved + (size_t)ts * pl.snap * eb;
                hw.e_cstride = pl.cstride;
            }
            if (pl.store_r) {
                hw.save_r = 1;
                if (int8) {
                    hw.r_kind = kF32;
                    hw.r_base = d_rscratch.p;
                    hw.r_cstride = pl.snap;
                } else {
                    hw.r_kind = pl.kind;
                    hw.r_base = (unsigned char*)P->R_saved + (size_t)ts * pl.snap * eb;
                    hw.r_cstride = pl.cstride;
                }
            }
            hw.active = hw.save_e || hw.save_r;
        }
        a.hist = hw;
        a.src.it = it;
        if (fused) {
            TIMED(tmr, kKTmStepFwd,
                  launch_tm_step_fwd(pl.order, tms.g, tms.f, tms.pp, a.src, trx, a.hist, it, stream));
            tms.flip();
        } else if (tm) {
            TIMED(tmr, kKFwdH, launch_tm_fwd_h(pl.order, tms.g, tms.f, stream));
            TIMED(tmr, kKFwdE, launch_tm_fwd_e(pl.order, tms.g, tms.f, a.src, a.hist, stream));
        } else if (xm) {
            TIMED(tmr, kKXmFwdH, launch_xm_fwd_h(pl.order, g, xt, kcp, a, stream));
            TIMED(tmr, kKXmFwdE, launch_xm_fwd_e(pl.order, g, xt, kcp, a, stream));
        } else {
            TIMED(tmr, kKFwdH, launch_fwd_h(pl.order, g, a, stream));
            TIMED(tmr, kKFwdE, launch_fwd_e(pl.order, g, a, stream));
        }
        if (save && pl.store_r && int8)
            TIMED(tmr, kKQuant,
                  launch_quant_int8(g, a.Ec, (const float*)d_rscratch.p, pl.snap,
                                    int8_writer(pl, P->R_saved, pl.nt_saved, ts), stream));
        if (P->nreceiver > 0 && !fused_rx) {
            if (tm && rcomp == 2)
                TIMED(tmr, kKSample,
                      launch_sample(g, tms.f.Ez, P->receiver_location, P->receiver_data, P->nreceiver, nt, it,
                                    stream, tms.g.plane, tms.g.NY, 1));
            else
                TIMED(tmr, kKSample,
                      launch_sample(g, P->E[rcomp], P->receiver_location, P->receiver_data, P->nreceiver, nt, it,
                                    stream));
        }
    }
    if (tm && tms.unpack(P, stream)) {
        set_error("forward: copy back of the 2D TM state");
        return -1;
    }
    if (pl.e_only) {
        if (int8) {
            launch_quant_int8(g, a.Ec, nullptr, 0, int8_writer(pl, P->R_saved, 1, 0), stream);
        } else {
            HistWrite hw{};
            hw.mode = pl.comps == 3 ? 3 : 2;
            hw.box = pl.box;
            hw.e_kind = pl.kind;
            hw.e_base = P->R_saved;
            hw.e_cstride = pl.snap;
            TIMED(tmr, kKSaveFrame, launch_save_frame(g, a.Ec, hw, stream));
        }
    }
    // Asynchronous like the reference: launch errors are reported here, the
    // caller's stream orders everything after (scratch is freed on it too).
    DGX_TRY(hipGetLastError(), "forward: kernel launch");
    if (tmr.on) {
        DGX_TRY(hipStreamSynchronize(stream), "forward: stream synchronize");
        tmr.collect();
    }
    sweep_graphs();
    return 0;
}

// History frames and weight of the saved step that starts at time step it.
MatGrad mat_step(const Plan& pl, const dgx_solve* P, MatGrad mg, int it, int nt, int S)
{
    const size_t eb = elem_bytes(pl.kind);
    int wgt = S;
    const int ts = it / S;
    if (it + wgt > nt) wgt = nt - it;
    mg.weight = wgt;
    if (pl.kind == kInt8) {
        mg.e8 = int8_view(pl, P->E_saved, pl.nt_saved, ts);
        if (!pl.e_only) mg.s8 = int8_view(pl, P->R_saved, pl.nt_saved, ts);
        else if (ts + 1 < pl.nt_saved) mg.s8 = int8_view(pl, P->E_saved, pl.nt_saved, ts + 1);
        else mg.s8 = int8_view(pl, P->R_saved, 1, 0);
    } else {
        mg.eframe = (const unsigned char*)P->E_saved + (size_t)ts * pl.snap * eb;
        mg.e_cstride = pl.cstride;
        mg.s_cstride = pl.cstride;
        if (!pl.e_only) {
            mg.sframe = (const unsigned char*)P->R_saved + (size_t)ts * pl.snap * eb;
        } else if (ts + 1 < pl.nt_saved) {
            mg.sframe = (const unsigned char*)P->E_saved + (size_t)(ts + 1) * pl.snap * eb;
        } else {
            mg.sframe = P->R_saved;
            mg.s_cstride = pl.snap;
        }
    }
    return mg;
}

int run_backward(const dgx_solve* P, hipStream_t stream)
{
    CallStream call(stream);
    Plan pl;
    if (make_plan(P, true, pl)) return -1;
    const Grid& g = pl.g;
    const int nt = P->nt;
    const int S = P->sampling_interval;
    float* const* co = P->update_coeff;
    const bool want_mat = P->eps_r_requires_grad || P->sigma_requires_grad;
    const bool want_src = P->source_requires_grad == 1 && P->nsource > 0 && P->source_component >= 0 &&
                          P->source_component <= 2;
    const bool fp64 = P->accumulate_fp64 != 0;

    std::vector<int> hrloc((size_t)g.nshot * std::max(P->nreceiver, 0) * 3);
    std::vector<int> hsloc;
    if (P->nreceiver > 0) {
        if (!P->receiver_location || !P->data_grad) {
            set_error("backward: null receiver arrays");
            return -1;
        }
        DGX_TRY(hipMemcpyAsync(hrloc.data(), P->receiver_location, hrloc.size() * sizeof(int),
                               hipMemcpyDeviceToHost, stream),
                "backward: copy receiver locations");
    }
    if (want_src && (!P->source_location || !P->grad_source)) {
        set_error("backward: null source arrays");
        return -1;
    }
    if (P->nsource > 0 && P->source_location) {
        hsloc.resize((size_t)g.nshot * P->nsource * 3);
        DGX_TRY(hipMemcpyAsync(hsloc.data(), P->source_location, hsloc.size() * sizeof(int),
                               hipMemcpyDeviceToHost, stream),
                "backward: copy source locations");
    }
    DGX_TRY(hipStreamSynchronize(stream), "backward: stream synchronize");
    for (size_t q = 0; q < hrloc.size(); q += 3)
        if (!loc_ok(g, &hrloc[q])) {
            set_error("backward: receiver location (%d, %d, %d) outside the field grid", hrloc[q], hrloc[q + 1],
                      hrloc[q + 2]);
            return -1;
        }
    for (size_t q = 0; q < hsloc.size(); q += 3)
        if (!loc_ok(g, &hsloc[q])) {
            set_error("backward: source location (%d, %d, %d) outside the field grid", hsloc[q], hsloc[q + 1],
                      hsloc[q + 2]);
            return -1;
        }
    if (want_mat) {
        if (!P->E_saved || !P->R_saved || !P->sigma_pad ||
            (P->eps_r_requires_grad && !P->grad_eps_r) || (P->sigma_requires_grad && !P->grad_sigma)) {
            set_error("backward: null history or gradient buffer");
            return -1;
        }
    }

    const bool tm = tm2d_applies(P, pl, hsloc, hrloc);
    const bool fused = tm && tm2d_fused();
    const Tm2dGrid tg = tm_grid_of(pl, fused ? tm2d_tile_rows(pl, P, stream) : kTmBI);
    Grid gc = g;  // compact 2D TM addressing for the receiver/source kernels
    if (tm) {
        gc.st0 = g.N[1];
        gc.N[2] = 1;
        gc.cells = (long long)g.N[0] * g.N[1];
    }

    // receiver groups: one per (shot, cell), members in receiver order; on the
    // 2D TM path only Ez receivers couple into the active fields
    const bool adj_rx = P->nreceiver > 0 && P->receiver_component >= 0 && P->receiver_component <= 2 &&
                        (!tm || P->receiver_component == 2);
    std::vector<int> goff(1, 0), gshot, gmem;
    std::vector<long long> gcell;
    if (adj_rx) receiver_groups(hrloc, g.nshot, P->nreceiver, gc.st0, gc.N[2], goff, gshot, gcell, gmem);
    DevBuf d_goff, d_gshot, d_gcell, d_gmem;
    ReceiverGroups rg{};
    if (adj_rx && !gshot.empty()) {
        if (!d_goff.alloc(goff.size() * 4) || !d_gshot.alloc(gshot.size() * 4) || !d_gcell.alloc(gcell.size() * 8) ||
            !d_gmem.alloc(gmem.size() * 4)) {
            set_error("backward: out of device memory for receiver groups");
            return -1;
        }
        DGX_TRY(hipMemcpyAsync(d_goff.p, goff.data(), goff.size() * 4, hipMemcpyHostToDevice, stream), "upload");
        DGX_TRY(hipMemcpyAsync(d_gshot.p, gshot.data(), gshot.size() * 4, hipMemcpyHostToDevice, stream), "upload");
        DGX_TRY(hipMemcpyAsync(d_gcell.p, gcell.data(), gcell.size() * 8, hipMemcpyHostToDevice, stream), "upload");
        DGX_TRY(hipMemcpyAsync(d_gmem.p, gmem.data(), gmem.size() * 4, hipMemcpyHostToDevice, stream), "upload");
        rg.off = (const int*)d_goff.p;
        rg.shot = (const int*)d_gshot.p;
        rg.cell = (const long long*)d_gcell.p;
        rg.member = (const int*)d_gmem.p;
        rg.ngroups = (int)gshot.size();
    }

    const bool xm = !tm && xmarch3d();
    const XmTiles xt = xm_tiles(g, pl.order);
    const bool fuse3 = !tm && !xm && adj3d_fused();

    // fused 2D TM kernels: the groups again, listed for every (tile, shot)
    // whose staged region (tile plus halo) holds the cell
    DevBuf d_xoff, d_xgrp;
    TmReceiversAdj tax{};
    if (fused && adj_rx && !gshot.empty()) {
        std::vector<int> xoff;
        std::vector<int4> xgrp;
        tm_receiver_groups(tg, pl.order, goff, gshot, gcell, xoff, xgrp);
        if (!upload(d_xoff, xoff, stream) || !upload(d_xgrp, xgrp, stream)) {
            set_error("backward: out of device memory for receiver lists");
            return -1;
        }
        tax.off = (const int*)d_xoff.p;
        tax.grp = (const int4*)d_xgrp.p;
        tax.member = rg.member;
        tax.dgrad = P->data_grad;
        tax.nrx = P->nreceiver;
        tax.nt = nt;
        tax.count = (int)xgrp.size();
    }
    // fused 2D TM kernels: source lists, per-shot source and material terms
    DevBuf d_soff, d_sent, d_svals, d_terms;
    SourceBlocks tsrc{};
    TmAdjGrad tgrad{};
    if (fused && want_src) {
        std::vector<int> soff;
        std::vector<int2> sent;
        tm_source_lists(tg, hsloc, P->nsource, soff, sent);
        if (!upload(d_soff, soff, stream) || !upload(d_sent, sent, stream) ||
            !d_svals.alloc((size_t)nt * g.nshot * P->nsource * sizeof(float))) {
            set_error("backward: out of device memory for source lists");
            return -1;
        }
        tsrc.block_off = (const int*)d_soff.p;
        tsrc.entries = (const int2*)d_sent.p;
        tsrc.count = (int)sent.size();
        tgrad.src_vals = (float*)d_svals.p;
        tgrad.nsrc = P->nsource;
    }
    // two buffers: a launch writes one while it sums the other (the previous saved step)
    const long long terms_len = 2 * (long long)g.nshot * pl.box.cells;
    if (fused && want_mat && !d_terms.alloc((size_t)2 * terms_len * sizeof(float))) {
        set_error("backward: out of device memory for per-shot gradient terms");
        return -1;
    }

    // CPML weight scratch, shaped like the psi arrays
    DevBuf d_w;
    Tm2dWeights tw{};
    if (pl.has_cpml && fused) {
        // the fused kernels keep the weights in LDS
    } else if (pl.has_cpml && tm) {
        long long tot = 0;
        for (int f = 0; f < 4; ++f) {
            long long across = f < 2 ? g.N[1] : g.N[0];
            tot += (long long)g.nshot * across * (2 * g.pml[f] + 1);
        }
        if (!d_w.alloc((size_t)tot * sizeof(float))) {
            set_error("backward: out of device memory for CPML weights");
            return -1;
        }
        float* p = (float*)d_w.p;
        for (int f = 0; f < 4; ++f) {
            long long across = (long long)g.nshot * (f < 2 ? g.N[1] : g.N[0]);
            tw.we[f] = p;
            p += across * (g.pml[f] + 1);
            tw.wh[f] = p;
            p += across * g.pml[f];
        }
    } else if (!tm && !xm) {
        // six channel-weight arrays: shared by the E and H halves of the
        // adjoint, or one set per half when the gathers write the next weights
        const long long wlen = (fuse3 ? 12LL : 6LL) * g.nshot * g.cells;
        if (!d_w.alloc((size_t)wlen * sizeof(float))) {
            set_error("backward: out of device memory for the adjoint channel weights (%lld bytes)", wlen * 4);
            return -1;
        }
    }

    // FP64 accumulators of the time-summed material gradients
    const long long ngrad = (long long)(g.N[0] - 1) * (g.N[1] - 1) * (g.N[2] - 1);
    DevBuf d_acc;
    void* geps = P->grad_eps_r;
    void* gsig = P->grad_sigma;
    if (want_mat && fp64) {
        if (!d_acc.alloc((size_t)ngrad * 2 * sizeof(double))) {
            set_error("backward: out of device memory for FP64 gradient accumulators");
            return -1;
        }
        DGX_TRY(hipMemsetAsync(d_acc.p, 0, (size_t)ngrad * 2 * sizeof(double), stream), "backward: zero accumulators");
        geps = d_acc.p;
        gsig = (double*)d_acc.p + ngrad;
    }

    LaunchTimer tmr(stream);
    TIMED(tmr, kKCoeff, launch_coefficients(g, P->eps_r_pad, P->sigma_pad, P->mu_r_pad, co, P->dt, stream));

    AdjArgs a{};
    a.ca = co[0];
    a.ccurl = co[1];
    a.cb = co[2];
    a.hcurl = co[4];
    a.hrhs = co[5];
    for (int c = 0; c < 3; ++c) {
        a.lE.c[c] = P->E[c];
        a.lH.c[c] = P->H[c];
        a.lEc.c[c] = P->E[c];
        a.lHc.c[c] = P->H[c];
    }
    a.cpe = pl.cpe;
    a.cph = pl.cph;
    a.has_cpml = pl.has_cpml;
    DevBuf d_kc;  // grid, CPML and coefficients of the x-marching sweeps
    if (xm) {
        const XmConst kc = xm_const(g, pl.cpe, pl.cph, co[0], co[1], co[2], co[4], co[5]);
        if (!d_kc.alloc(sizeof(XmConst))) {
            set_error("backward: out of device memory for the sweep constants");
            return -1;
        }
        DGX_TRY(hipMemcpyAsync(d_kc.p, &kc, sizeof(XmConst), hipMemcpyHostToDevice, stream),
                "backward: upload sweep constants");
    }
    const XmConst* kcp = (const XmConst*)d_kc.p;

    MatGrad mg{};
    mg.ca = co[0];
    mg.cb = co[2];
    mg.sig = P->sigma_pad;
    mg.kind = pl.kind;
    mg.e_only = pl.e_only;
    mg.mode = pl.comps == 3 ? 3 : 2;
    mg.want_eps = P->eps_r_requires_grad == 1;
    mg.want_sig = P->sigma_requires_grad == 1;
    mg.dt = P->dt;
    mg.box = pl.box;
    if (pl.box.nz > 0 && pl.box.ny > 0) {
        mg.div_z = make_div((unsigned)pl.box.nz);
        mg.div_y = make_div((unsigned)pl.box.ny);
    }
    const float geo = want_src ? g.d[P->source_component] / (g.d[0] * g.d[1] * g.d[2]) : 0.0f;
    Tm2dState tms;
    if (tm && tms.setup(pl, P, stream, fused, tg.bi)) {
        set_error("backward: out of device memory for the 2D TM state");
        return -1;
    }
    tgrad.geo = geo;
    tgrad.ez_comp = pl.comps == 3 ? 2 : 0;
    float* lam_rx = tm ? tms.f.Ez : (adj_rx ? P->E[P->receiver_component] : nullptr);
    float* lam_src = tm ? tms.f.Ez : (want_src ? P->E[P->source_component] : nullptr);
    const float* cb_src = tm ? tms.f.cb : co[2];

    tgrad.geps = geps;
    tgrad.gsig = gsig;
    tgrad.mg = mg;
    int terms_buf = 0;
    for (int it = nt - 1; it >= 0; --it) {
        if (fused) {
            tgrad.pending = tgrad.terms;
            tgrad.terms = nullptr;
            if (want_mat && it % S == 0) {
                tgrad.mg = mat_step(pl, P, mg, it, nt, S);
                tgrad.terms = (float*)d_terms.p + terms_buf * terms_len;
                terms_buf ^= 1;
            }
            TIMED(tmr, kKTmStepAdj,
                  launch_tm_step_adj(pl.order, tms.g, tms.f, tms.pp, tsrc, tax, tgrad, it, fp64, stream));
            tms.flip();
            continue;
        }
        if (adj_rx)
            TIMED(tmr, kKAdjRx, launch_adj_receivers(gc, rg, lam_rx, P->data_grad, P->nreceiver, nt, it, stream));
        if (want_src)
            TIMED(tmr, kKAdjSrc,
                  launch_adj_source(gc, P->source_location, cb_src, lam_src, P->grad_source, P->nsource, nt, it,
                                    geo, stream));
        if (want_mat && it % S == 0) {
            mg = mat_step(pl, P, mg, it, nt, S);
            if (tm)
                TIMED(tmr, kKMatGrad,
                      launch_tm_material_gradient(tms.g, tms.f, mg, pl.comps == 3 ? 2 : 0, geps, gsig, fp64, stream));
            else
                TIMED(tmr, kKMatGrad, launch_material_gradient(g, a.lEc, mg, geps, gsig, fp64, stream));
        }
        if (tm) {
            if (a.has_cpml) TIMED(tmr, kKAdjPmlE, launch_tm_adj_pml(true, tms.g, tms.f, tw, stream));
            TIMED(tmr, kKAdjE, launch_tm_adj_e(pl.order, tms.g, tms.f, tw, stream));
            if (a.has_cpml) TIMED(tmr, kKAdjPmlH, launch_tm_adj_pml(false, tms.g, tms.f, tw, stream));
            TIMED(tmr, kKAdjH, launch_tm_adj_h(pl.order, tms.g, tms.f, tw, stream));
        } else if (xm) {
            TIMED(tmr, kKXmAdjE, launch_xm_adj(pl.order, true, g, xt, kcp, a, it != nt - 1, stream));
            TIMED(tmr, kKXmAdjH, launch_xm_adj(pl.order, false, g, xt, kcp, a, false, stream));
        } else if (fuse3) {
            float* We = (float*)d_w.p;
            float* Wh = We + 6LL * g.nshot * g.cells;
            TIMED(tmr, kKAdjWeightsE, launch_adj_weights(pl.order, true, g, a, We, stream));
            TIMED(tmr, kKAdjEW, launch_adj_gather_e_w(pl.order, g, a, We, Wh, stream));
            TIMED(tmr, kKAdjH, launch_adj_gather_h(pl.order, g, a, Wh, stream));
        } else {
            float* W = (float*)d_w.p;
            TIMED(tmr, kKAdjWeightsE, launch_adj_weights(pl.order, true, g, a, W, stream));
            TIMED(tmr, kKAdjE, launch_adj_gather_e(pl.order, g, a, W, stream));
            TIMED(tmr, kKAdjWeightsH, launch_adj_weights(pl.order, false, g, a, W, stream));
            TIMED(tmr, kKAdjH, launch_adj_gather_h(pl.order, g, a, W, stream));
        }
    }
    if (xm && nt > 0) TIMED(tmr, kKXmPsiH, launch_xm_psi_h_final(g, kcp, a, stream));
    if (tm && tms.unpack(P, stream)) {
        set_error("backward: copy back of the 2D TM state");
        return -1;
    }
    if (fused && tgrad.terms)  // the last saved step has no later launch to sum it
        TIMED(tmr, kKTmGradReduce,
              launch_tm_grad_reduce(tms.g, pl.box, tgrad.terms, mg.want_eps, mg.want_sig, geps, gsig, fp64, stream));
    if (fused && want_src)
        TIMED(tmr, kKTmSrcReduce,
              launch_tm_src_reduce(tgrad.src_vals, P->grad_source, g.nshot, P->nsource, nt, stream));
    if (want_mat && fp64) {
        if (P->eps_r_requires_grad == 1) launch_add_f64((const double*)geps, P->grad_eps_r, ngrad, stream);
        if (P->sigma_requires_grad == 1) launch_add_f64((const double*)gsig, P->grad_sigma, ngrad, stream);
    }
    DGX_TRY(hipGetLastError(), "backward: kernel launch");
    if (tmr.on) {
        DGX_TRY(hipStreamSynchronize(stream), "backward: stream synchronize");
        tmr.collect();
    }
    return 0;
}

// --- FP16/BF16 codec self-test ----------------------------------------------

__global__ void k_codec(const float* in, void* enc, float* dec, long long count, int kind, int backend)
{
    long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= count) return;
    float v = in[n];
    uint16_t bits;
    if (kind == kF16) {
        if (backend == 0) {
            bits = f32_to_f16_bits(v);
        } else {
            _Float16 hv = (_Float16)v;  // hardware round-to-nearest-even conversion
            __builtin_memcpy(&bits, &hv, 2);
        }
        dec[n] = f16_bits_to_f32(bits);
    } else if (kind == kBF16) {
        if (backend == 0) {
            bits = f32_to_bf16_bits(v);
        } else {
            __bf16 bv = (__bf16)v;  // hardware round-to-nearest-even conversion
            __builtin_memcpy(&bits, &bv, 2);
        }
        dec[n] = bf16_bits_to_f32(bits);
    } else {
        ((float*)enc)[n] = v;
        dec[n] = v;
        return;
    }
    ((uint16_t*)enc)[n] = bits;
}

}  // namespace

// ---------------------------------------------------------------------------
// Exported C ABI
// ---------------------------------------------------------------------------

extern "C" {

DGX_API int dgx_abi_version(void) { return DGX_ABI_VERSION; }

DGX_API const char* dgx_last_error(void) { return g_error; }

// ---------------------------------------------------------------------------
// Batched 2D TM solves: instances of different sizes, orders and lengths
// advance together, one launch per (order, phase) per step.
// ---------------------------------------------------------------------------
struct BatchInstance {
    int id = 0;
    int order = 2;
    bool gradient = false;
    TmBatchInst d{};
    Tm2dState tms;
    DevBuf l_soff, l_sent, l_roff, l_rent, l_xoff, l_xgrp, l_mem;
    DevBuf dgrad, acc, terms;
    float* grad_eps = nullptr;
    float* grad_sig = nullptr;
    long long ngrad = 0;
    const float* rx_out = nullptr;
    long long rx_count = 0;
    // set up on another stream, used and finally freed on the batch stream
    void free_after(hipStream_t s)
    {
        for (DevBuf* d : {&tms.buf, &tms.alt, &l_soff, &l_sent, &l_roff, &l_rent, &l_xoff, &l_xgrp, &l_mem, &dgrad,
                          &acc, &terms})
            d->free_after(s);
    }
};

// Pinned host buffer that stages one descriptor upload; `done` marks the end
// of the copy that last read it, so the host waits only for that copy (a
// pageable copy would wait for everything queued on the stream before it).
struct Staging {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
};

// Device memory the batch stream may still read, kept until an event
// recorded after its last reader has completed. Only then is it freed, so a
// later stream-ordered allocation on any stream (instances are set up on
// another stream) can never receive memory a queued step still uses.
// Memory the batch stream may still read: a finished instance's buffers or a
// replaced descriptor table (released by their destructors).
struct Retiree {
    std::unique_ptr<BatchInstance> inst;
    std::unique_ptr<DevBuf> table;
};

struct HipRetireOps {
    using Event = hipEvent_t;
    using Stream = hipStream_t;
    bool create(Event* e) { return hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess; }
    bool record(Event e, Stream s) { return hipEventRecord(e, s) == hipSuccess; }
    bool done(Event e) { return hipEventQuery(e) == hipSuccess; }
    void destroy(Event e) { (void)hipEventDestroy(e); }
    bool sync(Stream s) { return hipStreamSynchronize(s) == hipSuccess; }
};

struct Batch {
    hipStream_t stream = nullptr;
    int gstep = 0;
    int next_id = 1;
    std::vector<std::unique_ptr<BatchInstance>> active;
    std::vector<std::pair<int, hipEvent_t>> finished;
    RetireList<HipRetireOps, Retiree> retired;
    bool dirty = true;
    std::unique_ptr<DevBuf> table;  // descriptors and prefixes of the 6 slots
    size_t table_cap = 0;
    Staging ring[8];
    int ring_next = 0;
    const TmBatchInst* desc[6] = {};
    const int* prefix[6] = {};
    int ninst[6] = {}, total[6] = {}, order_of[6] = {2, 4, 8, 2, 4, 8};
    // kBatchGraphSteps steps of the current table as one graph (kernel nodes
    // only); node q launches step *gbase + q
    DevBuf gbase;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool graph_failed = false;
    ~Batch()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (Staging& r : ring) {
            if (r.done) (void)hipEventDestroy(r.done);
            if (r.host) (void)hipHostFree(r.host);
        }
    }
};

constexpr int kBatchGraphSteps = 32;

// Groups of kBatchGraphSteps batch steps between phase ends replay one graph
// per descriptor table; DEEPGPR_AMD_BATCH_GRAPHS=0 launches every step one by
// one (same bits). Read at every call.
bool batch_graphs()
{
    const char* v = std::getenv("DEEPGPR_AMD_BATCH_GRAPHS");
    return !(v && std::strcmp(v, "0") == 0);
}

// The graph reads the descriptor table it was captured with: it is retired
// (destroyed once the batch stream has passed its last launch) whenever the
// table is rebuilt.
void batch_drop_graph(Batch* b)
{
    if (!b->gexec && !b->graph) return;
    hipEvent_t done = nullptr;
    if (hipEventCreateWithFlags(&done, hipEventDisableTiming) == hipSuccess &&
        hipEventRecord(done, b->stream) == hipSuccess) {
        std::lock_guard<std::mutex> lock(g_graph_mu);
        g_graphs.push_back({b->gexec, b->graph, done});
    } else {
        if (done) (void)hipEventDestroy(done);
        (void)hipStreamSynchronize(b->stream);
        if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
    }
    b->gexec = nullptr;
    b->graph = nullptr;
}

// Captures kBatchGraphSteps steps of the current table (not executed while
// captured); false leaves the caller to launch the steps one by one.
bool batch_graph_ready(Batch* b)
{
    if (b->gexec) return true;
    if (b->graph_failed) return false;
    hipStream_t cs = capture_stream();
    if (!cs || (!b->gbase.p && !b->gbase.alloc(sizeof(int)))) return false;
    const int* base = (const int*)b->gbase.p;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        for (int q = 0; q < kBatchGraphSteps; ++q)
            for (int slot = 0; slot < 6; ++slot)
                launch_tm_batch(b->order_of[slot], slot >= 3, b->desc[slot], b->prefix[slot], b->ninst[slot],
                                b->total[slot], q, cs, base);
        ok = hipStreamEndCapture(cs, &graph) == hipSuccess && graph;
    }
    ok = ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (!ok) {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();  // a failed capture falls back to launches one by one
        b->graph_failed = true;
        return false;
    }
    b->graph = graph;
    b->gexec = exec;
    return true;
}

int order_slot(int order) { return order == 8 ? 2 : (order == 4 ? 1 : 0); }

// Moves memory the batch stream may still read to the retired list, behind
// an event recorded on the batch stream now (after every launch queued so
// far). When no event can be recorded the list drains the stream first, so
// the memory is never released while queued steps may read it.
int batch_retire(Batch* b, std::unique_ptr<BatchInstance> inst, std::unique_ptr<DevBuf> table)
{
    if (b->retired.retire(b->stream, Retiree{std::move(inst), std::move(table)}) != 0) {
        set_error("batch: could not record a completion event for retired memory");
        return -1;
    }
    return 0;
}

void batch_sweep(Batch* b) { b->retired.sweep(); }

// Sets an instance up on `stream` (not the batch stream, so that the host
// waits only for the set-up work); the batch stream waits for it before the
// instance's first step.
int batch_add(Batch* b, const dgx_solve* P, const dgx_solve* G, hipStream_t stream)
{
    CallStream call(stream);
    batch_sweep(b);
    Plan pl;
    if (make_plan(P, false, pl)) return -1;
    const Grid& g = pl.g;
    if (P->nsource <= 0 || P->nreceiver <= 0 || !P->source_location || !P->receiver_location) {
        set_error("batch: an instance needs sources and receivers");
        return -1;
    }
    std::vector<int> hsloc((size_t)g.nshot * P->nsource * 3), hrloc((size_t)g.nshot * P->nreceiver * 3);
    DGX_TRY(hipMemcpyAsync(hsloc.data(), P->source_location, hsloc.size() * 4, hipMemcpyDeviceToHost, stream),
            "batch: copy source locations");
    DGX_TRY(hipMemcpyAsync(hrloc.data(), P->receiver_location, hrloc.size() * 4, hipMemcpyDeviceToHost, stream),
            "batch: copy receiver locations");
    DGX_TRY(hipStreamSynchronize(stream), "batch: stream synchronize");
    for (size_t q = 0; q < hsloc.size(); q += 3)
        if (!loc_ok(g, &hsloc[q])) {
            set_error("batch: source location outside the field grid");
            return -1;
        }
    for (size_t q = 0; q < hrloc.size(); q += 3)
        if (!loc_ok(g, &hrloc[q])) {
            set_error("batch: receiver location outside the field grid");
            return -1;
        }
    if (!tm2d_applies(P, pl, hsloc, hrloc) || P->receiver_component != 2) {
        set_error("batch: only 2D Ez-TM solves (SOLVER_TM2D, Ez source and receivers on k = 0) can be batched");
        return -1;
    }
    const bool gradient = G != nullptr;
    if (gradient && (!pl.history || !pl.e_only || P->sampling_interval != 1 || !P->E_saved || !P->R_saved)) {
        set_error("batch: gradient instances need an E-only history at sampling interval 1");
        return -1;
    }
    auto inst = std::make_unique<BatchInstance>();
    BatchInstance& I = *inst;
    I.id = b->next_id++;
    I.order = pl.order;
    I.gradient = gradient;
    launch_coefficients(g, P->eps_r_pad, P->sigma_pad, P->mu_r_pad, P->update_coeff, P->dt, stream);
    if (I.tms.setup(pl, P, stream, true, kTmBI)) {  // the batch kernels use kTmBI-row tiles
        set_error("batch: out of device memory for an instance");
        return -1;
    }
    const Tm2dGrid& tg = I.tms.g;
    TmBatchInst& d = I.d;
    d.g = tg;
    d.f = I.tms.f;
    d.pp = I.tms.pp;
    std::vector<int> soff, roff;
    std::vector<int2> sent, rent;
    tm_source_lists(tg, hsloc, P->nsource, soff, sent);
    tm_receiver_lists(tg, hrloc, P->nreceiver, roff, rent);
    if (!upload(I.l_soff, soff, stream) || !upload(I.l_sent, sent, stream) || !upload(I.l_roff, roff, stream) ||
        !upload(I.l_rent, rent, stream)) {
        set_error("batch: out of device memory for lists");
        return -1;
    }
    d.src.block_off = (const int*)I.l_soff.p;
    d.src.entries = (const int2*)I.l_sent.p;
    d.src.count = (int)sent.size();
    d.src.wave = P->source_waveform;
    d.src.nt = P->nt;
    d.src.comp = 2;
    d.src.dl = g.d[2];
    d.src.vol = g.d[0] * g.d[1] * g.d[2];
    d.rx.off = (const int*)I.l_roff.p;
    d.rx.ent = (const int2*)I.l_rent.p;
    d.rx.out = P->receiver_data;
    d.rx.nrx = P->nreceiver;
    d.rx.nt = P->nt;
    d.rx.count = (int)rent.size();
    d.mode = pl.comps == 3 ? 3 : 2;
    d.box = pl.box;
    d.nt = P->nt;
    d.dt = P->dt;
    I.rx_out = P->receiver_data;
    I.rx_count = (long long)g.nshot * P->nt * P->nreceiver;
    if (gradient) {
        d.e_saved = P->E_saved;
        d.e_final = P->R_saved;
        d.snap = pl.snap;
        d.cstride = pl.cstride;
        d.kind = pl.kind;
        d.nt_saved = pl.nt_saved;
        d.sig_full = P->sigma_pad;
        d.want_eps = G->eps_r_requires_grad == 1;
        d.want_sig = G->sigma_requires_grad == 1;
        I.grad_eps = G->grad_eps_r;
        I.grad_sig = G->grad_sigma;
        if ((d.want_eps && !I.grad_eps) || (d.want_sig && !I.grad_sig)) {
            set_error("batch: null gradient output");
            return -1;
        }
        std::vector<int> goff, gshot, gmem, xoff;
        std::vector<long long> gcell;
        std::vector<int4> xgrp;
        receiver_groups(hrloc, g.nshot, P->nreceiver, g.N[1], 1, goff, gshot, gcell, gmem);
        tm_receiver_groups(tg, pl.order, goff, gshot, gcell, xoff, xgrp);
        if (!upload(I.l_xoff, xoff, stream) || !upload(I.l_xgrp, xgrp, stream) || !upload(I.l_mem, gmem, stream) ||
            !I.dgrad.alloc((size_t)I.rx_count * sizeof(float))) {
            set_error("batch: out of device memory for receiver groups");
            return -1;
        }
        d.rxa.off = (const int*)I.l_xoff.p;
        d.rxa.grp = (const int4*)I.l_xgrp.p;
        d.rxa.member = (const int*)I.l_mem.p;
        d.rxa.dgrad = (const float*)I.dgrad.p;
        d.rxa.nrx = P->nreceiver;
        d.rxa.nt = P->nt;
        d.rxa.count = (int)xgrp.size();
        I.ngrad = (long long)(g.N[0] - 1) * (g.N[1] - 1) * (g.N[2] - 1);
        d.terms_len = 2 * (long long)g.nshot * pl.box.cells;
        if (!I.acc.alloc((size_t)I.ngrad * 2 * sizeof(double)) ||
            !I.terms.alloc((size_t)2 * d.terms_len * sizeof(float))) {
            set_error("batch: out of device memory for gradient accumulators");
            return -1;
        }
        DGX_TRY(hipMemsetAsync(I.acc.p, 0, (size_t)I.ngrad * 2 * sizeof(double), stream), "batch: zero accumulators");
        d.geps = (double*)I.acc.p;
        d.gsig = (double*)I.acc.p + I.ngrad;
        d.terms = (float*)I.terms.p;
    }
    d.phase = 0;
    d.start = b->gstep;
    DGX_TRY(hipGetLastError(), "batch: kernel launch");
    if (stream != b->stream) {
        hipEvent_t ready;
        DGX_TRY(hipEventCreateWithFlags(&ready, hipEventDisableTiming), "batch: event");
        hipError_t e = hipEventRecord(ready, stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, ready, 0);
        (void)hipEventDestroy(ready);
        DGX_TRY(e, "batch: order the set-up before the batch");
        I.free_after(b->stream);
    }
    b->active.push_back(std::move(inst));
    b->dirty = true;
    return b->active.back()->id;
}

int batch_rebuild(Batch* b)
{
    // one table: per slot, 256-byte aligned descriptors then prefix sums
    std::vector<TmBatchInst> desc[6];
    std::vector<int> prefix[6];
    size_t off_desc[6], off_prefix[6], bytes = 0;
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    for (int slot = 0; slot < 6; ++slot) {
        prefix[slot].assign(1, 0);
        for (auto& ip : b->active) {
            const BatchInstance& I = *ip;
            if (order_slot(I.order) + 3 * I.d.phase != slot) continue;
            desc[slot].push_back(I.d);
            prefix[slot].push_back(prefix[slot].back() + I.d.g.ftiles * I.d.g.nshot);
        }
        b->ninst[slot] = (int)desc[slot].size();
        b->total[slot] = prefix[slot].back();
        off_desc[slot] = bytes;
        bytes = align(bytes + desc[slot].size() * sizeof(TmBatchInst));
        off_prefix[slot] = bytes;
        bytes = align(bytes + prefix[slot].size() * sizeof(int));
    }
    if (bytes > b->table_cap) {
        // queued steps still read the old table: retire it behind an event
        if (b->table && batch_retire(b, nullptr, std::move(b->table))) return -1;
        const size_t cap = std::max(bytes, 2 * b->table_cap);
        b->table = std::make_unique<DevBuf>();
        if (!b->table->alloc(cap)) return -1;
        b->table_cap = cap;
    }
    Staging& r = b->ring[b->ring_next];
    b->ring_next = (b->ring_next + 1) % 8;
    if (r.done && hipEventSynchronize(r.done) != hipSuccess) return -1;
    if (bytes > r.cap) {
        if (r.host) (void)hipHostFree(r.host);
        r.host = nullptr;
        const size_t cap = std::max(bytes, (size_t)64 << 10);
        if (hipHostMalloc(&r.host, cap, hipHostMallocDefault) != hipSuccess) {
            r.cap = 0;
            return -1;
        }
        r.cap = cap;
    }
    if (!r.done && hipEventCreateWithFlags(&r.done, hipEventDisableTiming) != hipSuccess) return -1;
    char* h = (char*)r.host;
    char* dv = (char*)b->table->p;
    for (int slot = 0; slot < 6; ++slot) {
        std::memcpy(h + off_desc[slot], desc[slot].data(), desc[slot].size() * sizeof(TmBatchInst));
        std::memcpy(h + off_prefix[slot], prefix[slot].data(), prefix[slot].size() * sizeof(int));
        b->desc[slot] = (const TmBatchInst*)(dv + off_desc[slot]);
        b->prefix[slot] = (const int*)(dv + off_prefix[slot]);
    }
    if (hipMemcpyAsync(dv, h, bytes, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipEventRecord(r.done, b->stream) != hipSuccess)
        return -1;
    b->dirty = false;
    return 0;
}

// An instance whose phase ended at the current global step: start its
// adjoint, or finish it.
int batch_phase_end(Batch* b, BatchInstance& I, bool& finished)
{
    hipStream_t stream = b->stream;
    TmBatchInst& d = I.d;
    finished = false;
    if (d.phase == 0 && I.gradient) {
        const long long field = (long long)d.g.nshot * d.g.plane;
        const float* ez = (d.nt & 1) ? d.pp.Ez[1] : d.pp.Ez[0];
        HistWrite hw{};
        hw.mode = d.mode;
        hw.box = d.box;
        hw.e_kind = d.kind;
        hw.e_base = d.e_final;
        hw.e_cstride = d.snap;
        launch_tm_final_frame(d.g, ez, hw, stream);
        launch_scale(I.rx_out, (float*)I.dgrad.p, 2.0f, I.rx_count, stream);
        // the adjoint starts from zero cotangents in both ping-pong copies:
        // the six field copies are one block of the state buffer, the CPML
        // auxiliaries are the home arrays plus one block of second copies
        ClearSpans sp{};
        bool fits = sp.add((float*)I.tms.buf.p, 6 * field);
        long long alt_len = 0;
        for (int f4 = 0; f4 < 4; ++f4) {
            fits = fits && sp.add(I.tms.home_e[f4], I.tms.psi_len_e[f4]) && sp.add(I.tms.home_h[f4], I.tms.psi_len_h[f4]);
            alt_len += I.tms.psi_len_e[f4] + I.tms.psi_len_h[f4];
        }
        fits = fits && sp.add((float*)I.tms.alt.p, alt_len);
        if (!fits) {
            set_error("batch: too many arrays to clear");
            return -1;
        }
        launch_clear_spans(sp, stream);
        d.phase = 1;
        d.start = b->gstep;
        return 0;
    }
    if (d.phase == 1) {
        // the last saved step (it = 0) has no later launch to sum it
        const float* last = d.terms + (size_t)((d.nt - 1) & 1) * d.terms_len;
        launch_tm_grad_reduce(d.g, d.box, last, d.want_eps, d.want_sig, d.geps, d.gsig, true, stream);
        if (d.want_eps) launch_add_f64(d.geps, I.grad_eps, I.ngrad, stream);
        if (d.want_sig) launch_add_f64(d.gsig, I.grad_sig, I.ngrad, stream);
    }
    finished = true;
    return 0;
}

int batch_step(Batch* b, int steps)
{
    CallStream call(b->stream);
    sweep_graphs();
    for (int k = 0; k < steps && !b->active.empty();) {
        if (b->dirty) {
            batch_drop_graph(b);
            if (batch_rebuild(b)) {
                set_error("batch: out of device memory for descriptors");
                return -1;
            }
        }
        // steps before the first phase end: a whole graph of them replays
        // kBatchGraphSteps steps at once, the phase ends are handled after it
        long long room = std::numeric_limits<long long>::max();
        for (const auto& ip : b->active) room = std::min<long long>(room, (long long)ip->d.start + ip->d.nt - b->gstep);
        int adv = 1;
        if (batch_graphs() && room >= kBatchGraphSteps && steps - k >= kBatchGraphSteps && batch_graph_ready(b)) {
            launch_step_counter((int*)b->gbase.p, b->gstep, false, b->stream);
            DGX_TRY(hipGraphLaunch(b->gexec, b->stream), "batch: graph launch");
            adv = kBatchGraphSteps;
        } else {
            for (int slot = 0; slot < 6; ++slot)
                launch_tm_batch(b->order_of[slot], slot >= 3, b->desc[slot], b->prefix[slot], b->ninst[slot],
                                b->total[slot], b->gstep, b->stream);
        }
        b->gstep += adv;
        k += adv;
        for (size_t q = 0; q < b->active.size();) {
            BatchInstance& I = *b->active[q];
            if (b->gstep - I.d.start < I.d.nt) {
                ++q;
                continue;
            }
            bool finished = false;
            if (batch_phase_end(b, I, finished)) return -1;
            b->dirty = true;
            if (!finished) {
                ++q;
                continue;
            }
            hipEvent_t ev;
            DGX_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "batch: event");
            DGX_TRY(hipEventRecord(ev, b->stream), "batch: event");
            b->finished.push_back({I.id, ev});
            std::unique_ptr<BatchInstance> done = std::move(b->active[q]);
            b->active.erase(b->active.begin() + (long)q);
            if (batch_retire(b, std::move(done), nullptr)) return -1;
        }
    }
    batch_sweep(b);
    DGX_TRY(hipGetLastError(), "batch: kernel launch");
    return (int)b->active.size();
}

DGX_API int dgx_forward(const dgx_solve* solve, dgx_stream_t stream)
{
    g_error[0] = '\0';
    return run_forward(solve, (hipStream_t)stream);
}

DGX_API int dgx_forward_noalloc(const dgx_solve* solve, dgx_stream_t stream)
{
    g_error[0] = '\0';
    void* fixed = noalloc_scratch();
    if (!fixed) return -1;
    return run_forward(solve, (hipStream_t)stream, fixed, kNoAllocScratchBytes);
}

DGX_API long long dgx_forward_noalloc_scratch_bytes(void) { return (long long)kNoAllocScratchBytes; }

DGX_API int dgx_tm2d_tile_rows_tuned(int nx_fields, int ny_fields, int nshot, int order, int cpml)
{
    const TuneKey key{nx_fields, ny_fields, nshot, order, cpml ? 1 : 0};
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (const auto& kv : g_tuned)
        if (kv.first == key) return kv.second;
    return 0;
}

DGX_API int dgx_backward(const dgx_solve* solve, dgx_stream_t stream)
{
    g_error[0] = '\0';
    return run_backward(solve, (hipStream_t)stream);
}

DGX_API void* dgx_batch_create(dgx_stream_t stream)
{
    g_error[0] = '\0';
    Batch* b = new (std::nothrow) Batch();
    if (!b) {
        set_error("batch: out of host memory");
        return nullptr;
    }
    b->stream = (hipStream_t)stream;
    return b;
}

DGX_API int dgx_batch_add(void* batch, const dgx_solve* forward, const dgx_solve* gradient,
                          dgx_stream_t setup_stream)
{
    g_error[0] = '\0';
    if (!batch || !forward) {
        set_error("batch: null argument");
        return -1;
    }
    Batch* b = (Batch*)batch;
    return batch_add(b, forward, gradient, setup_stream ? (hipStream_t)setup_stream : b->stream);
}

DGX_API int dgx_batch_step(void* batch, int steps)
{
    g_error[0] = '\0';
    if (!batch) {
        set_error("batch: null argument");
        return -1;
    }
    return batch_step((Batch*)batch, steps);
}

DGX_API int dgx_batch_poll(void* batch, int id)
{
    g_error[0] = '\0';
    Batch* b = (Batch*)batch;
    if (!b) return -1;
    for (size_t q = 0; q < b->finished.size(); ++q) {
        if (b->finished[q].first != id) continue;
        const hipError_t e = hipEventQuery(b->finished[q].second);
        if (e == hipErrorNotReady) return 0;
        (void)hipEventDestroy(b->finished[q].second);
        b->finished.erase(b->finished.begin() + (long)q);
        if (e != hipSuccess) {
            set_error("batch: %s", hipGetErrorString(e));
            return -1;
        }
        return 1;
    }
    for (auto& ip : b->active)
        if (ip->id == id) return 0;
    set_error("batch: unknown or already collected instance %d", id);
    return -1;
}

DGX_API int dgx_batch_active(void* batch) { return batch ? (int)((Batch*)batch)->active.size() : 0; }

DGX_API void dgx_batch_destroy(void* batch)
{
    Batch* b = (Batch*)batch;
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    for (auto& f : b->finished) (void)hipEventDestroy(f.second);
    {
        CallStream call(b->stream);
        b->active.clear();
        b->table.reset();
    }
    delete b;  // the retired memory too: every event has completed
}

DGX_API int dgx_codec_roundtrip(const float* input, void* encoded, float* decoded, long long count,
                                int storage_kind, int conversion_backend, dgx_stream_t stream)
{
    g_error[0] = '\0';
    if (count <= 0) return 0;
    if (!input || !encoded || !decoded) {
        set_error("codec: null pointer");
        return -1;
    }
    hipLaunchKernelGGL(k_codec, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, (hipStream_t)stream, input,
                       encoded, decoded, count, storage_kind, conversion_backend);
    DGX_TRY(hipGetLastError(), "codec: kernel launch");
    DGX_TRY(hipStreamSynchronize((hipStream_t)stream), "codec: stream synchronize");
    return 0;
}

DGX_API long long dgx_scratch_bytes(const dgx_solve* solve, int backward)
{
    Plan pl;
    if (make_plan(solve, backward != 0, pl)) return -1;
    const Grid& g = pl.g;
    long long bytes = 0;
    if (!backward) {
        if (pl.history && pl.store_r && pl.kind == kInt8) bytes += (long long)pl.comps * pl.snap * 4;
        if (xmarch3d()) {
            const XmTiles xt = xm_tiles(g, pl.order);
            bytes += ((long long)xt.nty * xt.ntz * g.nshot + 1) * 4;
        } else {
            bytes += (cell_sweep(g, pl.order).blocks + 1) * 4;
        }
    } else {
        if (pl.has_cpml)
            for (int f = 0; f < 6; ++f)
                for (int w = 0; w < 2; ++w) bytes += 4 * (psi_elems(g, true, f, w) + psi_elems(g, false, f, w));
        if (solve->accumulate_fp64) bytes += 16LL * (g.N[0] - 1) * (g.N[1] - 1) * (g.N[2] - 1);
        // adjoint channel weights of the general (non-2D-TM) path
        const bool tm_shape = (solve->storage_type & kFlagTm2d) && g.N[2] == 2 && solve->source_component == 2 &&
                              pl.kind != kInt8;
        if (!tm_shape && !xmarch3d()) bytes += (adj3d_fused() ? 48LL : 24LL) * g.nshot * g.cells;
    }
    return bytes;
}

DGX_API void dgx_set_kernel_timing(int enable)
{
    std::lock_guard<std::mutex> lock(g_timing_mu);
    g_timing_on = enable ? 1 : 0;
    if (!enable) return;  // totals stay readable after recording stops
    for (int k = 0; k < kKCount; ++k) {
        g_timing_ms[k] = 0.0;
        g_timing_count[k] = 0;
    }
}

DGX_API int dgx_kernel_timing(int kernel, double* total_ms, long long* launches)
{
    if (kernel < 0 || kernel >= kKCount) return -1;
    std::lock_guard<std::mutex> lock(g_timing_mu);
    if (total_ms) *total_ms = g_timing_ms[kernel];
    if (launches) *launches = g_timing_count[kernel];
    return 0;
}

DGX_API const char* dgx_kernel_name(int kernel)
{
    return (kernel >= 0 && kernel < kKCount) ? kKernelNames[kernel] : "";
}

// --- reference ABI-6 export set ----------------------------------------------

DGX_API int deepgpr_supports_error_reporting(void) { return 1; }
DGX_API const char* deepgpr_last_error(void) { return g_error; }
DGX_API void deepgpr_clear_last_error(void) { g_error[0] = '\0'; }
DGX_API void deepgpr_clear_cuda_resource_cache(void) {}
DGX_API int deepgpr_deterministic_adjoint(void) { return 1; }
DGX_API int deepgpr_abi_version(void) { return DEEPGPR_ABI_VERSION; }
DGX_API int deepgpr_supports_external_pml(void) { return 1; }
DGX_API int deepgpr_supports_pml_material_gradients(void) { return 1; }
DGX_API int deepgpr_supports_int8_wavefield(void) { return 1; }
DGX_API int deepgpr_supports_conversion_backends(void) { return 1; }
DGX_API int deepgpr_supports_int8_reduction_backends(void) { return 1; }
DGX_API int deepgpr_supports_rhs_reconstruction(void) { return 1; }
DGX_API int deepgpr_supports_physical_history(void) { return 1; }
DGX_API int deepgpr_supports_tm2d_fast_path(void) { return 1; }

DGX_API void deepgpr_test_wavefield_conversion(const float* input, void* encoded, float* decoded, long long count,
                                               int storage_kind, int conversion_backend)
{
    (void)dgx_codec_roundtrip(input, encoded, decoded, count, storage_kind, conversion_backend, nullptr);
}

DGX_API void set_fdtd_order(int order) { g_abi6_order = (order == 4 || order == 8) ? order : 2; }

DGX_API void forward(
    const float* eps_r_pad, const float* sigma_pad, const float* mu_r_pad, void* E_saved, void* R_saved,
    float* Ex, float* Ey, float* Ez, float* Hx, float* Hy, float* Hz, float* ce_hist, float* ce_curl,
    float* ce_rhs, float* ch_hist, float* ch_curl, float* ch_rhs, float* x0_e_phi1, float* x0_e_phi2,
    float* x0_h_phi1, float* x0_h_phi2, float* xm_e_phi1, float* xm_e_phi2, float* xm_h_phi1, float* xm_h_phi2,
    float* y0_e_phi1, float* y0_e_phi2, float* y0_h_phi1, float* y0_h_phi2, float* ym_e_phi1, float* ym_e_phi2,
    float* ym_h_phi1, float* ym_h_phi2, float* z0_e_phi1, float* z0_e_phi2, float* z0_h_phi1, float* z0_h_phi2,
    float* zm_e_phi1, float* zm_e_phi2, float* zm_h_phi1, float* zm_h_phi2, int pml_x0, int pml_xm, int pml_y0,
    int pml_ym, int pml_z0, int pml_zm, const float* x0_e_coeff, const float* xm_e_coeff, const float* y0_e_coeff,
    const float* ym_e_coeff, const float* z0_e_coeff, const float* zm_e_coeff, const float* x0_h_coeff,
    const float* xm_h_coeff, const float* y0_h_coeff, const float* ym_h_coeff, const float* z0_h_coeff,
    const float* zm_h_coeff, float dt, int nt, int nshot, int nreceiver, float dx, float dy, float dz,
    const int* receiver_location, float* receiver_data, int receiver_component, int nx_fields, int ny_fields,
    int nz_fields, int nsource, const int* source_location, const float* source_waveform, int source_component,
    int sampling_interval, int fwi_mode, int storage_type, int save_model_history, int use_async_offload)
{
    dgx_solve s;
    std::memset(&s, 0, sizeof(s));
    s.abi_version = DGX_ABI_VERSION;
    s.fdtd_order = g_abi6_order;
    s.eps_r_pad = eps_r_pad;
    s.sigma_pad = sigma_pad;
    s.mu_r_pad = mu_r_pad;
    s.E_saved = E_saved;
    s.R_saved = R_saved;
    float* E[3] = {Ex, Ey, Ez};
    float* H[3] = {Hx, Hy, Hz};
    for (int c = 0; c < 3; ++c) {
        s.E[c] = E[c];
        s.H[c] = H[c];
    }
    float* uc[6] = {ce_hist, ce_curl, ce_rhs, ch_hist, ch_curl, ch_rhs};
    for (int c = 0; c < 6; ++c) s.update_coeff[c] = uc[c];
    float* phi[24] = {x0_e_phi1, x0_e_phi2, x0_h_phi1, x0_h_phi2, xm_e_phi1, xm_e_phi2, xm_h_phi1, xm_h_phi2,
                      y0_e_phi1, y0_e_phi2, y0_h_phi1, y0_h_phi2, ym_e_phi1, ym_e_phi2, ym_h_phi1, ym_h_phi2,
                      z0_e_phi1, z0_e_phi2, z0_h_phi1, z0_h_phi2, zm_e_phi1, zm_e_phi2, zm_h_phi1, zm_h_phi2};
    for (int c = 0; c < 24; ++c) s.phi[c] = phi[c];
    const float* pc[12] = {x0_e_coeff, xm_e_coeff, y0_e_coeff, ym_e_coeff, z0_e_coeff, zm_e_coeff,
                           x0_h_coeff, xm_h_coeff, y0_h_coeff, ym_h_coeff, z0_h_coeff, zm_h_coeff};
    for (int c = 0; c < 12; ++c) s.pml_coeff[c] = pc[c];
    int pml[6] = {pml_x0, pml_xm, pml_y0, pml_ym, pml_z0, pml_zm};
    for (int f = 0; f < 6; ++f) s.pml[f] = pml[f];
    s.dt = dt;
    s.nt = nt;
    s.nshot = nshot;
    s.nreceiver = nreceiver;
    s.dx = dx;
    s.dy = dy;
    s.dz = dz;
    s.receiver_location = receiver_location;
    s.receiver_data = receiver_data;
    s.receiver_component = receiver_component;
    s.nx_fields = nx_fields;
    s.ny_fields = ny_fields;
    s.nz_fields = nz_fields;
    s.nsource = nsource;
    s.source_location = source_location;
    s.source_waveform = source_waveform;
    s.source_component = source_component;
    s.sampling_interval = sampling_interval;
    s.fwi_mode = fwi_mode;
    s.storage_type = storage_type;
    s.save_model_history = save_model_history;
    s.use_async_offload = use_async_offload;
    // the reference export set returns with the null stream drained
    if (dgx_forward(&s, nullptr) == 0 && hipStreamSynchronize(nullptr) != hipSuccess)
        set_error("forward: stream synchronize");
}

DGX_API void backward(
    const float* eps_r_pad, const float* sigma_pad, const float* mu_r_pad, const void* E_saved, const void* R_saved,
    float* lambda_ex, float* lambda_ey, float* lambda_ez, float* lambda_hx, float* lambda_hy, float* lambda_hz,
    float* ce_hist, float* ce_curl, float* ce_rhs, float* ch_hist, float* ch_curl, float* ch_rhs,
    float* l_x0_e_phi1, float* l_x0_e_phi2, float* l_x0_h_phi1, float* l_x0_h_phi2, float* l_xm_e_phi1,
    float* l_xm_e_phi2, float* l_xm_h_phi1, float* l_xm_h_phi2, float* l_y0_e_phi1, float* l_y0_e_phi2,
    float* l_y0_h_phi1, float* l_y0_h_phi2, float* l_ym_e_phi1, float* l_ym_e_phi2, float* l_ym_h_phi1,
    float* l_ym_h_phi2, float* l_z0_e_phi1, float* l_z0_e_phi2, float* l_z0_h_phi1, float* l_z0_h_phi2,
    float* l_zm_e_phi1, float* l_zm_e_phi2, float* l_zm_h_phi1, float* l_zm_h_phi2, int pml_x0, int pml_xm,
    int pml_y0, int pml_ym, int pml_z0, int pml_zm, const float* x0_e_coeff, const float* xm_e_coeff,
    const float* y0_e_coeff, const float* ym_e_coeff, const float* z0_e_coeff, const float* zm_e_coeff,
    const float* x0_h_coeff, const float* xm_h_coeff, const float* y0_h_coeff, const float* ym_h_coeff,
    const float* z0_h_coeff, const float* zm_h_coeff, float dt, int nt, int nshot, int nreceiver, float dx,
    float dy, float dz, int nx_fields, int ny_fields, int nz_fields, int ndata_source, const int* receiver_location,
    const float* data_grad, int receiver_component, int nsource, const int* source_location, int source_component,
    float* grad_source, int source_requires_grad, float* grad_eps_r, float* grad_sigma, int eps_r_requires_grad,
    int sigma_requires_grad, int sampling_interval, int fwi_mode, int storage_type, int use_async_offload)
{
    (void)nreceiver;
    dgx_solve s;
    std::memset(&s, 0, sizeof(s));
    s.abi_version = DGX_ABI_VERSION;
    s.fdtd_order = g_abi6_order;
    s.eps_r_pad = eps_r_pad;
    s.sigma_pad = sigma_pad;
    s.mu_r_pad = mu_r_pad;
    s.E_saved = (void*)E_saved;
    s.R_saved = (void*)R_saved;
    float* E[3] = {lambda_ex, lambda_ey, lambda_ez};
    float* H[3] = {lambda_hx, lambda_hy, lambda_hz};
    for (int c = 0; c < 3; ++c) {
        s.E[c] = E[c];
        s.H[c] = H[c];
    }
    float* uc[6] = {ce_hist, ce_curl, ce_rhs, ch_hist, ch_curl, ch_rhs};
    for (int c = 0; c < 6; ++c) s.update_coeff[c] = uc[c];
    float* phi[24] = {l_x0_e_phi1, l_x0_e_phi2, l_x0_h_phi1, l_x0_h_phi2, l_xm_e_phi1, l_xm_e_phi2,
                      l_xm_h_phi1, l_xm_h_phi2, l_y0_e_phi1, l_y0_e_phi2, l_y0_h_phi1, l_y0_h_phi2,
                      l_ym_e_phi1, l_ym_e_phi2, l_ym_h_phi1, l_ym_h_phi2, l_z0_e_phi1, l_z0_e_phi2,
                      l_z0_h_phi1, l_z0_h_phi2, l_zm_e_phi1, l_zm_e_phi2, l_zm_h_phi1, l_zm_h_phi2};
    for (int c = 0; c < 24; ++c) s.phi[c] = phi[c];
    const float* pc[12] = {x0_e_coeff, xm_e_coeff, y0_e_coeff, ym_e_coeff, z0_e_coeff, zm_e_coeff,
                           x0_h_coeff, xm_h_coeff, y0_h_coeff, ym_h_coeff, z0_h_coeff, zm_h_coeff};
    for (int c = 0; c < 12; ++c) s.pml_coeff[c] = pc[c];
    int pml[6] = {pml_x0, pml_xm, pml_y0, pml_ym, pml_z0, pml_zm};
    for (int f = 0; f < 6; ++f) s.pml[f] = pml[f];
    s.dt = dt;
    s.nt = nt;
    s.nshot = nshot;
    s.nreceiver = ndata_source;
    s.dx = dx;
    s.dy = dy;
    s.dz = dz;
    s.nx_fields = nx_fields;
    s.ny_fields = ny_fields;
    s.nz_fields = nz_fields;
    s.receiver_location = receiver_location;
    s.data_grad = data_grad;
    s.receiver_component = receiver_component;
    s.nsource = nsource;
    s.source_location = source_location;
    s.source_component = source_component;
    s.grad_source = grad_source;
    s.source_requires_grad = source_requires_grad;
    s.grad_eps_r = grad_eps_r;
    s.grad_sigma = grad_sigma;
    s.eps_r_requires_grad = eps_r_requires_grad;
    s.sigma_requires_grad = sigma_requires_grad;
    s.sampling_interval = sampling_interval;
    s.fwi_mode = fwi_mode;
    s.storage_type = storage_type;
    s.use_async_offload = use_async_offload;
    // the reference export set returns with the null stream drained
    if (dgx_backward(&s, nullptr) == 0 && hipStreamSynchronize(nullptr) != hipSuccess)
        set_error("backward: stream synchronize");
}

}  // extern "C"
