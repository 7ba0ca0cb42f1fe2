/*
 * deepgpr_amd: MI355X (gfx950) FDTD hot path behind a C ABI.
 *
 * Two export sets, both implemented by the same HIP kernels:
 *
 *  1. The primary, versioned API (dgx_*): one descriptor struct per call, an
 *     explicit FDTD order and an explicit HIP stream. It replaces the
 *     81/86-argument forward/backward pair plus the process-global
 *     set_fdtd_order of the reference ABI (src/DeepGPR/lib/deepgpr.h:87-152)
 *     and the legacy-default-stream ordering done by the reference bridge
 *     (src/DeepGPR/native/bridge.py:200-248).
 *
 *  2. The reference ABI-6 export set (declared at the end of this file), so
 *     the library is a literal drop-in "deepgpr.so" for the reference ctypes
 *     loader (src/DeepGPR/native/loader.py:36-46, 138-171). Those entry points
 *     launch on the null stream, like the reference CUDA backend.
 *
 * All pointers are device pointers; every float array is contiguous FP32 in
 * the reference layouts (fields (nshot, NX, NY, NZ) z-fastest with NX = nx+1).
 * Functions return 0 on success and -1 on failure; the message is then
 * available from dgx_last_error() / deepgpr_last_error() on the calling
 * thread, as in the reference (deepgpr.h:41-50).
 */
#ifndef DEEPGPR_AMD_H
#define DEEPGPR_AMD_H

#include <stddef.h>

#define DGX_ABI_VERSION 1
#define DEEPGPR_ABI_VERSION 6

#if defined(_WIN32)
#define DGX_API __declspec(dllexport)
#else
#define DGX_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque HIP stream handle (hipStream_t); NULL is the null stream. */
typedef void* dgx_stream_t;

/*
 * One solve. Field meanings follow the reference ABI parameters of the same
 * names (src/DeepGPR/solver/autograd.py:164-331, deepgpr.h:89-152):
 *   forward:  E/H are the fields, advanced in place; receiver_data is written;
 *             E_saved/R_saved receive the history.
 *   backward: E/H are the adjoint fields lambda_E/lambda_H, advanced in place;
 *             data_grad is read; grad_* are accumulated with +=.
 * phi[24] are the CPML auxiliary arrays (or their cotangents) in the native
 * order x0_e_phi1, x0_e_phi2, x0_h_phi1, x0_h_phi2, xm_e_phi1, ... zm_h_phi2.
 * pml_coeff[12] are the CPML profiles in the native order x0_e, xm_e, y0_e,
 * ym_e, z0_e, zm_e, x0_h, ..., zm_h, each (4, 1, thickness).
 * update_coeff[6] are zero-initialised scratch arrays filled with ce_hist,
 * ce_curl, ce_rhs, ch_hist, ch_curl, ch_rhs.
 */
typedef struct dgx_solve {
    int abi_version;            /* must be DGX_ABI_VERSION */
    int fdtd_order;             /* 2, 4 or 8 */
    const float* eps_r_pad;
    const float* sigma_pad;
    const float* mu_r_pad;
    void* E_saved;
    void* R_saved;
    float* E[3];
    float* H[3];
    float* update_coeff[6];
    float* phi[24];
    const float* pml_coeff[12];
    int pml[6];
    float dt;
    int nt;
    int nshot;
    int nreceiver;
    float dx, dy, dz;
    int nx_fields, ny_fields, nz_fields;
    const int* receiver_location;
    float* receiver_data;       /* forward output (nshot, nt, nreceiver) */
    const float* data_grad;     /* backward input (nshot, nt, nreceiver) */
    int receiver_component;
    int nsource;
    const int* source_location;
    const float* source_waveform;
    int source_component;
    float* grad_source;
    float* grad_eps_r;
    float* grad_sigma;
    int source_requires_grad;
    int eps_r_requires_grad;
    int sigma_requires_grad;
    int sampling_interval;
    int fwi_mode;
    int storage_type;           /* bit layout of src/DeepGPR/config/constants.py:125-200 */
    int save_model_history;
    int use_async_offload;      /* accepted, histories stay in HBM */
    int accumulate_fp64;        /* 1: time-summed material gradients in FP64 */
} dgx_solve;

/* Library version of the dgx_* API. */
DGX_API int dgx_abi_version(void);
/* Message of the last failed call on this thread ("" after success). */
DGX_API const char* dgx_last_error(void);
/* Forward solve; replaces forward() of deepgpr.h:89-115. Stream-ordered:
 * the call queues its work (and the free of its scratch) on `stream` and
 * returns without waiting, unless kernel timing is on; the caller's arrays
 * must stay valid until that work has run. */
DGX_API int dgx_forward(const dgx_solve* solve, dgx_stream_t stream);
/* Exact discrete adjoint; replaces backward() of deepgpr.h:117-152.
 * Stream-ordered like dgx_forward. */
DGX_API int dgx_backward(const dgx_solve* solve, dgx_stream_t stream);
/* No counterpart in the reference. dgx_forward for callers that cannot
 * allocate device memory (serving loops that must not allocate per request,
 * harnesses that forbid hipMalloc): it calls no hipMalloc, hipFree or other
 * allocating routine. It runs the one-thread-per-cell sweeps (results are
 * those of dgx_forward, bit for bit) and keeps its only scratch, the source
 * lists of the electric sweep, in a fixed device array of
 * dgx_forward_noalloc_scratch_bytes() bytes (4 MiB) per GPU: a solve whose
 * lists need more (about 4 bytes per 256-cell workgroup of the sweep over all
 * shots, plus 8 bytes per source) is rejected with an error, and so is an
 * INT8 history (it needs a scratch frame). Calls that overlap on one GPU
 * share the array and race. */
DGX_API int dgx_forward_noalloc(const dgx_solve* solve, dgx_stream_t stream);
DGX_API long long dgx_forward_noalloc_scratch_bytes(void);
/* No counterpart in the reference. Tile rows (4, 8 or 16) that the run-time
 * tuner chose for the fused 2D Ez-TM step kernels of one shape (field grid
 * nx_fields x ny_fields, shots, FDTD order, CPML present), or 0 when no solve
 * of that shape has run yet in this process. The first solve of a shape times
 * the forward step kernel at each height on buffers of its own and keeps the
 * fastest; every height gives the same results. */
DGX_API int dgx_tm2d_tile_rows_tuned(int nx_fields, int ny_fields, int nshot, int order, int cpml);
/* FP16/BF16 history codec round trip on device arrays; replaces
 * deepgpr_test_wavefield_conversion (deepgpr.h:84-86). */
DGX_API int dgx_codec_roundtrip(const float* input, void* encoded, float* decoded, long long count,
                                int storage_kind, int conversion_backend, dgx_stream_t stream);
/* Device memory a solve needs beyond the caller's arrays (bytes), for
 * admission by HBM footprint. No counterpart in the reference. */
DGX_API long long dgx_scratch_bytes(const dgx_solve* solve, int backward);

/* Per-launch kernel timing with HIP events on the launch stream, for
 * measurement only (no counterpart in the reference). enable=1 clears the
 * totals and starts recording; every later dgx_forward/dgx_backward adds the
 * elapsed time of each launch to its kernel's total; enable=0 stops
 * recording and keeps the totals readable. */
DGX_API void dgx_set_kernel_timing(int enable);
/* Total milliseconds and launch count recorded for one kernel id. */
DGX_API int dgx_kernel_timing(int kernel, double* total_ms, long long* launches);
/* Name of a kernel id ("" past the last id). */
DGX_API const char* dgx_kernel_name(int kernel);

/* Batched 2D Ez-TM solves. No counterpart in the reference, which batches
 * only the shots of one model on one grid (deepgpr.cu:618, 3493). Instances of
 * different grids, FDTD orders and numbers of time steps advance together:
 * every dgx_batch_step issues one launch per (order, phase) for all of them.
 * An instance is a forward solve (the dgx_solve of dgx_forward, with the
 * SOLVER_TM2D bit, Ez source and receivers) plus, when `gradient` is not
 * NULL, the eps_r/sigma gradient of sum(receivers^2): `gradient` supplies
 * grad_eps_r, grad_sigma and their *_requires_grad flags, and the forward
 * must then keep an E-only history at sampling interval 1. The instance's
 * state arrays (E, H, phi) are consumed as scratch. Time steps run on the
 * batch's stream; one host thread drives a batch. */
DGX_API void* dgx_batch_create(dgx_stream_t stream);
/* Adds an instance; returns its id (> 0) or -1 (dgx_last_error). Its set-up
 * (coefficients, lists, buffers) runs on `setup_stream` (NULL: the batch's
 * stream), which must be the stream that produced the instance's inputs; the
 * batch's stream waits for it before the instance's first step, and the host
 * waits only for the set-up stream. */
DGX_API int dgx_batch_add(void* batch, const dgx_solve* forward, const dgx_solve* gradient,
                          dgx_stream_t setup_stream);
/* Advances every instance by up to `steps` time steps; returns the number of
 * instances still running, or -1. */
DGX_API int dgx_batch_step(void* batch, int steps);
/* 1 when instance `id` has finished and its outputs are complete on the
 * device (the id is then released), 0 while it runs, -1 on error. */
DGX_API int dgx_batch_poll(void* batch, int id);
/* Number of instances still running. */
DGX_API int dgx_batch_active(void* batch);
/* Waits for the batch's stream and releases everything. */
DGX_API void dgx_batch_destroy(void* batch);

/* ---------------------------------------------------------------------- */
/* Reference ABI-6 export set (src/DeepGPR/lib/deepgpr.h:48-152)          */
/* ---------------------------------------------------------------------- */

/* deepgpr.h:48 */
DGX_API int deepgpr_supports_error_reporting(void);
/* deepgpr.h:49 */
DGX_API const char* deepgpr_last_error(void);
/* deepgpr.h:50 */
DGX_API void deepgpr_clear_last_error(void);
/* deepgpr.h:53; there is no async offload cache to release: a no-op */
DGX_API void deepgpr_clear_cuda_resource_cache(void);
/* deepgpr.h:56; the gather adjoint uses no atomics: 1 */
DGX_API int deepgpr_deterministic_adjoint(void);
/* deepgpr.h:58 */
DGX_API int deepgpr_abi_version(void);
/* deepgpr.h:59 */
DGX_API int deepgpr_supports_external_pml(void);
/* deepgpr.h:61 */
DGX_API int deepgpr_supports_pml_material_gradients(void);
/* deepgpr.h:62 */
DGX_API int deepgpr_supports_int8_wavefield(void);
/* deepgpr.h:63 */
DGX_API int deepgpr_supports_conversion_backends(void);
/* deepgpr.h:64 */
DGX_API int deepgpr_supports_int8_reduction_backends(void);
/* deepgpr.h:70 */
DGX_API int deepgpr_supports_rhs_reconstruction(void);
/* deepgpr.h:76 */
DGX_API int deepgpr_supports_physical_history(void);
/* deepgpr.h:83 */
DGX_API int deepgpr_supports_tm2d_fast_path(void);
/* deepgpr.h:84-86 (null stream, synchronous) */
DGX_API void deepgpr_test_wavefield_conversion(const float* input, void* encoded, float* decoded,
                                               long long count, int storage_kind, int conversion_backend);
/* deepgpr.h:87 */
DGX_API void set_fdtd_order(int order);

/* deepgpr.h:89-115 */
DGX_API void forward(
    const float* eps_r_pad, const float* sigma_pad, const float* mu_r_pad,
    void* E_saved, void* R_saved,
    float* Ex, float* Ey, float* Ez,
    float* Hx, float* Hy, float* Hz,
    float* ce_hist, float* ce_curl, float* ce_rhs,
    float* ch_hist, float* ch_curl, float* ch_rhs,
    float* x0_e_phi1, float* x0_e_phi2, float* x0_h_phi1, float* x0_h_phi2,
    float* xm_e_phi1, float* xm_e_phi2, float* xm_h_phi1, float* xm_h_phi2,
    float* y0_e_phi1, float* y0_e_phi2, float* y0_h_phi1, float* y0_h_phi2,
    float* ym_e_phi1, float* ym_e_phi2, float* ym_h_phi1, float* ym_h_phi2,
    float* z0_e_phi1, float* z0_e_phi2, float* z0_h_phi1, float* z0_h_phi2,
    float* zm_e_phi1, float* zm_e_phi2, float* zm_h_phi1, float* zm_h_phi2,
    int pml_x0, int pml_xm, int pml_y0, int pml_ym, int pml_z0, int pml_zm,
    const float* x0_e_coeff, const float* xm_e_coeff,
    const float* y0_e_coeff, const float* ym_e_coeff,
    const float* z0_e_coeff, const float* zm_e_coeff,
    const float* x0_h_coeff, const float* xm_h_coeff,
    const float* y0_h_coeff, const float* ym_h_coeff,
    const float* z0_h_coeff, const float* zm_h_coeff,
    float dt, int nt, int nshot, int nreceiver,
    float dx, float dy, float dz,
    const int* receiver_location, float* receiver_data, int receiver_component,
    int nx_fields, int ny_fields, int nz_fields, int nsource,
    const int* source_location, const float* source_waveform,
    int source_component, int sampling_interval, int fwi_mode, int storage_type,
    int save_model_history, int use_async_offload);

/* deepgpr.h:117-152 */
DGX_API void backward(
    const float* eps_r_pad, const float* sigma_pad, const float* mu_r_pad,
    const void* E_saved, const void* R_saved,
    float* lambda_ex, float* lambda_ey, float* lambda_ez,
    float* lambda_hx, float* lambda_hy, float* lambda_hz,
    float* ce_hist, float* ce_curl, float* ce_rhs,
    float* ch_hist, float* ch_curl, float* ch_rhs,
    float* lambda_x0_e_phi1, float* lambda_x0_e_phi2,
    float* lambda_x0_h_phi1, float* lambda_x0_h_phi2,
    float* lambda_xm_e_phi1, float* lambda_xm_e_phi2,
    float* lambda_xm_h_phi1, float* lambda_xm_h_phi2,
    float* lambda_y0_e_phi1, float* lambda_y0_e_phi2,
    float* lambda_y0_h_phi1, float* lambda_y0_h_phi2,
    float* lambda_ym_e_phi1, float* lambda_ym_e_phi2,
    float* lambda_ym_h_phi1, float* lambda_ym_h_phi2,
    float* lambda_z0_e_phi1, float* lambda_z0_e_phi2,
    float* lambda_z0_h_phi1, float* lambda_z0_h_phi2,
    float* lambda_zm_e_phi1, float* lambda_zm_e_phi2,
    float* lambda_zm_h_phi1, float* lambda_zm_h_phi2,
    int pml_x0, int pml_xm, int pml_y0, int pml_ym, int pml_z0, int pml_zm,
    const float* x0_e_coeff, const float* xm_e_coeff,
    const float* y0_e_coeff, const float* ym_e_coeff,
    const float* z0_e_coeff, const float* zm_e_coeff,
    const float* x0_h_coeff, const float* xm_h_coeff,
    const float* y0_h_coeff, const float* ym_h_coeff,
    const float* z0_h_coeff, const float* zm_h_coeff,
    float dt, int nt, int nshot, int nreceiver,
    float dx, float dy, float dz,
    int nx_fields, int ny_fields, int nz_fields,
    int ndata_source, const int* receiver_location, const float* data_grad,
    int receiver_component,
    int nsource, const int* source_location, int source_component,
    float* grad_source, int source_requires_grad,
    float* grad_eps_r, float* grad_sigma,
    int eps_r_requires_grad, int sigma_requires_grad,
    int sampling_interval, int fwi_mode, int storage_type, int use_async_offload);

#ifdef __cplusplus
}
#endif

#endif /* DEEPGPR_AMD_H */
