/*
 * psw.h -- C ABI of the MI355X (gfx950) multiview-stereo library libpsw.so.
 *
 * Three stages of dkavolis/plane-sweep-cuda, rebuilt as hand-written HIP
 * kernels for CDNA4:
 *   1. plane sweep   (reference: src/kernels.cu, src/planesweep.cpp:74-259)
 *   2. TGV2 refine   (reference: src/TGV2_kernels.cu, src/planesweep.cpp:381-536)
 *   3. TV-L1 histogram fusion (reference: src/fusion.cu, src/inc/fusion.h,
 *      driver loop src/pclviewer.cpp:972-1083)
 * plus the TVL1 depth denoiser that feeds fusion in the reference
 * (src/kernels.cu:250-403, src/planesweep.cpp:297-379).
 *
 * Conventions shared by every entry point:
 *  - every function returns int status: PSW_OK (0) or a PSW_ERR_* code; no
 *    exception or abort crosses the ABI.  psw_last_error() gives the message
 *    of the last failure on the calling thread;
 *  - pointers named d_* are device pointers (dense row-major, width-indexed,
 *    exactly as the reference's Image<float> with pitch = width*sizeof(float),
 *    src/inc/memory.h:112-147); other pointers are host memory;
 *  - matrices are row-major 3x3 float[9] (the reference's Matrix3D rows,
 *    src/inc/structs.h:410-764), vectors float[3] (Vector3D);
 *  - work is enqueued on the given stream (NULL = the null stream) and is
 *    asynchronous unless the comment says otherwise;
 *  - psw_dim3 blocks/threads are the reference launchers' launch shape.  A
 *    zero shape lets the library choose; a given shape is used as is (pixels
 *    outside blocks*threads are left untouched, as in the reference).
 */
#ifndef PSW_H
#define PSW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* psw_stream_t; /* hipStream_t */
typedef void* psw_event_t;  /* hipEvent_t */

typedef struct { unsigned x, y, z; } psw_dim3;

enum {
    PSW_OK = 0,
    PSW_ERR_ARG = 1,      /* invalid argument (sizes, null pointers, even window, ...) */
    PSW_ERR_LAUNCH = 2,   /* kernel launch failed (e.g. block larger than the device allows) */
    PSW_ERR_ALLOC = 3,    /* device allocation failed */
    PSW_ERR_DEVICE = 4,   /* no device / device error (reference: NO_CUDA_DEVICE, defines.h:34) */
    PSW_ERR_COPY = 5,     /* memory copy failed */
    PSW_ERR_UNSUPPORTED = 6
};

#define PSW_MAX_SOURCES 16   /* source views per TGV2 solve / fused call */
#define PSW_FUSION_BINS 8    /* histogram bins of the fusion grid (fusionData<8>) */

/* ---------------------------------------------------------------------- */
/* Runtime plumbing (no reference counterpart: the reference uses the CUDA  */
/* runtime directly; these give FFI callers device memory and streams).      */
/* ---------------------------------------------------------------------- */
const char* psw_status_string(int status);
const char* psw_last_error(void);
int psw_device_count(int* count);             /* cudaDevInit, src/planesweep.cpp:30-58 */
int psw_set_device(int device);
int psw_device_name(char* buf, size_t n);
int psw_malloc(void** d_ptr, size_t bytes);
int psw_free(void* d_ptr);
int psw_host_alloc(void** h_ptr, size_t bytes);  /* pinned host memory */
int psw_host_free(void* h_ptr);
int psw_memcpy_htod(void* d_dst, const void* src, size_t bytes, psw_stream_t s); /* returns after the copy */
int psw_memcpy_dtoh(void* dst, const void* d_src, size_t bytes, psw_stream_t s); /* returns after the copy */
int psw_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes, psw_stream_t s);
int psw_fill_f32(float* d_ptr, float value, size_t count, psw_stream_t s);   /* kernel fill, graph-safe */
int psw_fill_u32(uint32_t* d_ptr, uint32_t value, size_t count, psw_stream_t s);
/* Kernel copy of count floats (count % 4 == 0, 16-byte aligned), graph-safe. */
int psw_copy_f32(float* d_dst, const float* d_src, size_t count, psw_stream_t s);
int psw_stream_create(psw_stream_t* s);
int psw_stream_destroy(psw_stream_t s);
int psw_stream_synchronize(psw_stream_t s);
int psw_device_synchronize(void);
int psw_event_create(psw_event_t* e);
int psw_event_destroy(psw_event_t e);
int psw_event_record(psw_event_t e, psw_stream_t s);
int psw_event_synchronize(psw_event_t e);
int psw_event_elapsed_ms(float* ms, psw_event_t start, psw_event_t stop);

/* ---------------------------------------------------------------------- */
/* Host helpers of the reference's drivers (pure host code).                */
/* ---------------------------------------------------------------------- */
/* Plane depths of the reference's loop dstep = (zfar-znear)/(nplanes-1),
 * for (float d = znear; d <= zfar; d += dstep) (src/planesweep.cpp:169,200):
 * nplanes or nplanes-1 values, written to out[0..*count). */
int psw_reference_planes(float znear, float zfar, unsigned nplanes, float* out, int maxn, int* count);
/* Matrix3D::inv (src/inc/structs.h:584-593), float. */
int psw_mat3_inverse(const float* m, float* out);
/* PlaneSweep::RelativeMatrices (src/planesweep.cpp:538-549), float. */
int psw_relative_matrices(const float* Rref, const float* tref, const float* Rsrc, const float* tsrc,
                          int alternative, float* Rrel, float* trel);
/* Per-(source, plane) homographies of PlaneSweepThread in the reference's
 * float Matrix3D arithmetic (src/planesweep.cpp:195-203), host output
 * H[nsrc][nplanes][9]; K in the reference's 1-based pixel convention. */
int psw_sweep_homographies_reference(float* H, const float* depths, int nplanes, int nsrc, const float* K,
                                     const float* Rref, const float* tref, const float* Rsrc, const float* tsrc,
                                     int alternative);

/* ---------------------------------------------------------------------- */
/* Reference launchers, one per kernel of src/inc/kernels.cu.h.  Same        */
/* argument order and meaning as the reference, plus stream, plus status.    */
/* These reproduce the reference's arithmetic operation by operation         */
/* (including its double temporaries) so that the reference's own host       */
/* sequencing run over them gives the reference's results.                   */
/* ---------------------------------------------------------------------- */
/* kernels.cu.h:44 / kernels.cu:431 */
int psw_transform_indexes(float* d_x, float* d_y, const float* h, int width, int height,
                          psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:62 / kernels.cu:440 */
int psw_bilinear_interpolation(float* d_result, const float* d_data, const float* d_xout, const float* d_yout,
                               int M1, int M2, int N1, int N2, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:87 / kernels.cu:448 */
int psw_calcNCC(float* d_ncc, const float* d_prod_mean, const float* d_mean1, const float* d_mean2,
                const float* d_std1, const float* d_std2, float stdthresh1, float stdthresh2, int width,
                int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:107 / kernels.cu:462 */
int psw_update_arrays(float* d_depthmap, float* d_bestncc, const float* d_currentncc, float current_depth,
                      int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:129 / kernels.cu:472 */
int psw_sum_depthmap_NCC(float* d_depthmap_out, float* d_count, const float* d_depthmap, const float* d_ncc,
                         float nccthreshold, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                         psw_stream_t s);
/* kernels.cu.h:146 / kernels.cu:484 */
int psw_calculate_STD(float* d_std, const float* d_mean, const float* d_mean_of_squares, int width, int height,
                      psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:160 / kernels.cu:494 */
int psw_set_value(float* d_output, float value, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                  psw_stream_t s);
/* kernels.cu.h:176 / kernels.cu:499 */
int psw_element_multiply(float* d_output, const float* d_input1, const float* d_input2, int width, int height,
                         psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:192 / kernels.cu:508 (count 0 -> quiet NaN) */
int psw_element_rdivide(float* d_output, const float* d_input1, const float* d_input2, int width, int height,
                        psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:226 / kernels.cu:525 */
int psw_windowed_mean_row(float* d_output, const float* d_input, unsigned winsize, int squared, int width,
                          int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:244 / kernels.cu:533 */
int psw_windowed_mean_column(float* d_output, const float* d_input, unsigned winsize, int squared, int width,
                             int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:311 / kernels.cu:565 */
int psw_element_scale(float* d_output, float scale, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                      psw_stream_t s);
/* kernels.cu.h:322 / kernels.cu:570 */
int psw_element_add(float* d_output, float value, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                    psw_stream_t s);
/* kernels.cu.h:333 / kernels.cu:575 */
int psw_set_QNAN_value(float* d_output, float value, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                       psw_stream_t s);
/* kernels.cu.h:266 / kernels.cu:547 */
int psw_denoising_TVL1_calculateP(float* d_Px, float* d_Py, const float* d_input, float sigma, int width,
                                  int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:288 / kernels.cu:555 */
int psw_denoising_TVL1_calculateP_tensor_weighed(float* d_Px, float* d_Py, const float* d_T11, const float* d_T12,
                                                 const float* d_T21, const float* d_T22, const float* d_input,
                                                 float sigma, int width, int height, psw_dim3 blocks,
                                                 psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:352 / kernels.cu:580 */
int psw_denoising_TVL1_update(float* d_output, float* d_R, const float* d_Px, const float* d_Py,
                              const float* d_origin, float tau, float theta, float lambda, float sigma, int width,
                              int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:379 / kernels.cu:589 */
int psw_denoising_TVL1_update_tensor_weighed(float* d_output, float* d_R, const float* d_Px, const float* d_Py,
                                             const float* d_origin, const float* d_T11, const float* d_T12,
                                             const float* d_T21, const float* d_T22, float tau, float theta,
                                             float lambda, float sigma, int width, int height, psw_dim3 blocks,
                                             psw_dim3 threads, psw_stream_t s);
/* kernels.cu.h:400 / kernels.cu:599 */
int psw_compute3D(float* d_x, float* d_y, float* d_z, const float* Rrel, const float* trel, const float* invK,
                  int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);

/* TGV2 launchers, kernels.cu.h:478-729 / TGV2_kernels.cu:447-567 */
int psw_TGV2_updateP(float* d_Px, float* d_Py, const float* d_u, const float* d_u1x, const float* d_u1y,
                     float alpha1, float sigma, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                     psw_stream_t s);
int psw_TGV2_updateQ(float* d_Qx, float* d_Qy, float* d_Qz, float* d_Qw, const float* d_u1x, const float* d_u1y,
                     float alpha0, float sigma, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                     psw_stream_t s);
int psw_TGV2_updateR(float* d_r, float* d_prodsum, const float* d_u, const float* d_u0, const float* d_It,
                     const float* d_Iu, float sigma, float lambda, int width, int height, psw_dim3 blocks,
                     psw_dim3 threads, psw_stream_t s);
int psw_TGV2_updateU(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar, float* d_u1ybar,
                     const float* d_Px, const float* d_Py, const float* d_Qx, const float* d_Qy, const float* d_Qz,
                     const float* d_Qw, const float* d_prodsum, float alpha0, float alpha1, float tau,
                     float lambda, int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
int psw_TGV2_transform_coordinates(float* d_x, float* d_y, float* d_X, float* d_Y, float* d_Z, const float* d_u,
                                   const float* K, const float* Rrel, const float* trel, const float* invK,
                                   int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
int psw_subtract(float* d_out, const float* d_in1, const float* d_in2, int width, int height, psw_dim3 blocks,
                 psw_dim3 threads, psw_stream_t s);
int psw_TGV2_calculate_coordinate_derivatives(float* d_dX, float* d_dY, float* d_dZ, const float* invK,
                                              const float* Rrel, int width, int height, psw_dim3 blocks,
                                              psw_dim3 threads, psw_stream_t s);
int psw_TGV2_calculate_derivativeF(float* d_dfx, float* d_dfy, const float* d_X, const float* d_dX,
                                   const float* d_Y, const float* d_dY, const float* d_Z, const float* d_dZ,
                                   float fx, float fy, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                                   psw_stream_t s);
int psw_TGV2_calculate_Iu(float* d_Iu, const float* d_I, const float* d_dfx, const float* d_dfy, int width,
                          int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
int psw_Anisotropic_diffusion_tensor(float* d_T11, float* d_T12, float* d_T21, float* d_T22, const float* d_Img,
                                     float beta, float gamma, int width, int height, psw_dim3 blocks,
                                     psw_dim3 threads, psw_stream_t s);
int psw_TGV2_updateU_tensor_weighed(float* d_u, float* d_u1x, float* d_u1y, const float* d_T11,
                                    const float* d_T12, const float* d_T21, const float* d_T22, float* d_ubar,
                                    float* d_u1xbar, float* d_u1ybar, const float* d_Px, const float* d_Py,
                                    const float* d_Qx, const float* d_Qy, const float* d_Qz, const float* d_Qw,
                                    const float* d_prodsum, float alpha0, float alpha1, float tau, float lambda,
                                    int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
int psw_TGV2_updateP_tensor_weighed(float* d_Px, float* d_Py, const float* d_T11, const float* d_T12,
                                    const float* d_T21, const float* d_T22, const float* d_u, const float* d_u1x,
                                    const float* d_u1y, float alpha1, float sigma, int width, int height,
                                    psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
/* sparse-depth TGV (TGV2_kernels.cu:367-445, work in progress in the reference) */
int psw_TGV2_updateU_sparseDepth(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar,
                                 float* d_u1ybar, const float* d_Px, const float* d_Py, const float* d_Qx,
                                 const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_w,
                                 const float* d_Ds, float alpha0, float alpha1, float tau, float theta, int width,
                                 int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
int psw_TGV2_updateU_sparseDepthTensor(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar,
                                       float* d_u1ybar, const float* d_T11, const float* d_T12, const float* d_T21,
                                       const float* d_T22, const float* d_Px, const float* d_Py, const float* d_Qx,
                                       const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_w,
                                       const float* d_Ds, float alpha0, float alpha1, float tau, float theta,
                                       int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s);
int psw_calculateWeights_sparseDepth(float* d_w, const float* d_Ds, int width, int height, psw_dim3 blocks,
                                     psw_dim3 threads, psw_stream_t s);

/* ---------------------------------------------------------------------- */
/* Stage 1: fused plane sweep.  One launch sweeps all planes and sources     */
/* over LDS-staged tiles; the cost volume never reaches HBM.                 */
/* ---------------------------------------------------------------------- */

/* Per-(source, plane) homographies H = K (Rrel + trel e3^T / d) K^-1 / H22
 * built in FP64 on the GPU and stored as float [nsrc][nplanes][9].
 * K, Rref, tref, Rsrc[nsrc*9], tsrc[nsrc*3] are host arrays; relative pose
 * as PlaneSweep::RelativeMatrices (src/planesweep.cpp:538-549) in double. */
int psw_sweep_homographies(float* d_H, const float* d_depths, int nplanes, int nsrc, const float* K,
                           const float* Rref, const float* tref, const float* Rsrc, const float* tsrc,
                           int alternative, psw_stream_t s);

/* Reference semantics (PlaneSweep::RunAlgorithm, src/planesweep.cpp:74-161):
 * per source the best NCC over planes (strict >, zero start), depth averaged
 * over sources whose best NCC > nccthresh, zfar where none.  1-based pixel
 * convention.  d_H [nsrc][nplanes][9], d_depths [nplanes].  Bit-identical to
 * the reference's float arithmetic (no FMA contraction). */
int psw_sweep_fused_reference(const float* d_ref, const float* d_src, int width, int height, int nsrc,
                              const float* d_H, const float* d_depths, int nplanes, int winsize,
                              float stdthresh, float nccthresh, float zfar, float* d_depth_out, psw_stream_t s);

/* Aggregated semantics: cost(p) = -(sum_s NCC_s(p)) / nsrc; outputs argmin
 * plane (first minimum), minimum cost and second-smallest cost per pixel.
 * coord_offset 0 = 0-based pixel convention, 1 = the reference's 1-based one.
 * fast = 0: reference-order arithmetic (bit-identical to the float oracle);
 * fast = 1: running window sums, reciprocal scaling and FMA.  With more
 * than one plane group (the default below about 9600 tiles of 32 x 16
 * pixels; PSW_SWEEP_GROUPS sets it) the groups' keys, 16 bytes per pixel,
 * come from the library's stream-ordered
 * workspace pool, which keeps freed blocks for later calls until the
 * process ends. */
int psw_sweep_fused_aggregate(const float* d_ref, const float* d_src, int width, int height, int nsrc,
                              const float* d_H, int nplanes, int winsize, float stdthresh, int coord_offset,
                              int fast, int32_t* d_argmin, float* d_cost_min, float* d_cost_second,
                              psw_stream_t s);

/* Depth of the argmin plane: d_depth[i] = d_depths[argmin[i]] (zfar if -1). */
int psw_sweep_argmin_depth(float* d_depth, const int32_t* d_argmin, const float* d_depths, int npixels,
                           float fallback, psw_stream_t s);

/* psw_sweep_fused_aggregate without any device allocation.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops).  The plane groups'
 * per-pixel keys (16 bytes per pixel at any group count) go to a static
 * device array of PSW_SWEEP_SCRATCH_BYTES (32 MiB, held by every GPU the
 * library runs on until the process ends: frames up to 2^21 pixels) instead
 * of a per-call allocation; outputs are bit-identical to
 * psw_sweep_fused_aggregate.  groups: plane groups (0 = the library's choice,
 * 1 for frames whose keys exceed the scratch; a count above 1 for such a
 * frame is rejected with PSW_ERR_ARG; 1 needs no scratch).  Every call on
 * one GPU uses that GPU's one array: calls that overlap in time on the same
 * GPU (different streams or threads) race.  Each call clears the keys it
 * uses (with a fill kernel), so nothing needs clearing between calls. */
#define PSW_SWEEP_SCRATCH_BYTES (32u << 20)
int psw_sweep_fused_aggregate_noalloc(const float* d_ref, const float* d_src, int width, int height, int nsrc,
                                      const float* d_H, int nplanes, int winsize, float stdthresh, int coord_offset,
                                      int fast, int groups, int32_t* d_argmin, float* d_cost_min,
                                      float* d_cost_second, psw_stream_t s);

/* Many frames of uneven size, source and plane counts in one persistent
 * launch (no reference counterpart): workgroups take (tile, plane group)
 * items of all frames from a device counter.  Each frame's results are
 * bit-identical to psw_sweep_fused_aggregate on that frame.  groups: plane
 * groups per tile (0 = chosen from the batch size); workgroups: resident
 * workgroups (0 = 2048).  Returns after the work is done. */
typedef struct {
    const float* d_ref;        /* [height][width], 0..255 */
    const float* d_src;        /* [nsrc][height][width] */
    int width, height, nsrc;
    const float* d_H;          /* [nsrc][nplanes][9], psw_sweep_homographies */
    int nplanes;
    int32_t* d_argmin;         /* outputs, [height][width] */
    float* d_cost_min;
    float* d_cost_second;
} psw_sweep_frame;
int psw_sweep_fused_aggregate_batch(const psw_sweep_frame* frames, int nframes, int winsize, float stdthresh,
                                    int coord_offset, int fast, int groups, int workgroups, psw_stream_t s);

/* ---------------------------------------------------------------------- */
/* Stage 2: TGV2 solver.  State lives in one device allocation; one fused    */
/* kernel per primal-dual iteration (P, Q, R, U of src/planesweep.cpp:485-509)*/
/* ---------------------------------------------------------------------- */
typedef struct psw_tgv psw_tgv; /* opaque */

typedef struct {
    double lambda, alpha0, alpha1, tau, sigma, beta, gamma; /* PlaneSweep::TGV arguments */
} psw_tgv_params;

/* Allocate a solver for width x height with nsrc source views. */
int psw_tgv_create(psw_tgv** out, int width, int height, int nsrc);
int psw_tgv_destroy(psw_tgv* t);
/* Device bytes held by a solver of this shape. */
size_t psw_tgv_state_bytes(int width, int height, int nsrc);
/* Load images (0..255 intensities, device), scale by 1/255 and build the
 * anisotropic diffusion tensor (src/planesweep.cpp:418-436). */
int psw_tgv_set_images(psw_tgv* t, const float* d_ref, const float* d_src, double beta, double gamma,
                       psw_stream_t s);
/* u := d_u (device) or, if d_u is NULL, u := value (cold start: 1.0). */
int psw_tgv_set_u(psw_tgv* t, const float* d_u, float value, psw_stream_t s);
/* Start a warp: u0 := u, duals and u1 := 0, It/Iu of every source at u0
 * (src/planesweep.cpp:445-483).  K (1-based convention), Rrel[nsrc*9],
 * trel[nsrc*3] host arrays. */
int psw_tgv_linearise(psw_tgv* t, const float* K, const float* Rrel, const float* trel, psw_stream_t s);
/* Warm start within the current linearisation (no counterpart in the
 * reference, which always restarts a warp from u0 with u1 = 0): u := d_u
 * (device, W*H) or u_value when d_u is NULL, u1 := d_u1 (device, 2*W*H:
 * u1x then u1y) or 0 when NULL; duals p, q, r := 0.  u0, It and Iu stay. */
int psw_tgv_set_primal(psw_tgv* t, const float* d_u, float u_value, const float* d_u1, psw_stream_t s);
/* Load the duals (no counterpart in the reference; for warm starts and
 * state restore): p (2*W*H: px, py), q (3*W*H: qx, qy, qz), r (nsrc*W*H),
 * device arrays; a NULL array sets that dual to 0. */
int psw_tgv_set_duals(psw_tgv* t, const float* d_p, const float* d_q, const float* d_r, psw_stream_t s);
/* Pixels counted by psw_tgv_energy / psw_tgv_solve (no reference
 * counterpart).  A pixel is left out when q = tau sigma lambda^2 sum_j Iu_j^2
 * exceeds qmax there or at a 4-neighbour: with the reference's fixed step
 * sizes the primal-dual iteration need not converge at such pixels (it can
 * oscillate without end), and their energy would hide the convergence of
 * the rest.  qmax <= 0 counts every pixel (the default, restored by each
 * psw_tgv_linearise).  excluded (nullable, synchronous when set) receives
 * the number of pixels left out. */
int psw_tgv_set_energy_mask(psw_tgv* t, double qmax, const psw_tgv_params* prm, int64_t* excluded,
                            psw_stream_t s);
/* Run niters fused primal-dual iterations. */
int psw_tgv_iterate(psw_tgv* t, int niters, const psw_tgv_params* prm, psw_stream_t s);
/* FP64 primal energy of the current iterate (synchronous). */
int psw_tgv_energy(psw_tgv* t, const psw_tgv_params* prm, double* energy, psw_stream_t s);
/* Iterate until (E - e_star) / e_star <= tol (checked every check_every
 * iterations, FP64 reductions on the GPU) or max_iters: e_star is the lowest
 * energy known for the problem, and an energy below it stops too (PDHG from
 * different starts can settle on fixed points whose energies differ by more
 * than tol).  e_star <= 0 disables the test.  iters_done / energy_out
 * (nullable) report the stop. */
int psw_tgv_solve(psw_tgv* t, int max_iters, int check_every, double tol, double e_star,
                  const psw_tgv_params* prm, int* iters_done, double* energy_out, psw_stream_t s);
/* Many solvers (frames of any size and source count) per launch (no
 * reference counterpart): every iteration is one launch over the tiles of all
 * frames, each frame with its own parameters prm[f], energy target e_star[f]
 * and stop; a stopped frame's tiles become no-ops.  With use_graph (and a
 * non-NULL stream) a chunk of check_every iterations plus the energy check is
 * captured once in a hipGraph and replayed.  variant: iteration kernel shape
 * (0, 1, 2 as PSW_TGV_VARIANT; -1 = that variable or the default).  Per frame,
 * results are bit-identical to psw_tgv_iterate / psw_tgv_solve with the same
 * check_every.  iters_done[f], energy_out[f] (nullable) report the stops;
 * both calls return after the work is done. */
int psw_tgv_iterate_batch(psw_tgv* const* solvers, int nsolvers, int niters, const psw_tgv_params* prm, int variant,
                          psw_stream_t s);
int psw_tgv_solve_batch(psw_tgv* const* solvers, int nsolvers, const psw_tgv_params* prm, const double* e_star,
                        int max_iters, int check_every, double tol, int variant, int use_graph, int* iters_done,
                        double* energy_out, psw_stream_t s);
/* Copy u (depth) / full state out.  Planes of the state: u, u1x, u1y, px, py,
 * qx, qy, qz, It[N], Iu[N], r[N], T11, T12, T22, u0. */
int psw_tgv_get_u(psw_tgv* t, float* d_out, psw_stream_t s);
int psw_tgv_get_state(psw_tgv* t, float* d_out, psw_stream_t s);
/* The energy mask's weights (W*H floats: 1 where the pixel is counted, 0
 * where psw_tgv_set_energy_mask left it out; no reference counterpart). */
int psw_tgv_get_energy_mask(psw_tgv* t, float* d_out, psw_stream_t s);
/* Whole reference driver PlaneSweep::TGV: warps x niters from u_init (NULL:
 * cold start u = 1). */
int psw_tgv_run(psw_tgv* t, const float* K, const float* Rrel, const float* trel, int niters, int warps,
                const psw_tgv_params* prm, const float* d_u_init, float* d_u_out, psw_stream_t s);

/* Sparse-depth TGV2 densification, resident or fused (several iterations per launch).
 * Replaces PlaneSweep::TGVdenoiseFromSparse (src/planesweep.cpp:596-674,
 * kernels TGV2_kernels.cu:340-365, :54-85, :399-445; work in progress in the
 * reference and not called by its GUI).  d_ref: reference image 0..255;
 * d_dense: the initial depthmap (the reference's `depthmap`); d_sparse:
 * depth where known (> 0), else <= 0; output depth in d_out.  Semantics of
 * the reference driver: u1 and u1bar are the same arrays in its U step, and
 * P reads u, so theta has no effect.  d_workspace: NULL (taken from the
 * library's stream-ordered workspace pool and given back in stream order;
 * the pool keeps freed blocks for later calls until the process ends) or at
 * least psw_tgv_sparse_workspace_bytes(width, height) bytes of device
 * memory.  Runs of two or more iterations on frames of at most 256 tiles of
 * 64 x 20 (and one per CU) go to the resident solver: all iterations in one
 * cooperative launch whose tiles trade only their edges' u, u1 each
 * iteration; such a run waits for its launch (it reads the launch's error
 * words) and for other resident runs on the device, and
 * PSW_TGV_SPARSE_RESIDENT=0 turns it off.  The results are the
 * same either way. */
size_t psw_tgv_sparse_workspace_bytes(int width, int height);
int psw_tgv_sparse_denoise(const float* d_ref, const float* d_dense, const float* d_sparse, int width, int height,
                           float zfar, int niters, double alpha0, double alpha1, double tau, double sigma,
                           double theta, double beta, double gamma, void* d_workspace, size_t workspace_bytes,
                           float* d_out, psw_stream_t s);

/* ---------------------------------------------------------------------- */
/* Stage 3: histogram volumetric fusion on an SoA voxel grid.               */
/* ---------------------------------------------------------------------- */
typedef struct {
    int width, height, depth;   /* voxels (x fastest, then y, then z) */
    float volume[6];            /* bounding box a.xyz, b.xyz (Rectangle3D) */
    float* u;                   /* [V] primal */
    float* v;                   /* [V] over-relaxed primal */
    float* p;                   /* [3][V] dual */
    uint16_t* hist;             /* [V][8] bin counts */
} psw_fusion_grid;

/* Allocate (zero-filled) / free the grid arrays. */
int psw_fusion_alloc(psw_fusion_grid* g, int width, int height, int depth, const float* volume);
int psw_fusion_free(psw_fusion_grid* g);
int psw_fusion_clear(const psw_fusion_grid* g, psw_stream_t s);
/* FusionUpdateHistogram<8> (fusion.cu:84-89): one depthmap, K, world->camera R, T. */
int psw_fusion_update_histogram(const psw_fusion_grid* g, const float* d_depthmap, int width, int height,
                                const float* K, const float* R, const float* T, float threshold, psw_stream_t s);
/* Many depthmaps in one launch: d_depthmaps [m][height][width], K/R/T host
 * arrays of m*9, m*9, m*3.  Bit-identical to m single updates. */
int psw_fusion_update_histogram_batch(const psw_fusion_grid* g, const float* d_depthmaps, int m, int width,
                                      int height, const float* K, const float* R, const float* T,
                                      float threshold, psw_stream_t s);
/* FusionUpdateU<8> (fusion.cu:72-76) and FusionUpdateP<8> (fusion.cu:78-82). */
int psw_fusion_update_u(const psw_fusion_grid* g, double tau, double lambda, psw_stream_t s);
int psw_fusion_update_p(const psw_fusion_grid* g, double sigma, psw_stream_t s);
/* FusionUpdateIteration<8> (fusion.cu:91-99): histogram, U, P. */
int psw_fusion_update_iteration(const psw_fusion_grid* g, const float* d_depthmap, int width, int height,
                                const float* K, const float* R, const float* T, float threshold, double tau,
                                double lambda, double sigma, psw_stream_t s);
/* niters primal-dual iterations (U then P each), fused kernels. */
int psw_fusion_solve(const psw_fusion_grid* g, int niters, double tau, double lambda, double sigma,
                     psw_stream_t s);
/* FP64 TV-L1 histogram energy of the current u (no reference counterpart):
 * sum |grad u| + lambda sum_k h_k |u - c_k| with the P step's forward
 * differences and the bin centres; synchronous. */
int psw_fusion_energy(const psw_fusion_grid* g, double lambda, double* energy, psw_stream_t s);
/* Fused-surface points (src/pclviewer.cpp:1058-1074): world coordinates of
 * the centres of the voxels with -1 < u < 0 (worldCoords, fusion.h:360-364),
 * in the reference's z, y, x loop order, into d_xyz [capacity][3] (device).
 * *count receives the total (which may exceed capacity; only capacity points
 * are written).  Synchronous. */
int psw_fusion_surface_points(const psw_fusion_grid* g, float* d_xyz, int64_t capacity, int64_t* count,
                              psw_stream_t s);
/* Voxels with -1 < u < 0 (src/pclviewer.cpp:1062); synchronous. */
int psw_fusion_surface_count(const psw_fusion_grid* g, int64_t* count, psw_stream_t s);

/* ---------------------------------------------------------------------- */
/* TVL1 depth denoising, PlaneSweep::CudaDenoise (src/planesweep.cpp:297-379) */
/* ---------------------------------------------------------------------- */
/* d_depth (device, W*H) is denoised into d_out with the reference's scaling
 * to [0,1] over [znear, zfar]; d_ref holds 0..255 intensities for the tensor.
 * d_work: 12*W*H floats of device workspace (the solver state is
 * double-buffered).  Runs of two or more iterations on frames of at most
 * 256 tiles of 64 x 20 (and one per CU) go to the resident solver: all
 * iterations in one cooperative launch whose tiles trade only their edges'
 * depth each iteration; such a run waits for its launch (it reads the
 * launch's error words) and for other resident runs on the device, and
 * PSW_TVL1_RESIDENT=0 turns it off.  Otherwise,
 * or when the launch is called off, the fused kernel runs up to PSW_TVL1_M
 * iterations per launch (2 to 6, default 4; at most 4 with the 64x16 tiles
 * of PSW_TVL1_TY=16), a shorter launch for the remainder; PSW_TVL1_FUSED=0
 * selects the reference's two launches per iteration instead.  The results
 * are identical every way. */
int psw_tvl1_denoise(const float* d_depth, const float* d_ref, int width, int height, float znear, float zfar,
                     int niters, double lambda, double tau, double sigma, double theta, double beta, double gamma,
                     float* d_out, float* d_work, psw_stream_t s);

/* ---------------------------------------------------------------------- */
/* Seeded synthetic scenes (no reference counterpart: the reference only     */
/* ships ten ICL-NUIM frames).  A room is an axis-aligned interior box with  */
/* axis-aligned boxes inside; every surface carries a 4-octave solid value-  */
/* noise texture in [0, 1].                                                  */
/* ---------------------------------------------------------------------- */
#define PSW_SCENE_MAX_BOXES 32
typedef struct {
    double lo[3], hi[3];                    /* interior box */
    int nboxes;
    double box_lo[PSW_SCENE_MAX_BOXES][3];
    double box_hi[PSW_SCENE_MAX_BOXES][3];
    uint32_t octave_salt[4];                /* texture hash salt per octave */
} psw_room;

/* Ray-cast one pinhole view in double precision: d_depth = z in the camera
 * frame, d_img = scale * (texture + N(0, noise_sigma^2)) with a counter-based
 * Gaussian keyed by noise_seed (noise_sigma = 0: no noise).  invK: inverse of
 * the 0-based K (row-major double[9]); R, t: world->camera (X_c = R X_w + t).
 * Either output may be NULL. */
int psw_scene_render(const psw_room* room, const double* invK, const double* R, const double* t, int width,
                     int height, double noise_sigma, uint64_t noise_seed, float scale, float* d_img, float* d_depth,
                     psw_stream_t s);

/* ---------------------------------------------------------------------- */
/* Learned TGV2 warm start (no reference counterpart).  A 4-level U-Net      */
/* (3x3 convolutions, 32/48/64/96 channels, 480,264 parameters) in BF16 on   */
/* the matrix cores with FP32 accumulation; forward, backward and Adam are   */
/* hand-written kernels.  Input: the aggregated sweep of a view (reference   */
/* image, argmin plane, min and second-best cost).  Output: u, u1, p, q to   */
/* start psw_tgv_set_primal / psw_tgv_set_duals from.                        */
/* Tensors live in "slots": slot b of a batch is image b.  width and height  */
/* must be multiples of 8; batch * width * height <= max_pixels.            */
/* ---------------------------------------------------------------------- */
typedef struct psw_unet psw_unet; /* opaque */

size_t psw_unet_param_count(void);
/* Allocate for up to max_pixels = batch * width * height; parameters are
 * initialised with psw_unet_init(seed = 1). */
int psw_unet_create(psw_unet** out, int64_t max_pixels);
int psw_unet_destroy(psw_unet* n);
/* He-normal weights from a seeded host stream; the head starts near zero so
 * the untrained warm start is the sweep depth.  Synchronous. */
int psw_unet_init(psw_unet* n, uint64_t seed);
/* Flat FP32 parameters: per layer W[3][3][c_in][c_out] then bias[c_out],
 * layers in network order, head W[32][8] and bias[8] last.  Setting resets
 * Adam's moments.  Synchronous. */
int psw_unet_set_params(psw_unet* n, const float* params, size_t count);
int psw_unet_get_params(psw_unet* n, float* params, size_t count);
int psw_unet_get_grads(psw_unet* n, float* grads, size_t count);
/* Slot input from a sweep: d_ref (0..255), d_argmin / d_cmin / d_csec of
 * psw_sweep_fused_aggregate, d_depths[nplanes] the plane depths; pixels with
 * no valid plane get depth `fallback`.  Also keeps the sweep depth of the slot. */
int psw_unet_features(psw_unet* n, int slot, int width, int height, const float* d_ref, const int32_t* d_argmin,
                      const float* d_cmin, const float* d_csec, const float* d_depths, int nplanes, float fallback,
                      psw_stream_t s);
/* Raw slot input (height, width, 8) FP32 and sweep depth (tests). */
int psw_unet_set_input(psw_unet* n, int slot, int width, int height, const float* d_x, const float* d_depth,
                       psw_stream_t s);
/* Training target of a slot from a converged TGV2 state: 8 planes of
 * width*height (u, u1x, u1y, px, py, qx, qy, qz, as psw_tgv_get_state's first
 * eight), scaled against the slot's sweep depth. */
int psw_unet_targets(psw_unet* n, int slot, int width, int height, const float* d_state, psw_stream_t s);
int psw_unet_set_target(psw_unet* n, int slot, int width, int height, const float* d_target, psw_stream_t s);
/* Forward pass over slots 0..batch-1. */
int psw_unet_forward(psw_unet* n, int batch, int width, int height, psw_stream_t s);
/* Output of a slot, (height, width, 8) FP32. */
int psw_unet_output(psw_unet* n, int slot, int width, int height, float* d_out, psw_stream_t s);
/* Warm-start fields of a slot (any may be NULL): d_u (W*H), d_u1 (2*W*H),
 * d_p (2*W*H), d_q (3*W*H), in psw_tgv_set_primal / set_duals layout. */
int psw_unet_warm_start(psw_unet* n, int slot, int width, int height, float* d_u, float* d_u1, float* d_p,
                        float* d_q, psw_stream_t s);
/* Forward, smooth-L1 loss (channel_weights[8]) against the slots' targets,
 * backward into the gradient buffer.  loss (nullable, synchronous when set):
 * mean over pixels of the weighted per-channel losses. */
int psw_unet_loss_grad(psw_unet* n, int batch, int width, int height, const float* channel_weights, double* loss,
                       psw_stream_t s);
/* One Adam update (bias-corrected, step >= 1) from the gradient buffer. */
int psw_unet_adam_step(psw_unet* n, double lr, double beta1, double beta2, double eps, int step, psw_stream_t s);
/* Forward FLOPs (2 per multiply-add) of one pass. */
int psw_unet_flops(int width, int height, int batch, double* fwd_flops);
/* Single layers in NHWC bf16 (uint16 bit patterns), for tests and tools;
 * these allocate temporaries and return after the work is done.
 * Forward: y = conv3x3([xa | nearest-up(xb)], w[9][ca+cb][cout]) + bias, as
 * bf16 ReLU (relu_bf16 = 1) or FP32 (0).  Data gradient: dx[.., cin] FP32 of
 * dz[.., cout].  Weight gradient: dwb[(9*(ca+cb)+1)*cout] FP32, bias last. */
int psw_conv3x3_bf16(const uint16_t* d_xa, int ca, const uint16_t* d_xb, int cb, const float* d_w,
                     const float* d_bias, int cout, int batch, int width, int height, int relu_bf16, void* d_y,
                     psw_stream_t s);
int psw_conv3x3_bf16_dgrad(const uint16_t* d_dz, int cout, const float* d_w, int cin, int batch, int width,
                           int height, float* d_dx, psw_stream_t s);
int psw_conv3x3_bf16_wgrad(const uint16_t* d_xa, int ca, const uint16_t* d_xb, int cb, const uint16_t* d_dz,
                           int cout, int batch, int width, int height, float* d_dwb, psw_stream_t s);

#ifdef __cplusplus
}
#endif

#endif /* PSW_H */
