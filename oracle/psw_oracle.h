/*
 * psw_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a CPU restatement of the reference's
 * multiview-stereo path (dkavolis/plane-sweep-cuda: src/kernels.cu,
 * src/TGV2_kernels.cu, src/fusion.cu, src/inc/fusion.h, src/inc/structs.h,
 * src/planesweep.cpp, src/pclviewer.cpp).  It is used only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, to compare the
 * GPU results against.
 * The product (libpsw.so) never links, loads or calls it.
 *
 * Every function exists in two instantiations:
 *   _f32 : the reference's own arithmetic -- float storage and float
 *          operations, with double temporaries exactly where the reference
 *          declares them (TGV2 P/Q updates, Iu, fusion U step);
 *   _f64 : the same algorithm with every operation in double (accuracy judge).
 * Inputs are float (the path's inputs are float images and float poses);
 * outputs of the _f64 functions are double.
 *
 * Matrices are row-major 3x3 (9 values), vectors 3 values.
 * Images are dense row-major W*H; stacks of N images are N*H*W.
 * Voxel grids are x-fastest, then y, then z; p is stored as 3 planes [3][V];
 * histograms as [V][8] uint16 counts.
 */
#ifndef PSW_ORACLE_H
#define PSW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- helpers ---------------------------------------------------------- */

/* Plane depths of the reference's float-accumulated loop
 * (src/planesweep.cpp:169,200).  Returns the number of planes written. */
int orc_reference_planes(float znear, float zfar, unsigned nplanes, float* out, int maxn);

/* Matrix3D::inv (src/inc/structs.h:584-593) in float. */
void orc_mat3_inv_f32(const float* m, float* out);

/* PlaneSweep::RelativeMatrices (src/planesweep.cpp:538-549) in float. */
void orc_relative_matrices_f32(const float* Rref, const float* tref, const float* Rsrc,
                               const float* tsrc, int alternative, float* Rrel, float* trel);

/* Per-plane homography of PlaneSweepThread (src/planesweep.cpp:195-203),
 * float Matrix3D arithmetic, normalised by H(2,2). */
void orc_homography_ref_f32(const float* K, const float* Rref, const float* tref, const float* Rsrc,
                            const float* tsrc, int alternative, float depth, float* H);

/* Same homography built in double from the float inputs (used by the
 * aggregated sweep), returned rounded to float and in double. */
void orc_homography_f64(const float* K, const float* Rref, const float* tref, const float* Rsrc,
                        const float* tsrc, int alternative, float depth, double* H);

/* ---- single reference kernels (unit-level known-answer tests) --------- */

void orc_windowed_mean_row_f32(float* out, const float* in, int winsize, int squared, int w, int h);
void orc_windowed_mean_col_f32(float* out, const float* in, int winsize, int squared, int w, int h);
void orc_bilinear_f32(float* out, const float* data, const float* xs, const float* ys,
                      int M1, int M2, int N1, int N2);
void orc_transform_indexes_f32(float* xs, float* ys, const float* H, int w, int h);

/* ---- stage 1: plane sweep --------------------------------------------- */

/* PlaneSweep::RunAlgorithm + PlaneSweepThread (src/planesweep.cpp:74-259)
 * with the reference's per-source argmax and thresholded depth average.
 * planes: the plane depths (orc_reference_planes gives the reference list).
 * best_ncc / best_depth (nullable): per-source N*H*W outputs. */
int orc_sweep_ref_f32(const float* ref, const float* src, int W, int H, int N,
                      const float* K, const float* Rref, const float* tref,
                      const float* Rsrc, const float* tsrc, int alternative,
                      const float* planes, int P, int winsize, float stdthresh,
                      float nccthresh, float zfar,
                      float* depth_out, float* best_ncc, float* best_depth);
int orc_sweep_ref_f64(const float* ref, const float* src, int W, int H, int N,
                      const float* K, const float* Rref, const float* tref,
                      const float* Rsrc, const float* tsrc, int alternative,
                      const float* planes, int P, int winsize, float stdthresh,
                      float nccthresh, float zfar,
                      double* depth_out, double* best_ncc, double* best_depth);

/* Aggregated sweep: cost(p) = -(sum over sources of NCC) / N with the
 * reference's NCC definition; argmin (first minimum), minimum and
 * second-smallest cost.  Homographies are built in double (0-based or
 * 1-based pixel convention via coord_offset) and, in the _f32 variant,
 * rounded to float before warping. */
int orc_sweep_agg_f32(const float* ref, const float* src, int W, int H, int N,
                      const float* K, const float* Rref, const float* tref,
                      const float* Rsrc, const float* tsrc, int alternative,
                      const float* planes, int P, int winsize, float stdthresh, int coord_offset,
                      int32_t* argmin, float* cost_min, float* cost_second);
int orc_sweep_agg_f64(const float* ref, const float* src, int W, int H, int N,
                      const float* K, const float* Rref, const float* tref,
                      const float* Rsrc, const float* tsrc, int alternative,
                      const float* planes, int P, int winsize, float stdthresh, int coord_offset,
                      int32_t* argmin, double* cost_min, double* cost_second);

/* Aggregated sweep in double that also writes the cost of every plane:
 * volume[P][H][W] = -(sum over sources of NCC) / N. */
int orc_sweep_agg_volume_f64(const float* ref, const float* src, int W, int H, int N,
                             const float* K, const float* Rref, const float* tref,
                             const float* Rsrc, const float* tsrc, int alternative,
                             const float* planes, int P, int winsize, float stdthresh, int coord_offset,
                             int32_t* argmin, double* cost_min, double* cost_second, double* volume);

/* The same in the reference's float arithmetic (float volume). */
int orc_sweep_agg_volume_f32(const float* ref, const float* src, int W, int H, int N,
                             const float* K, const float* Rref, const float* tref,
                             const float* Rsrc, const float* tsrc, int alternative,
                             const float* planes, int P, int winsize, float stdthresh, int coord_offset,
                             int32_t* argmin, float* cost_min, float* cost_second, float* volume);

/* The aggregated sweep's per-pixel reduction over a cost sequence in plane
 * order: argmin (first minimum), minimum and second-smallest cost. */
void orc_reduce_best2_f64(const double* costs, int P, int32_t* argmin, double* cost_min, double* cost_second);

/* ---- stage 2: TGV2 ----------------------------------------------------- */

/* PlaneSweep::TGV (src/planesweep.cpp:381-536).  ref/src hold 0..255
 * intensities; K uses the reference's 1-based pixel convention.
 * u_init (nullable) replaces the constant cold start u = 1.
 * state_out (nullable) receives the final state as planes of H*W:
 *   u, u1x, u1y, px, py, qx, qy, qz, then It[N], Iu[N], r[N]. */
int orc_tgv_f32(const float* ref, const float* src, int W, int H, int N,
                const float* K, const float* Rref, const float* tref,
                const float* Rsrc, const float* tsrc, int alternative,
                int niters, int warps, double lambda, double alpha0, double alpha1,
                double tau, double sigma, double beta, double gamma,
                const float* u_init, float* u_out, float* state_out);
int orc_tgv_f64(const float* ref, const float* src, int W, int H, int N,
                const float* K, const float* Rref, const float* tref,
                const float* Rsrc, const float* tsrc, int alternative,
                int niters, int warps, double lambda, double alpha0, double alpha1,
                double tau, double sigma, double beta, double gamma,
                const float* u_init, double* u_out, double* state_out);

/* Anisotropic_diffusion_tensor_kernel (src/TGV2_kernels.cu:257-298) on an
 * already normalised image. */
void orc_tensor_f32(const float* img, int W, int H, float beta, float gamma,
                    float* T11, float* T12, float* T21, float* T22);

/* Linearisation of one warp (src/planesweep.cpp:460-483): It and Iu of every
 * source at u0, and the tensor. Outputs It[N*H*W], Iu[N*H*W], T[4*H*W]. */
int orc_tgv_linearise_f32(const float* ref, const float* src, int W, int H, int N,
                          const float* K, const float* Rref, const float* tref,
                          const float* Rsrc, const float* tsrc, int alternative,
                          double beta, double gamma, const float* u0,
                          float* It, float* Iu, float* T);

int orc_tgv_linearise_f64(const float* ref, const float* src, int W, int H, int N,
                          const float* K, const float* Rref, const float* tref,
                          const float* Rsrc, const float* tsrc, int alternative,
                          double beta, double gamma, const float* u0,
                          double* It, double* Iu, double* T);

/* niters primal-dual iterations (src/planesweep.cpp:485-509) from a given
 * state inside a given linearisation (no warp reset).  state (in and out):
 * planes of H*W u, u1x, u1y, ubar, u1xbar, u1ybar, px, py, qx, qy, qz, qw,
 * r[N].  T = T11, T12, T21, T22 planes; It, Iu: N planes each. */
int orc_tgv_iterate_f32(int W, int H, int N, const float* T, const float* u0, const float* It,
                        const float* Iu, float* state, int niters, double lambda, double alpha0,
                        double alpha1, double tau, double sigma);
int orc_tgv_iterate_f64(int W, int H, int N, const double* T, const double* u0, const double* It,
                        const double* Iu, double* state, int niters, double lambda, double alpha0,
                        double alpha1, double tau, double sigma);

/* Primal TGV2 energy of a state, in double:
 *   alpha1 * sum |T (grad u - u1)| + alpha0 * sum |eps(u1)|_F
 *   + lambda * sum_j sum |It_j + (u - u0) Iu_j|
 * with the clamped forward differences the iteration uses. T = [T11,T12,T21,T22]. */
double orc_tgv_energy(const float* u, const float* u1x, const float* u1y, const float* u0,
                      const float* T, const float* It, const float* Iu, int W, int H, int N,
                      double lambda, double alpha0, double alpha1);
double orc_tgv_energy_f64(const double* u, const double* u1x, const double* u1y, const double* u0,
                          const double* T, const double* It, const double* Iu, int W, int H, int N,
                          double lambda, double alpha0, double alpha1);
/* The same sum over the pixels whose weight is nonzero (weight: H*W, as the
 * solver's energy mask; NULL counts every pixel). */
double orc_tgv_energy_masked_f64(const double* u, const double* u1x, const double* u1y, const double* u0,
                                 const double* T, const double* It, const double* Iu, int W, int H, int N,
                                 double lambda, double alpha0, double alpha1, const float* weight);

/* ---- stage 3: histogram fusion ---------------------------------------- */

/* fusionData::binParams (src/inc/fusion.h:1041-1054). */
void orc_fusion_bin_centers(int bins, float* out);

/* FusionUpdateHistogram_kernel (src/fusion.cu:5-43) for one depthmap,
 * 8 bins, uint16 counts [V][8]. vol = {ax, ay, az, bx, by, bz}. */
void orc_fusion_hist_f32(uint16_t* hist, int gw, int gh, int gd, const float* vol,
                         const float* depth, int W, int H, const float* K, const float* R,
                         const float* T, float threshold);
void orc_fusion_hist_f64(uint16_t* hist, int gw, int gh, int gd, const float* vol,
                         const float* depth, int W, int H, const float* K, const float* R,
                         const float* T, float threshold);

/* FusionUpdateU_kernel (src/fusion.cu:45-57), 8 bins. */
void orc_fusion_u_f32(float* u, float* v, const float* p, const uint16_t* hist,
                      int gw, int gh, int gd, double tau, double lambda);
void orc_fusion_u_f64(double* u, double* v, const double* p, const uint16_t* hist,
                      int gw, int gh, int gd, double tau, double lambda);

/* FusionUpdateP_kernel (src/fusion.cu:59-70). */
void orc_fusion_p_f32(float* p, const float* v, int gw, int gh, int gd, double sigma);
void orc_fusion_p_f64(double* p, const double* v, int gw, int gh, int gd, double sigma);

/* TV-L1 histogram energy sum |grad u| + lambda sum_k h_k |u - c_k| in FP64
 * (forward differences as FusionUpdateP_kernel; no reference counterpart). */
double orc_fusion_energy_f64(const double* u, const uint16_t* hist, int gw, int gh, int gd, double lambda);

/* proxHist through the reference's sortedHist insertion
 * (src/inc/fusion.h:851-858, src/inc/structs.h:212-284). W[8] = W_1..W_8. */
float orc_prox_hist_insertion_f32(float u, const int* Wi, float tau, float lambda);

/* Number of voxels with -1 < u < 0 (src/pclviewer.cpp:1058-1074). */
int64_t orc_surface_count_f32(const float* u, int64_t n);
int64_t orc_surface_count_f64(const double* u, int64_t n);

/* ---- TVL1 depth denoising (src/planesweep.cpp:297-379) ------------------ */

/* PlaneSweep::CudaDenoise: depth (float), reference image 0..255 for the
 * tensor, parameters as the reference.  Output denoised depth. */
int orc_tvl1_f32(const float* depth, const float* ref, int W, int H, float znear, float zfar,
                 int niters, double lambda, double tau, double sigma, double theta,
                 double beta, double gamma, float* out);
int orc_tvl1_f64(const float* depth, const float* ref, int W, int H, float znear, float zfar,
                 int niters, double lambda, double tau, double sigma, double theta,
                 double beta, double gamma, double* out);

/* ---- sparse-depth TGV2 and 3-D coordinates ----------------------------- */

/* PlaneSweep::TGVdenoiseFromSparse (src/planesweep.cpp:596-674): reference
 * image 0..255, dense depth (the initial u), sparse depth (> 0 where
 * known); output depth; u1_out (2 x W*H, may be NULL) gets u1x, u1y. */
int orc_tgv_sparse_f32(const float* ref, const float* dense, const float* sparse, int W, int H, float zfar,
                       int niters, double alpha0, double alpha1, double tau, double sigma, double theta,
                       double beta, double gamma, float* out, float* u1_out);
int orc_tgv_sparse_f64(const float* ref, const float* dense, const float* sparse, int W, int H, float zfar,
                       int niters, double alpha0, double alpha1, double tau, double sigma, double theta,
                       double beta, double gamma, double* out, double* u1_out);

/* PlaneSweep::get3Dcoordinates + compute3D_kernel (src/planesweep.cpp:
 * 551-586, src/kernels.cu:405-429): world coordinates of a depthmap. */
void orc_compute3d_f32(const float* depth, int W, int H, const float* K, const float* Rref, const float* tref,
                       int alt, float* X, float* Y, float* Z);
void orc_compute3d_f64(const float* depth, int W, int H, const float* K, const float* Rref, const float* tref,
                       int alt, double* X, double* Y, double* Z);

#ifdef __cplusplus
}
#endif

#endif /* PSW_ORACLE_H */
