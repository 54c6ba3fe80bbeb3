// kernels.h -- the reference's kernel launchers (src/inc/kernels.cu.h:37-729)
// with their C++ signatures (Matrix3D / Vector3D by value, dim3 launch shape),
// forwarding to libpsw's C ABI on the null stream.  The reference's
// launchers return void and report errors only through cudaPeekAtLastError;
// these return the libpsw status (0 = ok), which callers may ignore.
#pragma once

#include "../../include/psw.h"
#include "structs.h"

namespace psw {

inline psw_dim3 to_dim(dim3 d) { return psw_dim3{d.x, d.y, d.z}; }
inline void arr(const Matrix3D& m, float o[9]) { m.toArray(o); }

#define PSW_M(name) float name##_[9]; arr(name, name##_)
#define PSW_V(name) const float name##_[3] = {name.x, name.y, name.z}
#define PSW_D to_dim(blocks), to_dim(threads), nullptr

inline int transform_indexes(float* d_x, float* d_y, const Matrix3D h, const int width, const int height, dim3 blocks,
                             dim3 threads) {
    PSW_M(h);
    return psw_transform_indexes(d_x, d_y, h_, width, height, PSW_D);
}
inline int bilinear_interpolation(float* d_result, const float* d_data, const float* d_xout, const float* d_yout,
                                  const int M1, const int M2, const int N1, const int N2, dim3 blocks, dim3 threads) {
    return psw_bilinear_interpolation(d_result, d_data, d_xout, d_yout, M1, M2, N1, N2, PSW_D);
}
inline int calcNCC(float* d_ncc, const float* d_prod_mean, const float* d_mean1, const float* d_mean2,
                   const float* d_std1, const float* d_std2, const float stdthresh1, const float stdthresh2,
                   const int width, const int height, dim3 blocks, dim3 threads) {
    return psw_calcNCC(d_ncc, d_prod_mean, d_mean1, d_mean2, d_std1, d_std2, stdthresh1, stdthresh2, width, height,
                       PSW_D);
}
inline int update_arrays(float* d_depthmap, float* d_bestncc, const float* d_currentncc, const float current_depth,
                         const int width, const int height, dim3 blocks, dim3 threads) {
    return psw_update_arrays(d_depthmap, d_bestncc, d_currentncc, current_depth, width, height, PSW_D);
}
inline int sum_depthmap_NCC(float* d_depthmap_out, float* d_count, const float* d_depthmap, const float* d_ncc,
                            const float nccthreshold, const int width, const int height, dim3 blocks, dim3 threads) {
    return psw_sum_depthmap_NCC(d_depthmap_out, d_count, d_depthmap, d_ncc, nccthreshold, width, height, PSW_D);
}
inline int calculate_STD(float* d_std, const float* d_mean, const float* d_mean_of_squares, const int width,
                         const int height, dim3 blocks, dim3 threads) {
    return psw_calculate_STD(d_std, d_mean, d_mean_of_squares, width, height, PSW_D);
}
inline int set_value(float* d_output, const float value, const int width, const int height, dim3 blocks,
                     dim3 threads) {
    return psw_set_value(d_output, value, width, height, PSW_D);
}
inline int element_multiply(float* d_output, const float* d_input1, const float* d_input2, const int width,
                            const int height, dim3 blocks, dim3 threads) {
    return psw_element_multiply(d_output, d_input1, d_input2, width, height, PSW_D);
}
inline int element_rdivide(float* d_output, const float* d_input1, const float* d_input2, const int width,
                           const int height, dim3 blocks, dim3 threads) {
    return psw_element_rdivide(d_output, d_input1, d_input2, width, height, PSW_D);
}
inline int windowed_mean_row(float* d_output, const float* d_input, const unsigned int winsize, const bool squared,
                             const int width, const int height, dim3 blocks, dim3 threads) {
    return psw_windowed_mean_row(d_output, d_input, winsize, squared ? 1 : 0, width, height, PSW_D);
}
inline int windowed_mean_column(float* d_output, const float* d_input, const unsigned int winsize, const bool squared,
                                const int width, const int height, dim3 blocks, dim3 threads) {
    return psw_windowed_mean_column(d_output, d_input, winsize, squared ? 1 : 0, width, height, PSW_D);
}
inline int element_scale(float* d_output, const float scale, const int width, const int height, dim3 blocks,
                         dim3 threads) {
    return psw_element_scale(d_output, scale, width, height, PSW_D);
}
inline int element_add(float* d_output, const float value, const int width, const int height, dim3 blocks,
                       dim3 threads) {
    return psw_element_add(d_output, value, width, height, PSW_D);
}
inline int set_QNAN_value(float* d_output, const float value, const int width, const int height, dim3 blocks,
                          dim3 threads) {
    return psw_set_QNAN_value(d_output, value, width, height, PSW_D);
}
inline int denoising_TVL1_calculateP(float* d_Px, float* d_Py, const float* d_input, const float sigma, const int width,
                                     const int height, dim3 blocks, dim3 threads) {
    return psw_denoising_TVL1_calculateP(d_Px, d_Py, d_input, sigma, width, height, PSW_D);
}
inline int denoising_TVL1_calculateP_tensor_weighed(float* d_Px, float* d_Py, const float* d_T11, const float* d_T12,
                                                    const float* d_T21, const float* d_T22, const float* d_input,
                                                    const float sigma, const int width, const int height, dim3 blocks,
                                                    dim3 threads) {
    return psw_denoising_TVL1_calculateP_tensor_weighed(d_Px, d_Py, d_T11, d_T12, d_T21, d_T22, d_input, sigma, width,
                                                        height, PSW_D);
}
inline int denoising_TVL1_update(float* d_output, float* d_R, const float* d_Px, const float* d_Py,
                                 const float* d_origin, const float tau, const float theta, const float lambda,
                                 const float sigma, const int width, const int height, dim3 blocks, dim3 threads) {
    return psw_denoising_TVL1_update(d_output, d_R, d_Px, d_Py, d_origin, tau, theta, lambda, sigma, width, height,
                                     PSW_D);
}
inline int denoising_TVL1_update_tensor_weighed(float* d_output, float* d_R, const float* d_Px, const float* d_Py,
                                                const float* d_origin, const float* d_T11, const float* d_T12,
                                                const float* d_T21, const float* d_T22, const float tau,
                                                const float theta, const float lambda, const float sigma,
                                                const int width, const int height, dim3 blocks, dim3 threads) {
    return psw_denoising_TVL1_update_tensor_weighed(d_output, d_R, d_Px, d_Py, d_origin, d_T11, d_T12, d_T21, d_T22,
                                                    tau, theta, lambda, sigma, width, height, PSW_D);
}
inline int compute3D(float* d_x, float* d_y, float* d_z, const Matrix3D Rrel, const Vector3D trel,
                     const Matrix3D invK, const int width, const int height, dim3 blocks, dim3 threads) {
    PSW_M(Rrel);
    PSW_V(trel);
    PSW_M(invK);
    return psw_compute3D(d_x, d_y, d_z, Rrel_, trel_, invK_, width, height, PSW_D);
}

// ---- TGV2 ----
inline int TGV2_updateP(float* d_Px, float* d_Py, const float* d_u, const float* d_u1x, const float* d_u1y,
                        const float alpha1, const float sigma, const int width, const int height, dim3 blocks,
                        dim3 threads) {
    return psw_TGV2_updateP(d_Px, d_Py, d_u, d_u1x, d_u1y, alpha1, sigma, width, height, PSW_D);
}
inline int TGV2_updateQ(float* d_Qx, float* d_Qy, float* d_Qz, float* d_Qw, const float* d_u1x, const float* d_u1y,
                        const float alpha0, const float sigma, const int width, const int height, dim3 blocks,
                        dim3 threads) {
    return psw_TGV2_updateQ(d_Qx, d_Qy, d_Qz, d_Qw, d_u1x, d_u1y, alpha0, sigma, width, height, PSW_D);
}
inline int TGV2_updateR(float* d_r, float* d_prodsum, const float* d_u, const float* d_u0, const float* d_It,
                        const float* d_Iu, const float sigma, const float lambda, const int width, const int height,
                        dim3 blocks, dim3 threads) {
    return psw_TGV2_updateR(d_r, d_prodsum, d_u, d_u0, d_It, d_Iu, sigma, lambda, width, height, PSW_D);
}
inline int TGV2_updateU(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar, float* d_u1ybar,
                        const float* d_Px, const float* d_Py, const float* d_Qx, const float* d_Qy, const float* d_Qz,
                        const float* d_Qw, const float* d_prodsum, const float alpha0, const float alpha1,
                        const float tau, const float lambda, const int width, const int height, dim3 blocks,
                        dim3 threads) {
    return psw_TGV2_updateU(d_u, d_u1x, d_u1y, d_ubar, d_u1xbar, d_u1ybar, d_Px, d_Py, d_Qx, d_Qy, d_Qz, d_Qw,
                            d_prodsum, alpha0, alpha1, tau, lambda, width, height, PSW_D);
}
inline int TGV2_transform_coordinates(float* d_x, float* d_y, float* d_X, float* d_Y, float* d_Z, const float* d_u,
                                      const Matrix3D K, const Matrix3D Rrel, const Vector3D trel, const Matrix3D invK,
                                      const int width, const int height, dim3 blocks, dim3 threads) {
    PSW_M(K);
    PSW_M(Rrel);
    PSW_V(trel);
    PSW_M(invK);
    return psw_TGV2_transform_coordinates(d_x, d_y, d_X, d_Y, d_Z, d_u, K_, Rrel_, trel_, invK_, width, height, PSW_D);
}
inline int subtract(float* d_out, const float* d_in1, const float* d_in2, const int width, const int height,
                    dim3 blocks, dim3 threads) {
    return psw_subtract(d_out, d_in1, d_in2, width, height, PSW_D);
}
inline int TGV2_calculate_coordinate_derivatives(float* d_dX, float* d_dY, float* d_dZ, const Matrix3D invK,
                                                 const Matrix3D Rrel, const int width, const int height, dim3 blocks,
                                                 dim3 threads) {
    PSW_M(invK);
    PSW_M(Rrel);
    return psw_TGV2_calculate_coordinate_derivatives(d_dX, d_dY, d_dZ, invK_, Rrel_, width, height, PSW_D);
}
inline int TGV2_calculate_derivativeF(float* d_dfx, float* d_dfy, const float* d_X, const float* d_dX,
                                      const float* d_Y, const float* d_dY, const float* d_Z, const float* d_dZ,
                                      const float fx, const float fy, const int width, const int height, dim3 blocks,
                                      dim3 threads) {
    return psw_TGV2_calculate_derivativeF(d_dfx, d_dfy, d_X, d_dX, d_Y, d_dY, d_Z, d_dZ, fx, fy, width, height, PSW_D);
}
inline int TGV2_calculate_Iu(float* d_Iu, const float* d_I, const float* d_dfx, const float* d_dfy, const int width,
                             const int height, dim3 blocks, dim3 threads) {
    return psw_TGV2_calculate_Iu(d_Iu, d_I, d_dfx, d_dfy, width, height, PSW_D);
}
inline int Anisotropic_diffusion_tensor(float* d_T11, float* d_T12, float* d_T21, float* d_T22, const float* d_Img,
                                        const float beta, const float gamma, const int width, const int height,
                                        dim3 blocks, dim3 threads) {
    return psw_Anisotropic_diffusion_tensor(d_T11, d_T12, d_T21, d_T22, d_Img, beta, gamma, width, height, PSW_D);
}
inline int TGV2_updateU_tensor_weighed(float* d_u, float* d_u1x, float* d_u1y, const float* d_T11, const float* d_T12,
                                       const float* d_T21, const float* d_T22, float* d_ubar, float* d_u1xbar,
                                       float* d_u1ybar, const float* d_Px, const float* d_Py, const float* d_Qx,
                                       const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_prodsum,
                                       const float alpha0, const float alpha1, const float tau, const float lambda,
                                       const int width, const int height, dim3 blocks, dim3 threads) {
    return psw_TGV2_updateU_tensor_weighed(d_u, d_u1x, d_u1y, d_T11, d_T12, d_T21, d_T22, d_ubar, d_u1xbar, d_u1ybar,
                                           d_Px, d_Py, d_Qx, d_Qy, d_Qz, d_Qw, d_prodsum, alpha0, alpha1, tau, lambda,
                                           width, height, PSW_D);
}
inline int TGV2_updateP_tensor_weighed(float* d_Px, float* d_Py, const float* d_T11, const float* d_T12,
                                       const float* d_T21, const float* d_T22, const float* d_u, const float* d_u1x,
                                       const float* d_u1y, const float alpha1, const float sigma, const int width,
                                       const int height, dim3 blocks, dim3 threads) {
    return psw_TGV2_updateP_tensor_weighed(d_Px, d_Py, d_T11, d_T12, d_T21, d_T22, d_u, d_u1x, d_u1y, alpha1, sigma,
                                           width, height, PSW_D);
}
inline int TGV2_updateU_sparseDepth(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar,
                                    float* d_u1ybar, const float* d_Px, const float* d_Py, const float* d_Qx,
                                    const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_w,
                                    const float* d_Ds, const float alpha0, const float alpha1, const float tau,
                                    const float theta, const int width, const int height, dim3 blocks, dim3 threads) {
    return psw_TGV2_updateU_sparseDepth(d_u, d_u1x, d_u1y, d_ubar, d_u1xbar, d_u1ybar, d_Px, d_Py, d_Qx, d_Qy, d_Qz,
                                        d_Qw, d_w, d_Ds, alpha0, alpha1, tau, theta, width, height, PSW_D);
}
inline int TGV2_updateU_sparseDepthTensor(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar,
                                          float* d_u1ybar, const float* d_T11, const float* d_T12, const float* d_T21,
                                          const float* d_T22, const float* d_Px, const float* d_Py, const float* d_Qx,
                                          const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_w,
                                          const float* d_Ds, const float alpha0, const float alpha1, const float tau,
                                          const float theta, const int width, const int height, dim3 blocks,
                                          dim3 threads) {
    return psw_TGV2_updateU_sparseDepthTensor(d_u, d_u1x, d_u1y, d_ubar, d_u1xbar, d_u1ybar, d_T11, d_T12, d_T21,
                                              d_T22, d_Px, d_Py, d_Qx, d_Qy, d_Qz, d_Qw, d_w, d_Ds, alpha0, alpha1,
                                              tau, theta, width, height, PSW_D);
}
inline int calculateWeights_sparseDepth(float* d_w, const float* d_Ds, const int width, const int height, dim3 blocks,
                                        dim3 threads) {
    return psw_calculateWeights_sparseDepth(d_w, d_Ds, width, height, PSW_D);
}

#undef PSW_M
#undef PSW_V
#undef PSW_D

}  // namespace psw
