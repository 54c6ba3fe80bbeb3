// ref_kernels.hip -- the reference's per-step kernels (src/inc/kernels.cu.h),
// one HIP kernel each, behind C-ABI launchers with the reference's argument
// order plus a stream and a status.
//
// They reproduce the reference's arithmetic operation by operation: float
// where it computes in float, double where it declares double temporaries,
// no FMA contraction.  Driving them in the reference's own host sequence
// therefore yields the reference's results; the fused kernels of
// sweep_fused.hip / tgv_fused.hip / fusion.hip / tvl1_fused.hip are the fast
// path.
//
// Pointer parameters carry no __restrict__: the reference's host code passes
// aliased arguments in places (calculate_STD in place, planesweep.cpp:118;
// element_rdivide in place, :130), and those must stay well defined.
#include "psw_common.h"

namespace psw {

#define PIX2D                                              \
    const int ind_x = threadIdx.x + blockDim.x * blockIdx.x; \
    const int ind_y = threadIdx.y + blockDim.y * blockIdx.y; \
    if (!((ind_x < width) && (ind_y < height))) return;      \
    const int ind = ind_y * width + ind_x;                   \
    (void)ind;

// kernels.cu:31-44
__global__ void transform_indexes_kernel(float* d_x, float* d_y, M3f h, int width, int height) {
    PIX2D
    float ox, oy, oz;
    mv3(h, (float)(ind_x + 1), (float)(ind_y + 1), 1.f, ox, oy, oz);
    d_x[ind] = ox / oz - 1.f;
    d_y[ind] = oy / oz - 1.f;
}

// kernels.cu:4-29 (N1 x N2 output, M1 x M2 data)
__global__ void bilinear_kernel(float* d_result, const float* d_data, const float* d_xout, const float* d_yout,
                                int M1, int M2, int N1, int N2) {
    const int l = threadIdx.x + blockDim.x * blockIdx.x;
    const int k = threadIdx.y + blockDim.y * blockIdx.y;
    if ((l < N1) && (k < N2)) {
        const size_t i = (size_t)k * N1 + l;
        d_result[i] = bilinear_sample(d_data, M1, M2, d_xout[i], d_yout[i]);
    }
}

// kernels.cu:46-64
__global__ void calcNCC_kernel(float* d_ncc, const float* d_prod_mean, const float* d_mean1, const float* d_mean2,
                               const float* d_std1, const float* d_std2, float th1, float th2, int width,
                               int height) {
    PIX2D
    if ((d_std1[ind] < th1) || (d_std2[ind] < th2)) d_ncc[ind] = 0.f;
    else d_ncc[ind] = (d_prod_mean[ind] - d_mean1[ind] * d_mean2[ind]) / (d_std1[ind] * d_std2[ind]);
}

// kernels.cu:66-82
__global__ void update_arrays_kernel(float* d_depthmap, float* d_bestncc, const float* d_currentncc, float depth,
                                     int width, int height) {
    PIX2D
    if (d_currentncc[ind] > d_bestncc[ind]) {
        d_bestncc[ind] = d_currentncc[ind];
        d_depthmap[ind] = depth;
    }
}

// kernels.cu:84-101
__global__ void sum_depthmap_NCC_kernel(float* d_out, float* d_count, const float* d_depthmap, const float* d_ncc,
                                        float thr, int width, int height) {
    PIX2D
    if (d_ncc[ind] > thr) {
        d_out[ind] += d_depthmap[ind];
        d_count[ind] += 1.f;
    }
}

// kernels.cu:103-120
__global__ void calculate_STD_kernel(float* d_std, const float* d_mean, const float* d_msq, int width, int height) {
    PIX2D
    const float var = d_msq[ind] - d_mean[ind] * d_mean[ind];
    d_std[ind] = var > 0.f ? sqrtf(var) : 0.f;
}

// kernels.cu:122-131
__global__ void set_value_kernel(float* d_out, float value, int width, int height) {
    PIX2D
    d_out[ind] = value;
}

// kernels.cu:133-143
__global__ void element_multiply_kernel(float* d_out, const float* a, const float* b, int width, int height) {
    PIX2D
    d_out[ind] = a[ind] * b[ind];
}

// kernels.cu:145-157
__global__ void element_rdivide_kernel(float* d_out, const float* a, const float* b, int width, int height) {
    PIX2D
    const float den = b[ind];
    d_out[ind] = (den != 0.f) ? a[ind] / den : __builtin_nanf("");
}

// kernels.cu:187-210
__global__ void windowed_mean_row_kernel(float* d_out, const float* d_in, unsigned winsize, bool squared, int width,
                                         int height) {
    PIX2D
    float mean = 0.f;
    const int n = winsize / 2;
    const float* row = d_in + (size_t)ind_y * width;
    for (int i = -n; i <= n; i++) {
        const float v = row[reflect101(ind_x + i, width)];
        if (squared) mean += v * v;
        else mean += v;
    }
    d_out[ind] = mean / (float)winsize;
}

// kernels.cu:212-235
__global__ void windowed_mean_column_kernel(float* d_out, const float* d_in, unsigned winsize, bool squared,
                                            int width, int height) {
    PIX2D
    float mean = 0.f;
    const int n = winsize / 2;
    for (int i = -n; i <= n; i++) {
        const float v = d_in[(size_t)reflect101(ind_y + i, height) * width + ind_x];
        if (squared) mean += v * v;
        else mean += v;
    }
    d_out[ind] = mean / (float)winsize;
}

// kernels.cu:309-318
__global__ void element_scale_kernel(float* d_out, float scale, int width, int height) {
    PIX2D
    d_out[ind] = d_out[ind] * scale;
}

// kernels.cu:320-329
__global__ void element_add_kernel(float* d_out, float value, int width, int height) {
    PIX2D
    d_out[ind] += value;
}

// kernels.cu:331-340
__global__ void set_QNAN_value_kernel(float* d_out, float value, int width, int height) {
    PIX2D
    if (d_out[ind] != d_out[ind]) d_out[ind] = value;
}

// kernels.cu:250-281
__global__ void tvl1_calculateP_kernel(float* d_Px, float* d_Py, const float* d_input, float sigma, int width,
                                       int height) {
    PIX2D
    const int ny = min(ind_y + 1, height - 1);
    double dx, dy, m;
    if (ind_x == width - 1) {
        dy = (double)((d_input[ny * width + ind_x] - d_input[ind]) * sigma + d_Py[ind]);
        m = 1.0 / dy;
        if (m < 0.0) m = -m;
        if (m > 1.0) m = 1.0;
        d_Px[ind] = 0.f;
        d_Py[ind] = (float)(dy * m);
    } else {
        dx = (double)((d_input[ind + 1] - d_input[ind]) * sigma + d_Px[ind]);
        dy = (double)((d_input[ny * width + ind_x] - d_input[ind]) * sigma + d_Py[ind]);
        m = 1.0 / sqrt(dx * dx + dy * dy);
        if (m > 1.0) m = 1.0;
        d_Px[ind] = (float)(dx * m);
        d_Py[ind] = (float)(dy * m);
    }
}

// kernels.cu:283-307
__global__ void tvl1_calculateP_tensor_kernel(float* d_Px, float* d_Py, const float* T11, const float* T12,
                                              const float* T21, const float* T22, const float* d_input, float sigma,
                                              int width, int height) {
    PIX2D
    const int xn = min(ind_x + 1, width - 1);
    const int yn = min(ind_y + 1, height - 1);
    const double x = (double)(d_input[ind_y * width + xn] - d_input[ind]);
    const double y = (double)(d_input[yn * width + ind_x] - d_input[ind]);
    const double dx = (double)d_Px[ind] + (double)sigma * ((double)T11[ind] * x + (double)T12[ind] * y);
    const double dy = (double)d_Py[ind] + (double)sigma * ((double)T21[ind] * x + (double)T22[ind] * y);
    const double d = (double)fmaxf(1.f, (float)sqrt(dx * dx + dy * dy));
    d_Px[ind] = (float)(dx / d);
    d_Py[ind] = (float)(dy / d);
}

// kernels.cu:342-370
__global__ void tvl1_update_kernel(float* d_out, float* d_R, const float* d_Px, const float* d_Py,
                                   const float* d_origin, float tau, float theta, float lambda, float sigma,
                                   int width, int height) {
    PIX2D
    const int yp = max(ind_y - 1, 0);
    float r = d_R[ind];
    r += d_origin[ind];
    r += sigma * d_out[ind];
    if (r > lambda) r = lambda;
    if (r < -lambda) r = -lambda;
    d_R[ind] = r;
    const float o = d_out[ind];
    double x_new;
    if (ind_x == 0) x_new = (double)(o + tau * (d_Py[ind] - d_Py[yp * width + ind_x]) - tau * r);
    else x_new = (double)(o + tau * (d_Px[ind] - d_Px[ind - 1] + d_Py[ind] - d_Py[yp * width + ind_x]) - tau * r);
    d_out[ind] = (float)(x_new + (double)theta * (x_new - (double)o));
}

// kernels.cu:372-403
__global__ void tvl1_update_tensor_kernel(float* d_out, float* d_R, const float* d_Px, const float* d_Py,
                                          const float* d_origin, const float* T11, const float* T12,
                                          const float* T21, const float* T22, float tau, float theta, float lambda,
                                          float sigma, int width, int height) {
    PIX2D
    const int xp = max(ind_x - 1, 0);
    const int yp = max(ind_y - 1, 0);
    float r = d_R[ind];
    r += d_origin[ind];
    r += sigma * d_out[ind];
    if (r > lambda) r = lambda;
    if (r < -lambda) r = -lambda;
    d_R[ind] = r;
    const float c_px = d_Px[ind], c_py = d_Py[ind], xp_px = d_Px[ind_y * width + xp],
                yp_px = d_Px[yp * width + ind_x], xp_py = d_Py[ind_y * width + xp],
                yp_py = d_Py[yp * width + ind_x];
    const float o = d_out[ind];
    const double x_new = (double)(o + tau * ((T11[ind] * (c_px - xp_px) + T12[ind] * (c_py - xp_py) +
                                              T21[ind] * (c_px - yp_px) + T22[ind] * (c_py - yp_py)) -
                                             r));
    d_out[ind] = (float)(x_new + (double)theta * (x_new - (double)o));
}

// kernels.cu:405-429
__global__ void compute3D_kernel(float* d_x, float* d_y, float* d_z, M3f Rrel, V3f trel, M3f invK, int width,
                                 int height) {
    PIX2D
    const float z = d_z[ind];
    if ((double)z < 5.5) {
        M3f zK;
        for (int j = 0; j < 9; ++j) zK.m[j] = invK.m[j] * z;
        float a0, a1, a2, b0, b1, b2;
        mv3(zK, (float)(ind_x + 1), (float)(ind_y + 1), 1.f, a0, a1, a2);
        mv3(Rrel, a0, a1, a2, b0, b1, b2);
        d_x[ind] = b0 + trel.x;
        d_y[ind] = b1 + trel.y;
        d_z[ind] = b2 + trel.z;
    } else {
        d_z[ind] = -9.f;
    }
}

// ---- TGV2 (src/TGV2_kernels.cu) --------------------------------------------

// :31-52
__global__ void tgv_updateP_kernel(float* d_Px, float* d_Py, const float* d_u, const float* d_u1x,
                                   const float* d_u1y, float alpha1, float sigma, int width, int height) {
    PIX2D
    const int xn = min(ind_x + 1, width - 1);
    const int yn = min(ind_y + 1, height - 1);
    const float as = alpha1 * sigma;
    const double dx = (double)(d_Px[ind] + as * (d_u[ind_y * width + xn] - d_u[ind] - d_u1x[ind]));
    const double dy = (double)(d_Py[ind] + as * (d_u[yn * width + ind_x] - d_u[ind] - d_u1y[ind]));
    const double d = (double)fmaxf(1.f, (float)sqrt(dx * dx + dy * dy));
    d_Px[ind] = (float)(dx / d);
    d_Py[ind] = (float)(dy / d);
}

// :54-85
__global__ void tgv_updateQ_kernel(float* d_Qx, float* d_Qy, float* d_Qz, float* d_Qw, const float* d_u1x,
                                   const float* d_u1y, float alpha0, float sigma, int width, int height) {
    PIX2D
    const int xn = min(ind_x + 1, width - 1);
    const int yn = min(ind_y + 1, height - 1);
    const float dx_u1x = d_u1x[ind_y * width + xn] - d_u1x[ind];
    const float dy_u1x = d_u1x[yn * width + ind_x] - d_u1x[ind];
    const float dx_u1y = d_u1y[ind_y * width + xn] - d_u1y[ind];
    const float dy_u1y = d_u1y[yn * width + ind_x] - d_u1y[ind];
    const float as = alpha0 * sigma;
    const double dx = (double)(d_Qx[ind] + as * dx_u1x);
    const double dy = (double)(d_Qy[ind] + as * dy_u1y);
    const double dz = (double)(d_Qz[ind] + as * (dy_u1x + dx_u1y) / 2.0f);
    const double dw = (double)(d_Qw[ind] + as * (dy_u1x + dx_u1y) / 2.0f);
    const double d = (double)fmaxf(1.f, (float)sqrt(dx * dx + dy * dy + dz * dz + dw * dw));
    d_Qx[ind] = (float)(dx / d);
    d_Qy[ind] = (float)(dy / d);
    d_Qz[ind] = (float)(dz / d);
    d_Qw[ind] = (float)(dw / d);
}

// :87-105
__global__ void tgv_updateR_kernel(float* d_r, float* d_prodsum, const float* d_u, const float* d_u0,
                                   const float* d_It, const float* d_Iu, float sigma, float lambda, int width,
                                   int height) {
    PIX2D
    float r = d_r[ind] + sigma * lambda * (d_It[ind] + (d_u[ind] - d_u0[ind]) * d_Iu[ind]);
    r = r / fmaxf(1.f, fabsf(r));
    d_r[ind] = r;
    d_prodsum[ind] += r * d_Iu[ind];
}

// :107-138
__global__ void tgv_updateU_kernel(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar,
                                   float* d_u1ybar, const float* d_Px, const float* d_Py, const float* d_Qx,
                                   const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_prodsum,
                                   float alpha0, float alpha1, float tau, float lambda, int width, int height) {
    PIX2D
    const int xp = max(ind_x - 1, 0);
    const int yp = max(ind_y - 1, 0);
    const float uprev = d_u[ind], u1xprev = d_u1x[ind], u1yprev = d_u1y[ind];
    const float u = uprev - tau * (-alpha1 * (d_Px[ind] - d_Px[ind_y * width + xp] + d_Py[ind] - d_Py[yp * width + ind_x]) +
                                   lambda * d_prodsum[ind]);
    const float u1x = u1xprev - tau * (-alpha1 * d_Px[ind] - alpha0 * (d_Qx[ind] - d_Qx[ind_y * width + xp] + d_Qz[ind] -
                                                                       d_Qz[yp * width + ind_x]));
    const float u1y = u1yprev - tau * (-alpha1 * d_Py[ind] - alpha0 * (d_Qz[ind] - d_Qz[ind_y * width + xp] + d_Qy[ind] -
                                                                       d_Qy[yp * width + ind_x]));
    d_u[ind] = u;
    d_u1x[ind] = u1x;
    d_u1y[ind] = u1y;
    d_ubar[ind] = 2.f * u - uprev;
    d_u1xbar[ind] = 2.f * u1x - u1xprev;
    d_u1ybar[ind] = 2.f * u1y - u1yprev;
}

struct TgvPose {
    M3f K, Rrel, invK;
    V3f trel;
};

// :141-179
__global__ void tgv_transform_coordinates_kernel(float* d_x, float* d_y, float* d_X, float* d_Y, float* d_Z,
                                                 const float* d_u, TgvPose pz, int width, int height) {
    PIX2D
    float a0, a1, a2;
    mv3(pz.invK, (float)(ind_x + 1), (float)(ind_y + 1), 1.f, a0, a1, a2);
    const float u = d_u[ind];
    a0 = u * a0; a1 = u * a1; a2 = u * a2;
    float b0, b1, b2;
    mv3(pz.Rrel, a0, a1, a2, b0, b1, b2);
    b0 = b0 + pz.trel.x; b1 = b1 + pz.trel.y; b2 = b2 + pz.trel.z;
    d_X[ind] = b0; d_Y[ind] = b1; d_Z[ind] = b2;
    float c0, c1, c2;
    mv3(pz.K, b0, b1, b2, c0, c1, c2);
    d_x[ind] = c0 / c2 - 1.f;
    d_y[ind] = c1 / c2 - 1.f;
}

// :181-192
__global__ void subtract_kernel(float* d_out, const float* a, const float* b, int width, int height) {
    PIX2D
    d_out[ind] = a[ind] - b[ind];
}

// :194-219
__global__ void tgv_coord_derivatives_kernel(float* d_dX, float* d_dY, float* d_dZ, M3f RK, int width, int height) {
    PIX2D
    float a0, a1, a2;
    mv3(RK, (float)(ind_x + 1), (float)(ind_y + 1), 1.f, a0, a1, a2);
    d_dX[ind] = a0; d_dY[ind] = a1; d_dZ[ind] = a2;
}

// :221-236
__global__ void tgv_derivativeF_kernel(float* d_dfx, float* d_dfy, const float* d_X, const float* d_dX,
                                       const float* d_Y, const float* d_dY, const float* d_Z, const float* d_dZ,
                                       float fx, float fy, int width, int height) {
    PIX2D
    const float Z = d_Z[ind];
    d_dfx[ind] = fx * (d_dX[ind] * Z - d_X[ind] * d_dZ[ind]) / (Z * Z);
    d_dfy[ind] = fy * (d_dY[ind] * Z - d_Y[ind] * d_dZ[ind]) / (Z * Z);
}

// :238-255
__global__ void tgv_Iu_kernel(float* d_Iu, const float* d_I, const float* d_dfx, const float* d_dfy, int width,
                              int height) {
    PIX2D
    const int xn = min(ind_x + 1, width - 1);
    const int yn = min(ind_y + 1, height - 1);
    const double dx = (double)(d_I[ind_y * width + xn] - d_I[ind]);
    const double dy = (double)(d_I[yn * width + ind_x] - d_I[ind]);
    d_Iu[ind] = (float)(dx * (double)d_dfx[ind] + dy * (double)d_dfy[ind]);
}

// :257-298
__global__ void tensor_kernel(float* d_T11, float* d_T12, float* d_T21, float* d_T22, const float* d_Img, float beta,
                              float gamma, int width, int height) {
    PIX2D
    const int xn = min(ind_x + 1, width - 1);
    const int yn = min(ind_y + 1, height - 1);
    float x = d_Img[ind_y * width + xn] - d_Img[ind];
    float y = d_Img[yn * width + ind_x] - d_Img[ind];
    const float d = sqrtf(x * x + y * y);
    if (d > 0.f) {
        x = x / d;
        y = y / d;
        const float k = expf(-beta * powf(d, gamma));
        d_T11[ind] = k * x * x + y * y;
        d_T12[ind] = (k - 1.f) * x * y;
        d_T21[ind] = (k - 1.f) * x * y;
        d_T22[ind] = k * y * y + x * x;
    } else {
        d_T11[ind] = 1.f;
        d_T12[ind] = 0.f;
        d_T21[ind] = 0.f;
        d_T22[ind] = 1.f;
    }
}

// :300-338
__global__ void tgv_updateU_tensor_kernel(float* d_u, float* d_u1x, float* d_u1y, const float* T11,
                                          const float* T12, const float* T21, const float* T22, float* d_ubar,
                                          float* d_u1xbar, float* d_u1ybar, const float* d_Px, const float* d_Py,
                                          const float* d_Qx, const float* d_Qy, const float* d_Qz, const float* d_Qw,
                                          const float* d_prodsum, float alpha0, float alpha1, float tau,
                                          float lambda, int width, int height) {
    PIX2D
    const int xp = max(ind_x - 1, 0);
    const int yp = max(ind_y - 1, 0);
    const float uprev = d_u[ind], u1xprev = d_u1x[ind], u1yprev = d_u1y[ind];
    const float c_px = d_Px[ind], c_py = d_Py[ind], xp_px = d_Px[ind_y * width + xp],
                yp_px = d_Px[yp * width + ind_x], xp_py = d_Py[ind_y * width + xp],
                yp_py = d_Py[yp * width + ind_x];
    const float t11 = T11[ind], t12 = T12[ind], t21 = T21[ind], t22 = T22[ind];
    const float u = uprev - tau * (-alpha1 * (t11 * (c_px - xp_px) + t12 * (c_py - xp_py) + t21 * (c_px - yp_px) +
                                              t22 * (c_py - yp_py)) +
                                   lambda * d_prodsum[ind]);
    const float u1x = u1xprev - tau * (-alpha1 * (t11 * c_px + t12 * c_py) -
                                       alpha0 * (d_Qx[ind] - d_Qx[ind_y * width + xp] + d_Qz[ind] -
                                                 d_Qz[yp * width + ind_x]));
    const float u1y = u1yprev - tau * (-alpha1 * (t21 * c_px + t22 * c_py) -
                                       alpha0 * (d_Qz[ind] - d_Qz[ind_y * width + xp] + d_Qy[ind] -
                                                 d_Qy[yp * width + ind_x]));
    d_u[ind] = u;
    d_u1x[ind] = u1x;
    d_u1y[ind] = u1y;
    d_ubar[ind] = 2.f * u - uprev;
    d_u1xbar[ind] = 2.f * u1x - u1xprev;
    d_u1ybar[ind] = 2.f * u1y - u1yprev;
}

// :340-365
__global__ void tgv_updateP_tensor_kernel(float* d_Px, float* d_Py, const float* T11, const float* T12,
                                          const float* T21, const float* T22, const float* d_u, const float* d_u1x,
                                          const float* d_u1y, float alpha1, float sigma, int width, int height) {
    PIX2D
    const int xn = min(ind_x + 1, width - 1);
    const int yn = min(ind_y + 1, height - 1);
    const double x = (double)(d_u[ind_y * width + xn] - d_u[ind] - d_u1x[ind]);
    const double y = (double)(d_u[yn * width + ind_x] - d_u[ind] - d_u1y[ind]);
    const double as = (double)(alpha1 * sigma);
    const double dx = (double)d_Px[ind] + as * ((double)T11[ind] * x + (double)T12[ind] * y);
    const double dy = (double)d_Py[ind] + as * ((double)T21[ind] * x + (double)T22[ind] * y);
    const double d = (double)fmaxf(1.f, (float)sqrt(dx * dx + dy * dy));
    d_Px[ind] = (float)(dx / d);
    d_Py[ind] = (float)(dy / d);
}

// :367-397
__global__ void tgv_updateU_sparse_kernel(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar,
                                          float* d_u1ybar, const float* d_Px, const float* d_Py, const float* d_Qx,
                                          const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_w,
                                          const float* d_Ds, float alpha0, float alpha1, float tau, float theta,
                                          int width, int height) {
    PIX2D
    const int xp = max(ind_x - 1, 0);
    const int yp = max(ind_y - 1, 0);
    const float c_px = d_Px[ind], c_py = d_Py[ind], xp_px = d_Px[ind_y * width + xp], yp_py = d_Py[yp * width + ind_x];
    d_u[ind] = (d_u[ind] + tau * (alpha1 * ((c_px - xp_px) + (c_py - yp_py)) + d_w[ind] * d_Ds[ind])) /
               (1.f + tau * d_w[ind]);
    d_u1x[ind] = d_u1x[ind] - tau * (-alpha1 * (c_px) - alpha0 * (d_Qx[ind] - d_Qx[ind_y * width + xp] + d_Qz[ind] -
                                                                  d_Qz[yp * width + ind_x]));
    d_u1y[ind] = d_u1y[ind] - tau * (-alpha1 * (c_py) - alpha0 * (d_Qz[ind] - d_Qz[ind_y * width + xp] + d_Qy[ind] -
                                                                  d_Qy[yp * width + ind_x]));
    d_ubar[ind] = d_u[ind] + theta * (d_u[ind] - d_ubar[ind]);
    d_u1xbar[ind] = d_u1x[ind] + theta * (d_u1x[ind] - d_u1xbar[ind]);
    d_u1ybar[ind] = d_u1y[ind] + theta * (d_u1y[ind] - d_u1ybar[ind]);
}

// :399-432
__global__ void tgv_updateU_sparse_tensor_kernel(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar,
                                                 float* d_u1xbar, float* d_u1ybar, const float* T11,
                                                 const float* T12, const float* T21, const float* T22,
                                                 const float* d_Px, const float* d_Py, const float* d_Qx,
                                                 const float* d_Qy, const float* d_Qz, const float* d_Qw,
                                                 const float* d_w, const float* d_Ds, float alpha0, float alpha1,
                                                 float tau, float theta, int width, int height) {
    PIX2D
    const int xp = max(ind_x - 1, 0);
    const int yp = max(ind_y - 1, 0);
    const float c_px = d_Px[ind], c_py = d_Py[ind], xp_px = d_Px[ind_y * width + xp],
                yp_px = d_Px[yp * width + ind_x], xp_py = d_Py[ind_y * width + xp],
                yp_py = d_Py[yp * width + ind_x];
    d_u[ind] = (d_u[ind] + tau * (alpha1 * (T11[ind] * (c_px - xp_px) + T12[ind] * (c_py - xp_py) +
                                            T21[ind] * (c_px - yp_px) + T22[ind] * (c_py - yp_py)) +
                                  d_w[ind] * d_Ds[ind])) /
               (1.f + tau * d_w[ind]);
    d_u1x[ind] = d_u1x[ind] - tau * (-alpha1 * (T11[ind] * c_px + T12[ind] * c_py) -
                                     alpha0 * (d_Qx[ind] - d_Qx[ind_y * width + xp] + d_Qz[ind] -
                                               d_Qz[yp * width + ind_x]));
    d_u1y[ind] = d_u1y[ind] - tau * (-alpha1 * (T21[ind] * c_px + T22[ind] * c_py) -
                                     alpha0 * (d_Qz[ind] - d_Qz[ind_y * width + xp] + d_Qy[ind] -
                                               d_Qy[yp * width + ind_x]));
    d_ubar[ind] = d_u[ind] + theta * (d_u[ind] - d_ubar[ind]);
    d_u1xbar[ind] = d_u1x[ind] + theta * (d_u1x[ind] - d_u1xbar[ind]);
    d_u1ybar[ind] = d_u1y[ind] + theta * (d_u1y[ind] - d_u1ybar[ind]);
}

// :434-445
__global__ void sparse_weights_kernel(float* d_w, const float* d_Ds, int width, int height) {
    PIX2D
    d_w[ind] = d_Ds[ind] > 0.f ? 1.f : 0.f;
}

}  // namespace psw

using namespace psw;

#define LAUNCH_CHECK(name)                                  \
    Shape sh;                                               \
    {                                                       \
        int st_ = make_shape(blocks, threads, width, height, sh); \
        if (st_ != PSW_OK) return st_;                      \
    }                                                       \
    if (width == 0 || height == 0) return PSW_OK;

extern "C" {

int psw_transform_indexes(float* d_x, float* d_y, const float* h, int width, int height, psw_dim3 blocks,
                          psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_x && d_y && h, "NULL pointer");
    LAUNCH_CHECK();
    transform_indexes_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_x, d_y, m3_from(h), width, height);
    return launch_status("transform_indexes_kernel");
}

int psw_bilinear_interpolation(float* d_result, const float* d_data, const float* d_xout, const float* d_yout, int M1,
                               int M2, int N1, int N2, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_result && d_data && d_xout && d_yout, "NULL pointer");
    PSW_CHECK_ARG(M1 >= 0 && M2 >= 0, "negative data size");
    PSW_CHECK_ARG((uint64_t)M1 * (uint64_t)M2 < (1ull << 32), "source image above 2^32 pixels");
    const int width = N1, height = N2;
    LAUNCH_CHECK();
    bilinear_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_result, d_data, d_xout, d_yout, M1, M2, N1, N2);
    return launch_status("bilinear_kernel");
}

int psw_calcNCC(float* d_ncc, const float* d_prod_mean, const float* d_mean1, const float* d_mean2,
                const float* d_std1, const float* d_std2, float stdthresh1, float stdthresh2, int width, int height,
                psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_ncc && d_prod_mean && d_mean1 && d_mean2 && d_std1 && d_std2, "NULL pointer");
    LAUNCH_CHECK();
    calcNCC_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_ncc, d_prod_mean, d_mean1, d_mean2, d_std1, d_std2, stdthresh1,
                                                   stdthresh2, width, height);
    return launch_status("calcNCC_kernel");
}

int psw_update_arrays(float* d_depthmap, float* d_bestncc, const float* d_currentncc, float current_depth, int width,
                      int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_depthmap && d_bestncc && d_currentncc, "NULL pointer");
    LAUNCH_CHECK();
    update_arrays_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_depthmap, d_bestncc, d_currentncc, current_depth, width,
                                                         height);
    return launch_status("update_arrays_kernel");
}

int psw_sum_depthmap_NCC(float* d_depthmap_out, float* d_count, const float* d_depthmap, const float* d_ncc,
                         float nccthreshold, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                         psw_stream_t s) {
    PSW_CHECK_ARG(d_depthmap_out && d_count && d_depthmap && d_ncc, "NULL pointer");
    LAUNCH_CHECK();
    sum_depthmap_NCC_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_depthmap_out, d_count, d_depthmap, d_ncc, nccthreshold,
                                                            width, height);
    return launch_status("sum_depthmap_NCC_kernel");
}

int psw_calculate_STD(float* d_std, const float* d_mean, const float* d_mean_of_squares, int width, int height,
                      psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_std && d_mean && d_mean_of_squares, "NULL pointer");
    LAUNCH_CHECK();
    calculate_STD_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_std, d_mean, d_mean_of_squares, width, height);
    return launch_status("calculate_STD_kernel");
}

int psw_set_value(float* d_output, float value, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                  psw_stream_t s) {
    PSW_CHECK_ARG(d_output, "NULL pointer");
    LAUNCH_CHECK();
    set_value_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, value, width, height);
    return launch_status("set_value_kernel");
}

int psw_element_multiply(float* d_output, const float* d_input1, const float* d_input2, int width, int height,
                         psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_output && d_input1 && d_input2, "NULL pointer");
    LAUNCH_CHECK();
    element_multiply_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, d_input1, d_input2, width, height);
    return launch_status("element_multiply_kernel");
}

int psw_element_rdivide(float* d_output, const float* d_input1, const float* d_input2, int width, int height,
                        psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_output && d_input1 && d_input2, "NULL pointer");
    LAUNCH_CHECK();
    element_rdivide_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, d_input1, d_input2, width, height);
    return launch_status("element_rdivide_kernel");
}

int psw_windowed_mean_row(float* d_output, const float* d_input, unsigned winsize, int squared, int width, int height,
                          psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_output && d_input, "NULL pointer");
    PSW_CHECK_ARG(winsize >= 1 && (int)(winsize / 2) < width, "window half-size must be below the width");
    LAUNCH_CHECK();
    windowed_mean_row_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, d_input, winsize, squared != 0, width, height);
    return launch_status("windowed_mean_row_kernel");
}

int psw_windowed_mean_column(float* d_output, const float* d_input, unsigned winsize, int squared, int width,
                             int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_output && d_input, "NULL pointer");
    PSW_CHECK_ARG(winsize >= 1 && (int)(winsize / 2) < height, "window half-size must be below the height");
    LAUNCH_CHECK();
    windowed_mean_column_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, d_input, winsize, squared != 0, width,
                                                                height);
    return launch_status("windowed_mean_column_kernel");
}

int psw_element_scale(float* d_output, float scale, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                      psw_stream_t s) {
    PSW_CHECK_ARG(d_output, "NULL pointer");
    LAUNCH_CHECK();
    element_scale_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, scale, width, height);
    return launch_status("element_scale_kernel");
}

int psw_element_add(float* d_output, float value, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                    psw_stream_t s) {
    PSW_CHECK_ARG(d_output, "NULL pointer");
    LAUNCH_CHECK();
    element_add_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, value, width, height);
    return launch_status("element_add_kernel");
}

int psw_set_QNAN_value(float* d_output, float value, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                       psw_stream_t s) {
    PSW_CHECK_ARG(d_output, "NULL pointer");
    LAUNCH_CHECK();
    set_QNAN_value_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, value, width, height);
    return launch_status("set_QNAN_value_kernel");
}

int psw_denoising_TVL1_calculateP(float* d_Px, float* d_Py, const float* d_input, float sigma, int width, int height,
                                  psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_Px && d_Py && d_input, "NULL pointer");
    LAUNCH_CHECK();
    tvl1_calculateP_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_Px, d_Py, d_input, sigma, width, height);
    return launch_status("tvl1_calculateP_kernel");
}

int psw_denoising_TVL1_calculateP_tensor_weighed(float* d_Px, float* d_Py, const float* d_T11, const float* d_T12,
                                                 const float* d_T21, const float* d_T22, const float* d_input,
                                                 float sigma, int width, int height, psw_dim3 blocks,
                                                 psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_Px && d_Py && d_T11 && d_T12 && d_T21 && d_T22 && d_input, "NULL pointer");
    LAUNCH_CHECK();
    tvl1_calculateP_tensor_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_Px, d_Py, d_T11, d_T12, d_T21, d_T22, d_input,
                                                                  sigma, width, height);
    return launch_status("tvl1_calculateP_tensor_kernel");
}

int psw_denoising_TVL1_update(float* d_output, float* d_R, const float* d_Px, const float* d_Py,
                              const float* d_origin, float tau, float theta, float lambda, float sigma, int width,
                              int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_output && d_R && d_Px && d_Py && d_origin, "NULL pointer");
    LAUNCH_CHECK();
    tvl1_update_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, d_R, d_Px, d_Py, d_origin, tau, theta, lambda, sigma,
                                                       width, height);
    return launch_status("tvl1_update_kernel");
}

int psw_denoising_TVL1_update_tensor_weighed(float* d_output, float* d_R, const float* d_Px, const float* d_Py,
                                             const float* d_origin, const float* d_T11, const float* d_T12,
                                             const float* d_T21, const float* d_T22, float tau, float theta,
                                             float lambda, float sigma, int width, int height, psw_dim3 blocks,
                                             psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_output && d_R && d_Px && d_Py && d_origin && d_T11 && d_T12 && d_T21 && d_T22, "NULL pointer");
    LAUNCH_CHECK();
    tvl1_update_tensor_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_output, d_R, d_Px, d_Py, d_origin, d_T11, d_T12,
                                                              d_T21, d_T22, tau, theta, lambda, sigma, width, height);
    return launch_status("tvl1_update_tensor_kernel");
}

int psw_compute3D(float* d_x, float* d_y, float* d_z, const float* Rrel, const float* trel, const float* invK,
                  int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_x && d_y && d_z && Rrel && trel && invK, "NULL pointer");
    LAUNCH_CHECK();
    compute3D_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_x, d_y, d_z, m3_from(Rrel), v3_from(trel), m3_from(invK), width,
                                                     height);
    return launch_status("compute3D_kernel");
}

int psw_TGV2_updateP(float* d_Px, float* d_Py, const float* d_u, const float* d_u1x, const float* d_u1y, float alpha1,
                     float sigma, int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_Px && d_Py && d_u && d_u1x && d_u1y, "NULL pointer");
    LAUNCH_CHECK();
    tgv_updateP_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_Px, d_Py, d_u, d_u1x, d_u1y, alpha1, sigma, width, height);
    return launch_status("tgv_updateP_kernel");
}

int psw_TGV2_updateQ(float* d_Qx, float* d_Qy, float* d_Qz, float* d_Qw, const float* d_u1x, const float* d_u1y,
                     float alpha0, float sigma, int width, int height, psw_dim3 blocks, psw_dim3 threads,
                     psw_stream_t s) {
    PSW_CHECK_ARG(d_Qx && d_Qy && d_Qz && d_Qw && d_u1x && d_u1y, "NULL pointer");
    LAUNCH_CHECK();
    tgv_updateQ_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_Qx, d_Qy, d_Qz, d_Qw, d_u1x, d_u1y, alpha0, sigma, width,
                                                       height);
    return launch_status("tgv_updateQ_kernel");
}

int psw_TGV2_updateR(float* d_r, float* d_prodsum, const float* d_u, const float* d_u0, const float* d_It,
                     const float* d_Iu, float sigma, float lambda, int width, int height, psw_dim3 blocks,
                     psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_r && d_prodsum && d_u && d_u0 && d_It && d_Iu, "NULL pointer");
    LAUNCH_CHECK();
    tgv_updateR_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_r, d_prodsum, d_u, d_u0, d_It, d_Iu, sigma, lambda, width,
                                                       height);
    return launch_status("tgv_updateR_kernel");
}

int psw_TGV2_updateU(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar, float* d_u1ybar,
                     const float* d_Px, const float* d_Py, const float* d_Qx, const float* d_Qy, const float* d_Qz,
                     const float* d_Qw, const float* d_prodsum, float alpha0, float alpha1, float tau, float lambda,
                     int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_u && d_u1x && d_u1y && d_ubar && d_u1xbar && d_u1ybar && d_Px && d_Py && d_Qx && d_Qy && d_Qz &&
                      d_Qw && d_prodsum,
                  "NULL pointer");
    LAUNCH_CHECK();
    tgv_updateU_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_u, d_u1x, d_u1y, d_ubar, d_u1xbar, d_u1ybar, d_Px, d_Py, d_Qx,
                                                       d_Qy, d_Qz, d_Qw, d_prodsum, alpha0, alpha1, tau, lambda, width,
                                                       height);
    return launch_status("tgv_updateU_kernel");
}

int psw_TGV2_transform_coordinates(float* d_x, float* d_y, float* d_X, float* d_Y, float* d_Z, const float* d_u,
                                   const float* K, const float* Rrel, const float* trel, const float* invK, int width,
                                   int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_x && d_y && d_X && d_Y && d_Z && d_u && K && Rrel && trel && invK, "NULL pointer");
    LAUNCH_CHECK();
    TgvPose pz;
    pz.K = m3_from(K);
    pz.Rrel = m3_from(Rrel);
    pz.invK = m3_from(invK);
    pz.trel = v3_from(trel);
    tgv_transform_coordinates_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_x, d_y, d_X, d_Y, d_Z, d_u, pz, width, height);
    return launch_status("tgv_transform_coordinates_kernel");
}

int psw_subtract(float* d_out, const float* d_in1, const float* d_in2, int width, int height, psw_dim3 blocks,
                 psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_out && d_in1 && d_in2, "NULL pointer");
    LAUNCH_CHECK();
    subtract_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_out, d_in1, d_in2, width, height);
    return launch_status("subtract_kernel");
}

int psw_TGV2_calculate_coordinate_derivatives(float* d_dX, float* d_dY, float* d_dZ, const float* invK,
                                              const float* Rrel, int width, int height, psw_dim3 blocks,
                                              psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_dX && d_dY && d_dZ && invK && Rrel, "NULL pointer");
    LAUNCH_CHECK();
    // Rrel * invK * (x+1, y+1, 1) associates left: the matrix product first.
    const M3f RK = mm3(m3_from(Rrel), m3_from(invK));
    tgv_coord_derivatives_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_dX, d_dY, d_dZ, RK, width, height);
    return launch_status("tgv_coord_derivatives_kernel");
}

int psw_TGV2_calculate_derivativeF(float* d_dfx, float* d_dfy, const float* d_X, const float* d_dX, const float* d_Y,
                                   const float* d_dY, const float* d_Z, const float* d_dZ, float fx, float fy,
                                   int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_dfx && d_dfy && d_X && d_dX && d_Y && d_dY && d_Z && d_dZ, "NULL pointer");
    LAUNCH_CHECK();
    tgv_derivativeF_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_dfx, d_dfy, d_X, d_dX, d_Y, d_dY, d_Z, d_dZ, fx, fy, width,
                                                           height);
    return launch_status("tgv_derivativeF_kernel");
}

int psw_TGV2_calculate_Iu(float* d_Iu, const float* d_I, const float* d_dfx, const float* d_dfy, int width,
                          int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_Iu && d_I && d_dfx && d_dfy, "NULL pointer");
    LAUNCH_CHECK();
    tgv_Iu_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_Iu, d_I, d_dfx, d_dfy, width, height);
    return launch_status("tgv_Iu_kernel");
}

int psw_Anisotropic_diffusion_tensor(float* d_T11, float* d_T12, float* d_T21, float* d_T22, const float* d_Img,
                                     float beta, float gamma, int width, int height, psw_dim3 blocks,
                                     psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_T11 && d_T12 && d_T21 && d_T22 && d_Img, "NULL pointer");
    LAUNCH_CHECK();
    tensor_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_T11, d_T12, d_T21, d_T22, d_Img, beta, gamma, width, height);
    return launch_status("tensor_kernel");
}

int psw_TGV2_updateU_tensor_weighed(float* d_u, float* d_u1x, float* d_u1y, const float* d_T11, const float* d_T12,
                                    const float* d_T21, const float* d_T22, float* d_ubar, float* d_u1xbar,
                                    float* d_u1ybar, const float* d_Px, const float* d_Py, const float* d_Qx,
                                    const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_prodsum,
                                    float alpha0, float alpha1, float tau, float lambda, int width, int height,
                                    psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_u && d_u1x && d_u1y && d_T11 && d_T12 && d_T21 && d_T22 && d_ubar && d_u1xbar && d_u1ybar &&
                      d_Px && d_Py && d_Qx && d_Qy && d_Qz && d_Qw && d_prodsum,
                  "NULL pointer");
    LAUNCH_CHECK();
    tgv_updateU_tensor_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_u, d_u1x, d_u1y, d_T11, d_T12, d_T21, d_T22, d_ubar,
                                                              d_u1xbar, d_u1ybar, d_Px, d_Py, d_Qx, d_Qy, d_Qz, d_Qw,
                                                              d_prodsum, alpha0, alpha1, tau, lambda, width, height);
    return launch_status("tgv_updateU_tensor_kernel");
}

int psw_TGV2_updateP_tensor_weighed(float* d_Px, float* d_Py, const float* d_T11, const float* d_T12,
                                    const float* d_T21, const float* d_T22, const float* d_u, const float* d_u1x,
                                    const float* d_u1y, float alpha1, float sigma, int width, int height,
                                    psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_Px && d_Py && d_T11 && d_T12 && d_T21 && d_T22 && d_u && d_u1x && d_u1y, "NULL pointer");
    LAUNCH_CHECK();
    tgv_updateP_tensor_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_Px, d_Py, d_T11, d_T12, d_T21, d_T22, d_u, d_u1x,
                                                              d_u1y, alpha1, sigma, width, height);
    return launch_status("tgv_updateP_tensor_kernel");
}

int psw_TGV2_updateU_sparseDepth(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar,
                                 float* d_u1ybar, const float* d_Px, const float* d_Py, const float* d_Qx,
                                 const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_w,
                                 const float* d_Ds, float alpha0, float alpha1, float tau, float theta, int width,
                                 int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_u && d_u1x && d_u1y && d_ubar && d_u1xbar && d_u1ybar && d_Px && d_Py && d_Qx && d_Qy && d_Qz &&
                      d_Qw && d_w && d_Ds,
                  "NULL pointer");
    LAUNCH_CHECK();
    tgv_updateU_sparse_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_u, d_u1x, d_u1y, d_ubar, d_u1xbar, d_u1ybar, d_Px, d_Py,
                                                              d_Qx, d_Qy, d_Qz, d_Qw, d_w, d_Ds, alpha0, alpha1, tau,
                                                              theta, width, height);
    return launch_status("tgv_updateU_sparse_kernel");
}

int psw_TGV2_updateU_sparseDepthTensor(float* d_u, float* d_u1x, float* d_u1y, float* d_ubar, float* d_u1xbar,
                                       float* d_u1ybar, const float* d_T11, const float* d_T12, const float* d_T21,
                                       const float* d_T22, const float* d_Px, const float* d_Py, const float* d_Qx,
                                       const float* d_Qy, const float* d_Qz, const float* d_Qw, const float* d_w,
                                       const float* d_Ds, float alpha0, float alpha1, float tau, float theta,
                                       int width, int height, psw_dim3 blocks, psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_u && d_u1x && d_u1y && d_ubar && d_u1xbar && d_u1ybar && d_T11 && d_T12 && d_T21 && d_T22 &&
                      d_Px && d_Py && d_Qx && d_Qy && d_Qz && d_Qw && d_w && d_Ds,
                  "NULL pointer");
    LAUNCH_CHECK();
    tgv_updateU_sparse_tensor_kernel<<<sh.grid, sh.block, 0, S(s)>>>(
        d_u, d_u1x, d_u1y, d_ubar, d_u1xbar, d_u1ybar, d_T11, d_T12, d_T21, d_T22, d_Px, d_Py, d_Qx, d_Qy, d_Qz, d_Qw,
        d_w, d_Ds, alpha0, alpha1, tau, theta, width, height);
    return launch_status("tgv_updateU_sparse_tensor_kernel");
}

int psw_calculateWeights_sparseDepth(float* d_w, const float* d_Ds, int width, int height, psw_dim3 blocks,
                                     psw_dim3 threads, psw_stream_t s) {
    PSW_CHECK_ARG(d_w && d_Ds, "NULL pointer");
    LAUNCH_CHECK();
    sparse_weights_kernel<<<sh.grid, sh.block, 0, S(s)>>>(d_w, d_Ds, width, height);
    return launch_status("sparse_weights_kernel");
}

}  // extern "C"

