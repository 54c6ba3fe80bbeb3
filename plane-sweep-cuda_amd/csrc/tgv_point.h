// tgv_point.h -- TGV2 point updates (diffusion tensor, duals) shared by the dense
// (tgv_fused.hip) and sparse-depth (tgv_sparse.hip) fused kernels.
#pragma once

#include "rn_math.h"

namespace psw {

// TGV2_updateP_tensor_weighed + TGV2_updateQ (q_w == q_z) at one position:
// ub*/u1xb*/u1yb* are ubar at the position (c), its right (r) and lower (d)
// neighbour (the position itself at the image's last column / row).  Each
// product that feeds an addition is fused into it (one rounding); /2 is the
// exact *0.5 moved inside the product.
__device__ __forceinline__ void pq_update(float& px, float& py, float& qx, float& qy, float& qz, float t11, float t12,
                                          float t22, float ubc, float ubr, float ubd, float u1xc, float u1xr,
                                          float u1xd, float u1yc, float u1yr, float u1yd, float as1, float as0) {
    const float gxv = ubr - ubc - u1xc;
    const float gyv = ubd - ubc - u1yc;
    px = __builtin_fmaf(as1, __builtin_fmaf(t11, gxv, t12 * gyv), px);
    py = __builtin_fmaf(as1, __builtin_fmaf(t12, gxv, t22 * gyv), py);
    // x / fmaxf(1, sqrt(s)) is x itself for s <= 1 (and for NaN s), so the
    // root and the divisions are only needed outside the unit ball.
    project_ball2(__builtin_fmaf(px, px, py * py), px, py);
    const float dx_u1x = u1xr - u1xc;
    const float dy_u1x = u1xd - u1xc;
    const float dx_u1y = u1yr - u1yc;
    const float dy_u1y = u1yd - u1yc;
    qx = __builtin_fmaf(as0, dx_u1x, qx);
    qy = __builtin_fmaf(as0, dy_u1y, qy);
    qz = __builtin_fmaf(as0, 0.5f * (dy_u1x + dx_u1y), qz);
    const float sq = __builtin_fmaf(qz, qz, __builtin_fmaf(qz, qz, __builtin_fmaf(qy, qy, qx * qx)));
    project_ball3(sq, qx, qy, qz);
}

// Anisotropic_diffusion_tensor_kernel (TGV2_kernels.cu:257-298) at one
// pixel from the forward differences (gx, gy) of the scaled image; T21 = T12.
__device__ __forceinline__ void tensor_point(float gx, float gy, float beta, float gamma, float& t11, float& t12,
                                             float& t22) {
    const float d = sqrtf(gx * gx + gy * gy);
    t11 = 1.f;
    t12 = 0.f;
    t22 = 1.f;
    if (d > 0.f) {
        gx = gx / d;
        gy = gy / d;
        const float k = expf(-beta * powf(d, gamma));
        t11 = k * gx * gx + gy * gy;
        t12 = (k - 1.f) * gx * gy;
        t22 = k * gy * gy + gx * gx;
    }
}

}  // namespace psw
