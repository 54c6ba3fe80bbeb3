// tgv_point.h -- TGV2 point updates (diffusion tensor, duals) shared by the dense
// (tgv_fused.hip) and sparse-depth (tgv_sparse.hip) fused kernels.
#pragma once

namespace psw {

// ---- correctly rounded square root and division for the dual projections
//
// sqrtf and '/' compile to general sequences that first rescale operands
// that could under- or overflow and then patch special values.  Inside a
// projection the argument of the root is s > 1 and the quotients are
// |x| / sqrt(s) <= 1, so most of that is dead weight.  The helpers below run
// the same core steps without the rescaling; every value they return is
// the correctly rounded result (tested bit for bit against sqrtf and '/').

// sqrt(s), correctly rounded, for s > 1 (finite or +inf): the hardware
// estimate y (within 1 ulp) is replaced by its lower neighbour when
// y_lo * y >= s and by its upper neighbour when y_hi * y < s, both decided
// on exact FMA residuals (Markstein's rounding test for the square root).
__device__ __forceinline__ float sqrt_rn_gt1(float s) {
    const float y = __builtin_amdgcn_sqrtf(s);
    const float ylo = __int_as_float(__float_as_int(y) - 1);
    const float yhi = __int_as_float(__float_as_int(y) + 1);
    float r = y;
    if (__builtin_fmaf(-ylo, y, s) <= 0.f) r = ylo;
    if (__builtin_fmaf(-yhi, y, s) > 0.f) r = yhi;
    return r;
}

// 1/d to about half an ulp: the hardware reciprocal and one Newton step.
__device__ __forceinline__ float rcp_nr(float d) {
    const float y = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(__builtin_fmaf(-d, y, 1.f), y, y);
}

// x / d, correctly rounded, given y = rcp_nr(d): two residual corrections
// of the quotient x * y (Markstein).  Exact steps, hence the correctly
// rounded quotient, while no intermediate leaves the normal range; the
// callers guarantee 1 < d < 2^20 and 2^-100 <= |x| (and take '/' otherwise).
__device__ __forceinline__ float div_nr(float x, float d, float y) {
    const float q0 = x * y;
    const float q1 = __builtin_fmaf(__builtin_fmaf(-d, q0, x), y, q0);
    return __builtin_fmaf(__builtin_fmaf(-d, q1, x), y, q1);
}

// True when div_nr(x_i, sqrt(s), .) is exact for every |x_i| >= m.
__device__ __forceinline__ bool div_nr_ok(float s, float m) { return s < 0x1p40f && m >= 0x1p-100f; }

// TGV2_updateP_tensor_weighed + TGV2_updateQ (q_w == q_z) at one position:
// ub*/u1xb*/u1yb* are ubar at the position (c), its right (r) and lower (d)
// neighbour (the position itself at the image's last column / row).
__device__ __forceinline__ void pq_update(float& px, float& py, float& qx, float& qy, float& qz, float t11, float t12,
                                          float t22, float ubc, float ubr, float ubd, float u1xc, float u1xr,
                                          float u1xd, float u1yc, float u1yr, float u1yd, float as1, float as0) {
    const float gxv = ubr - ubc - u1xc;
    const float gyv = ubd - ubc - u1yc;
    px = px + as1 * (t11 * gxv + t12 * gyv);
    py = py + as1 * (t12 * gxv + t22 * gyv);
    // x / fmaxf(1, sqrt(s)) is x itself for s <= 1 (and for NaN s), so the
    // root and the divisions are only needed outside the unit ball.
    const float sp = px * px + py * py;
    if (sp > 1.f) {
        const float dp = sqrt_rn_gt1(sp);
        if (div_nr_ok(sp, fminf(fabsf(px), fabsf(py)))) {
            const float y = rcp_nr(dp);
            px = div_nr(px, dp, y);
            py = div_nr(py, dp, y);
        } else {
            px = px / dp;
            py = py / dp;
        }
    }
    const float dx_u1x = u1xr - u1xc;
    const float dy_u1x = u1xd - u1xc;
    const float dx_u1y = u1yr - u1yc;
    const float dy_u1y = u1yd - u1yc;
    qx = qx + as0 * dx_u1x;
    qy = qy + as0 * dy_u1y;
    qz = qz + as0 * (dy_u1x + dx_u1y) / 2.0f;
    const float sq = qx * qx + qy * qy + qz * qz + qz * qz;
    if (sq > 1.f) {
        const float dq = sqrt_rn_gt1(sq);
        if (div_nr_ok(sq, fminf(fminf(fabsf(qx), fabsf(qy)), fabsf(qz)))) {
            const float y = rcp_nr(dq);
            qx = div_nr(qx, dq, y);
            qy = div_nr(qy, dq, y);
            qz = div_nr(qz, dq, y);
        } else {
            qx = qx / dq;
            qy = qy / dq;
            qz = qz / dq;
        }
    }
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
