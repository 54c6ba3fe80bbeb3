// rn_math.h -- correctly rounded square root and division for the
// unit-ball projections of the primal-dual kernels (TGV2 duals in
// tgv_point.h, fusion duals in fusion.hip).
#pragma once

namespace psw {

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

// (x, y) / fmaxf(1, sqrt(s)) with s the caller's squared norm, correctly
// rounded: (x, y) itself for s <= 1 (and NaN s), no root or division there.
__device__ __forceinline__ void project_ball2(float s, float& x, float& y) {
    if (s > 1.f) {
        const float d = sqrt_rn_gt1(s);
        if (div_nr_ok(s, fminf(fabsf(x), fabsf(y)))) {
            const float r = rcp_nr(d);
            x = div_nr(x, d, r);
            y = div_nr(y, d, r);
        } else {
            x = x / d;
            y = y / d;
        }
    }
}

// (x, y, z) / fmaxf(1, sqrt(s)), as project_ball2.
__device__ __forceinline__ void project_ball3(float s, float& x, float& y, float& z) {
    if (s > 1.f) {
        const float d = sqrt_rn_gt1(s);
        if (div_nr_ok(s, fminf(fminf(fabsf(x), fabsf(y)), fabsf(z)))) {
            const float r = rcp_nr(d);
            x = div_nr(x, d, r);
            y = div_nr(y, d, r);
            z = div_nr(z, d, r);
        } else {
            x = x / d;
            y = y / d;
            z = z / d;
        }
    }
}

}  // namespace psw
