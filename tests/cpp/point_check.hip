// Bit-for-bit check of the dual-projection arithmetic of the fused TGV2
// kernels (plane-sweep-cuda_amd/csrc/tgv_point.h) against the plain
// expressions it replaces: sqrtf and IEEE '/'.  Test infrastructure: built
// into tests/cpp/_build/libpoint_check.so and driven by
// tests/test_point_math.py on the GPU.
//
// Inputs are generated on the device from a counter-based hash, so one call
// covers millions of cases without host buffers:
//   - roots: s > 1 over every binade up to 2^127 (and +inf);
//   - quotients x / sqrt(s) with |x| from 2^-149 up to sqrt(s), +-0;
//   - whole dual updates (pq_update) against the formula with sqrtf and '/',
//     and the fusion dual's 3-vector projection.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../plane-sweep-cuda_amd/csrc/tgv_point.h"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {   // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// a float with the given unbiased exponent and a hashed mantissa
__device__ __forceinline__ float with_exp(uint64_t h, int e) {
    const uint32_t bits = (uint32_t)((e + 127) << 23) | (uint32_t)(h & 0x7fffff);
    return __uint_as_float(bits);
}

// The projection as the reference writes it (TGV2_updateP_tensor_weighed,
// TGV2_updateQ): x / fmaxf(1, sqrt(s)), with sqrtf and IEEE division; the
// steps before it are the kernels' (products fused into the additions).
__device__ bool pq_plain(float& px, float& py, float& qx, float& qy, float& qz, float t11, float t12, float t22,
                         float ubc, float ubr, float ubd, float u1xc, float u1xr, float u1xd, float u1yc, float u1yr,
                         float u1yd, float as1, float as0) {
    const float gxv = ubr - ubc - u1xc;
    const float gyv = ubd - ubc - u1yc;
    px = __builtin_fmaf(as1, __builtin_fmaf(t11, gxv, t12 * gyv), px);
    py = __builtin_fmaf(as1, __builtin_fmaf(t12, gxv, t22 * gyv), py);
    const float dp = fmaxf(1.f, sqrtf(__builtin_fmaf(px, px, py * py)));
    px = px / dp;
    py = py / dp;
    const float dx_u1x = u1xr - u1xc;
    const float dy_u1x = u1xd - u1xc;
    const float dx_u1y = u1yr - u1yc;
    const float dy_u1y = u1yd - u1yc;
    qx = __builtin_fmaf(as0, dx_u1x, qx);
    qy = __builtin_fmaf(as0, dy_u1y, qy);
    qz = __builtin_fmaf(as0, 0.5f * (dy_u1x + dx_u1y), qz);
    const float dq = fmaxf(1.f, sqrtf(__builtin_fmaf(qz, qz, __builtin_fmaf(qz, qz, __builtin_fmaf(qy, qy, qx * qx)))));
    qx = qx / dq;
    qy = qy / dq;
    qz = qz / dq;
    return dp > 1.f || dq > 1.f;
}

__device__ __forceinline__ bool same(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }

// counts: 0 root mismatches, 1 quotient mismatches, 2 quotients on the fast
// path, 3 dual-update and fusion-projection mismatches, 4 dual updates that
// projected
__global__ void check_kernel(uint64_t n, uint64_t seed, unsigned long long* counts) {
    unsigned long long bad_root = 0, bad_div = 0, fast = 0, bad_pq = 0, proj = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t h0 = mix64(seed * 0x9e3779b97f4a7c15ULL + i);
        const uint64_t h1 = mix64(h0), h2 = mix64(h1), h3 = mix64(h2);
        // root: s > 1 in binade 0..127, plus +inf and the smallest s above 1
        float s = with_exp(h0, (int)(h1 % 128));
        if ((h2 & 1023) == 0) s = __int_as_float(0x7f800000);
        if ((h2 & 1023) == 1) s = __uint_as_float(0x3f800001);
        if (!(s > 1.f)) s = 1.5f;
        const float d = psw::sqrt_rn_gt1(s);
        bad_root += !same(d, sqrtf(s));
        // quotient x / sqrt(s): |x| <= sqrt(s) from any binade down to denormals
        if (d < __int_as_float(0x7f800000)) {
            const int ed = (int)((__float_as_uint(d) >> 23) & 255) - 127;
            const int ex = ed - (int)(h3 % 280);
            float x;
            if (ex >= -126) x = with_exp(h2 >> 8, ex);
            else if (ex >= -149) x = __uint_as_float((uint32_t)(h2 >> 8) & ((1u << (ex + 150)) - 1));   // denormal
            else x = 0.f;
            if (h3 & (1ULL << 40)) x = -x;
            if (psw::div_nr_ok(s, fabsf(x))) {
                ++fast;
                bad_div += !same(psw::div_nr(x, d, psw::rcp_nr(d)), x / d);
            }
        }
        // whole dual update at magnitudes where the projections trigger
        const float sc = (float)(1 + (h3 & 7)) * 0.15f;
        auto r = [&](uint64_t h, int k) { return ((float)((h >> (k * 8)) & 255) - 127.5f) / 64.f * sc; };
        float px = r(h0, 0), py = r(h0, 1), qx = r(h0, 2), qy = r(h0, 3), qz = r(h0, 4);
        float ax = px, ay = py, bx = qx, by = qy, bz = qz;
        const float t11 = 0.5f + r(h1, 0) * 0.1f, t12 = r(h1, 1) * 0.1f, t22 = 0.5f + r(h1, 2) * 0.1f;
        const float ubc = r(h1, 3), ubr = r(h1, 4), ubd = r(h1, 5), u1xc = r(h1, 6), u1xr = r(h1, 7);
        const float u1xd = r(h2, 0), u1yc = r(h2, 1), u1yr = r(h2, 2), u1yd = r(h2, 3);
        const float as1 = 0.25f + (float)(h2 >> 56) / 256.f, as0 = 0.25f + (float)((h3 >> 48) & 255) / 256.f;
        psw::pq_update(px, py, qx, qy, qz, t11, t12, t22, ubc, ubr, ubd, u1xc, u1xr, u1xd, u1yc, u1yr, u1yd, as1, as0);
        proj += pq_plain(ax, ay, bx, by, bz, t11, t12, t22, ubc, ubr, ubd, u1xc, u1xr, u1xd, u1yc, u1yr, u1yd, as1,
                         as0);
        bad_pq += !(same(px, ax) && same(py, ay) && same(qx, bx) && same(qy, by) && same(qz, bz));
        // the fusion dual step's projection (projectUnitBall, fusion.h:866-870)
        float fx = r(h3, 0), fy = r(h3, 1), fz = r(h3, 2);
        const float s3 = fx * fx + fy * fy + fz * fz, d3 = fmaxf(1.f, sqrtf(s3));
        const float gx = fx / d3, gy = fy / d3, gz = fz / d3;
        psw::project_ball3(s3, fx, fy, fz);
        bad_pq += !(same(fx, gx) && same(fy, gy) && same(fz, gz));
    }
    atomicAdd(&counts[0], bad_root);
    atomicAdd(&counts[1], bad_div);
    atomicAdd(&counts[2], fast);
    atomicAdd(&counts[3], bad_pq);
    atomicAdd(&counts[4], proj);
}

}  // namespace

extern "C" int point_check(uint64_t n, uint64_t seed, unsigned long long* out) {
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, 5 * sizeof(unsigned long long)) != hipSuccess) return -1;
    int rc = 0;
    if (hipMemset(d, 0, 5 * sizeof(unsigned long long)) != hipSuccess) rc = -2;
    if (rc == 0) {
        check_kernel<<<2048, 256>>>(n, seed, d);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -3;
    }
    if (rc == 0 && hipMemcpy(out, d, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess) rc = -4;
    (void)hipFree(d);
    return rc;
}
