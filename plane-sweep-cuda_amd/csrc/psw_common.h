// psw_common.h -- shared helpers of libpsw: status/error capture, launch
// shapes, and float math with the reference's operation order.
//
// All HIP sources are compiled with -ffp-contract=off: every + - * / rounds
// on its own (no fused multiply-add), which is what makes the reference-order
// kernels bit-identical to the reference's float arithmetic.  Kernels that
// want FMA (the fast sweep) opt in locally with #pragma clang fp contract.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>

#include "../../include/psw.h"

namespace psw {

// Thread-local message of the last failure (psw_last_error).
void set_error(const char* fmt, ...);
const char* get_error();

inline hipStream_t S(psw_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Map a HIP error to a status, recording its text.
int hip_status(hipError_t e, const char* what, int code = PSW_ERR_DEVICE);

// Check the most recent launch (hipGetLastError), as the reference's
// cudaPeekAtLastError check (src/planesweep.cpp:134).
int launch_status(const char* kernel);

// Stream-ordered workspace of one call (freed on the same stream when the
// call's kernels are queued).  It comes from a memory pool of the library's
// own per device that keeps freed blocks for later calls: the default pool
// hands them back to the device at every synchronisation, and the next call
// then maps fresh memory, which costs more than the kernels of a small call.
int workspace_alloc_async(void** p, size_t bytes, hipStream_t st);
int workspace_free_async(void* p, hipStream_t st);

#define PSW_CHECK_ARG(cond, msg)                                         \
    do {                                                                 \
        if (!(cond)) {                                                   \
            psw::set_error("%s: %s", __func__, msg);                     \
            return PSW_ERR_ARG;                                          \
        }                                                                \
    } while (0)

#define PSW_TRY_HIP(expr, code)                                          \
    do {                                                                 \
        hipError_t e_ = (expr);                                          \
        if (e_ != hipSuccess) return psw::hip_status(e_, #expr, code);   \
    } while (0)

// Launch shape of a reference launcher: caller's dim3 if non-zero, else a
// 64x4 block (one wave per image row segment) covering width x height.
struct Shape {
    dim3 grid, block;
};
int make_shape(psw_dim3 blocks, psw_dim3 threads, int width, int height, Shape& out);

// Row-major 3x3 float matrix passed by value to kernels (Matrix3D).
struct M3f {
    float m[9];
};
inline M3f m3_from(const float* p) {
    M3f r;
    for (int i = 0; i < 9; ++i) r.m[i] = p[i];
    return r;
}
struct V3f {
    float x, y, z;
};
inline V3f v3_from(const float* p) { return V3f{p[0], p[1], p[2]}; }

// Reference dot product: a.x*b.x + a.y*b.y + a.z*b.z, left to right.
__host__ __device__ inline float dot3(float a0, float a1, float a2, float b0, float b1, float b2) {
    return a0 * b0 + a1 * b1 + a2 * b2;
}
// Matrix3D * float3 (helper_structs.h:130-133).
__host__ __device__ inline void mv3(const M3f& M, float x, float y, float z, float& ox, float& oy, float& oz) {
    ox = dot3(M.m[0], M.m[1], M.m[2], x, y, z);
    oy = dot3(M.m[3], M.m[4], M.m[5], x, y, z);
    oz = dot3(M.m[6], M.m[7], M.m[8], x, y, z);
}
// Matrix3D * Matrix3D (helper_structs.h:136-144): C[i][j] = row_i(A) . col_j(B).
__host__ __device__ inline M3f mm3(const M3f& A, const M3f& B) {
    M3f C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C.m[3 * i + j] = dot3(A.m[3 * i], A.m[3 * i + 1], A.m[3 * i + 2], B.m[j], B.m[3 + j], B.m[6 + j]);
    return C;
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8).  Map
// block b to work item t so that each XCD gets a contiguous range of items:
// with items in raster order, neighbouring tiles then share an XCD and the
// halo cache lines one tile reads were usually just fetched by its neighbour
// into the same L2.  A bijection of [0, nitems).
__device__ inline int xcd_tile(int b, int nitems) {
    const int per = (nitems + 7) / 8;
    const int xcd = b & 7, k = b >> 3;
    const int full = nitems - (per - 1) * 8;  // XCDs that get 'per' items
    if (xcd < full) return xcd * per + k;
    return full * per + (xcd - full) * (per - 1) + k;
}

// Reflect-101 border of the reference's box filters (kernels.cu:203-204).
__device__ inline int reflect101(int k, int n) {
    if (k < 0) k = -k;
    if (k > n - 1) k = 2 * (n - 1) - k;
    return k;
}

// Bilinear sample with the reference's out-of-image rule (kernels.cu:15-27):
// the sample is 0 unless 0 <= floor(x) <= M1-2 and 0 <= floor(y) <= M2-2,
// i.e. unless 0 <= x < M1-1 and 0 <= y < M2-1.  Testing the float range
// first also sends NaN and huge coordinates to 0 without an int conversion.
__device__ inline float bilinear_sample(const float* __restrict__ d, int M1, int M2, float x, float y) {
    if (!(x >= 0.f && x < (float)(M1 - 1) && y >= 0.f && y < (float)(M2 - 1))) return 0.f;
    // 0 <= x < M1 - 1: floorf(x) is an exact small integer, so x - floorf(x)
    // equals x - (float)(int)floorf(x); the offset is below M1 * M2, which the
    // entry points keep under 2^32 (32-bit index arithmetic, one 64-bit add).
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const uint32_t o = (uint32_t)(int)fy * (uint32_t)M1 + (uint32_t)(int)fx;
    const float* r0 = d + o;
    const float* r1 = r0 + M1;
    const float t1 = a * r0[1] + (1.f - a) * r0[0];
    const float t2 = a * r1[1] + (1.f - a) * r1[0];
    return b * t2 + (1.f - b) * t1;
}

// The same sample with each interpolation as one fused multiply-add,
// v0 + a (v1 - v0): two instructions per interpolation instead of four, and
// one rounding of the product-sum instead of three (the fast sweep).  Past
// the bounds test x, y >= 0, so the hardware fraction (x - floor(x), exact
// here) and truncation give the same a, b and cell as floorf above, in two
// instructions per coordinate instead of three.
// The cell's byte offset is formed in 32 bits (a 24-bit multiply: rows and
// widths stay below 2^24, and the entry points keep images below 2^30
// pixels), so both rows load from a uniform base plus that offset.
__device__ inline float bilinear_sample_fma(const float* __restrict__ d, int M1, int M2, float x, float y) {
    if (!(x >= 0.f && x < (float)(M1 - 1) && y >= 0.f && y < (float)(M2 - 1))) return 0.f;
    const float a = __builtin_amdgcn_fractf(x), b = __builtin_amdgcn_fractf(y);
    const uint32_t ob = __umul24((uint32_t)(int)y, 4u * (uint32_t)M1) + 4u * (uint32_t)(int)x;
    const float* r0 = reinterpret_cast<const float*>(reinterpret_cast<const char*>(d) + ob);
    const float* r1 = reinterpret_cast<const float*>(reinterpret_cast<const char*>(d + M1) + ob);
    const float t1 = __builtin_fmaf(a, r0[1] - r0[0], r0[0]);
    const float t2 = __builtin_fmaf(a, r1[1] - r1[0], r1[0]);
    return __builtin_fmaf(b, t2 - t1, t1);
}

}  // namespace psw
