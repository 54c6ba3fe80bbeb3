// sweep_fused.hip -- fused plane sweep for gfx950.
//
// One workgroup owns a 32x16 tile of reference pixels and sweeps every
// (plane, source) pair over it without leaving the CU:
//   warp   : homography + bilinear sample of the source on the tile plus a
//            window-radius halo, into LDS (reflect-101 halo positions, as the
//            reference's box filters, kernels.cu:203-204 / :228-229);
//   column : per column, register-blocked sums of I, I^2 and Ref*I over the
//            window (4 output rows per thread), into LDS;
//   row    : per pixel pair, sums of the column means -> mean, E[I^2],
//            E[Ref*I] -> std, NCC (kernels.cu:46-64, :103-120), reduction.
// The reference instead writes eleven full-frame images to HBM per
// (plane, source) (src/planesweep.cpp:206-249).
//
// Two reductions share the per-(pixel, plane, source) NCC:
//   reference : per source, best NCC over planes (strict >, zero start) and
//               its depth; depth averaged over sources whose best NCC >
//               nccthresh; zfar where none (src/planesweep.cpp:125-131,
//               kernels.cu:66-101, :145-157, :331-340);
//   aggregate : cost(p) = -(sum_s NCC_s) / N; argmin (first), minimum and
//               second-smallest cost.
// With FAST=false every sum, product and quotient is taken in the reference's
// order (column mean = sum/win, row mean of column means = sum/win), so the
// NCC is bit-identical to the reference's float arithmetic.  FAST=true
// multiplies by 1/win, keeps running window sums down each column segment
// (TY/4 rows) and along each row's pixels, and uses explicit fused
// multiply-adds and the hardware reciprocal: its results are the same in
// every launch of one tile height (single frame, batched, any plane groups),
// and differ in the last bits between tile heights.
#include "psw_common.h"

#include <mutex>
#include <vector>

namespace psw {

// Tile: SW_TX x TY pixels, TY = 16 or 32 (a template parameter; the
// aggregated sweep's launches choose it, the reference reduction uses 16).
// A taller tile warps fewer halo positions per pixel (the window radius is
// added on all four sides) and gives the row pass more pixels per thread.
constexpr int SW_TX = 32;       // tile width  (pixels)
constexpr int SW_TY = 16;       // default tile height (pixels)
constexpr int SW_THREADS = 256; // 4 waves
constexpr int SW_NCG = 4;       // column-pass row groups: TY / 4 rows per thread

struct SweepArgs {
    const float* ref;
    const float* src;     // [N][H][W]
    const float* Hs;      // [N][P][9]
    const float* depths;  // [P] (reference mode)
    int W, H, N, P;
    int off;              // pixel-coordinate offset of the homography (1 = reference)
    int tiles_x, ntiles;
    int ty;               // tile height (16 or 32)
    int G;                // plane groups (aggregate mode): workgroups per tile
    float stdthresh, nccthresh, zfar;
    float* depth_out;     // reference mode
    int32_t* argmin;      // aggregate mode: written directly when G == 1 ...
    float* cmin;
    float* csec;
    unsigned long long* top2;  // ... else every group's best two go to these per-pixel keys [2][H*W]
};

// The aggregated sweep's plane groups of one tile run in different
// workgroups.  Each folds its best (cost, plane) and its second-best cost
// into two 64-bit words per pixel with atomic minima: the key is the cost's
// order-preserving bits in the high half and the plane in the low half, so
// the smallest key is the smallest cost at its first plane, which is what the
// sequential strict-< scan over all planes picks.  -0 is folded into +0 so
// that equal costs are equal keys (the output then reads +0 where the scan
// could read -0: the same value).
__device__ __forceinline__ unsigned long long sweep_key(float c, int plane) {
    const unsigned u = __float_as_uint(c + 0.0f);
    const unsigned o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((unsigned long long)o << 32) | (unsigned)plane;
}

__device__ __forceinline__ float sweep_key_cost(unsigned long long k) {
    const unsigned o = (unsigned)(k >> 32);
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// One group's best (c1 at plane p1) and second-best c2 of pixel i into the
// keys: the first word keeps the smallest key of all groups; every other
// key offered ends in the second word (a key displaced from the first word
// is offered there), so the second word keeps the second-smallest key of the
// whole plane loop.  The two keys a group offers the second word (the
// larger of its best and the displaced one, and its second-best) go as
// their minimum in one atomic.  Costs that are not below +inf count as
// absent, as in the scan.  The atomics go past the L2 (the write counters
// read 16 bytes per atomic), so each one is a memory transaction: two per
// pixel and group.
__device__ __forceinline__ void sweep_top2_insert(unsigned long long* top2, size_t npix, size_t i, float c1, int p1,
                                                  float c2) {
    const float inf = __builtin_inff();
    if (!(c1 < inf)) return;
    const unsigned long long k1 = sweep_key(c1, p1);
    const unsigned long long old = atomicMin(&top2[i], k1);
    const unsigned long long hi = k1 < old ? old : k1;
    const unsigned long long k2 = c2 < inf ? sweep_key(c2, p1) : ~0ull;
    const unsigned long long offer = k2 < hi ? k2 : hi;
    if (offer != ~0ull) atomicMin(&top2[npix + i], offer);
}

template <bool FAST>
__device__ inline float scale_mean(float s, float win, float rwin) {
    if constexpr (FAST) return s * rwin;
    else return s / win;
}

// One work item: plane group g of tile t (item = t * G + g).  Called by all
// threads of a workgroup with the same item (it contains barriers).
template <int R, int MODE, bool FAST, int TY = SW_TY>
__device__ __forceinline__ void sweep_item(const SweepArgs& a, const int item) {
    constexpr int TW = SW_TX + 2 * R;
    constexpr int TH = TY + 2 * R;
    constexpr int WIN = 2 * R + 1;
    constexpr int NCG = SW_NCG;                 // column-pass row groups
    constexpr int SW_CR = TY / SW_NCG;          // rows per thread in the column pass
    constexpr int SW_PX = SW_TX * TY / SW_THREADS;  // pixels per thread in the row pass
    constexpr int NCOL = NCG * TW;              // column-pass work items
    constexpr int CV = SW_CR + 2 * R;           // column values per column item
    constexpr int RV = SW_PX + 2 * R;           // row values per row item
    static_assert(NCOL <= SW_THREADS, "window too large for the tile");
    static_assert(TY % SW_NCG == 0 && (SW_TX * TY) % SW_THREADS == 0, "tile height");
    __shared__ float sR[TH * TW];
    __shared__ __align__(16) float sW[TH * TW];
    __shared__ float sC[3][TY * TW];

    const int t = item / a.G, g = item - (item / a.G) * a.G;
    const int p0 = (g * a.P) / a.G, p1 = ((g + 1) * a.P) / a.G;
    const int x0 = (t % a.tiles_x) * SW_TX;
    const int y0 = (t / a.tiles_x) * TY;
    const int tid = threadIdx.x;
    const int W = a.W, H = a.H;
    const float win = (float)WIN, rwin = 1.f / (float)WIN;

    // Actual (reflected) image coordinate of each halo-tile position this
    // thread warps; positions past W-1+R / H-1+R only feed discarded pixels.
    // The reference-order sweep takes positions i = tid + k * threads; the
    // fast one takes pairs (2q, 2q+1) of one row (TW is even), which share
    // the homography's row terms.
    constexpr int NWARP = (TH * TW + SW_THREADS - 1) / SW_THREADS;
    constexpr int NPAIR = (TH * TW / 2 + SW_THREADS - 1) / SW_THREADS;
    static_assert(TW % 2 == 0, "pairs of one row");
    constexpr int NPOS = FAST ? 2 * NPAIR : NWARP;
    int wgx[NPOS], wgy[NPOS];
#pragma unroll
    for (int k = 0; k < NPOS; ++k) {
        const int i = FAST ? 2 * (tid + (k / 2) * SW_THREADS) + (k & 1) : tid + k * SW_THREADS;
        const int ty = i / TW, tx = i - (i / TW) * TW;
        wgx[k] = reflect101(min(x0 - R + tx, W - 1 + R), W);
        wgy[k] = reflect101(min(y0 - R + ty, H - 1 + R), H);
    }
#pragma unroll
    for (int k = 0; k < NPOS; ++k) {
        const int i = FAST ? 2 * (tid + (k / 2) * SW_THREADS) + (k & 1) : tid + k * SW_THREADS;
        if (i < TH * TW) sR[i] = a.ref[(size_t)wgy[k] * W + wgx[k]];
    }
    __syncthreads();

    // Column-pass item of this thread and its reference values (constant over
    // the whole sweep, so they live in registers).
    const bool col_active = tid < NCOL;
    const int ccx = tid % TW, cg = min(tid / TW, NCG - 1);
    float rv[CV];
#pragma unroll
    for (int i = 0; i < CV; ++i) rv[i] = sR[(cg * SW_CR + i) * TW + ccx];
    // Row-pass pixels of this thread: (ox, oy), (ox+1, oy).
    const int rox = SW_PX * (tid % (SW_TX / SW_PX));
    const int roy = tid / (SW_TX / SW_PX);

    // ---- reference window statistics (src/planesweep.cpp:108-119) ----
    if (col_active) {
#pragma unroll
        for (int j = 0; j < SW_CR; ++j) {
            float s1 = 0.f, s2 = 0.f;
#pragma unroll
            for (int i = 0; i < WIN; ++i) {
                s1 += rv[j + i];
                s2 += rv[j + i] * rv[j + i];
            }
            sC[0][(cg * SW_CR + j) * TW + ccx] = scale_mean<FAST>(s1, win, rwin);
            sC[1][(cg * SW_CR + j) * TW + ccx] = scale_mean<FAST>(s2, win, rwin);
        }
    }
    __syncthreads();
    float rmean[SW_PX], rstd[SW_PX];
    {
        float c0[RV], c1[RV];
#pragma unroll
        for (int i = 0; i < RV; ++i) {
            c0[i] = sC[0][roy * TW + rox + i];
            c1[i] = sC[1][roy * TW + rox + i];
        }
#pragma unroll
        for (int k = 0; k < SW_PX; ++k) {
            float m = 0.f, q = 0.f;
#pragma unroll
            for (int i = 0; i < WIN; ++i) {
                m += c0[k + i];
                q += c1[k + i];
            }
            m = scale_mean<FAST>(m, win, rwin);
            q = scale_mean<FAST>(q, win, rwin);
            const float var = q - m * m;
            rmean[k] = m;
            rstd[k] = var > 0.f ? sqrtf(var) : 0.f;
        }
    }
    __syncthreads();

    // One (plane, source) step: warp into sW, window sums, NCC of my pixels.
    auto plane_ncc = [&](const float* __restrict__ src, const float* __restrict__ Hm, float ncc[SW_PX]) {
        const float h0 = Hm[0], h1 = Hm[1], h2 = Hm[2], h3 = Hm[3], h4 = Hm[4], h5 = Hm[5], h6 = Hm[6],
                    h7 = Hm[7], h8 = Hm[8];
        const float offf = (float)a.off;
        if constexpr (FAST) {
#pragma unroll
            for (int k = 0; k < NPAIR; ++k) {
                const int q = tid + k * SW_THREADS;
                if (q < TH * TW / 2) {
                    // explicit fused multiply-adds (no contraction left to the
                    // compiler), so every kernel that inlines this computes
                    // the same bits; the row terms are the pair's
                    const float py = (float)(wgy[2 * k] + a.off);
                    const float cx = __builtin_fmaf(h1, py, h2), cy = __builtin_fmaf(h4, py, h5),
                                cz = __builtin_fmaf(h7, py, h8);
                    float out[2];
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const float px = (float)(wgx[2 * k + e] + a.off);
                        const float vx = __builtin_fmaf(h0, px, cx);
                        const float vy = __builtin_fmaf(h3, px, cy);
                        const float vz = __builtin_fmaf(h6, px, cz);
                        const float rz = __builtin_amdgcn_rcpf(vz);
                        out[e] = bilinear_sample_fma(src, W, H, __builtin_fmaf(vx, rz, -offf),
                                                     __builtin_fmaf(vy, rz, -offf));
                    }
                    *reinterpret_cast<float2*>(&sW[2 * q]) = make_float2(out[0], out[1]);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < NWARP; ++k) {
                const int i = tid + k * SW_THREADS;
                if (i < TH * TW) {
                    const float px = (float)(wgx[k] + a.off), py = (float)(wgy[k] + a.off);
                    const float vx = dot3(h0, h1, h2, px, py, 1.f);
                    const float vy = dot3(h3, h4, h5, px, py, 1.f);
                    const float vz = dot3(h6, h7, h8, px, py, 1.f);
                    sW[i] = bilinear_sample(src, W, H, vx / vz - offf, vy / vz - offf);
                }
            }
        }
        __syncthreads();
        if (col_active) {
            float wv[CV];
#pragma unroll
            for (int i = 0; i < CV; ++i) wv[i] = sW[(cg * SW_CR + i) * TW + ccx];
            if constexpr (FAST) {
                // running window sums down the column
                float s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
                for (int i = 0; i < WIN; ++i) {
                    s1 += wv[i];
                    s2 = __builtin_fmaf(wv[i], wv[i], s2);
                    s3 = __builtin_fmaf(rv[i], wv[i], s3);
                }
#pragma unroll
                for (int j = 0; j < SW_CR; ++j) {
                    if (j > 0) {
                        const float wa = wv[j + 2 * R], wd = wv[j - 1];
                        s1 += wa - wd;
                        s2 = __builtin_fmaf(wa, wa, __builtin_fmaf(-wd, wd, s2));
                        s3 = __builtin_fmaf(rv[j + 2 * R], wa, __builtin_fmaf(-rv[j - 1], wd, s3));
                    }
                    sC[0][(cg * SW_CR + j) * TW + ccx] = s1 * rwin;
                    sC[1][(cg * SW_CR + j) * TW + ccx] = s2 * rwin;
                    sC[2][(cg * SW_CR + j) * TW + ccx] = s3 * rwin;
                }
            } else {
#pragma unroll
                for (int j = 0; j < SW_CR; ++j) {
                    float s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
                    for (int i = 0; i < WIN; ++i) {
                        const float w = wv[j + i];
                        s1 += w;
                        s2 += w * w;
                        s3 += rv[j + i] * w;
                    }
                    sC[0][(cg * SW_CR + j) * TW + ccx] = s1 / win;
                    sC[1][(cg * SW_CR + j) * TW + ccx] = s2 / win;
                    sC[2][(cg * SW_CR + j) * TW + ccx] = s3 / win;
                }
            }
        }
        __syncthreads();
        float c0[RV], c1[RV], c2[RV];
#pragma unroll
        for (int i = 0; i < RV; ++i) {
            c0[i] = sC[0][roy * TW + rox + i];
            c1[i] = sC[1][roy * TW + rox + i];
            c2[i] = sC[2][roy * TW + rox + i];
        }
        if constexpr (FAST) {
            float m = 0.f, q = 0.f, pm = 0.f;
#pragma unroll
            for (int k = 0; k < SW_PX; ++k) {
                if (k == 0) {
#pragma unroll
                    for (int i = 0; i < WIN; ++i) {
                        m += c0[i];
                        q += c1[i];
                        pm += c2[i];
                    }
                } else {
                    m += c0[k + 2 * R] - c0[k - 1];
                    q += c1[k + 2 * R] - c1[k - 1];
                    pm += c2[k + 2 * R] - c2[k - 1];
                }
                const float mm = m * rwin, qq = q * rwin, pp = pm * rwin;
                const float var = __builtin_fmaf(-mm, mm, qq);
                const float sd = var > 0.f ? __builtin_amdgcn_sqrtf(var) : 0.f;
                if ((rstd[k] < a.stdthresh) || (sd < a.stdthresh)) ncc[k] = 0.f;
                else ncc[k] = __builtin_fmaf(-rmean[k], mm, pp) * __builtin_amdgcn_rcpf(rstd[k] * sd);
            }
        } else {
#pragma unroll
            for (int k = 0; k < SW_PX; ++k) {
                float m = 0.f, q = 0.f, pm = 0.f;
#pragma unroll
                for (int i = 0; i < WIN; ++i) {
                    m += c0[k + i];
                    q += c1[k + i];
                    pm += c2[k + i];
                }
                m = m / win;
                q = q / win;
                pm = pm / win;
                const float var = q - m * m;
                const float sd = var > 0.f ? sqrtf(var) : 0.f;
                if ((rstd[k] < a.stdthresh) || (sd < a.stdthresh)) ncc[k] = 0.f;
                else ncc[k] = (pm - rmean[k] * m) / (rstd[k] * sd);
            }
        }
        // No barrier here: the next warp writes sW (its readers passed the
        // barrier above); the next column pass writes sC only after the next
        // barrier, which every thread reaches after finishing this row pass.
    };

    const size_t npix = (size_t)W * H;
    if constexpr (MODE == 0) {
        float gsum[SW_PX], gcnt[SW_PX];
#pragma unroll
        for (int k = 0; k < SW_PX; ++k) { gsum[k] = 0.f; gcnt[k] = 0.f; }
        for (int s = 0; s < a.N; ++s) {
            const float* src = a.src + (size_t)s * npix;
            float best[SW_PX], bd[SW_PX];
#pragma unroll
            for (int k = 0; k < SW_PX; ++k) { best[k] = 0.f; bd[k] = 0.f; }
            for (int p = 0; p < a.P; ++p) {
                float ncc[SW_PX];
                plane_ncc(src, a.Hs + ((size_t)s * a.P + p) * 9, ncc);
                const float d = a.depths[p];
#pragma unroll
                for (int k = 0; k < SW_PX; ++k)
                    if (ncc[k] > best[k]) { best[k] = ncc[k]; bd[k] = d; }
            }
#pragma unroll
            for (int k = 0; k < SW_PX; ++k)
                if (best[k] > a.nccthresh) { gsum[k] += bd[k]; gcnt[k] += 1.f; }
        }
#pragma unroll
        for (int k = 0; k < SW_PX; ++k) {
            const int x = x0 + rox + k, y = y0 + roy;
            if (x < W && y < H) {
                float v = (gcnt[k] != 0.f) ? gsum[k] / gcnt[k] : __builtin_nanf("");
                if (v != v) v = a.zfar;
                a.depth_out[(size_t)y * W + x] = v;
            }
        }
    } else {
        float cmin[SW_PX], csec[SW_PX];
        int arg[SW_PX];
        const float inf = __builtin_inff();
#pragma unroll
        for (int k = 0; k < SW_PX; ++k) { cmin[k] = inf; csec[k] = inf; arg[k] = -1; }
        const float fN = (float)a.N;
        for (int p = p0; p < p1; ++p) {
            float acc[SW_PX];
#pragma unroll
            for (int k = 0; k < SW_PX; ++k) acc[k] = 0.f;
            for (int s = 0; s < a.N; ++s) {
                float ncc[SW_PX];
                plane_ncc(a.src + (size_t)s * npix, a.Hs + ((size_t)s * a.P + p) * 9, ncc);
#pragma unroll
                for (int k = 0; k < SW_PX; ++k) acc[k] += ncc[k];
            }
#pragma unroll
            for (int k = 0; k < SW_PX; ++k) {
                const float c = -acc[k] / fN;
                if (c < cmin[k]) { csec[k] = cmin[k]; cmin[k] = c; arg[k] = p; }
                else if (c < csec[k]) csec[k] = c;
            }
        }
#pragma unroll
        for (int k = 0; k < SW_PX; ++k) {
            const int x = x0 + rox + k, y = y0 + roy;
            if (x < W && y < H) {
                const size_t i = (size_t)y * W + x;
                if (a.G > 1) {
                    sweep_top2_insert(a.top2, npix, i, cmin[k], arg[k], csec[k]);
                } else {
                    a.argmin[i] = arg[k];
                    a.cmin[i] = cmin[k];
                    a.csec[i] = csec[k];
                }
            }
        }
    }
}

// The fast variant is held to 72 VGPRs (7 waves per SIMD instead of 6;
// measured 1.73 against 1.76 ms at 640x480 x 8 x 256, 9x9: two registers
// spill; at 8 waves sixteen do and it is slower).
template <int R, int MODE, bool FAST, int TY>
__global__ __launch_bounds__(SW_THREADS) __attribute__((amdgpu_waves_per_eu(FAST ? 7 : 1))) void sweep_fused_kernel(SweepArgs a) {
    if ((int)blockIdx.x >= a.ntiles * a.G) return;
    sweep_item<R, MODE, FAST, TY>(a, xcd_tile(blockIdx.x, a.ntiles * a.G));
}

// Many frames (uneven sizes, plane and source counts) in one persistent
// launch: workgroups take items from a device counter until all frames'
// items are done.  Item g belongs to the frame f with off[f] <= g < off[f+1].
// Every item is computed exactly as in sweep_fused_kernel.
template <int R, int MODE, bool FAST, int TY>
__global__ __launch_bounds__(SW_THREADS) void sweep_batch_kernel(const SweepArgs* __restrict__ frames,
                                                                 const int* __restrict__ off, int nframes, int total,
                                                                 unsigned* __restrict__ counter) {
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int gi = s_item;
        __syncthreads();  // all threads hold gi before thread 0 takes the next one
        if (gi >= total) break;
        int lo = 0, hi = nframes - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= gi) lo = mid;
            else hi = mid - 1;
        }
        const SweepArgs a = frames[lo];
        sweep_item<R, MODE, FAST, TY>(a, gi - off[lo]);
    }
}

// The outputs from the per-pixel keys of the plane groups (all-ones: no
// finite cost, as the scan's -1 / +inf / +inf).
__global__ void sweep_top2_kernel(const unsigned long long* __restrict__ top2, size_t npix, int32_t* __restrict__ arg,
                                  float* __restrict__ cmin, float* __restrict__ csec) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const unsigned long long k0 = top2[i], k1 = top2[npix + i];
    const float inf = __builtin_inff();
    arg[i] = k0 == ~0ull ? -1 : (int32_t)(unsigned)k0;
    cmin[i] = k0 == ~0ull ? inf : sweep_key_cost(k0);
    csec[i] = k1 == ~0ull ? inf : sweep_key_cost(k1);
}

// ---- FP64 homographies -----------------------------------------------------

struct HomArgs {
    double K[9], Kinv[9];
    double Rrel[PSW_MAX_SOURCES][9];
    double trel[PSW_MAX_SOURCES][3];
    int N, P;
    const float* depths;
    float* H;
};

__device__ __host__ inline double d_dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
    return a0 * b0 + a1 * b1 + a2 * b2;
}
__device__ __host__ inline void d_mm3(const double* A, const double* B, double* C) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[3 * i + j] = d_dot3(A[3 * i], A[3 * i + 1], A[3 * i + 2], B[j], B[3 + j], B[6 + j]);
}

// H = K (Rrel + trel e3^T / d) K^-1, then H / H(2,2) (src/planesweep.cpp:196-203),
// in double; one thread per (source, plane).
__global__ void sweep_homography_kernel(HomArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N * a.P) return;
    const int s = i / a.P, p = i - (i / a.P) * a.P;
    const double d = (double)a.depths[p];
    double M[9];
    for (int j = 0; j < 9; ++j) M[j] = a.Rrel[s][j];
    // trel e3^T has trel in its third column; the rest is 0 / d = 0.
    M[2] = M[2] + a.trel[s][0] / d;
    M[5] = M[5] + a.trel[s][1] / d;
    M[8] = M[8] + a.trel[s][2] / d;
    M[0] = M[0] + 0.0 / d; M[1] = M[1] + 0.0 / d;
    M[3] = M[3] + 0.0 / d; M[4] = M[4] + 0.0 / d;
    M[6] = M[6] + 0.0 / d; M[7] = M[7] + 0.0 / d;
    double KM[9], Hm[9];
    d_mm3(a.K, M, KM);
    d_mm3(KM, a.Kinv, Hm);
    const double h22 = Hm[8];
    float* out = a.H + (size_t)i * 9;
    for (int j = 0; j < 9; ++j) out[j] = (float)(Hm[j] / h22);
}

// Matrix3D::inv in double (src/inc/structs.h:565-593 order).
static void d_inv3(const double* m, double* o) {
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    const double r[9] = {m[4] * m[8] - m[5] * m[7], m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
                         m[5] * m[6] - m[3] * m[8], m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
                         m[3] * m[7] - m[4] * m[6], m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3]};
    for (int j = 0; j < 9; ++j) o[j] = r[j] / det;
}
static void d_trans3(const double* m, double* o) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o[3 * j + i] = m[3 * i + j];
}
static void d_mv3(const double* M, const double* v, double* o) {
    for (int i = 0; i < 3; ++i) o[i] = d_dot3(M[3 * i], M[3 * i + 1], M[3 * i + 2], v[0], v[1], v[2]);
}

// PlaneSweep::RelativeMatrices (src/planesweep.cpp:538-549) in double.
static void d_relative(const float* Rref, const float* tref, const float* Rsrc, const float* tsrc, bool alt,
                       double* Rrel, double* trel) {
    double Rr[9], Rs[9], tr[3], ts[3];
    for (int j = 0; j < 9; ++j) { Rr[j] = Rref[j]; Rs[j] = Rsrc[j]; }
    for (int j = 0; j < 3; ++j) { tr[j] = tref[j]; ts[j] = tsrc[j]; }
    if (!alt) {
        double Ri[9], Rt[3];
        d_inv3(Rr, Ri);
        d_mm3(Rs, Ri, Rrel);
        d_mv3(Rrel, tr, Rt);
        for (int j = 0; j < 3; ++j) trel[j] = ts[j] - Rt[j];
    } else {
        double RsT[9], dt[3];
        d_trans3(Rs, RsT);
        d_mm3(RsT, Rr, Rrel);
        for (int j = 0; j < 3; ++j) dt[j] = tr[j] - ts[j];
        d_mv3(RsT, dt, trel);
    }
}

template <int R>
static int launch_sweep_r(const SweepArgs& a, int mode, bool fast, hipStream_t st) {
    const dim3 grid((unsigned)(a.ntiles * a.G)), block(SW_THREADS);
    if (mode == 0) {
        if (fast) sweep_fused_kernel<R, 0, true, 16><<<grid, block, 0, st>>>(a);
        else sweep_fused_kernel<R, 0, false, 16><<<grid, block, 0, st>>>(a);
    } else if (a.ty == 32) {
        if (fast) sweep_fused_kernel<R, 1, true, 32><<<grid, block, 0, st>>>(a);
        else sweep_fused_kernel<R, 1, false, 32><<<grid, block, 0, st>>>(a);
    } else {
        if (fast) sweep_fused_kernel<R, 1, true, 16><<<grid, block, 0, st>>>(a);
        else sweep_fused_kernel<R, 1, false, 16><<<grid, block, 0, st>>>(a);
    }
    return launch_status("sweep_fused_kernel");
}

static int launch_sweep(SweepArgs a, int winsize, int mode, bool fast, hipStream_t st) {
    switch (winsize / 2) {
        case 0: return launch_sweep_r<0>(a, mode, fast, st);
        case 1: return launch_sweep_r<1>(a, mode, fast, st);
        case 2: return launch_sweep_r<2>(a, mode, fast, st);
        case 3: return launch_sweep_r<3>(a, mode, fast, st);
        case 4: return launch_sweep_r<4>(a, mode, fast, st);
        case 5: return launch_sweep_r<5>(a, mode, fast, st);
        case 6: return launch_sweep_r<6>(a, mode, fast, st);
        case 7: return launch_sweep_r<7>(a, mode, fast, st);
        default:
            set_error("window size %d not supported (odd, 1..15)", winsize);
            return PSW_ERR_UNSUPPORTED;
    }
}

// Tiles of the image, and plane groups per tile so that a launch has enough
// workgroups to keep 256 CUs busy to the end (about 14400 at 32 x 16 tiles:
// 24 groups of the 600 tiles of a 640 x 480 frame, measured the fastest of
// 1..64 groups).  PSW_SWEEP_GROUPS overrides the choice.
constexpr long SW_TARGET_ITEMS = 14400;
// The aggregated sweep's tile height: PSW_SWEEP_TY (16 or 32) or 16.
static int sweep_tile_height() {
    const char* e = getenv("PSW_SWEEP_TY");
    return (e && atoi(e) == 32) ? 32 : 16;
}

static void sweep_layout(SweepArgs& a, bool allow_groups, int ty = SW_TY) {
    a.ty = ty;
    a.tiles_x = (a.W + SW_TX - 1) / SW_TX;
    const int tiles_y = (a.H + a.ty - 1) / a.ty;
    a.ntiles = a.tiles_x * tiles_y;
    a.G = 1;
    if (!allow_groups) return;
    int G = (int)((SW_TARGET_ITEMS * 16 / a.ty + a.ntiles / 2) / a.ntiles);
    if (const char* e = getenv("PSW_SWEEP_GROUPS")) G = atoi(e);
    a.G = std::max(1, std::min(G, std::min(a.P, 64)));
}

static int check_sweep_args(const float* d_ref, const float* d_src, int width, int height, int nsrc,
                            const float* d_H, int nplanes, int winsize) {
    if (!d_ref || !d_src || !d_H) { set_error("NULL pointer"); return PSW_ERR_ARG; }
    if (nsrc < 1 || nplanes < 1) { set_error("need at least one source and one plane"); return PSW_ERR_ARG; }
    if (winsize < 1 || winsize % 2 == 0) {  // PlaneSweep::setWindowSize, planesweep.h:234
        set_error("Window size must be an odd number (got %d)", winsize);
        return PSW_ERR_ARG;
    }
    if (width < 2 || height < 2 || width <= winsize / 2 || height <= winsize / 2) {
        set_error("image %d x %d too small for window %d", width, height, winsize);
        return PSW_ERR_ARG;
    }
    if ((uint64_t)width * (uint64_t)height >= (1ull << 32)) {
        set_error("image %d x %d above 2^32 pixels", width, height);
        return PSW_ERR_ARG;
    }
    return PSW_OK;
}

// Device memory of psw_sweep_fused_aggregate_noalloc: the plane groups'
// per-pixel keys (16 bytes per pixel at any group count) live in this static
// array, one copy per GPU for the life of the process, instead of a per-call
// allocation.  Host code never uses the name itself (on the host it is a
// stand-in, not a device address): scratch_address() looks up this GPU's
// copy once with hipGetSymbolAddress and keeps it per device.
static __device__ __align__(256) unsigned char scratch[PSW_SWEEP_SCRATCH_BYTES];

constexpr int SW_MAX_DEVICES = 64;

static int scratch_address(void** out) {
    static std::mutex mu;
    static void* addr[SW_MAX_DEVICES] = {};
    int dev = 0;
    PSW_TRY_HIP(hipGetDevice(&dev), PSW_ERR_DEVICE);
    PSW_CHECK_ARG(dev >= 0 && dev < SW_MAX_DEVICES, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_status(e, "hipGetSymbolAddress(scratch)", PSW_ERR_DEVICE);
        addr[dev] = p;
    }
    *out = addr[dev];
    return PSW_OK;
}

// The aggregated sweep's launch and, with plane groups, the per-pixel keys
// in `ws` (SW_KEY_BYTES per pixel: cleared first, turned into the outputs
// after).
constexpr size_t SW_KEY_BYTES = 2 * sizeof(unsigned long long);
static int sweep_aggregate_run(SweepArgs& a, int winsize, int fast, void* ws, hipStream_t st) {
    const size_t npix = (size_t)a.W * a.H;
    int rc = PSW_OK;
    if (a.G > 1) {
        a.top2 = static_cast<unsigned long long*>(ws);
        rc = psw_fill_u32(reinterpret_cast<uint32_t*>(ws), 0xFFFFFFFFu, 4 * npix, st);
    }
    if (rc == PSW_OK) rc = launch_sweep(a, winsize, 1, fast != 0, st);
    if (rc == PSW_OK && a.G > 1) {
        sweep_top2_kernel<<<(unsigned)((npix + 255) / 256), 256, 0, st>>>(a.top2, npix, a.argmin, a.cmin, a.csec);
        rc = launch_status("sweep_top2_kernel");
    }
    return rc;
}

__global__ void argmin_depth_kernel(float* out, const int32_t* arg, const float* depths, int n, float fallback) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int a = arg[i];
        out[i] = a >= 0 ? depths[a] : fallback;
    }
}

}  // namespace psw

using namespace psw;

extern "C" {

int psw_sweep_homographies(float* d_H, const float* d_depths, int nplanes, int nsrc, const float* K,
                           const float* Rref, const float* tref, const float* Rsrc, const float* tsrc,
                           int alternative, psw_stream_t s) {
    PSW_CHECK_ARG(d_H && d_depths && K && Rref && tref && Rsrc && tsrc, "NULL pointer");
    PSW_CHECK_ARG(nsrc >= 1 && nsrc <= PSW_MAX_SOURCES, "nsrc must be 1..PSW_MAX_SOURCES");
    PSW_CHECK_ARG(nplanes >= 1, "nplanes must be positive");
    HomArgs a;
    for (int j = 0; j < 9; ++j) a.K[j] = K[j];
    d_inv3(a.K, a.Kinv);
    for (int i = 0; i < nsrc; ++i) d_relative(Rref, tref, Rsrc + 9 * i, tsrc + 3 * i, alternative != 0, a.Rrel[i], a.trel[i]);
    a.N = nsrc;
    a.P = nplanes;
    a.depths = d_depths;
    a.H = d_H;
    const int n = nsrc * nplanes;
    sweep_homography_kernel<<<(n + 63) / 64, 64, 0, S(s)>>>(a);
    return launch_status("sweep_homography_kernel");
}

int psw_sweep_fused_reference(const float* d_ref, const float* d_src, int width, int height, int nsrc,
                              const float* d_H, const float* d_depths, int nplanes, int winsize, float stdthresh,
                              float nccthresh, float zfar, float* d_depth_out, psw_stream_t s) {
    int st = check_sweep_args(d_ref, d_src, width, height, nsrc, d_H, nplanes, winsize);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(d_depths && d_depth_out, "NULL pointer");
    SweepArgs a = {};
    a.ref = d_ref; a.src = d_src; a.Hs = d_H; a.depths = d_depths;
    a.W = width; a.H = height; a.N = nsrc; a.P = nplanes; a.off = 1;
    a.stdthresh = stdthresh; a.nccthresh = nccthresh; a.zfar = zfar;
    a.depth_out = d_depth_out;
    sweep_layout(a, false);
    return launch_sweep(a, winsize, 0, false, S(s));
}

int psw_sweep_fused_aggregate(const float* d_ref, const float* d_src, int width, int height, int nsrc,
                              const float* d_H, int nplanes, int winsize, float stdthresh, int coord_offset, int fast,
                              int32_t* d_argmin, float* d_cost_min, float* d_cost_second, psw_stream_t s) {
    int st = check_sweep_args(d_ref, d_src, width, height, nsrc, d_H, nplanes, winsize);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(d_argmin && d_cost_min && d_cost_second, "NULL pointer");
    PSW_CHECK_ARG(coord_offset == 0 || coord_offset == 1, "coord_offset must be 0 or 1");
    SweepArgs a = {};
    a.ref = d_ref; a.src = d_src; a.Hs = d_H;
    a.W = width; a.H = height; a.N = nsrc; a.P = nplanes; a.off = coord_offset;
    a.stdthresh = stdthresh;
    a.argmin = d_argmin; a.cmin = d_cost_min; a.csec = d_cost_second;
    sweep_layout(a, true, sweep_tile_height());
    const size_t npix = (size_t)width * height;
    void* ws = nullptr;
    if (a.G > 1) {   // the plane groups' keys, from the library's workspace pool
        st = workspace_alloc_async(&ws, npix * SW_KEY_BYTES, S(s));
        if (st != PSW_OK) return st;
    }
    st = sweep_aggregate_run(a, winsize, fast, ws, S(s));
    if (ws) (void)workspace_free_async(ws, S(s));
    return st;
}

int psw_sweep_fused_aggregate_noalloc(const float* d_ref, const float* d_src, int width, int height, int nsrc,
                                      const float* d_H, int nplanes, int winsize, float stdthresh, int coord_offset,
                                      int fast, int groups, int32_t* d_argmin, float* d_cost_min,
                                      float* d_cost_second, psw_stream_t s) {
    int st = check_sweep_args(d_ref, d_src, width, height, nsrc, d_H, nplanes, winsize);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(d_argmin && d_cost_min && d_cost_second, "NULL pointer");
    PSW_CHECK_ARG(coord_offset == 0 || coord_offset == 1, "coord_offset must be 0 or 1");
    PSW_CHECK_ARG(groups >= 0, "groups must not be negative");
    SweepArgs a = {};
    a.ref = d_ref; a.src = d_src; a.Hs = d_H;
    a.W = width; a.H = height; a.N = nsrc; a.P = nplanes; a.off = coord_offset;
    a.stdthresh = stdthresh;
    a.argmin = d_argmin; a.cmin = d_cost_min; a.csec = d_cost_second;
    sweep_layout(a, true, sweep_tile_height());
    const size_t key_bytes = (size_t)width * height * SW_KEY_BYTES;
    if (groups > 0) {
        a.G = std::min(groups, std::min(a.P, 64));
        if (a.G > 1 && key_bytes > PSW_SWEEP_SCRATCH_BYTES) {
            set_error("plane groups of a %d x %d image need %zu bytes of per-pixel keys, above the %zu-byte static "
                      "scratch (groups = 1 needs none, groups = 0 picks 1 here)",
                      width, height, key_bytes, (size_t)PSW_SWEEP_SCRATCH_BYTES);
            return PSW_ERR_ARG;
        }
    } else if (key_bytes > PSW_SWEEP_SCRATCH_BYTES) {
        a.G = 1;
    }
    void* ws = nullptr;
    if (a.G > 1 && (st = scratch_address(&ws)) != PSW_OK) return st;
    return sweep_aggregate_run(a, winsize, fast, ws, S(s));
}

int psw_sweep_fused_aggregate_batch(const psw_sweep_frame* frames, int nframes, int winsize, float stdthresh,
                                    int coord_offset, int fast, int groups, int workgroups, psw_stream_t s) {
    PSW_CHECK_ARG(frames && nframes >= 1, "no frames");
    PSW_CHECK_ARG(coord_offset == 0 || coord_offset == 1, "coord_offset must be 0 or 1");
    PSW_CHECK_ARG(groups >= 0 && workgroups >= 0, "groups and workgroups must not be negative");
    std::vector<SweepArgs> args((size_t)nframes);
    std::vector<int> off((size_t)nframes + 1, 0);
    long total_tiles = 0;
    const int ty = sweep_tile_height();
    for (int f = 0; f < nframes; ++f) {
        const psw_sweep_frame& fr = frames[f];
        int st = check_sweep_args(fr.d_ref, fr.d_src, fr.width, fr.height, fr.nsrc, fr.d_H, fr.nplanes, winsize);
        if (st != PSW_OK) {
            const std::string msg = get_error();
            set_error("frame %d: %s", f, msg.c_str());
            return st;
        }
        PSW_CHECK_ARG(fr.d_argmin && fr.d_cost_min && fr.d_cost_second, "NULL output pointer");
        SweepArgs& a = args[(size_t)f];
        a = SweepArgs{};
        a.ref = fr.d_ref; a.src = fr.d_src; a.Hs = fr.d_H;
        a.W = fr.width; a.H = fr.height; a.N = fr.nsrc; a.P = fr.nplanes; a.off = coord_offset;
        a.stdthresh = stdthresh;
        a.argmin = fr.d_argmin; a.cmin = fr.d_cost_min; a.csec = fr.d_cost_second;
        sweep_layout(a, false, ty);
        total_tiles += a.ntiles;
    }
    // Plane groups: enough items across the whole batch to fill the GPU
    // (about 3000 workgroups' worth) unless given.
    int G = groups > 0 ? groups
                       : (int)std::max(1L, (SW_TARGET_ITEMS * 16 / ty + total_tiles / 2) / std::max(1L, total_tiles));
    size_t ws_bytes = 0;
    for (int f = 0; f < nframes; ++f) {
        SweepArgs& a = args[(size_t)f];
        a.G = std::max(1, std::min(G, std::min(a.P, 64)));
        off[(size_t)f + 1] = off[(size_t)f] + a.ntiles * a.G;
        if (a.G > 1) ws_bytes += (size_t)a.W * a.H * SW_KEY_BYTES;
    }
    const int total = off[(size_t)nframes];
    const size_t args_bytes = sizeof(SweepArgs) * (size_t)nframes, off_bytes = sizeof(int) * ((size_t)nframes + 1);
    const size_t head = (args_bytes + off_bytes + 16 + 255) / 256 * 256;
    char* dev = nullptr;
    {
        void* w = nullptr;
        const int rc0 = workspace_alloc_async(&w, head + ws_bytes, S(s));
        if (rc0 != PSW_OK) return rc0;
        dev = static_cast<char*>(w);
    }
    // per-frame keys of the plane groups
    size_t o = head;
    for (int f = 0; f < nframes; ++f) {
        SweepArgs& a = args[(size_t)f];
        if (a.G > 1) {
            a.top2 = reinterpret_cast<unsigned long long*>(dev + o);
            o += (size_t)a.W * a.H * SW_KEY_BYTES;
        }
    }
    SweepArgs* d_args = reinterpret_cast<SweepArgs*>(dev);
    int* d_off = reinterpret_cast<int*>(dev + args_bytes);
    unsigned* d_counter = reinterpret_cast<unsigned*>(dev + args_bytes + off_bytes);
    hipStream_t st = S(s);
    int rc = PSW_OK;
    hipError_t e = hipMemcpyAsync(d_args, args.data(), args_bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_off, off.data(), off_bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) rc = hip_status(e, "psw_sweep_fused_aggregate_batch", PSW_ERR_COPY);
    if (rc == PSW_OK) rc = psw_fill_u32(d_counter, 0u, 1, s);
    if (rc == PSW_OK && ws_bytes > 0)
        rc = psw_fill_u32(reinterpret_cast<uint32_t*>(dev + head), 0xFFFFFFFFu, ws_bytes / 4, s);
    const int wg = workgroups > 0 ? std::min(workgroups, total) : std::min(total, 2048);
    if (rc == PSW_OK) {
        const dim3 grid((unsigned)wg), block(SW_THREADS);
        switch (winsize / 2) {
#define PSW_SB(r)                                                                                             \
    case r:                                                                                                   \
        if (ty == 32) {                                                                                       \
            if (fast) sweep_batch_kernel<r, 1, true, 32><<<grid, block, 0, st>>>(d_args, d_off, nframes, total, d_counter); \
            else sweep_batch_kernel<r, 1, false, 32><<<grid, block, 0, st>>>(d_args, d_off, nframes, total, d_counter); \
        } else {                                                                                              \
            if (fast) sweep_batch_kernel<r, 1, true, 16><<<grid, block, 0, st>>>(d_args, d_off, nframes, total, d_counter); \
            else sweep_batch_kernel<r, 1, false, 16><<<grid, block, 0, st>>>(d_args, d_off, nframes, total, d_counter); \
        }                                                                                                     \
        break;
            PSW_SB(0) PSW_SB(1) PSW_SB(2) PSW_SB(3) PSW_SB(4) PSW_SB(5) PSW_SB(6) PSW_SB(7)
#undef PSW_SB
            default:
                set_error("window size %d not supported (odd, 1..15)", winsize);
                rc = PSW_ERR_UNSUPPORTED;
        }
        if (rc == PSW_OK) rc = launch_status("sweep_batch_kernel");
    }
    for (int f = 0; f < nframes && rc == PSW_OK; ++f) {
        const SweepArgs& a = args[(size_t)f];
        if (a.G <= 1) continue;
        const size_t npix = (size_t)a.W * a.H;
        sweep_top2_kernel<<<(unsigned)((npix + 255) / 256), 256, 0, st>>>(a.top2, npix, a.argmin, a.cmin, a.csec);
        rc = launch_status("sweep_top2_kernel");
    }
    // the host arrays live until the copies are done; the workspace goes
    // back to the pool in stream order
    (void)workspace_free_async(dev, st);
    const hipError_t se = hipStreamSynchronize(st);
    if (rc == PSW_OK && se != hipSuccess) rc = hip_status(se, "psw_sweep_fused_aggregate_batch");
    return rc;
}

int psw_sweep_argmin_depth(float* d_depth, const int32_t* d_argmin, const float* d_depths, int npixels,
                           float fallback, psw_stream_t s) {
    PSW_CHECK_ARG(d_depth && d_argmin && d_depths, "NULL pointer");
    if (npixels <= 0) return PSW_OK;
    argmin_depth_kernel<<<(npixels + 255) / 256, 256, 0, S(s)>>>(d_depth, d_argmin, d_depths, npixels, fallback);
    return launch_status("argmin_depth_kernel");
}

}  // extern "C"
