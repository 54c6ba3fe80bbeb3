// tgv_fused.hip -- TGV2 multiview primal-dual depth refinement for gfx950.
//
// The reference runs, per iteration, P, Q, prodsum:=0, N x R and U as 4+N
// launches over full-frame images (src/planesweep.cpp:485-509).  Here one
// kernel does the whole iteration for a 64x8 tile:
//   phase 1: ubar, u1bar on the tile plus a one-pixel halo into LDS, formed
//            on the fly as 2*u_k - u_{k-1} (the value the reference's U step
//            stores as ubar, TGV2_kernels.cu:334-336);
//   phase 2: new p (TGV2_kernels.cu:340-365) and q (:54-85) on the tile plus
//            its left/top halo, into LDS and, for the tile, to HBM;
//   phase 3: r of every source and prodsum (:87-105) in registers, then the
//            tensor-weighted u / u1 step (:300-338) from the LDS duals.
// Buffers read at neighbours by other tiles are never written in place:
// the primal (u, u1x, u1y) cycles through three buffers (k-1, k, k+1) and
// the duals through two.  q_w is not stored: the reference updates it with
// exactly the q_z formula from the same zero start, so it always equals q_z.
//
// Arithmetic is FP32 throughout (the reference's double temporaries in the
// P and Q steps become float); FP64 is used only in the energy reductions
// that decide stopping.
#include "psw_common.h"
#include "tgv_point.h"

#include <cstring>
#include <mutex>
#include <algorithm>
#include <vector>

namespace psw {

constexpr int TG_TX = 64;
constexpr int TG_THREADS = 256;   // energy kernels
// u cycles through 8 slots, p, q, r through 5.  A launch of m = 1, 2 or 3
// iterations from state j reads u_j, u_{j-1}, p_j, q_j, r_j and writes u of
// j+1..j+m (at least the last two) and p, q, r of j+m only, so it never
// writes a slot it reads (8 does not divide 1..4, 5 does not divide 1..3).
// The energy launch of a check at state k and the finalising launch after
// it have the same m; a stop at k is seen while they write u of k+1..k+2m
// and p, q, r of k+m and k+2m, so u_k, u_{k-1}, p_k, q_k, r_k stay in place
// (8 divides none of 1..7, 5 none of 1..6).
constexpr int TG_USLOTS = 8;
constexpr int TG_PSLOTS = 5;

struct TgvLayout {
    size_t n;
    size_t U, P, Q, r, It, Iu, T, u0, ref, src, mask, total;
    void init(int W, int H, int N) {
        n = (size_t)W * H;
        size_t o = 0;
        U = o;  o += 3 * TG_USLOTS * n;  // [TG_USLOTS slots][u, u1x, u1y]: u_k in slot k % TG_USLOTS
        P = o;  o += 2 * TG_PSLOTS * n;  // [TG_PSLOTS slots][px, py]: p_k in slot k % TG_PSLOTS
        Q = o;  o += 3 * TG_PSLOTS * n;  // [TG_PSLOTS slots][qx, qy, qz]
        r = o;  o += (size_t)TG_PSLOTS * N * n;  // [TG_PSLOTS slots][N][n]
        It = o; o += (size_t)N * n;
        Iu = o; o += (size_t)N * n;
        T = o;  o += 3 * n;      // T11, T12 (= T21), T22
        u0 = o; o += n;
        ref = o; o += n;         // normalised reference
        src = o; o += (size_t)N * n;
        mask = o; o += n;        // energy weight per pixel (1 = counted)
        total = o;
    }
};

struct TgvCtrl {
    int done;
    int iters_at_done;
    int pad[2];
    double energy;
    double energy_at_done;
};

}  // namespace psw

struct psw_tgv {
    int W, H, N;
    psw::TgvLayout L;
    float* base;
    double* partials;     // per-block energy partial sums
    int npartials;
    psw::TgvCtrl* ctrl;   // device
    unsigned long long* masked;  // device counter of psw_tgv_set_energy_mask
    psw::TgvCtrl* hctrl;  // pinned host copy (filled by copies)
    psw::TgvCtrl* hmirror;  // pinned, mapped host memory written by checking launches
    psw::TgvCtrl* dmirror;  // its device address
    unsigned long long* rx; // exchange words of the resident solver (NULL: not available), uncached, pooled
    size_t rx_words;
    int rx_dev;
    int k;                // iterations since the warp started
    bool have_images;
};

namespace psw {

struct TgvPoses {
    M3f K, invK;
    M3f Rrel[PSW_MAX_SOURCES];
    M3f RK[PSW_MAX_SOURCES];  // Rrel * invK
    V3f trel[PSW_MAX_SOURCES];
    float fx, fy;
};

// Slots by iteration count k (so that launches doing one or two iterations
// mix freely and a no-op launch after the stop leaves the mapping intact):
// u_k in U slot k % TG_USLOTS, p_k / q_k / r_k in slot k % TG_PSLOTS.  A launch reads the
// slots of k (and k-1 for u) and writes those of k+1 (and k+2).
struct TgvPtrs {
    const float* ucur;   // [3][n] u_k
    const float* uprev;  // [3][n] u_{k-1}
    float* unext;        // [3][n] u_{k+1}
    float* unext2;       // [3][n] u_{k+2}
    float* unext3;       // [3][n] u_{k+3}
    const float* pcur;   // [2][n]
    float* pnext;
    float* pnext2;
    float* pnext3;
    const float* qcur;   // [3][n]
    float* qnext;
    float* qnext2;
    float* qnext3;
    const float* rcur;   // [N][n]
    float* rnext;
    float* rnext2;
    float* rnext3;
    const float* It;
    const float* Iu;
    const float* T;      // [3][n]
    const float* u0;
    const TgvCtrl* ctrl;
    int W, H, N;
    float lambda, alpha0, alpha1, tau, sigma;
};

// ---- set-up kernels --------------------------------------------------------

__global__ void tgv_scale_kernel(float* __restrict__ out, const float* __restrict__ in, size_t n) {
    const float sc = 1.f / 255.f;  // element_scale(Ref, 1/255.f), planesweep.cpp:419
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i] * sc;
}

// Anisotropic_diffusion_tensor_kernel (TGV2_kernels.cu:257-298); T21 = T12.
__global__ void tgv_tensor_kernel(float* __restrict__ T, const float* __restrict__ img, int W, int H, float beta,
                                  float gamma) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t n = (size_t)W * H, i = (size_t)y * W + x;
    const int xn = min(x + 1, W - 1), yn = min(y + 1, H - 1);
    float t11, t12, t22;
    tensor_point(img[(size_t)y * W + xn] - img[i], img[(size_t)yn * W + x] - img[i], beta, gamma, t11, t12, t22);
    T[i] = t11;
    T[n + i] = t12;
    T[2 * n + i] = t22;
}

// Warped source intensity at pixel (x, y) for depth u (transform_coordinates
// TGV2_kernels.cu:156-177 + bilinear kernels.cu:4-29), 1-based convention.
__device__ inline float tgv_warp(const TgvPoses& pz, int j, const float* __restrict__ src, int W, int H, int x, int y,
                                 float u, float& X, float& Y, float& Z) {
    float a0, a1, a2;
    mv3(pz.invK, (float)(x + 1), (float)(y + 1), 1.f, a0, a1, a2);
    a0 = u * a0; a1 = u * a1; a2 = u * a2;
    float b0, b1, b2;
    mv3(pz.Rrel[j], a0, a1, a2, b0, b1, b2);
    X = b0 + pz.trel[j].x;
    Y = b1 + pz.trel[j].y;
    Z = b2 + pz.trel[j].z;
    float c0, c1, c2;
    mv3(pz.K, X, Y, Z, c0, c1, c2);
    return bilinear_sample(src, W, H, c0 / c2 - 1.f, c1 / c2 - 1.f);
}

// One warp's linearisation for every source (planesweep.cpp:460-483):
// It_j = I_j(warp(u0)) - Ref, Iu_j = dI_j/du (TGV2_kernels.cu:194-255), r_j = 0.
__global__ void tgv_linearise_kernel(TgvPoses pz, const float* __restrict__ src, const float* __restrict__ ref,
                                     const float* __restrict__ u0, float* __restrict__ It, float* __restrict__ Iu,
                                     float* __restrict__ r, int W, int H, int N) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t n = (size_t)W * H, i = (size_t)y * W + x;
    const int xn = min(x + 1, W - 1), yn = min(y + 1, H - 1);
    const float uc = u0[i], ux = u0[(size_t)y * W + xn], uy = u0[(size_t)yn * W + x];
    for (int j = 0; j < N; ++j) {
        const float* S = src + (size_t)j * n;
        float X, Y, Z, Xd, Yd, Zd;
        const float Ic = tgv_warp(pz, j, S, W, H, x, y, uc, X, Y, Z);
        const float Ix = tgv_warp(pz, j, S, W, H, xn, y, ux, Xd, Yd, Zd);
        const float Iy = tgv_warp(pz, j, S, W, H, x, yn, uy, Xd, Yd, Zd);
        float dX, dY, dZ;
        mv3(pz.RK[j], (float)(x + 1), (float)(y + 1), 1.f, dX, dY, dZ);
        const float dfx = pz.fx * (dX * Z - X * dZ) / (Z * Z);
        const float dfy = pz.fy * (dY * Z - Y * dZ) / (Z * Z);
        Iu[(size_t)j * n + i] = (Ix - Ic) * dfx + (Iy - Ic) * dfy;
        It[(size_t)j * n + i] = Ic - ref[i];
        r[(size_t)j * n + i] = 0.f;
    }
}

// Start of a warp: u_k = u_{k-1} = u (so ubar = u), u1 = 0 in both, duals 0.
__global__ void tgv_warp_reset_kernel(float* __restrict__ U, float* __restrict__ P, float* __restrict__ Q,
                                      float* __restrict__ u0, size_t n, int cur, int prev) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float u = U[(size_t)cur * 3 * n + i];
        u0[i] = u;
        U[(size_t)prev * 3 * n + i] = u;
        U[(size_t)cur * 3 * n + n + i] = 0.f;
        U[(size_t)cur * 3 * n + 2 * n + i] = 0.f;
        U[(size_t)prev * 3 * n + n + i] = 0.f;
        U[(size_t)prev * 3 * n + 2 * n + i] = 0.f;
        P[i] = 0.f; P[n + i] = 0.f;
        Q[i] = 0.f; Q[n + i] = 0.f; Q[2 * n + i] = 0.f;
    }
}

// Warm start inside a linearisation: both time levels of u and u1 take the
// given fields (u1 = 0 when none is given), duals p, q, r restart from 0; u0,
// It and Iu (the linearisation) are left alone.
__global__ void tgv_set_primal_kernel(float* __restrict__ U, float* __restrict__ P, float* __restrict__ Q,
                                      float* __restrict__ r, const float* __restrict__ du, float uval,
                                      const float* __restrict__ du1, size_t n, int N, int cur, int prev) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float u = du ? du[i] : uval;
        const float ax = du1 ? du1[i] : 0.f, ay = du1 ? du1[n + i] : 0.f;
        U[(size_t)cur * 3 * n + i] = u;
        U[(size_t)prev * 3 * n + i] = u;
        U[(size_t)cur * 3 * n + n + i] = ax;
        U[(size_t)cur * 3 * n + 2 * n + i] = ay;
        U[(size_t)prev * 3 * n + n + i] = ax;
        U[(size_t)prev * 3 * n + 2 * n + i] = ay;
        P[i] = 0.f; P[n + i] = 0.f;
        Q[i] = 0.f; Q[n + i] = 0.f; Q[2 * n + i] = 0.f;
        for (int j = 0; j < N; ++j) r[(size_t)j * n + i] = 0.f;
    }
}

// Energy weights: 0 where the data block of the primal-dual step is far from
// stable -- q = tau sigma lambda^2 sum_j Iu_j^2 above qmax at the pixel or a
// 4-neighbour (the reference's fixed steps let such pixels oscillate without
// end) -- and 1 elsewhere.  qmax <= 0 counts every pixel.
__global__ void tgv_mask_kernel(const float* __restrict__ Iu, int W, int H, int N, double scale, double qmax,
                                float* __restrict__ mask, unsigned long long* __restrict__ excluded) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    bool keep = true;
    if (x < W && y < H && qmax > 0.0) {
        const size_t n = (size_t)W * H;
        const int xs[5] = {x, x - 1, x + 1, x, x};
        const int ys[5] = {y, y, y, y - 1, y + 1};
        for (int k = 0; k < 5; ++k) {
            if (xs[k] < 0 || xs[k] >= W || ys[k] < 0 || ys[k] >= H) continue;
            const size_t i = (size_t)ys[k] * W + xs[k];
            double q = 0.0;
            for (int j = 0; j < N; ++j) {
                const double v = Iu[(size_t)j * n + i];
                q += v * v;
            }
            if (q * scale > qmax) keep = false;
        }
    }
    if (x < W && y < H) {
        mask[(size_t)y * W + x] = keep ? 1.f : 0.f;
        if (!keep && excluded) atomicAdd(excluded, 1ull);
    }
}

// ---- the fused primal-dual iteration --------------------------------------

// Point updates shared by every dense TGV2 kernel (the same operations in
// the same order, so all give identical results): the reference's
// expressions with each product that feeds an addition fused into it (one
// rounding, as a contracting compiler such as nvcc by default emits them);
// the dual update pq_update is in tgv_point.h.

// TGV2_updateR at one pixel and source (uses u_k, not ubar).
__device__ __forceinline__ float r_update(float r, float it, float du, float iu, float sl) {
    const float rr = __builtin_fmaf(sl, __builtin_fmaf(du, iu, it), r);
    // rr / fmaxf(1, |rr|): rr itself for |rr| <= 1, exactly +-1 above, which
    // is the median of (rr, -1, 1) for every finite rr (one instruction)
    return __builtin_amdgcn_fmed3f(rr, -1.f, 1.f);
}

// The primal steps of TGV2_updateU_tensor_weighed with the products fused
// into the following additions (as a compiler that contracts them would):
// u - tau (-alpha1 (T (c - l) + T (c - up)) + lambda ps), and
// u1 - tau (-alpha1 (T c) - alpha0 div q) for one component of u1.
__device__ __forceinline__ float u_step(float u, float t11, float t12, float t22, float c_px, float c_py, float l_px,
                                        float l_py, float u_px, float u_py, float ps, float tau, float alpha1,
                                        float lambda) {
    float acc = t11 * (c_px - l_px);
    acc = __builtin_fmaf(t12, c_py - l_py, acc);
    acc = __builtin_fmaf(t12, c_px - u_px, acc);
    acc = __builtin_fmaf(t22, c_py - u_py, acc);
    return __builtin_fmaf(-tau, __builtin_fmaf(-alpha1, acc, lambda * ps), u);
}

__device__ __forceinline__ float u1_step(float u1, float ta, float tb, float c_px, float c_py, float divq, float tau,
                                         float alpha1, float alpha0) {
    return __builtin_fmaf(-tau, __builtin_fmaf(-alpha1, __builtin_fmaf(ta, c_px, tb * c_py), -alpha0 * divq), u1);
}

// The data dual's share of the u step: ps + r Iu.
__device__ __forceinline__ float ps_add(float ps, float r, float iu) { return __builtin_fmaf(r, iu, ps); }

// TGV2_updateU_tensor_weighed at one pixel: c = own duals, l = left
// neighbour's, u = upper neighbour's (own at the first column / row).
__device__ __forceinline__ void u_update(float u, float u1x, float u1y, float t11, float t12, float t22, float c_px,
                                         float c_py, float l_px, float l_py, float u_px, float u_py, float q0c,
                                         float q0l, float q1c, float q1u, float q2c, float q2l, float q2u, float ps,
                                         const TgvPtrs& a, float& un, float& u1xn, float& u1yn) {
    un = u_step(u, t11, t12, t22, c_px, c_py, l_px, l_py, u_px, u_py, ps, a.tau, a.alpha1, a.lambda);
    u1xn = u1_step(u1x, t11, t12, c_px, c_py, q0c - q0l + q2c - q2u, a.tau, a.alpha1, a.alpha0);
    u1yn = u1_step(u1y, t12, t22, c_px, c_py, q2c - q2l + q1c - q1u, a.tau, a.alpha1, a.alpha0);
}

// Energy stop test inside the iteration launches.  The launch that starts
// from state k also sums the energy of state k (it holds u_k in LDS with the
// halo the forward differences need) and stores one partial per tile; one
// extra workgroup of the next launch adds those partials in tile order
// (after the kernel boundary, so with plain loads) and applies the stop
// test.  A stop at k thus takes effect one launch later; the two launches
// wrote later states into other slots (see TG_USLOTS), so state k stays
// intact, and the launches after them are no-ops.
struct TgvCheck {
    double* partials;            // this launch's partials, one per tile (energy launches)
    const double* fin_partials;  // the partials to finalise (finalising launches)
    int fin_count;
    int fin_k;                   // the state they belong to
    TgvCtrl* ctrl;               // device
    TgvCtrl* mirror;             // pinned host memory: the host polls it
    const float* mask;
    double lambda, alpha0, alpha1, tol, e_star;
};

// alpha1 |T(grad u - u1)| + alpha0 |eps(u1)|_F + lambda ed at one pixel, with
// the iteration's clamped forward differences (x / y = right / lower
// neighbour); ed = sum_j |It_j + (u - u0) Iu_j|.
__device__ __forceinline__ double energy_point(double u, double ux, double uy, double a1, double a1x, double a1y,
                                               double b1, double b1x, double b1y, double t11, double t12, double t22,
                                               double ed, double lambda, double alpha0, double alpha1) {
    const double gx = (ux - u) - a1, gy = (uy - u) - b1;
    const double p = t11 * gx + t12 * gy, q = t12 * gx + t22 * gy;
    const double ex = a1x - a1, ey = b1y - b1, ez = 0.5 * ((a1y - a1) + (b1x - b1));
    return alpha1 * sqrt(p * p + q * q) + alpha0 * sqrt(ex * ex + ey * ey + 2.0 * ez * ez) + lambda * ed;
}

// Sum over the NT threads of a workgroup, in a fixed order; valid on thread 0.
template <int NT>
__device__ inline double block_sum(double v) {
    __shared__ double wsum[NT / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / 64; ++w) t += wsum[w];
    return t;
}

// Sum of the partials (the first 256 threads in a fixed pattern, so the
// result does not depend on NT), then the stop test.
constexpr int FIN_THREADS = 256;

// The stop rule: the energy is at most tol (relative) above e_star, the
// lowest energy known for the problem (an energy below it stops too).
__device__ inline bool tgv_at_gap(double E, double e_star, double tol) {
    return e_star > 0.0 && E - e_star <= tol * e_star;
}

// Returns (on thread 0) whether this check stopped the solve.  The host
// mirror is optional (the batched solve counts stops on the device instead).
template <int NT>
__device__ bool check_finalize(const TgvCheck& c) {
    double v = 0.0;
    if (threadIdx.x < FIN_THREADS)
        for (int i = threadIdx.x; i < c.fin_count; i += FIN_THREADS) v += c.fin_partials[i];
    const double E = block_sum<NT>(v);
    bool stopped = false;
    if (threadIdx.x == 0) {
        TgvCtrl* d = c.ctrl;
        volatile TgvCtrl* m = c.mirror;
        d->energy = E;
        if (m) m->energy = E;
        if (!d->done && tgv_at_gap(E, c.e_star, c.tol)) {
            d->iters_at_done = c.fin_k;
            d->energy_at_done = E;
            d->done = 1;
            stopped = true;
            if (m) {
                m->iters_at_done = c.fin_k;
                m->energy_at_done = E;
                __threadfence_system();
                m->done = 1;
            }
        }
    }
    return stopped;
}

template <int NT>
__global__ __launch_bounds__(NT) void tgv_check_finalize_kernel(TgvCheck c) {
    (void)check_finalize<NT>(c);
}

// Tile TG_TX x TY pixels, NT threads.  Variants (selected by tgv_variant()):
// 64x8 with 256 or 512 threads, 64x4 with 256 threads.
// One tile of one iteration; called by every thread of the workgroup (it
// contains barriers).
// NS: the most sources the instance handles (a.N <= NS); all sources' r, It,
// Iu of a pixel are loaded before any r is stored, so the loads overlap.
template <int TY, int NT, int NS, bool CHECK = false>
__device__ __forceinline__ void tgv_iter_tile(const TgvPtrs& a, const int tile, const TgvCheck* chk = nullptr) {
    constexpr int TG_TY = TY, TG_THREADS = NT;
    constexpr int HX = TG_TX + 2, HY = TG_TY + 2;   // ubar region: x0-1..x0+TX, y0-1..y0+TY
    constexpr int DX = TG_TX + 1, DY = TG_TY + 1;   // dual region: x0-1..x0+TX-1, y0-1..y0+TY-1
    __shared__ float sUb[3][HY][HX];
    __shared__ float sP[2][DY][DX];
    __shared__ float sQ[3][DY][DX];
    __shared__ float sUc[CHECK ? 3 : 1][CHECK ? HY : 1][CHECK ? HX : 1];   // u_k (checking launches)

    const int W = a.W, H = a.H;
    const size_t n = (size_t)W * H;
    const int tiles_x = (W + TG_TX - 1) / TG_TX;
    const int x0 = (tile % tiles_x) * TG_TX, y0 = (tile / tiles_x) * TG_TY;
    const int tid = threadIdx.x;

    // phase 1: ubar / u1bar at clamped coordinates (clamping reproduces the
    // reference's min(x+1, W-1) neighbours, TGV2_kernels.cu:352-353).
    for (int i = tid; i < HX * HY; i += TG_THREADS) {
        const int hy = i / HX, hx = i - (i / HX) * HX;
        const int gx = min(max(x0 - 1 + hx, 0), W - 1);
        const int gy = min(max(y0 - 1 + hy, 0), H - 1);
        const size_t g = (size_t)gy * W + gx;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const float uc = a.ucur[f * n + g];
            sUb[f][hy][hx] = 2.f * uc - a.uprev[f * n + g];
            if (CHECK) sUc[f][hy][hx] = uc;
        }
    }
    __syncthreads();

    // phase 2: new p, q on the tile plus left/top halo.
    const float as1 = a.alpha1 * a.sigma, as0 = a.alpha0 * a.sigma;
    for (int i = tid; i < DX * DY; i += TG_THREADS) {
        const int dy = i / DX, dx = i - (i / DX) * DX;
        const int vx = x0 - 1 + dx, vy = y0 - 1 + dy;
        const int gx = min(max(vx, 0), W - 1), gy = min(max(vy, 0), H - 1);
        const size_t g = (size_t)gy * W + gx;
        const int hx = dx, hy = dy;  // ubar index of this position (+1 -> right / down neighbour)
        const float t11 = a.T[g], t12 = a.T[n + g], t22 = a.T[2 * n + g];
        float px = a.pcur[g], py = a.pcur[n + g];
        float qx = a.qcur[g], qy = a.qcur[n + g], qz = a.qcur[2 * n + g];
        pq_update(px, py, qx, qy, qz, t11, t12, t22, sUb[0][hy][hx], sUb[0][hy][hx + 1], sUb[0][hy + 1][hx],
                  sUb[1][hy][hx], sUb[1][hy][hx + 1], sUb[1][hy + 1][hx], sUb[2][hy][hx], sUb[2][hy][hx + 1],
                  sUb[2][hy + 1][hx], as1, as0);
        sP[0][dy][dx] = px; sP[1][dy][dx] = py;
        sQ[0][dy][dx] = qx; sQ[1][dy][dx] = qy; sQ[2][dy][dx] = qz;
        if (dx > 0 && dy > 0 && vx < W && vy < H) {
            a.pnext[g] = px; a.pnext[n + g] = py;
            a.qnext[g] = qx; a.qnext[n + g] = qy; a.qnext[2 * n + g] = qz;
        }
    }
    __syncthreads();

    // phase 3: r, prodsum and the primal step for the tile pixels (and, in
    // a checking launch, the energy of state k).
    const float sl = a.sigma * a.lambda;
    double eacc = 0.0;
    for (int i = tid; i < TG_TX * TG_TY; i += TG_THREADS) {
        const int ty = i / TG_TX, tx = i - (i / TG_TX) * TG_TX;
        const int x = x0 + tx, y = y0 + ty;
        if (x >= W || y >= H) continue;
        const size_t g = (size_t)y * W + x;
        const float u = a.ucur[g], u1x = a.ucur[n + g], u1y = a.ucur[2 * n + g];
        const float u0g = a.u0[g];
        const float du = u - u0g;
        const double dud = (double)u - (double)u0g;
        float rv[NS], itv[NS], iuv[NS];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                const size_t gj = (size_t)j * n + g;
                rv[j] = a.rcur[gj];
                itv[j] = a.It[gj];
                iuv[j] = a.Iu[gj];
            }
        }
        float ps = 0.f;
        double ed = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                const float rr = r_update(rv[j], itv[j], du, iuv[j], sl);
                a.rnext[(size_t)j * n + g] = rr;
                ps = ps_add(ps, rr, iuv[j]);
                if (CHECK) ed += fabs((double)itv[j] + dud * (double)iuv[j]);
            }
        }
        if (CHECK && chk->mask[g] != 0.f) {
            const int hx = tx + 1, hy = ty + 1;          // u_k index; +1 is the clamped right / lower neighbour
            eacc += energy_point(sUc[0][hy][hx], sUc[0][hy][hx + 1], sUc[0][hy + 1][hx], sUc[1][hy][hx],
                                 sUc[1][hy][hx + 1], sUc[1][hy + 1][hx], sUc[2][hy][hx], sUc[2][hy][hx + 1],
                                 sUc[2][hy + 1][hx], a.T[g], a.T[n + g], a.T[2 * n + g], ed, chk->lambda,
                                 chk->alpha0, chk->alpha1);
        }
        const int cx = tx + 1, cy = ty + 1;              // dual index of this pixel
        const int lx = x > 0 ? cx - 1 : cx;              // max(x-1, 0)
        const int ty_ = y > 0 ? cy - 1 : cy;             // max(y-1, 0)
        const float t11 = a.T[g], t12 = a.T[n + g], t22 = a.T[2 * n + g];
        const float c_px = sP[0][cy][cx], c_py = sP[1][cy][cx];
        const float xp_px = sP[0][cy][lx], xp_py = sP[1][cy][lx];
        const float yp_px = sP[0][ty_][cx], yp_py = sP[1][ty_][cx];
        float un, u1xn, u1yn;
        u_update(u, u1x, u1y, t11, t12, t22, c_px, c_py, xp_px, xp_py, yp_px, yp_py, sQ[0][cy][cx], sQ[0][cy][lx],
                 sQ[1][cy][cx], sQ[1][ty_][cx], sQ[2][cy][cx], sQ[2][cy][lx], sQ[2][ty_][cx], ps, a, un, u1xn, u1yn);
        a.unext[g] = un;
        a.unext[n + g] = u1xn;
        a.unext[2 * n + g] = u1yn;
    }
    if (CHECK) {
        const double e = block_sum<NT>(eacc);
        if (threadIdx.x == 0) chk->partials[tile] = e;
    }
}

// Two iterations (k -> k+1 -> k+2) of a 64 x TY tile in one launch.  Each
// iteration widens the dependence by one pixel, so iteration k+1 runs on the
// tile plus a 1-pixel ring (computing p, q on a 2-pixel ring and reading
// ubar_k on a 2-pixel ring plus the right / lower neighbours), entirely in
// LDS; only the tile's u_{k+1}, u_{k+2}, p_{k+2}, q_{k+2}, r_{k+2} reach HBM.
// Regions (global offsets from the tile origin x0, y0; positions outside the
// image read clamped coordinates and are never used by pixels inside it,
// because left/up neighbours at the first column/row and right/down
// neighbours at the last are the position itself, as in the reference),
// sizes for TY = 8:
//   A  ubar_k          x0-2 .. x0+65, y0-2 .. y0+9    (68 x 12)
//   B  p, q_{k+1}      x0-2 .. x0+64, y0-2 .. y0+8    (67 x 11)
//   C  u_{k+1}, u_k    x0-1 .. x0+64, y0-1 .. y0+8    (66 x 10)
//   D  p, q_{k+2}      x0-1 .. x0+63, y0-1 .. y0+7    (65 x  9)
//   tile u, r_{k+2}    x0   .. x0+63, y0   .. y0+7    (64 x  8)
// T2_MAXN: the most sources of the two- and three-iteration kernels.
constexpr int T2_TY = 10, T2_MAXN = 16;

__device__ __forceinline__ int clampi(int v, int hi) { return min(max(v, 0), hi); }

// k-th position of the one-pixel ring around a TX x TY tile inside the
// CX x CY region C: top row, bottom row, left column, right column.
template <int CX, int CY, int TY>
__device__ __forceinline__ void ring_pos(int k, int& ix, int& iy) {
    if (k < CX) { iy = 0; ix = k; }
    else if (k < 2 * CX) { iy = CY - 1; ix = k - CX; }
    else if (k < 2 * CX + TY) { iy = k - 2 * CX + 1; ix = 0; }
    else { iy = k - 2 * CX - TY + 1; ix = CX - 1; }
}

// One tile pixel per thread for the first TX * TY threads: thread t owns
// pixel (t / TX, t % TX) in both iterations and keeps its r_{k+1}, It, Iu,
// u0 and T in registers (NS >= a.N sources), so no source plane is read
// twice.  NT > TX * TY threads give the halo phases more lanes (one pass for
// more of them) and take the one-pixel ring of iteration k+1 on the threads
// past the tile, in parallel with the tile pixels.
template <int NS, bool ENERGY = false, int TY = T2_TY, int NT = TG_TX * TY>
__device__ __forceinline__ void tgv_iter2_tile(const TgvPtrs& a, const int tile, const TgvCheck* chk = nullptr) {
    constexpr int TX = TG_TX, TP = TG_TX * TY;
    static_assert(NT >= TP, "at least one thread per tile pixel");
    constexpr int AX = TX + 4, AY = TY + 4;
    constexpr int BX = TX + 3, BY = TY + 3;
    constexpr int CX = TX + 2, CY = TY + 2;
    constexpr int DX = TX + 1, DY = TY + 1;
    __shared__ float sA[3][AY][AX];   // ubar_k on A; then ubar_{k+1} on C (A index = C index + 1), formed
                                      // by the thread that computes u_{k+1} at the position
    __shared__ float sUk[3][CY][CX];  // u_k on C
    __shared__ float sP[2][BY][BX];   // p_{k+1} on B; then p_{k+2} on D in place (B index = D index + 1)
    __shared__ float sQ[3][BY][BX];
    __shared__ float sT[3][BY][BX];   // T11, T12, T22 on B (read again on D and the C ring)

    const int W = a.W, H = a.H;
    const size_t n = (size_t)W * H;
    const int tiles_x = (W + TX - 1) / TX;
    const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
    const int tid = threadIdx.x;
    const float as1 = a.alpha1 * a.sigma, as0 = a.alpha0 * a.sigma;
    const float sl = a.sigma * a.lambda;

    // T, p_k, q_k on B (p, q are then updated in place)
    for (int i = tid; i < BX * BY; i += NT) {
        const int iy = i / BX, ix = i - iy * BX;
        const size_t g = (size_t)clampi(y0 - 2 + iy, H - 1) * W + clampi(x0 - 2 + ix, W - 1);
#pragma unroll
        for (int f = 0; f < 3; ++f) sT[f][iy][ix] = a.T[f * n + g];
        sP[0][iy][ix] = a.pcur[g];
        sP[1][iy][ix] = a.pcur[n + g];
#pragma unroll
        for (int f = 0; f < 3; ++f) sQ[f][iy][ix] = a.qcur[f * n + g];
    }
    // A: ubar_k; its inner part C (one pixel in from each side) also keeps u_k
    for (int i = tid; i < AX * AY; i += NT) {
        const int iy = i / AX, ix = i - iy * AX;
        const size_t g = (size_t)clampi(y0 - 2 + iy, H - 1) * W + clampi(x0 - 2 + ix, W - 1);
        const bool inC = ix >= 1 && ix <= CX && iy >= 1 && iy <= CY;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const float uk = a.ucur[f * n + g];
            sA[f][iy][ix] = 2.f * uk - a.uprev[f * n + g];
            if (inC) sUk[f][iy - 1][ix - 1] = uk;
        }
    }
    // this thread's tile pixel (clamped coordinates outside the image: computed, never stored)
    const bool tp = NT == TP || tid < TP;
    const int ty = tid / TX, tx = tid - (tid / TX) * TX;
    const int x = x0 + tx, y = y0 + ty;
    const size_t gt = (size_t)clampi(y, H - 1) * W + clampi(x, W - 1);
    float rv[NS], itv[NS], iuv[NS];
    float u0t = 0.f;
    if (tp) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                const size_t gj = (size_t)j * n + gt;
                rv[j] = a.rcur[gj];
                itv[j] = a.It[gj];
                iuv[j] = a.Iu[gj];
            }
        }
        u0t = a.u0[gt];
    }
    __syncthreads();
    // the tile pixel's tensor, from the B region staged above (B index = tile index + 2)
    float t11 = 0.f, t12 = 0.f, t22 = 0.f;
    if (tp) {
        t11 = sT[0][ty + 2][tx + 2]; t12 = sT[1][ty + 2][tx + 2]; t22 = sT[2][ty + 2][tx + 2];
    }

    // B: p, q of iteration k+1
    for (int i = tid; i < BX * BY; i += NT) {
        const int iy = i / BX, ix = i - iy * BX;
        const int gx = x0 - 2 + ix, gy = y0 - 2 + iy;
        const int rx = gx < W - 1 ? ix + 1 : ix, dy = gy < H - 1 ? iy + 1 : iy;
        float px = sP[0][iy][ix], py = sP[1][iy][ix];
        float qx = sQ[0][iy][ix], qy = sQ[1][iy][ix], qz = sQ[2][iy][ix];
        pq_update(px, py, qx, qy, qz, sT[0][iy][ix], sT[1][iy][ix], sT[2][iy][ix], sA[0][iy][ix], sA[0][iy][rx],
                  sA[0][dy][ix], sA[1][iy][ix], sA[1][iy][rx], sA[1][dy][ix], sA[2][iy][ix], sA[2][iy][rx],
                  sA[2][dy][ix], as1, as0);
        sP[0][iy][ix] = px; sP[1][iy][ix] = py;
        sQ[0][iy][ix] = qx; sQ[1][iy][ix] = qy; sQ[2][iy][ix] = qz;
    }
    __syncthreads();

    // C: r and u of iteration k+1 and ubar_{k+1} = 2 u_{k+1} - u_k into sA
    // (ubar_k is no longer read); the tile pixel from registers (its r_{k+1}
    // and u_{k+1} stay there), then the ring.  An energy launch also sums the
    // energy of state k at the tile pixel (u_k with its right / lower
    // neighbours in sUk).
    double eacc = 0.0;
    float uk1 = 0.f, u1xk1 = 0.f, u1yk1 = 0.f;   // u_{k+1} of the tile pixel
    if (tp) {
        const int ix = tx + 1, iy = ty + 1;
        const float u = sUk[0][iy][ix], u1x = sUk[1][iy][ix], u1y = sUk[2][iy][ix];
        const float du = u - u0t;
        const double dud = (double)u - (double)u0t;
        float ps = 0.f;
        double ed = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                if (ENERGY) ed += fabs((double)itv[j] + dud * (double)iuv[j]);
                rv[j] = r_update(rv[j], itv[j], du, iuv[j], sl);
                ps = ps_add(ps, rv[j], iuv[j]);
            }
        }
        if (ENERGY && x < W && y < H && chk->mask[gt] != 0.f)
            eacc = energy_point(u, sUk[0][iy][ix + 1], sUk[0][iy + 1][ix], u1x, sUk[1][iy][ix + 1], sUk[1][iy + 1][ix],
                                u1y, sUk[2][iy][ix + 1], sUk[2][iy + 1][ix], t11, t12, t22, ed, chk->lambda,
                                chk->alpha0, chk->alpha1);
        const int bx = ix + 1, by = iy + 1;               // B index of this position
        const int lx = x > 0 ? bx - 1 : bx, uy = y > 0 ? by - 1 : by;
        float un, u1xn, u1yn;
        u_update(u, u1x, u1y, t11, t12, t22, sP[0][by][bx], sP[1][by][bx], sP[0][by][lx], sP[1][by][lx],
                 sP[0][uy][bx], sP[1][uy][bx], sQ[0][by][bx], sQ[0][by][lx], sQ[1][by][bx], sQ[1][uy][bx],
                 sQ[2][by][bx], sQ[2][by][lx], sQ[2][uy][bx], ps, a, un, u1xn, u1yn);
        sA[0][iy + 1][ix + 1] = 2.f * un - u;
        sA[1][iy + 1][ix + 1] = 2.f * u1xn - u1x;
        sA[2][iy + 1][ix + 1] = 2.f * u1yn - u1y;
        uk1 = un; u1xk1 = u1xn; u1yk1 = u1yn;
    }
    // the ring: on the threads past the tile if there are any, else after the tile pixel
    constexpr int RING = CX * CY - TX * TY;
    constexpr int RT = NT > TP ? NT - TP : NT;   // ring threads
    for (int k = NT > TP ? tid - TP : tid; k >= 0 && k < RING; k += RT) {
        int ix, iy;
        ring_pos<CX, CY, TY>(k, ix, iy);
        const int gx = x0 - 1 + ix, gy = y0 - 1 + iy;
        const float u = sUk[0][iy][ix], u1x = sUk[1][iy][ix], u1y = sUk[2][iy][ix];
        const size_t g = (size_t)clampi(gy, H - 1) * W + clampi(gx, W - 1);
        const float du = u - a.u0[g];
        float ps = 0.f;
        for (int j = 0; j < a.N; ++j) {
            const size_t gj = (size_t)j * n + g;
            const float iu = a.Iu[gj];
            ps = ps_add(ps, r_update(a.rcur[gj], a.It[gj], du, iu, sl), iu);
        }
        const int bx = ix + 1, by = iy + 1;
        const int lx = gx > 0 ? bx - 1 : bx, uy = gy > 0 ? by - 1 : by;
        float un, u1xn, u1yn;
        u_update(u, u1x, u1y, sT[0][by][bx], sT[1][by][bx], sT[2][by][bx], sP[0][by][bx], sP[1][by][bx],
                 sP[0][by][lx], sP[1][by][lx], sP[0][uy][bx], sP[1][uy][bx], sQ[0][by][bx], sQ[0][by][lx],
                 sQ[1][by][bx], sQ[1][uy][bx], sQ[2][by][bx], sQ[2][by][lx], sQ[2][uy][bx], ps, a, un, u1xn, u1yn);
        sA[0][iy + 1][ix + 1] = 2.f * un - u;
        sA[1][iy + 1][ix + 1] = 2.f * u1xn - u1x;
        sA[2][iy + 1][ix + 1] = 2.f * u1yn - u1y;
    }
    __syncthreads();

    // D: p, q of iteration k+2 (in place over p, q_{k+1} at the same positions)
    for (int i = tid; i < DX * DY; i += NT) {
        const int iy = i / DX, ix = i - iy * DX;
        const int gx = x0 - 1 + ix, gy = y0 - 1 + iy;
        const size_t g = (size_t)clampi(gy, H - 1) * W + clampi(gx, W - 1);
        const int ax = ix + 1, ay = iy + 1;               // A index (ubar_{k+1})
        const int rx = gx < W - 1 ? ax + 1 : ax, dy = gy < H - 1 ? ay + 1 : ay;
        const int bx = ix + 1, by = iy + 1;               // B index
        float px = sP[0][by][bx], py = sP[1][by][bx];
        float qx = sQ[0][by][bx], qy = sQ[1][by][bx], qz = sQ[2][by][bx];
        pq_update(px, py, qx, qy, qz, sT[0][by][bx], sT[1][by][bx], sT[2][by][bx], sA[0][ay][ax], sA[0][ay][rx],
                  sA[0][dy][ax], sA[1][ay][ax], sA[1][ay][rx], sA[1][dy][ax], sA[2][ay][ax], sA[2][ay][rx],
                  sA[2][dy][ax], as1, as0);
        sP[0][by][bx] = px; sP[1][by][bx] = py;
        sQ[0][by][bx] = qx; sQ[1][by][bx] = qy; sQ[2][by][bx] = qz;
        if (ix > 0 && iy > 0 && gx < W && gy < H) {
            a.pnext2[g] = px; a.pnext2[n + g] = py;
            a.qnext2[g] = qx; a.qnext2[n + g] = qy; a.qnext2[2 * n + g] = qz;
        }
    }
    __syncthreads();

    // tile: r and u of iteration k+2
    if (tp && x < W && y < H) {
        const size_t g = (size_t)y * W + x;
        const float u = uk1, u1x = u1xk1, u1y = u1yk1;
        const float du = u - u0t;
        float ps = 0.f;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                const float rr = r_update(rv[j], itv[j], du, iuv[j], sl);
                a.rnext2[(size_t)j * n + g] = rr;
                ps = ps_add(ps, rr, iuv[j]);
            }
        }
        const int bx = tx + 2, by = ty + 2;               // B index of this pixel (D index + 1)
        const int lx = x > 0 ? bx - 1 : bx, uy = y > 0 ? by - 1 : by;
        float un, u1xn, u1yn;
        u_update(u, u1x, u1y, t11, t12, t22, sP[0][by][bx], sP[1][by][bx], sP[0][by][lx], sP[1][by][lx],
                 sP[0][uy][bx], sP[1][uy][bx], sQ[0][by][bx], sQ[0][by][lx], sQ[1][by][bx], sQ[1][uy][bx],
                 sQ[2][by][bx], sQ[2][by][lx], sQ[2][uy][bx], ps, a, un, u1xn, u1yn);
        a.unext2[g] = un; a.unext2[n + g] = u1xn; a.unext2[2 * n + g] = u1yn;
        a.unext[g] = u; a.unext[n + g] = u1x; a.unext[2 * n + g] = u1y;
    }
    if (ENERGY) {
        const double e = block_sum<NT>(eacc);
        if (tid == 0) chk->partials[tile] = e;
    }
}

// Three iterations (k -> k+3) of a 64 x TY tile in one launch: iteration
// k+1 runs on the tile plus a two-pixel ring (C0), k+2 on the tile plus a
// one-pixel ring (C1), k+3 on the tile.  Regions, as offsets from the tile
// origin (positions outside the image read clamped coordinates and never
// feed positions inside it, as in tgv_iter2_tile):
//   A  ubar_k, then ubar_{k+1} (on C0), ubar_{k+2} (on C1)   -3 .. +TX+2
//   B  T and p, q: p, q of k+1 on all of B, of k+2 on -2.., of k+3 on -1..
//                                                              -3 .. +TX+1
//   U  u_k (C0)                                                -2 .. +TX+1
// One tile pixel per thread (registers: r, It, Iu, u0, T through all three
// iterations); one ring position per thread besides (the C1 ring first, its
// threads keep r_{k+1} and u_{k+1} for iteration k+2, then the outer ring
// of C0).  Same per-position arithmetic as tgv_iter_tile, so identical
// results.  Only u_{k+2}, u_{k+3}, p, q, r of k+3 reach HBM.
template <int NS, bool ENERGY = false, int TY = 10>
__device__ __forceinline__ void tgv_iter3_tile(const TgvPtrs& a, const int tile, const TgvCheck* chk = nullptr) {
    constexpr int TX = TG_TX, NT = TG_TX * TY;
    constexpr int AX = TX + 6, AY = TY + 6;
    constexpr int BX = TX + 5, BY = TY + 5;
    constexpr int UX = TX + 4, UY = TY + 4;
    constexpr int R1 = 2 * (TX + 2) + 2 * TY;          // C1 ring
    constexpr int R0 = 2 * (TX + 4) + 2 * (TY + 2);    // C0 outer ring
    static_assert(R1 + R0 <= NT, "one ring position per thread");
    __shared__ float sA[3][AY][AX];   // ubar; A index = U index + 1 = B index
    __shared__ float sU[3][UY][UX];   // u_k
    __shared__ float sP[2][BY][BX];
    __shared__ float sQ[3][BY][BX];
    __shared__ float sT[3][BY][BX];

    const int W = a.W, H = a.H;
    const size_t n = (size_t)W * H;
    const int tiles_x = (W + TX - 1) / TX;
    const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
    const int tid = threadIdx.x;
    const float as1 = a.alpha1 * a.sigma, as0 = a.alpha0 * a.sigma;
    const float sl = a.sigma * a.lambda;

    // ---- loads: T, p_k, q_k on B; ubar_k on A; u_k on U; the tile pixel
    for (int i = tid; i < BX * BY; i += NT) {
        const int iy = i / BX, ix = i - iy * BX;
        const size_t g = (size_t)clampi(y0 - 3 + iy, H - 1) * W + clampi(x0 - 3 + ix, W - 1);
#pragma unroll
        for (int f = 0; f < 3; ++f) sT[f][iy][ix] = a.T[f * n + g];
        sP[0][iy][ix] = a.pcur[g];
        sP[1][iy][ix] = a.pcur[n + g];
#pragma unroll
        for (int f = 0; f < 3; ++f) sQ[f][iy][ix] = a.qcur[f * n + g];
    }
    for (int i = tid; i < AX * AY; i += NT) {
        const int iy = i / AX, ix = i - iy * AX;
        const size_t g = (size_t)clampi(y0 - 3 + iy, H - 1) * W + clampi(x0 - 3 + ix, W - 1);
#pragma unroll
        for (int f = 0; f < 3; ++f) sA[f][iy][ix] = 2.f * a.ucur[f * n + g] - a.uprev[f * n + g];
    }
    for (int i = tid; i < UX * UY; i += NT) {
        const int iy = i / UX, ix = i - iy * UX;
        const size_t g = (size_t)clampi(y0 - 2 + iy, H - 1) * W + clampi(x0 - 2 + ix, W - 1);
#pragma unroll
        for (int f = 0; f < 3; ++f) sU[f][iy][ix] = a.ucur[f * n + g];
    }
    const int ty = tid / TX, tx = tid - (tid / TX) * TX;
    const int x = x0 + tx, y = y0 + ty;
    const size_t gt = (size_t)clampi(y, H - 1) * W + clampi(x, W - 1);
    float rv[NS], itv[NS], iuv[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        if (j < a.N) {
            const size_t gj = (size_t)j * n + gt;
            rv[j] = a.rcur[gj];
            itv[j] = a.It[gj];
            iuv[j] = a.Iu[gj];
        }
    }
    const float u0t = a.u0[gt];
    const float t11 = a.T[gt], t12 = a.T[n + gt], t22 = a.T[2 * n + gt];
    __syncthreads();

    // ---- energy of state k at the tile pixel (energy launches)
    double eacc = 0.0;
    if (ENERGY && x < W && y < H && chk->mask[gt] != 0.f) {
        const int ux = tx + 2, uy = ty + 2;
        const float u = sU[0][uy][ux];
        const double dud = (double)u - (double)u0t;
        double ed = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j)
            if (j < a.N) ed += fabs((double)itv[j] + dud * (double)iuv[j]);
        eacc = energy_point(u, sU[0][uy][ux + 1], sU[0][uy + 1][ux], sU[1][uy][ux], sU[1][uy][ux + 1],
                            sU[1][uy + 1][ux], sU[2][uy][ux], sU[2][uy][ux + 1], sU[2][uy + 1][ux], t11, t12, t22, ed,
                            chk->lambda, chk->alpha0, chk->alpha1);
    }

    // ---- p, q of k+1 on B (ubar_k on A, same offset)
    for (int i = tid; i < BX * BY; i += NT) {
        const int iy = i / BX, ix = i - iy * BX;
        const int gx = x0 - 3 + ix, gy = y0 - 3 + iy;
        const int rx = gx < W - 1 ? ix + 1 : ix, dy = gy < H - 1 ? iy + 1 : iy;
        float px = sP[0][iy][ix], py = sP[1][iy][ix];
        float qx = sQ[0][iy][ix], qy = sQ[1][iy][ix], qz = sQ[2][iy][ix];
        pq_update(px, py, qx, qy, qz, sT[0][iy][ix], sT[1][iy][ix], sT[2][iy][ix], sA[0][iy][ix], sA[0][iy][rx],
                  sA[0][dy][ix], sA[1][iy][ix], sA[1][iy][rx], sA[1][dy][ix], sA[2][iy][ix], sA[2][iy][rx],
                  sA[2][dy][ix], as1, as0);
        sP[0][iy][ix] = px; sP[1][iy][ix] = py;
        sQ[0][iy][ix] = qx; sQ[1][iy][ix] = qy; sQ[2][iy][ix] = qz;
    }
    __syncthreads();

    // ---- r, u of k+1 on C0 and ubar_{k+1} into A (ubar_k is no longer read)
    float uk1[3];                      // tile pixel's u_{k+1}
    {
        const int ux = tx + 2, uy = ty + 2;
        const float u = sU[0][uy][ux], u1x = sU[1][uy][ux], u1y = sU[2][uy][ux];
        const float du = u - u0t;
        float ps = 0.f;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                rv[j] = r_update(rv[j], itv[j], du, iuv[j], sl);
                ps = ps_add(ps, rv[j], iuv[j]);
            }
        }
        const int bx = ux + 1, by = uy + 1;
        const int lx = x > 0 ? bx - 1 : bx, uyb = y > 0 ? by - 1 : by;
        u_update(u, u1x, u1y, t11, t12, t22, sP[0][by][bx], sP[1][by][bx], sP[0][by][lx], sP[1][by][lx],
                 sP[0][uyb][bx], sP[1][uyb][bx], sQ[0][by][bx], sQ[0][by][lx], sQ[1][by][bx], sQ[1][uyb][bx],
                 sQ[2][by][bx], sQ[2][by][lx], sQ[2][uyb][bx], ps, a, uk1[0], uk1[1], uk1[2]);
        sA[0][by][bx] = 2.f * uk1[0] - u;
        sA[1][by][bx] = 2.f * uk1[1] - u1x;
        sA[2][by][bx] = 2.f * uk1[2] - u1y;
    }
    // the ring position of this thread (C1 ring for tid < R1, then the outer ring)
    int rux = 0, ruy = 0;
    const bool ring = tid < R1 + R0;
    if (tid < R1) {
        ring_pos<TX + 2, TY + 2, TY>(tid, rux, ruy);
        rux += 1;
        ruy += 1;
    } else if (ring) {
        ring_pos<TX + 4, TY + 4, TY + 2>(tid - R1, rux, ruy);
    }
    const int rgx = x0 - 2 + rux, rgy = y0 - 2 + ruy;
    const size_t rg = (size_t)clampi(rgy, H - 1) * W + clampi(rgx, W - 1);
    float rr1[NS], ru1[3] = {0.f, 0.f, 0.f}, ru0 = 0.f;
    if (ring) {
        const float u = sU[0][ruy][rux], u1x = sU[1][ruy][rux], u1y = sU[2][ruy][rux];
        ru0 = a.u0[rg];
        const float du = u - ru0;
        float ps = 0.f;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                const size_t gj = (size_t)j * n + rg;
                const float iu = a.Iu[gj];
                rr1[j] = r_update(a.rcur[gj], a.It[gj], du, iu, sl);
                ps = ps_add(ps, rr1[j], iu);
            }
        }
        const int bx = rux + 1, by = ruy + 1;
        const int lx = rgx > 0 ? bx - 1 : bx, uyb = rgy > 0 ? by - 1 : by;
        u_update(u, u1x, u1y, sT[0][by][bx], sT[1][by][bx], sT[2][by][bx], sP[0][by][bx], sP[1][by][bx],
                 sP[0][by][lx], sP[1][by][lx], sP[0][uyb][bx], sP[1][uyb][bx], sQ[0][by][bx], sQ[0][by][lx],
                 sQ[1][by][bx], sQ[1][uyb][bx], sQ[2][by][bx], sQ[2][by][lx], sQ[2][uyb][bx], ps, a, ru1[0], ru1[1],
                 ru1[2]);
        sA[0][by][bx] = 2.f * ru1[0] - u;
        sA[1][by][bx] = 2.f * ru1[1] - u1x;
        sA[2][by][bx] = 2.f * ru1[2] - u1y;
    }
    __syncthreads();

    // ---- p, q of k+2 on B indices [1, BX-1) x [1, BY-1) (ubar_{k+1})
    for (int i = tid; i < (BX - 2) * (BY - 2); i += NT) {
        const int iy = i / (BX - 2) + 1, ix = i - (i / (BX - 2)) * (BX - 2) + 1;
        const int gx = x0 - 3 + ix, gy = y0 - 3 + iy;
        const int rx = gx < W - 1 ? ix + 1 : ix, dy = gy < H - 1 ? iy + 1 : iy;
        float px = sP[0][iy][ix], py = sP[1][iy][ix];
        float qx = sQ[0][iy][ix], qy = sQ[1][iy][ix], qz = sQ[2][iy][ix];
        pq_update(px, py, qx, qy, qz, sT[0][iy][ix], sT[1][iy][ix], sT[2][iy][ix], sA[0][iy][ix], sA[0][iy][rx],
                  sA[0][dy][ix], sA[1][iy][ix], sA[1][iy][rx], sA[1][dy][ix], sA[2][iy][ix], sA[2][iy][rx],
                  sA[2][dy][ix], as1, as0);
        sP[0][iy][ix] = px; sP[1][iy][ix] = py;
        sQ[0][iy][ix] = qx; sQ[1][iy][ix] = qy; sQ[2][iy][ix] = qz;
    }
    __syncthreads();

    // ---- r, u of k+2 on C1 and ubar_{k+2} into A
    float uk2[3];
    {
        const int ux = tx + 2, uy = ty + 2;
        const float du = uk1[0] - u0t;
        float ps = 0.f;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                rv[j] = r_update(rv[j], itv[j], du, iuv[j], sl);
                ps = ps_add(ps, rv[j], iuv[j]);
            }
        }
        const int bx = ux + 1, by = uy + 1;
        const int lx = x > 0 ? bx - 1 : bx, uyb = y > 0 ? by - 1 : by;
        u_update(uk1[0], uk1[1], uk1[2], t11, t12, t22, sP[0][by][bx], sP[1][by][bx], sP[0][by][lx], sP[1][by][lx],
                 sP[0][uyb][bx], sP[1][uyb][bx], sQ[0][by][bx], sQ[0][by][lx], sQ[1][by][bx], sQ[1][uyb][bx],
                 sQ[2][by][bx], sQ[2][by][lx], sQ[2][uyb][bx], ps, a, uk2[0], uk2[1], uk2[2]);
        sA[0][by][bx] = 2.f * uk2[0] - uk1[0];
        sA[1][by][bx] = 2.f * uk2[1] - uk1[1];
        sA[2][by][bx] = 2.f * uk2[2] - uk1[2];
    }
    if (tid < R1) {
        const float du = ru1[0] - ru0;
        float ps = 0.f;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                const float iu = a.Iu[(size_t)j * n + rg];
                ps = ps_add(ps, r_update(rr1[j], a.It[(size_t)j * n + rg], du, iu, sl), iu);
            }
        }
        const int bx = rux + 1, by = ruy + 1;
        const int lx = rgx > 0 ? bx - 1 : bx, uyb = rgy > 0 ? by - 1 : by;
        float un, u1xn, u1yn;
        u_update(ru1[0], ru1[1], ru1[2], sT[0][by][bx], sT[1][by][bx], sT[2][by][bx], sP[0][by][bx], sP[1][by][bx],
                 sP[0][by][lx], sP[1][by][lx], sP[0][uyb][bx], sP[1][uyb][bx], sQ[0][by][bx], sQ[0][by][lx],
                 sQ[1][by][bx], sQ[1][uyb][bx], sQ[2][by][bx], sQ[2][by][lx], sQ[2][uyb][bx], ps, a, un, u1xn,
                 u1yn);
        sA[0][by][bx] = 2.f * un - ru1[0];
        sA[1][by][bx] = 2.f * u1xn - ru1[1];
        sA[2][by][bx] = 2.f * u1yn - ru1[2];
    }
    __syncthreads();

    // ---- p, q of k+3 on B indices [2, BX-2) x [2, BY-2) (ubar_{k+2}); the tile's go to HBM
    for (int i = tid; i < (BX - 4) * (BY - 4); i += NT) {
        const int iy = i / (BX - 4) + 2, ix = i - (i / (BX - 4)) * (BX - 4) + 2;
        const int gx = x0 - 3 + ix, gy = y0 - 3 + iy;
        const int rx = gx < W - 1 ? ix + 1 : ix, dy = gy < H - 1 ? iy + 1 : iy;
        float px = sP[0][iy][ix], py = sP[1][iy][ix];
        float qx = sQ[0][iy][ix], qy = sQ[1][iy][ix], qz = sQ[2][iy][ix];
        pq_update(px, py, qx, qy, qz, sT[0][iy][ix], sT[1][iy][ix], sT[2][iy][ix], sA[0][iy][ix], sA[0][iy][rx],
                  sA[0][dy][ix], sA[1][iy][ix], sA[1][iy][rx], sA[1][dy][ix], sA[2][iy][ix], sA[2][iy][rx],
                  sA[2][dy][ix], as1, as0);
        sP[0][iy][ix] = px; sP[1][iy][ix] = py;
        sQ[0][iy][ix] = qx; sQ[1][iy][ix] = qy; sQ[2][iy][ix] = qz;
        if (ix >= 3 && iy >= 3 && gx < W && gy < H) {
            const size_t g = (size_t)gy * W + gx;
            a.pnext3[g] = px; a.pnext3[n + g] = py;
            a.qnext3[g] = qx; a.qnext3[n + g] = qy; a.qnext3[2 * n + g] = qz;
        }
    }
    __syncthreads();

    // ---- tile: r and u of k+3
    if (x < W && y < H) {
        const size_t g = (size_t)y * W + x;
        const float du = uk2[0] - u0t;
        float ps = 0.f;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < a.N) {
                const float rr = r_update(rv[j], itv[j], du, iuv[j], sl);
                a.rnext3[(size_t)j * n + g] = rr;
                ps = ps_add(ps, rr, iuv[j]);
            }
        }
        const int bx = tx + 3, by = ty + 3;
        const int lx = x > 0 ? bx - 1 : bx, uyb = y > 0 ? by - 1 : by;
        float un, u1xn, u1yn;
        u_update(uk2[0], uk2[1], uk2[2], t11, t12, t22, sP[0][by][bx], sP[1][by][bx], sP[0][by][lx], sP[1][by][lx],
                 sP[0][uyb][bx], sP[1][uyb][bx], sQ[0][by][bx], sQ[0][by][lx], sQ[1][by][bx], sQ[1][uyb][bx],
                 sQ[2][by][bx], sQ[2][by][lx], sQ[2][uyb][bx], ps, a, un, u1xn, u1yn);
        a.unext3[g] = un; a.unext3[n + g] = u1xn; a.unext3[2 * n + g] = u1yn;
        a.unext2[g] = uk2[0]; a.unext2[n + g] = uk2[1]; a.unext2[2 * n + g] = uk2[2];
    }
    if (ENERGY) {
        const double e = block_sum<NT>(eacc);
        if (tid == 0) chk->partials[tile] = e;
    }
}

template <int NS, int TY>
__global__ __launch_bounds__(TG_TX * TY) void tgv_iter3_kernel(TgvPtrs a) {
    if (a.ctrl->done) return;
    const int ntiles = ((a.W + TG_TX - 1) / TG_TX) * ((a.H + TY - 1) / TY);
    tgv_iter3_tile<NS, false, TY>(a, xcd_tile(blockIdx.x, ntiles));
}

template <int NS, int TY, bool ENERGY, bool FINAL>
__global__ __launch_bounds__(TG_TX * TY) void tgv_iter3_check_kernel(TgvPtrs a, TgvCheck c) {
    if (a.ctrl->done) return;
    const int ntiles = ((a.W + TG_TX - 1) / TG_TX) * ((a.H + TY - 1) / TY);
    if (FINAL && (int)blockIdx.x == ntiles) {
        (void)check_finalize<TG_TX * TY>(c);
        return;
    }
    tgv_iter3_tile<NS, ENERGY, TY>(a, xcd_tile(blockIdx.x, ntiles), &c);
}

template <int NS, int TY, int NT = TG_TX * TY>
__global__ __launch_bounds__(NT) void tgv_iter2_kernel(TgvPtrs a) {
    if (a.ctrl->done) return;
    const int ntiles = ((a.W + TG_TX - 1) / TG_TX) * ((a.H + TY - 1) / TY);
    tgv_iter2_tile<NS, false, TY, NT>(a, xcd_tile(blockIdx.x, ntiles));
}

// Two iterations from state k that also sum the energy of state k (ENERGY)
// and / or finalise the previous energy launch's sum in one extra
// workgroup (FINAL).
template <int NS, int TY, bool ENERGY, bool FINAL, int NT = TG_TX * TY>
__global__ __launch_bounds__(NT) void tgv_iter2_check_kernel(TgvPtrs a, TgvCheck c) {
    if (a.ctrl->done) return;
    const int ntiles = ((a.W + TG_TX - 1) / TG_TX) * ((a.H + TY - 1) / TY);
    if (FINAL && (int)blockIdx.x == ntiles) {
        (void)check_finalize<NT>(c);
        return;
    }
    tgv_iter2_tile<NS, ENERGY, TY, NT>(a, xcd_tile(blockIdx.x, ntiles), &c);
}

template <int TY, int NT, int NS>
__global__ __launch_bounds__(NT) void tgv_iter_kernel(TgvPtrs a) {
    if (a.ctrl->done) return;
    const int ntiles = ((a.W + TG_TX - 1) / TG_TX) * ((a.H + TY - 1) / TY);
    tgv_iter_tile<TY, NT, NS>(a, xcd_tile(blockIdx.x, ntiles));
}

// One iteration k -> k+1 that also sums the energy of state k (ENERGY)
// and / or finalises the previous energy launch's sum in one extra
// workgroup (FINAL).
template <int TY, int NT, int NS, bool ENERGY, bool FINAL>
__global__ __launch_bounds__(NT) void tgv_iter_check_kernel(TgvPtrs a, TgvCheck c) {
    if (a.ctrl->done) return;
    const int ntiles = ((a.W + TG_TX - 1) / TG_TX) * ((a.H + TY - 1) / TY);
    if (FINAL && (int)blockIdx.x == ntiles) {
        (void)check_finalize<NT>(c);
        return;
    }
    tgv_iter_tile<TY, NT, NS, ENERGY>(a, xcd_tile(blockIdx.x, ntiles), &c);
}

// ---- FP64 energy -------------------------------------------------------------

__device__ inline double block_reduce_sum(double v) {
    __shared__ double wsum[TG_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < TG_THREADS / 64; ++k) t += wsum[k];
    return t;
}

// Per-block partial sums of alpha1 |T(grad u - u1)| + alpha0 |eps(u1)|_F +
// lambda sum_j |It_j + (u - u0) Iu_j| with the iteration's clamped forward
// differences.
// Block bid of nblk: grid-stride over the frame's pixels; returns the block
// sum on thread 0.
__device__ inline double tgv_energy_block(const float* __restrict__ U, const float* __restrict__ T,
                                          const float* __restrict__ u0, const float* __restrict__ It,
                                          const float* __restrict__ Iu, const float* __restrict__ mask, int W, int H,
                                          int N, double lambda, double alpha0, double alpha1, int bid, int nblk) {
    const size_t n = (size_t)W * H;
    double e = 0.0;
    for (size_t i = (size_t)bid * TG_THREADS + threadIdx.x; i < n; i += (size_t)nblk * TG_THREADS) {
        const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
        const size_t ix = (size_t)y * W + min(x + 1, W - 1), iy = (size_t)min(y + 1, H - 1) * W + x;
        if (mask[i] == 0.f) continue;
        double ed = 0.0;
        const double du = (double)U[i] - (double)u0[i];
        for (int j = 0; j < N; ++j) ed += fabs((double)It[(size_t)j * n + i] + du * (double)Iu[(size_t)j * n + i]);
        e += energy_point(U[i], U[ix], U[iy], U[n + i], U[n + ix], U[n + iy], U[2 * n + i], U[2 * n + ix],
                          U[2 * n + iy], T[i], T[n + i], T[2 * n + i], ed, lambda, alpha0, alpha1);
    }
    return block_reduce_sum(e);
}

__global__ __launch_bounds__(TG_THREADS) void tgv_energy_kernel(const float* __restrict__ U, const float* __restrict__ T,
                                                             const float* __restrict__ u0, const float* __restrict__ It,
                                                             const float* __restrict__ Iu,
                                                             const float* __restrict__ mask, int W, int H, int N,
                                                             double lambda, double alpha0, double alpha1,
                                                             double* __restrict__ partials) {
    const double t = tgv_energy_block(U, T, u0, It, Iu, mask, W, H, N, lambda, alpha0, alpha1, blockIdx.x, gridDim.x);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Sum of the partials in a fixed order, then the stop test; returns (on
// thread 0) whether this check stopped the solve.
__device__ inline bool tgv_finalize_block(const double* __restrict__ partials, int np, TgvCtrl* ctrl, int k, double tol,
                                          double e_star) {
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += TG_THREADS) v += partials[i];
    const double E = block_reduce_sum(v);
    bool stopped = false;
    if (threadIdx.x == 0) {
        ctrl->energy = E;
        if (!ctrl->done && tgv_at_gap(E, e_star, tol)) {
            ctrl->done = 1;
            ctrl->iters_at_done = k;
            ctrl->energy_at_done = E;
            stopped = true;
        }
    }
    return stopped;
}

__global__ __launch_bounds__(TG_THREADS) void tgv_energy_finalize_kernel(const double* __restrict__ partials, int np,
                                                                      TgvCtrl* ctrl, int k, double tol, double e_star) {
    (void)tgv_finalize_block(partials, np, ctrl, k, tol, e_star);
}

constexpr int TG_ENERGY_BLOCKS = 1024;

__host__ __device__ inline TgvPtrs tgv_ptrs_at(float* B, const TgvLayout& L, int W, int H, int N, int k,
                                                const TgvCtrl* ctrl, float lambda, float alpha0, float alpha1,
                                                float tau, float sigma) {
    const size_t n = L.n;
    const size_t rn = (size_t)N * n;
    TgvPtrs p;
    p.ucur = B + L.U + (size_t)(k % TG_USLOTS) * 3 * n;
    p.uprev = B + L.U + (size_t)((k + TG_USLOTS - 1) % TG_USLOTS) * 3 * n;
    p.unext = B + L.U + (size_t)((k + 1) % TG_USLOTS) * 3 * n;
    p.unext2 = B + L.U + (size_t)((k + 2) % TG_USLOTS) * 3 * n;
    p.unext3 = B + L.U + (size_t)((k + 3) % TG_USLOTS) * 3 * n;
    p.pcur = B + L.P + (size_t)(k % TG_PSLOTS) * 2 * n;
    p.pnext = B + L.P + (size_t)((k + 1) % TG_PSLOTS) * 2 * n;
    p.pnext2 = B + L.P + (size_t)((k + 2) % TG_PSLOTS) * 2 * n;
    p.pnext3 = B + L.P + (size_t)((k + 3) % TG_PSLOTS) * 2 * n;
    p.qcur = B + L.Q + (size_t)(k % TG_PSLOTS) * 3 * n;
    p.qnext = B + L.Q + (size_t)((k + 1) % TG_PSLOTS) * 3 * n;
    p.qnext2 = B + L.Q + (size_t)((k + 2) % TG_PSLOTS) * 3 * n;
    p.qnext3 = B + L.Q + (size_t)((k + 3) % TG_PSLOTS) * 3 * n;
    p.rcur = B + L.r + (size_t)(k % TG_PSLOTS) * rn;
    p.rnext = B + L.r + (size_t)((k + 1) % TG_PSLOTS) * rn;
    p.rnext2 = B + L.r + (size_t)((k + 2) % TG_PSLOTS) * rn;
    p.rnext3 = B + L.r + (size_t)((k + 3) % TG_PSLOTS) * rn;
    p.It = B + L.It;
    p.Iu = B + L.Iu;
    p.T = B + L.T;
    p.u0 = B + L.u0;
    p.ctrl = ctrl;
    p.W = W; p.H = H; p.N = N;
    p.lambda = lambda; p.alpha0 = alpha0; p.alpha1 = alpha1; p.tau = tau; p.sigma = sigma;
    return p;
}

static inline TgvPtrs tgv_ptrs(psw_tgv* t, const psw_tgv_params* prm) {
    return tgv_ptrs_at(t->base, t->L, t->W, t->H, t->N, t->k, t->ctrl, (float)prm->lambda, (float)prm->alpha0,
                       (float)prm->alpha1, (float)prm->tau, (float)prm->sigma);
}

// Launch shape of the iteration kernel: 0 = 64x8 tile / 256 threads,
// 1 = 64x8 / 512, 2 = 64x4 / 256 (PSW_TGV_VARIANT overrides the default).
static int tgv_variant() {
    const char* e = getenv("PSW_TGV_VARIANT");
    const int x = e ? atoi(e) : 1;
    return (x < 0 || x > 2) ? 1 : x;
}

// Chunks of check_every iterations the single-frame solver keeps in flight
// before it waits for the oldest (PSW_TGV_INFLIGHT, 2..8; identical results:
// the chunks enqueued after the stop run as no-ops).  At 640x480 x 8 sources
// depths 2, 3, 4 and 6 measured the same (53.04-53.09 ms to the gap from the
// learned start): two chunks already keep the device busy.
constexpr int TG_MAX_INFLIGHT = 8;
static int tgv_inflight() {
    const char* e = getenv("PSW_TGV_INFLIGHT");
    const int v = e ? atoi(e) : 2;
    return v < 2 ? 2 : v > TG_MAX_INFLIGHT ? TG_MAX_INFLIGHT : v;
}

// Two iterations per launch (tgv_iter2_kernel) unless PSW_TGV_DOUBLE=0;
// identical results either way.
static bool tgv_double() {
    const char* e = getenv("PSW_TGV_DOUBLE");
    return !(e && atoi(e) == 0);
}

// Tile height of the two-iteration kernel (identical results): 10 (640
// threads, the default), 12 (768), 8 (512) or 4 (256), PSW_TGV2_TY.
static int tgv2_ty() {
    const char* e = getenv("PSW_TGV2_TY");
    const int v = e ? atoi(e) : 10;
    return v == 4 || v == 8 || v == 12 ? v : 10;
}

static int tgv_tiles2(const psw_tgv* t) {
    const int ty = tgv2_ty();
    return ((t->W + TG_TX - 1) / TG_TX) * ((t->H + ty - 1) / ty);
}

template <int NS, int TY, int NT>
static void tgv_launch_double_nt(const TgvPtrs& p, const TgvCheck* c, bool energy, bool fin, int ntiles,
                                 hipStream_t st) {
    const unsigned g = (unsigned)(ntiles + (fin ? 1 : 0));
    if (energy && fin) tgv_iter2_check_kernel<NS, TY, true, true, NT><<<g, NT, 0, st>>>(p, *c);
    else if (energy) tgv_iter2_check_kernel<NS, TY, true, false, NT><<<g, NT, 0, st>>>(p, *c);
    else if (fin) tgv_iter2_check_kernel<NS, TY, false, true, NT><<<g, NT, 0, st>>>(p, *c);
    else tgv_iter2_kernel<NS, TY, NT><<<g, NT, 0, st>>>(p);
}

template <int NS, int TY>
static void tgv_launch_double_ns(const TgvPtrs& p, const TgvCheck* c, bool energy, bool fin, int ntiles,
                                 hipStream_t st) {
    // One thread per tile pixel: 896 threads on the 64x10 tile (halo phases in
    // one pass, the ring beside the tile pixels; identical results) measured
    // 21.4 us per launch against 18.0 at 640x480 with 8 sources.
    tgv_launch_double_nt<NS, TY, TG_TX * TY>(p, c, energy, fin, ntiles, st);
}

template <int TY>
static void tgv_launch_double_ty(const TgvPtrs& p, int N, const TgvCheck* c, bool energy, bool fin, int ntiles,
                                 hipStream_t st) {
    if (N <= 8) tgv_launch_double_ns<8, TY>(p, c, energy, fin, ntiles, st);
    else tgv_launch_double_ns<T2_MAXN, TY>(p, c, energy, fin, ntiles, st);
}

// Two iterations in one launch; with a check (c non-NULL) the launch may
// also sum the energy of the current state and / or finalise a previous sum.
static int tgv_launch_double(psw_tgv* t, const psw_tgv_params* prm, hipStream_t st, const TgvCheck* c = nullptr,
                             bool energy = false, bool fin = false) {
    const TgvPtrs p = tgv_ptrs(t, prm);
    const int nt2 = tgv_tiles2(t);
    const int ty = tgv2_ty();
    if (ty == 4) tgv_launch_double_ty<4>(p, t->N, c, energy, fin, nt2, st);
    else if (ty == 8) tgv_launch_double_ty<8>(p, t->N, c, energy, fin, nt2, st);
    else if (ty == 12) tgv_launch_double_ty<12>(p, t->N, c, energy, fin, nt2, st);
    else tgv_launch_double_ty<10>(p, t->N, c, energy, fin, nt2, st);
    const int s = launch_status(energy || fin ? "tgv_iter2_check_kernel" : "tgv_iter2_kernel");
    if (s == PSW_OK) t->k += 2;
    return s;
}

// Three iterations per launch (tgv_iter3_kernel, 64 x 10 tiles of 640
// threads) where a launch has three iterations to do, with PSW_TGV_TRIPLE=1;
// identical results.  Off by default: at 640x480 with 8 sources it moves
// fewer bytes per iteration than tgv_iter2_kernel but takes 10.0 us per
// iteration against 9.05 (its extra halo work and a second round of
// scattered ring loads cost more VALU time than the HBM time it saves).
constexpr int T3_TY = 10;

static bool tgv_triple() {
    const char* e = getenv("PSW_TGV_TRIPLE");
    return e && atoi(e) == 1;
}

static int tgv_tiles3(const psw_tgv* t) {
    return ((t->W + TG_TX - 1) / TG_TX) * ((t->H + T3_TY - 1) / T3_TY);
}

// Most iterations one launch may do for this solver: 3, 2 or 1.
static int tgv_max_per_launch(const psw_tgv* t) {
    if (t->N > T2_MAXN || !tgv_double()) return 1;
    return tgv_triple() ? 3 : 2;
}

template <int NS>
static void tgv_launch_triple_ns(const TgvPtrs& p, const TgvCheck* c, bool energy, bool fin, int ntiles,
                                 hipStream_t st) {
    const unsigned g = (unsigned)(ntiles + (fin ? 1 : 0));
    constexpr int NT = TG_TX * T3_TY;
    if (energy && fin) tgv_iter3_check_kernel<NS, T3_TY, true, true><<<g, NT, 0, st>>>(p, *c);
    else if (energy) tgv_iter3_check_kernel<NS, T3_TY, true, false><<<g, NT, 0, st>>>(p, *c);
    else if (fin) tgv_iter3_check_kernel<NS, T3_TY, false, true><<<g, NT, 0, st>>>(p, *c);
    else tgv_iter3_kernel<NS, T3_TY><<<g, NT, 0, st>>>(p);
}

// Three iterations in one launch (check arguments as tgv_launch_double).
static int tgv_launch_triple(psw_tgv* t, const psw_tgv_params* prm, hipStream_t st, const TgvCheck* c = nullptr,
                             bool energy = false, bool fin = false) {
    const TgvPtrs p = tgv_ptrs(t, prm);
    const int nt3 = tgv_tiles3(t);
    if (t->N <= 8) tgv_launch_triple_ns<8>(p, c, energy, fin, nt3, st);
    else tgv_launch_triple_ns<T2_MAXN>(p, c, energy, fin, nt3, st);
    const int s = launch_status(energy || fin ? "tgv_iter3_check_kernel" : "tgv_iter3_kernel");
    if (s == PSW_OK) t->k += 3;
    return s;
}

// Iterations of the next launch when `left` remain and at most `mmax` fit
// one launch: three while they last, but 4 as 2 + 2 (no single launch).
static int tgv_launch_size(int left, int mmax) {
    int m = std::min(left, mmax);
    if (m == 3 && left == 4) m = 2;
    return m;
}

// Partials of the checking launches: one per tile of the smallest tile shape.
static int tgv_tile_partials(int W, int H) { return ((W + TG_TX - 1) / TG_TX) * ((H + 3) / 4); }

// One iteration launch of the solve loop: ENERGY sums the energy of the
// current state into c.partials, FINAL finalises c.fin_partials in one
// extra workgroup.
template <int TY, int NT, int NS>
static void tgv_launch_check_ns(const TgvPtrs& p, const TgvCheck& c, bool energy, bool fin, int ntiles,
                                hipStream_t st) {
    const unsigned g = (unsigned)(ntiles + (fin ? 1 : 0));
    if (energy && fin) tgv_iter_check_kernel<TY, NT, NS, true, true><<<g, NT, 0, st>>>(p, c);
    else if (energy) tgv_iter_check_kernel<TY, NT, NS, true, false><<<g, NT, 0, st>>>(p, c);
    else if (fin) tgv_iter_check_kernel<TY, NT, NS, false, true><<<g, NT, 0, st>>>(p, c);
    else tgv_iter_kernel<TY, NT, NS><<<g, NT, 0, st>>>(p);
}

template <int TY, int NT>
static void tgv_launch_check(const TgvPtrs& p, const TgvCheck& c, bool energy, bool fin, int ntiles, hipStream_t st) {
    if (p.N <= 8) tgv_launch_check_ns<TY, NT, 8>(p, c, energy, fin, ntiles, st);
    else tgv_launch_check_ns<TY, NT, PSW_MAX_SOURCES>(p, c, energy, fin, ntiles, st);
}

static int tgv_launch_one(psw_tgv* t, const psw_tgv_params* prm, const TgvCheck& c, bool energy, bool fin,
                          hipStream_t st) {
    const int var = tgv_variant();
    const int ty = var == 2 ? 4 : 8;
    const int ntiles = ((t->W + TG_TX - 1) / TG_TX) * ((t->H + ty - 1) / ty);
    const TgvPtrs p = tgv_ptrs(t, prm);
    if (var == 0) tgv_launch_check<8, 256>(p, c, energy, fin, ntiles, st);
    else if (var == 1) tgv_launch_check<8, 512>(p, c, energy, fin, ntiles, st);
    else tgv_launch_check<4, 256>(p, c, energy, fin, ntiles, st);
    const int s = launch_status(energy || fin ? "tgv_iter_check_kernel" : "tgv_iter_kernel");
    if (s == PSW_OK) t->k++;
    return s;
}

static int tgv_launch_iters(psw_tgv* t, int niters, const psw_tgv_params* prm, hipStream_t st) {
    const int var = tgv_variant();
    const int ty = var == 2 ? 4 : 8;
    const dim3 grid((unsigned)(((t->W + TG_TX - 1) / TG_TX) * ((t->H + ty - 1) / ty)));
    const int mmax = tgv_max_per_launch(t);
    while (niters >= 2 && mmax >= 2) {
        const int m = tgv_launch_size(niters, mmax);
        const int s = m == 3 ? tgv_launch_triple(t, prm, st) : tgv_launch_double(t, prm, st);
        if (s != PSW_OK) return s;
        niters -= m;
    }
    TgvCheck none{};
    for (int i = 0; i < niters; ++i) {
        const TgvPtrs p = tgv_ptrs(t, prm);
        if (var == 0) tgv_launch_check<8, 256>(p, none, false, false, (int)grid.x, st);
        else if (var == 1) tgv_launch_check<8, 512>(p, none, false, false, (int)grid.x, st);
        else tgv_launch_check<4, 256>(p, none, false, false, (int)grid.x, st);
        int s = launch_status("tgv_iter_kernel");
        if (s != PSW_OK) return s;
        t->k++;
    }
    return PSW_OK;
}

static int tgv_launch_energy(psw_tgv* t, const psw_tgv_params* prm, double tol, double e_star, hipStream_t st) {
    const TgvLayout& L = t->L;
    const float* U = t->base + L.U + (size_t)(t->k % TG_USLOTS) * 3 * L.n;
    int blocks = (int)((L.n + TG_THREADS - 1) / TG_THREADS);
    if (blocks > t->npartials) blocks = t->npartials;
    tgv_energy_kernel<<<blocks, TG_THREADS, 0, st>>>(U, t->base + L.T, t->base + L.u0, t->base + L.It, t->base + L.Iu,
                                                     t->base + L.mask, t->W, t->H, t->N, prm->lambda, prm->alpha0,
                                                     prm->alpha1, t->partials);
    int s = launch_status("tgv_energy_kernel");
    if (s != PSW_OK) return s;
    tgv_energy_finalize_kernel<<<1, TG_THREADS, 0, st>>>(t->partials, blocks, t->ctrl, t->k, tol, e_star);
    return launch_status("tgv_energy_finalize_kernel");
}

__global__ void tgv_ctrl_reset_kernel(TgvCtrl* c) {
    c->done = 0;
    c->iters_at_done = 0;
    c->energy = 0.0;
    c->energy_at_done = 0.0;
}

// ---- many frames per launch ---------------------------------------------------

// One frame of a batch as the device sees it; the iteration count of every
// frame is its own start k0 plus a batch counter.  The counter lives in two
// device copies: an iteration launch reads copy rd and (block 0) writes the
// advanced count to copy wr != rd, so no launch writes what its own blocks
// read, and a captured chunk of launches can be replayed; the finalising
// kernel of a chunk moves the count back to copy 0.
struct TgvFrameDev {
    float* base;
    TgvLayout L;
    int W, H, N, k0;
    TgvCtrl* ctrl;
    TgvCtrl* mirror;             // the solver's pinned, mapped host copy (device address)
    double* partials;
    int eblocks;
    int tiles2;                  // tiles of the two-iteration launches (their energy partials)
    float lambda, alpha0, alpha1, tau, sigma;
    double dlambda, dalpha0, dalpha1, e_star;
};

// The frame of block b: the number of frames whose block range ends at or
// before b (off[f] = first block of frame f, off[nf] = total).  Each lane
// tests one frame end and the wave counts them with a ballot, so a new
// workgroup finds its frame after one load round instead of a dependent
// binary search.
__device__ inline int batch_frame(const int* __restrict__ off, int nf, int b) {
    const int lane = (int)(threadIdx.x & 63);
    int f = 0;
    for (int base = 0; base < nf; base += 64) {
        const int j = base + lane;
        const bool past = j < nf && off[j + 1] <= b;
        const int c = __popcll(__ballot(past));
        f += c;
        if (c < 64) break;
    }
    return f;
}

template <int TY, int NT>
__global__ __launch_bounds__(NT) void tgv_iter_batch_kernel(const TgvFrameDev* __restrict__ fr,
                                                            const int* __restrict__ toff, int nf, int* kd, int rd,
                                                            int wr) {
    const int kb = kd[rd];
    if (blockIdx.x == 0 && threadIdx.x == 0) kd[wr] = kb + 1;
    const int vb = xcd_tile((int)blockIdx.x, toff[nf]);   // XCD-local tiles
    const int f = batch_frame(toff, nf, vb);
    const TgvFrameDev& d = fr[f];
    if (d.ctrl->done) return;
    const TgvPtrs a = tgv_ptrs_at(d.base, d.L, d.W, d.H, d.N, d.k0 + kb, d.ctrl, d.lambda, d.alpha0, d.alpha1, d.tau,
                                  d.sigma);
    tgv_iter_tile<TY, NT, PSW_MAX_SOURCES>(a, vb - toff[f]);
}

__device__ inline TgvCheck batch_check(const TgvFrameDev& d, double tol) {
    TgvCheck c;
    c.partials = d.partials;
    c.fin_partials = d.partials;
    c.fin_count = d.tiles2;
    c.fin_k = 0;
    c.ctrl = d.ctrl;
    c.mirror = d.mirror;
    c.mask = d.base + d.L.mask;
    c.lambda = d.dlambda;
    c.alpha0 = d.dalpha0;
    c.alpha1 = d.dalpha1;
    c.tol = tol;
    c.e_star = d.e_star;
    return c;
}

// The frames of a two-iteration batch launch travel in the kernel argument
// (up to TG_BATCH_MAXF; the solver groups frames accordingly): the compiler
// knows pointers read from kernel arguments to be global and emits global
// loads through them, while pointers read from a table in memory would be
// generic (FLAT accesses, which also hold up every LDS wait and cost
// registers), and a workgroup finds its frame without a memory round trip.
constexpr int TG_BATCH_MAXF = 8;

struct TgvBatchArgs {
    int nf;
    int toff[TG_BATCH_MAXF + 1];   // first tile of each frame, toff[nf] = all tiles
    TgvFrameDev fr[TG_BATCH_MAXF];
};

// Two iterations of every frame of A (64 x TY tiles).  ENERGY: the launch
// also sums each frame's energy of its state at the launch into per-tile
// partials (the frame's partial buffer, tile order, as the single-frame
// solve's energy launch); FINAL: one extra workgroup per frame finalises the
// previous launch's partials (the state two iterations back); a stop is
// also written to the frame's mapped host copy.
template <int NS, int TY, bool ENERGY, bool FINAL>
__global__ __launch_bounds__(TG_TX * TY) void tgv_iter2_batch_kernel(const TgvBatchArgs A, int* kd, int rd, int wr,
                                                                    double tol) {
    const int kb = kd[rd];
    if (blockIdx.x == 0 && threadIdx.x == 0) kd[wr] = kb + 2;
    const int ntiles = A.toff[A.nf];
    if (FINAL && (int)blockIdx.x >= ntiles) {
        const TgvFrameDev& d = A.fr[(int)blockIdx.x - ntiles];
        if (d.ctrl->done) return;
        TgvCheck c = batch_check(d, tol);
        c.fin_k = d.k0 + kb - 2;
        (void)check_finalize<TG_TX * TY>(c);
        return;
    }
    // XCD-local tiles (as the single-frame kernels): each XCD takes a
    // contiguous range of the launch's tiles, so halo rows stay in its L2
    const int vb = xcd_tile((int)blockIdx.x, ntiles);
    int f = 0;
#pragma unroll
    for (int j = 1; j < TG_BATCH_MAXF; ++j) f += (j < A.nf && A.toff[j] <= vb) ? 1 : 0;
    const TgvFrameDev& d = A.fr[f];
    if (d.ctrl->done) return;
    const TgvPtrs a = tgv_ptrs_at(d.base, d.L, d.W, d.H, d.N, d.k0 + kb, d.ctrl, d.lambda, d.alpha0, d.alpha1, d.tau,
                                  d.sigma);
    if constexpr (ENERGY) {
        const TgvCheck c = batch_check(d, tol);
        tgv_iter2_tile<NS, true, TY>(a, vb - A.toff[f], &c);
    } else {
        tgv_iter2_tile<NS, false, TY>(a, vb - A.toff[f]);
    }
}

// Moves the batch count back to copy 0 at the end of a chunk.
__global__ void tgv_batch_count_home_kernel(int* kd) { kd[0] = kd[1]; }

// kdev: the copy of the batch counter that holds the current count.
__global__ __launch_bounds__(TG_THREADS) void tgv_energy_batch_kernel(const TgvFrameDev* __restrict__ fr,
                                                                   const int* __restrict__ eoff, int nf,
                                                                   const int* __restrict__ kdev) {
    const int f = batch_frame(eoff, nf, blockIdx.x);
    const TgvFrameDev& d = fr[f];
    const int bid = (int)blockIdx.x - eoff[f];
    const size_t n = d.L.n;
    const float* U = d.base + d.L.U + (size_t)((d.k0 + *kdev) % TG_USLOTS) * 3 * n;
    const double t = tgv_energy_block(U, d.base + d.L.T, d.base + d.L.u0, d.base + d.L.It, d.base + d.L.Iu,
                                      d.base + d.L.mask, d.W, d.H, d.N, d.dlambda, d.dalpha0, d.dalpha1, bid,
                                      d.eblocks);
    if (threadIdx.x == 0) d.partials[bid] = t;
}

__global__ __launch_bounds__(TG_THREADS) void tgv_finalize_batch_kernel(const TgvFrameDev* __restrict__ fr,
                                                                     int* kd, int rd, double tol) {
    const TgvFrameDev& d = fr[blockIdx.x];
    const int kb = kd[rd];
    const bool stopped = tgv_finalize_block(d.partials, d.eblocks, d.ctrl, d.k0 + kb, tol, d.e_star);
    if (threadIdx.x == 0 && stopped) {
        volatile TgvCtrl* m = d.mirror;
        m->iters_at_done = d.ctrl->iters_at_done;
        m->energy_at_done = d.ctrl->energy_at_done;
        __threadfence_system();
        m->done = 1;
    }
    if (rd != 0 && blockIdx.x == 0 && threadIdx.x == 0) kd[0] = kb;   // the next chunk starts from copy 0
}

// Frames of a batch in groups, solved one group after another.  Batching
// pays for frames that cannot fill the GPU alone: a launch's tiles run as
// one round of resident workgroups (two 64x10 tiles per CU), and a small
// frame's launch leaves most CUs idle, so frames are packed until their
// tiles fill one round.  Frames that fill a round by themselves run alone:
// batching them gains nothing, and their state would no longer fit the
// Infinity Cache together (the iteration kernels stream the whole state
// every launch and run about 15 % slower per pixel once it does not fit:
// 29.4 ps per pixel-iteration at 640x480 with 8 sources, 146 MB of state,
// against 33.7 ps at 1920x1440, 1.3 GB).  A group also stays within
// PSW_TGV_BATCH_MB of state (default 192 MB of the MI355X's 256 MB; 0 = no
// limit) and TG_BATCH_MAXF frames.  Largest frames first, each into the
// first group with room.
static std::vector<std::vector<int>> tgv_batch_groups(psw_tgv* const* ts, int n) {
    const char* e = getenv("PSW_TGV_BATCH_MB");
    double budget = (e ? atof(e) : 192.0) * 1e6;
    if (budget <= 0.0) budget = 1e300;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    const int round = 2 * cus;                  // resident 64x10 workgroups
    const int ty = tgv2_ty() == 8 ? 8 : 10;
    std::vector<std::vector<int>> groups;
    std::vector<double> bytes((size_t)n, 0.0);
    std::vector<int> tiles((size_t)n, 0), order((size_t)n);
    for (int f = 0; f < n; ++f) {
        order[(size_t)f] = f;
        if (ts[f]) {
            bytes[(size_t)f] = (double)ts[f]->L.total * sizeof(float);
            tiles[(size_t)f] = ((ts[f]->W + TG_TX - 1) / TG_TX) * ((ts[f]->H + ty - 1) / ty);
        }
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bytes[(size_t)a] > bytes[(size_t)b]; });
    std::vector<double> used;
    std::vector<int> used_tiles;
    for (int f : order) {
        size_t g = 0;
        while (g < groups.size() &&
               (used[g] + bytes[(size_t)f] > budget || used_tiles[g] + tiles[(size_t)f] > round ||
                groups[g].size() >= TG_BATCH_MAXF))
            ++g;
        if (g == groups.size()) {
            groups.emplace_back();
            used.push_back(0.0);
            used_tiles.push_back(0);
        }
        groups[g].push_back(f);
        used[g] += bytes[(size_t)f];
        used_tiles[g] += tiles[(size_t)f];
    }
    for (auto& g : groups) std::sort(g.begin(), g.end());
    return groups;
}

// ---- resident solver --------------------------------------------------------
//
// One cooperative launch runs the whole solve (or a long psw_tgv_iterate):
// every 64 x 20 tile is owned by one workgroup of 640 threads (one per CU)
// for the whole launch, two pixels per thread (rows r and r + 10), and its
// state (u_k, u_{k-1}, p, q, r and the constants It, Iu, T, u0) stays in
// registers.  HBM is read once at the start and written once at the end.
// Per iteration a tile needs from its neighbours only the ubar of the pixels
// next to its border (and, before an energy check, u of the right / lower
// neighbours): every tile publishes its edge rows and columns as tagged
// 64-bit words (iteration in the high half, the float in the low half), and
// its neighbours poll those words until the tag matches, so no fence or grid
// barrier is involved.  The duals p, q of the one-pixel ring left of and
// above the tile (which the u step reads) are computed by the tile itself,
// from ring copies kept in LDS; they are bit-identical to the neighbour's
// own values.  Words are double-buffered by iteration parity: a tile cannot
// publish iteration j+2 before every neighbour has read its iteration j,
// because it needs their iteration j+1.
//
// The launch waits until every workgroup has started (spin-waiting
// neighbours need all tiles resident at once); if that does not happen in
// time the launch is called off before anything is written and the caller
// uses the launch sequence.  Energy checks use the two-iteration launches'
// partials (one per 64 x 10 half tile, per-wave butterflies and the waves in
// order, as a 640-thread block) and the same finalising sum (tile 0's
// workgroup), so a check gives the same bits and the solve stops at the same
// iteration with the same state.  Every wait is bounded; a wait that runs
// out sets an error word, after which every workgroup leaves at its next
// wait, and the host reports the failure.
constexpr int TR_SUB = 10;                 // the launch sequence's tile height (energy partials)
constexpr int TR_TY = 2 * TR_SUB, TR_NT = TG_TX * TR_SUB, TR_MAXN = 8, TR_MAXTILES = 256;
constexpr int TR_MIN_ITERS = 4;            // shorter psw_tgv_iterate calls use the launch sequence
// words per tile and parity: ubar (3 fields) on the top row, bottom row, left
// and right columns, then u (3 fields) on the top row and left column
constexpr int TR_UB_TOP = 0, TR_UB_BOT = 3 * TG_TX, TR_UB_LEFT = 6 * TG_TX, TR_UB_RIGHT = 6 * TG_TX + 3 * TR_TY;
constexpr int TR_U_TOP = 6 * TG_TX + 6 * TR_TY, TR_U_LEFT = 9 * TG_TX + 6 * TR_TY;
constexpr int TR_XW = 9 * TG_TX + 9 * TR_TY;
constexpr int TR_HALO_UB = 3 * (2 * TG_TX + 2 * TR_TY + 2);   // ubar words a tile reads per iteration
constexpr int TR_HALO_U = 3 * (TG_TX + TR_TY);                 // u words before a check
static_assert(TR_HALO_UB <= TR_NT && TR_HALO_U <= TR_NT, "one halo word per thread");

struct TgvResArgs {
    const float* ucur;               // state k0
    const float* uprev;
    const float* pcur;
    const float* qcur;
    const float* rcur;
    const float* It;
    const float* Iu;
    const float* T;
    const float* u0;
    const float* mask;
    float* U;                        // the slot arrays (the end state is stored by slot)
    float* P;
    float* Q;
    float* Rs;
    unsigned long long* xch;         // [2][ntiles][TR_XW]
    unsigned long long* part;        // [2][nsub][2] energy partials (two 32-bit halves)
    unsigned long long* dec;         // [2] stop decisions
    unsigned* err;                   // [0] error, [1] started, [2] go (1) / called off (2), [3..7] first failure:
                                     // tile, word offset from xch, awaited tag, seen tag, iteration
    TgvCtrl* ctrl;
    int W, H, N;
    int k0, niters, check_every;     // check_every 0: no checks
    int tiles_x, tiles_y, ntiles, sub_y, nsub;
    unsigned spin, start_spin;
    float lambda, alpha0, alpha1, tau, sigma;
    double dlambda, dalpha0, dalpha1, e_star;
    double gap;                      // tol * e_star (the stop test's bound, the same double product)
};

__device__ __forceinline__ unsigned long long tr_word(unsigned tag, unsigned bits) {
    return ((unsigned long long)tag << 32) | bits;
}

__device__ __forceinline__ void tr_put(unsigned long long* p, unsigned tag, float v) {
    __hip_atomic_store(p, tr_word(tag, __float_as_uint(v)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long tr_load(unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned tr_load32(unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Records the first failure (tile, word offset from the exchange words,
// awaited tag, seen tag, iteration) and sets the error word.
__device__ inline void tr_fail(const TgvResArgs& R, unsigned tile, unsigned word, unsigned tag, unsigned seen,
                               unsigned j) {
    unsigned expected = 0u;
    if (__hip_atomic_compare_exchange_strong(R.err, &expected, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)) {
        const unsigned vals[5] = {tile, word, tag, seen, j};
        for (int k = 0; k < 5; ++k) __hip_atomic_store(R.err + 3 + k, vals[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Poll up to K words (p[i] == nullptr: none) until each carries `tag`;
// their low halves go to bits[i].  The loads of one round are issued
// together.  At most R.spin rounds (0 fails every wait at once: the host's
// way to exercise the failure path); the error word, set by any wait that
// runs out, ends the others too.
template <int K>
__device__ inline bool tr_getn(unsigned long long* const p[K], unsigned tag, const TgvResArgs& R, unsigned bits[K],
                               int tile, int j) {
    bool ok[K];
    bool none = true;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        ok[i] = p[i] == nullptr;
        none = none && ok[i];
    }
    if (none) return true;   // nothing to wait for (also with a zero budget)
    for (unsigned it = 0; it < R.spin; ++it) {
        unsigned long long w[K];
#pragma unroll
        for (int i = 0; i < K; ++i) w[i] = ok[i] ? 0ull : tr_load(p[i]);
        bool all = true;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if (!ok[i] && (unsigned)(w[i] >> 32) == tag) {
                ok[i] = true;
                bits[i] = (unsigned)w[i];
            }
            all = all && ok[i];
        }
        if (all) return true;
        if ((it & 63) == 63 && tr_load32(R.err) != 0u) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    // some word is still missing (the early return above), so p[i] is set
    int i = 0;
    while (ok[i]) ++i;
    tr_fail(R, (unsigned)tile, (unsigned)(p[i] - R.xch), tag, (unsigned)(tr_load(p[i]) >> 32), (unsigned)j);
    return false;
}

// A halo word of the tile (tcx, tcy): w < TR_HALO_UB are ubar words (cell
// c = w / 3: the row above, the above-right corner, the column left, the
// below-left corner, the column right, the row below), the rest u words
// before a check (cells: the column right, the row below).  Returns false
// if the cell is outside the image or has no neighbour tile; else the
// source tile and word, the LDS cell (row, column of the ubar array, or of
// the u array) and the cell's image position.
__device__ inline bool tr_halo(const TgvResArgs& R, int tcx, int tcy, int w, int& src, int& word, int& ly, int& lx,
                               int& gx, int& gy) {
    constexpr int TX = TG_TX, TY = TR_TY;
    const int x0 = tcx * TX, y0 = tcy * TY;
    const int tile = tcy * R.tiles_x + tcx;
    const bool has_l = tcx > 0, has_r = tcx < R.tiles_x - 1, has_u = tcy > 0, has_d = tcy < R.tiles_y - 1;
    if (w < TR_HALO_UB) {
        const int c = w / 3, f = w - (w / 3) * 3;
        if (c < TX) {
            gx = x0 + c; gy = y0 - 1; src = tile - R.tiles_x; word = TR_UB_BOT + f * TX + c; ly = 0; lx = c + 1;
            return has_u && gx < R.W;
        }
        if (c == TX) {
            gx = x0 + TX; gy = y0 - 1; src = tile - R.tiles_x + 1; word = TR_UB_BOT + f * TX; ly = 0; lx = TX + 1;
            return has_u && has_r;
        }
        if (c < TX + 1 + TY) {
            const int r = c - (TX + 1);
            gx = x0 - 1; gy = y0 + r; src = tile - 1; word = TR_UB_RIGHT + f * TY + r; ly = r + 1; lx = 0;
            return has_l && gy < R.H;
        }
        if (c == TX + 1 + TY) {
            gx = x0 - 1; gy = y0 + TY; src = tile + R.tiles_x - 1; word = TR_UB_TOP + f * TX + TX - 1; ly = TY + 1;
            lx = 0;
            return has_l && has_d;
        }
        if (c < TX + 2 + 2 * TY) {
            const int r = c - (TX + 2 + TY);
            gx = x0 + TX; gy = y0 + r; src = tile + 1; word = TR_UB_LEFT + f * TY + r; ly = r + 1; lx = TX + 1;
            return has_r && gy < R.H;
        }
        const int cc = c - (TX + 2 + 2 * TY);
        gx = x0 + cc; gy = y0 + TY; src = tile + R.tiles_x; word = TR_UB_TOP + f * TX + cc; ly = TY + 1; lx = cc + 1;
        return has_d && gx < R.W;
    }
    const int wu = w - TR_HALO_UB, c = wu / 3, f = wu - (wu / 3) * 3;
    if (c < TY) {
        gx = x0 + TX; gy = y0 + c; src = tile + 1; word = TR_U_LEFT + f * TY + c; ly = c; lx = TX;
        return has_r && gy < R.H;
    }
    const int cc = c - TY;
    gx = x0 + cc; gy = y0 + TY; src = tile + R.tiles_x; word = TR_U_TOP + f * TX + cc; ly = TY; lx = cc;
    return has_d && gx < R.W;
}

// The ring position of ring thread rid: 0..63 the row above the tile,
// 64..83 the column left of it (false where there is no such neighbour, and
// for rid outside 0..83).  Its image position and its cell of the
// ring-indexed LDS arrays.
__device__ __forceinline__ bool tr_ring(const TgvResArgs& R, int tcx, int tcy, int tid, int& gx, int& gy, int& lx,
                                        int& ly) {
    const int x0 = tcx * TG_TX, y0 = tcy * TR_TY;
    if (tid < 0) return false;
    if (tid < TG_TX) {
        gx = x0 + tid; gy = y0 - 1; lx = tid + 1; ly = 0;
        return tcy > 0 && gx < R.W;
    }
    const int r = tid - TG_TX;
    gx = x0 - 1; gy = y0 + r; lx = 0; ly = r + 1;
    return r < TR_TY && tcx > 0 && gy < R.H;
}

// Waves of the resident workgroup with extra duties.  Ten waves share four
// SIMDs three, three, two and two, so the per-iteration extras go to the
// waves of the two lighter SIMDs: the ring's duals (waves 2 and 3), and the
// energy partials' stores; the stop decision's polls and sums are spread
// over waves 4..7.
constexpr int TR_RING0 = 2 * TG_TX;        // first thread of the ring (waves 2 and 3)
constexpr int TR_FIN0 = 4;                 // waves 4..7 sum the energy partials, a quarter each
constexpr int TR_PART_WAVE = 7;            // lanes 0 and 1 store the tile's energy partials
static_assert(2 * TR_MAXTILES <= 2 * FIN_THREADS && FIN_THREADS == 4 * 64, "partials: two per finalising thread");

// The energy of a check from every half tile's partial, summed as the
// launch sequence's finalising block sums them (threads below FIN_THREADS
// add partials i and i + FIN_THREADS, per-wave butterflies, then the wave
// sums in order): this wave does the butterfly of finalising wave q, whose
// sum goes to fin[q].  Polls the partials of check nchk.
__device__ inline bool tr_fin_quarter(const TgvResArgs& R, int q, int nchk, double* fin, int tile, int j) {
    const int l = threadIdx.x & 63;
    const int i0 = 64 * q + l, i1 = i0 + FIN_THREADS;
    unsigned long long* base = R.part + (size_t)(nchk & 1) * R.nsub * 2;
    unsigned long long* const pp[4] = {i0 < R.nsub ? base + 2 * i0 : nullptr, i0 < R.nsub ? base + 2 * i0 + 1 : nullptr,
                                       i1 < R.nsub ? base + 2 * i1 : nullptr, i1 < R.nsub ? base + 2 * i1 + 1 : nullptr};
    unsigned hl[4] = {0u, 0u, 0u, 0u};
    const bool ok = tr_getn<4>(pp, (unsigned)nchk, R, hl, tile, j);
    double v = 0.0;
    if (i0 < R.nsub) v += __longlong_as_double((long long)(((unsigned long long)hl[0] << 32) | hl[1]));
    if (i1 < R.nsub) v += __longlong_as_double((long long)(((unsigned long long)hl[2] << 32) | hl[3]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (l == 0) fin[q] = v;
    return ok;
}

// NS: the number of sources (the frame's N exactly).  Thread (wave w, lane
// tx) owns the pixels of rows 2w and 2w+1 in column tx: its own values, and
// the pair's vertical neighbours, come from registers; each image row is one
// wave slot in column order, so the energy's per-row butterflies are those
// of a 640-thread block over a 64 x 10 half tile.
// The per-tile work of a resident launch: `tile` is the tile of R's frame
// this workgroup owns; `start` counts the launch's workgroups in (start[0])
// and carries the go (start[1]: 1 go, 2 called off) once all nstart have.
// An energy check at iteration j saves the state j (u_{j-1}, p, q, r; u_j
// is still in registers one iteration later) in LDS, stores the tile's
// partials and goes on; every tile sums all partials during iteration j+1
// (the same bits as the launch sequence's finalising block) and, if the
// check stops the solve, writes out the saved state j.  So no iteration
// waits for the whole grid.  Needs check_every >= 2 (the host's rule): the
// next check's save must not come before this decision.
template <int NS>
__device__ __forceinline__ void tgv_resident_body(const TgvResArgs& R, const int tile, unsigned* start,
                                                  const unsigned nstart) {
    constexpr int TX = TG_TX, TY = TR_TY, SUB = TR_SUB;
    constexpr int NSAVE = 2 * (5 + NS);        // per thread: (px, py, qx, qy, qz, r) x 2 pixels
    __shared__ float sUb[3][TY + 2][TX + 2];   // ubar_j: ring row / column 0, own 1..TY / 1..TX, halo TY+1 / TX+1
    __shared__ float sPQ[5][TY + 1][TX + 1];   // p, q of j+1: ring 0, own 1..
    __shared__ float sUk[3][TY + 1][TX + 1];   // u_j before a check: own 0.., right / lower halo TY / TX
    __shared__ float sT[3][TY + 1][TX + 1];    // T: ring 0, own 1.. (as sPQ)
    __shared__ float sSave[NSAVE][TR_NT];      // the state of the check whose decision is pending (but u_{j-1})
    __shared__ float sU0[TY][TX];              // u0 of the own pixels
    // u_{j-1} of the pending check (v, v1x, v1y x 2 pixels), saved once the
    // check's energy has read sUk, in sUk's memory (the next check writes
    // sUk only after this one's decision)
    float* const sSaveV = &sUk[0][0][0];
    static_assert(6 * TR_NT <= 3 * (TY + 1) * (TX + 1), "sSaveV fits sUk");
    __shared__ double sWave[TY];               // per-row energy sums
    __shared__ double sFin[4];                 // the finalising waves' sums
    __shared__ int sFlag[1];                   // a wait failed

    const int W = R.W, H = R.H;
    const size_t n = (size_t)W * H;
    const int tcx = tile % R.tiles_x, tcy = tile / R.tiles_x;
    const int x0 = tcx * TX, y0 = tcy * TY;
    const int tid = threadIdx.x;
    const int tx = tid % TX;
    const int wv = __builtin_amdgcn_readfirstlane(tid / TX);   // pixels (tx, 2 wv) and (tx, 2 wv + 1) of the tile
    const int x = x0 + tx;
    const int rid = tid - TR_RING0;           // ring thread (waves 2 and 3)
    const float as1 = R.alpha1 * R.sigma, as0 = R.alpha0 * R.sigma, sl = R.sigma * R.lambda;
    if (tid == 0) {
        sFlag[0] = 0;
        // the last workgroup to start gives the go
        const unsigned before = __hip_atomic_fetch_add(start, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (before + 1 == nstart) {
            unsigned expected = 0u;
            __hip_atomic_compare_exchange_strong(start + 1, &expected, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    // own pixels: state k0 and the constants
    float u[2], u1x[2], u1y[2], v[2], v1x[2], v1y[2], px[2], py[2], qx[2], qy[2], qz[2];
    float rv[2][NS], itv[2][NS], iuv[2][NS];
    unsigned counted = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int y = y0 + 2 * wv + h;
        const size_t g = (size_t)min(y, H - 1) * W + min(x, W - 1);
        u[h] = R.ucur[g]; u1x[h] = R.ucur[n + g]; u1y[h] = R.ucur[2 * n + g];
        v[h] = R.uprev[g]; v1x[h] = R.uprev[n + g]; v1y[h] = R.uprev[2 * n + g];
        px[h] = R.pcur[g]; py[h] = R.pcur[n + g];
        qx[h] = R.qcur[g]; qy[h] = R.qcur[n + g]; qz[h] = R.qcur[2 * n + g];
        sT[0][2 * wv + h + 1][tx + 1] = R.T[g]; sT[1][2 * wv + h + 1][tx + 1] = R.T[n + g];
        sT[2][2 * wv + h + 1][tx + 1] = R.T[2 * n + g];
        sU0[2 * wv + h][tx] = R.u0[g];
        if (x < W && y < H && R.mask[g] != 0.f) counted |= 1u << h;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const size_t gs = (size_t)s * n + g;
            rv[h][s] = R.rcur[gs];
            itv[h][s] = R.It[gs];
            iuv[h][s] = R.Iu[gs];
        }
    }
    // the ring: p, q of state k0 into its sPQ cells (updated there in place
    // every iteration), T into sRT
    {
        int rgx, rgy, rix, riy;
        if (tr_ring(R, tcx, tcy, rid, rgx, rgy, rix, riy)) {
            const size_t gr = (size_t)rgy * W + rgx;
            sPQ[0][riy][rix] = R.pcur[gr]; sPQ[1][riy][rix] = R.pcur[n + gr];
            sPQ[2][riy][rix] = R.qcur[gr]; sPQ[3][riy][rix] = R.qcur[n + gr]; sPQ[4][riy][rix] = R.qcur[2 * n + gr];
            sT[0][riy][rix] = R.T[gr]; sT[1][riy][rix] = R.T[n + gr]; sT[2][riy][rix] = R.T[2 * n + gr];
        }
    }
    // the halo of state k0 from the state buffers (later states come from
    // the neighbours' words)
    int hoff = -1, hcell = 0, uoff = -1, ucell = 0;
    {
        int src, word, ly, lx, gx, gy;
        if (tid < TR_HALO_UB && tr_halo(R, tcx, tcy, tid, src, word, ly, lx, gx, gy)) {
            const int f = tid % 3;
            const size_t gh = (size_t)gy * W + gx;
            sUb[f][ly][lx] = 2.f * R.ucur[f * n + gh] - R.uprev[f * n + gh];
            hoff = src * TR_XW + word;
            hcell = (f * (TY + 2) + ly) * (TX + 2) + lx;
        }
        if (tid < TR_HALO_U && tr_halo(R, tcx, tcy, TR_HALO_UB + tid, src, word, ly, lx, gx, gy)) {
            uoff = src * TR_XW + word;
            ucell = ((tid % 3) * (TY + 1) + ly) * (TX + 1) + lx;
        }
    }
    // every workgroup resident before any waits on a neighbour
    if (tid == 0) {
        unsigned go = tr_load32(start + 1);
        for (unsigned it = 0; go == 0u && it < R.start_spin; ++it) {
            __builtin_amdgcn_s_sleep(2);
            go = tr_load32(start + 1);
        }
        if (go == 0u) {
            unsigned expected = 0u;
            __hip_atomic_compare_exchange_strong(start + 1, &expected, 2u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            go = tr_load32(start + 1);
        }
        if (go != 1u) sFlag[0] = 1;
    }
    __syncthreads();
    if (sFlag[0]) return;   // called off: nothing written

    // clamped neighbours (the reference's min(x+1, W-1) and max(x-1, 0))
    const bool has_r = x < W - 1, has_l = x > 0;
    const bool has_d0 = y0 + 2 * wv < H - 1, has_d1 = y0 + 2 * wv + 1 < H - 1, has_u0 = y0 + 2 * wv > 0;
    int j = 0, nchk = 0, jpend = -1;          // jpend: the check whose decision is pending
    int jnext = R.check_every > 0 ? R.check_every : R.niters;   // the next check
    bool failed = false, stopped = false;
    // The pending check's energy from the finalising waves' sums (after the
    // barrier that follows them), as the finalising block adds its waves'
    // (waves 4.. add nothing); tile 0 records it.  True if it stops the solve.
    auto stop_at_pending = [&]() {
        double E = 0.0;
#pragma unroll
        for (int w = 0; w < TR_NT / 64; ++w) E += w < 4 ? sFin[w] : 0.0;
        const bool stop = R.e_star > 0.0 && E - R.e_star <= R.gap;   // tgv_at_gap (ctrl was reset)
        if (tile == 0 && tid == 0) {
            TgvCtrl* d = R.ctrl;
            d->energy = E;
            if (stop) {
                d->iters_at_done = R.k0 + jpend;
                d->energy_at_done = E;
                d->done = 1;
            }
        }
        return stop;
    };
    for (;; ++j) {
        const bool check = j == jnext && j < R.niters;
        if (check) jnext += R.check_every;
        const bool decide = jpend >= 0 && j == jpend + 1;
        if (j == R.niters) {
            // a check at the last iteration: its decision before the end
            if (decide) {
                if (wv >= TR_FIN0 && wv < TR_FIN0 + 4 && !tr_fin_quarter(R, wv - TR_FIN0, nchk, sFin, tile, j))
                    sFlag[0] = 1;
                __syncthreads();
                if (sFlag[0]) failed = true;
                else stopped = stop_at_pending();
            }
            break;
        }
        if (check) {
            // the state j, written out instead of the last one if this check
            // stops the solve
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float st[5] = {px[h], py[h], qx[h], qy[h], qz[h]};
#pragma unroll
                for (int f = 0; f < 5; ++f) sSave[2 * f + h][tid] = st[f];
#pragma unroll
                for (int s = 0; s < NS; ++s) sSave[10 + 2 * s + h][tid] = rv[h][s];
            }
        }
        // r of j+1 and its sum with Iu need only the pixel itself: formed
        // before the wait for the neighbours' edges, in its shadow (after the
        // energy at a check, which leaves fewer values live there)
        float ps[2] = {0.f, 0.f};
        auto r_step = [&]() {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float du = u[h] - sU0[2 * wv + h][tx];
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    rv[h][s] = r_update(rv[h][s], itv[h][s], du, iuv[h][s], sl);
                    ps[h] = ps_add(ps[h], rv[h][s], iuv[h][s]);
                }
            }
        };
        if (!check) r_step();
        // ubar_j of the pixels (registers and LDS), u_j before a check
        float ub[2][3];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = 2 * wv + h;
            ub[h][0] = 2.f * u[h] - v[h];
            ub[h][1] = 2.f * u1x[h] - v1x[h];
            ub[h][2] = 2.f * u1y[h] - v1y[h];
#pragma unroll
            for (int f = 0; f < 3; ++f) sUb[f][row + 1][tx + 1] = ub[h][f];
            if (check) { sUk[0][row][tx] = u[h]; sUk[1][row][tx] = u1x[h]; sUk[2][row][tx] = u1y[h]; }
        }
        if (j > 0) {
            unsigned long long* blk = R.xch + (size_t)(j & 1) * R.ntiles * TR_XW;
            unsigned long long* const hp[3] = {hoff >= 0 ? blk + hoff : nullptr,
                                               check && uoff >= 0 ? blk + uoff : nullptr, nullptr};
            unsigned bits[3] = {0u, 0u, 0u};
            if (!tr_getn<3>(hp, (unsigned)j, R, bits, tile, j)) sFlag[0] = 1;
            if (hp[0]) (&sUb[0][0][0])[hcell] = __uint_as_float(bits[0]);
            if (hp[1]) (&sUk[0][0][0])[ucell] = __uint_as_float(bits[1]);
        }
        __syncthreads();
        if (sFlag[0]) { failed = true; break; }
        if (check) {
            // energy of state j: per-row butterflies (the rows of each 64 x 10
            // half tile are summed in order once every row is in)
            double e2[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int row = 2 * wv + h, y = y0 + row;
                e2[h] = 0.0;
                if (counted & (1u << h)) {
                    const double dud = (double)u[h] - (double)sU0[row][tx];
                    double ed = 0.0;
#pragma unroll
                    for (int s = 0; s < NS; ++s) ed += fabs((double)itv[h][s] + dud * (double)iuv[h][s]);
                    const int kr = has_r ? tx + 1 : tx, kd = y < H - 1 ? row + 1 : row;
                    e2[h] = energy_point(u[h], sUk[0][row][kr], sUk[0][kd][tx], u1x[h], sUk[1][row][kr],
                                         sUk[1][kd][tx], u1y[h], sUk[2][row][kr], sUk[2][kd][tx], sT[0][row + 1][tx + 1], sT[1][row + 1][tx + 1],
                                         sT[2][row + 1][tx + 1], ed, R.dlambda, R.dalpha0, R.dalpha1);
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                e2[0] += __shfl_xor(e2[0], off, 64);
                e2[1] += __shfl_xor(e2[1], off, 64);
            }
            if (tx == 0) { sWave[2 * wv] = e2[0]; sWave[2 * wv + 1] = e2[1]; }
            r_step();
        }
        // p, q of j+1 on the pixels (the lower pixel of the pair is the upper
        // one's down neighbour) and the ring
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = 2 * wv + h;
            const bool has_d = h == 0 ? has_d0 : has_d1;
            float rt[3], dn[3];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                // loaded by every lane (the cell exists; unused at the last
                // column), then selected: no branch on the lane's column
                const float right = sUb[f][row + 1][tx + 2];
                rt[f] = has_r ? right : ub[h][f];
                dn[f] = !has_d ? ub[h][f] : (h == 0 ? ub[1][f] : sUb[f][row + 2][tx + 1]);
            }
            pq_update(px[h], py[h], qx[h], qy[h], qz[h], sT[0][row + 1][tx + 1], sT[1][row + 1][tx + 1],
                      sT[2][row + 1][tx + 1], ub[h][0], rt[0], dn[0], ub[h][1],
                      rt[1], dn[1], ub[h][2], rt[2], dn[2], as1, as0);
            sPQ[0][row + 1][tx + 1] = px[h]; sPQ[1][row + 1][tx + 1] = py[h];
            sPQ[2][row + 1][tx + 1] = qx[h]; sPQ[3][row + 1][tx + 1] = qy[h]; sPQ[4][row + 1][tx + 1] = qz[h];
        }
        {
            int rgx, rgy, rix, riy;
            if (tr_ring(R, tcx, tcy, rid, rgx, rgy, rix, riy)) {
                const int rr = rgx < W - 1 ? rix + 1 : rix, rd = rgy < H - 1 ? riy + 1 : riy;
                float rpx = sPQ[0][riy][rix], rpy = sPQ[1][riy][rix];
                float rqx = sPQ[2][riy][rix], rqy = sPQ[3][riy][rix], rqz = sPQ[4][riy][rix];
                pq_update(rpx, rpy, rqx, rqy, rqz, sT[0][riy][rix], sT[1][riy][rix], sT[2][riy][rix], sUb[0][riy][rix],
                          sUb[0][riy][rr], sUb[0][rd][rix], sUb[1][riy][rix], sUb[1][riy][rr], sUb[1][rd][rix],
                          sUb[2][riy][rix], sUb[2][riy][rr], sUb[2][rd][rix], as1, as0);
                sPQ[0][riy][rix] = rpx; sPQ[1][riy][rix] = rpy;
                sPQ[2][riy][rix] = rqx; sPQ[3][riy][rix] = rqy; sPQ[4][riy][rix] = rqz;
            }
        }
        // the pending check's energy: every tile sums all partials
        if (decide && wv >= TR_FIN0 && wv < TR_FIN0 + 4 && !tr_fin_quarter(R, wv - TR_FIN0, nchk, sFin, tile, j))
            sFlag[0] = 1;
        __syncthreads();
        if (sFlag[0]) { failed = true; break; }
        if (decide) {
            if (stop_at_pending()) { stopped = true; break; }
            jpend = -1;
        }
        if (check) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                sSaveV[(0 + h) * TR_NT + tid] = v[h];
                sSaveV[(2 + h) * TR_NT + tid] = v1x[h];
                sSaveV[(4 + h) * TR_NT + tid] = v1y[h];
            }
            // the tile's two half-tile partials (a 640-thread block's rows in order)
            ++nchk;
            jpend = j;
            const int pl = tid - TR_PART_WAVE * 64;
            if (pl >= 0 && pl < 2) {
                const int sy = 2 * tcy + pl;   // the half tile's row in the launch sequence's tiles
                if (sy < R.sub_y) {
                    double e = 0.0;
                    for (int k = 0; k < SUB; ++k) e += sWave[pl * SUB + k];
                    const unsigned long long eb = (unsigned long long)__double_as_longlong(e);
                    unsigned long long* pw = R.part + ((size_t)(nchk & 1) * R.nsub + sy * R.tiles_x + tcx) * 2;
                    __hip_atomic_store(pw, tr_word((unsigned)nchk, (unsigned)(eb >> 32)), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(pw + 1, tr_word((unsigned)nchk, (unsigned)eb), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        // u of j+1 (u_step and u1_step, as u_update; the upper
        // pixel of the pair is the lower one's up neighbour), then the
        // tile's edges of state j+1 for the neighbours
        const bool publish = j + 1 < R.niters;
        unsigned long long* xb = R.xch + ((size_t)((j + 1) & 1) * R.ntiles + tile) * TR_XW;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = 2 * wv + h, y = y0 + row;
            const float lp0 = sPQ[0][row + 1][tx], lp1 = sPQ[1][row + 1][tx];   // every lane, then select
            const float lq0 = sPQ[2][row + 1][tx], lq2 = sPQ[4][row + 1][tx];
            const float l_px = has_l ? lp0 : px[h], l_py = has_l ? lp1 : py[h];
            const float l_qx = has_l ? lq0 : qx[h], l_qz = has_l ? lq2 : qz[h];
            const bool has_u = h == 1 || has_u0;
            const float u_px = !has_u ? px[h] : (h == 1 ? px[0] : sPQ[0][row][tx + 1]);
            const float u_py = !has_u ? py[h] : (h == 1 ? py[0] : sPQ[1][row][tx + 1]);
            const float u_qy = !has_u ? qy[h] : (h == 1 ? qy[0] : sPQ[3][row][tx + 1]);
            const float u_qz = !has_u ? qz[h] : (h == 1 ? qz[0] : sPQ[4][row][tx + 1]);
            const float c_px = px[h], c_py = py[h];
            const float t11h = sT[0][row + 1][tx + 1], t12h = sT[1][row + 1][tx + 1], t22h = sT[2][row + 1][tx + 1];
            const float un = u_step(u[h], t11h, t12h, t22h, c_px, c_py, l_px, l_py, u_px, u_py, ps[h], R.tau, R.alpha1,
                                    R.lambda);
            const float u1xn = u1_step(u1x[h], t11h, t12h, c_px, c_py, qx[h] - l_qx + qz[h] - u_qz, R.tau, R.alpha1,
                                       R.alpha0);
            const float u1yn = u1_step(u1y[h], t12h, t22h, c_px, c_py, qz[h] - l_qz + qy[h] - u_qy, R.tau, R.alpha1,
                                       R.alpha0);
            v[h] = u[h]; v1x[h] = u1x[h]; v1y[h] = u1y[h];
            u[h] = un; u1x[h] = u1xn; u1y[h] = u1yn;
            if (publish && x < W && y < H) {
                const unsigned tag = (unsigned)(j + 1);
                const float pub[3] = {2.f * u[h] - v[h], 2.f * u1x[h] - v1x[h], 2.f * u1y[h] - v1y[h]};
                const float uu[3] = {u[h], u1x[h], u1y[h]};
                // the top and bottom rows belong to whole waves (wv is uniform)
                if (row == 0) {
#pragma unroll
                    for (int f = 0; f < 3; ++f) {
                        tr_put(xb + TR_UB_TOP + f * TX + tx, tag, pub[f]);
                        tr_put(xb + TR_U_TOP + f * TX + tx, tag, uu[f]);
                    }
                }
                if (row == TY - 1) {
#pragma unroll
                    for (int f = 0; f < 3; ++f) tr_put(xb + TR_UB_BOT + f * TX + tx, tag, pub[f]);
                }
                // the left and right columns: lanes 0 and 63 of every wave
                if (tx == 0 || tx == TX - 1) {
                    unsigned long long* cb = xb + (tx == 0 ? TR_UB_LEFT : TR_UB_RIGHT) + row;
#pragma unroll
                    for (int f = 0; f < 3; ++f) tr_put(cb + f * TY, tag, pub[f]);
                    if (tx == 0) {
#pragma unroll
                        for (int f = 0; f < 3; ++f) tr_put(xb + TR_U_LEFT + f * TY + row, tag, uu[f]);
                    }
                }
            }
        }
    }
    if (failed) return;
    int ke = R.k0 + j;
    if (stopped) {
        // the state of the check that stopped the solve: u_j is in v, the
        // rest was saved
        ke = R.k0 + jpend;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            u[h] = v[h]; u1x[h] = v1x[h]; u1y[h] = v1y[h];
            v[h] = sSaveV[(0 + h) * TR_NT + tid]; v1x[h] = sSaveV[(2 + h) * TR_NT + tid];
            v1y[h] = sSaveV[(4 + h) * TR_NT + tid];
            px[h] = sSave[0 + h][tid]; py[h] = sSave[2 + h][tid];
            qx[h] = sSave[4 + h][tid]; qy[h] = sSave[6 + h][tid]; qz[h] = sSave[8 + h][tid];
#pragma unroll
            for (int s = 0; s < NS; ++s) rv[h][s] = sSave[10 + 2 * s + h][tid];
        }
    }
    // the end state into its slots
    float* U1 = R.U + (size_t)(ke % TG_USLOTS) * 3 * n;
    float* U0 = R.U + (size_t)((ke + TG_USLOTS - 1) % TG_USLOTS) * 3 * n;
    float* P = R.P + (size_t)(ke % TG_PSLOTS) * 2 * n;
    float* Q = R.Q + (size_t)(ke % TG_PSLOTS) * 3 * n;
    float* Rr = R.Rs + (size_t)(ke % TG_PSLOTS) * NS * n;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int y = y0 + 2 * wv + h;
        if (x >= W || y >= H) continue;
        const size_t g = (size_t)y * W + x;
        U1[g] = u[h]; U1[n + g] = u1x[h]; U1[2 * n + g] = u1y[h];
        U0[g] = v[h]; U0[n + g] = v1x[h]; U0[2 * n + g] = v1y[h];
        P[g] = px[h]; P[n + g] = py[h];
        Q[g] = qx[h]; Q[n + g] = qy[h]; Q[2 * n + g] = qz[h];
#pragma unroll
        for (int s = 0; s < NS; ++s) Rr[(size_t)s * n + g] = rv[h][s];
    }
}

// One frame: its tiles in XCD-local order; the start words are the frame's.
template <int NS>
__global__ __launch_bounds__(TR_NT, 3) void tgv_resident_kernel(TgvResArgs R) {
    tgv_resident_body<NS>(R, xcd_tile(blockIdx.x, R.ntiles), R.err + 1, (unsigned)R.ntiles);
}

// Several frames of the same source count in one launch (at most one tile
// per CU in all): frame f owns the launch's tiles tile0[f] .. tile0[f+1]-1.
// Each frame keeps its own exchange words, checks and stop; only the start
// count is the launch's.
template <int NS>
__global__ __launch_bounds__(TR_NT, 3) void tgv_resident_batch_kernel(const TgvResArgs* __restrict__ frames,
                                                                    const int* __restrict__ tile0, int nframes,
                                                                    int total, unsigned* start) {
    const int g = xcd_tile(blockIdx.x, total);
    int lo = 0, hi = nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile0[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    const TgvResArgs R = frames[lo];
    tgv_resident_body<NS>(R, g - tile0[lo], start, (unsigned)total);
}

// Exchange words live in uncached device memory, so that a word a tile on
// one XCD stores never waits in that XCD's L2 for a poller on another
// (ordinary device memory with device-scope atomics measured the same
// speed).  Uncached memory is never freed once a kernel has used it (later
// allocations at its address have held wrong values on these systems):
// solvers take buffers from a per-device pool and give them back on
// destroy.
struct TrPool {
    std::mutex mu;
    std::vector<std::pair<unsigned long long*, size_t>> free_bufs[64];
};
static TrPool& tr_pool() {
    static TrPool* p = new TrPool();   // never destroyed (the buffers outlive it)
    return *p;
}

static unsigned long long* tr_take(size_t words, int* dev_out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    TrPool& P = tr_pool();
    std::lock_guard<std::mutex> lock(P.mu);
    auto& fl = P.free_bufs[dev];
    for (size_t i = 0; i < fl.size(); ++i) {
        if (fl[i].second >= words) {
            unsigned long long* b = fl[i].first;
            fl.erase(fl.begin() + (long)i);
            *dev_out = dev;
            return b;
        }
    }
    void* b = nullptr;
    if (hipExtMallocWithFlags(&b, words * sizeof(unsigned long long), hipDeviceMallocUncached) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    *dev_out = dev;
    return static_cast<unsigned long long*>(b);
}

static void tr_give(unsigned long long* b, size_t words, int dev) {
    if (!b || dev < 0 || dev >= 64) return;
    TrPool& P = tr_pool();
    std::lock_guard<std::mutex> lock(P.mu);
    P.free_bufs[dev].emplace_back(b, words);
}

// Workgroups of the resident kernel the device holds at once: one per CU
// (per device, cached; 0 if the runtime cannot say or the kernel does not
// fit a CU), and the tile counts whose launches were called off because not
// every workgroup became resident in time (those use the launch sequence).
struct TrSlots {
    std::mutex mu;
    int cus[64] = {};
    std::vector<int> declined[64];
};
static TrSlots& tr_slots() {
    static TrSlots* s = new TrSlots();
    return *s;
}

static int tgv_resident_slots(int nt) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    TrSlots& S = tr_slots();
    std::lock_guard<std::mutex> lock(S.mu);
    if (S.cus[dev] == 0) {
        int per = 0, ncu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)tgv_resident_kernel<TR_MAXN>, TR_NT, 0) !=
                hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            (void)hipGetLastError();
            per = 0;
        }
        S.cus[dev] = per >= 1 && ncu > 0 ? ncu : -1;
    }
    for (int d : S.declined[dev])
        if (d == nt) return 0;
    return S.cus[dev] > 0 ? S.cus[dev] : 0;
}

static void tgv_resident_decline(int nt) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    TrSlots& S = tr_slots();
    std::lock_guard<std::mutex> lock(S.mu);
    S.declined[dev].push_back(nt);
}

static int tgv_resident_tiles(int W, int H) { return ((W + TG_TX - 1) / TG_TX) * ((H + TR_TY - 1) / TR_TY); }
static int tgv_resident_halves(int W, int H) { return ((W + TG_TX - 1) / TG_TX) * ((H + TR_SUB - 1) / TR_SUB); }

// 64-bit words of the resident solver's exchange buffer: edges, partials,
// decisions, error words.
static size_t tgv_resident_words(int W, int H) {
    const size_t nt = (size_t)tgv_resident_tiles(W, H), nh = (size_t)tgv_resident_halves(W, H);
    return 2 * nt * TR_XW + 4 * nh + 2 + 4;
}

// The resident kernel runs when it gives what the default launch sequence
// gives (PSW_TGV_VARIANT, PSW_TGV_DOUBLE, PSW_TGV2_TY and PSW_TGV_TRIPLE at
// their defaults: the energy checks then use the same 64 x 10 tile partials),
// the frame has at most TR_MAXN sources and at most one tile per CU.
// PSW_TGV_RESIDENT=0 turns it off.
static bool tgv_resident_use(const psw_tgv* t) {
    const char* e = getenv("PSW_TGV_RESIDENT");
    if (e && atoi(e) == 0) return false;
    if (!t->rx || t->N > TR_MAXN) return false;
    if (tgv_variant() != 1 || !tgv_double() || tgv_triple() || tgv2_ty() != TR_SUB) return false;
    const int nt = tgv_resident_tiles(t->W, t->H);
    return nt <= TR_MAXTILES && nt <= tgv_resident_slots(nt);
}

// Polls per wait (PSW_TGV_RESIDENT_SPIN; default 2^22, seconds at the
// latency of a device-scope load; 0 fails every wait), and polls of the
// start wait (PSW_TGV_RESIDENT_START_SPIN, default 2^16; set, it also keeps
// a called-off launch from being remembered).
static unsigned tgv_env_u32(const char* name, unsigned dflt) {
    const char* e = getenv(name);
    return e ? (unsigned)strtoul(e, nullptr, 10) : dflt;
}

// The arguments of a resident launch of t from state t->k (energy checks
// every check_every iterations when > 0, as in psw_tgv_solve), after
// clearing its exchange words (tags that match no iteration) and its error
// words.
static int tgv_resident_args(psw_tgv* t, const psw_tgv_params* prm, int niters, int check_every, double tol,
                             double e_star, hipStream_t st, TgvResArgs& R) {
    const int tiles_x = (t->W + TG_TX - 1) / TG_TX, tiles_y = (t->H + TR_TY - 1) / TR_TY;
    const int nt = tiles_x * tiles_y, nh = tgv_resident_halves(t->W, t->H);
    const size_t nx = (size_t)2 * nt * TR_XW, np = (size_t)4 * nh;
    int rc = psw_fill_u32(reinterpret_cast<uint32_t*>(t->rx), 0xFFFFFFFFu, 2 * (nx + np + 2), st);
    unsigned* err = reinterpret_cast<unsigned*>(t->rx + nx + np + 2);
    if (rc == PSW_OK) rc = psw_fill_u32(err, 0u, 8, st);
    if (rc != PSW_OK) return rc;
    const TgvPtrs a = tgv_ptrs(t, prm);
    R = TgvResArgs{};
    R.ucur = a.ucur; R.uprev = a.uprev; R.pcur = a.pcur; R.qcur = a.qcur; R.rcur = a.rcur;
    R.It = a.It; R.Iu = a.Iu; R.T = a.T; R.u0 = a.u0;
    R.mask = t->base + t->L.mask;
    R.U = t->base + t->L.U; R.P = t->base + t->L.P; R.Q = t->base + t->L.Q; R.Rs = t->base + t->L.r;
    R.xch = t->rx;
    R.part = t->rx + nx;
    R.dec = t->rx + nx + np;
    R.err = err;
    R.ctrl = t->ctrl;
    R.W = t->W; R.H = t->H; R.N = t->N;
    R.k0 = t->k;
    R.niters = niters;
    R.check_every = check_every;
    R.tiles_x = tiles_x;
    R.tiles_y = tiles_y;
    R.ntiles = nt;
    R.sub_y = (t->H + TR_SUB - 1) / TR_SUB;
    R.nsub = nh;
    R.spin = tgv_env_u32("PSW_TGV_RESIDENT_SPIN", 1u << 22);
    R.start_spin = tgv_env_u32("PSW_TGV_RESIDENT_START_SPIN", 1u << 16);
    R.lambda = a.lambda; R.alpha0 = a.alpha0; R.alpha1 = a.alpha1; R.tau = a.tau; R.sigma = a.sigma;
    R.dlambda = prm->lambda;
    R.dalpha0 = prm->alpha0;
    R.dalpha1 = prm->alpha1;
    R.e_star = e_star;
    R.gap = tol * e_star;
    return PSW_OK;
}

// The first failed wait of a resident launch of t, as the error message
// (PSW_ERR_DEVICE), or PSW_OK.  `he` holds the frame's error words.
static int tgv_resident_report(const TgvResArgs& R, const unsigned* he) {
    if (!he[0]) return PSW_OK;
    const unsigned blk = (unsigned)R.ntiles * TR_XW;
    if (he[4] < 2 * blk)
        set_error("resident TGV2 solver: tile %u waited %u polls for word %u of tile %u (iteration tag %u, saw %u) "
                  "at iteration %u; the state is undefined",
                  he[3], R.spin, he[4] % TR_XW, (he[4] % blk) / TR_XW, he[5], he[6], he[7]);
    else
        set_error("resident TGV2 solver: tile %u waited %u polls for an energy word (offset %u, check %u, saw %u) "
                  "at iteration %u; the state is undefined",
                  he[3], R.spin, he[4], he[5], he[6], he[7]);
    return PSW_ERR_DEVICE;
}

static const void* tgv_resident_kernel_for(int N, bool batch) {
#define PSW_RK(n) (batch ? (const void*)tgv_resident_batch_kernel<n> : (const void*)tgv_resident_kernel<n>)
    switch (N) {
        case 1: return PSW_RK(1);
        case 2: return PSW_RK(2);
        case 3: return PSW_RK(3);
        case 4: return PSW_RK(4);
        case 5: return PSW_RK(5);
        case 6: return PSW_RK(6);
        case 7: return PSW_RK(7);
        default: return PSW_RK(8);
    }
#undef PSW_RK
}

// niters iterations of t from state t->k in one cooperative launch.  *ran is
// false, with nothing run or written, when the runtime refuses the
// cooperative launch or the launch is called off.  Synchronous: it reads
// the error words.
static int tgv_resident_run(psw_tgv* t, const psw_tgv_params* prm, int niters, int check_every, double tol,
                            double e_star, hipStream_t st, bool* ran) {
    *ran = false;
    TgvResArgs R;
    int rc = tgv_resident_args(t, prm, niters, check_every, tol, e_star, st, R);
    if (rc != PSW_OK) return rc;
    void* args[] = {&R};
    const hipError_t e = hipLaunchCooperativeKernel(tgv_resident_kernel_for(t->N, false), dim3((unsigned)R.ntiles),
                                                    dim3(TR_NT), args, 0, st);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return PSW_OK;
    }
    static_assert(sizeof(TgvCtrl) >= 8 * sizeof(unsigned), "the error words are read through hctrl");
    PSW_TRY_HIP(hipMemcpyAsync(t->hctrl, R.err, 8 * sizeof(unsigned), hipMemcpyDeviceToHost, st), PSW_ERR_COPY);
    PSW_TRY_HIP(hipStreamSynchronize(st), PSW_ERR_DEVICE);
    unsigned he[8];
    std::memcpy(he, t->hctrl, sizeof(he));
    if (he[2] != 1u) {   // called off before any wait on a neighbour: nothing was written
        if (!getenv("PSW_TGV_RESIDENT_START_SPIN")) tgv_resident_decline(R.ntiles);
        return PSW_OK;
    }
    *ran = true;
    return tgv_resident_report(R, he);
}

// Several frames of one source count (together at most one tile per CU) in
// one cooperative launch, each from its own state with its own parameters
// and energy target.  *ran as for tgv_resident_run.  Synchronous.
static int tgv_resident_batch_run(psw_tgv* const* ts, const psw_tgv_params* prms, int nf, int niters, int check_every,
                                  double tol, const double* e_stars, hipStream_t st, bool* ran) {
    *ran = false;
    std::vector<TgvResArgs> Rs((size_t)nf);
    std::vector<int> tile0((size_t)nf + 1, 0);
    for (int f = 0; f < nf; ++f) {
        const int rc = tgv_resident_args(ts[f], &prms[f], niters, check_every, tol, e_stars ? e_stars[f] : 0.0, st,
                                         Rs[(size_t)f]);
        if (rc != PSW_OK) return rc;
        tile0[(size_t)f + 1] = tile0[(size_t)f] + Rs[(size_t)f].ntiles;
    }
    const int total = tile0[(size_t)nf];
    const size_t fb = sizeof(TgvResArgs) * (size_t)nf, tb = (sizeof(int) * ((size_t)nf + 1) + 15) / 16 * 16;
    char* dev = nullptr;
    PSW_TRY_HIP(hipMalloc((void**)&dev, fb + tb + 16), PSW_ERR_ALLOC);
    TgvResArgs* d_frames = reinterpret_cast<TgvResArgs*>(dev);
    int* d_tile0 = reinterpret_cast<int*>(dev + fb);
    unsigned* d_start = reinterpret_cast<unsigned*>(dev + fb + tb);
    int rc = PSW_OK;
    hipError_t e = hipMemcpyAsync(d_frames, Rs.data(), fb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tile0, tile0.data(), sizeof(int) * ((size_t)nf + 1), hipMemcpyHostToDevice,
                                            st);
    if (e != hipSuccess) rc = hip_status(e, "tgv_resident_batch_run", PSW_ERR_COPY);
    if (rc == PSW_OK) rc = psw_fill_u32(d_start, 0u, 2, st);
    bool launched = false;
    if (rc == PSW_OK) {
        int nfi = nf;
        void* args[] = {&d_frames, &d_tile0, &nfi, (void*)&total, &d_start};
        e = hipLaunchCooperativeKernel(tgv_resident_kernel_for(ts[0]->N, true), dim3((unsigned)total), dim3(TR_NT),
                                       args, 0, st);
        if (e != hipSuccess) (void)hipGetLastError();
        else launched = true;
    }
    unsigned go = 0u;
    std::vector<unsigned> he((size_t)nf * 8, 0u);
    if (rc == PSW_OK && launched) {
        e = hipMemcpyAsync(&go, d_start + 1, sizeof(unsigned), hipMemcpyDeviceToHost, st);
        for (int f = 0; f < nf && e == hipSuccess; ++f)
            e = hipMemcpyAsync(&he[(size_t)f * 8], Rs[(size_t)f].err, 8 * sizeof(unsigned), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = hip_status(e, "tgv_resident_batch_run", PSW_ERR_DEVICE);
    } else {
        (void)hipStreamSynchronize(st);
    }
    (void)hipFree(dev);
    if (rc != PSW_OK || !launched) return rc;
    if (go != 1u) {   // called off: nothing was written
        if (!getenv("PSW_TGV_RESIDENT_START_SPIN")) tgv_resident_decline(total);
        return PSW_OK;
    }
    *ran = true;
    for (int f = 0; f < nf; ++f) {
        const int r = tgv_resident_report(Rs[(size_t)f], &he[(size_t)f * 8]);
        if (r != PSW_OK) return r;
    }
    return PSW_OK;
}

static int tgv_reset_ctrl(psw_tgv* t, hipStream_t st) {
    tgv_ctrl_reset_kernel<<<1, 1, 0, st>>>(t->ctrl);
    return launch_status("tgv_ctrl_reset_kernel");
}

// After a resident solve of t from k_start: the stop, else k_start +
// max_iters and the energy of that state tested as the launch sequence tests
// the state its last chunk ends at; the iterations and the energy out; ctrl
// cleared.  Synchronous.
static int tgv_resident_finish(psw_tgv* t, int k_start, int max_iters, double tol, double e_star,
                               const psw_tgv_params* prm, hipStream_t st, int* iters_done, double* energy_out) {
    PSW_TRY_HIP(hipMemcpy(t->hctrl, t->ctrl, sizeof(TgvCtrl), hipMemcpyDeviceToHost), PSW_ERR_COPY);
    if (t->hctrl->done) {
        t->k = t->hctrl->iters_at_done;
    } else {
        t->k = k_start + max_iters;
        const int rc = tgv_launch_energy(t, prm, tol, e_star, st);
        if (rc != PSW_OK) return rc;
        PSW_TRY_HIP(hipStreamSynchronize(st), PSW_ERR_DEVICE);
        PSW_TRY_HIP(hipMemcpy(t->hctrl, t->ctrl, sizeof(TgvCtrl), hipMemcpyDeviceToHost), PSW_ERR_COPY);
    }
    if (iters_done) *iters_done = t->k - k_start;
    if (energy_out) *energy_out = t->hctrl->done ? t->hctrl->energy_at_done : t->hctrl->energy;
    return tgv_reset_ctrl(t, st) == PSW_OK && hipStreamSynchronize(st) == hipSuccess ? PSW_OK : PSW_ERR_DEVICE;
}

// The frames of a batched call that the resident solver takes: launches of
// one source count and at most one tile per CU in all, filled largest frame
// first (first fit).  The other frames go to `rest`.  The launch shape
// `variant` must be the default (the resident checks use its partials).
static std::vector<std::vector<int>> tgv_resident_packs(psw_tgv* const* ts, int n, int variant,
                                                        std::vector<int>& rest) {
    std::vector<std::vector<int>> packs;
    std::vector<int> ok;
    for (int f = 0; f < n; ++f) {
        if ((variant < 0 || variant == 1) && ts[f] && tgv_resident_use(ts[f])) ok.push_back(f);
        else rest.push_back(f);
    }
    const int cus = tgv_resident_slots(0);
    std::stable_sort(ok.begin(), ok.end(), [&](int a, int b) {
        return tgv_resident_tiles(ts[a]->W, ts[a]->H) > tgv_resident_tiles(ts[b]->W, ts[b]->H);
    });
    std::vector<int> fill, srcs;
    for (int f : ok) {
        const int nt = tgv_resident_tiles(ts[f]->W, ts[f]->H);
        size_t b = 0;
        while (b < packs.size() && (srcs[b] != ts[f]->N || fill[b] + nt > cus)) ++b;
        if (b == packs.size()) {
            packs.emplace_back();
            fill.push_back(0);
            srcs.push_back(ts[f]->N);
        }
        packs[b].push_back(f);
        fill[b] += nt;
    }
    return packs;
}

}  // namespace psw

using namespace psw;

extern "C" {

size_t psw_tgv_state_bytes(int width, int height, int nsrc) {
    TgvLayout L;
    L.init(width, height, nsrc);
    const size_t rx = nsrc <= TR_MAXN && tgv_resident_tiles(width, height) <= TR_MAXTILES
                          ? tgv_resident_words(width, height) * sizeof(unsigned long long)
                          : 0;
    return L.total * sizeof(float) +
           2 * std::max(TG_ENERGY_BLOCKS, tgv_tile_partials(width, height)) * sizeof(double) + sizeof(TgvCtrl) + rx;
}

int psw_tgv_create(psw_tgv** out, int width, int height, int nsrc) {
    PSW_CHECK_ARG(out, "out is NULL");
    *out = nullptr;
    PSW_CHECK_ARG(width >= 2 && height >= 2, "image must be at least 2 x 2");
    PSW_CHECK_ARG((uint64_t)width * (uint64_t)height < (1ull << 32), "image above 2^32 pixels");
    PSW_CHECK_ARG(nsrc >= 1 && nsrc <= PSW_MAX_SOURCES, "nsrc must be 1..PSW_MAX_SOURCES");
    psw_tgv* t = new psw_tgv();
    t->W = width; t->H = height; t->N = nsrc;
    t->L.init(width, height, nsrc);
    t->npartials = TG_ENERGY_BLOCKS;
    t->k = 0;
    t->have_images = false;
    hipError_t e = hipMalloc((void**)&t->base, t->L.total * sizeof(float));
    // the stand-alone energy pass uses the first TG_ENERGY_BLOCKS; the solve
    // loop alternates between two halves of tile partials
    const int np = 2 * std::max(TG_ENERGY_BLOCKS, tgv_tile_partials(width, height));
    if (e == hipSuccess) e = hipMalloc((void**)&t->partials, np * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&t->ctrl, sizeof(TgvCtrl));
    if (e == hipSuccess) e = hipMalloc((void**)&t->masked, sizeof(unsigned long long));
    if (e == hipSuccess && nsrc <= TR_MAXN && tgv_resident_tiles(width, height) <= TR_MAXTILES) {
        t->rx_words = tgv_resident_words(width, height);
        t->rx = tr_take(t->rx_words, &t->rx_dev);   // NULL: the solver uses the launch sequence
    }
    if (e == hipSuccess) e = hipHostMalloc((void**)&t->hctrl, sizeof(TgvCtrl), hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&t->hmirror, sizeof(TgvCtrl), hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&t->dmirror, t->hmirror, 0);
    if (e == hipSuccess) std::memset(t->hmirror, 0, sizeof(TgvCtrl));
    if (e == hipSuccess) e = hipMemset(t->base, 0, t->L.total * sizeof(float));
    if (e == hipSuccess) e = hipMemset(t->ctrl, 0, sizeof(TgvCtrl));
    if (e != hipSuccess) {
        int s = hip_status(e, "psw_tgv_create", PSW_ERR_ALLOC);
        psw_tgv_destroy(t);
        return s;
    }
    int fs = psw_fill_f32(t->base + t->L.mask, 1.f, t->L.n, nullptr);
    if (fs == PSW_OK) fs = hipDeviceSynchronize() == hipSuccess ? PSW_OK : PSW_ERR_DEVICE;
    if (fs != PSW_OK) {
        psw_tgv_destroy(t);
        return fs;
    }
    *out = t;
    return PSW_OK;
}

int psw_tgv_destroy(psw_tgv* t) {
    if (!t) return PSW_OK;
    if (t->base) (void)hipFree(t->base);
    if (t->partials) (void)hipFree(t->partials);
    if (t->ctrl) (void)hipFree(t->ctrl);
    if (t->masked) (void)hipFree(t->masked);
    if (t->rx) tr_give(t->rx, t->rx_words, t->rx_dev);
    if (t->hctrl) (void)hipHostFree(t->hctrl);
    if (t->hmirror) (void)hipHostFree(t->hmirror);
    delete t;
    return PSW_OK;
}

int psw_tgv_set_images(psw_tgv* t, const float* d_ref, const float* d_src, double beta, double gamma,
                       psw_stream_t s) {
    PSW_CHECK_ARG(t && d_ref && d_src, "NULL pointer");
    const TgvLayout& L = t->L;
    tgv_scale_kernel<<<512, 256, 0, S(s)>>>(t->base + L.ref, d_ref, L.n);
    int st = launch_status("tgv_scale_kernel");
    if (st != PSW_OK) return st;
    tgv_scale_kernel<<<2048, 256, 0, S(s)>>>(t->base + L.src, d_src, L.n * t->N);
    st = launch_status("tgv_scale_kernel");
    if (st != PSW_OK) return st;
    const dim3 blk(64, 4), grd((unsigned)((t->W + 63) / 64), (unsigned)((t->H + 3) / 4));
    tgv_tensor_kernel<<<grd, blk, 0, S(s)>>>(t->base + L.T, t->base + L.ref, t->W, t->H, (float)beta, (float)gamma);
    st = launch_status("tgv_tensor_kernel");
    if (st != PSW_OK) return st;
    t->have_images = true;
    return PSW_OK;
}

int psw_tgv_set_u(psw_tgv* t, const float* d_u, float value, psw_stream_t s) {
    PSW_CHECK_ARG(t, "NULL solver");
    float* u = t->base + t->L.U + (size_t)(t->k % TG_USLOTS) * 3 * t->L.n;
    if (d_u) {
        PSW_TRY_HIP(hipMemcpyAsync(u, d_u, t->L.n * sizeof(float), hipMemcpyDeviceToDevice, S(s)), PSW_ERR_COPY);
        return PSW_OK;
    }
    return psw_fill_f32(u, value, t->L.n, s);
}

int psw_tgv_linearise(psw_tgv* t, const float* K, const float* Rrel, const float* trel, psw_stream_t s) {
    PSW_CHECK_ARG(t && K && Rrel && trel, "NULL pointer");
    PSW_CHECK_ARG(t->have_images, "psw_tgv_set_images must come first");
    const TgvLayout& L = t->L;
    const size_t n = L.n;
    const int cur = t->k % TG_USLOTS, prev = (t->k + TG_USLOTS - 1) % TG_USLOTS;
    tgv_warp_reset_kernel<<<1024, 256, 0, S(s)>>>(t->base + L.U, t->base + L.P + (size_t)(t->k % TG_PSLOTS) * 2 * n,
                                                  t->base + L.Q + (size_t)(t->k % TG_PSLOTS) * 3 * n, t->base + L.u0, n, cur,
                                                  prev);
    int st = launch_status("tgv_warp_reset_kernel");
    if (st != PSW_OK) return st;
    TgvPoses pz;
    pz.K = m3_from(K);
    // Matrix3D::inv (structs.h:584-593) in float, as PlaneSweep::setK stores it.
    {
        const float* m = K;
        const float det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                          m[2] * (m[3] * m[7] - m[4] * m[6]);
        const float r[9] = {m[4] * m[8] - m[5] * m[7], m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
                            m[5] * m[6] - m[3] * m[8], m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
                            m[3] * m[7] - m[4] * m[6], m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3]};
        for (int j = 0; j < 9; ++j) pz.invK.m[j] = r[j] / det;
    }
    for (int j = 0; j < t->N; ++j) {
        pz.Rrel[j] = m3_from(Rrel + 9 * j);
        pz.trel[j] = v3_from(trel + 3 * j);
        pz.RK[j] = mm3(pz.Rrel[j], pz.invK);
    }
    pz.fx = (float)(double)K[0];
    pz.fy = (float)(double)K[4];
    const dim3 blk(64, 4), grd((unsigned)((t->W + 63) / 64), (unsigned)((t->H + 3) / 4));
    tgv_linearise_kernel<<<grd, blk, 0, S(s)>>>(pz, t->base + L.src, t->base + L.ref, t->base + L.u0, t->base + L.It,
                                                t->base + L.Iu, t->base + L.r + (size_t)(t->k % TG_PSLOTS) * t->N * n, t->W,
                                                t->H, t->N);
    st = launch_status("tgv_linearise_kernel");
    if (st != PSW_OK) return st;
    if ((st = psw_fill_f32(t->base + L.mask, 1.f, n, s)) != PSW_OK) return st;
    return tgv_reset_ctrl(t, S(s));
}

int psw_tgv_set_primal(psw_tgv* t, const float* d_u, float u_value, const float* d_u1, psw_stream_t s) {
    PSW_CHECK_ARG(t, "NULL solver");
    const TgvLayout& L = t->L;
    const size_t n = L.n;
    const int cur = t->k % TG_USLOTS, prev = (t->k + TG_USLOTS - 1) % TG_USLOTS;
    tgv_set_primal_kernel<<<1024, 256, 0, S(s)>>>(t->base + L.U, t->base + L.P + (size_t)(t->k % TG_PSLOTS) * 2 * n,
                                                  t->base + L.Q + (size_t)(t->k % TG_PSLOTS) * 3 * n,
                                                  t->base + L.r + (size_t)(t->k % TG_PSLOTS) * t->N * n, d_u,
                                                  u_value, d_u1, n, t->N, cur, prev);
    int st = launch_status("tgv_set_primal_kernel");
    if (st != PSW_OK) return st;
    return tgv_reset_ctrl(t, S(s));
}

int psw_tgv_set_duals(psw_tgv* t, const float* d_p, const float* d_q, const float* d_r, psw_stream_t s) {
    PSW_CHECK_ARG(t, "NULL solver");
    const TgvLayout& L = t->L;
    const size_t n = L.n;
    float* P = t->base + L.P + (size_t)(t->k % TG_PSLOTS) * 2 * n;
    float* Q = t->base + L.Q + (size_t)(t->k % TG_PSLOTS) * 3 * n;
    float* R = t->base + L.r + (size_t)(t->k % TG_PSLOTS) * t->N * n;
    int st = PSW_OK;
    if (d_p) PSW_TRY_HIP(hipMemcpyAsync(P, d_p, 2 * n * sizeof(float), hipMemcpyDeviceToDevice, S(s)), PSW_ERR_COPY);
    else st = psw_fill_f32(P, 0.f, 2 * n, s);
    if (st != PSW_OK) return st;
    if (d_q) PSW_TRY_HIP(hipMemcpyAsync(Q, d_q, 3 * n * sizeof(float), hipMemcpyDeviceToDevice, S(s)), PSW_ERR_COPY);
    else st = psw_fill_f32(Q, 0.f, 3 * n, s);
    if (st != PSW_OK) return st;
    if (d_r) PSW_TRY_HIP(hipMemcpyAsync(R, d_r, (size_t)t->N * n * sizeof(float), hipMemcpyDeviceToDevice, S(s)),
                         PSW_ERR_COPY);
    else st = psw_fill_f32(R, 0.f, (size_t)t->N * n, s);
    if (st != PSW_OK) return st;
    return tgv_reset_ctrl(t, S(s));
}

int psw_tgv_set_energy_mask(psw_tgv* t, double qmax, const psw_tgv_params* prm, int64_t* excluded,
                            psw_stream_t s) {
    PSW_CHECK_ARG(t && prm, "NULL pointer");
    const TgvLayout& L = t->L;
    unsigned long long* d = excluded ? t->masked : nullptr;   // the solver's own counter
    if (d) {
        const int fs = psw_fill_u32(reinterpret_cast<uint32_t*>(d), 0u, 2, s);
        if (fs != PSW_OK) return fs;
    }
    const dim3 blk(64, 4), grd((unsigned)((t->W + 63) / 64), (unsigned)((t->H + 3) / 4));
    const double scale = prm->tau * prm->sigma * prm->lambda * prm->lambda;
    tgv_mask_kernel<<<grd, blk, 0, S(s)>>>(t->base + L.Iu, t->W, t->H, t->N, scale, qmax, t->base + L.mask, d);
    int st = launch_status("tgv_mask_kernel");
    if (excluded) {
        unsigned long long h = 0;
        if (st == PSW_OK) {
            hipError_t e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, S(s));
            if (e == hipSuccess) e = hipStreamSynchronize(S(s));
            if (e != hipSuccess) st = hip_status(e, "psw_tgv_set_energy_mask");
        }
        *excluded = (int64_t)h;
    }
    return st;
}

int psw_tgv_iterate(psw_tgv* t, int niters, const psw_tgv_params* prm, psw_stream_t s) {
    PSW_CHECK_ARG(t && prm, "NULL pointer");
    PSW_CHECK_ARG(niters >= 0, "niters must be >= 0");
    if (niters >= TR_MIN_ITERS && tgv_resident_use(t)) {
        bool ran = false;
        const int rc = tgv_resident_run(t, prm, niters, 0, 0.0, 0.0, S(s), &ran);
        if (rc != PSW_OK) return rc;
        if (ran) {
            t->k += niters;
            return PSW_OK;
        }
    }
    return tgv_launch_iters(t, niters, prm, S(s));
}

int psw_tgv_energy(psw_tgv* t, const psw_tgv_params* prm, double* energy, psw_stream_t s) {
    PSW_CHECK_ARG(t && prm && energy, "NULL pointer");
    int st = tgv_launch_energy(t, prm, 0.0, 0.0, S(s));
    if (st != PSW_OK) return st;
    PSW_TRY_HIP(hipMemcpyAsync(t->hctrl, t->ctrl, sizeof(TgvCtrl), hipMemcpyDeviceToHost, S(s)), PSW_ERR_COPY);
    PSW_TRY_HIP(hipStreamSynchronize(S(s)), PSW_ERR_DEVICE);
    *energy = t->hctrl->energy;
    return PSW_OK;
}

int psw_tgv_solve(psw_tgv* t, int max_iters, int check_every, double tol, double e_star, const psw_tgv_params* prm,
                  int* iters_done, double* energy_out, psw_stream_t s) {
    PSW_CHECK_ARG(t && prm, "NULL pointer");
    PSW_CHECK_ARG(max_iters >= 0 && check_every >= 1, "bad iteration counts");
    hipStream_t st = S(s);
    int rc = tgv_reset_ctrl(t, st);
    if (rc != PSW_OK) return rc;
    const int k_start = t->k;
    if (max_iters >= 1 && check_every >= 2 && tgv_resident_use(t)) {
        // the whole solve in one resident launch (whose checks need
        // check_every >= 2); the state it ends at (the stop, else k_start +
        // max_iters) is tested afterwards as below
        bool ran = false;
        rc = tgv_resident_run(t, prm, max_iters, check_every, tol, e_star, st, &ran);
        if (rc != PSW_OK) return rc;
        if (ran) return tgv_resident_finish(t, k_start, max_iters, tol, e_star, prm, st, iters_done, energy_out);
    }
    std::memset(t->hmirror, 0, sizeof(TgvCtrl));   // no launch of this solver writes it now
    volatile const TgvCtrl* mirror = t->hmirror;
    const int half = std::max(TG_ENERGY_BLOCKS, tgv_tile_partials(t->W, t->H));
    const int ntiles = ((t->W + TG_TX - 1) / TG_TX) * ((t->H + (tgv_variant() == 2 ? 4 : 8) - 1) /
                                                        (tgv_variant() == 2 ? 4 : 8));
    TgvCheck chk;
    chk.partials = nullptr;
    chk.fin_partials = nullptr;
    chk.fin_count = ntiles;
    chk.fin_k = 0;
    chk.ctrl = t->ctrl;
    chk.mirror = t->dmirror;
    chk.mask = t->base + t->L.mask;
    chk.lambda = prm->lambda;
    chk.alpha0 = prm->alpha0;
    chk.alpha1 = prm->alpha1;
    chk.tol = tol;
    chk.e_star = e_star;
    // Chunks of check_every iterations.  The first launch of every chunk
    // after the first sums the energy of the state the previous chunk ended
    // at (into one of two partial buffers, alternately); the next launch
    // finalises it and writes the outcome to pinned host memory.  At most
    // `depth` chunks are in flight: the host waits for the chunk `depth` back,
    // so the device never idles and at most `depth` chunks run as no-ops (they
    // test ctrl->done and return) after the stop.
    const int depth = tgv_inflight();
    hipEvent_t ev[TG_MAX_INFLIGHT] = {};
    for (int i = 0; i < depth && rc == PSW_OK; ++i)
        if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) rc = PSW_ERR_DEVICE;
    // A check is an energy launch from state k and the finalising launch
    // after it; both do the same number of iterations, so a stop at k leaves
    // u_k, u_{k-1}, p_k, q_k, r_k in place (slot counts at TG_USLOTS).  The
    // other launches do up to three iterations each (tgv_launch_size).
    const int mmax = tgv_max_per_launch(t);
    const int ntiles2 = tgv_tiles2(t), ntiles3 = tgv_tiles3(t);
    int done_iters = -1, chunk_idx = 0, nchecks = 0;
    int launched = 0;
    bool stopped = false, pending = false;   // pending: an energy sum not yet finalised
    int pend_m = 1;                          // iterations of the pending energy launch
    double* pend_buf = nullptr;
    int pend_k = 0, pend_count = 0;
    while (launched < max_iters && rc == PSW_OK) {
        const int c = std::min(check_every, max_iters - launched);
        for (int i = 0; i < c && rc == PSW_OK;) {
            const bool energy = i == 0 && launched > 0;
            const bool fin = pending;
            int m;
            if (fin) m = pend_m;                                      // the same kind as its energy launch
            else if (energy) m = std::max(1, std::min(mmax, (c - i) / 2));   // room for its finalising launch
            else m = tgv_launch_size(c - i, mmax);
            if (fin) {
                chk.fin_partials = pend_buf;
                chk.fin_k = pend_k;
                chk.fin_count = pend_count;
            }
            if (energy) chk.partials = t->partials + (size_t)(nchecks & 1) * half;
            const int k_before = t->k;
            rc = m == 3   ? tgv_launch_triple(t, prm, st, &chk, energy, fin)
                 : m == 2 ? tgv_launch_double(t, prm, st, &chk, energy, fin)
                          : tgv_launch_one(t, prm, chk, energy, fin, st);
            if (energy) {
                pend_buf = chk.partials;
                pend_k = k_before;
                pend_count = m == 3 ? ntiles3 : m == 2 ? ntiles2 : ntiles;
                pend_m = m;
                nchecks++;
            }
            pending = energy;
            i += m;
        }
        if (rc != PSW_OK) break;
        launched += c;
        hipEvent_t e = ev[chunk_idx % depth];
        if (chunk_idx >= depth) {
            // the event slot holds the chunk `depth` back: wait for it
            if (hipEventSynchronize(e) != hipSuccess) { rc = PSW_ERR_DEVICE; break; }
            if (mirror->done) { done_iters = mirror->iters_at_done; stopped = true; break; }
        }
        (void)hipEventRecord(e, st);
        chunk_idx++;
    }
    if (rc == PSW_OK && pending && !stopped) {
        // the last energy sum has no following launch: finalise it alone
        chk.fin_partials = pend_buf;
        chk.fin_k = pend_k;
        chk.fin_count = pend_count;
        if (tgv_variant() == 1) tgv_check_finalize_kernel<512><<<1, 512, 0, st>>>(chk);
        else tgv_check_finalize_kernel<256><<<1, 256, 0, st>>>(chk);
        rc = launch_status("tgv_check_finalize_kernel");
    }
    // the state the last chunk ended at has not been tested yet
    if (rc == PSW_OK && !stopped) rc = tgv_launch_energy(t, prm, tol, e_star, st);
    (void)hipStreamSynchronize(st);
    (void)hipMemcpy(t->hctrl, t->ctrl, sizeof(TgvCtrl), hipMemcpyDeviceToHost);
    for (int i = 0; i < depth; ++i)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    if (rc != PSW_OK) return rc;
    if (t->hctrl->done) done_iters = t->hctrl->iters_at_done;
    if (done_iters >= 0) t->k = done_iters;  // later kernels were no-ops
    if (iters_done) *iters_done = t->k - k_start;
    if (energy_out) *energy_out = t->hctrl->done ? t->hctrl->energy_at_done : t->hctrl->energy;
    // leave ctrl cleared so that plain psw_tgv_iterate calls run
    return tgv_reset_ctrl(t, st) == PSW_OK && hipStreamSynchronize(st) == hipSuccess ? PSW_OK : PSW_ERR_DEVICE;
}

int psw_tgv_get_u(psw_tgv* t, float* d_out, psw_stream_t s) {
    PSW_CHECK_ARG(t && d_out, "NULL pointer");
    const float* u = t->base + t->L.U + (size_t)(t->k % TG_USLOTS) * 3 * t->L.n;
    PSW_TRY_HIP(hipMemcpyAsync(d_out, u, t->L.n * sizeof(float), hipMemcpyDeviceToDevice, S(s)), PSW_ERR_COPY);
    return PSW_OK;
}

int psw_tgv_get_energy_mask(psw_tgv* t, float* d_out, psw_stream_t s) {
    PSW_CHECK_ARG(t && d_out, "NULL pointer");
    PSW_TRY_HIP(hipMemcpyAsync(d_out, t->base + t->L.mask, t->L.n * sizeof(float), hipMemcpyDeviceToDevice, S(s)),
                PSW_ERR_COPY);
    return PSW_OK;
}

int psw_tgv_get_state(psw_tgv* t, float* d_out, psw_stream_t s) {
    PSW_CHECK_ARG(t && d_out, "NULL pointer");
    const TgvLayout& L = t->L;
    const size_t n = L.n, N = (size_t)t->N;
    const float* B = t->base;
    auto cp = [&](float* dst, const float* src, size_t cnt) {
        return hipMemcpyAsync(dst, src, cnt * sizeof(float), hipMemcpyDeviceToDevice, S(s));
    };
    const float* U = B + L.U + (size_t)(t->k % TG_USLOTS) * 3 * n;
    const float* P = B + L.P + (size_t)(t->k % TG_PSLOTS) * 2 * n;
    const float* Q = B + L.Q + (size_t)(t->k % TG_PSLOTS) * 3 * n;
    float* o = d_out;
    PSW_TRY_HIP(cp(o, U, 3 * n), PSW_ERR_COPY); o += 3 * n;
    PSW_TRY_HIP(cp(o, P, 2 * n), PSW_ERR_COPY); o += 2 * n;
    PSW_TRY_HIP(cp(o, Q, 3 * n), PSW_ERR_COPY); o += 3 * n;
    PSW_TRY_HIP(cp(o, B + L.It, N * n), PSW_ERR_COPY); o += N * n;
    PSW_TRY_HIP(cp(o, B + L.Iu, N * n), PSW_ERR_COPY); o += N * n;
    PSW_TRY_HIP(cp(o, B + L.r + (size_t)(t->k % TG_PSLOTS) * N * n, N * n), PSW_ERR_COPY); o += N * n;
    PSW_TRY_HIP(cp(o, B + L.T, 3 * n), PSW_ERR_COPY); o += 3 * n;
    PSW_TRY_HIP(cp(o, B + L.u0, n), PSW_ERR_COPY);
    return PSW_OK;
}

int psw_tgv_run(psw_tgv* t, const float* K, const float* Rrel, const float* trel, int niters, int warps,
                const psw_tgv_params* prm, const float* d_u_init, float* d_u_out, psw_stream_t s) {
    PSW_CHECK_ARG(t && K && Rrel && trel && prm && d_u_out, "NULL pointer");
    PSW_CHECK_ARG(niters >= 0 && warps >= 0, "negative iteration count");
    int st = psw_tgv_set_u(t, d_u_init, 1.f, s);   // cold start u = 1 (planesweep.cpp:414)
    if (st != PSW_OK) return st;
    for (int l = 0; l < warps; ++l) {
        st = psw_tgv_linearise(t, K, Rrel, trel, s);
        if (st != PSW_OK) return st;
        st = psw_tgv_iterate(t, niters, prm, s);   // resident when the frame allows it
        if (st != PSW_OK) return st;
    }
    return psw_tgv_get_u(t, d_u_out, s);
}


// Device tables of a batch: frames, tile and energy-block offsets, the
// iteration counter and the stopped-frame counter.
struct TgvBatch {
    char* dev = nullptr;
    TgvFrameDev* frames = nullptr;
    int* toff = nullptr;
    int* eoff = nullptr;
    int* kdev = nullptr;                // three copies of the batch counter
    int tiles = 0, eblocks = 0, ty = 8, nt = 512;
    int tiles2 = 0, ty2 = 10;
    TgvBatchArgs args{};                // the active frames for the two-iteration launches
    bool dbl = false;                   // two iterations per launch
    int nact = 0;                       // frames in the device tables
    std::vector<TgvFrameDev> all;       // every frame of the batch
    std::vector<int> active;            // indices (into all) still iterating
    ~TgvBatch() {
        if (dev) (void)hipFree(dev);
    }
    // Device tables for the active frames only (stopped frames no longer
    // cost a dispatch or an energy pass).  Synchronous.
    int upload(hipStream_t st) {
        nact = (int)active.size();
        std::vector<TgvFrameDev> fr;
        std::vector<int> toff(1, 0), eoff(1, 0), toff2(1, 0);
        for (int f : active) {
            const TgvFrameDev& d = all[(size_t)f];
            fr.push_back(d);
            toff.push_back(toff.back() + ((d.W + TG_TX - 1) / TG_TX) * ((d.H + ty - 1) / ty));
            eoff.push_back(eoff.back() + d.eblocks);
            toff2.push_back(toff2.back() + d.tiles2);
        }
        tiles = toff.back();
        eblocks = eoff.back();
        tiles2 = toff2.back();
        if (dbl && nact > TG_BATCH_MAXF) return PSW_ERR_ARG;   // the solver groups frames
        args = TgvBatchArgs{};
        args.nf = dbl ? nact : 0;
        for (int k = 0; dbl && k < nact; ++k) {
            args.fr[k] = fr[(size_t)k];
            args.toff[k] = toff2[(size_t)k];
        }
        if (dbl) args.toff[nact] = toff2.back();
        if (nact == 0) return PSW_OK;
        PSW_TRY_HIP(hipMemcpyAsync(frames, fr.data(), sizeof(TgvFrameDev) * fr.size(), hipMemcpyHostToDevice, st),
                    PSW_ERR_COPY);
        PSW_TRY_HIP(hipMemcpyAsync(toff_d(), toff.data(), sizeof(int) * toff.size(), hipMemcpyHostToDevice, st),
                    PSW_ERR_COPY);
        PSW_TRY_HIP(hipMemcpyAsync(eoff_d(), eoff.data(), sizeof(int) * eoff.size(), hipMemcpyHostToDevice, st),
                    PSW_ERR_COPY);
        PSW_TRY_HIP(hipStreamSynchronize(st), PSW_ERR_DEVICE);
        return PSW_OK;
    }
    int* toff_d() const { return toff; }
    int* eoff_d() const { return eoff; }
};

static int tgv_batch_setup(psw_tgv* const* ts, int nf, const psw_tgv_params* prm, const double* e_star, int variant,
                           TgvBatch& b, hipStream_t st) {
    PSW_CHECK_ARG(ts && prm && nf >= 1, "NULL pointer or no frames");
    const int var = variant >= 0 && variant <= 2 ? variant : tgv_variant();
    b.dbl = tgv_double();
    for (int f = 0; f < nf; ++f) b.dbl = b.dbl && ts[f] && ts[f]->N <= T2_MAXN;
    b.ty = var == 2 ? 4 : 8;              // the one-iteration launches
    b.nt = var == 1 ? 512 : 256;
    // The two-iteration launches use the single-frame solve's tiles, so the
    // checking launches' energies (and the stops) are the single-frame ones
    // bit for bit.
    b.ty2 = tgv2_ty() == 8 ? 8 : 10;
    b.all.assign((size_t)nf, TgvFrameDev{});
    b.active.clear();
    for (int f = 0; f < nf; ++f) {
        psw_tgv* t = ts[f];
        PSW_CHECK_ARG(t, "NULL solver in the batch");
        TgvFrameDev& d = b.all[(size_t)f];
        d.base = t->base; d.L = t->L; d.W = t->W; d.H = t->H; d.N = t->N; d.k0 = t->k;
        d.ctrl = t->ctrl; d.mirror = t->dmirror; d.partials = t->partials;
        int blocks = (int)((t->L.n + TG_THREADS - 1) / TG_THREADS);
        d.eblocks = blocks > t->npartials ? t->npartials : blocks;   // as tgv_launch_energy
        d.tiles2 = ((t->W + TG_TX - 1) / TG_TX) * ((t->H + b.ty2 - 1) / b.ty2);
        const psw_tgv_params& q = prm[f];
        d.lambda = (float)q.lambda; d.alpha0 = (float)q.alpha0; d.alpha1 = (float)q.alpha1;
        d.tau = (float)q.tau; d.sigma = (float)q.sigma;
        d.dlambda = q.lambda; d.dalpha0 = q.alpha0; d.dalpha1 = q.alpha1;
        d.e_star = e_star ? e_star[f] : 0.0;
        b.active.push_back(f);
    }
    const size_t fb = sizeof(TgvFrameDev) * (size_t)nf, ob = sizeof(int) * ((size_t)nf + 1);
    PSW_TRY_HIP(hipMalloc((void**)&b.dev, fb + 2 * ob + 3 * sizeof(int)), PSW_ERR_ALLOC);
    b.frames = reinterpret_cast<TgvFrameDev*>(b.dev);
    b.toff = reinterpret_cast<int*>(b.dev + fb);
    b.eoff = reinterpret_cast<int*>(b.dev + fb + ob);
    b.kdev = reinterpret_cast<int*>(b.dev + fb + 2 * ob);
    int rc = psw_fill_u32(reinterpret_cast<uint32_t*>(b.kdev), 0u, 3, st);
    for (int f = 0; f < nf && rc == PSW_OK; ++f) {
        rc = tgv_reset_ctrl(ts[f], st);
        std::memset(ts[f]->hmirror, 0, sizeof(TgvCtrl));   // no launch of this solver writes it now
    }
    if (rc == PSW_OK) rc = b.upload(st);
    return rc;
}

extern "C++" template <int NS, int TY>
static void tgv_batch_launch2_ns(const TgvBatch& b, int rd, int wr, bool energy, bool fin, double tol,
                                 hipStream_t st) {
    constexpr int NT = TG_TX * TY;
    const unsigned g = (unsigned)(b.tiles2 + (fin ? b.nact : 0));
    const TgvBatchArgs& A = b.args;
    if (!energy && !fin) tgv_iter2_batch_kernel<NS, TY, false, false><<<g, NT, 0, st>>>(A, b.kdev, rd, wr, tol);
    else if (energy && fin) tgv_iter2_batch_kernel<NS, TY, true, true><<<g, NT, 0, st>>>(A, b.kdev, rd, wr, tol);
    else if (energy) tgv_iter2_batch_kernel<NS, TY, true, false><<<g, NT, 0, st>>>(A, b.kdev, rd, wr, tol);
    else tgv_iter2_batch_kernel<NS, TY, false, true><<<g, NT, 0, st>>>(A, b.kdev, rd, wr, tol);
}

// One two-iteration launch over the active frames, reading the count from
// copy rd and writing it advanced to copy wr (wr != rd, both in 0..2).
static void tgv_batch_launch2(const TgvBatch& b, int rd, int wr, bool energy, bool fin, double tol,
                              hipStream_t st) {
    bool all8 = true;
    for (int f : b.active) all8 = all8 && b.all[(size_t)f].N <= 8;
    if (b.ty2 == 8) {
        if (all8) tgv_batch_launch2_ns<8, 8>(b, rd, wr, energy, fin, tol, st);
        else tgv_batch_launch2_ns<T2_MAXN, 8>(b, rd, wr, energy, fin, tol, st);
    } else {
        if (all8) tgv_batch_launch2_ns<8, 10>(b, rd, wr, energy, fin, tol, st);
        else tgv_batch_launch2_ns<T2_MAXN, 10>(b, rd, wr, energy, fin, tol, st);
    }
}

// c iterations of every frame, starting from counter copy 0; *cur gets the
// copy that holds the count afterwards.  With check (c >= 4, two-iteration
// launches), the first launch also sums each frame's energy of the state it
// starts from and the second finalises it (the stop test), as in the
// single-frame solve.
static int tgv_batch_iters(const TgvBatch& b, int c, hipStream_t st, int* cur, bool check = false,
                           double tol = 0.0) {
    const int nf = b.nact;
    // Launch j reads the count from copy rd and its block 0 writes the
    // advanced count to copy wr != rd.  The copies go 0 -> 1 -> 2 -> 1 -> 2
    // ... and the last launch writes copy 0, so that a chunk of two or more
    // launches ends where it started (a captured chunk replays unchanged).
    int nl = 0;
    for (int i = 0; i < c;) {
        i += b.dbl && c - i >= 2 ? 2 : 1;
        ++nl;
    }
    int rd = 0, j = 0;
    for (int i = 0; i < c; ++j) {
        const int wr = j == nl - 1 && nl >= 2 ? 0 : (rd == 1 ? 2 : 1);
        if (b.dbl && c - i >= 2) {
            tgv_batch_launch2(b, rd, wr, check && i == 0, check && i == 2, tol, st);
            i += 2;
        } else {
            if (b.ty == 4)
                tgv_iter_batch_kernel<4, 256><<<b.tiles, 256, 0, st>>>(b.frames, b.toff, nf, b.kdev, rd, wr);
            else if (b.nt == 512)
                tgv_iter_batch_kernel<8, 512><<<b.tiles, 512, 0, st>>>(b.frames, b.toff, nf, b.kdev, rd, wr);
            else
                tgv_iter_batch_kernel<8, 256><<<b.tiles, 256, 0, st>>>(b.frames, b.toff, nf, b.kdev, rd, wr);
            i += 1;
        }
        rd = wr;
    }
    *cur = rd;
    return launch_status("tgv_iter_batch_kernel");
}

// A chunk that leaves the count in copy 0 (the next chunk, or a replay of a
// captured chunk, reads it there).
static int tgv_batch_chunk(const TgvBatch& b, int c, hipStream_t st, bool check, double tol) {
    int cur = 0;
    int rc = tgv_batch_iters(b, c, st, &cur, check, tol);
    if (rc == PSW_OK && cur != 0) {
        tgv_batch_count_home_kernel<<<1, 1, 0, st>>>(b.kdev);
        rc = launch_status("tgv_batch_count_home_kernel");
    }
    return rc;
}

// The energy test of every active frame at the count in copy cur; the
// finaliser moves the count to copy 0.
static int tgv_batch_check(const TgvBatch& b, double tol, int cur, hipStream_t st) {
    const int nf = b.nact;
    tgv_energy_batch_kernel<<<b.eblocks, TG_THREADS, 0, st>>>(b.frames, b.eoff, nf, b.kdev + cur);
    tgv_finalize_batch_kernel<<<nf, TG_THREADS, 0, st>>>(b.frames, b.kdev, cur, tol);
    return launch_status("tgv_energy_batch_kernel");
}

static int tgv_iterate_batch_group(psw_tgv* const* solvers, int nsolvers, int niters, const psw_tgv_params* prm,
                                   int variant, hipStream_t st) {
    TgvBatch b;
    int rc = tgv_batch_setup(solvers, nsolvers, prm, nullptr, variant, b, st);
    int cur = 0;
    if (rc == PSW_OK) rc = tgv_batch_iters(b, niters, st, &cur);
    if (rc == PSW_OK) PSW_TRY_HIP(hipStreamSynchronize(st), PSW_ERR_DEVICE);
    if (rc != PSW_OK) return rc;
    for (int f = 0; f < nsolvers; ++f) solvers[f]->k += niters;
    return PSW_OK;
}

static int tgv_solve_batch_group(psw_tgv* const* solvers, int nsolvers, const psw_tgv_params* prm,
                                 const double* e_star, int max_iters, int check_every, double tol, int variant,
                                 int use_graph, int* iters_done, double* energy_out, hipStream_t st) {
    TgvBatch b;
    int rc = tgv_batch_setup(solvers, nsolvers, prm, e_star, variant, b, st);
    if (rc != PSW_OK) return rc;
    // frames stopped so far, from the solvers' mapped host copies (written by
    // the finalising workgroups; no copy in the stream)
    auto ndone = [&]() {
        int n = 0;
        for (int f = 0; f < nsolvers; ++f) n += ((volatile const TgvCtrl*)solvers[f]->hmirror)->done != 0;
        return n;
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    if (e != hipSuccess) rc = hip_status(e, "hipEventCreate");
    const int full = max_iters / check_every;
    const bool graphs = use_graph && full > 0 && st != nullptr;   // the null stream cannot be captured
    // In-launch checks (two-iteration launches, chunks of at least 4): a
    // chunk's first launch sums the energy of the state the chunk starts
    // from and its second launch applies the stop test, as in the
    // single-frame solve (same tiles, so the same energies and stops); the
    // state after the last chunk is tested by the stand-alone pass at the end.
    // Otherwise each chunk ends with the stand-alone energy pass and test.
    const bool inl = b.dbl && check_every >= 4;
    // one chunk: check_every iterations and the check
    auto capture = [&]() -> int {
        if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
        hipError_t ce = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (ce != hipSuccess) return hip_status(ce, "hipStreamBeginCapture");
        int cur = 0;
        int crc;
        if (inl) {
            crc = tgv_batch_chunk(b, check_every, st, true, tol);
        } else {
            crc = tgv_batch_iters(b, check_every, st, &cur);
            if (crc == PSW_OK) crc = tgv_batch_check(b, tol, cur, st);
        }
        const hipError_t ee = hipStreamEndCapture(st, &graph);
        if (ce == hipSuccess) ce = ee;
        if (ce == hipSuccess && crc == PSW_OK) ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (crc != PSW_OK) return crc;
        return ce == hipSuccess ? PSW_OK : hip_status(ce, "hipGraph capture");
    };
    if (rc == PSW_OK && graphs) rc = capture();
    int launched = 0, chunk = 0, seen_done = 0;
    while (rc == PSW_OK && launched < max_iters) {
        const int c = std::min(check_every, max_iters - launched);
        if (exec && c == check_every && (!inl || launched > 0)) {
            e = hipGraphLaunch(exec, st);
            if (e != hipSuccess) { rc = hip_status(e, "hipGraphLaunch"); break; }
        } else {
            if (inl && launched == 0) {
                rc = tgv_batch_chunk(b, c, st, false, tol);        // nothing to test yet
            } else if (inl && c >= 4) {
                rc = tgv_batch_chunk(b, c, st, true, tol);
            } else if (inl) {
                // a short last chunk: the state it starts from is tested on its own
                rc = tgv_batch_check(b, tol, 0, st);
                if (rc == PSW_OK) rc = tgv_batch_chunk(b, c, st, false, tol);
            } else {
                int cur = 0;
                rc = tgv_batch_iters(b, c, st, &cur);
                if (rc == PSW_OK) rc = tgv_batch_check(b, tol, cur, st);
            }
            if (rc != PSW_OK) break;
        }
        launched += c;
        // two chunks in flight: the event slot holds the chunk two back; wait
        // for it and read the stopped frames (a later chunk may already have
        // stopped more; a stop is never undone), then re-arm the slot
        hipEvent_t slot = ev[chunk & 1];
        if (chunk >= 2) {
            if (hipEventSynchronize(slot) != hipSuccess) { rc = PSW_ERR_DEVICE; break; }
            const int nd = ndone();
            if (nd >= nsolvers) break;
            if (nd > seen_done) {
                // frames stopped: let the queued chunk finish, drop them from
                // the tables (their launches would only return) and recapture
                if (hipStreamSynchronize(st) != hipSuccess) { rc = PSW_ERR_DEVICE; break; }
                seen_done = ndone();
                if (seen_done >= nsolvers) break;
                std::vector<int> still;
                for (int f : b.active)
                    if (!((volatile const TgvCtrl*)solvers[f]->hmirror)->done) still.push_back(f);
                if (still.empty()) break;
                b.active.swap(still);
                rc = b.upload(st);
                if (rc == PSW_OK && graphs) rc = capture();
            }
        }
        if (rc != PSW_OK) break;
        e = hipEventRecord(slot, st);
        if (e != hipSuccess) { rc = hip_status(e, "hipEventRecord"); break; }
        chunk++;
    }
    (void)hipStreamSynchronize(st);
    if (rc == PSW_OK && inl && ndone() < nsolvers && !b.active.empty()) {
        // the state after the last chunk (count in copy 0)
        rc = tgv_batch_check(b, tol, 0, st);
        (void)hipStreamSynchronize(st);
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int k = 0; k < 2; ++k)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    if (rc != PSW_OK) return rc;
    for (int f = 0; f < nsolvers; ++f) {
        psw_tgv* t = solvers[f];
        const int k0 = t->k;
        PSW_TRY_HIP(hipMemcpy(t->hctrl, t->ctrl, sizeof(TgvCtrl), hipMemcpyDeviceToHost), PSW_ERR_COPY);
        t->k = t->hctrl->done ? t->hctrl->iters_at_done : k0 + launched;
        if (iters_done) iters_done[f] = t->k - k0;
        if (energy_out) energy_out[f] = t->hctrl->done ? t->hctrl->energy_at_done : t->hctrl->energy;
        rc = tgv_reset_ctrl(t, st);
        if (rc != PSW_OK) return rc;
    }
    PSW_TRY_HIP(hipStreamSynchronize(st), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_tgv_iterate_batch(psw_tgv* const* solvers, int nsolvers, int niters, const psw_tgv_params* prm, int variant,
                          psw_stream_t s) {
    PSW_CHECK_ARG(niters >= 0, "niters must be >= 0");
    PSW_CHECK_ARG(solvers && prm && nsolvers >= 1, "NULL pointer or no frames");
    // frames the resident solver takes, several per launch
    std::vector<int> rest;
    std::vector<std::vector<int>> packs;
    if (niters >= TR_MIN_ITERS) packs = tgv_resident_packs(solvers, nsolvers, variant, rest);
    else for (int f = 0; f < nsolvers; ++f) rest.push_back(f);
    for (const auto& pk : packs) {
        std::vector<psw_tgv*> ts;
        std::vector<psw_tgv_params> ps;
        for (int f : pk) {
            ts.push_back(solvers[f]);
            ps.push_back(prm[f]);
        }
        bool ran = false;
        const int rc = tgv_resident_batch_run(ts.data(), ps.data(), (int)ts.size(), niters, 0, 0.0, nullptr, S(s), &ran);
        if (rc != PSW_OK) return rc;
        if (!ran) {
            rest.insert(rest.end(), pk.begin(), pk.end());
            continue;
        }
        for (psw_tgv* t : ts) t->k += niters;
    }
    std::sort(rest.begin(), rest.end());
    std::vector<psw_tgv*> rs;
    std::vector<psw_tgv_params> rp;
    for (int f : rest) {
        rs.push_back(solvers[f]);
        rp.push_back(prm[f]);
    }
    if (rs.empty()) return PSW_OK;
    for (const auto& g : tgv_batch_groups(rs.data(), (int)rs.size())) {
        int rc;
        if (g.size() == 1) {
            // a group of one frame is the single-frame iteration
            PSW_CHECK_ARG(rs[g[0]], "NULL solver in the batch");
            rc = psw_tgv_iterate(rs[g[0]], niters, &rp[g[0]], s);
            if (rc == PSW_OK) PSW_TRY_HIP(hipStreamSynchronize(S(s)), PSW_ERR_DEVICE);
        } else {
            std::vector<psw_tgv*> ts;
            std::vector<psw_tgv_params> ps;
            for (int f : g) {
                ts.push_back(rs[f]);
                ps.push_back(rp[f]);
            }
            rc = tgv_iterate_batch_group(ts.data(), (int)ts.size(), niters, ps.data(), variant, S(s));
        }
        if (rc != PSW_OK) return rc;
    }
    return PSW_OK;
}

int psw_tgv_solve_batch(psw_tgv* const* solvers, int nsolvers, const psw_tgv_params* prm, const double* e_star,
                        int max_iters, int check_every, double tol, int variant, int use_graph, int* iters_done,
                        double* energy_out, psw_stream_t s) {
    PSW_CHECK_ARG(e_star, "NULL e_star");
    PSW_CHECK_ARG(max_iters >= 0 && check_every >= 1, "bad iteration counts");
    PSW_CHECK_ARG(solvers && prm && nsolvers >= 1, "NULL pointer or no frames");
    hipStream_t st = S(s);
    // frames the resident solver takes, several per launch (same stops and
    // states as one at a time)
    std::vector<int> rest;
    std::vector<std::vector<int>> packs;
    if (max_iters >= 1 && check_every >= 2) packs = tgv_resident_packs(solvers, nsolvers, variant, rest);
    else for (int f = 0; f < nsolvers; ++f) rest.push_back(f);
    for (const auto& pk : packs) {
        std::vector<psw_tgv*> ts;
        std::vector<psw_tgv_params> ps;
        std::vector<double> es;
        std::vector<int> k0;
        for (int f : pk) {
            int rc = tgv_reset_ctrl(solvers[f], st);
            if (rc != PSW_OK) return rc;
            ts.push_back(solvers[f]);
            ps.push_back(prm[f]);
            es.push_back(e_star[f]);
            k0.push_back(solvers[f]->k);
        }
        bool ran = false;
        int rc = tgv_resident_batch_run(ts.data(), ps.data(), (int)ts.size(), max_iters, check_every, tol, es.data(), st,
                                        &ran);
        if (rc != PSW_OK) return rc;
        if (!ran) {
            rest.insert(rest.end(), pk.begin(), pk.end());
            continue;
        }
        for (size_t k = 0; k < pk.size(); ++k) {
            int it = 0;
            double en = 0.0;
            rc = tgv_resident_finish(ts[k], k0[k], max_iters, tol, es[k], &ps[k], st, &it, &en);
            if (rc != PSW_OK) return rc;
            if (iters_done) iters_done[pk[k]] = it;
            if (energy_out) energy_out[pk[k]] = en;
        }
    }
    std::sort(rest.begin(), rest.end());
    std::vector<psw_tgv*> rs;
    std::vector<psw_tgv_params> rp;
    std::vector<double> re;
    for (int f : rest) {
        rs.push_back(solvers[f]);
        rp.push_back(prm[f]);
        re.push_back(e_star[f]);
    }
    if (rs.empty()) return PSW_OK;
    for (const auto& g : tgv_batch_groups(rs.data(), (int)rs.size())) {
        if (g.size() == 1) {
            // a group of one frame is the single-frame solve (same stops and states)
            const int f = g[0];
            PSW_CHECK_ARG(rs[f], "NULL solver in the batch");
            int it = 0;
            double en = 0.0;
            const int rc = psw_tgv_solve(rs[f], max_iters, check_every, tol, re[f], &rp[f], &it, &en, s);
            if (rc != PSW_OK) return rc;
            if (iters_done) iters_done[rest[f]] = it;
            if (energy_out) energy_out[rest[f]] = en;
            continue;
        }
        std::vector<psw_tgv*> ts;
        std::vector<psw_tgv_params> ps;
        std::vector<double> es;
        for (int f : g) {
            ts.push_back(rs[f]);
            ps.push_back(rp[f]);
            es.push_back(re[f]);
        }
        std::vector<int> it(g.size(), 0);
        std::vector<double> en(g.size(), 0.0);
        const int rc = tgv_solve_batch_group(ts.data(), (int)ts.size(), ps.data(), es.data(), max_iters, check_every,
                                             tol, variant, use_graph, it.data(), en.data(), st);
        if (rc != PSW_OK) return rc;
        for (size_t k = 0; k < g.size(); ++k) {
            if (iters_done) iters_done[rest[g[k]]] = it[k];
            if (energy_out) energy_out[rest[g[k]]] = en[k];
        }
    }
    return PSW_OK;
}

}  // extern "C"
