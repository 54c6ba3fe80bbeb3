// tgv_sparse.hip -- sparse-depth TGV2 densification for gfx950
// (PlaneSweep::TGVdenoiseFromSparse, src/planesweep.cpp:596-674).
//
// The reference runs three launches per iteration -- P (TGV2_kernels.cu:
// 340-365), Q (:54-85) and the sparse-depth U step (:399-432) -- over 19
// full-frame images.  Here one launch per iteration does all three for a
// 64xTY tile:
//   phase 1: u, u1x, u1y of iteration k on the tile plus a one-pixel halo
//            into LDS;
//   phase 2: p, q of iteration k+1 on the tile plus its left/top halo, into
//            LDS and, for the tile, to HBM;
//   phase 3: the primal step of the tile pixels from the LDS duals.
// Buffers read at neighbours by other tiles are never written in place: u
// (3 planes), p (2) and q (3) alternate between two slots by iteration
// parity.
//
// Semantics are those of the reference driver.  It passes the same arrays
// as u1 and u1bar to its U kernel, so in program order u1bar is the fresh u1
// after every step, and its P step reads u, not ubar; ubar is written but
// never read, so theta has no effect on the result (the entry point takes it
// for the interface).  The data weight w = (sparse > 0) and the data term
// Ds = sparse * (1.f / zfar) come from the caller's sparse depth on the fly
// (calculateWeights_sparseDepth and element_scale, same roundings).
//
// Arithmetic is FP32 throughout (the reference's double temporaries in P and
// Q become float, as in the dense solver).  HBM per pixel and iteration:
// reads u 12 B, p 8, q 12, T 12, sparse 4 (48 B); writes u 12, p 8, q 12
// (32 B): 80 B in all.
#include "psw_common.h"
#include "tgv_point.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace psw {

constexpr int TS_TX = 64;

struct SparsePtrs {
    const float* ucur;    // [3][n]: u, u1x, u1y at iteration k
    float* unext;         // [3][n] at k+1
    const float* pcur;    // [2][n]
    float* pnext;
    const float* qcur;    // [3][n]: qx, qy, qz (q_w == q_z)
    float* qnext;
    const float* T;       // [3][n]: T11, T12 (= T21), T22
    const float* sparse;  // [n] the caller's sparse depth, unscaled
    int W, H;
    float alpha0, alpha1, tau, sigma, sz;
};

// Per-pixel set-up: the tensor of ref / 255 (Anisotropic_diffusion_tensor,
// TGV2_kernels.cu:257-298, with the image scaled on load: the same float
// products the reference's element_scale stores), u = dense / zfar and zero
// u1, p, q in slot 0.
__global__ void tgv_sparse_setup_kernel(float* __restrict__ T, float* __restrict__ U, float* __restrict__ P,
                                        float* __restrict__ Q, const float* __restrict__ ref,
                                        const float* __restrict__ dense, int W, int H, float beta, float gamma,
                                        float sz) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t n = (size_t)W * H, i = (size_t)y * W + x;
    const float sc = 1.f / 255.f;
    const int xn = min(x + 1, W - 1), yn = min(y + 1, H - 1);
    const float c = ref[i] * sc;
    float t11, t12, t22;
    tensor_point(ref[(size_t)y * W + xn] * sc - c, ref[(size_t)yn * W + x] * sc - c, beta, gamma, t11, t12, t22);
    T[i] = t11;
    T[n + i] = t12;
    T[2 * n + i] = t22;
    U[i] = dense[i] * sz;
    U[n + i] = 0.f;
    U[2 * n + i] = 0.f;
    P[i] = 0.f;
    P[n + i] = 0.f;
    Q[i] = 0.f;
    Q[n + i] = 0.f;
    Q[2 * n + i] = 0.f;
}

// TGV2_updateU_sparseDepthTensor at one pixel (both fused kernels, same
// expression order): c = own duals, l / u = left / upper neighbour's (own at
// the first column / row); sp = the caller's sparse depth there.
__device__ __forceinline__ void sparse_u_step(const SparsePtrs& a, float u, float u1x, float u1y, float t11, float t12,
                                              float t22, float c_px, float c_py, float l_px, float l_py, float u_px,
                                              float u_py, float q0c, float q0l, float q1c, float q1u, float q2c,
                                              float q2l, float q2u, float sp, float& un, float& u1xn, float& u1yn) {
    const float wv = sp > 0.f ? 1.f : 0.f;
    const float ds = sp * a.sz;
    un = (u + a.tau * (a.alpha1 * (t11 * (c_px - l_px) + t12 * (c_py - l_py) + t12 * (c_px - u_px) + t22 * (c_py - u_py)) +
                       wv * ds)) /
         (1.f + a.tau * wv);
    u1xn = u1x - a.tau * (-a.alpha1 * (t11 * c_px + t12 * c_py) - a.alpha0 * (q0c - q0l + q2c - q2u));
    u1yn = u1y - a.tau * (-a.alpha1 * (t12 * c_px + t22 * c_py) - a.alpha0 * (q2c - q2l + q1c - q1u));
}

// Tile 64 x TY pixels, NT threads (variants chosen by sparse_variant()).
template <int TS_TY, int TS_THREADS>
__global__ __launch_bounds__(TS_THREADS) void tgv_sparse_iter_kernel(SparsePtrs a) {
    constexpr int HX = TS_TX + 2, HY = TS_TY + 2;   // u region: x0-1..x0+TX, y0-1..y0+TY
    constexpr int DX = TS_TX + 1, DY = TS_TY + 1;   // dual region: x0-1..x0+TX-1, y0-1..y0+TY-1
    __shared__ float sU[3][HY][HX];
    __shared__ float sP[2][DY][DX];
    __shared__ float sQ[3][DY][DX];

    const int W = a.W, H = a.H;
    const size_t n = (size_t)W * H;
    const int tiles_x = (W + TS_TX - 1) / TS_TX;
    const int ntiles = tiles_x * ((H + TS_TY - 1) / TS_TY);
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int x0 = (tile % tiles_x) * TS_TX, y0 = (tile / tiles_x) * TS_TY;
    const int tid = threadIdx.x;

    // phase 1: u, u1 at clamped coordinates (the reference's min(x+1, W-1)
    // neighbours, TGV2_kernels.cu:352-353).
    for (int i = tid; i < HX * HY; i += TS_THREADS) {
        const int hy = i / HX, hx = i - hy * HX;
        const int gx = min(max(x0 - 1 + hx, 0), W - 1);
        const int gy = min(max(y0 - 1 + hy, 0), H - 1);
        const size_t g = (size_t)gy * W + gx;
#pragma unroll
        for (int f = 0; f < 3; ++f) sU[f][hy][hx] = a.ucur[f * n + g];
    }
    __syncthreads();

    // phase 2: p, q on the tile plus its left/top halo.
    const float as1 = a.alpha1 * a.sigma, as0 = a.alpha0 * a.sigma;
    for (int i = tid; i < DX * DY; i += TS_THREADS) {
        const int dy = i / DX, dx = i - dy * DX;
        const int vx = x0 - 1 + dx, vy = y0 - 1 + dy;
        const int gx = min(max(vx, 0), W - 1), gy = min(max(vy, 0), H - 1);
        const size_t g = (size_t)gy * W + gx;
        float px = a.pcur[g], py = a.pcur[n + g];
        float qx = a.qcur[g], qy = a.qcur[n + g], qz = a.qcur[2 * n + g];
        pq_update(px, py, qx, qy, qz, a.T[g], a.T[n + g], a.T[2 * n + g], sU[0][dy][dx], sU[0][dy][dx + 1],
                  sU[0][dy + 1][dx], sU[1][dy][dx], sU[1][dy][dx + 1], sU[1][dy + 1][dx], sU[2][dy][dx],
                  sU[2][dy][dx + 1], sU[2][dy + 1][dx], as1, as0);
        sP[0][dy][dx] = px;
        sP[1][dy][dx] = py;
        sQ[0][dy][dx] = qx;
        sQ[1][dy][dx] = qy;
        sQ[2][dy][dx] = qz;
        if (dx > 0 && dy > 0 && vx < W && vy < H) {
            a.pnext[g] = px;
            a.pnext[n + g] = py;
            a.qnext[g] = qx;
            a.qnext[n + g] = qy;
            a.qnext[2 * n + g] = qz;
        }
    }
    __syncthreads();

    // phase 3: TGV2_updateU_sparseDepthTensor at the tile pixels.
    for (int i = tid; i < TS_TX * TS_TY; i += TS_THREADS) {
        const int ty = i / TS_TX, tx = i - ty * TS_TX;
        const int x = x0 + tx, y = y0 + ty;
        if (x >= W || y >= H) continue;
        const size_t g = (size_t)y * W + x;
        const int cx = tx + 1, cy = ty + 1;       // dual (and u) index of this pixel
        const int lx = x > 0 ? cx - 1 : cx;       // max(x-1, 0)
        const int uy = y > 0 ? cy - 1 : cy;       // max(y-1, 0)
        float un, u1xn, u1yn;
        sparse_u_step(a, sU[0][cy][cx], sU[1][cy][cx], sU[2][cy][cx], a.T[g], a.T[n + g], a.T[2 * n + g],
                      sP[0][cy][cx], sP[1][cy][cx], sP[0][cy][lx], sP[1][cy][lx], sP[0][uy][cx], sP[1][uy][cx],
                      sQ[0][cy][cx], sQ[0][cy][lx], sQ[1][cy][cx], sQ[1][uy][cx], sQ[2][cy][cx], sQ[2][cy][lx],
                      sQ[2][uy][cx], a.sparse[g], un, u1xn, u1yn);
        a.unext[g] = un;
        a.unext[n + g] = u1xn;
        a.unext[2 * n + g] = u1yn;
    }
}

// Two iterations (k -> k+2) of a 64 x TY tile in one launch.  Iteration k+1
// runs on the tile plus a one-pixel ring, entirely in LDS (the iteration has
// no extrapolated primal: P and Q read u_k itself); only u, p, q of k+2
// reach HBM, 80 B per pixel for both iterations.  Regions (offsets from the
// tile origin; positions outside the image read clamped coordinates and
// never feed positions inside it, whose right / lower and left / upper
// neighbours at the image border are the position itself):
//   A  u_k, then u_{k+1} on C in place      -2 .. TX+1   (TX+4)
//   B  T; p, q of k+1, then of k+2 on D     -2 .. TX     (TX+3)
//   C  u_{k+1}                              -1 .. TX
//   D  p, q of k+2                          -1 .. TX-1
// LDS index = offset + 2 in every array.  Each position's arithmetic is
// the one-iteration kernel's (pq_update, sparse_u_step), so the results
// are identical.
//
// Every phase walks whole rows of the AX-wide window and idles the positions
// outside its own columns.  With TX = 60 the window is 64 wide, so a
// position's row and column are a shift and a mask of its index.
template <int TY, int NT, int TX>
__global__ __launch_bounds__(NT) void tgv_sparse_iter2_kernel(SparsePtrs a) {
    constexpr int AX = TX + 4, AY = TY + 4, BX = TX + 3, BY = TY + 3;
    __shared__ float sU[3][AY][AX];
    __shared__ float sP[2][BY][BX];
    __shared__ float sQ[3][BY][BX];
    __shared__ float sT[3][BY][BX];

    const int W = a.W, H = a.H;
    const size_t n = (size_t)W * H;
    const int tiles_x = (W + TX - 1) / TX;
    const int ntiles = tiles_x * ((H + TY - 1) / TY);
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
    const int tid = threadIdx.x;
    auto gidx = [&](int ix, int iy) {   // clamped global index of LDS position (ix, iy)
        return (size_t)min(max(y0 - 2 + iy, 0), H - 1) * W + min(max(x0 - 2 + ix, 0), W - 1);
    };

    for (int i = tid; i < AX * AY; i += NT) {
        const int iy = i / AX, ix = i - iy * AX;
        const size_t g = gidx(ix, iy);
#pragma unroll
        for (int f = 0; f < 3; ++f) sU[f][iy][ix] = a.ucur[f * n + g];
    }
    for (int i = tid; i < AX * BY; i += NT) {
        const int iy = i / AX, ix = i % AX;
        if (ix >= BX) continue;
        const size_t g = gidx(ix, iy);
        sP[0][iy][ix] = a.pcur[g];
        sP[1][iy][ix] = a.pcur[n + g];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            sQ[f][iy][ix] = a.qcur[f * n + g];
            sT[f][iy][ix] = a.T[f * n + g];
        }
    }
    __syncthreads();

    const float as1 = a.alpha1 * a.sigma, as0 = a.alpha0 * a.sigma;
    // p, q of k+1 on B: [0, BX) x [0, BY)
    for (int i = tid; i < AX * BY; i += NT) {
        const int iy = i / AX, ix = i % AX;
        if (ix >= BX) continue;
        const int gx = x0 - 2 + ix, gy = y0 - 2 + iy;
        const int rx = gx < W - 1 ? ix + 1 : ix, dy = gy < H - 1 ? iy + 1 : iy;
        float px = sP[0][iy][ix], py = sP[1][iy][ix];
        float qx = sQ[0][iy][ix], qy = sQ[1][iy][ix], qz = sQ[2][iy][ix];
        pq_update(px, py, qx, qy, qz, sT[0][iy][ix], sT[1][iy][ix], sT[2][iy][ix], sU[0][iy][ix], sU[0][iy][rx],
                  sU[0][dy][ix], sU[1][iy][ix], sU[1][iy][rx], sU[1][dy][ix], sU[2][iy][ix], sU[2][iy][rx],
                  sU[2][dy][ix], as1, as0);
        sP[0][iy][ix] = px; sP[1][iy][ix] = py;
        sQ[0][iy][ix] = qx; sQ[1][iy][ix] = qy; sQ[2][iy][ix] = qz;
    }
    __syncthreads();

    // u of k+1 on C: [1, TX+3) x [1, TY+3), in place (a position reads only
    // its own u)
    constexpr int CX = TX + 2, CY = TY + 2;
    for (int i = tid; i < AX * CY; i += NT) {
        const int iy = i / AX + 1, ix = i % AX;
        if (ix < 1 || ix > CX) continue;
        const int gx = x0 - 2 + ix, gy = y0 - 2 + iy;
        const int lx = gx > 0 ? ix - 1 : ix, uy = gy > 0 ? iy - 1 : iy;
        float un, u1xn, u1yn;
        sparse_u_step(a, sU[0][iy][ix], sU[1][iy][ix], sU[2][iy][ix], sT[0][iy][ix], sT[1][iy][ix], sT[2][iy][ix],
                      sP[0][iy][ix], sP[1][iy][ix], sP[0][iy][lx], sP[1][iy][lx], sP[0][uy][ix], sP[1][uy][ix],
                      sQ[0][iy][ix], sQ[0][iy][lx], sQ[1][iy][ix], sQ[1][uy][ix], sQ[2][iy][ix], sQ[2][iy][lx],
                      sQ[2][uy][ix], a.sparse[gidx(ix, iy)], un, u1xn, u1yn);
        sU[0][iy][ix] = un; sU[1][iy][ix] = u1xn; sU[2][iy][ix] = u1yn;
    }
    __syncthreads();

    // p, q of k+2 on D: [1, TX+2) x [1, TY+2), in place; the tile's to HBM
    constexpr int DX = TX + 1, DY = TY + 1;
    for (int i = tid; i < AX * DY; i += NT) {
        const int iy = i / AX + 1, ix = i % AX;
        if (ix < 1 || ix > DX) continue;
        const int gx = x0 - 2 + ix, gy = y0 - 2 + iy;
        const int rx = gx < W - 1 ? ix + 1 : ix, dy = gy < H - 1 ? iy + 1 : iy;
        float px = sP[0][iy][ix], py = sP[1][iy][ix];
        float qx = sQ[0][iy][ix], qy = sQ[1][iy][ix], qz = sQ[2][iy][ix];
        pq_update(px, py, qx, qy, qz, sT[0][iy][ix], sT[1][iy][ix], sT[2][iy][ix], sU[0][iy][ix], sU[0][iy][rx],
                  sU[0][dy][ix], sU[1][iy][ix], sU[1][iy][rx], sU[1][dy][ix], sU[2][iy][ix], sU[2][iy][rx],
                  sU[2][dy][ix], as1, as0);
        sP[0][iy][ix] = px; sP[1][iy][ix] = py;
        sQ[0][iy][ix] = qx; sQ[1][iy][ix] = qy; sQ[2][iy][ix] = qz;
        if (ix >= 2 && iy >= 2 && gx < W && gy < H) {
            const size_t g = (size_t)gy * W + gx;
            a.pnext[g] = px; a.pnext[n + g] = py;
            a.qnext[g] = qx; a.qnext[n + g] = qy; a.qnext[2 * n + g] = qz;
        }
    }
    __syncthreads();

    // u of k+2 on the tile
    for (int i = tid; i < AX * TY; i += NT) {
        const int ty = i / AX, tx = i % AX;
        const int x = x0 + tx, y = y0 + ty;
        if (tx >= TX || x >= W || y >= H) continue;
        const size_t g = (size_t)y * W + x;
        const int ix = tx + 2, iy = ty + 2;
        const int lx = x > 0 ? ix - 1 : ix, uy = y > 0 ? iy - 1 : iy;
        float un, u1xn, u1yn;
        sparse_u_step(a, sU[0][iy][ix], sU[1][iy][ix], sU[2][iy][ix], sT[0][iy][ix], sT[1][iy][ix], sT[2][iy][ix],
                      sP[0][iy][ix], sP[1][iy][ix], sP[0][iy][lx], sP[1][iy][lx], sP[0][uy][ix], sP[1][uy][ix],
                      sQ[0][iy][ix], sQ[0][iy][lx], sQ[1][iy][ix], sQ[1][uy][ix], sQ[2][iy][ix], sQ[2][iy][lx],
                      sQ[2][uy][ix], a.sparse[g], un, u1xn, u1yn);
        a.unext[g] = un;
        a.unext[n + g] = u1xn;
        a.unext[2 * n + g] = u1yn;
    }
}

// M iterations per launch (M = 3..6) on a (64 - 2M) x TY tile, so that the
// staging window is 64 wide (a position's window row and column are a shift
// and a mask of its index).  The window starts M pixels up and left of the
// tile; the dual step of iteration k covers [k, R-2-k], the primal step
// [k+1, R-2-k], so the last primal step covers exactly the tile.  Positions
// outside the image are skipped, and none inside it reads one (forward
// differences clamp to the pixel at the last column / row, backward ones to
// the pixel at the first).  Each position's arithmetic is the one-iteration
// kernel's (pq_update, sparse_u_step), so the results are identical.
template <int M, int TY>
__global__ __launch_bounds__(64 * TY) void tgv_sparse_iterm_kernel(SparsePtrs a) {
    constexpr int NT = 64 * TY, TX = 64 - 2 * M;
    constexpr int RX = TX + 2 * M, RY = TY + 2 * M, RN = RX * RY;
    __shared__ float sU[3][RN];
    __shared__ float sP[2][RN];
    __shared__ float sQ[3][RN];
    __shared__ float sT[3][RN];
    const int W = a.W, H = a.H;
    const size_t n = (size_t)W * H;
    const int tiles_x = (W + TX - 1) / TX;
    const int ntiles = tiles_x * ((H + TY - 1) / TY);
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int gx0 = (tile % tiles_x) * TX - M, gy0 = (tile / tiles_x) * TY - M;
    for (int i = threadIdx.x; i < RN; i += NT) {
        const int ly = i / RX, lx = i % RX;
        const int gx = gx0 + lx, gy = gy0 + ly;
        if (gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
        const size_t g = (size_t)gy * W + gx;
#pragma unroll
        for (int f = 0; f < 3; ++f) sU[f][i] = a.ucur[f * n + g];
        if (lx <= RX - 2 && ly <= RY - 2) {   // the first dual step's region
            sP[0][i] = a.pcur[g];
            sP[1][i] = a.pcur[n + g];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                sQ[f][i] = a.qcur[f * n + g];
                sT[f][i] = a.T[f * n + g];
            }
        }
    }
    __syncthreads();
    const float as1 = a.alpha1 * a.sigma, as0 = a.alpha0 * a.sigma;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        {   // p, q of k+1 over [k, RX-2-k] x [k, RY-2-k]
            const int PY = RY - 1 - 2 * k;
            for (int i = threadIdx.x; i < PY * RX; i += NT) {
                const int ly = k + i / RX, lx = i % RX;
                const int gx = gx0 + lx, gy = gy0 + ly;
                if (lx < k || lx > RX - 2 - k || gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
                const int li = ly * RX + lx;
                const int r = gx < W - 1 ? li + 1 : li, d = gy < H - 1 ? li + RX : li;
                float px = sP[0][li], py = sP[1][li];
                float qx = sQ[0][li], qy = sQ[1][li], qz = sQ[2][li];
                pq_update(px, py, qx, qy, qz, sT[0][li], sT[1][li], sT[2][li], sU[0][li], sU[0][r], sU[0][d],
                          sU[1][li], sU[1][r], sU[1][d], sU[2][li], sU[2][r], sU[2][d], as1, as0);
                sP[0][li] = px; sP[1][li] = py;
                sQ[0][li] = qx; sQ[1][li] = qy; sQ[2][li] = qz;
                if (k == M - 1 && lx >= M && lx < M + TX && ly >= M && ly < M + TY) {
                    const size_t g = (size_t)gy * W + gx;
                    a.pnext[g] = px; a.pnext[n + g] = py;
                    a.qnext[g] = qx; a.qnext[n + g] = qy; a.qnext[2 * n + g] = qz;
                }
            }
        }
        __syncthreads();
        {   // u, u1 of k+1 over [k+1, RX-2-k] x [k+1, RY-2-k] (a position reads only its own u)
            const int UY = RY - 2 - 2 * k;
            for (int i = threadIdx.x; i < UY * RX; i += NT) {
                const int ly = k + 1 + i / RX, lx = i % RX;
                const int gx = gx0 + lx, gy = gy0 + ly;
                if (lx < k + 1 || lx > RX - 2 - k || gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
                const int li = ly * RX + lx;
                const int l = gx > 0 ? li - 1 : li, u = gy > 0 ? li - RX : li;
                const size_t g = (size_t)gy * W + gx;
                float un, u1xn, u1yn;
                sparse_u_step(a, sU[0][li], sU[1][li], sU[2][li], sT[0][li], sT[1][li], sT[2][li], sP[0][li],
                              sP[1][li], sP[0][l], sP[1][l], sP[0][u], sP[1][u], sQ[0][li], sQ[0][l], sQ[1][li],
                              sQ[1][u], sQ[2][li], sQ[2][l], sQ[2][u], a.sparse[g], un, u1xn, u1yn);
                if (k == M - 1) {
                    a.unext[g] = un;
                    a.unext[n + g] = u1xn;
                    a.unext[2 * n + g] = u1yn;
                } else {
                    sU[0][li] = un; sU[1][li] = u1xn; sU[2][li] = u1yn;
                }
            }
        }
        if (k < M - 1) __syncthreads();
    }
}

__global__ void tgv_sparse_output_kernel(float* __restrict__ out, const float* __restrict__ u, size_t n, float zfar) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = u[i] * zfar;
}

// Launch shapes: 0 = 64x4 / 256 threads, 1 = 64x8 / 512 (default),
// 2 = 64x8 / 256, 3 = 64x16 / 512.  Read from PSW_TGV_SPARSE_VARIANT on each
// call; results do not depend on it.
static int sparse_variant() {
    const char* e = std::getenv("PSW_TGV_SPARSE_VARIANT");
    const int v = e ? std::atoi(e) : 1;
    return (v >= 0 && v <= 3) ? v : 1;
}

template <int TY, int NT>
static hipError_t sparse_launch(const SparsePtrs& a, hipStream_t hs) {
    const int ntiles = ((a.W + TS_TX - 1) / TS_TX) * ((a.H + TY - 1) / TY);
    tgv_sparse_iter_kernel<TY, NT><<<ntiles, NT, 0, hs>>>(a);
    return hipGetLastError();
}

// TX x TY tiles with NT threads (NT >= TX * TY: the halo regions, up to
// (TX + 4) x (TY + 4) positions, take fewer passes).
template <int TY, int NT, int TX>
static hipError_t sparse_launch2(const SparsePtrs& a, hipStream_t hs) {
    static_assert(NT >= TX * TY, "at least one thread per tile pixel");
    const int ntiles = ((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY);
    tgv_sparse_iter2_kernel<TY, NT, TX><<<ntiles, NT, 0, hs>>>(a);
    return hipGetLastError();
}

// 60-wide tiles for the two-iteration kernel (a 64-wide window) unless
// PSW_TGV_SPARSE_NARROW=0 (64-wide tiles); identical results.
static bool sparse_narrow() {
    const char* e = std::getenv("PSW_TGV_SPARSE_NARROW");
    return !(e && std::atoi(e) == 0);
}

// Iterations per launch at most (PSW_TGV_SPARSE_M: 2, the two-iteration
// kernel only, 3 to 6, 4 the default; identical results).  At 640x480:
// 4.75, 4.06-4.14, 3.72-3.80, 3.72-3.80 and 5.0 us per iteration for 2..6.
static int sparse_max_per_launch() {
    const char* e = std::getenv("PSW_TGV_SPARSE_M");
    const int v = e ? std::atoi(e) : 4;
    return v >= 2 && v <= 6 ? v : 4;
}

template <int M>
static hipError_t sparse_launchm(const SparsePtrs& a, hipStream_t hs) {
    constexpr int TX = 64 - 2 * M, TY = 8;
    const int ntiles = ((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY);
    tgv_sparse_iterm_kernel<M, TY><<<ntiles, 64 * TY, 0, hs>>>(a);
    return hipGetLastError();
}

// Two iterations per launch unless PSW_TGV_SPARSE_DOUBLE=0 (identical results).
static bool sparse_double() {
    const char* e = std::getenv("PSW_TGV_SPARSE_DOUBLE");
    return !(e && std::atoi(e) == 0);
}

// Workspace: T 3n, then two slots of (u 3n, p 2n, q 3n).
static size_t sparse_floats(size_t n) { return 3 * n + 2 * 8 * n; }

// ---- resident solver ---------------------------------------------------------
// All iterations in one cooperative launch, on the scheme of the dense and
// TVL1 resident solvers: workgroup t owns a 64 x 20 tile (ten waves, thread
// (wave w, lane x) the pixels of rows 2w and 2w+1 in column x) and keeps its
// pixels' u, u1x, u1y, p, q, tensor and sparse depth in registers for the
// whole run; the tiles trade only their edges' u, u1x, u1y each iteration.
// The dual step of the row above the tile and of the column left of it (the
// ring, whose p and q the primal step reads) is recomputed here from the
// halo, so one exchange per iteration suffices.  The per-pixel arithmetic is
// pq_update and sparse_u_step, so the results are those of the fused
// launches.  Exchange words are 64-bit (iteration tag, float bits) in
// uncached device memory, double-buffered by the iteration's parity; every
// wait is bounded, and a launch whose workgroups are not all resident in
// time is called off before any wait and any write.
constexpr int TSR_TX = 64, TSR_TY = 20, TSR_NT = TSR_TX * (TSR_TY / 2);
constexpr int TSR_F = 3;                                          // u, u1x, u1y
constexpr int TSR_TOP = 0, TSR_BOT = TSR_F * TSR_TX, TSR_LEFT = 2 * TSR_F * TSR_TX,
              TSR_RIGHT = 2 * TSR_F * TSR_TX + TSR_F * TSR_TY;     // word f * (64 | 20) + c of each edge
constexpr int TSR_XW = 2 * TSR_F * (TSR_TX + TSR_TY);            // words a tile publishes per iteration
constexpr int TSR_HALO = TSR_F * (2 * TSR_TX + 2 * TSR_TY + 2);  // halo words a tile reads per iteration
constexpr int TSR_RING0 = 2 * TSR_TX, TSR_RING = TSR_TX + TSR_TY;  // ring threads: waves 2 and 3
constexpr int TSR_MAXTILES = 256;
constexpr int TSR_PAD = 6144;                                     // 24 KB, see tgv_sparse_resident_kernel
static_assert(TSR_HALO <= TSR_NT && TSR_RING0 + TSR_RING + TSR_TY <= 4 * TSR_TX,
              "one halo word per thread; the ring and the right column's workers in waves 2 and 3");

struct SparseResArgs {
    SparsePtrs a;               // the start state in ucur / pcur / qcur, the end state to unext / pnext / qnext
    int niters;
    int tiles_x, tiles_y, ntiles;
    unsigned long long* xch;    // [2][ntiles][TSR_XW]
    unsigned* err;              // [0] error, [1] started, [2] go (1) / called off (2), [3..7] first failure:
                                // tile, word offset from xch, awaited tag, seen tag, iteration
    unsigned spin, start_spin;
};

__device__ __forceinline__ unsigned long long tsr_load(unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned tsr_load32(unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void tsr_put(unsigned long long* p, unsigned tag, float v) {
    __hip_atomic_store(p, ((unsigned long long)tag << 32) | __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Poll word p until it carries `tag`; its float bits go to v.  False when
// the spin budget runs out (the first failure is recorded) or another wait
// has failed.
__device__ bool tsr_get(const SparseResArgs& A, unsigned long long* p, unsigned tag, float& v, int tile, int j) {
    for (unsigned it = 0; it < A.spin; ++it) {
        const unsigned long long w = tsr_load(p);
        if ((unsigned)(w >> 32) == tag) {
            v = __uint_as_float((unsigned)w);
            return true;
        }
        if ((it & 63) == 63 && tsr_load32(A.err) != 0u) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    unsigned expected = 0u;
    if (__hip_atomic_compare_exchange_strong(A.err, &expected, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)) {
        const unsigned vals[5] = {(unsigned)tile, (unsigned)(p - A.xch), tag, (unsigned)(tsr_load(p) >> 32),
                                  (unsigned)j};
        for (int k = 0; k < 5; ++k) __hip_atomic_store(A.err + 3 + k, vals[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return false;
}

// Halo word w of tile (tcx, tcy): field f = w % 3 of cell c = w / 3, the
// cells being the row above (c < 64), the above-right corner, the column
// left, the below-left corner, the column right and the row below.  False
// where the cell is outside the image or has no neighbour tile; else the
// source tile's word, the LDS cell (row ly, column lx of the 22 x 66 arrays)
// and the image position.
__device__ inline bool tsr_halo(const SparseResArgs& A, int tcx, int tcy, int w, int& off, int& ly, int& lx, int& gx,
                                int& gy) {
    constexpr int TX = TSR_TX, TY = TSR_TY;
    const int W = A.a.W, H = A.a.H;
    const int x0 = tcx * TX, y0 = tcy * TY, tile = tcy * A.tiles_x + tcx;
    const bool has_l = tcx > 0, has_r = tcx < A.tiles_x - 1, has_u = tcy > 0, has_d = tcy < A.tiles_y - 1;
    const int c = w / TSR_F, f = w - (w / TSR_F) * TSR_F;
    int src, word;
    bool ok;
    if (c < TX) {
        gx = x0 + c; gy = y0 - 1; src = tile - A.tiles_x; word = TSR_BOT + f * TX + c; ly = 0; lx = c + 1;
        ok = has_u && gx < W;
    } else if (c == TX) {
        gx = x0 + TX; gy = y0 - 1; src = tile - A.tiles_x + 1; word = TSR_BOT + f * TX; ly = 0; lx = TX + 1;
        ok = has_u && has_r;
    } else if (c < TX + 1 + TY) {
        const int r = c - (TX + 1);
        gx = x0 - 1; gy = y0 + r; src = tile - 1; word = TSR_RIGHT + f * TY + r; ly = r + 1; lx = 0;
        ok = has_l && gy < H;
    } else if (c == TX + 1 + TY) {
        gx = x0 - 1; gy = y0 + TY; src = tile + A.tiles_x - 1; word = TSR_TOP + f * TX + TX - 1; ly = TY + 1; lx = 0;
        ok = has_l && has_d;
    } else if (c < TX + 2 + 2 * TY) {
        const int r = c - (TX + 2 + TY);
        gx = x0 + TX; gy = y0 + r; src = tile + 1; word = TSR_LEFT + f * TY + r; ly = r + 1; lx = TX + 1;
        ok = has_r && gy < H;
    } else {
        const int cc = c - (TX + 2 + 2 * TY);
        gx = x0 + cc; gy = y0 + TY; src = tile + A.tiles_x; word = TSR_TOP + f * TX + cc; ly = TY + 1; lx = cc + 1;
        ok = has_d && gx < W;
    }
    off = src * TSR_XW + word;
    return ok;
}

__global__ __launch_bounds__(TSR_NT) void tgv_sparse_resident_kernel(const SparseResArgs A) {
    constexpr int TX = TSR_TX, TY = TSR_TY;
    __shared__ float sU[TSR_F][TY + 2][TX + 2];   // u, u1x, u1y: the tile and its halo
    __shared__ float sPQ[5][TY + 1][TX + 1];      // px, py, qx, qy, qz: the tile, the row above, the column left
    __shared__ float sT[3][TY + 1][TX + 1];       // the tensor of the ring and of the right column
    __shared__ int sFlag[1];
    // LDS held only so that a CU takes one workgroup (the tiles run in
    // lockstep, so two on one CU would halve the whole solve's speed)
    __shared__ float sPad[TSR_PAD];
    const SparsePtrs& a = A.a;
    const int W = a.W, H = a.H;
    const size_t n = (size_t)W * H;
    const int tile = xcd_tile(blockIdx.x, A.ntiles);
    const int tcx = tile % A.tiles_x, tcy = tile / A.tiles_x;
    const int x0 = tcx * TX, y0 = tcy * TY;
    const int tid = threadIdx.x, tx = tid % TX;
    const int wv = __builtin_amdgcn_readfirstlane(tid / TX);
    const int x = x0 + tx;
    const float as1 = a.alpha1 * a.sigma, as0 = a.alpha0 * a.sigma;
    if (tid == 0) {
        sFlag[0] = 0;
        // the last workgroup to start gives the go
        const unsigned before = __hip_atomic_fetch_add(A.err + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (before + 1 == (unsigned)A.ntiles) {
            unsigned expected = 0u;
            __hip_atomic_compare_exchange_strong(A.err + 2, &expected, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // own pixels: the start state and the constants
    float uf[2][TSR_F], px[2], py[2], qx[2], qy[2], qz[2], t11[2], t12[2], t22[2], sp[2];
    bool in[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int y = y0 + 2 * wv + h;
        in[h] = x < W && y < H;
        const size_t g = (size_t)min(y, H - 1) * W + min(x, W - 1);
#pragma unroll
        for (int f = 0; f < TSR_F; ++f) {
            uf[h][f] = a.ucur[f * n + g];
            sU[f][2 * wv + h + 1][tx + 1] = uf[h][f];
        }
        px[h] = a.pcur[g]; py[h] = a.pcur[n + g];
        qx[h] = a.qcur[g]; qy[h] = a.qcur[n + g]; qz[h] = a.qcur[2 * n + g];
        t11[h] = a.T[g]; t12[h] = a.T[n + g]; t22[h] = a.T[2 * n + g];
        sp[h] = a.sparse[g];
        if (tx == TX - 1) {   // the right column's dual step runs from LDS (see the loop)
            const int cy = 2 * wv + h + 1;
            sPQ[0][cy][TX] = px[h]; sPQ[1][cy][TX] = py[h];
            sPQ[2][cy][TX] = qx[h]; sPQ[3][cy][TX] = qy[h]; sPQ[4][cy][TX] = qz[h];
            sT[0][cy][TX] = t11[h]; sT[1][cy][TX] = t12[h]; sT[2][cy][TX] = t22[h];
        }
    }
    // the ring: its duals and tensor of the start state
    const int rid = tid - TSR_RING0;
    bool ring = false;
    int rgx = 0, rgy = 0, rlx = 0, rly = 0;
    if (rid >= 0 && rid < TSR_RING) {
        if (rid < TX) { rgx = x0 + rid; rgy = y0 - 1; rlx = rid + 1; rly = 0; ring = tcy > 0 && rgx < W; }
        else { rgx = x0 - 1; rgy = y0 + rid - TX; rlx = 0; rly = rid - TX + 1; ring = tcx > 0 && rgy < H; }
        if (ring) {
            const size_t g = (size_t)rgy * W + rgx;
            sPQ[0][rly][rlx] = a.pcur[g]; sPQ[1][rly][rlx] = a.pcur[n + g];
            sPQ[2][rly][rlx] = a.qcur[g]; sPQ[3][rly][rlx] = a.qcur[n + g]; sPQ[4][rly][rlx] = a.qcur[2 * n + g];
            sT[0][rly][rlx] = a.T[g]; sT[1][rly][rlx] = a.T[n + g]; sT[2][rly][rlx] = a.T[2 * n + g];
        }
    }
    // the halo of the start state; later ones come from the neighbours
    int hoff = -1, hly = 0, hlx = 0;
    const int hf = tid % TSR_F;
    if (tid < TSR_HALO) {
        int gx, gy;
        if (tsr_halo(A, tcx, tcy, tid, hoff, hly, hlx, gx, gy)) sU[hf][hly][hlx] = a.ucur[hf * n + (size_t)gy * W + gx];
        else hoff = -1;
    }
    // every workgroup resident before any waits on a neighbour
    if (tid == 0) {
        unsigned go = tsr_load32(A.err + 2);
        for (unsigned it = 0; go == 0u && it < A.start_spin; ++it) {
            __builtin_amdgcn_s_sleep(2);
            go = tsr_load32(A.err + 2);
        }
        if (go == 0u) {
            unsigned expected = 0u;
            __hip_atomic_compare_exchange_strong(A.err + 2, &expected, 2u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            go = tsr_load32(A.err + 2);
        }
        if (go != 1u) sFlag[0] = 1;
    }
    __syncthreads();
    if (sFlag[0]) return;   // called off: nothing written

    unsigned long long* const xb0 = A.xch + (size_t)tile * TSR_XW;   // this tile's words (parity 0)
    bool failed = false;
    // The dual step of a pixel reads u, u1 right of it and below it, so only
    // the tile's right column and bottom row (and the ring) need the halo:
    // the others are done before the wait for the neighbours' edges, in its
    // shadow.  The bottom row is wave 9's second row, done by that wave after
    // the wait; the right column (one lane of every wave) is done by lanes
    // 20..39 of wave 3 from LDS, and its owners take their duals back from
    // there for the primal step.
    const int rc = tid - (TSR_RING0 + TSR_RING);   // right-column worker: pixel (x0 + 63, y0 + rc)
    const bool rcol = rc >= 0 && rc < TY && x0 + TX - 1 < W && y0 + rc < H;
    const bool last_col = tx == TX - 1;
    for (int j = 0; j < A.niters; ++j) {
        // p, q of j + 1 (from u of state j) off the right column and the bottom row
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = 2 * wv + h, y = y0 + row;
            if (row == TY - 1) continue;   // wave-uniform
            float rt[TSR_F], dn[TSR_F];
#pragma unroll
            for (int f = 0; f < TSR_F; ++f) {   // read by every lane, then chosen
                const float right = sU[f][row + 1][tx + 2], down = sU[f][row + 2][tx + 1];
                rt[f] = x < W - 1 ? right : uf[h][f];
                dn[f] = y < H - 1 ? down : uf[h][f];
            }
            if (in[h] && !last_col) {
                pq_update(px[h], py[h], qx[h], qy[h], qz[h], t11[h], t12[h], t22[h], uf[h][0], rt[0], dn[0], uf[h][1],
                          rt[1], dn[1], uf[h][2], rt[2], dn[2], as1, as0);
                sPQ[0][row + 1][tx + 1] = px[h]; sPQ[1][row + 1][tx + 1] = py[h];
                sPQ[2][row + 1][tx + 1] = qx[h]; sPQ[3][row + 1][tx + 1] = qy[h]; sPQ[4][row + 1][tx + 1] = qz[h];
            }
        }
        // the neighbours' edges of state j (from the start state at j = 0)
        if (j > 0 && hoff >= 0) {
            float v;
            if (tsr_get(A, A.xch + (size_t)(j & 1) * A.ntiles * TSR_XW + hoff, (unsigned)j, v, tile, j))
                sU[hf][hly][hlx] = v;
            else sFlag[0] = 1;
        }
        __syncthreads();
        if (sFlag[0]) { failed = true; break; }
        // p, q of the bottom row, the ring and the right column
        if (2 * wv + 1 == TY - 1) {   // wave-uniform
            const int row = TY - 1, y = y0 + row;
            float rt[TSR_F], dn[TSR_F];
#pragma unroll
            for (int f = 0; f < TSR_F; ++f) {
                const float right = sU[f][row + 1][tx + 2], down = sU[f][row + 2][tx + 1];
                rt[f] = x < W - 1 ? right : uf[1][f];
                dn[f] = y < H - 1 ? down : uf[1][f];
            }
            if (in[1] && !last_col) {
                pq_update(px[1], py[1], qx[1], qy[1], qz[1], t11[1], t12[1], t22[1], uf[1][0], rt[0], dn[0], uf[1][1],
                          rt[1], dn[1], uf[1][2], rt[2], dn[2], as1, as0);
                sPQ[0][row + 1][tx + 1] = px[1]; sPQ[1][row + 1][tx + 1] = py[1];
                sPQ[2][row + 1][tx + 1] = qx[1]; sPQ[3][row + 1][tx + 1] = qy[1]; sPQ[4][row + 1][tx + 1] = qz[1];
            }
        }
        if (ring || rcol) {
            // the ring's cell, or the right column's (the same dual step from LDS)
            const int cly = ring ? rly : rc + 1, clx = ring ? rlx : TX;
            const int cgx = ring ? rgx : x0 + TX - 1, cgy = ring ? rgy : y0 + rc;
            const int rr = cgx < W - 1 ? clx + 1 : clx, rdn = cgy < H - 1 ? cly + 1 : cly;
            float rpx = sPQ[0][cly][clx], rpy = sPQ[1][cly][clx];
            float rqx = sPQ[2][cly][clx], rqy = sPQ[3][cly][clx], rqz = sPQ[4][cly][clx];
            pq_update(rpx, rpy, rqx, rqy, rqz, sT[0][cly][clx], sT[1][cly][clx], sT[2][cly][clx], sU[0][cly][clx],
                      sU[0][cly][rr], sU[0][rdn][clx], sU[1][cly][clx], sU[1][cly][rr], sU[1][rdn][clx],
                      sU[2][cly][clx], sU[2][cly][rr], sU[2][rdn][clx], as1, as0);
            sPQ[0][cly][clx] = rpx; sPQ[1][cly][clx] = rpy;
            sPQ[2][cly][clx] = rqx; sPQ[3][cly][clx] = rqy; sPQ[4][cly][clx] = rqz;
        }
        __syncthreads();
        // u, u1 of j + 1; the tile's edges for the neighbours
        const bool publish = j + 1 < A.niters;
        unsigned long long* const xb = xb0 + (size_t)((j + 1) & 1) * A.ntiles * TSR_XW;
        const unsigned tag = (unsigned)(j + 1);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = 2 * wv + h, y = y0 + row;
            float lp[5], up[5], op[5];
#pragma unroll
            for (int f = 0; f < 5; ++f) {   // read by every lane, then chosen
                lp[f] = sPQ[f][row + 1][tx];
                up[f] = sPQ[f][row][tx + 1];
                op[f] = sPQ[f][row + 1][tx + 1];
            }
            if (in[h]) {
                if (last_col) {   // the right column's duals, from its worker
                    px[h] = op[0]; py[h] = op[1]; qx[h] = op[2]; qy[h] = op[3]; qz[h] = op[4];
                }
                const bool hl = x > 0, hu = y > 0;
                float un, u1xn, u1yn;
                sparse_u_step(a, uf[h][0], uf[h][1], uf[h][2], t11[h], t12[h], t22[h], px[h], py[h],
                              hl ? lp[0] : px[h], hl ? lp[1] : py[h], hu ? up[0] : px[h], hu ? up[1] : py[h], qx[h],
                              hl ? lp[2] : qx[h], qy[h], hu ? up[3] : qy[h], qz[h], hl ? lp[4] : qz[h],
                              hu ? up[4] : qz[h], sp[h], un, u1xn, u1yn);
                uf[h][0] = un; uf[h][1] = u1xn; uf[h][2] = u1yn;
                if (publish) {
#pragma unroll
                    for (int f = 0; f < TSR_F; ++f) {
                        if (row == 0) tsr_put(xb + TSR_TOP + f * TX + tx, tag, uf[h][f]);
                        if (row == TY - 1) tsr_put(xb + TSR_BOT + f * TX + tx, tag, uf[h][f]);
                        if (tx == 0) tsr_put(xb + TSR_LEFT + f * TY + row, tag, uf[h][f]);
                        if (tx == TX - 1) tsr_put(xb + TSR_RIGHT + f * TY + row, tag, uf[h][f]);
                    }
                }
            }
        }
        if (!publish) break;
        // the own u, u1 of state j + 1 for the next dual step (its readers
        // have passed the barrier above; the next one orders it before them)
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (in[h]) {
#pragma unroll
                for (int f = 0; f < TSR_F; ++f) sU[f][2 * wv + h + 1][tx + 1] = uf[h][f];
            }
        __syncthreads();
    }
    if (failed) return;
    // the pad is read on a branch never taken (the host passes niters >= 2),
    // a select with no arithmetic, so that it stays allocated
    if (A.niters < 0) {
        sPad[tid] = uf[0][0];
        __syncthreads();
        uf[0][0] = sPad[(tid + 1) % TSR_NT];
    }
    // the end state
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (!in[h]) continue;
        const size_t g = (size_t)(y0 + 2 * wv + h) * W + x;
#pragma unroll
        for (int f = 0; f < TSR_F; ++f) a.unext[f * n + g] = uf[h][f];
        a.pnext[g] = px[h]; a.pnext[n + g] = py[h];
        a.qnext[g] = qx[h]; a.qnext[n + g] = qy[h]; a.qnext[2 * n + g] = qz[h];
    }
}

// The exchange buffer of the resident solver on each device: uncached
// memory for the largest launch (2 x 256 tiles x 504 edge words, 8 error
// words), allocated once and kept for the process (uncached memory that a
// kernel has read is never freed).  Also the CU count, and the tile counts
// whose launches were called off (those keep using the fused launches).
struct TsrDev {
    std::mutex mu;
    unsigned long long* buf[64] = {};
    int cus[64] = {};
    std::vector<int> declined[64];
    // one resident run per device at a time (they share the exchange buffer
    // and would compete for the CUs), and its error words on the host
    std::mutex run[64];
    unsigned* herr[64] = {};
};
static TsrDev& tsr_dev() {
    static TsrDev* d = new TsrDev();
    return *d;
}
constexpr size_t TSR_WORDS = 2 * (size_t)TSR_MAXTILES * TSR_XW + 4;

// The resident solver runs unless PSW_TGV_SPARSE_RESIDENT=0, for frames of
// at most one 64 x 20 tile per CU (and at most 256 tiles) and runs of at
// least two iterations; *buf receives the device's exchange buffer.
// PSW_TGV_SPARSE_RESIDENT=2 makes a run of two or more iterations that
// cannot use it an error (tests).
static bool sparse_resident_use(int W, int H, int niters, unsigned long long** buf) {
    const char* e = std::getenv("PSW_TGV_SPARSE_RESIDENT");
    if ((e && std::atoi(e) == 0) || niters < 2) return false;
    const int nt = ((W + TSR_TX - 1) / TSR_TX) * ((H + TSR_TY - 1) / TSR_TY);
    if (nt > TSR_MAXTILES) return false;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return false;
    }
    TsrDev& D = tsr_dev();
    std::lock_guard<std::mutex> lock(D.mu);
    if (D.cus[dev] == 0) {
        int per = 0, ncu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)tgv_sparse_resident_kernel, TSR_NT, 0) !=
                hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            (void)hipGetLastError();
            per = 0;
        }
        D.cus[dev] = per >= 1 && ncu > 0 ? ncu : -1;
    }
    if (D.cus[dev] < nt) return false;
    for (int d : D.declined[dev])
        if (d == nt) return false;
    if (!D.buf[dev]) {
        void* b = nullptr;
        if (hipExtMallocWithFlags(&b, TSR_WORDS * sizeof(unsigned long long), hipDeviceMallocUncached) != hipSuccess) {
            (void)hipGetLastError();
            D.cus[dev] = -1;
            return false;
        }
        D.buf[dev] = static_cast<unsigned long long*>(b);
    }
    *buf = D.buf[dev];
    return true;
}

static bool sparse_resident_required() {
    const char* e = std::getenv("PSW_TGV_SPARSE_RESIDENT");
    return e && std::atoi(e) == 2;
}

static void sparse_resident_decline(int nt) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    TsrDev& D = tsr_dev();
    std::lock_guard<std::mutex> lock(D.mu);
    D.declined[dev].push_back(nt);
}

static unsigned tsr_env_u32(const char* name, unsigned dflt) {
    const char* e = std::getenv(name);
    return e ? (unsigned)std::strtoul(e, nullptr, 10) : dflt;
}

// niters iterations from a's current state to its next one in one
// cooperative launch.  *ran is false, with nothing written, when the runtime
// refuses the launch or the launch is called off.  Synchronous (it reads the
// error words); runs on one device take turns, since they share its
// exchange buffer.  Polls per wait: PSW_TGV_SPARSE_RESIDENT_SPIN (2^22), of the
// start: PSW_TGV_SPARSE_RESIDENT_START_SPIN (2^16; set, a called-off launch
// is not remembered).
static int sparse_resident_run(const SparsePtrs& a, unsigned long long* buf, int niters, hipStream_t st, bool* ran) {
    *ran = false;
    SparseResArgs A;
    A.a = a;
    A.niters = niters;
    A.tiles_x = (a.W + TSR_TX - 1) / TSR_TX;
    A.tiles_y = (a.H + TSR_TY - 1) / TSR_TY;
    A.ntiles = A.tiles_x * A.tiles_y;
    A.xch = buf;
    A.err = reinterpret_cast<unsigned*>(buf + 2 * (size_t)A.ntiles * TSR_XW);
    A.spin = tsr_env_u32("PSW_TGV_SPARSE_RESIDENT_SPIN", 1u << 22);
    A.start_spin = tsr_env_u32("PSW_TGV_SPARSE_RESIDENT_START_SPIN", 1u << 16);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return PSW_OK;
    }
    TsrDev& D = tsr_dev();
    std::lock_guard<std::mutex> run_lock(D.run[dev]);
    if (!D.herr[dev]) {
        std::lock_guard<std::mutex> lock(D.mu);
        PSW_TRY_HIP(hipHostMalloc((void**)&D.herr[dev], 8 * sizeof(unsigned), hipHostMallocDefault), PSW_ERR_ALLOC);
    }
    unsigned* const he = D.herr[dev];
    // tags that match no iteration, and clear error words
    int rc = psw_fill_u32(reinterpret_cast<uint32_t*>(buf), 0xFFFFFFFFu, 4 * (size_t)A.ntiles * TSR_XW, st);
    if (rc == PSW_OK) rc = psw_fill_u32(A.err, 0u, 8, st);
    if (rc != PSW_OK) return rc;
    void* args[] = {&A};
    if (hipLaunchCooperativeKernel((const void*)tgv_sparse_resident_kernel, dim3((unsigned)A.ntiles), dim3(TSR_NT),
                                   args, 0, st) != hipSuccess) {
        (void)hipGetLastError();
        return PSW_OK;
    }
    int out = PSW_OK;
    if (hipMemcpyAsync(he, A.err, 8 * sizeof(unsigned), hipMemcpyDeviceToHost, st) != hipSuccess) out = PSW_ERR_COPY;
    else if (hipStreamSynchronize(st) != hipSuccess) out = PSW_ERR_DEVICE;
    if (out == PSW_OK) {
        if (he[2] != 1u) {   // called off before any wait on a neighbour: nothing was written
            if (!std::getenv("PSW_TGV_SPARSE_RESIDENT_START_SPIN")) sparse_resident_decline(A.ntiles);
        } else {
            *ran = true;
            if (he[0]) {
                set_error("resident sparse-depth TGV2 solver: tile %u waited %u polls for word %u of tile %u "
                          "(iteration tag %u, saw %u) at iteration %u; the result is undefined",
                          he[3], A.spin, he[4] % TSR_XW, (he[4] % ((unsigned)A.ntiles * TSR_XW)) / TSR_XW, he[5],
                          he[6], he[7]);
                out = PSW_ERR_DEVICE;
            }
        }
    } else {
        (void)hipGetLastError();
    }
    return out;
}

}  // namespace psw

using namespace psw;

extern "C" {

size_t psw_tgv_sparse_workspace_bytes(int width, int height) {
    if (width <= 0 || height <= 0) return 0;
    return sparse_floats((size_t)width * height) * sizeof(float);
}

int psw_tgv_sparse_denoise(const float* d_ref, const float* d_dense, const float* d_sparse, int width, int height,
                           float zfar, int niters, double alpha0, double alpha1, double tau, double sigma,
                           double theta, double beta, double gamma, void* d_workspace, size_t workspace_bytes,
                           float* d_out, psw_stream_t s) {
    (void)theta;
    PSW_CHECK_ARG(d_ref && d_dense && d_sparse && d_out, "NULL image pointer");
    PSW_CHECK_ARG(width > 0 && height > 0 && (size_t)width * height < (1ull << 31), "bad image size");
    PSW_CHECK_ARG(niters >= 0, "niters < 0");
    PSW_CHECK_ARG(zfar > 0.f && std::isfinite(zfar), "zfar must be positive and finite");
    const size_t n = (size_t)width * height;
    const size_t need = sparse_floats(n) * sizeof(float);
    PSW_CHECK_ARG(!d_workspace || workspace_bytes >= need, "workspace too small");
    float* ws = static_cast<float*>(d_workspace);
    bool owned = false;
    if (!ws) {   // stream-ordered, from the library's workspace pool
        void* w = nullptr;
        const int rc = workspace_alloc_async(&w, need, S(s));
        if (rc != PSW_OK) return rc;
        ws = static_cast<float*>(w);
        owned = true;
    }
    float* T = ws;
    float* slot[2] = {ws + 3 * n, ws + 11 * n};   // each: u 3n, p 2n, q 3n
    hipStream_t hs = S(s);
    const float sz = 1.f / zfar;
    const dim3 blk(64, 4), grd((width + 63) / 64, (height + 3) / 4);
    tgv_sparse_setup_kernel<<<grd, blk, 0, hs>>>(T, slot[0], slot[0] + 3 * n, slot[0] + 5 * n, d_ref, d_dense, width,
                                                  height, (float)beta, (float)gamma, sz);
    hipError_t e = hipGetLastError();
    const int variant = sparse_variant();
    SparsePtrs a;
    a.T = T;
    a.sparse = d_sparse;
    a.W = width;
    a.H = height;
    a.alpha0 = (float)alpha0;
    a.alpha1 = (float)alpha1;
    a.tau = (float)tau;
    a.sigma = (float)sigma;
    a.sz = sz;
    // Each launch reads one slot and writes the other (a two-iteration launch
    // too), so the current state is in slot (launches & 1).
    const bool dbl = sparse_double();
    const bool narrow = sparse_narrow();
    const int mmax = sparse_max_per_launch();
    int launches = 0, k0 = 0;
    if (e == hipSuccess) {
        unsigned long long* xbuf = nullptr;
        bool ran = false;
        if (sparse_resident_use(width, height, niters, &xbuf)) {
            a.ucur = slot[0];
            a.unext = slot[1];
            a.pcur = slot[0] + 3 * n;
            a.pnext = slot[1] + 3 * n;
            a.qcur = slot[0] + 5 * n;
            a.qnext = slot[1] + 5 * n;
            const int rc = sparse_resident_run(a, xbuf, niters, hs, &ran);
            if (rc != PSW_OK) {
                if (owned) (void)workspace_free_async(ws, hs);
                return rc;
            }
            if (ran) {
                launches = 1;
                k0 = niters;
            }
        }
        if (!ran && niters >= 2 && sparse_resident_required()) {
            if (owned) (void)workspace_free_async(ws, hs);
            set_error("psw_tgv_sparse_denoise: PSW_TGV_SPARSE_RESIDENT=2 and the resident solver did not run "
                      "(%dx%d, %d iterations)", width, height, niters);
            return PSW_ERR_DEVICE;
        }
    }
    for (int k = k0; k < niters && e == hipSuccess; ++launches) {
        float* c = slot[launches & 1];
        float* nx = slot[(launches + 1) & 1];
        a.ucur = c;
        a.unext = nx;
        a.pcur = c + 3 * n;
        a.pnext = nx + 3 * n;
        a.qcur = c + 5 * n;
        a.qnext = nx + 5 * n;
        if (dbl && mmax > 2 && niters - k >= 3) {
            const int m = std::min(mmax, niters - k);
            e = m == 6 ? sparse_launchm<6>(a, hs)
                : m == 5 ? sparse_launchm<5>(a, hs)
                : m == 4 ? sparse_launchm<4>(a, hs)
                         : sparse_launchm<3>(a, hs);
            k += m;
            continue;
        }
        if (dbl && niters - k >= 2) {
            // 768 threads measured 5.4 against 5.1 us
            e = narrow ? sparse_launch2<8, 512, 60>(a, hs) : sparse_launch2<8, 512, TS_TX>(a, hs);
            k += 2;
            continue;
        }
        switch (variant) {
            case 0: e = sparse_launch<4, 256>(a, hs); break;
            case 2: e = sparse_launch<8, 256>(a, hs); break;
            case 3: e = sparse_launch<16, 512>(a, hs); break;
            default: e = sparse_launch<8, 512>(a, hs); break;
        }
        k += 1;
    }
    if (e == hipSuccess) {
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
        tgv_sparse_output_kernel<<<blocks, 256, 0, hs>>>(d_out, slot[launches & 1], n, zfar);
        e = hipGetLastError();
    }
    if (owned) (void)workspace_free_async(ws, hs);   // after the launches above, in stream order
    if (e != hipSuccess) return hip_status(e, "psw_tgv_sparse_denoise");
    return PSW_OK;
}

}  // extern "C"
