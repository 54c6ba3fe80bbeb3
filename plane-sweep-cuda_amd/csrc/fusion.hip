// fusion.hip -- TV-L1 histogram volumetric fusion for gfx950.
//
// Grid storage is SoA: u[V], v[V], p[3][V] (px, py, pz planes) and hist[V][8]
// uint16 counts, x fastest, then y, then z.  The reference keeps one 28-byte
// AoS fusionvoxel<8> per voxel (src/inc/structs.h:131-189) and reads all of
// it in every kernel; SoA lets each step read only what it uses with
// coalesced 4- and 16-byte accesses.
//
// Histogram votes reproduce FusionUpdateHistogram_kernel (src/fusion.cu:5-43)
// bit for bit: voxel centre with the double division of worldCoords
// (src/inc/fusion.h:360-364), float projection, bilinear depth lookup
// (src/inc/dev_functions.h:33-38) and the bin rule of updateHist
// (src/inc/fusion.h:883-907), all without FMA contraction.
//
// The U step's proxHist (src/inc/fusion.h:851-858) inserts u and p_1..p_8
// into the sorted bin centres with sortedHist::insert (src/inc/structs.h:
// 212-284), which drops values below the running minimum; since the minimum
// is always the bin centre -1, the result equals max(-1, median of the 17
// values).  Here the median is taken in closed form: the 8 values p_i are
// already ordered, so inserting u sorts the 9 data values, then the 9th
// smallest of the union with the 8 (sorted, constant) bin centres is
// selected.
#include "psw_common.h"
#include "rn_math.h"

#include <type_traits>
#include <vector>

namespace psw {

constexpr int FB = PSW_FUSION_BINS;

struct FusionView {
    int w, h, d;
    float a[3], b[3];
    float* u;
    float* v;
    float* p;
    uint16_t* hist;
    size_t V;
};

static inline FusionView view_of(const psw_fusion_grid* g) {
    FusionView f;
    f.w = g->width; f.h = g->height; f.d = g->depth;
    for (int i = 0; i < 3; ++i) { f.a[i] = g->volume[i]; f.b[i] = g->volume[3 + i]; }
    f.u = g->u; f.v = g->v; f.p = g->p; f.hist = g->hist;
    f.V = (size_t)g->width * g->height * g->depth;
    return f;
}

// Voxel centre in world coordinates (worldCoords, fusion.h:360-364).
__device__ inline void voxel_centre(const FusionView& f, int x, int y, int z, float& cx, float& cy, float& cz) {
    const float fx = (float)((x + .5) / (double)f.w);
    const float fy = (float)((y + .5) / (double)f.h);
    const float fz = (float)((z + .5) / (double)f.d);
    const float sx = f.b[0] - f.a[0], sy = f.b[1] - f.a[1], sz = f.b[2] - f.a[2];
    cx = f.a[0] + sx * fx;
    cy = f.a[1] + sy * fy;
    cz = f.a[2] + sz * fz;
}

struct Cam {
    M3f K, R;
    V3f T;
};

// Bin of one depthmap vote for a voxel centre c, or -1 if it projects
// outside the image (fusion.cu:17-41, fusion.h:883-907).
__device__ inline int vote_bin(const Cam& cam, const float* __restrict__ depth, int W, int H, float thr, float cx,
                               float cy, float cz) {
    float c0, c1, c2;
    mv3(cam.R, cx, cy, cz, c0, c1, c2);
    c0 = c0 + cam.T.x; c1 = c1 + cam.T.y; c2 = c2 + cam.T.z;
    float k0, k1, k2;
    mv3(cam.K, c0, c1, c2, k0, k1, k2);
    const float px = k0 / k2, py = k1 / k2;
    if ((px < 0.f) || (px > (float)(W - 1)) || (py < 0.f) || (py > (float)(H - 1))) return -1;
    const float flx = floorf(px), fly = floorf(py);
    const int x0 = (int)fmaxf(flx, 0.f), y0 = (int)fmaxf(fly, 0.f);
    const int x1 = (int)fminf((float)(x0 + 1), (float)(W - 1)), y1 = (int)fminf((float)(y0 + 1), (float)(H - 1));
    const float frx = px - flx, fry = py - fly;
    const float a0 = depth[x0 + (size_t)y0 * W], a1 = depth[x1 + (size_t)y0 * W];
    const float b0 = depth[x0 + (size_t)y1 * W], b1 = depth[x1 + (size_t)y1 * W];
    const float l0 = a0 + fry * (b0 - a0), l1 = a1 + fry * (b1 - a1);
    const float dep = l0 + frx * (l1 - l0);
    const float sd = dep - k2;
    if (sd >= thr) return FB - 1;
    if (sd <= -thr) return 0;
    const float q = fminf(roundf((sd + thr) / (2.f * thr) * (float)(FB - 3) + 1.f), (float)(FB - 2));
    // q is in [1, 6] for thr > 0 (and 6 for NaN, as fminf returns the number);
    // the lower clamp only keeps the index inside the histogram.
    return (int)fmaxf(q, 0.f);
}

__global__ void fusion_hist_kernel(FusionView f, Cam cam, const float* __restrict__ depth, int W, int H, float thr) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= f.w || y >= f.h || z >= f.d) return;
    float cx, cy, cz;
    voxel_centre(f, x, y, z, cx, cy, cz);
    const int bin = vote_bin(cam, depth, W, H, thr, cx, cy, cz);
    if (bin < 0) return;
    const size_t i = ((size_t)z * f.h + y) * f.w + x;
    f.hist[i * FB + bin] = (uint16_t)(f.hist[i * FB + bin] + 1);
}

constexpr int FH_BATCH = 24;  // depthmaps per launch (kernel-argument space)

struct CamBatch {
    Cam cam[FH_BATCH];
    int m;
};

// Votes of up to FH_BATCH depthmaps per voxel, counted in registers and added
// to the stored histogram once.  Counts are integers, so the order of the
// votes does not matter and the result equals m single updates (mod 2^16).
__global__ void fusion_hist_batch_kernel(FusionView f, CamBatch cb, const float* __restrict__ depths, int W, int H,
                                         float thr) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= f.w || y >= f.h || z >= f.d) return;
    float cx, cy, cz;
    voxel_centre(f, x, y, z, cx, cy, cz);
    uint32_t cnt[FB];
#pragma unroll
    for (int k = 0; k < FB; ++k) cnt[k] = 0;
    const size_t plane = (size_t)W * H;
    for (int j = 0; j < cb.m; ++j) {
        const int bin = vote_bin(cb.cam[j], depths + (size_t)j * plane, W, H, thr, cx, cy, cz);
#pragma unroll
        for (int k = 0; k < FB; ++k) cnt[k] += (bin == k) ? 1u : 0u;
    }
    const size_t i = ((size_t)z * f.h + y) * f.w + x;
    uint4* hp = reinterpret_cast<uint4*>(f.hist + i * FB);
    uint4 h = *hp;
    uint32_t w4[4] = {h.x, h.y, h.z, h.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t lo = (w4[k] + cnt[2 * k]) & 0xffffu;
        const uint32_t hi = ((w4[k] >> 16) + cnt[2 * k + 1]) & 0xffffu;
        w4[k] = lo | (hi << 16);
    }
    *hp = make_uint4(w4[0], w4[1], w4[2], w4[3]);
}

// ---- primal-dual ----------------------------------------------------------

// max(-1, median of {centres} U {u, p_1..p_8}), p_i = u + (tau*lambda)*W_i,
// W_i = -sum_{j<i} h_j + sum_{j>=i} h_j (fusion.h:813-838).
__device__ inline float prox_hist_h(float u, uint4 hv, float tl);

__device__ inline float prox_hist(float u, const uint16_t* __restrict__ hp, float tl) {
    return prox_hist_h(u, *reinterpret_cast<const uint4*>(hp), tl);
}

__device__ inline float prox_hist_counts(float u, const int* h, float tl);

__device__ inline float prox_hist_h(float u, uint4 hv, float tl) {
    int h[FB];
    h[0] = hv.x & 0xffff; h[1] = hv.x >> 16; h[2] = hv.y & 0xffff; h[3] = hv.y >> 16;
    h[4] = hv.z & 0xffff; h[5] = hv.z >> 16; h[6] = hv.w & 0xffff; h[7] = hv.w >> 16;
    return prox_hist_counts(u, h, tl);
}

// The same from the byte-packed copy of the counts (all counts <= 255).
__device__ inline float prox_hist_h8(float u, uint2 hv, float tl) {
    int h[FB];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        h[k] = (hv.x >> (8 * k)) & 0xff;
        h[4 + k] = (hv.y >> (8 * k)) & 0xff;
    }
    return prox_hist_counts(u, h, tl);
}

__device__ inline float prox_hist_counts(float u, const int* h, float tl) {
    int total = 0;
#pragma unroll
    for (int k = 0; k < FB; ++k) total += h[k];
    float a[9];
    a[0] = u;
    int below = 0;
#pragma unroll
    for (int i = 1; i <= FB; ++i) {
        below += h[i - 1];
        const int Wi = total - 2 * below;
        a[i] = u + tl * (float)Wi;
    }
    // Sort the 9 values.  W_i falls with i, and u + tl*W is monotone in W in
    // float, so p_8 <= ... <= p_1 already: q = (p_8, ..., p_1) is sorted and
    // only u has to be inserted, s[j] = max(q[j-1], min(q[j], u)).
    {
        const float q0 = a[8], q1 = a[7], q2 = a[6], q3 = a[5], q4 = a[4], q5 = a[3], q6 = a[2], q7 = a[1];
        const float uu = a[0];
        a[0] = fminf(q0, uu);
        a[1] = fmaxf(q0, fminf(q1, uu));
        a[2] = fmaxf(q1, fminf(q2, uu));
        a[3] = fmaxf(q2, fminf(q3, uu));
        a[4] = fmaxf(q3, fminf(q4, uu));
        a[5] = fmaxf(q4, fminf(q5, uu));
        a[6] = fmaxf(q5, fminf(q6, uu));
        a[7] = fmaxf(q6, fminf(q7, uu));
        a[8] = fmaxf(q7, uu);
    }
    // 9th smallest of a (sorted, 9) and the centres c (sorted, 8):
    // min over splits i = 1..9 of max(a[i-1], c[8-i]).  Centres of 8 bins
    // (binParams, fusion.h:1041-1054): -1, -1, -0.6, -0.2, 0.2, 0.6, 1, 1.
    const float c1 = 2.f * 1.f / 5.f - 1.f, c2 = 2.f * 2.f / 5.f - 1.f, c3 = 2.f * 3.f / 5.f - 1.f,
                c4 = 2.f * 4.f / 5.f - 1.f;
    float m = a[8];
    m = fminf(m, fmaxf(a[7], -1.f));
    m = fminf(m, fmaxf(a[6], -1.f));
    m = fminf(m, fmaxf(a[5], c1));
    m = fminf(m, fmaxf(a[4], c2));
    m = fminf(m, fmaxf(a[3], c3));
    m = fminf(m, fmaxf(a[2], c4));
    m = fminf(m, fmaxf(a[1], 1.f));
    m = fminf(m, fmaxf(a[0], 1.f));
    return fmaxf(m, -1.f);
}

struct PdArgs {
    FusionView f;
    const float* pin;   // dual read
    float* pout;        // dual written
    const float* uin;
    float* uout;
    const float* vin;
    float* vout;
    const uint2* hist8;   // byte-packed counts (fusion_pd_zpipe_kernel<true>)
    float tau, lambda, tl, sigma;
};

// FusionUpdateU_kernel (fusion.cu:45-57) with divPBwd (fusion.h:793-802).
__global__ void fusion_u_kernel(FusionView f, float tau, float tl) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= f.w || y >= f.h || z >= f.d) return;
    const size_t V = f.V, i = ((size_t)z * f.h + y) * f.w + x;
    const float* px = f.p;
    const float* py = f.p + V;
    const float* pz = f.p + 2 * V;
    float div = px[i] + py[i] + pz[i];
    if (x > 0) div -= px[i - 1];
    if (y > 0) div -= py[i - f.w];
    if (z > 0) div -= pz[i - (size_t)f.w * f.h];
    const float un = f.u[i];
    const float uu = un - tau * (-div);
    const float unew = prox_hist(uu, f.hist + i * FB, tl);
    f.u[i] = unew;
    f.v[i] = 2.f * unew - un;
}

// FusionUpdateP_kernel (fusion.cu:59-70): p <- project(p + sigma grad+ v).
__global__ void fusion_p_kernel(FusionView f, float sigma) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= f.w || y >= f.h || z >= f.d) return;
    const size_t V = f.V, i = ((size_t)z * f.h + y) * f.w + x;
    const float vv = f.v[i];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    if (x < f.w - 1) gx = f.v[i + 1] - vv;
    if (y < f.h - 1) gy = f.v[i + f.w] - vv;
    if (z < f.d - 1) gz = f.v[i + (size_t)f.w * f.h] - vv;
    float ax = f.p[i] + sigma * gx, ay = f.p[V + i] + sigma * gy, az = f.p[2 * V + i] + sigma * gz;
    const float s2 = ax * ax + ay * ay + az * az;
    if (s2 > 1.f) {   // x / fmaxf(1, sqrt(s)) is x itself for s <= 1
        const float dn = sqrtf(s2);
        ax = ax / dn;
        ay = ay / dn;
        az = az / dn;
    }
    f.p[i] = ax;
    f.p[V + i] = ay;
    f.p[2 * V + i] = az;
}

// Fused iteration: P step of iteration k (from v_k) followed by the U step of
// iteration k+1 (from the new p).  The U step at voxel i needs the new p at
// i-x, i-y, i-z, which this thread recomputes from p_k and v_k (same
// arithmetic, same result as the owner thread).  p and v are ping-ponged
// because neighbouring voxels read their old values; u is read and written
// only at the own voxel.
// p + sigma * g projected onto the unit ball (projectUnitBall, fusion.h:
// 866-870); x / fmaxf(1, sqrt(s)) is x itself for s <= 1, so there is no
// root or division inside the ball.
__device__ __forceinline__ void p_step(float p0, float p1, float p2, float gx, float gy, float gz, float sigma,
                                       float& ox, float& oy, float& oz) {
    ox = p0 + sigma * gx;
    oy = p1 + sigma * gy;
    oz = p2 + sigma * gz;
    project_ball3(ox * ox + oy * oy + oz * oz, ox, oy, oz);
}

__device__ inline void p_new_at(const PdArgs& a, size_t V, int x, int y, int z, size_t i, float& ox, float& oy,
                                float& oz) {
    const FusionView& f = a.f;
    const float vv = a.vin[i];
    float gx = 0.f, gy = 0.f, gz = 0.f;
    if (x < f.w - 1) gx = a.vin[i + 1] - vv;
    if (y < f.h - 1) gy = a.vin[i + f.w] - vv;
    if (z < f.d - 1) gz = a.vin[i + (size_t)f.w * f.h] - vv;
    p_step(a.pin[i], a.pin[V + i], a.pin[2 * V + i], gx, gy, gz, a.sigma, ox, oy, oz);
}

__global__ void fusion_pd_fused_kernel(PdArgs a) {
    const FusionView& f = a.f;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= f.w || y >= f.h || z >= f.d) return;
    const size_t V = f.V, i = ((size_t)z * f.h + y) * f.w + x;
    const size_t sxy = (size_t)f.w * f.h;
    float cx, cy, cz;
    p_new_at(a, V, x, y, z, i, cx, cy, cz);
    a.pout[i] = cx;
    a.pout[V + i] = cy;
    a.pout[2 * V + i] = cz;
    float div = cx + cy + cz;
    float tx, ty, tz;
    if (x > 0) { p_new_at(a, V, x - 1, y, z, i - 1, tx, ty, tz); div -= tx; }
    if (y > 0) { p_new_at(a, V, x, y - 1, z, i - f.w, tx, ty, tz); div -= ty; }
    if (z > 0) { p_new_at(a, V, x, y, z - 1, i - sxy, tx, ty, tz); div -= tz; }
    const float un = a.uin[i];
    const float uu = un - a.tau * (-div);
    const float unew = prox_hist(uu, f.hist + i * FB, a.tl);
    a.uout[i] = unew;
    a.vout[i] = 2.f * unew - un;
}

// The same fused iteration with each thread marching FZ_CHUNK consecutive z
// planes of one (x, y) column: the new p of plane z-1 stays in a register, so
// only the first plane of a chunk recomputes its z-1 neighbour.  Every value
// goes through the same arithmetic in the same order as in
// fusion_pd_fused_kernel, so both kernels give identical results.
constexpr int FZ_CHUNK = 16;

// Blocks of 64x4 columns, one-dimensional grid: block b takes tile
// xcd_tile(b) in x, y, z-chunk order, so that each XCD works through a
// contiguous range of tiles and the y-halo rows a block reads were mostly
// just fetched into the same L2 by its neighbour.
constexpr int FZ_BX = 64, FZ_BY = 4;

__global__ __launch_bounds__(FZ_BX * FZ_BY) void fusion_pd_zmarch_kernel(PdArgs a) {
    const FusionView& f = a.f;
    const int tx_n = (f.w + FZ_BX - 1) / FZ_BX, ty_n = (f.h + FZ_BY - 1) / FZ_BY;
    const int ntiles = tx_n * ty_n * ((f.d + FZ_CHUNK - 1) / FZ_CHUNK);
    const int t = xcd_tile(blockIdx.x, ntiles);
    const int bx = t % tx_n, by = (t / tx_n) % ty_n, bz = t / (tx_n * ty_n);
    const int x = bx * FZ_BX + (int)(threadIdx.x % FZ_BX);
    const int y = by * FZ_BY + (int)(threadIdx.x / FZ_BX);
    const int z0 = bz * FZ_CHUNK;
    if (x >= f.w || y >= f.h || z0 >= f.d) return;
    const int z1 = min(z0 + FZ_CHUNK, f.d);
    const size_t V = f.V, sxy = (size_t)f.w * f.h;
    size_t i = ((size_t)z0 * f.h + y) * f.w + x;
    float pz_below = 0.f;
    if (z0 > 0) {
        float tx, ty;
        p_new_at(a, V, x, y, z0 - 1, i - sxy, tx, ty, pz_below);
    }
    for (int z = z0; z < z1; ++z, i += sxy) {
        float cx, cy, cz;
        p_new_at(a, V, x, y, z, i, cx, cy, cz);
        a.pout[i] = cx;
        a.pout[V + i] = cy;
        a.pout[2 * V + i] = cz;
        float div = cx + cy + cz;
        float tx, ty, tz;
        if (x > 0) { p_new_at(a, V, x - 1, y, z, i - 1, tx, ty, tz); div -= tx; }
        if (y > 0) { p_new_at(a, V, x, y - 1, z, i - f.w, tx, ty, tz); div -= ty; }
        if (z > 0) div -= pz_below;
        pz_below = cz;
        const float un = a.uin[i];
        const float uu = un - a.tau * (-div);
        const float unew = prox_hist(uu, f.hist + i * FB, a.tl);
        a.uout[i] = unew;
        a.vout[i] = 2.f * unew - un;
    }
}

// The z-march with its loads software-pipelined: everything plane z+1 needs
// (v at the voxel and at x+-1, y+-1 neighbours, p of the voxel and of its x-1
// and y-1 neighbours, the histogram, u) is loaded before plane z's results
// are stored, and the z+1 values of v that plane z's gradients need come
// from that bundle.  (The compiler cannot hoist those loads itself: u is read
// and written in place.)  Same arithmetic and order as fusion_pd_fused_kernel.
// H8: the histogram comes from the byte-packed copy (8 B per voxel instead
// of 16); same counts, same results.
template <bool H8>
struct ZBundle {
    float v0, vxp, vyp, vxm, vxmyp, vym, vymxp;
    float p0, p1, p2, pxm0, pxm1, pxm2, pym0, pym1, pym2;
    typename std::conditional<H8, uint2, uint4>::type h;
    float u;
};

template <bool H8>
__device__ __forceinline__ void zbundle_load(const PdArgs& a, size_t V, int x, int y, size_t i, ZBundle<H8>& b) {
    const FusionView& f = a.f;
    const bool xp = x < f.w - 1, yp = y < f.h - 1, xm = x > 0, ym = y > 0;
    b.v0 = a.vin[i];
    b.vxp = xp ? a.vin[i + 1] : 0.f;
    b.vyp = yp ? a.vin[i + f.w] : 0.f;
    b.vxm = xm ? a.vin[i - 1] : 0.f;
    b.vxmyp = xm && yp ? a.vin[i - 1 + f.w] : 0.f;
    b.vym = ym ? a.vin[i - f.w] : 0.f;
    b.vymxp = ym && xp ? a.vin[i - f.w + 1] : 0.f;
    b.p0 = a.pin[i];
    b.p1 = a.pin[V + i];
    b.p2 = a.pin[2 * V + i];
    b.pxm0 = xm ? a.pin[i - 1] : 0.f;
    b.pxm1 = xm ? a.pin[V + i - 1] : 0.f;
    b.pxm2 = xm ? a.pin[2 * V + i - 1] : 0.f;
    b.pym0 = ym ? a.pin[i - f.w] : 0.f;
    b.pym1 = ym ? a.pin[V + i - f.w] : 0.f;
    b.pym2 = ym ? a.pin[2 * V + i - f.w] : 0.f;
    if constexpr (H8) b.h = a.hist8[i];
    else b.h = *reinterpret_cast<const uint4*>(f.hist + i * FB);
    b.u = a.uin[i];
}

// Blocks of the pipelined z-march: 64 x FZP_BY columns (the y-halo rows a
// block reads, v at y-1 and y+1 and p at y-1, are a smaller share of its
// loads than with the 64 x 4 blocks of the unpipelined kernel).
constexpr int FZP_BY = 8, FZP_CHUNK = 16;

template <bool H8>
__global__ __launch_bounds__(FZ_BX * FZP_BY) void fusion_pd_zpipe_kernel(PdArgs a) {
    const FusionView& f = a.f;
    constexpr int FZ_BY = FZP_BY, FZ_CHUNK = FZP_CHUNK;
    const int tx_n = (f.w + FZ_BX - 1) / FZ_BX, ty_n = (f.h + FZ_BY - 1) / FZ_BY;
    const int ntiles = tx_n * ty_n * ((f.d + FZ_CHUNK - 1) / FZ_CHUNK);
    const int t = xcd_tile(blockIdx.x, ntiles);
    const int bx = t % tx_n, by = (t / tx_n) % ty_n, bz = t / (tx_n * ty_n);
    const int x = bx * FZ_BX + (int)(threadIdx.x % FZ_BX);
    const int y = by * FZ_BY + (int)(threadIdx.x / FZ_BX);
    const int z0 = bz * FZ_CHUNK;
    if (x >= f.w || y >= f.h || z0 >= f.d) return;
    const int z1 = min(z0 + FZ_CHUNK, f.d);
    const size_t V = f.V, sxy = (size_t)f.w * f.h;
    const bool xp = x < f.w - 1, yp = y < f.h - 1, xm = x > 0, ym = y > 0;
    size_t i = ((size_t)z0 * f.h + y) * f.w + x;
    float pz_below = 0.f;
    if (z0 > 0) {
        float tx, ty;
        p_new_at(a, V, x, y, z0 - 1, i - sxy, tx, ty, pz_below);
    }
    ZBundle<H8> cur, nxt;
    zbundle_load(a, V, x, y, i, cur);
    for (int z = z0; z < z1; ++z, i += sxy) {
        const bool zp = z < f.d - 1;
        if (zp) zbundle_load(a, V, x, y, i + sxy, nxt);   // the next plane (also when it is the next chunk's)
        float cx, cy, cz;
        p_step(cur.p0, cur.p1, cur.p2, xp ? cur.vxp - cur.v0 : 0.f, yp ? cur.vyp - cur.v0 : 0.f,
               zp ? nxt.v0 - cur.v0 : 0.f, a.sigma, cx, cy, cz);
        a.pout[i] = cx;
        a.pout[V + i] = cy;
        a.pout[2 * V + i] = cz;
        float div = cx + cy + cz;
        float tx, ty, tz;
        if (xm) {
            p_step(cur.pxm0, cur.pxm1, cur.pxm2, cur.v0 - cur.vxm, yp ? cur.vxmyp - cur.vxm : 0.f,
                   zp ? nxt.vxm - cur.vxm : 0.f, a.sigma, tx, ty, tz);
            div -= tx;
        }
        if (ym) {
            p_step(cur.pym0, cur.pym1, cur.pym2, xp ? cur.vymxp - cur.vym : 0.f, cur.v0 - cur.vym,
                   zp ? nxt.vym - cur.vym : 0.f, a.sigma, tx, ty, tz);
            div -= ty;
        }
        if (z > 0) div -= pz_below;
        pz_below = cz;
        const float un = cur.u;
        const float uu = un - a.tau * (-div);
        float unew;
        if constexpr (H8) unew = prox_hist_h8(uu, cur.h, a.tl);
        else unew = prox_hist_h(uu, cur.h, a.tl);
        a.uout[i] = unew;
        a.vout[i] = 2.f * unew - un;
        if (zp) cur = nxt;
    }
}

// Two fused steps per launch: (P_k, U_{k+1}, P_{k+1}, U_{k+2}).  A block of
// 64 x 8 columns marches FT_CHUNK planes; at step s it computes, in LDS,
//   A  p of k+1 at plane s     on the tile + 2 columns left/up, 1 right/down
//   B  u, v of k+1 at plane s  on the tile + a 1-column ring
//   C  p of k+2 at plane s-1   on the tile + 1 column left/up
//   D  u, v of k+2 at plane s-1 on the tile, written to HBM with p of k+2,
// so a voxel's p, u, v, counts are read once per two steps and its new
// values written once.  Planes alternate between two LDS buffers by parity;
// a chunk starts two planes early (stage A) to have the z-1 duals.  The
// per-position arithmetic and its order are those of fusion_pd_fused_kernel
// (p_step, the same divergence, prox), so the grids are identical.  u is
// read on the halo, so unlike the one-step kernels it needs uin != uout.
// LDS index of a column: (x - x0 + 2, y - y0 + 2); positions outside the
// grid read clamped coordinates and never feed positions inside it.
// 768 threads: every halo stage (at most 67 x 11 = 737 columns) takes one
// pass; stage D uses the first 512.  With at most 80 VGPRs (launch bounds)
// two blocks fit per CU.
constexpr int FT_X = 64, FT_Y = 8, FT_NT = 768, FT_CHUNK = 32;
constexpr int FT_LX = FT_X + 4, FT_LY = FT_Y + 4;

template <bool H8>
__global__ __launch_bounds__(FT_NT, 2) void fusion_pd2_kernel(PdArgs a) {
    using HT = typename std::conditional<H8, uint2, uint4>::type;
    __shared__ float sV[2][FT_LY][FT_LX];       // v of k, planes by parity
    __shared__ float sP1[2][3][FT_LY][FT_LX];   // p of k+1
    __shared__ float sU1[2][FT_LY][FT_LX];      // u of k+1
    __shared__ float sV1[2][FT_LY][FT_LX];      // v of k+1
    __shared__ float sP2[2][3][FT_LY][FT_LX];   // p of k+2
    const FusionView& f = a.f;
    const int W = f.w, H = f.h, D = f.d;
    const size_t V = f.V, sxy = (size_t)W * H;
    const int tx_n = (W + FT_X - 1) / FT_X, ty_n = (H + FT_Y - 1) / FT_Y;
    const int ntiles = tx_n * ty_n * ((D + FT_CHUNK - 1) / FT_CHUNK);
    const int t = xcd_tile(blockIdx.x, ntiles);
    const int bx = t % tx_n, by = (t / tx_n) % ty_n, bz = t / (tx_n * ty_n);
    const int x0 = bx * FT_X, y0 = by * FT_Y;
    const int z0 = bz * FT_CHUNK, z1 = min(z0 + FT_CHUNK, D);
    const int tid = threadIdx.x;
    auto gxy = [&](int ix, int iy) {   // clamped in-plane index of LDS column (ix, iy)
        return (size_t)min(max(y0 - 2 + iy, 0), H - 1) * W + min(max(x0 - 2 + ix, 0), W - 1);
    };
    auto hist_at = [&](size_t i) -> HT {
        if constexpr (H8) return a.hist8[i];
        else return *reinterpret_cast<const uint4*>(f.hist + i * FB);
    };
    auto prox = [&](float uu, const HT& h) -> float {
        if constexpr (H8) return prox_hist_h8(uu, h, a.tl);
        else return prox_hist_h(uu, h, a.tl);
    };
    const int sa = max(z0 - 2, 0), sb = max(z0 - 1, 0);   // first planes of stages A and B (and C)
    const int slast = min(z1, D - 1);                       // last plane of stages A and B
    // Each thread's positions are the same on every plane: j-th position of
    // a region = tid + j * FT_NT.  Their global inputs for the next plane are
    // loaded into registers right after the current ones are used.
    constexpr int NV = (FT_LX * FT_LY + FT_NT - 1) / FT_NT;               // v of k (plane s+1)
    constexpr int NA = ((FT_LX - 1) * (FT_LY - 1) + FT_NT - 1) / FT_NT;   // stage A: p of k
    constexpr int NB = ((FT_LX - 2) * (FT_LY - 2) + FT_NT - 1) / FT_NT;   // stage B: u of k, counts
    float rv[NV], rp[NA][3], ru[NB];
    HT rh[NB], rhd{};
    const int tcx = tid % FT_X, tcy = tid / FT_X;   // stage D: threads below FT_X * FT_Y
    const bool dlive = tid < FT_X * FT_Y && x0 + tcx < W && y0 + tcy < H;
    const size_t dxy = dlive ? (size_t)(y0 + tcy) * W + (x0 + tcx) : 0;
    auto load_v = [&](int plane) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int i = tid + j * FT_NT;
            if (i < FT_LX * FT_LY) rv[j] = a.vin[(size_t)plane * sxy + gxy(i % FT_LX, i / FT_LX)];
        }
    };
    auto load_a = [&](int plane) {
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            const int i = tid + j * FT_NT;
            if (i < (FT_LX - 1) * (FT_LY - 1)) {
                const size_t g = (size_t)plane * sxy + gxy(i % (FT_LX - 1), i / (FT_LX - 1));
                rp[j][0] = a.pin[g];
                rp[j][1] = a.pin[V + g];
                rp[j][2] = a.pin[2 * V + g];
            }
        }
    };
    auto load_b = [&](int plane) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int i = tid + j * FT_NT;
            if (i < (FT_LX - 2) * (FT_LY - 2)) {
                const size_t g = (size_t)plane * sxy + gxy(i % (FT_LX - 2) + 1, i / (FT_LX - 2) + 1);
                ru[j] = a.uin[g];
                rh[j] = hist_at(g);
            }
        }
    };
    for (int i = tid; i < FT_LX * FT_LY; i += FT_NT) {
        const int iy = i / FT_LX, ix = i - iy * FT_LX;
        sV[sa & 1][iy][ix] = a.vin[(size_t)sa * sxy + gxy(ix, iy)];
    }
    if (sa + 1 < D) load_v(sa + 1);
    load_a(sa);
    load_b(sa);
    for (int s = sa; s <= slast + 1; ++s) {
        const bool ab = s <= slast;
        const int cs = s & 1, ps = (s + 1) & 1;   // buffers of planes s and s-1 (and s+1)
        if (ab && s + 1 < D) {
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                const int i = tid + j * FT_NT;
                if (i < FT_LX * FT_LY) sV[ps][i / FT_LX][i % FT_LX] = rv[j];
            }
        }
        __syncthreads();
        // A: p of k+1 at plane s on [0, LX-1) x [0, LY-1)
        if (ab) {
#pragma unroll
            for (int j = 0; j < NA; ++j) {
                const int i = tid + j * FT_NT;
                if (i < (FT_LX - 1) * (FT_LY - 1)) {
                    const int iy = i / (FT_LX - 1), ix = i - iy * (FT_LX - 1);
                    const int gx = x0 - 2 + ix, gy = y0 - 2 + iy;
                    const float v0 = sV[cs][iy][ix];
                    const float gxv = gx < W - 1 ? sV[cs][iy][ix + 1] - v0 : 0.f;
                    const float gyv = gy < H - 1 ? sV[cs][iy + 1][ix] - v0 : 0.f;
                    const float gzv = s < D - 1 ? sV[ps][iy][ix] - v0 : 0.f;
                    float ox, oy, oz;
                    p_step(rp[j][0], rp[j][1], rp[j][2], gxv, gyv, gzv, a.sigma, ox, oy, oz);
                    sP1[cs][0][iy][ix] = ox;
                    sP1[cs][1][iy][ix] = oy;
                    sP1[cs][2][iy][ix] = oz;
                }
            }
            // the next plane's v of k and p of k
            if (s + 2 < D && s + 1 <= slast) load_v(s + 2);
            if (s + 1 <= slast) load_a(s + 1);
        }
        __syncthreads();
        // B: u, v of k+1 at plane s on [1, LX-1) x [1, LY-1)
        if (ab && s >= sb) {
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int i = tid + j * FT_NT;
                if (i < (FT_LX - 2) * (FT_LY - 2)) {
                    const int iy = i / (FT_LX - 2) + 1, ix = i - (i / (FT_LX - 2)) * (FT_LX - 2) + 1;
                    const int gx = x0 - 2 + ix, gy = y0 - 2 + iy;
                    float div = sP1[cs][0][iy][ix] + sP1[cs][1][iy][ix] + sP1[cs][2][iy][ix];
                    if (gx > 0) div -= sP1[cs][0][iy][ix - 1];
                    if (gy > 0) div -= sP1[cs][1][iy - 1][ix];
                    if (s > 0) div -= sP1[ps][2][iy][ix];
                    const float un = ru[j];
                    const float uu = un - a.tau * (-div);
                    const float unew = prox(uu, rh[j]);
                    sU1[cs][iy][ix] = unew;
                    sV1[cs][iy][ix] = 2.f * unew - un;
                }
            }
        }
        if (ab && s + 1 <= slast) load_b(s + 1);
        __syncthreads();
        // C: p of k+2 at plane c = s-1 on [1, LX-2) x [1, LY-2)
        const int c = s - 1;
        if (c >= sb && c < z1) {
            for (int i = tid; i < (FT_LX - 3) * (FT_LY - 3); i += FT_NT) {
                const int iy = i / (FT_LX - 3) + 1, ix = i - (i / (FT_LX - 3)) * (FT_LX - 3) + 1;
                const int gx = x0 - 2 + ix, gy = y0 - 2 + iy;
                const float v0 = sV1[ps][iy][ix];
                const float gxv = gx < W - 1 ? sV1[ps][iy][ix + 1] - v0 : 0.f;
                const float gyv = gy < H - 1 ? sV1[ps][iy + 1][ix] - v0 : 0.f;
                const float gzv = c < D - 1 ? sV1[cs][iy][ix] - v0 : 0.f;
                float ox, oy, oz;
                p_step(sP1[ps][0][iy][ix], sP1[ps][1][iy][ix], sP1[ps][2][iy][ix], gxv, gyv, gzv, a.sigma, ox, oy,
                       oz);
                sP2[ps][0][iy][ix] = ox;
                sP2[ps][1][iy][ix] = oy;
                sP2[ps][2][iy][ix] = oz;
            }
        }
        __syncthreads();
        // D: u, v of k+2 at plane c on the tile; p, u, v of k+2 to HBM
        if (c >= z0 && c < z1 && dlive) {
            const int x = x0 + tcx, y = y0 + tcy;
            const int ix = tcx + 2, iy = tcy + 2;
            const size_t g = (size_t)c * sxy + dxy;
            const float px = sP2[ps][0][iy][ix], py = sP2[ps][1][iy][ix], pz = sP2[ps][2][iy][ix];
            float div = px + py + pz;
            if (x > 0) div -= sP2[ps][0][iy][ix - 1];
            if (y > 0) div -= sP2[ps][1][iy - 1][ix];
            if (c > 0) div -= sP2[cs][2][iy][ix];
            const float un = sU1[ps][iy][ix];
            const float uu = un - a.tau * (-div);
            const float unew = prox(uu, rhd);
            a.pout[g] = px;
            a.pout[V + g] = py;
            a.pout[2 * V + g] = pz;
            a.uout[g] = unew;
            a.vout[g] = 2.f * unew - un;
        }
        // the counts of the next step's stage D (plane s)
        if (dlive && s >= z0 && s < z1) rhd = hist_at((size_t)s * sxy + dxy);
    }
}

// z-marching with the x-1 and y-1 neighbours' new duals taken from the
// threads that own them (through LDS) instead of recomputed: a 64x8 block
// recomputes only its left column and top row halo.  Same arithmetic and
// order as fusion_pd_fused_kernel, so the results are identical.
constexpr int FZL_TX = 64, FZL_TY = 8;

__global__ __launch_bounds__(FZL_TX * FZL_TY) void fusion_pd_zlds_kernel(PdArgs a) {
    __shared__ float sPx[2][FZL_TY][FZL_TX];
    __shared__ float sPy[2][FZL_TY][FZL_TX];
    const FusionView& f = a.f;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = blockIdx.x * FZL_TX + tx;
    const int y = blockIdx.y * FZL_TY + ty;
    const int z0 = blockIdx.z * FZ_CHUNK;
    const int z1 = min(z0 + FZ_CHUNK, f.d);   // uniform over the block
    const bool live = x < f.w && y < f.h;
    const size_t V = f.V, sxy = (size_t)f.w * f.h;
    size_t i = live ? ((size_t)z0 * f.h + y) * f.w + x : 0;
    float pz_below = 0.f;
    if (live && z0 > 0) {
        float tx_, ty_;
        p_new_at(a, V, x, y, z0 - 1, i - sxy, tx_, ty_, pz_below);
    }
    for (int z = z0, buf = 0; z < z1; ++z, i += sxy, buf ^= 1) {
        float cx = 0.f, cy = 0.f, cz = 0.f;
        if (live) p_new_at(a, V, x, y, z, i, cx, cy, cz);
        sPx[buf][ty][tx] = cx;
        sPy[buf][ty][tx] = cy;
        __syncthreads();   // one barrier per plane: the other buffer is written next plane
        if (live) {
            a.pout[i] = cx;
            a.pout[V + i] = cy;
            a.pout[2 * V + i] = cz;
            float div = cx + cy + cz;
            if (x > 0) {
                float nx = 0.f;
                if (tx > 0) {
                    nx = sPx[buf][ty][tx - 1];
                } else {
                    float t1, t2;
                    p_new_at(a, V, x - 1, y, z, i - 1, nx, t1, t2);
                }
                div -= nx;
            }
            if (y > 0) {
                float ny = 0.f;
                if (ty > 0) {
                    ny = sPy[buf][ty - 1][tx];
                } else {
                    float t1, t2;
                    p_new_at(a, V, x, y - 1, z, i - f.w, t1, ny, t2);
                }
                div -= ny;
            }
            if (z > 0) div -= pz_below;
            pz_below = cz;
            const float un = a.uin[i];
            const float uu = un - a.tau * (-div);
            const float unew = prox_hist(uu, f.hist + i * FB, a.tl);
            a.uout[i] = unew;
            a.vout[i] = 2.f * unew - un;
        }
    }
}

// FP64 TV-L1 histogram energy: per-block partials of
// sum_i |grad u_i| + lambda sum_k h_k(i) |u_i - c_k| over a grid-stride loop
// (fixed assignment of voxels to blocks, so the result is reproducible).
constexpr int FE_BLOCKS = 1024;

__global__ __launch_bounds__(256) void fusion_energy_kernel(FusionView f, double lambda, double* __restrict__ part) {
    __shared__ double red[4];
    const size_t sxy = (size_t)f.w * f.h;
    const double c[FB] = {-1.0, -1.0, (double)(2.f * 1.f / 5.f - 1.f), (double)(2.f * 2.f / 5.f - 1.f),
                          (double)(2.f * 3.f / 5.f - 1.f), (double)(2.f * 4.f / 5.f - 1.f), 1.0, 1.0};
    double e = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.V; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % f.w);
        const size_t t = i / f.w;
        const int y = (int)(t % f.h), z = (int)(t / f.h);
        const double u = f.u[i];
        const double gx = x < f.w - 1 ? (double)f.u[i + 1] - u : 0.0;
        const double gy = y < f.h - 1 ? (double)f.u[i + f.w] - u : 0.0;
        const double gz = z < f.d - 1 ? (double)f.u[i + sxy] - u : 0.0;
        const uint4 hv = *reinterpret_cast<const uint4*>(f.hist + i * FB);
        const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w};
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < FB; ++k) d += (double)((hw[k >> 1] >> (16 * (k & 1))) & 0xffffu) * fabs(u - c[k]);
        e += sqrt(gx * gx + gy * gy + gz * gz) + lambda * d;
    }
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Fused-surface points (src/pclviewer.cpp:1058-1074): voxels with
// -1 < u < 0, as world coordinates of the voxel centres, in the reference's
// z, y, x loop order.  Two passes over fixed voxel ranges: counts per
// workgroup, then (after an exclusive scan of the counts) an ordered write
// with a wave ballot prefix inside each workgroup.
constexpr int SP_BLOCK = 256, SP_PER_THREAD = 16;

__device__ inline bool on_surface(float u) { return u < 0.f && u > -1.f; }

__global__ __launch_bounds__(SP_BLOCK) void surface_points_count_kernel(const float* __restrict__ u, size_t n,
                                                                        unsigned* __restrict__ counts) {
    __shared__ unsigned wsum[SP_BLOCK / 64];
    const size_t base = (size_t)blockIdx.x * SP_BLOCK * SP_PER_THREAD;
    unsigned c = 0;
    for (int k = 0; k < SP_PER_THREAD; ++k) {
        const size_t i = base + (size_t)k * SP_BLOCK + threadIdx.x;
        if (i < n && on_surface(u[i])) ++c;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(SP_BLOCK) void surface_points_write_kernel(FusionView f, const unsigned* __restrict__ start,
                                                                        float* __restrict__ xyz, size_t cap) {
    __shared__ unsigned wcount[SP_BLOCK / 64];
    const size_t base = (size_t)blockIdx.x * SP_BLOCK * SP_PER_THREAD;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned running = start[blockIdx.x];
    for (int k = 0; k < SP_PER_THREAD; ++k) {
        const size_t i = base + (size_t)k * SP_BLOCK + threadIdx.x;
        const bool hit = i < f.V && on_surface(f.u[i]);
        const unsigned long long mask = __ballot(hit);
        if (lane == 0) wcount[wave] = (unsigned)__popcll(mask);
        __syncthreads();
        unsigned before = running;
        for (int w = 0; w < wave; ++w) before += wcount[w];
        before += (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
        if (hit && before < cap) {
            const int x = (int)(i % f.w);
            const size_t t = i / f.w;
            const int y = (int)(t % f.h), z = (int)(t / f.h);
            float cx, cy, cz;
            voxel_centre(f, x, y, z, cx, cy, cz);
            xyz[3 * (size_t)before] = cx;
            xyz[3 * (size_t)before + 1] = cy;
            xyz[3 * (size_t)before + 2] = cz;
        }
        running += wcount[0] + wcount[1] + wcount[2] + wcount[3];
        __syncthreads();
    }
}

__global__ void surface_count_kernel(const float* __restrict__ u, size_t n, unsigned long long* out) {
    __shared__ unsigned int part[256 / 64];
    unsigned int c = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = u[i];
        c += (v < 0.f && v > -1.f) ? 1u : 0u;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < 4; ++k) t += part[k];
        atomicAdd(out, t);
    }
}

static dim3 vox_grid(const FusionView& f, dim3 blk) {
    return dim3((unsigned)((f.w + blk.x - 1) / blk.x), (unsigned)((f.h + blk.y - 1) / blk.y), (unsigned)f.d);
}

static int check_grid(const psw_fusion_grid* g) {
    if (!g || !g->u || !g->v || !g->p || !g->hist) { set_error("fusion grid not allocated"); return PSW_ERR_ARG; }
    if (g->width < 1 || g->height < 1 || g->depth < 1) { set_error("empty fusion grid"); return PSW_ERR_ARG; }
    if (g->depth > 65535) { set_error("grid depth above 65535"); return PSW_ERR_ARG; }
    return PSW_OK;
}

static Cam cam_of(const float* K, const float* R, const float* T) {
    Cam c;
    c.K = m3_from(K);
    c.R = m3_from(R);
    c.T = v3_from(T);
    return c;
}

}  // namespace psw

using namespace psw;

extern "C" {

int psw_fusion_alloc(psw_fusion_grid* g, int width, int height, int depth, const float* volume) {
    PSW_CHECK_ARG(g && volume, "NULL pointer");
    PSW_CHECK_ARG(width >= 1 && height >= 1 && depth >= 1 && depth <= 65535, "bad grid size");
    std::memset(g, 0, sizeof(*g));
    g->width = width; g->height = height; g->depth = depth;
    for (int i = 0; i < 6; ++i) g->volume[i] = volume[i];
    const size_t V = (size_t)width * height * depth;
    hipError_t e = hipMalloc((void**)&g->u, V * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&g->v, V * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&g->p, 3 * V * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&g->hist, V * FB * sizeof(uint16_t));
    if (e != hipSuccess) {
        int s = hip_status(e, "psw_fusion_alloc", PSW_ERR_ALLOC);
        psw_fusion_free(g);
        return s;
    }
    int s = psw_fusion_clear(g, nullptr);
    if (s == PSW_OK) PSW_TRY_HIP(hipDeviceSynchronize(), PSW_ERR_DEVICE);
    return s;
}

int psw_fusion_free(psw_fusion_grid* g) {
    if (!g) return PSW_OK;
    if (g->u) (void)hipFree(g->u);
    if (g->v) (void)hipFree(g->v);
    if (g->p) (void)hipFree(g->p);
    if (g->hist) (void)hipFree(g->hist);
    g->u = g->v = g->p = nullptr;
    g->hist = nullptr;
    return PSW_OK;
}

int psw_fusion_clear(const psw_fusion_grid* g, psw_stream_t s) {
    int st = check_grid(g);
    if (st != PSW_OK) return st;
    const size_t V = (size_t)g->width * g->height * g->depth;
    if ((st = psw_fill_f32(g->u, 0.f, V, s)) != PSW_OK) return st;
    if ((st = psw_fill_f32(g->v, 0.f, V, s)) != PSW_OK) return st;
    if ((st = psw_fill_f32(g->p, 0.f, 3 * V, s)) != PSW_OK) return st;
    return psw_fill_u32(reinterpret_cast<uint32_t*>(g->hist), 0u, V * FB / 2, s);
}

int psw_fusion_update_histogram(const psw_fusion_grid* g, const float* d_depthmap, int width, int height,
                                const float* K, const float* R, const float* T, float threshold, psw_stream_t s) {
    int st = check_grid(g);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(!(threshold < 0.f), "threshold must not be negative");
    PSW_CHECK_ARG(d_depthmap && K && R && T, "NULL pointer");
    PSW_CHECK_ARG(width >= 1 && height >= 1, "bad depthmap size");
    const FusionView f = view_of(g);
    const dim3 blk(64, 4, 1);
    fusion_hist_kernel<<<vox_grid(f, blk), blk, 0, S(s)>>>(f, cam_of(K, R, T), d_depthmap, width, height, threshold);
    return launch_status("fusion_hist_kernel");
}

int psw_fusion_update_histogram_batch(const psw_fusion_grid* g, const float* d_depthmaps, int m, int width, int height,
                                      const float* K, const float* R, const float* T, float threshold,
                                      psw_stream_t s) {
    int st = check_grid(g);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(!(threshold < 0.f), "threshold must not be negative");
    PSW_CHECK_ARG(d_depthmaps && K && R && T, "NULL pointer");
    PSW_CHECK_ARG(m >= 0 && width >= 1 && height >= 1, "bad depthmap batch");
    const FusionView f = view_of(g);
    const dim3 blk(64, 4, 1);
    for (int j0 = 0; j0 < m; j0 += FH_BATCH) {
        CamBatch cb;
        cb.m = std::min(FH_BATCH, m - j0);
        for (int j = 0; j < cb.m; ++j) cb.cam[j] = cam_of(K + 9 * (j0 + j), R + 9 * (j0 + j), T + 3 * (j0 + j));
        fusion_hist_batch_kernel<<<vox_grid(f, blk), blk, 0, S(s)>>>(f, cb, d_depthmaps + (size_t)j0 * width * height,
                                                                     width, height, threshold);
        st = launch_status("fusion_hist_batch_kernel");
        if (st != PSW_OK) return st;
    }
    return PSW_OK;
}

int psw_fusion_update_u(const psw_fusion_grid* g, double tau, double lambda, psw_stream_t s) {
    int st = check_grid(g);
    if (st != PSW_OK) return st;
    const FusionView f = view_of(g);
    const float tf = (float)tau, lf = (float)lambda;
    const dim3 blk(64, 4, 1);
    fusion_u_kernel<<<vox_grid(f, blk), blk, 0, S(s)>>>(f, tf, tf * lf);
    return launch_status("fusion_u_kernel");
}

int psw_fusion_update_p(const psw_fusion_grid* g, double sigma, psw_stream_t s) {
    int st = check_grid(g);
    if (st != PSW_OK) return st;
    const FusionView f = view_of(g);
    const dim3 blk(64, 4, 1);
    fusion_p_kernel<<<vox_grid(f, blk), blk, 0, S(s)>>>(f, (float)sigma);
    return launch_status("fusion_p_kernel");
}

int psw_fusion_update_iteration(const psw_fusion_grid* g, const float* d_depthmap, int width, int height,
                                const float* K, const float* R, const float* T, float threshold, double tau,
                                double lambda, double sigma, psw_stream_t s) {
    int st = psw_fusion_update_histogram(g, d_depthmap, width, height, K, R, T, threshold, s);
    if (st != PSW_OK) return st;
    st = psw_fusion_update_u(g, tau, lambda, s);
    if (st != PSW_OK) return st;
    return psw_fusion_update_p(g, sigma, s);
}

// Byte-packed copy of the histogram for the PD iterations, and the largest
// count (atomicMax per block): the copy is used only if that is <= 255.
__global__ __launch_bounds__(256) void fusion_hist_pack_kernel(const uint16_t* __restrict__ hist,
                                                               uint2* __restrict__ out, size_t V,
                                                               unsigned* __restrict__ maxcount) {
    __shared__ unsigned sm[256 / 64];
    unsigned m = 0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < V; i += (size_t)gridDim.x * 256) {
        const uint4 h = *reinterpret_cast<const uint4*>(hist + i * FB);
        const uint32_t w[4] = {h.x, h.y, h.z, h.w};
        uint32_t lo = 0, hi = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t a = w[k] & 0xffffu, b = w[k] >> 16;
            m = max(m, max(a, b));
            const uint32_t two = (a & 0xffu) | ((b & 0xffu) << 8);
            if (k < 2) lo |= two << (16 * k);
            else hi |= two << (16 * (k - 2));
        }
        out[i] = make_uint2(lo, hi);
    }
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned b = sm[0];
        for (int k = 1; k < 256 / 64; ++k) b = max(b, sm[k]);
        atomicMax(maxcount, b);
    }
}

__global__ void fusion_zero_u32_kernel(unsigned* p) { *p = 0u; }

int psw_fusion_solve(const psw_fusion_grid* g, int niters, double tau, double lambda, double sigma, psw_stream_t s) {
    int st = check_grid(g);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(niters >= 0, "niters must be >= 0");
    if (niters == 0) return PSW_OK;
    const FusionView f = view_of(g);
    const size_t V = f.V;
    // Iteration 1's U step on its own, then niters-1 fused (P_k, U_{k+1})
    // launches, then the last P step: the same sequence U,P,U,P,...,U,P.
    st = psw_fusion_update_u(g, tau, lambda, s);
    if (st != PSW_OK) return st;
    if (niters > 1) {
        // PSW_FUSION_PD selects the kernel (identical results): 0 one voxel per
        // thread, 1 z-marching, 2 z-marching with LDS-shared neighbour duals,
        // 3 z-marching with software-pipelined loads, 4 two steps per launch
        // (the default; an odd last step runs on mode 3's kernel).
        const char* env = getenv("PSW_FUSION_PD");
        int mode = env ? atoi(env) : 4;
        if (mode < 0 || mode > 4) mode = 4;
        // Second dual/primal buffers for the ping-pong (u too for the
        // two-step kernel, which reads u on a halo).
        float *p2 = nullptr, *v2 = nullptr, *u2 = nullptr;
        uint2* h8 = nullptr;
        auto release = [&]() {
            if (p2) (void)hipFree(p2);
            if (v2) (void)hipFree(v2);
            if (u2) (void)hipFree(u2);
            if (h8) (void)hipFree(h8);
        };
        hipError_t e = hipMalloc((void**)&p2, 3 * V * sizeof(float));
        if (e == hipSuccess) e = hipMalloc((void**)&v2, V * sizeof(float));
        if (e == hipSuccess && mode == 4) e = hipMalloc((void**)&u2, V * sizeof(float));
        if (e != hipSuccess) {
            release();
            return hip_status(e, "hipMalloc", PSW_ERR_ALLOC);
        }
        PdArgs a;
        a.f = f;
        a.hist8 = nullptr;
        a.tau = (float)tau; a.lambda = (float)lambda; a.tl = a.tau * a.lambda; a.sigma = (float)sigma;
        float* pb[2] = {g->p, p2};
        float* vb[2] = {g->v, v2};
        float* ub[2] = {g->u, mode == 4 ? u2 : g->u};
        const bool zmarch = mode >= 1;
        const dim3 blk = mode == 2 ? dim3(FZL_TX, FZL_TY, 1) : dim3(64, 4, 1);
        const dim3 grd = zmarch ? dim3((unsigned)((f.w + blk.x - 1) / blk.x), (unsigned)((f.h + blk.y - 1) / blk.y),
                                       (unsigned)((f.d + FZ_CHUNK - 1) / FZ_CHUNK))
                                : vox_grid(f, blk);
        const unsigned zm_blocks = (unsigned)(((f.w + FZ_BX - 1) / FZ_BX) * ((f.h + FZ_BY - 1) / FZ_BY) *
                                              ((f.d + FZ_CHUNK - 1) / FZ_CHUNK));
        const unsigned zp_blocks = (unsigned)(((f.w + FZ_BX - 1) / FZ_BX) * ((f.h + FZP_BY - 1) / FZP_BY) *
                                              ((f.d + FZP_CHUNK - 1) / FZP_CHUNK));
        const unsigned ft_blocks = (unsigned)(((f.w + FT_X - 1) / FT_X) * ((f.h + FT_Y - 1) / FT_Y) *
                                              ((f.d + FT_CHUNK - 1) / FT_CHUNK));
        // The pipelined kernels read the counts from a byte-packed copy when
        // every count fits in a byte (8 B per voxel-iteration less traffic),
        // else from the grid's own 16-bit histogram.
        if (mode >= 3 && hipMalloc((void**)&h8, V * sizeof(uint2) + 256) == hipSuccess) {
            unsigned* d_max = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(h8) + V * sizeof(uint2));
            unsigned hmax = 0xffffffffu;
            fusion_zero_u32_kernel<<<1, 1, 0, S(s)>>>(d_max);
            fusion_hist_pack_kernel<<<(unsigned)std::min<size_t>((V + 255) / 256, 8192), 256, 0, S(s)>>>(
                f.hist, h8, V, d_max);
            st = launch_status("fusion_hist_pack_kernel");
            if (st == PSW_OK &&
                (hipMemcpyAsync(&hmax, d_max, sizeof(unsigned), hipMemcpyDeviceToHost, S(s)) != hipSuccess ||
                 hipStreamSynchronize(S(s)) != hipSuccess))
                st = PSW_ERR_DEVICE;
            if (st == PSW_OK && hmax <= 255u) a.hist8 = h8;
        }
        int cur = 0;
        for (int k = 1; k < niters && st == PSW_OK;) {
            a.pin = pb[cur]; a.pout = pb[cur ^ 1];
            a.vin = vb[cur]; a.vout = vb[cur ^ 1];
            a.uin = ub[cur]; a.uout = ub[cur ^ 1];
            if (mode == 4 && niters - k >= 2) {
                if (a.hist8) fusion_pd2_kernel<true><<<ft_blocks, FT_NT, 0, S(s)>>>(a);
                else fusion_pd2_kernel<false><<<ft_blocks, FT_NT, 0, S(s)>>>(a);
                st = launch_status("fusion_pd2_kernel");
                k += 2;
            } else {
                if (mode == 2) {
                    fusion_pd_zlds_kernel<<<grd, blk, 0, S(s)>>>(a);
                    st = launch_status("fusion_pd_zlds_kernel");
                } else if (mode >= 3) {
                    if (a.hist8) fusion_pd_zpipe_kernel<true><<<zp_blocks, FZ_BX * FZP_BY, 0, S(s)>>>(a);
                    else fusion_pd_zpipe_kernel<false><<<zp_blocks, FZ_BX * FZP_BY, 0, S(s)>>>(a);
                    st = launch_status("fusion_pd_zpipe_kernel");
                } else if (zmarch) {
                    fusion_pd_zmarch_kernel<<<zm_blocks, FZ_BX * FZ_BY, 0, S(s)>>>(a);
                    st = launch_status("fusion_pd_zmarch_kernel");
                } else {
                    fusion_pd_fused_kernel<<<grd, blk, 0, S(s)>>>(a);
                    st = launch_status("fusion_pd_fused_kernel");
                }
                k += 1;
            }
            cur ^= 1;
        }
        if (st == PSW_OK && cur == 1) {
            // results live in the second buffers: move them home
            e = hipMemcpyAsync(g->p, p2, 3 * V * sizeof(float), hipMemcpyDeviceToDevice, S(s));
            if (e == hipSuccess) e = hipMemcpyAsync(g->v, v2, V * sizeof(float), hipMemcpyDeviceToDevice, S(s));
            if (e == hipSuccess && ub[1] != g->u)
                e = hipMemcpyAsync(g->u, ub[1], V * sizeof(float), hipMemcpyDeviceToDevice, S(s));
            if (e != hipSuccess) st = hip_status(e, "hipMemcpyAsync", PSW_ERR_COPY);
        }
        (void)hipStreamSynchronize(S(s));
        release();
        if (st != PSW_OK) return st;
    }
    return psw_fusion_update_p(g, sigma, s);
}

int psw_fusion_energy(const psw_fusion_grid* g, double lambda, double* energy, psw_stream_t s) {
    int st = check_grid(g);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(energy, "energy is NULL");
    double* d = nullptr;
    PSW_TRY_HIP(hipMalloc((void**)&d, FE_BLOCKS * sizeof(double)), PSW_ERR_ALLOC);
    fusion_energy_kernel<<<FE_BLOCKS, 256, 0, S(s)>>>(view_of(g), lambda, d);
    hipError_t e = hipGetLastError();
    double h[FE_BLOCKS];
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, S(s));
    if (e == hipSuccess) e = hipStreamSynchronize(S(s));
    (void)hipFree(d);
    if (e != hipSuccess) return hip_status(e, "psw_fusion_energy");
    double t = 0.0;
    for (int i = 0; i < FE_BLOCKS; ++i) t += h[i];
    *energy = t;
    return PSW_OK;
}

int psw_fusion_surface_points(const psw_fusion_grid* g, float* d_xyz, int64_t capacity, int64_t* count,
                              psw_stream_t s) {
    int st = check_grid(g);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(count, "count is NULL");
    PSW_CHECK_ARG(capacity >= 0 && (capacity == 0 || d_xyz), "bad output buffer");
    const FusionView f = view_of(g);
    const size_t per = (size_t)SP_BLOCK * SP_PER_THREAD;
    const size_t nb = (f.V + per - 1) / per;
    PSW_CHECK_ARG(nb < (1u << 31), "grid too large");
    unsigned* d = nullptr;
    PSW_TRY_HIP(hipMalloc((void**)&d, 2 * nb * sizeof(unsigned)), PSW_ERR_ALLOC);
    std::vector<unsigned> h(nb);
    hipStream_t hs = S(s);
    surface_points_count_kernel<<<(unsigned)nb, SP_BLOCK, 0, hs>>>(f.u, f.V, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, nb * sizeof(unsigned), hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    uint64_t total = 0;
    if (e == hipSuccess) {
        for (size_t b = 0; b < nb; ++b) {   // exclusive scan (counts fit: at most 4096 per block)
            const unsigned c = h[b];
            h[b] = (unsigned)std::min<uint64_t>(total, 0xffffffffu);
            total += c;
        }
        if (total >= 0xffffffffull) {
            (void)hipFree(d);
            PSW_CHECK_ARG(false, "too many surface voxels");
        }
        if (capacity > 0 && total > 0) {
            e = hipMemcpyAsync(d + nb, h.data(), nb * sizeof(unsigned), hipMemcpyHostToDevice, hs);
            if (e == hipSuccess) {
                surface_points_write_kernel<<<(unsigned)nb, SP_BLOCK, 0, hs>>>(f, d + nb, d_xyz, (size_t)capacity);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipStreamSynchronize(hs);
        }
    }
    (void)hipFree(d);
    if (e != hipSuccess) return hip_status(e, "psw_fusion_surface_points");
    *count = (int64_t)total;
    return PSW_OK;
}

int psw_fusion_surface_count(const psw_fusion_grid* g, int64_t* count, psw_stream_t s) {
    int st = check_grid(g);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(count, "count is NULL");
    const size_t V = (size_t)g->width * g->height * g->depth;
    unsigned long long* d = nullptr;
    PSW_TRY_HIP(hipMalloc((void**)&d, sizeof(unsigned long long)), PSW_ERR_ALLOC);
    hipError_t e = hipMemsetAsync(d, 0, sizeof(unsigned long long), S(s));
    if (e == hipSuccess) {
        size_t blocks = (V + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        surface_count_kernel<<<(unsigned)blocks, 256, 0, S(s)>>>(g->u, V, d);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, S(s));
    if (e == hipSuccess) e = hipStreamSynchronize(S(s));
    (void)hipFree(d);
    if (e != hipSuccess) return hip_status(e, "psw_fusion_surface_count");
    *count = (int64_t)h;
    return PSW_OK;
}

}  // extern "C"
