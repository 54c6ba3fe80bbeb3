// tvl1_fused.hip -- TVL1 depth denoising, PlaneSweep::CudaDenoise
// (src/planesweep.cpp:297-379): the tensor-weighted P step
// (src/kernels.cu:283-307) and the plain update (src/kernels.cu:342-370),
// iterated niters times on a depthmap scaled to [0, 1] over [znear, zfar].
//
// The reference launches the two steps separately for every iteration.  Here
// one launch does up to four whole iterations of a tile 8 rows high (4 and
// 16 selectable) and 64 - 2M columns wide for M iterations: the tile's depth
// and duals are staged in LDS with a halo of one pixel per iteration, the P step and the update alternate between
// barriers in LDS, and only the last iteration's state is written back.  Neighbouring tiles read
// each other's halo, so the state is double-buffered: a launch reads slot rd
// and writes slot 1 - rd.
//
// The per-pixel arithmetic is the launchers' (ref_kernels.hip) operation by
// operation, double where the reference declares double temporaries, no FMA
// contraction, so the result is bit-identical to driving the launchers in
// the reference's host sequence (PSW_TVL1_FUSED=0 does exactly that).
//
// Per pixel and launch the kernel reads depth, Px, Py, R, the scaled input
// and the four tensor entries (36 B) and writes depth, Px, Py, R (16 B).
#include "psw_common.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace psw {
namespace {

constexpr int TV_X = 64;

struct TvlState {
    float* out[2];
    float* px[2];
    float* py[2];
    float* r[2];
    const float* raw;  // scaled input depth, -sigma (d - znear) / (zfar - znear)
    const float* t11;
    const float* t12;
    const float* t21;
    const float* t22;
    int w, h;
    float tau, theta, lambda, sigma;
};

// P step at one pixel (kernels.cu:283-307).  c is the depth at the pixel, e
// and s its right and lower neighbours, already clamped to the image.
__device__ __forceinline__ void tvl1_p(float& px, float& py, float c, float e, float s, float t11, float t12,
                                       float t21, float t22, float sg) {
    const double x = (double)(e - c);
    const double y = (double)(s - c);
    const double dx = (double)px + (double)sg * ((double)t11 * x + (double)t12 * y);
    const double dy = (double)py + (double)sg * ((double)t21 * x + (double)t22 * y);
    const double n2 = dx * dx + dy * dy;
    if (n2 <= 1.0) {   // sqrt(n2) <= 1, so the divisor is exactly 1 and the quotients are dx, dy
        px = (float)dx;
        py = (float)dy;
        return;
    }
    const double d = (double)fmaxf(1.f, (float)sqrt(n2));
    px = (float)(dx / d);
    py = (float)(dy / d);
}

// Update at one pixel (kernels.cu:342-370): the data dual r, then the
// over-relaxed depth.  pxl / pyu are Px left and Py above (Py above is the
// pixel's own on the first row); the first column has no x term.
__device__ __forceinline__ float tvl1_u(float o, float& r, float raw, float pxc, float pxl, float pyc, float pyu,
                                        bool first_col, float tau, float theta, float lambda, float sigma) {
    r += raw;
    r += sigma * o;
    if (r > lambda) r = lambda;
    if (r < -lambda) r = -lambda;
    double x_new;
    if (first_col) x_new = (double)(o + tau * (pyc - pyu) - tau * r);
    else x_new = (double)(o + tau * (pxc - pxl + pyc - pyu) - tau * r);
    return (float)(x_new + (double)theta * (x_new - (double)o));
}

// M iterations (1 to 6) on one tile.  Region coordinates are local to the
// tile's staging window, which starts M pixels up and left of the tile: the
// P step of iteration k covers [k, R-2-k], the update [k+1, R-2-k], so the
// last update covers exactly the tile.  Pixels outside the image are never
// computed, and no pixel inside it reads one (the reference clamps its
// forward differences and drops the backward ones at the first row/column).
// The tensor is loaded by the first P step and kept in LDS for the others,
// R by the first update; the scaled input is read where it is used.
//
// TX is the tile width; the block has 64 x TV_Y threads either way.  With
// TX = 64 - 2M (the default) the staging window is 64 wide, so a position's
// window row and column are a shift and a mask of its index, and each pass of
// the block covers TV_Y whole rows (positions outside a phase's columns idle).
//
// Measured at 640x480, us per iteration: 3.49 for M = 4 on 56x8 tiles, 3.62
// for M = 3 on 58x8, 3.99 for M = 2 on 60x8, 4.29 for M = 2 on 64x8.  A
// four-iteration launch moves 17.3 MB (12.4 MB read, 4.9 MB written), far
// below what HBM allows in its 14 us: the kernel is bound by instruction
// issue.  Slower restructurings of the two-iteration case, with identical
// results: every input staged in the first phase (9.5 us per launch against
// 8.9; 9.4 us with all loads issued before the first LDS store) and
// branch-free fixed passes that let the compiler interleave a thread's two
// positions (9.3 us).
template <int M, int TV_Y, int TX>
__global__ __launch_bounds__(TV_X * TV_Y) void tvl1_iter_kernel(const TvlState a, int rd, float sp0, float sp1) {
    constexpr int TV_NT = TV_X * TV_Y;
    constexpr int RX = TX + 2 * M, RY = TV_Y + 2 * M, RN = RX * RY;
    __shared__ float s_o[RN], s_px[RN], s_py[RN];
    __shared__ float s_r[M > 1 ? RN : 1];
    __shared__ float s_t[M > 1 ? 4 * RN : 1];
    const int W = a.w, H = a.h;
    const int tx = (W + TX - 1) / TX, ty = (H + TV_Y - 1) / TV_Y;
    const int t = xcd_tile(blockIdx.x, tx * ty);
    const int gx0 = (t % tx) * TX - M, gy0 = (t / tx) * TV_Y - M;  // image position of local (0, 0)
    const int wr = 1 - rd;
    {
        const float* __restrict__ go = a.out[rd];
        const float* __restrict__ gpx = a.px[rd];
        const float* __restrict__ gpy = a.py[rd];
        for (int i = threadIdx.x; i < RN; i += TV_NT) {
            const int ly = i / RX, lx = i - ly * RX;
            const int gx = gx0 + lx, gy = gy0 + ly;
            if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
                const size_t g = (size_t)gy * W + gx;
                s_o[i] = go[g];
                s_px[i] = gpx[g];
                s_py[i] = gpy[g];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const float sg = k == 0 ? sp0 : sp1;
        {   // P step over [k, RX-2-k] x [k, RY-2-k], by whole window rows
            const int PY = RY - 1 - 2 * k;
            for (int i = threadIdx.x; i < PY * RX; i += TV_NT) {
                const int ly = k + i / RX, lx = i % RX;
                const int gx = gx0 + lx, gy = gy0 + ly;
                if (lx < k || lx > RX - 2 - k || gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
                const int li = ly * RX + lx;
                const float c = s_o[li];
                const float e = gx + 1 < W ? s_o[li + 1] : c;
                const float s = gy + 1 < H ? s_o[li + RX] : c;
                float t11, t12, t21, t22;
                if (k == 0) {
                    const size_t g = (size_t)gy * W + gx;
                    t11 = a.t11[g];
                    t12 = a.t12[g];
                    t21 = a.t21[g];
                    t22 = a.t22[g];
                    if (M > 1) {
                        s_t[li] = t11;
                        s_t[RN + li] = t12;
                        s_t[2 * RN + li] = t21;
                        s_t[3 * RN + li] = t22;
                    }
                } else {
                    t11 = s_t[li];
                    t12 = s_t[RN + li];
                    t21 = s_t[2 * RN + li];
                    t22 = s_t[3 * RN + li];
                }
                tvl1_p(s_px[li], s_py[li], c, e, s, t11, t12, t21, t22, sg);
            }
        }
        __syncthreads();
        {   // update over [k+1, RX-2-k] x [k+1, RY-2-k], by whole window rows
            const int UY = RY - 2 - 2 * k;
            for (int i = threadIdx.x; i < UY * RX; i += TV_NT) {
                const int ly = k + 1 + i / RX, lx = i % RX;
                const int gx = gx0 + lx, gy = gy0 + ly;
                if (lx < k + 1 || lx > RX - 2 - k || gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
                const int li = ly * RX + lx;
                const size_t g = (size_t)gy * W + gx;
                float r = k == 0 ? a.r[rd][g] : s_r[li];
                const float pxc = s_px[li], pyc = s_py[li];
                const float pxl = gx > 0 ? s_px[li - 1] : pxc;
                const float pyu = gy > 0 ? s_py[li - RX] : pyc;
                const float nv = tvl1_u(s_o[li], r, a.raw[g], pxc, pxl, pyc, pyu, gx == 0, a.tau, a.theta, a.lambda,
                                        a.sigma);
                if (k == M - 1) {
                    a.out[wr][g] = nv;
                    a.r[wr][g] = r;
                    a.px[wr][g] = pxc;
                    a.py[wr][g] = pyc;
                } else {
                    s_o[li] = nv;
                    s_r[li] = r;
                }
            }
        }
        if (k < M - 1) __syncthreads();
    }
}

// Two iterations in three phases.  The update is cheap next to the P step
// (no square root or division), so the first update is not given a phase of
// its own: the thread of each position of the second P step's region
// evaluates it at its position and at its right and lower neighbours, the
// three values the second P step reads, and keeps its own.  The first P step
// reads its neighbours' depth from global memory (cache hits), so no staging
// phase precedes it either.  Phases: P1 (global -> LDS), U1 x3 + P2, U2
// (LDS -> global).  Every evaluation is the per-pixel arithmetic above, so
// the results are those of tvl1_iter_kernel<2>.  Local (0, 0) is two pixels
// up and left of the tile.  Used for two-iteration launches on 64-wide tiles
// (PSW_TVL1_NARROW=0), where it measured 4.14 against 4.29 us per iteration.
template <int TV_Y>
__global__ __launch_bounds__(TV_X * TV_Y) void tvl1_iter2m_kernel(const TvlState a, int rd, float sp0, float sp1) {
    constexpr int TV_NT = TV_X * TV_Y;
    constexpr int RX = TV_X + 4, RY = TV_Y + 4, RN = RX * RY;
    __shared__ float s_o[RN], s_px[RN], s_py[RN], s_r[RN], s_in[RN];    // state k, P1, inputs: [0, R-2]
    __shared__ float s_t[4][RN];
    __shared__ float s_px2[RN], s_py2[RN], s_o1[RN], s_r1[RN];          // P2, depth and R of k+1: [1, R-3]
    const int W = a.w, H = a.h;
    const int tx = (W + TV_X - 1) / TV_X, ty = (H + TV_Y - 1) / TV_Y;
    const int t = xcd_tile(blockIdx.x, tx * ty);
    const int gx0 = (t % tx) * TV_X - 2, gy0 = (t / tx) * TV_Y - 2;
    const int wr = 1 - rd;
    {   // P1 over [0, RX-2] x [0, RY-2]
        const float* __restrict__ go = a.out[rd];
        constexpr int PX = RX - 1, PY = RY - 1;
        for (int i = threadIdx.x; i < PX * PY; i += TV_NT) {
            const int ly = i / PX, lx = i - ly * PX;
            const int gx = gx0 + lx, gy = gy0 + ly;
            if (gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
            const size_t g = (size_t)gy * W + gx;
            const float c = go[g];
            const float e = gx + 1 < W ? go[g + 1] : c;
            const float s = gy + 1 < H ? go[g + W] : c;
            const float t11 = a.t11[g], t12 = a.t12[g], t21 = a.t21[g], t22 = a.t22[g];
            float px = a.px[rd][g], py = a.py[rd][g];
            tvl1_p(px, py, c, e, s, t11, t12, t21, t22, sp0);
            const int li = ly * RX + lx;
            s_o[li] = c;
            s_px[li] = px;
            s_py[li] = py;
            s_t[0][li] = t11;
            s_t[1][li] = t12;
            s_t[2][li] = t21;
            s_t[3][li] = t22;
            s_r[li] = a.r[rd][g];
            s_in[li] = a.raw[g];
        }
    }
    __syncthreads();
    {   // U1 at each position and its right and lower neighbours, P2 at the position: [1, RX-3] x [1, RY-3]
        constexpr int PX = RX - 3, PY = RY - 3;
        auto u1 = [&](int lj, int gxj, int gyj, float& r) {
            r = s_r[lj];
            const float pxc = s_px[lj], pyc = s_py[lj];
            const float pxl = gxj > 0 ? s_px[lj - 1] : pxc;
            const float pyu = gyj > 0 ? s_py[lj - RX] : pyc;
            return tvl1_u(s_o[lj], r, s_in[lj], pxc, pxl, pyc, pyu, gxj == 0, a.tau, a.theta, a.lambda, a.sigma);
        };
        for (int i = threadIdx.x; i < PX * PY; i += TV_NT) {
            const int ly = 1 + i / PX, lx = 1 + i % PX;
            const int gx = gx0 + lx, gy = gy0 + ly;
            if (gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
            const int li = ly * RX + lx;
            float r1, rn;
            const float c1 = u1(li, gx, gy, r1);
            const float e1 = gx + 1 < W ? u1(li + 1, gx + 1, gy, rn) : c1;
            const float s1 = gy + 1 < H ? u1(li + RX, gx, gy + 1, rn) : c1;
            float px = s_px[li], py = s_py[li];
            tvl1_p(px, py, c1, e1, s1, s_t[0][li], s_t[1][li], s_t[2][li], s_t[3][li], sp1);
            s_px2[li] = px;
            s_py2[li] = py;
            s_o1[li] = c1;
            s_r1[li] = r1;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TV_NT; i += TV_NT) {   // U2 over the tile, to global memory
        const int ly = 2 + i / TV_X, lx = 2 + i % TV_X;
        const int gx = gx0 + lx, gy = gy0 + ly;
        if (gx >= W || gy >= H) continue;
        const int li = ly * RX + lx;
        float r = s_r1[li];
        const float pxc = s_px2[li], pyc = s_py2[li];
        const float pxl = gx > 0 ? s_px2[li - 1] : pxc;
        const float pyu = gy > 0 ? s_py2[li - RX] : pyc;
        const float nv = tvl1_u(s_o1[li], r, s_in[li], pxc, pxl, pyc, pyu, gx == 0, a.tau, a.theta, a.lambda,
                                a.sigma);
        const size_t g = (size_t)gy * W + gx;
        a.out[wr][g] = nv;
        a.r[wr][g] = r;
        a.px[wr][g] = pxc;
        a.py[wr][g] = pyc;
    }
}

template <int TY, int M>
int tvl1_launch_m(const TvlState& tv, bool narrow, int rd, float sp0, float sp1, hipStream_t s) {
    if (narrow) {
        constexpr int TX = TV_X - 2 * M;   // a 64-wide staging window
        const int tiles = ((tv.w + TX - 1) / TX) * ((tv.h + TY - 1) / TY);
        tvl1_iter_kernel<M, TY, TX><<<tiles, TV_X * TY, 0, s>>>(tv, rd, sp0, sp1);
    } else {
        const int tiles = ((tv.w + TV_X - 1) / TV_X) * ((tv.h + TY - 1) / TY);
        tvl1_iter_kernel<M, TY, TV_X><<<tiles, TV_X * TY, 0, s>>>(tv, rd, sp0, sp1);
    }
    return launch_status("tvl1_iter_kernel");
}

template <int TY>
int tvl1_launch_ty(const TvlState& tv, bool merged, bool narrow, int m, int rd, float sp0, float sp1,
                   hipStream_t s) {
    if constexpr (TY <= 8) {   // the three-phase kernel's LDS (13 arrays) fits 64 KB up to 64x8 tiles
        if (m == 2 && merged && !narrow) {
            const int tiles = ((tv.w + TV_X - 1) / TV_X) * ((tv.h + TY - 1) / TY);
            tvl1_iter2m_kernel<TY><<<tiles, TV_X * TY, 0, s>>>(tv, rd, sp0, sp1);
            return launch_status("tvl1_iter2m_kernel");
        }
    }
    switch (m) {
        case 6: return tvl1_launch_m<TY, 6>(tv, narrow, rd, sp0, sp1, s);
        case 5: return tvl1_launch_m<TY, 5>(tv, narrow, rd, sp0, sp1, s);
        case 4: return tvl1_launch_m<TY, 4>(tv, narrow, rd, sp0, sp1, s);
        case 3: return tvl1_launch_m<TY, 3>(tv, narrow, rd, sp0, sp1, s);
        case 2: return tvl1_launch_m<TY, 2>(tv, narrow, rd, sp0, sp1, s);
        default: return tvl1_launch_m<TY, 1>(tv, narrow, rd, sp0, sp1, s);
    }
}

// Tile height (identical results): 8 (512 threads, the default), 4 or 16,
// PSW_TVL1_TY.
int tvl1_ty() {
    const char* e = std::getenv("PSW_TVL1_TY");
    const int v = e ? std::atoi(e) : 8;
    return v == 4 || v == 16 ? v : 8;
}

// Iterations per launch at most (PSW_TVL1_M: 2 to 6, 4 the default, at most 4
// on 64x16 tiles; identical results).  At 640x480 with 64-wide staging
// windows: 3.99, 3.62, 3.44-3.49, 3.35-3.39 and 4.05-4.11 us per iteration
// for 2..6.
int tvl1_max_per_launch() {
    const char* e = std::getenv("PSW_TVL1_M");
    const int v = e ? std::atoi(e) : 4;
    return v >= 2 && v <= 6 ? v : 4;
}

// Two-iteration launches in three phases (tvl1_iter2m_kernel) unless
// PSW_TVL1_MERGED=0; identical results.
bool tvl1_merged() {
    const char* e = std::getenv("PSW_TVL1_MERGED");
    return !(e && std::atoi(e) == 0);
}

// Tiles 64 - 2M wide, so that the staging window is 64 wide, unless
// PSW_TVL1_NARROW=0 (64-wide tiles); identical results.  Two iterations per
// launch measured 3.99 us per iteration on 60-wide tiles against 4.20 on
// 64-wide ones (three-phase kernel), three 3.62 against 4.18.
bool tvl1_narrow() {
    const char* e = std::getenv("PSW_TVL1_NARROW");
    return !(e && std::atoi(e) == 0);
}

int tvl1_launch(const TvlState& tv, int ty, bool merged, bool narrow, int m, int rd, float sp0, float sp1,
                hipStream_t s) {
    switch (ty) {
        case 4: return tvl1_launch_ty<4>(tv, merged, narrow, m, rd, sp0, sp1, s);
        case 16: return tvl1_launch_ty<16>(tv, merged, narrow, m, rd, sp0, sp1, s);
        default: return tvl1_launch_ty<8>(tv, merged, narrow, m, rd, sp0, sp1, s);
    }
}

// Fused iterations unless PSW_TVL1_FUSED=0 (the reference's launch sequence;
// identical results).
bool tvl1_fused() {
    const char* e = std::getenv("PSW_TVL1_FUSED");
    return !(e && std::atoi(e) == 0);
}

// ---- resident solver ---------------------------------------------------------
// All iterations in one cooperative launch: workgroup t owns a 64 x 20 tile
// (ten waves, thread (wave w, lane x) the pixels of rows 2w and 2w+1 in
// column x), keeps its pixels' depth o, duals Px, Py, R, input and tensor in
// registers for the whole run, and trades only the depth of its edges with
// its eight neighbours each iteration.  The P step of the row above the tile
// and of the column left of it (the ring: the update reads Px on the left
// and Py above) is recomputed here from the depth halo, so one exchange per
// iteration suffices.  The per-pixel arithmetic is tvl1_p and tvl1_u, so the
// results are those of the fused launches and of the reference's sequence.
//
// Exchange words are 64-bit (iteration tag, float bits) in uncached device
// memory, double-buffered by the parity of the iteration: a tile overwrites
// its words of iteration j only after its neighbours have published j + 1,
// which they do after reading j.  Every wait is bounded (spin polls); one
// that runs out sets the error words, which end the other waits too.  A
// launch whose workgroups are not all resident within start_spin polls is
// called off before any wait on a neighbour and before anything is written.
constexpr int TVR_TX = 64, TVR_TY = 20, TVR_NT = TVR_TX * (TVR_TY / 2);
constexpr int TVR_TOP = 0, TVR_BOT = TVR_TX, TVR_LEFT = 2 * TVR_TX, TVR_RIGHT = 2 * TVR_TX + TVR_TY;
constexpr int TVR_XW = 2 * TVR_TX + 2 * TVR_TY;       // words a tile publishes per iteration
constexpr int TVR_HALO = 2 * TVR_TX + 2 * TVR_TY + 2;  // halo cells a tile reads per iteration
constexpr int TVR_RING0 = 2 * TVR_TX;                  // ring threads: waves 2 and 3
constexpr int TVR_RING = TVR_TX + TVR_TY;
constexpr int TVR_MAXTILES = 256;
constexpr int TVR_PAD = 20480;                         // 80 KB, see tvl1_resident_kernel
static_assert(TVR_HALO <= TVR_NT && TVR_RING0 + TVR_RING + TVR_TY <= 4 * TVR_TX,
              "one halo cell per thread; the ring and the right column's workers in waves 2 and 3");

struct TvrArgs {
    TvlState tv;
    int rd;                     // slot of the start state; the end state goes to slot 1 - rd
    int niters;
    float sp0, sp1;             // sigma of the first P step and of the others
    int tiles_x, tiles_y, ntiles;
    unsigned long long* xch;    // [2][ntiles][TVR_XW]
    unsigned* err;              // [0] error, [1] started, [2] go (1) / called off (2), [3..7] first failure:
                                // tile, word offset from xch, awaited tag, seen tag, iteration
    unsigned spin, start_spin;
};

__device__ __forceinline__ unsigned long long tvr_load(unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned tvr_load32(unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void tvr_put(unsigned long long* p, unsigned tag, float v) {
    __hip_atomic_store(p, ((unsigned long long)tag << 32) | __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Poll word p until it carries `tag`; its float bits go to v.  False when
// the spin budget runs out (the first failure is recorded) or another wait
// has failed.
__device__ bool tvr_get(const TvrArgs& A, unsigned long long* p, unsigned tag, float& v, int tile, int j) {
    for (unsigned it = 0; it < A.spin; ++it) {
        const unsigned long long w = tvr_load(p);
        if ((unsigned)(w >> 32) == tag) {
            v = __uint_as_float((unsigned)w);
            return true;
        }
        if ((it & 63) == 63 && tvr_load32(A.err) != 0u) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    unsigned expected = 0u;
    if (__hip_atomic_compare_exchange_strong(A.err, &expected, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)) {
        const unsigned vals[5] = {(unsigned)tile, (unsigned)(p - A.xch), tag, (unsigned)(tvr_load(p) >> 32),
                                  (unsigned)j};
        for (int k = 0; k < 5; ++k) __hip_atomic_store(A.err + 3 + k, vals[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return false;
}

// Halo cell c of tile (tcx, tcy): the row above (c < 64), the above-right
// corner, the column left, the below-left corner, the column right, the row
// below.  False where the cell is outside the image or has no neighbour
// tile; else the source tile's word, the LDS cell (row ly, column lx of the
// 22 x 66 depth array) and the image position.
__device__ inline bool tvr_halo(const TvrArgs& A, int tcx, int tcy, int c, int& off, int& ly, int& lx, int& gx,
                                int& gy) {
    constexpr int TX = TVR_TX, TY = TVR_TY;
    const int W = A.tv.w, H = A.tv.h;
    const int x0 = tcx * TX, y0 = tcy * TY, tile = tcy * A.tiles_x + tcx;
    const bool has_l = tcx > 0, has_r = tcx < A.tiles_x - 1, has_u = tcy > 0, has_d = tcy < A.tiles_y - 1;
    int src, word;
    bool ok;
    if (c < TX) {
        gx = x0 + c; gy = y0 - 1; src = tile - A.tiles_x; word = TVR_BOT + c; ly = 0; lx = c + 1;
        ok = has_u && gx < W;
    } else if (c == TX) {
        gx = x0 + TX; gy = y0 - 1; src = tile - A.tiles_x + 1; word = TVR_BOT; ly = 0; lx = TX + 1;
        ok = has_u && has_r;
    } else if (c < TX + 1 + TY) {
        const int r = c - (TX + 1);
        gx = x0 - 1; gy = y0 + r; src = tile - 1; word = TVR_RIGHT + r; ly = r + 1; lx = 0;
        ok = has_l && gy < H;
    } else if (c == TX + 1 + TY) {
        gx = x0 - 1; gy = y0 + TY; src = tile + A.tiles_x - 1; word = TVR_TOP + TX - 1; ly = TY + 1; lx = 0;
        ok = has_l && has_d;
    } else if (c < TX + 2 + 2 * TY) {
        const int r = c - (TX + 2 + TY);
        gx = x0 + TX; gy = y0 + r; src = tile + 1; word = TVR_LEFT + r; ly = r + 1; lx = TX + 1;
        ok = has_r && gy < H;
    } else {
        const int cc = c - (TX + 2 + 2 * TY);
        gx = x0 + cc; gy = y0 + TY; src = tile + A.tiles_x; word = TVR_TOP + cc; ly = TY + 1; lx = cc + 1;
        ok = has_d && gx < W;
    }
    off = src * TVR_XW + word;
    return ok;
}

__global__ __launch_bounds__(TVR_NT) void tvl1_resident_kernel(const TvrArgs A) {
    constexpr int TX = TVR_TX, TY = TVR_TY;
    __shared__ float sO[TY + 2][TX + 2];         // depth o: the tile and its halo
    __shared__ float sP[2][TY + 1][TX + 1];      // Px, Py: the tile, the row above and the column left
    __shared__ float sT[4][TY + 1][TX + 1];      // the tensor of the ring and of the right column
    __shared__ int sFlag[1];
    // LDS held only so that a CU takes one workgroup (the tiles run in
    // lockstep, so two on one CU would halve the whole solve's speed)
    __shared__ float sPad[TVR_PAD];
    const TvlState& a = A.tv;
    const int W = a.w, H = a.h;
    const int tile = xcd_tile(blockIdx.x, A.ntiles);
    const int tcx = tile % A.tiles_x, tcy = tile / A.tiles_x;
    const int x0 = tcx * TX, y0 = tcy * TY;
    const int tid = threadIdx.x, tx = tid % TX;
    const int wv = __builtin_amdgcn_readfirstlane(tid / TX);
    const int x = x0 + tx;
    const int rd = A.rd, wr = 1 - rd;
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
    float o[2], px[2], py[2], r[2], raw[2], t11[2], t12[2], t21[2], t22[2];
    bool in[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int y = y0 + 2 * wv + h;
        in[h] = x < W && y < H;
        const size_t g = (size_t)min(y, H - 1) * W + min(x, W - 1);
        o[h] = a.out[rd][g]; px[h] = a.px[rd][g]; py[h] = a.py[rd][g]; r[h] = a.r[rd][g];
        raw[h] = a.raw[g]; t11[h] = a.t11[g]; t12[h] = a.t12[g]; t21[h] = a.t21[g]; t22[h] = a.t22[g];
        sO[2 * wv + h + 1][tx + 1] = o[h];
        if (tx == TX - 1) {   // the right column's P step runs from LDS (see the loop)
            sP[0][2 * wv + h + 1][TX] = px[h]; sP[1][2 * wv + h + 1][TX] = py[h];
            sT[0][2 * wv + h + 1][TX] = t11[h]; sT[1][2 * wv + h + 1][TX] = t12[h];
            sT[2][2 * wv + h + 1][TX] = t21[h]; sT[3][2 * wv + h + 1][TX] = t22[h];
        }
    }
    // the ring: its duals and tensor of the start state
    const int rid = tid - TVR_RING0;
    bool ring = false;
    int rgx = 0, rgy = 0, rlx = 0, rly = 0;
    if (rid >= 0 && rid < TVR_RING) {
        if (rid < TX) { rgx = x0 + rid; rgy = y0 - 1; rlx = rid + 1; rly = 0; ring = tcy > 0 && rgx < W; }
        else { rgx = x0 - 1; rgy = y0 + rid - TX; rlx = 0; rly = rid - TX + 1; ring = tcx > 0 && rgy < H; }
        if (ring) {
            const size_t g = (size_t)rgy * W + rgx;
            sP[0][rly][rlx] = a.px[rd][g]; sP[1][rly][rlx] = a.py[rd][g];
            sT[0][rly][rlx] = a.t11[g]; sT[1][rly][rlx] = a.t12[g];
            sT[2][rly][rlx] = a.t21[g]; sT[3][rly][rlx] = a.t22[g];
        }
    }
    // the depth halo of the start state; later ones come from the neighbours
    int hoff = -1, hly = 0, hlx = 0;
    if (tid < TVR_HALO) {
        int gx, gy;
        if (tvr_halo(A, tcx, tcy, tid, hoff, hly, hlx, gx, gy)) sO[hly][hlx] = a.out[rd][(size_t)gy * W + gx];
        else hoff = -1;
    }
    // every workgroup resident before any waits on a neighbour
    if (tid == 0) {
        unsigned go = tvr_load32(A.err + 2);
        for (unsigned it = 0; go == 0u && it < A.start_spin; ++it) {
            __builtin_amdgcn_s_sleep(2);
            go = tvr_load32(A.err + 2);
        }
        if (go == 0u) {
            unsigned expected = 0u;
            __hip_atomic_compare_exchange_strong(A.err + 2, &expected, 2u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            go = tvr_load32(A.err + 2);
        }
        if (go != 1u) sFlag[0] = 1;
    }
    __syncthreads();
    if (sFlag[0]) return;   // called off: nothing written

    unsigned long long* const xb0 = A.xch + (size_t)tile * TVR_XW;   // this tile's words (parity 0)
    bool failed = false;
    // The P step of a pixel reads the depth right of it and below it, so
    // only the tile's right column and bottom row (and the ring) need the
    // halo: the others are done before the wait for the neighbours' edges,
    // in its shadow.  The bottom row is wave 9's second row, done by that
    // wave after the wait; the right column (one lane of every wave) is done
    // by lanes 20..39 of wave 3 from LDS, and its owners take their duals
    // back from there for the update.
    const int rc = tid - (TVR_RING0 + TVR_RING);   // right-column worker: pixel (x0 + 63, y0 + rc)
    const bool rcol = rc >= 0 && rc < TY && x0 + TX - 1 < W && y0 + rc < H;
    const bool last_col = tx == TX - 1;
    for (int j = 0; j < A.niters; ++j) {
        const float sg = j == 0 ? A.sp0 : A.sp1;
        // P step of the pixels off the right column and the bottom row
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = 2 * wv + h, y = y0 + row;
            if (row == TY - 1) continue;   // wave-uniform
            const float c = sO[row + 1][tx + 1];
            const float e_ = sO[row + 1][tx + 2], s_ = sO[row + 2][tx + 1];   // read by every lane, then chosen
            if (in[h] && !last_col) {
                tvl1_p(px[h], py[h], c, x + 1 < W ? e_ : c, y + 1 < H ? s_ : c, t11[h], t12[h], t21[h], t22[h], sg);
                sP[0][row + 1][tx + 1] = px[h];
                sP[1][row + 1][tx + 1] = py[h];
            }
        }
        // the neighbours' edges of state j (from the start state at j = 0)
        if (j > 0 && hoff >= 0) {
            float v;
            if (tvr_get(A, A.xch + (size_t)(j & 1) * A.ntiles * TVR_XW + hoff, (unsigned)j, v, tile, j)) sO[hly][hlx] = v;
            else sFlag[0] = 1;
        }
        __syncthreads();
        if (sFlag[0]) { failed = true; break; }
        // P step of the bottom row, the ring and the right column
        if (2 * wv + 1 == TY - 1) {   // wave-uniform
            const int row = TY - 1, y = y0 + row;
            const float c = sO[row + 1][tx + 1];
            const float e_ = sO[row + 1][tx + 2], s_ = sO[row + 2][tx + 1];
            if (in[1] && !last_col) {
                tvl1_p(px[1], py[1], c, x + 1 < W ? e_ : c, y + 1 < H ? s_ : c, t11[1], t12[1], t21[1], t22[1], sg);
                sP[0][row + 1][tx + 1] = px[1];
                sP[1][row + 1][tx + 1] = py[1];
            }
        }
        if (ring) {
            const float c = sO[rly][rlx];
            const float e_ = rgx + 1 < W ? sO[rly][rlx + 1] : c;
            const float s_ = rgy + 1 < H ? sO[rly + 1][rlx] : c;
            float rpx = sP[0][rly][rlx], rpy = sP[1][rly][rlx];
            tvl1_p(rpx, rpy, c, e_, s_, sT[0][rly][rlx], sT[1][rly][rlx], sT[2][rly][rlx], sT[3][rly][rlx], sg);
            sP[0][rly][rlx] = rpx;
            sP[1][rly][rlx] = rpy;
        }
        if (rcol) {
            const int gx = x0 + TX - 1, gy = y0 + rc;
            const float c = sO[rc + 1][TX];
            const float e_ = gx + 1 < W ? sO[rc + 1][TX + 1] : c;
            const float s_ = gy + 1 < H ? sO[rc + 2][TX] : c;
            float cpx = sP[0][rc + 1][TX], cpy = sP[1][rc + 1][TX];
            tvl1_p(cpx, cpy, c, e_, s_, sT[0][rc + 1][TX], sT[1][rc + 1][TX], sT[2][rc + 1][TX], sT[3][rc + 1][TX], sg);
            sP[0][rc + 1][TX] = cpx;
            sP[1][rc + 1][TX] = cpy;
        }
        __syncthreads();
        // update: R and the over-relaxed depth of state j + 1; the tile's
        // edges for the neighbours
        const bool publish = j + 1 < A.niters;
        unsigned long long* const xb = xb0 + (size_t)((j + 1) & 1) * A.ntiles * TVR_XW;
        const unsigned tag = (unsigned)(j + 1);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = 2 * wv + h, y = y0 + row;
            const float pxl_ = sP[0][row + 1][tx], pyu_ = sP[1][row][tx + 1];
            const float pxo = sP[0][row + 1][tx + 1], pyo = sP[1][row + 1][tx + 1];
            if (in[h]) {
                if (last_col) { px[h] = pxo; py[h] = pyo; }   // the right column's duals, from its worker
                const float pxl = x > 0 ? pxl_ : px[h], pyu = y > 0 ? pyu_ : py[h];
                o[h] = tvl1_u(o[h], r[h], raw[h], px[h], pxl, py[h], pyu, x == 0, a.tau, a.theta, a.lambda, a.sigma);
                if (publish) {
                    if (row == 0) tvr_put(xb + TVR_TOP + tx, tag, o[h]);
                    if (row == TY - 1) tvr_put(xb + TVR_BOT + tx, tag, o[h]);
                    if (tx == 0) tvr_put(xb + TVR_LEFT + row, tag, o[h]);
                    if (tx == TX - 1) tvr_put(xb + TVR_RIGHT + row, tag, o[h]);
                }
            }
        }
        if (!publish) break;
        // the own depth of state j + 1 for the next P step (its readers have
        // passed the barrier above; the next one orders it before them)
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (in[h]) sO[2 * wv + h + 1][tx + 1] = o[h];
        __syncthreads();
    }
    if (failed) return;
    // the pad is read on a branch never taken (the host passes niters >= 2),
    // a select with no arithmetic, so that it stays allocated
    if (A.niters < 0) {
        sPad[tid] = o[0];
        __syncthreads();
        o[0] = sPad[(tid + 1) % TVR_NT];
    }
    // the end state into slot wr
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (!in[h]) continue;
        const size_t g = (size_t)(y0 + 2 * wv + h) * W + x;
        a.out[wr][g] = o[h]; a.px[wr][g] = px[h]; a.py[wr][g] = py[h]; a.r[wr][g] = r[h];
    }
}

// The exchange buffer of the resident solver on each device: uncached
// memory for the largest launch (2 x 256 tiles x 168 edge words, 8 error
// words), allocated once and kept for the process (uncached memory that a
// kernel has read is never freed).  Also the CU count, and the tile counts
// whose launches were called off (those keep using the fused launches).
struct TvrDev {
    std::mutex mu;
    unsigned long long* buf[64] = {};
    int cus[64] = {};
    std::vector<int> declined[64];
    // one resident run per device at a time (they share the exchange buffer
    // and would compete for the CUs), and its error words on the host
    std::mutex run[64];
    unsigned* herr[64] = {};
};
TvrDev& tvr_dev() {
    static TvrDev* d = new TvrDev();
    return *d;
}
constexpr size_t TVR_WORDS = 2 * (size_t)TVR_MAXTILES * TVR_XW + 4;

// The resident solver runs unless PSW_TVL1_RESIDENT=0, for frames of at
// most one 64 x 20 tile per CU (and at most 256 tiles) and runs of at least
// two iterations; *buf receives the device's exchange buffer.
// PSW_TVL1_RESIDENT=2 makes a run of two or more iterations that cannot use
// it an error (tests).
bool tvl1_resident_use(int W, int H, int niters, unsigned long long** buf) {
    const char* e = std::getenv("PSW_TVL1_RESIDENT");
    if ((e && std::atoi(e) == 0) || niters < 2) return false;
    const int nt = ((W + TVR_TX - 1) / TVR_TX) * ((H + TVR_TY - 1) / TVR_TY);
    if (nt > TVR_MAXTILES) return false;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return false;
    }
    TvrDev& D = tvr_dev();
    std::lock_guard<std::mutex> lock(D.mu);
    if (D.cus[dev] == 0) {
        int per = 0, ncu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)tvl1_resident_kernel, TVR_NT, 0) !=
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
        if (hipExtMallocWithFlags(&b, TVR_WORDS * sizeof(unsigned long long), hipDeviceMallocUncached) != hipSuccess) {
            (void)hipGetLastError();
            D.cus[dev] = -1;
            return false;
        }
        D.buf[dev] = static_cast<unsigned long long*>(b);
    }
    *buf = D.buf[dev];
    return true;
}

void tvl1_resident_decline(int nt) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    TvrDev& D = tvr_dev();
    std::lock_guard<std::mutex> lock(D.mu);
    D.declined[dev].push_back(nt);
}

bool tvl1_resident_required() {
    const char* e = std::getenv("PSW_TVL1_RESIDENT");
    return e && std::atoi(e) == 2;
}

unsigned tvr_env_u32(const char* name, unsigned dflt) {
    const char* e = std::getenv(name);
    return e ? (unsigned)std::strtoul(e, nullptr, 10) : dflt;
}

// niters iterations from slot rd into slot 1 - rd in one cooperative
// launch.  *ran is false, with nothing written, when the runtime refuses
// the launch or the launch is called off.  Synchronous (it reads the error
// words); runs on one device take turns, since they share its exchange
// buffer.  Polls per wait: PSW_TVL1_RESIDENT_SPIN (2^22), of the start:
// PSW_TVL1_RESIDENT_START_SPIN (2^16; set, a called-off launch is not
// remembered).
int tvl1_resident_run(const TvlState& tv, unsigned long long* buf, int rd, int niters, float sp0, float sp1,
                      hipStream_t st, bool* ran) {
    *ran = false;
    TvrArgs A;
    A.tv = tv;
    A.rd = rd;
    A.niters = niters;
    A.sp0 = sp0;
    A.sp1 = sp1;
    A.tiles_x = (tv.w + TVR_TX - 1) / TVR_TX;
    A.tiles_y = (tv.h + TVR_TY - 1) / TVR_TY;
    A.ntiles = A.tiles_x * A.tiles_y;
    A.xch = buf;
    A.err = reinterpret_cast<unsigned*>(buf + 2 * (size_t)A.ntiles * TVR_XW);
    A.spin = tvr_env_u32("PSW_TVL1_RESIDENT_SPIN", 1u << 22);
    A.start_spin = tvr_env_u32("PSW_TVL1_RESIDENT_START_SPIN", 1u << 16);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return PSW_OK;
    }
    TvrDev& D = tvr_dev();
    std::lock_guard<std::mutex> run_lock(D.run[dev]);
    if (!D.herr[dev]) {
        std::lock_guard<std::mutex> lock(D.mu);
        PSW_TRY_HIP(hipHostMalloc((void**)&D.herr[dev], 8 * sizeof(unsigned), hipHostMallocDefault), PSW_ERR_ALLOC);
    }
    unsigned* const he = D.herr[dev];
    // tags that match no iteration, and clear error words
    int rc = psw_fill_u32(reinterpret_cast<uint32_t*>(buf), 0xFFFFFFFFu, 4 * (size_t)A.ntiles * TVR_XW, st);
    if (rc == PSW_OK) rc = psw_fill_u32(A.err, 0u, 8, st);
    if (rc != PSW_OK) return rc;
    void* args[] = {&A};
    if (hipLaunchCooperativeKernel((const void*)tvl1_resident_kernel, dim3((unsigned)A.ntiles), dim3(TVR_NT), args, 0,
                                   st) != hipSuccess) {
        (void)hipGetLastError();
        return PSW_OK;
    }
    int out = PSW_OK;
    if (hipMemcpyAsync(he, A.err, 8 * sizeof(unsigned), hipMemcpyDeviceToHost, st) != hipSuccess) out = PSW_ERR_COPY;
    else if (hipStreamSynchronize(st) != hipSuccess) out = PSW_ERR_DEVICE;
    if (out == PSW_OK) {
        if (he[2] != 1u) {   // called off before any wait on a neighbour: nothing was written
            if (!std::getenv("PSW_TVL1_RESIDENT_START_SPIN")) tvl1_resident_decline(A.ntiles);
        } else {
            *ran = true;
            if (he[0]) {
                set_error("resident TVL1 solver: tile %u waited %u polls for word %u of tile %u (iteration tag %u, "
                          "saw %u) at iteration %u; the result is undefined",
                          he[3], A.spin, he[4] % TVR_XW, (he[4] % ((unsigned)A.ntiles * TVR_XW)) / TVR_XW, he[5],
                          he[6], he[7]);
                out = PSW_ERR_DEVICE;
            }
        }
    } else {
        (void)hipGetLastError();
    }
    return out;
}
}  // namespace
}  // namespace psw

using namespace psw;

// PlaneSweep::CudaDenoise (src/planesweep.cpp:297-379).  Workspace layout
// (n = width * height floats each): R, Px, Py, scaled input, T11, T12, T21,
// T22, then the second slot of R, Px, Py and the depth.
extern "C" int psw_tvl1_denoise(const float* d_depth, const float* d_ref, int width, int height, float znear,
                                float zfar, int niters, double lambda, double tau, double sigma, double theta,
                                double beta, double gamma, float* d_out, float* d_work, psw_stream_t s) {
    PSW_CHECK_ARG(d_depth && d_ref && d_out && d_work, "NULL pointer");
    PSW_CHECK_ARG(width >= 2 && height >= 2 && niters >= 0, "bad size or iteration count");
    const size_t n = (size_t)width * height;
    float* R = d_work;
    float* Px = d_work + n;
    float* Py = d_work + 2 * n;
    float* raw = d_work + 3 * n;
    float* T11 = d_work + 4 * n;
    float* T12 = d_work + 5 * n;
    float* T21 = d_work + 6 * n;
    float* T22 = d_work + 7 * n;
    const psw_dim3 z = {0, 0, 0};
    int st;
#define TRY(x) do { if ((st = (x)) != PSW_OK) return st; } while (0)
    TRY(psw_fill_f32(R, 0.f, 3 * n, s));  // R, Px, Py start at zero
    TRY(psw_memcpy_dtod(d_out, d_depth, n * sizeof(float), s));
    TRY(psw_memcpy_dtod(raw, d_depth, n * sizeof(float), s));
    // normalised reference image in Px for the tensor, then Px := 0 again
    TRY(psw_memcpy_dtod(Px, d_ref, n * sizeof(float), s));
    TRY(psw_element_scale(Px, 1 / 255.f, width, height, z, z, s));
    TRY(psw_Anisotropic_diffusion_tensor(T11, T12, T21, T22, Px, (float)beta, (float)gamma, width, height, z, z, s));
    TRY(psw_fill_f32(Px, 0.f, n, s));
    TRY(psw_element_add(d_out, -znear, width, height, z, z, s));
    TRY(psw_element_add(raw, -znear, width, height, z, z, s));
    const double xscale = 1.f / (zfar - znear);
    const double inputscale = -sigma / (double)(zfar - znear);
    TRY(psw_element_scale(d_out, (float)xscale, width, height, z, z, s));
    TRY(psw_element_scale(raw, (float)inputscale, width, height, z, z, s));
    if (tvl1_fused()) {
        TvlState tv;
        tv.out[0] = d_out;
        tv.out[1] = d_work + 11 * n;
        tv.px[0] = Px;
        tv.px[1] = d_work + 9 * n;
        tv.py[0] = Py;
        tv.py[1] = d_work + 10 * n;
        tv.r[0] = R;
        tv.r[1] = d_work + 8 * n;
        tv.raw = raw;
        tv.t11 = T11;
        tv.t12 = T12;
        tv.t21 = T21;
        tv.t22 = T22;
        tv.w = width;
        tv.h = height;
        tv.tau = (float)tau;
        tv.theta = (float)theta;
        tv.lambda = (float)lambda;
        tv.sigma = (float)sigma;
        const int ty = tvl1_ty();
        const bool merged = tvl1_merged();
        const int mmax = std::min(tvl1_max_per_launch(), ty == 16 ? 4 : 6);   // LDS of 64x16 tiles
        const bool narrow = tvl1_narrow();
        int rd = 0, i = 0;
        unsigned long long* xbuf = nullptr;
        bool ran = false;
        if (tvl1_resident_use(width, height, niters, &xbuf)) {
            TRY(tvl1_resident_run(tv, xbuf, 0, niters, (float)(1 + sigma), (float)sigma, S(s), &ran));
            if (ran) {
                rd = 1;
                i = niters;
            }
        }
        if (!ran && niters >= 2 && tvl1_resident_required()) {
            set_error("psw_tvl1_denoise: PSW_TVL1_RESIDENT=2 and the resident solver did not run (%dx%d, %d iterations)",
                      width, height, niters);
            return PSW_ERR_DEVICE;
        }
        for (; i < niters;) {
            const int m = std::min(mmax, niters - i);
            const float sp0 = (float)(i == 0 ? 1 + sigma : sigma);
            TRY(tvl1_launch(tv, ty, merged, narrow, m, rd, sp0, (float)sigma, S(s)));
            rd = 1 - rd;
            i += m;
        }
        if (rd == 1) TRY(psw_memcpy_dtod(d_out, tv.out[1], n * sizeof(float), s));
    } else {
        for (int i = 0; i < niters; i++) {
            const double currsigma = i == 0 ? 1 + sigma : sigma;
            TRY(psw_denoising_TVL1_calculateP_tensor_weighed(Px, Py, T11, T12, T21, T22, d_out, (float)currsigma,
                                                             width, height, z, z, s));
            TRY(psw_denoising_TVL1_update(d_out, R, Px, Py, raw, (float)tau, (float)theta, (float)lambda,
                                          (float)sigma, width, height, z, z, s));
        }
    }
    TRY(psw_element_scale(d_out, zfar - znear, width, height, z, z, s));
    TRY(psw_element_add(d_out, znear, width, height, z, z, s));
#undef TRY
    return PSW_OK;
}
