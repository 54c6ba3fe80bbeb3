// unet.hip -- the learned TGV2 warm start: a 4-level U-Net in BF16 on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation), with
// hand-written forward, backward and Adam kernels.  No reference counterpart.
//
// Network (3x3 convolutions, ReLU, 2x2 average pooling, nearest upsampling,
// skip connections by channel concatenation; 480,264 parameters):
//   level 0 (W x H)      : e0a = conv(x, 8->32), e0 = conv(e0a, 32->32)
//   level 1 (W/2 x H/2)  : e1a = conv(pool e0, 32->48), e1 = conv(e1a, 48->48)
//   level 2 (W/4 x H/4)  : e2a = conv(pool e1, 48->64), e2 = conv(e2a, 64->64)
//   level 3 (W/8 x H/8)  : b3a = conv(pool e2, 64->96), b3 = conv(b3a, 96->96)
//   level 2              : d2a = conv([e2 | up b3], 160->64), d2 = conv(d2a, 64->64)
//   level 1              : d1a = conv([e1 | up d2], 112->48), d1 = conv(d1a, 48->48)
//   level 0              : d0a = conv([e0 | up d1], 80->32),  d0 = conv(d0a, 32->32)
//   head                 : out = d0 W_h + b_h (1x1, 32->8, FP32)
// Input channels (built from the sweep): ref/255 - 1/2, 1/depth of the argmin
// plane, min cost, second-best cost, argmin validity, 3 zero channels.
// Output channels: relative depth correction, u1 (scaled), p (2), q (3), which
// become the TGV2 warm start (u, u1, p, q).
//
// Layout: activations NHWC bf16 (channels contiguous, a multiple of 8, so one
// 16-byte load feeds one MFMA operand lane); activation gradients NHWC FP32.
// A 3x3 convolution is an implicit GEMM: M = pixels, N = output channels,
// K = 9 * input channels, k = tap * C + c.  Weights are stored FP32 as
// [tap][c_in][c_out] followed by the bias (the flat parameter vector), and
// packed to bf16 in the MFMA B-operand lane order for the forward pass and,
// flipped and transposed, for the data gradient.
#include "psw_common.h"

#include <cmath>
#include <vector>

namespace psw {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

constexpr int UN_LAYERS = 14;
constexpr int UN_IN = 8;     // input channels (5 used)
constexpr int UN_OUT = 8;    // output channels
constexpr int UN_HEAD_IN = 32;

struct LayerDef {
    int ca, cb, cout, level;  // cb: channels of the nearest-upsampled coarse input
};
static const LayerDef kLayers[UN_LAYERS] = {
    {8, 0, 32, 0},  {32, 0, 32, 0}, {32, 0, 48, 1}, {48, 0, 48, 1},  {48, 0, 64, 2},
    {64, 0, 64, 2}, {64, 0, 96, 3}, {96, 0, 96, 3}, {64, 96, 64, 2}, {64, 0, 64, 2},
    {48, 64, 48, 1}, {48, 0, 48, 1}, {32, 48, 32, 0}, {32, 0, 32, 0},
};

// ---- kernels ---------------------------------------------------------------

struct ConvK {
    const uint16_t* xa;   // [B][H][W][ca]
    const uint16_t* xb;   // [B][H/2][W/2][cb] (nearest-upsampled), or null
    const uint16_t* wp;   // packed weights [conv_chunks(ca + cb)][NB][64][8]
    const float* bias;    // [cout] or null
    void* y;              // [B][H][W][cout]: bf16 (relu) or fp32
    int ca, cb, cout, B, H, W, kchunks;
};

// Output tile of one block: 8 rows x 32 columns, four waves of 2 rows each
// (4 MFMA row blocks of 16 pixels per wave).  The input tile with its
// one-pixel halo (10 x 34 pixels) is staged in LDS one 32-channel slab at a
// time; the 9 taps then read their shifted windows from LDS, so every input
// value is fetched from memory once per slab (1.33x for the halo) instead of
// once per tap.  LDS pixel stride 40 bf16 (80 bytes): the 16 lanes of one
// 16-byte read hit 16 disjoint groups of 4 banks.
constexpr int CT_TH = 8, CT_TW = 32;
constexpr int CT_HH = CT_TH + 2, CT_HW = CT_TW + 2;
constexpr int CT_PS = 40;

// K order of the packed weights: for each 32-channel slab of the input, for
// each tap, the slab's 8-channel groups; four groups (32 k values) make one
// MFMA chunk, the last chunk of a slab padded with zero groups.
__host__ __device__ inline int slab_chunks(int slab_channels) { return (9 * (slab_channels >> 3) + 3) >> 2; }
__host__ __device__ inline int conv_chunks(int C) {
    return 9 * (C >> 5) + ((C & 31) ? slab_chunks(C & 31) : 0);
}

// i / d for the slab's 8-channel group count d in 1..4 and 0 <= i < 32768,
// with m = ceil(65536 / d): one 24-bit multiply and a shift instead of an
// integer division (the error i (m d - 65536) / 65536 d stays below 1/3).
__device__ __forceinline__ int div_small(int i, int m) { return (int)(((unsigned)i * (unsigned)m) >> 16); }

// MFMA 16x16x32 lane layout: A row (lane & 15), k (lane >> 4) * 8 + j;
// B column (lane & 15), same k; D rows 4 * (lane >> 4) + i, column (lane & 15).
// One slab's products for the wave's 4 row blocks and the block's NBL output
// blocks, accumulated chunk by chunk in K order; lane group jg of chunk kc
// holds k group 4 kc + jg = (tap, 8-channel group) of the slab's g8 groups.
// B fragments come from LDS (sw, the slab's chunks) or, for WLDS false,
// from memory.  FULL: the slab has 32 channels (the unrolled loop).
template <int NBL, bool WLDS, bool FULL>
__device__ __forceinline__ void conv_slab(f32x4 (&acc)[4][NBL], const uint16_t* sx, const uint16_t* sw,
                                          const uint16_t* wp, int g8, int mg, int nck, int nbtot, int nb0,
                                          int lane, int wave) {
    const int jg = lane >> 4, col = lane & 15;
    const uint16_t* xb = sx + ((wave * 2) * CT_HW + col) * CT_PS;
    auto step = [&](int kc, bool gv, int dy, int dx, int cg) {
        bf16x8 av[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (gv)
                v = *reinterpret_cast<const uint4*>(xb + (((m >> 1) + dy) * CT_HW + (m & 1) * 16 + dx) * CT_PS +
                                                    cg * 8);
            av[m] = __builtin_bit_cast(bf16x8, v);
        }
#pragma unroll
        for (int nb = 0; nb < NBL; ++nb) {
            uint4 w;
            if constexpr (WLDS) w = *reinterpret_cast<const uint4*>(&sw[((kc * NBL + nb) * 64 + lane) * 8]);
            else w = *reinterpret_cast<const uint4*>(wp + ((long)(kc * nbtot + nb0 + nb) * 64 + lane) * 8);
            const bf16x8 bv = __builtin_bit_cast(bf16x8, w);
#pragma unroll
            for (int m = 0; m < 4; ++m)
                acc[m][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv, acc[m][nb], 0, 0, 0);
        }
    };
    if (FULL) {
        // a full slab (g8 = 4): chunk kc is tap kc, lane group jg is channel
        // group jg, and every LDS offset is a constant
#pragma unroll
        for (int kc = 0; kc < 9; ++kc) step(kc, true, kc / 3, kc % 3, jg);
    } else {
        for (int kc = 0; kc < nck; ++kc) {
            const int gi = kc * 4 + jg;  // this lane's 8-channel k group
            const bool gv = gi < 9 * g8;
            const int tap = gv ? div_small(gi, mg) : 0;
            const int cg = gv ? gi - tap * g8 : 0;
            const int dy = div_small(tap, 21846);  // tap / 3
            step(kc, gv, dy, tap - dy * 3, cg);
        }
    }
}

// A block computes NBL of the layer's nbtot 16-channel output blocks
// (starting at blockIdx.y * NBL): small levels split their output channels
// over several blocks so that a launch still has enough of them.  The
// results go through LDS, 16 channels at a time, so that every thread
// stores whole 16-byte runs of one pixel's channels.
template <int NBL, bool RELU_BF16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NBL == 1 ? 4 : 1)))
void conv3x3_kernel(ConvK a) {
    constexpr int NPIX = CT_HH * CT_HW;               // staged pixels (tile + halo)
    constexpr int SX_N = (NPIX * 4 + 255) / 256;      // 16-byte input pieces per thread (4 groups)
    // Blocks of one 16-channel output block (the small levels, where the K
    // loop is long against the tile) also stage the slab's packed weights in
    // LDS (9.2 KB); wider blocks read their B fragments from memory.
    constexpr bool WLDS = NBL == 1;
    constexpr int SW_N = WLDS ? (9 * NBL * 64 + 255) / 256 : 1;  // 16-byte weight pieces per thread
    __shared__ __align__(16) uint16_t sx[NPIX * CT_PS];
    __shared__ __align__(16) uint16_t sw[WLDS ? 9 * NBL * 64 * 8 : 8];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int C = a.ca + a.cb;
    const int nbtot = a.cout >> 4, nb0 = blockIdx.y * NBL;
    const int tiles_x = (a.W + CT_TW - 1) / CT_TW, tiles_y = (a.H + CT_TH - 1) / CT_TH;
    int blk = blockIdx.x;
    const int tx0 = (blk % tiles_x) * CT_TW;
    blk /= tiles_x;
    const int ty0 = (blk % tiles_y) * CT_TH;
    const int b = blk / tiles_y;
    const int Hh = a.H >> 1, Wh = a.W >> 1;
    const int col = lane & 15, jg = lane >> 4;
    f32x4 acc[4][NBL];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int nb = 0; nb < NBL; ++nb) acc[m][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

    // One slab's global reads go to registers first, all of them in flight
    // together, and the next slab's are issued before this slab's matrix
    // products, so that their latency overlaps the work of the current one.
    uint4 rx[SX_N], rw[SW_N];
    auto load_slab = [&](int c0, const uint16_t* wsl) {
        const int g8 = (C - c0 < 32 ? C - c0 : 32) >> 3;  // 8-channel groups in the slab
        const int mg = (65536 + g8 - 1) / g8;
#pragma unroll
        for (int k = 0; k < SX_N; ++k) {
            const int i = t + k * 256;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (i < NPIX * g8) {
                const int pix = div_small(i, mg), g = i - pix * g8;
                const int hy = pix / CT_HW, hx = pix - hy * CT_HW;
                const int yy = ty0 + hy - 1, xx = tx0 + hx - 1;
                if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
                    const int c = c0 + g * 8;
                    if (c < a.ca)
                        v = *reinterpret_cast<const uint4*>(a.xa + (((long)b * a.H + yy) * a.W + xx) * a.ca + c);
                    else
                        v = *reinterpret_cast<const uint4*>(a.xb + (((long)b * Hh + (yy >> 1)) * Wh + (xx >> 1)) *
                                                                       a.cb + (c - a.ca));
                }
            }
            rx[k] = v;
        }
        if constexpr (WLDS) {
            // the slab's chunks kc < nck of output blocks nb0 .. nb0 + NBL - 1
            const int nck = slab_chunks(g8 * 8);
#pragma unroll
            for (int k = 0; k < SW_N; ++k) {
                const int j = t + k * 256;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (j < nck * NBL * 64) {
                    const int kc = j / (NBL * 64), r = j - kc * (NBL * 64);
                    v = *reinterpret_cast<const uint4*>(wsl + ((long)(kc * nbtot + nb0 + (r >> 6)) * 64 + (r & 63)) * 8);
                }
                rw[k] = v;
            }
        }
    };
    const uint16_t* wp = a.wp;   // packed weights of the current slab
    load_slab(0, wp);
    for (int c0 = 0; c0 < C; c0 += 32) {
        const int g8 = (C - c0 < 32 ? C - c0 : 32) >> 3;  // 8-channel groups in this slab
        const int mg = (65536 + g8 - 1) / g8;
        const int nck = slab_chunks(g8 * 8);
        __syncthreads();                                    // the previous slab's reads are done
#pragma unroll
        for (int k = 0; k < SX_N; ++k) {
            const int i = t + k * 256;
            if (i < NPIX * g8) {
                const int pix = div_small(i, mg), g = i - pix * g8;
                *reinterpret_cast<uint4*>(&sx[pix * CT_PS + g * 8]) = rx[k];
            }
        }
        if constexpr (WLDS) {
#pragma unroll
            for (int k = 0; k < SW_N; ++k) {
                const int j = t + k * 256;
                if (j < nck * NBL * 64) *reinterpret_cast<uint4*>(&sw[j * 8]) = rw[k];
            }
        }
        __syncthreads();
        if (c0 + 32 < C) load_slab(c0 + 32, wp + (long)nck * nbtot * 512);
        // Narrow blocks (inference) take the unrolled loop for full slabs;
        // a second MFMA loop gets accumulators of its own, which wide blocks
        // (training) cannot afford, so they keep the general loop only.
        if (NBL <= 2 && g8 == 4) conv_slab<NBL, WLDS, NBL <= 2>(acc, sx, sw, wp, g8, mg, nck, nbtot, nb0, lane, wave);
        else conv_slab<NBL, WLDS, false>(acc, sx, sw, wp, g8, mg, nck, nbtot, nb0, lane, wave);
        wp += (long)nck * nbtot * 512;
    }
    // Epilogue, 16 channels at a time: bias (+ ReLU, bf16) into LDS as
    // [pixel][channel] (row stride 24 bf16 / 20 floats: conflict-free 16-byte
    // reads), then thread t stores tile pixel t's 16 channels.
    const int ty = t / CT_TW, tx = t - ty * CT_TW;
    const int y = ty0 + ty, x = tx0 + tx;
    const long p = ((long)b * a.H + y) * a.W + x;
    const bool out = y < a.H && x < a.W;
#pragma unroll
    for (int nb = 0; nb < NBL; ++nb) {
        const int ch0 = (nb0 + nb) * 16;
        const float bias = a.bias ? a.bias[ch0 + col] : 0.f;
        __syncthreads();                                   // sx is free (K loop or previous block of channels)
        if (RELU_BF16) {
            uint16_t* so = sx;
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int q = (wave * 2 + (m >> 1)) * CT_TW + (m & 1) * 16 + 4 * jg + i;
                    so[q * 24 + col] = f2bf(fmaxf(acc[m][nb][i] + bias, 0.f));
                }
            __syncthreads();
            if (out) {
                const uint4* src = reinterpret_cast<const uint4*>(&so[t * 24]);
                uint4* dst = reinterpret_cast<uint4*>(reinterpret_cast<uint16_t*>(a.y) + p * a.cout + ch0);
                dst[0] = src[0];
                dst[1] = src[1];
            }
        } else {
            float* so = reinterpret_cast<float*>(sx);
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int q = (wave * 2 + (m >> 1)) * CT_TW + (m & 1) * 16 + 4 * jg + i;
                    so[q * 20 + col] = acc[m][nb][i] + bias;
                }
            __syncthreads();
            if (out) {
                const uint4* src = reinterpret_cast<const uint4*>(&so[t * 20]);
                uint4* dst = reinterpret_cast<uint4*>(reinterpret_cast<float*>(a.y) + p * a.cout + ch0);
#pragma unroll
                for (int k = 0; k < 4; ++k) dst[k] = src[k];
            }
        }
    }
}

// Pack FP32 weights [9][cin][cout] into the bf16 B-operand order of
// conv3x3_kernel (slab / tap / 8-channel group, see conv_chunks).  flip = 0:
// K over the cin input channels, N = cout (forward).  flip = 1: K over the
// cout channels of dZ, N = cin, with W'[tap'][co][ci] = W[8 - tap'][ci][co]
// (the data gradient is the forward convolution of dZ with the
// 180-degree-rotated, transposed kernel).
__global__ void pack_w_kernel(const float* __restrict__ w, int cin, int cout, int flip, int kchunks, int NB,
                              uint16_t* __restrict__ wp) {
    const long n_el = (long)kchunks * NB * 512;
    const int C = flip ? cout : cin, NO = flip ? cin : cout;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n_el; idx += (long)gridDim.x * blockDim.x) {
        const int j = (int)(idx & 7), l = (int)((idx >> 3) & 63);
        const long r = idx >> 9;
        const int nb = (int)(r % NB), kc = (int)(r / NB);
        const int slab = kc / 9, local = kc - slab * 9;
        const int c0 = 32 * slab;
        const int g8 = (C - c0 < 32 ? C - c0 : 32) >> 3;
        const int gi = 4 * local + (l >> 4), n = nb * 16 + (l & 15);
        float v = 0.f;
        if (c0 < C && gi < 9 * g8 && n < NO) {
            const int tap = gi / g8, c = c0 + (gi - tap * g8) * 8 + j;
            if (!flip) v = w[((long)tap * cin + c) * cout + n];
            else v = w[((long)(8 - tap) * cin + n) * cout + c];
        }
        wp[idx] = f2bf(v);
    }
}

// dZ = bf16(g * (y > 0)) for a ReLU output y.
__global__ void relu_bwd_kernel(const uint16_t* __restrict__ y, const float* __restrict__ g, uint16_t* __restrict__ dz,
                                long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dz[i] = f2bf(bf2f(y[i]) > 0.f ? g[i] : 0.f);
}

struct WgK {
    const uint16_t* xa;
    const uint16_t* xb;
    const uint16_t* dz;  // [NP][cout]
    float* part;         // [nchunks][Kext][cout]
    int ca, cb, cout, B, H, W, Kext, chunk_px;
};

// Weight (and bias) gradient: dW[k][n] = sum_p Xcol[p][k] dZ[p][n], with
// k = 9 C the bias row (Xcol = 1).  Block: 64 k-rows x cout over one chunk
// of pixels, 32 pixels per step, both operands transposed through LDS so
// that each MFMA lane reads 8 consecutive pixels.  Chunk partials are summed
// in a fixed order by wgrad_reduce_kernel (deterministic).
template <int NB>
__global__ __launch_bounds__(256) void wgrad_kernel(WgK a) {
    __shared__ __align__(16) uint16_t sA[64][40];
    __shared__ __align__(16) uint16_t sB[NB * 16][40];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int C = a.ca + a.cb, Kw = 9 * C;
    const long NP = (long)a.B * a.H * a.W;
    const int ktile = blockIdx.x, chunk = blockIdx.y;
    const long p_begin = (long)chunk * a.chunk_px;
    const long p_end = p_begin + a.chunk_px < NP ? p_begin + a.chunk_px : NP;
    const int Hh = a.H >> 1, Wh = a.W >> 1;
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long p0 = p_begin; p0 < p_end; p0 += 32) {
        {
            const long p = p0 + (t & 31);
            const int k0 = ktile * 64 + (t >> 5) * 8;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (p < p_end) {
                if (k0 < Kw) {
                    const int tap = k0 / C, c0 = k0 - (k0 / C) * C;
                    const int x = (int)(p % a.W);
                    const long tt = p / a.W;
                    const int y = (int)(tt % a.H), b = (int)(tt / a.H);
                    const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
                    if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
                        if (c0 < a.ca)
                            v = *reinterpret_cast<const uint4*>(a.xa + (((long)b * a.H + yy) * a.W + xx) * a.ca + c0);
                        else
                            v = *reinterpret_cast<const uint4*>(
                                a.xb + (((long)b * Hh + (yy >> 1)) * Wh + (xx >> 1)) * a.cb + (c0 - a.ca));
                    }
                } else if (k0 == Kw) {
                    v.x = 0x3f80u;  // bias row: 1.0 in bf16
                }
            }
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int j = 0; j < 8; ++j) sA[(t >> 5) * 8 + j][t & 31] = e[j];
        }
        for (int q = t; q < NB * 2 * 32; q += 256) {
            const int pxl = q & 31, cg = q >> 5;
            const long p = p0 + pxl;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (p < p_end) v = *reinterpret_cast<const uint4*>(a.dz + p * a.cout + cg * 8);
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int j = 0; j < 8; ++j) sB[cg * 8 + j][pxl] = e[j];
        }
        __syncthreads();
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(&sA[wave * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&sB[nb * 16 + (lane & 15)][(lane >> 4) * 8]);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[nb], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = ktile * 64 + wave * 16 + 4 * (lane >> 4) + i;
            if (k < a.Kext) a.part[((long)chunk * a.Kext + k) * a.cout + nb * 16 + (lane & 15)] = acc[nb][i];
        }
}

__global__ void wgrad_reduce_kernel(const float* __restrict__ part, int nchunks, long n, float* __restrict__ out) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < nchunks; ++c) s += part[(long)c * n + i];
        out[i] = s;
    }
}

// Data gradient of a concatenated input [A | up B]: g_a = tmp[:, :ca];
// g_b (coarse) = sum over the 2x2 children of tmp[:, ca:].
__global__ void concat_bwd_kernel(const float* __restrict__ tmp, int ca, int cb, int B, int H, int W,
                                  float* __restrict__ ga, float* __restrict__ gb) {
    const int C = ca + cb;
    const long na = (long)B * H * W * ca;
    const int Hh = H >> 1, Wh = W >> 1;
    const long nb = (long)B * Hh * Wh * cb;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += (long)gridDim.x * blockDim.x) {
        if (i < na) {
            const long p = i / ca;
            const int c = (int)(i - p * ca);
            ga[i] = tmp[p * C + c];
        } else {
            const long j = i - na;
            const long q = j / cb;
            const int c = (int)(j - q * cb);
            const int x = (int)(q % Wh);
            const long tt = q / Wh;
            const int y = (int)(tt % Hh), b = (int)(tt / Hh);
            float s = 0.f;
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const long p = ((long)b * H + 2 * y + dy) * W + 2 * x + dx;
                    s += tmp[p * C + ca + c];
                }
            gb[j] = s;
        }
    }
}

// 2x2 average pooling (bf16 -> bf16, FP32 sum) and its backward (accumulate).
__global__ void pool_fwd_kernel(const uint16_t* __restrict__ x, int C, int B, int H, int W, uint16_t* __restrict__ y) {
    const int Hh = H >> 1, Wh = W >> 1;
    const long n = (long)B * Hh * Wh * C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long q = i / C;
        const int c = (int)(i - q * C);
        const int xx = (int)(q % Wh);
        const long tt = q / Wh;
        const int yy = (int)(tt % Hh), b = (int)(tt / Hh);
        const long p = ((long)b * H + 2 * yy) * W + 2 * xx;
        const float s = (bf2f(x[p * C + c]) + bf2f(x[(p + 1) * C + c])) + (bf2f(x[(p + W) * C + c]) + bf2f(x[(p + W + 1) * C + c]));
        y[i] = f2bf(0.25f * s);
    }
}

__global__ void pool_bwd_kernel(const float* __restrict__ gy, int C, int B, int H, int W, float* __restrict__ gx) {
    const long n = (long)B * H * W * C;
    const int Hh = H >> 1, Wh = W >> 1;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long p = i / C;
        const int c = (int)(i - p * C);
        const int x = (int)(p % W);
        const long tt = p / W;
        const int y = (int)(tt % H), b = (int)(tt / H);
        gx[i] += 0.25f * gy[(((long)b * Hh + (y >> 1)) * Wh + (x >> 1)) * C + c];
    }
}

// Head: out[p][o] = b[o] + sum_c d0[p][c] W[c][o] (FP32).
__global__ void head_fwd_kernel(const uint16_t* __restrict__ x, const float* __restrict__ w, long NP,
                                float* __restrict__ out) {
    __shared__ float sw[UN_HEAD_IN * UN_OUT + UN_OUT];
    for (int i = threadIdx.x; i < UN_HEAD_IN * UN_OUT + UN_OUT; i += blockDim.x) sw[i] = w[i];
    __syncthreads();
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    float o[UN_OUT];
#pragma unroll
    for (int k = 0; k < UN_OUT; ++k) o[k] = sw[UN_HEAD_IN * UN_OUT + k];
    const uint4* xp = reinterpret_cast<const uint4*>(x + p * UN_HEAD_IN);
#pragma unroll
    for (int q = 0; q < UN_HEAD_IN / 8; ++q) {
        const uint4 v = xp[q];
        const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float xv = bf2f(e[j]);
#pragma unroll
            for (int k = 0; k < UN_OUT; ++k) o[k] += xv * sw[(q * 8 + j) * UN_OUT + k];
        }
    }
#pragma unroll
    for (int k = 0; k < UN_OUT; ++k) out[p * UN_OUT + k] = o[k];
}

// Head backward: g_x[p][c] = sum_o g[p][o] W[c][o]; per-chunk partials of
// dW[c][o] = sum_p x[p][c] g[p][o] and db[o] = sum_p g[p][o].
__global__ __launch_bounds__(256) void head_bwd_kernel(const uint16_t* __restrict__ x, const float* __restrict__ g,
                                                       const float* __restrict__ w, long NP, int chunk_px,
                                                       float* __restrict__ gx, float* __restrict__ part) {
    __shared__ float sw[UN_HEAD_IN * UN_OUT];
    __shared__ float sx[64][UN_HEAD_IN + 1];
    __shared__ float sg[64][UN_OUT + 1];
    const int t = threadIdx.x;
    for (int i = t; i < UN_HEAD_IN * UN_OUT; i += 256) sw[i] = w[i];
    const long p_begin = (long)blockIdx.x * chunk_px;
    const long p_end = p_begin + chunk_px < NP ? p_begin + chunk_px : NP;
    const int c = t >> 3, o = t & 7;  // this thread's dW entry
    float accw = 0.f, accb = 0.f;
    __syncthreads();
    for (long p0 = p_begin; p0 < p_end; p0 += 64) {
        // stage 64 pixels of x and g
        for (int i = t; i < 64 * UN_HEAD_IN; i += 256) {
            const long p = p0 + i / UN_HEAD_IN;
            sx[i / UN_HEAD_IN][i % UN_HEAD_IN] = p < p_end ? bf2f(x[p * UN_HEAD_IN + i % UN_HEAD_IN]) : 0.f;
        }
        for (int i = t; i < 64 * UN_OUT; i += 256) {
            const long p = p0 + i / UN_OUT;
            sg[i / UN_OUT][i % UN_OUT] = p < p_end ? g[p * UN_OUT + i % UN_OUT] : 0.f;
        }
        __syncthreads();
        for (int r = 0; r < 64; ++r) accw += sx[r][c] * sg[r][o];
        if (t < UN_OUT)
            for (int r = 0; r < 64; ++r) accb += sg[r][t];
        // g_x for these pixels: thread handles (pixel r, channel pair)
        for (int i = t; i < 64 * UN_HEAD_IN; i += 256) {
            const int r = i / UN_HEAD_IN, cc = i % UN_HEAD_IN;
            const long p = p0 + r;
            if (p < p_end) {
                float s = 0.f;
#pragma unroll
                for (int k = 0; k < UN_OUT; ++k) s += sg[r][k] * sw[cc * UN_OUT + k];
                gx[p * UN_HEAD_IN + cc] = s;
            }
        }
        __syncthreads();
    }
    float* pp = part + (long)blockIdx.x * (UN_HEAD_IN * UN_OUT + UN_OUT);
    pp[c * UN_OUT + o] = accw;
    if (t < UN_OUT) pp[UN_HEAD_IN * UN_OUT + t] = accb;
}

// Smooth-L1 (Huber, delta 1) loss per channel with weights wc; g = dL/dout.
struct LossK {
    const float* out;
    const float* tgt;
    float* g;
    double* part;  // per block
    long NP;
    float wc[UN_OUT];
    float inv_np;
};
__global__ __launch_bounds__(256) void loss_kernel(LossK a) {
    __shared__ double red[4];
    double s = 0.0;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < a.NP; p += (long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < UN_OUT; ++c) {
            const float d = a.out[p * UN_OUT + c] - a.tgt[p * UN_OUT + c];
            const float ad = fabsf(d);
            s += (double)(a.wc[c] * (ad <= 1.f ? 0.5f * d * d : ad - 0.5f));
            a.g[p * UN_OUT + c] = a.wc[c] * fminf(fmaxf(d, -1.f), 1.f) * a.inv_np;
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) a.part[blockIdx.x] = ((red[0] + red[1]) + (red[2] + red[3]));
}

__global__ void adam_kernel(float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long n, float lr_t, float b1, float b2, float eps) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        th[i] -= lr_t * mi / (sqrtf(vi) + eps);
    }
}

// Network input from the aggregated sweep of one view.
__global__ void features_kernel(const float* __restrict__ ref, const int32_t* __restrict__ arg,
                                const float* __restrict__ cmin, const float* __restrict__ csec,
                                const float* __restrict__ depths, int P, float fallback, long n,
                                uint16_t* __restrict__ x, float* __restrict__ ds) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int a = arg[p];
    const bool ok = a >= 0 && a < P;
    const float d = ok ? depths[a] : fallback;
    ds[p] = d;
    uint16_t e[8];
    e[0] = f2bf(ref[p] * (1.f / 255.f) - 0.5f);
    e[1] = f2bf(1.f / d);
    e[2] = f2bf(ok ? cmin[p] : 0.f);
    e[3] = f2bf(ok ? csec[p] : 0.f);
    e[4] = f2bf(ok ? 1.f : 0.f);
    e[5] = e[6] = e[7] = 0;
    uint4 v;
    v.x = e[0] | ((uint32_t)e[1] << 16);
    v.y = e[2] | ((uint32_t)e[3] << 16);
    v.z = e[4] | ((uint32_t)e[5] << 16);
    v.w = e[6] | ((uint32_t)e[7] << 16);
    reinterpret_cast<uint4*>(x)[p] = v;
}

// Output <-> TGV2 state scaling: u = ds (1 + S_U o0), u1 = S_U1 ds (o1, o2),
// p = (o3, o4), q = (o5, o6, o7) projected onto the duals' unit balls.
constexpr float S_U = 0.1f, S_U1 = 0.01f, T_CLAMP = 10.f;

__global__ void targets_kernel(const float* __restrict__ st, const float* __restrict__ ds, long n,
                               float* __restrict__ tgt) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float d = ds[p];
    float t[8];
    t[0] = fminf(fmaxf((st[p] / d - 1.f) / S_U, -T_CLAMP), T_CLAMP);
    t[1] = fminf(fmaxf(st[n + p] / (S_U1 * d), -T_CLAMP), T_CLAMP);
    t[2] = fminf(fmaxf(st[2 * n + p] / (S_U1 * d), -T_CLAMP), T_CLAMP);
    for (int c = 3; c < 8; ++c) t[c] = st[c * n + p];
    for (int c = 0; c < 8; ++c) tgt[p * 8 + c] = t[c];
}

__global__ void warm_start_kernel(const float* __restrict__ out, const float* __restrict__ ds, long n,
                                  float* __restrict__ u, float* __restrict__ u1, float* __restrict__ pp,
                                  float* __restrict__ qq) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float* o = out + p * 8;
    const float d = ds[p];
    if (u) u[p] = d * (1.f + S_U * fminf(fmaxf(o[0], -T_CLAMP), T_CLAMP));
    if (u1) {
        u1[p] = S_U1 * d * fminf(fmaxf(o[1], -T_CLAMP), T_CLAMP);
        u1[n + p] = S_U1 * d * fminf(fmaxf(o[2], -T_CLAMP), T_CLAMP);
    }
    if (pp) {
        const float dn = fmaxf(1.f, sqrtf(o[3] * o[3] + o[4] * o[4]));
        pp[p] = o[3] / dn;
        pp[n + p] = o[4] / dn;
    }
    if (qq) {
        const float dn = fmaxf(1.f, sqrtf(o[5] * o[5] + o[6] * o[6] + 2.f * o[7] * o[7]));
        qq[p] = o[5] / dn;
        qq[n + p] = o[6] / dn;
        qq[2 * n + p] = o[7] / dn;
    }
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ x, long n, uint16_t* __restrict__ y) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

// ---- host side -------------------------------------------------------------

static inline unsigned grid_for(long n, int per = 256) {
    long g = (n + per - 1) / per;
    if (g > 65535L * 16) g = 65535L * 16;
    return (unsigned)(g < 1 ? 1 : g);
}

// Output-channel split: levels with few tiles give each block fewer
// 16-channel output blocks (NBL), so that a launch has about 1000 blocks.
template <bool RELU>
static void launch_conv(const ConvK& a, int NB, hipStream_t st) {
    const long tiles = (long)a.B * ((a.H + CT_TH - 1) / CT_TH) * ((a.W + CT_TW - 1) / CT_TW);
    int nbl = NB;
    if (tiles * 2 < 1024 && NB % 2 == 0 && NB > 2) nbl = 2;
    if (tiles * (NB / nbl) < 1024) nbl = 1;
    const dim3 grd((unsigned)tiles, (unsigned)(NB / nbl));
    switch (nbl) {
#define PSW_CONV_CASE(n) \
    case n: conv3x3_kernel<n, RELU><<<grd, 256, 0, st>>>(a); break;
        PSW_CONV_CASE(1) PSW_CONV_CASE(2) PSW_CONV_CASE(3) PSW_CONV_CASE(4) PSW_CONV_CASE(5) PSW_CONV_CASE(6)
        PSW_CONV_CASE(7) PSW_CONV_CASE(10)
#undef PSW_CONV_CASE
        default: break;
    }
}

static bool nb_supported(int c) {
    return c % 16 == 0 && (c / 16 == 2 || c / 16 == 3 || c / 16 == 4 || c / 16 == 5 || c / 16 == 6 ||
                           c / 16 == 7 || c / 16 == 10);
}

// chunk size so that one level gets at most 256 pixel chunks
static int chunk_px_for(long NP) {
    long c = (NP + 255) / 256;
    c = (c + 31) / 32 * 32;
    return (int)(c < 512 ? 512 : c);
}

static void launch_wgrad(const WgK& a0, int nchunks, float* dwb, hipStream_t st) {
    WgK a = a0;
    const int NB = a.cout / 16;
    const dim3 grd((unsigned)((a.Kext + 63) / 64), (unsigned)nchunks);
    switch (NB) {
        case 2: wgrad_kernel<2><<<grd, 256, 0, st>>>(a); break;
        case 3: wgrad_kernel<3><<<grd, 256, 0, st>>>(a); break;
        case 4: wgrad_kernel<4><<<grd, 256, 0, st>>>(a); break;
        case 6: wgrad_kernel<6><<<grd, 256, 0, st>>>(a); break;
        default: break;
    }
    const long n = (long)a.Kext * a.cout;
    wgrad_reduce_kernel<<<grid_for(n), 256, 0, st>>>(a.part, nchunks, n, dwb);
}

}  // namespace psw

using namespace psw;

struct psw_unet {
    long max_px = 0;
    size_t nparams = 0;
    size_t off[UN_LAYERS + 1];  // start of W (then bias) of each layer; [UN_LAYERS] = head
    float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    uint16_t* wpf[UN_LAYERS];
    uint16_t* wpb[UN_LAYERS];
    uint16_t* act[UN_LAYERS + 1];
    uint16_t* pool[3];
    float* gact[UN_LAYERS + 1];
    float* gpool[3];
    float* tmp = nullptr;
    uint16_t* dz = nullptr;
    float *out = nullptr, *gout = nullptr, *tgt = nullptr, *ds = nullptr;
    float* part = nullptr;
    double* lpart = nullptr;
    std::vector<void*> allocs;
};

static const int kAdamBlocks = 1024;
static const int kLossBlocks = 512;

static size_t unet_param_count_impl(size_t* off) {
    size_t o = 0;
    for (int L = 0; L < UN_LAYERS; ++L) {
        if (off) off[L] = o;
        const LayerDef& d = kLayers[L];
        o += (size_t)(9 * (d.ca + d.cb) + 1) * d.cout;
    }
    if (off) off[UN_LAYERS] = o;
    o += UN_HEAD_IN * UN_OUT + UN_OUT;
    return o;
}

static int unet_alloc(psw_unet* n, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes < 256 ? 256 : bytes);
    if (e != hipSuccess) return hip_status(e, "hipMalloc", PSW_ERR_ALLOC);
    n->allocs.push_back(*p);
    return PSW_OK;
}

// bf16 weight packs of layer L from theta
static int unet_pack(psw_unet* n, int L, hipStream_t st) {
    const LayerDef& d = kLayers[L];
    const int cin = d.ca + d.cb;
    const int kf = conv_chunks(cin), kb = conv_chunks(d.cout);
    pack_w_kernel<<<grid_for((long)kf * (d.cout / 16) * 512), 256, 0, st>>>(n->theta + n->off[L], cin, d.cout, 0, kf,
                                                                            d.cout / 16, n->wpf[L]);
    if (L > 0)
        pack_w_kernel<<<grid_for((long)kb * (cin / 16) * 512), 256, 0, st>>>(n->theta + n->off[L], cin, d.cout, 1,
                                                                             kb, cin / 16, n->wpb[L]);
    return launch_status("pack_w_kernel");
}

static int unet_pack_all(psw_unet* n, hipStream_t st) {
    for (int L = 0; L < UN_LAYERS; ++L) {
        int s = unet_pack(n, L, st);
        if (s != PSW_OK) return s;
    }
    return PSW_OK;
}

static int unet_check_shape(psw_unet* n, int batch, int width, int height) {
    PSW_CHECK_ARG(n, "NULL network");
    PSW_CHECK_ARG(batch >= 1 && width >= 8 && height >= 8, "bad batch or image size");
    PSW_CHECK_ARG(width % 8 == 0 && height % 8 == 0, "width and height must be multiples of 8");
    PSW_CHECK_ARG((long)batch * width * height <= n->max_px, "batch * width * height above the network's capacity");
    return PSW_OK;
}

// level-l input pointers of layer L
static void layer_inputs(psw_unet* n, int L, const uint16_t** xa, const uint16_t** xb) {
    static const int src_a[UN_LAYERS] = {0, 1, -1, 3, -2, 5, -3, 7, 6, 9, 4, 11, 2, 13};
    static const int src_b[UN_LAYERS] = {-9, -9, -9, -9, -9, -9, -9, -9, 8, -9, 10, -9, 12, -9};
    const int a = src_a[L];
    *xa = a >= 0 ? n->act[a] : n->pool[-a - 1];
    *xb = src_b[L] >= 0 ? n->act[src_b[L]] : nullptr;
}

static int unet_forward_impl(psw_unet* n, int B, int W, int H, hipStream_t st) {
    for (int L = 0; L < UN_LAYERS; ++L) {
        const LayerDef& d = kLayers[L];
        ConvK a;
        layer_inputs(n, L, &a.xa, &a.xb);
        a.wp = n->wpf[L];
        a.bias = n->theta + n->off[L] + (size_t)9 * (d.ca + d.cb) * d.cout;
        a.y = n->act[L + 1];
        a.ca = d.ca; a.cb = d.cb; a.cout = d.cout;
        a.B = B; a.H = H >> d.level; a.W = W >> d.level;
        a.kchunks = conv_chunks(d.ca + d.cb);
        launch_conv<true>(a, d.cout / 16, st);
        int s = launch_status("conv3x3_kernel");
        if (s != PSW_OK) return s;
        if (L == 1 || L == 3 || L == 5) {
            const int pl = (L - 1) / 2;
            const long np = (long)B * (H >> (d.level + 1)) * (W >> (d.level + 1)) * d.cout;
            pool_fwd_kernel<<<grid_for(np), 256, 0, st>>>(n->act[L + 1], d.cout, B, H >> d.level, W >> d.level,
                                                          n->pool[pl]);
            s = launch_status("pool_fwd_kernel");
            if (s != PSW_OK) return s;
        }
    }
    const long NP = (long)B * W * H;
    head_fwd_kernel<<<(unsigned)((NP + 255) / 256), 256, 0, st>>>(n->act[UN_LAYERS], n->theta + n->off[UN_LAYERS], NP,
                                                                  n->out);
    return launch_status("head_fwd_kernel");
}

static int unet_backward_impl(psw_unet* n, int B, int W, int H, hipStream_t st) {
    const long NP = (long)B * W * H;
    {
        const int cpx = chunk_px_for(NP);
        const int nch = (int)((NP + cpx - 1) / cpx);
        head_bwd_kernel<<<nch, 256, 0, st>>>(n->act[UN_LAYERS], n->gout, n->theta + n->off[UN_LAYERS], NP, cpx,
                                             n->gact[UN_LAYERS], n->part);
        const long nh = UN_HEAD_IN * UN_OUT + UN_OUT;
        wgrad_reduce_kernel<<<grid_for(nh), 256, 0, st>>>(n->part, nch, nh, n->grad + n->off[UN_LAYERS]);
        int s = launch_status("head_bwd_kernel");
        if (s != PSW_OK) return s;
    }
    for (int L = UN_LAYERS - 1; L >= 0; --L) {
        const LayerDef& d = kLayers[L];
        const int h = H >> d.level, w = W >> d.level;
        const long np = (long)B * h * w;
        relu_bwd_kernel<<<grid_for(np * d.cout), 256, 0, st>>>(n->act[L + 1], n->gact[L + 1], n->dz, np * d.cout);
        WgK g;
        layer_inputs(n, L, &g.xa, &g.xb);
        g.dz = n->dz;
        g.part = n->part;
        g.ca = d.ca; g.cb = d.cb; g.cout = d.cout;
        g.B = B; g.H = h; g.W = w;
        g.Kext = 9 * (d.ca + d.cb) + 1;
        g.chunk_px = chunk_px_for(np);
        const int nch = (int)((np + g.chunk_px - 1) / g.chunk_px);
        launch_wgrad(g, nch, n->grad + n->off[L], st);
        int s = launch_status("wgrad_kernel");
        if (s != PSW_OK) return s;
        if (L == 0) break;
        ConvK a;
        a.xa = n->dz; a.xb = nullptr; a.wp = n->wpb[L]; a.bias = nullptr;
        a.ca = d.cout; a.cb = 0; a.cout = d.ca + d.cb;
        a.B = B; a.H = h; a.W = w;
        a.kchunks = conv_chunks(d.cout);
        static const int dst_a[UN_LAYERS] = {0, 1, -1, 3, -2, 5, -3, 7, 6, 9, 4, 11, 2, 13};
        if (d.cb == 0) {
            const int t = dst_a[L];
            a.y = t >= 0 ? (void*)n->gact[t] : (void*)n->gpool[-t - 1];
            launch_conv<false>(a, a.cout / 16, st);
            if ((s = launch_status("conv3x3_kernel (dgrad)")) != PSW_OK) return s;
            if (t < 0) {
                // pooled input: distribute into the pooled activation's gradient
                const int pl = -t - 1;
                const int src = 2 * pl + 2;  // act index that was pooled (e0, e1, e2)
                const int lv = kLayers[src - 1].level;
                const long nfine = (long)B * (H >> lv) * (W >> lv) * kLayers[src - 1].cout;
                pool_bwd_kernel<<<grid_for(nfine), 256, 0, st>>>(n->gpool[pl], kLayers[src - 1].cout, B, H >> lv,
                                                                 W >> lv, n->gact[src]);
                if ((s = launch_status("pool_bwd_kernel")) != PSW_OK) return s;
            }
        } else {
            a.y = n->tmp;
            launch_conv<false>(a, a.cout / 16, st);
            if ((s = launch_status("conv3x3_kernel (dgrad)")) != PSW_OK) return s;
            static const int dst_b[UN_LAYERS] = {-9, -9, -9, -9, -9, -9, -9, -9, 8, -9, 10, -9, 12, -9};
            const long tot = np * d.ca + (long)B * (h / 2) * (w / 2) * d.cb;
            concat_bwd_kernel<<<grid_for(tot), 256, 0, st>>>(n->tmp, d.ca, d.cb, B, h, w, n->gact[dst_a[L]],
                                                             n->gact[dst_b[L]]);
            if ((s = launch_status("concat_bwd_kernel")) != PSW_OK) return s;
        }
    }
    return PSW_OK;
}

extern "C" {

size_t psw_unet_param_count(void) { return unet_param_count_impl(nullptr); }

int psw_unet_destroy(psw_unet* n) {
    if (!n) return PSW_OK;
    for (void* p : n->allocs) (void)hipFree(p);
    delete n;
    return PSW_OK;
}

int psw_unet_create(psw_unet** out, int64_t max_pixels) {
    PSW_CHECK_ARG(out, "NULL output");
    PSW_CHECK_ARG(max_pixels >= 64, "max_pixels too small");
    psw_unet* n = new psw_unet();
    n->max_px = max_pixels;
    n->nparams = unet_param_count_impl(n->off);
    const size_t P = n->nparams;
    int s = PSW_OK;
    auto A = [&](void** p, size_t bytes) {
        if (s == PSW_OK) s = unet_alloc(n, p, bytes);
    };
    A((void**)&n->theta, P * 4);
    A((void**)&n->grad, P * 4);
    A((void**)&n->m, P * 4);
    A((void**)&n->v, P * 4);
    size_t part_floats = (size_t)257 * (UN_HEAD_IN * UN_OUT + UN_OUT);
    size_t tmp_floats = 0, dz_el = 0;
    for (int L = 0; L < UN_LAYERS; ++L) {
        const LayerDef& d = kLayers[L];
        const int cin = d.ca + d.cb;
        A((void**)&n->wpf[L], (size_t)conv_chunks(cin) * (d.cout / 16) * 512 * 2);
        A((void**)&n->wpb[L], (size_t)conv_chunks(d.cout) * (cin / 16) * 512 * 2);
        const size_t np = (size_t)(max_pixels >> (2 * d.level));
        A((void**)&n->act[L + 1], np * d.cout * 2);
        A((void**)&n->gact[L + 1], np * d.cout * 4);
        part_floats = std::max(part_floats, (size_t)257 * (9 * cin + 1) * d.cout);
        if (d.cb) tmp_floats = std::max(tmp_floats, np * cin);
        dz_el = std::max(dz_el, np * d.cout);
    }
    A((void**)&n->act[0], (size_t)max_pixels * UN_IN * 2);
    n->gact[0] = nullptr;
    for (int pl = 0; pl < 3; ++pl) {
        const int c = kLayers[2 * pl + 1].cout;
        const size_t np = (size_t)(max_pixels >> (2 * (pl + 1)));
        A((void**)&n->pool[pl], np * c * 2);
        A((void**)&n->gpool[pl], np * c * 4);
    }
    A((void**)&n->tmp, tmp_floats * 4);
    A((void**)&n->dz, dz_el * 2);
    A((void**)&n->out, (size_t)max_pixels * UN_OUT * 4);
    A((void**)&n->gout, (size_t)max_pixels * UN_OUT * 4);
    A((void**)&n->tgt, (size_t)max_pixels * UN_OUT * 4);
    A((void**)&n->ds, (size_t)max_pixels * 4);
    A((void**)&n->part, part_floats * 4);
    A((void**)&n->lpart, kLossBlocks * sizeof(double));
    if (s != PSW_OK) {
        psw_unet_destroy(n);
        return s;
    }
    *out = n;
    return psw_unet_init(n, 1);
}

// He-normal weights (std sqrt(2 / fan_in)) from a splitmix64 / Box-Muller
// stream on the host, zero biases, a small head so that the initial output
// is near 0 (warm start = the sweep depth).  Synchronous.
int psw_unet_init(psw_unet* n, uint64_t seed) {
    PSW_CHECK_ARG(n, "NULL network");
    std::vector<float> th(n->nparams, 0.f);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 12345;
    auto next = [&st]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {
        const double u1 = ((double)(next() >> 11) + 1.0) / 9007199254740993.0;
        const double u2 = (double)(next() >> 11) / 9007199254740992.0;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int L = 0; L < UN_LAYERS; ++L) {
        const LayerDef& d = kLayers[L];
        const int fan = 9 * (d.ca + d.cb);
        const double sd = std::sqrt(2.0 / fan);
        for (size_t i = 0; i < (size_t)fan * d.cout; ++i) th[n->off[L] + i] = (float)(sd * normal());
    }
    for (int i = 0; i < UN_HEAD_IN * UN_OUT; ++i) th[n->off[UN_LAYERS] + i] = (float)(1e-3 * normal());
    return psw_unet_set_params(n, th.data(), th.size());
}

int psw_unet_set_params(psw_unet* n, const float* params, size_t count) {
    PSW_CHECK_ARG(n && params, "NULL pointer");
    PSW_CHECK_ARG(count == n->nparams, "parameter count mismatch");
    PSW_TRY_HIP(hipMemcpy(n->theta, params, count * 4, hipMemcpyHostToDevice), PSW_ERR_COPY);
    PSW_TRY_HIP(hipMemset(n->m, 0, count * 4), PSW_ERR_DEVICE);
    PSW_TRY_HIP(hipMemset(n->v, 0, count * 4), PSW_ERR_DEVICE);
    PSW_TRY_HIP(hipMemset(n->grad, 0, count * 4), PSW_ERR_DEVICE);
    int s = unet_pack_all(n, nullptr);
    if (s == PSW_OK) PSW_TRY_HIP(hipDeviceSynchronize(), PSW_ERR_DEVICE);
    return s;
}

int psw_unet_get_params(psw_unet* n, float* params, size_t count) {
    PSW_CHECK_ARG(n && params, "NULL pointer");
    PSW_CHECK_ARG(count == n->nparams, "parameter count mismatch");
    PSW_TRY_HIP(hipDeviceSynchronize(), PSW_ERR_DEVICE);
    PSW_TRY_HIP(hipMemcpy(params, n->theta, count * 4, hipMemcpyDeviceToHost), PSW_ERR_COPY);
    return PSW_OK;
}

int psw_unet_get_grads(psw_unet* n, float* grads, size_t count) {
    PSW_CHECK_ARG(n && grads, "NULL pointer");
    PSW_CHECK_ARG(count == n->nparams, "parameter count mismatch");
    PSW_TRY_HIP(hipDeviceSynchronize(), PSW_ERR_DEVICE);
    PSW_TRY_HIP(hipMemcpy(grads, n->grad, count * 4, hipMemcpyDeviceToHost), PSW_ERR_COPY);
    return PSW_OK;
}

int psw_unet_features(psw_unet* n, int slot, int width, int height, const float* d_ref, const int32_t* d_argmin,
                      const float* d_cmin, const float* d_csec, const float* d_depths, int nplanes, float fallback,
                      psw_stream_t s) {
    int st = unet_check_shape(n, slot + 1, width, height);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(slot >= 0, "negative slot");
    PSW_CHECK_ARG(d_ref && d_argmin && d_cmin && d_csec && d_depths && nplanes >= 1, "bad sweep inputs");
    const long np = (long)width * height;
    features_kernel<<<(unsigned)((np + 255) / 256), 256, 0, S(s)>>>(d_ref, d_argmin, d_cmin, d_csec, d_depths, nplanes,
                                                                    fallback, np, n->act[0] + slot * np * UN_IN,
                                                                    n->ds + slot * np);
    return launch_status("features_kernel");
}

int psw_unet_set_input(psw_unet* n, int slot, int width, int height, const float* d_x, const float* d_ds,
                       psw_stream_t s) {
    int st = unet_check_shape(n, slot + 1, width, height);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(d_x && d_ds, "NULL pointer");
    const long np = (long)width * height;
    f32_to_bf16_kernel<<<grid_for(np * UN_IN), 256, 0, S(s)>>>(d_x, np * UN_IN, n->act[0] + slot * np * UN_IN);
    PSW_TRY_HIP(hipMemcpyAsync(n->ds + slot * np, d_ds, np * 4, hipMemcpyDeviceToDevice, S(s)), PSW_ERR_COPY);
    return launch_status("f32_to_bf16_kernel");
}

int psw_unet_targets(psw_unet* n, int slot, int width, int height, const float* d_state, psw_stream_t s) {
    int st = unet_check_shape(n, slot + 1, width, height);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(d_state, "NULL state");
    const long np = (long)width * height;
    targets_kernel<<<(unsigned)((np + 255) / 256), 256, 0, S(s)>>>(d_state, n->ds + slot * np, np,
                                                                   n->tgt + slot * np * UN_OUT);
    return launch_status("targets_kernel");
}

int psw_unet_set_target(psw_unet* n, int slot, int width, int height, const float* d_t, psw_stream_t s) {
    int st = unet_check_shape(n, slot + 1, width, height);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(d_t, "NULL pointer");
    const long np = (long)width * height;
    PSW_TRY_HIP(hipMemcpyAsync(n->tgt + slot * np * UN_OUT, d_t, np * UN_OUT * 4, hipMemcpyDeviceToDevice, S(s)),
                PSW_ERR_COPY);
    return PSW_OK;
}

int psw_unet_forward(psw_unet* n, int batch, int width, int height, psw_stream_t s) {
    int st = unet_check_shape(n, batch, width, height);
    if (st != PSW_OK) return st;
    return unet_forward_impl(n, batch, width, height, S(s));
}

int psw_unet_output(psw_unet* n, int slot, int width, int height, float* d_out, psw_stream_t s) {
    int st = unet_check_shape(n, slot + 1, width, height);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(d_out, "NULL output");
    const long np = (long)width * height;
    PSW_TRY_HIP(hipMemcpyAsync(d_out, n->out + slot * np * UN_OUT, np * UN_OUT * 4, hipMemcpyDeviceToDevice, S(s)),
                PSW_ERR_COPY);
    return PSW_OK;
}

int psw_unet_warm_start(psw_unet* n, int slot, int width, int height, float* d_u, float* d_u1, float* d_p,
                        float* d_q, psw_stream_t s) {
    int st = unet_check_shape(n, slot + 1, width, height);
    if (st != PSW_OK) return st;
    const long np = (long)width * height;
    warm_start_kernel<<<(unsigned)((np + 255) / 256), 256, 0, S(s)>>>(n->out + slot * np * UN_OUT, n->ds + slot * np,
                                                                      np, d_u, d_u1, d_p, d_q);
    return launch_status("warm_start_kernel");
}

int psw_unet_loss_grad(psw_unet* n, int batch, int width, int height, const float* channel_weights, double* loss,
                       psw_stream_t s) {
    int st = unet_check_shape(n, batch, width, height);
    if (st != PSW_OK) return st;
    PSW_CHECK_ARG(channel_weights, "NULL channel weights");
    hipStream_t hs = S(s);
    if ((st = unet_forward_impl(n, batch, width, height, hs)) != PSW_OK) return st;
    LossK a;
    a.out = n->out; a.tgt = n->tgt; a.g = n->gout; a.part = n->lpart;
    a.NP = (long)batch * width * height;
    for (int c = 0; c < UN_OUT; ++c) a.wc[c] = channel_weights[c];
    a.inv_np = 1.f / (float)a.NP;
    loss_kernel<<<kLossBlocks, 256, 0, hs>>>(a);
    if ((st = launch_status("loss_kernel")) != PSW_OK) return st;
    if ((st = unet_backward_impl(n, batch, width, height, hs)) != PSW_OK) return st;
    if (loss) {
        double h[kLossBlocks];
        PSW_TRY_HIP(hipMemcpyAsync(h, n->lpart, sizeof(h), hipMemcpyDeviceToHost, hs), PSW_ERR_COPY);
        PSW_TRY_HIP(hipStreamSynchronize(hs), PSW_ERR_DEVICE);
        double t = 0.0;
        for (int i = 0; i < kLossBlocks; ++i) t += h[i];
        *loss = t / (double)a.NP;
    }
    return PSW_OK;
}

int psw_unet_adam_step(psw_unet* n, double lr, double beta1, double beta2, double eps, int step, psw_stream_t s) {
    PSW_CHECK_ARG(n, "NULL network");
    PSW_CHECK_ARG(step >= 1 && lr > 0 && beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1, "bad Adam settings");
    const double lr_t = lr * std::sqrt(1.0 - std::pow(beta2, step)) / (1.0 - std::pow(beta1, step));
    adam_kernel<<<kAdamBlocks, 256, 0, S(s)>>>(n->theta, n->grad, n->m, n->v, (long)n->nparams, (float)lr_t,
                                               (float)beta1, (float)beta2, (float)eps);
    int st = launch_status("adam_kernel");
    if (st != PSW_OK) return st;
    return unet_pack_all(n, S(s));
}

int psw_unet_flops(int width, int height, int batch, double* fwd_flops) {
    PSW_CHECK_ARG(fwd_flops && width > 0 && height > 0 && batch > 0, "bad arguments");
    double f = 0.0;
    for (int L = 0; L < UN_LAYERS; ++L) {
        const LayerDef& d = kLayers[L];
        const double np = (double)batch * (width >> d.level) * (height >> d.level);
        f += 2.0 * 9.0 * (d.ca + d.cb) * d.cout * np;
    }
    f += 2.0 * UN_HEAD_IN * UN_OUT * (double)batch * width * height;
    *fwd_flops = f;
    return PSW_OK;
}

// ---- single-layer entry points (tests, tools): temporaries are allocated --

int psw_conv3x3_bf16(const uint16_t* d_xa, int ca, const uint16_t* d_xb, int cb, const float* d_w,
                     const float* d_bias, int cout, int batch, int width, int height, int relu_bf16, void* d_y,
                     psw_stream_t s) {
    PSW_CHECK_ARG(d_xa && d_w && d_y, "NULL pointer");
    PSW_CHECK_ARG(ca > 0 && ca % 8 == 0 && cb >= 0 && cb % 8 == 0 && (cb == 0) == (d_xb == nullptr),
                  "channels must be multiples of 8");
    PSW_CHECK_ARG(nb_supported(cout), "cout must be 16 x {2,3,4,5,6,7,10}");
    PSW_CHECK_ARG(batch >= 1 && width >= 2 && height >= 2 && (cb == 0 || (width % 2 == 0 && height % 2 == 0)),
                  "bad size");
    const int cin = ca + cb, kc = conv_chunks(cin);
    uint16_t* wp = nullptr;
    PSW_TRY_HIP(hipMalloc((void**)&wp, (size_t)kc * (cout / 16) * 512 * 2), PSW_ERR_ALLOC);
    hipStream_t st = S(s);
    pack_w_kernel<<<grid_for((long)kc * (cout / 16) * 512), 256, 0, st>>>(d_w, cin, cout, 0, kc, cout / 16, wp);
    ConvK a;
    a.xa = d_xa; a.xb = d_xb; a.wp = wp; a.bias = d_bias; a.y = d_y;
    a.ca = ca; a.cb = cb; a.cout = cout; a.B = batch; a.H = height; a.W = width; a.kchunks = kc;
    if (relu_bf16) launch_conv<true>(a, cout / 16, st);
    else launch_conv<false>(a, cout / 16, st);
    int r = launch_status("conv3x3_kernel");
    (void)hipStreamSynchronize(st);
    (void)hipFree(wp);
    return r;
}

int psw_conv3x3_bf16_dgrad(const uint16_t* d_dz, int cout, const float* d_w, int cin, int batch, int width,
                           int height, float* d_dx, psw_stream_t s) {
    PSW_CHECK_ARG(d_dz && d_w && d_dx, "NULL pointer");
    PSW_CHECK_ARG(cout % 8 == 0 && cout > 0 && nb_supported(cin), "unsupported channel counts");
    PSW_CHECK_ARG(batch >= 1 && width >= 1 && height >= 1, "bad size");
    const int kc = conv_chunks(cout);
    uint16_t* wp = nullptr;
    PSW_TRY_HIP(hipMalloc((void**)&wp, (size_t)kc * (cin / 16) * 512 * 2), PSW_ERR_ALLOC);
    hipStream_t st = S(s);
    pack_w_kernel<<<grid_for((long)kc * (cin / 16) * 512), 256, 0, st>>>(d_w, cin, cout, 1, kc, cin / 16, wp);
    ConvK a;
    a.xa = d_dz; a.xb = nullptr; a.wp = wp; a.bias = nullptr; a.y = d_dx;
    a.ca = cout; a.cb = 0; a.cout = cin; a.B = batch; a.H = height; a.W = width; a.kchunks = kc;
    launch_conv<false>(a, cin / 16, st);
    int r = launch_status("conv3x3_kernel (dgrad)");
    (void)hipStreamSynchronize(st);
    (void)hipFree(wp);
    return r;
}

int psw_conv3x3_bf16_wgrad(const uint16_t* d_xa, int ca, const uint16_t* d_xb, int cb, const uint16_t* d_dz,
                           int cout, int batch, int width, int height, float* d_dwb, psw_stream_t s) {
    PSW_CHECK_ARG(d_xa && d_dz && d_dwb, "NULL pointer");
    PSW_CHECK_ARG(ca > 0 && ca % 8 == 0 && cb >= 0 && cb % 8 == 0 && (cb == 0) == (d_xb == nullptr),
                  "channels must be multiples of 8");
    PSW_CHECK_ARG(cout == 32 || cout == 48 || cout == 64 || cout == 96, "cout must be 32, 48, 64 or 96");
    PSW_CHECK_ARG(batch >= 1 && width >= 2 && height >= 2 && (cb == 0 || (width % 2 == 0 && height % 2 == 0)),
                  "bad size");
    WgK g;
    g.xa = d_xa; g.xb = d_xb; g.dz = d_dz;
    g.ca = ca; g.cb = cb; g.cout = cout; g.B = batch; g.H = height; g.W = width;
    g.Kext = 9 * (ca + cb) + 1;
    const long np = (long)batch * width * height;
    g.chunk_px = chunk_px_for(np);
    const int nch = (int)((np + g.chunk_px - 1) / g.chunk_px);
    PSW_TRY_HIP(hipMalloc((void**)&g.part, (size_t)nch * g.Kext * cout * 4), PSW_ERR_ALLOC);
    hipStream_t st = S(s);
    launch_wgrad(g, nch, d_dwb, st);
    int r = launch_status("wgrad_kernel");
    (void)hipStreamSynchronize(st);
    (void)hipFree(g.part);
    return r;
}

}  // extern "C"
