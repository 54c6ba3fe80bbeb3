// scene.hip -- seeded synthetic room scenes rendered on the GPU.
//
// One thread per pixel ray-casts a room (interior box plus axis-aligned
// boxes) in double precision and shades the hit point with a 4-octave solid
// value-noise texture.  Every operation is written in the same order as the
// numpy generator (python/psw_amd/scene.py, Room.render / texture /
// value_noise / _hash) and the library is built without FMA contraction, so
// depth and the noise-free texture agree with numpy to the last bit except
// where libm-free IEEE arithmetic differs (none on this path).  Pixel noise
// is a counter-based Gaussian (splitmix64 -> Box-Muller) keyed by a seed.
#include "psw_common.h"

namespace psw {

struct SceneCam {
    double invK[9], R[9], t[3];
};

__device__ inline uint32_t tex_hash(int64_t ix, int64_t iy, int64_t iz, uint32_t salt) {
    uint32_t h = ((uint32_t)ix * 0x8DA6B343u) ^ ((uint32_t)iy * 0xD8163841u) ^ ((uint32_t)iz * 0xCB1AB31Fu) ^ salt;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

// Trilinear value noise with smoothstep weights at integer lattice points.
__device__ inline double value_noise(double px, double py, double pz, uint32_t salt) {
    const double fx = floor(px), fy = floor(py), fz = floor(pz);
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    const double tx = px - fx, ty = py - fy, tz = pz - fz;
    const double wx1 = tx * tx * (3.0 - 2.0 * tx), wy1 = ty * ty * (3.0 - 2.0 * ty), wz1 = tz * tz * (3.0 - 2.0 * tz);
    double out = 0.0;
    for (int dx = 0; dx < 2; ++dx) {
        const double wx = dx ? wx1 : 1.0 - wx1;
        for (int dy = 0; dy < 2; ++dy) {
            const double wy = dy ? wy1 : 1.0 - wy1;
            for (int dz = 0; dz < 2; ++dz) {
                const double wz = dz ? wz1 : 1.0 - wz1;
                const uint32_t h = tex_hash(ix + dx, iy + dy, iz + dz, salt);
                out += wx * wy * wz * ((double)(h >> 8) / 16777216.0);
            }
        }
    }
    return out;
}

__device__ inline double texture_at(double x, double y, double z, const psw_room& room) {
    const double freq[4] = {4.0, 9.0, 21.0, 47.0};
    const double amp[4] = {0.40, 0.30, 0.20, 0.10};
    double acc = 0.0, total = 0.0;
    for (int k = 0; k < 4; ++k) {
        const double off = 17.0 * k;
        acc += amp[k] * value_noise(x * freq[k] + off, y * freq[k] + off, z * freq[k] + off, room.octave_salt[k]);
        total += amp[k];
    }
    const double v = acc / total;
    return fmin(fmax((v - 0.5) * 1.8 + 0.5, 0.0), 1.0);
}

__device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void scene_render_kernel(psw_room room, SceneCam cam, int W, int H, double sigma, uint64_t seed,
                                    float scale, float* __restrict__ img, float* __restrict__ depth) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const double* iK = cam.invK;
    const double* R = cam.R;
    const double u = (double)x, v = (double)y;
    const double rc[3] = {iK[0] * u + iK[1] * v + iK[2], iK[3] * u + iK[4] * v + iK[5], iK[6] * u + iK[7] * v + iK[8]};
    double d[3], o[3], inv[3];
    for (int j = 0; j < 3; ++j) {
        d[j] = rc[0] * R[j] + rc[1] * R[3 + j] + rc[2] * R[6 + j];             // R^T applied to the ray
        o[j] = -(R[j] * cam.t[0] + R[3 + j] * cam.t[1] + R[6 + j] * cam.t[2]);  // camera centre -R^T t
        const double dd = fabs(d[j]) < 1e-12 ? 1e-12 : d[j];
        inv[j] = 1.0 / dd;
    }
    double troom = INFINITY;
    for (int j = 0; j < 3; ++j) {
        const double t1 = (room.lo[j] - o[j]) * inv[j], t2 = (room.hi[j] - o[j]) * inv[j];
        troom = fmin(troom, fmax(t1, t2));
    }
    double s = troom;
    for (int b = 0; b < room.nboxes; ++b) {
        double tn = -INFINITY, tf = INFINITY;
        for (int j = 0; j < 3; ++j) {
            const double a = (room.box_lo[b][j] - o[j]) * inv[j], c = (room.box_hi[b][j] - o[j]) * inv[j];
            tn = fmax(tn, fmin(a, c));
            tf = fmin(tf, fmax(a, c));
        }
        if (tn <= tf && tn > 1e-6 && tn < s) s = tn;
    }
    const size_t i = (size_t)y * W + x;
    if (depth) depth[i] = (float)s;
    if (img) {
        double val = texture_at(o[0] + s * d[0], o[1] + s * d[1], o[2] + s * d[2], room);
        if (sigma > 0.0) {
            const uint64_t h1 = splitmix64(seed ^ (2 * (uint64_t)i)), h2 = splitmix64(seed ^ (2 * (uint64_t)i + 1));
            const double u1 = ((double)(h1 >> 11) + 1.0) / 9007199254740993.0;  // (0, 1]
            const double u2 = (double)(h2 >> 11) / 9007199254740992.0;          // [0, 1)
            val += sigma * (sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
        }
        img[i] = (float)val * scale;
    }
}

}  // namespace psw

using namespace psw;

extern "C" int psw_scene_render(const psw_room* room, const double* invK, const double* R, const double* t, int width,
                                int height, double noise_sigma, uint64_t noise_seed, float scale, float* d_img,
                                float* d_depth, psw_stream_t s) {
    PSW_CHECK_ARG(room && invK && R && t, "NULL pointer");
    PSW_CHECK_ARG(width >= 1 && height >= 1, "bad image size");
    PSW_CHECK_ARG(room->nboxes >= 0 && room->nboxes <= PSW_SCENE_MAX_BOXES, "nboxes out of range");
    PSW_CHECK_ARG(!(noise_sigma < 0.0), "noise_sigma must not be negative");
    if (!d_img && !d_depth) return PSW_OK;
    SceneCam cam;
    for (int j = 0; j < 9; ++j) { cam.invK[j] = invK[j]; cam.R[j] = R[j]; }
    for (int j = 0; j < 3; ++j) cam.t[j] = t[j];
    const dim3 blk(64, 4), grd((unsigned)((width + 63) / 64), (unsigned)((height + 3) / 4));
    scene_render_kernel<<<grd, blk, 0, S(s)>>>(*room, cam, width, height, noise_sigma, noise_seed, scale, d_img,
                                               d_depth);
    return launch_status("scene_render_kernel");
}
