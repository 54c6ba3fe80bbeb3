/*
 * psw_oracle.cpp -- CPU oracle for the multiview-stereo path.
 *
 * TEST INFRASTRUCTURE ONLY: a line-by-line CPU restatement of the reference
 * kernels and of their host sequencing, used by tests/, smoke() and the
 * bench's cpu_baseline leg to compare the GPU results against.  Nothing in
 * the product links it.
 *
 * Build: g++ -O2 -fopenmp -ffp-contract=off -fno-fast-math (see oracle/build.sh).
 * Contraction is off so that every + - * / rounds on its own, which is the
 * arithmetic the GPU kernels reproduce where bit-exactness is claimed.
 *
 * Template parameter T is the reference's "float": T = float reproduces the
 * reference (with its explicit double temporaries kept as double), T = double
 * is the FP64 accuracy judge.
 *
 * Semantics pinned here where the reference is undefined:
 *  - device buffers the reference leaves uninitialised (cudaMallocPitch
 *    without memset: bestNCC/depth per source, global depth sum/count, TGV2
 *    p/q/r/prodsum, fusion voxels, TVL1 R/Px/Py) start at zero;
 *  - bilinear sample coordinates that are NaN or beyond int range are
 *    "outside the image" (the reference converts them to int, which is UB).
 */
#include "psw_oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#include <limits>
#include <algorithm>

namespace {

/* ---- small math types with the reference's operation order -------------
 * Matrix3D is three float3 rows (src/inc/structs.h:410-764); products are
 * row-dot-column with left-to-right sums (src/inc/helper_structs.h:159-175,
 * helper_math dot = a.x*b.x + a.y*b.y + a.z*b.z). */
template <class T> struct V3 { T x, y, z; };
template <class T> struct M3 { V3<T> r[3]; };

template <class T> inline V3<T> mk(T x, T y, T z) { V3<T> v; v.x = x; v.y = y; v.z = z; return v; }
template <class T> inline T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> operator+(const V3<T>& a, const V3<T>& b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(const V3<T>& a, const V3<T>& b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator*(T s, const V3<T>& a) { return mk(s * a.x, s * a.y, s * a.z); }
template <class T> inline V3<T> operator/(const V3<T>& a, T s) { return mk(a.x / s, a.y / s, a.z / s); }

template <class T> inline M3<T> load_m(const float* m) {
    M3<T> r;
    for (int i = 0; i < 3; ++i) r.r[i] = mk<T>((T)m[3 * i], (T)m[3 * i + 1], (T)m[3 * i + 2]);
    return r;
}
template <class T> inline V3<T> load_v(const float* v) { return mk<T>((T)v[0], (T)v[1], (T)v[2]); }
template <class T> inline V3<T> mv(const M3<T>& R, const V3<T>& v) { return mk(dot(R.r[0], v), dot(R.r[1], v), dot(R.r[2], v)); }
template <class T> inline M3<T> trans(const M3<T>& m) {
    M3<T> t;
    t.r[0] = mk(m.r[0].x, m.r[1].x, m.r[2].x);
    t.r[1] = mk(m.r[0].y, m.r[1].y, m.r[2].y);
    t.r[2] = mk(m.r[0].z, m.r[1].z, m.r[2].z);
    return t;
}
template <class T> inline M3<T> mm(const M3<T>& A, M3<T> B) {   /* helper_structs.h:136-144 */
    B = trans(B);
    M3<T> r;
    for (int i = 0; i < 3; ++i) r.r[i] = mk(dot(A.r[i], B.r[0]), dot(A.r[i], B.r[1]), dot(A.r[i], B.r[2]));
    return r;
}
template <class T> inline M3<T> madd(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
    for (int i = 0; i < 3; ++i) r.r[i] = a.r[i] + b.r[i];
    return r;
}
template <class T> inline M3<T> mdiv(const M3<T>& a, T s) {
    M3<T> r;
    for (int i = 0; i < 3; ++i) r.r[i] = a.r[i] / s;
    return r;
}
template <class T> inline T det(const M3<T>& m) {                /* structs.h:565-570 */
    const V3<T>* r = m.r;
    return r[0].x * (r[1].y * r[2].z - r[1].z * r[2].y) -
           r[0].y * (r[1].x * r[2].z - r[1].z * r[2].x) +
           r[0].z * (r[1].x * r[2].y - r[1].y * r[2].x);
}
template <class T> inline M3<T> inv(const M3<T>& m) {            /* structs.h:584-593 */
    const V3<T>* r = m.r;
    T d = det(m);
    M3<T> o;
    o.r[0] = mk(r[1].y * r[2].z - r[1].z * r[2].y, r[0].z * r[2].y - r[0].y * r[2].z, r[0].y * r[1].z - r[0].z * r[1].y) / d;
    o.r[1] = mk(r[1].z * r[2].x - r[1].x * r[2].z, r[0].x * r[2].z - r[0].z * r[2].x, r[0].z * r[1].x - r[0].x * r[1].z) / d;
    o.r[2] = mk(r[1].x * r[2].y - r[1].y * r[2].x, r[0].y * r[2].x - r[0].x * r[2].y, r[0].x * r[1].y - r[0].y * r[1].x) / d;
    return o;
}

/* PlaneSweep::RelativeMatrices, src/planesweep.cpp:538-549 */
template <class T>
void relative(const M3<T>& Rref, const V3<T>& tref, const M3<T>& Rsrc, const V3<T>& tsrc, bool alt,
              M3<T>& Rrel, V3<T>& trel) {
    if (!alt) {
        Rrel = mm(Rsrc, inv(Rref));
        trel = tsrc - mv(Rrel, tref);
    } else {
        Rrel = mm(trans(Rsrc), Rref);
        trel = mv(trans(Rsrc), tref - tsrc);
    }
}

/* H = K (Rrel + trel e3^T / d) K^-1, H /= H(2,2): src/planesweep.cpp:195-203 */
template <class T>
M3<T> homography(const float* Kf, const float* Rref, const float* tref, const float* Rsrc,
                 const float* tsrc, bool alt, T d) {
    M3<T> K = load_m<T>(Kf), Kinv = inv(K), Rrel;
    V3<T> trel;
    relative(load_m<T>(Rref), load_v<T>(tref), load_m<T>(Rsrc), load_v<T>(tsrc), alt, Rrel, trel);
    M3<T> tr;
    tr.r[0] = mk<T>(0, 0, 0); tr.r[1] = mk<T>(0, 0, 0); tr.r[2] = trel;
    tr = trans(tr);
    M3<T> H = mm(mm(K, madd(Rrel, mdiv(tr, d))), Kinv);
    T h22 = H.r[2].z;
    return mdiv(H, h22);
}

template <class T> inline T tmax(T a, T b) { return std::fmax(a, b); }
template <class T> inline T tmin(T a, T b) { return std::fmin(a, b); }

/* ---- reference sweep kernels ------------------------------------------ */

/* windowed_mean_column_kernel, src/kernels.cu:212-235 (reflect-101 border) */
template <class T>
void box_col(T* out, const T* in, int win, bool sq, int w, int h) {
    const int n = win / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            T mean = 0;
            for (int i = -n; i <= n; ++i) {
                int k = y + i;
                if (k < 0) k = -k;
                if (k > h - 1) k = 2 * (h - 1) - k;
                T v = in[(size_t)k * w + x];
                if (sq) mean += v * v; else mean += v;
            }
            out[(size_t)y * w + x] = mean / (T)win;
        }
}

/* windowed_mean_row_kernel, src/kernels.cu:187-210 */
template <class T>
void box_row(T* out, const T* in, int win, bool sq, int w, int h) {
    const int n = win / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            T mean = 0;
            for (int i = -n; i <= n; ++i) {
                int k = x + i;
                if (k < 0) k = -k;
                if (k > w - 1) k = 2 * (w - 1) - k;
                T v = in[(size_t)y * w + k];
                if (sq) mean += v * v; else mean += v;
            }
            out[(size_t)y * w + x] = mean / (T)win;
        }
}

/* bilinear_interpolation_kernel_GPU, src/kernels.cu:4-29.  Valid iff
 * 0 <= floor(x) and floor(x)+1 <= M1-1, i.e. 0 <= x < M1-1 (same for y). */
template <class T>
inline T bilinear_at(const T* d, int M1, int M2, T x, T y) {
    if (!(x >= (T)0 && x < (T)(M1 - 1) && y >= (T)0 && y < (T)(M2 - 1))) return (T)0;
    const int ix = (int)std::floor(x);
    const T a = x - (T)ix;
    const int iy = (int)std::floor(y);
    const T b = y - (T)iy;
    T t1 = a * d[(size_t)iy * M1 + ix + 1] + ((T)1 - a) * d[(size_t)iy * M1 + ix];
    T t2 = a * d[(size_t)(iy + 1) * M1 + ix + 1] + ((T)1 - a) * d[(size_t)(iy + 1) * M1 + ix];
    return b * t2 + ((T)1 - b) * t1;
}

/* transform_indexes_kernel, src/kernels.cu:31-44, with the pixel-coordinate
 * offset made explicit (1 in the reference). */
template <class T>
inline void warp_coord(const M3<T>& H, int l, int k, int off, T& xo, T& yo) {
    V3<T> p = mk<T>((T)(l + off), (T)(k + off), (T)1);
    V3<T> x = mv(H, p);
    xo = x.x / x.z - (T)off;
    yo = x.y / x.z - (T)off;
}

/* calculate_STD_kernel, src/kernels.cu:103-120 */
template <class T> inline T stdev(T mean, T msq) {
    T var = msq - mean * mean;
    return var > (T)0 ? (T)std::sqrt(var) : (T)0;
}

/* calcNCC_kernel, src/kernels.cu:46-64 */
template <class T> inline T ncc(T pm, T m1, T m2, T s1, T s2, T th1, T th2) {
    if (s1 < th1 || s2 < th2) return (T)0;
    return (pm - m1 * m2) / (s1 * s2);
}

/* Window statistics of one warped source at one plane: the nine launches of
 * src/planesweep.cpp:206-243 producing the per-pixel NCC. */
template <class T>
struct SweepScratch {
    std::vector<T> warped, inter1, mw, sw, msq, prod, pm, tmp;
    explicit SweepScratch(size_t n) : warped(n), inter1(n), mw(n), sw(n), msq(n), prod(n), pm(n), tmp(n) {}
};

template <class T>
void plane_ncc(const M3<T>& Hm, const T* src, const T* ref, const T* refmean, const T* refstd,
               int W, int Hh, int win, T stdthresh, int off, SweepScratch<T>& s, T* nccout) {
    const size_t n = (size_t)W * Hh;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Hh; ++k)
        for (int l = 0; l < W; ++l) {
            T xo, yo;
            warp_coord(Hm, l, k, off, xo, yo);
            s.warped[(size_t)k * W + l] = bilinear_at(src, W, Hh, xo, yo);
        }
    box_col(s.inter1.data(), s.warped.data(), win, false, W, Hh);
    box_row(s.mw.data(), s.inter1.data(), win, false, W, Hh);
    box_col(s.tmp.data(), s.warped.data(), win, true, W, Hh);
    box_row(s.msq.data(), s.tmp.data(), win, false, W, Hh);
    for (size_t i = 0; i < n; ++i) s.sw[i] = stdev(s.mw[i], s.msq[i]);
    for (size_t i = 0; i < n; ++i) s.prod[i] = ref[i] * s.warped[i];     /* element_multiply */
    box_col(s.tmp.data(), s.prod.data(), win, false, W, Hh);
    box_row(s.pm.data(), s.tmp.data(), win, false, W, Hh);
    for (size_t i = 0; i < n; ++i)
        nccout[i] = ncc(s.pm[i], refmean[i], s.mw[i], refstd[i], s.sw[i], stdthresh, stdthresh);
}

template <class T>
void ref_stats(const T* ref, int W, int H, int win, std::vector<T>& mean, std::vector<T>& sd) {
    const size_t n = (size_t)W * H;
    std::vector<T> tmp(n), msq(n);
    mean.assign(n, 0); sd.assign(n, 0);
    box_col(tmp.data(), ref, win, false, W, H);          /* src/planesweep.cpp:108-119 */
    box_row(mean.data(), tmp.data(), win, false, W, H);
    box_col(tmp.data(), ref, win, true, W, H);
    box_row(msq.data(), tmp.data(), win, false, W, H);
    for (size_t i = 0; i < n; ++i) sd[i] = stdev(mean[i], msq[i]);
}

inline bool window_ok(int W, int H, int win) {
    return win >= 1 && (win % 2) == 1 && W > win / 2 && H > win / 2 && W >= 2 && H >= 2;
}

template <class T>
int sweep_ref(const float* reff, const float* srcf, int W, int H, int N, const float* K,
              const float* Rref, const float* tref, const float* Rsrc, const float* tsrc, int alt,
              const float* planes, int P, int win, float stdthresh, float nccthresh, float zfar,
              T* depth_out, T* best_ncc_out, T* best_depth_out) {
    if (!window_ok(W, H, win) || N < 1 || P < 1) return 1;
    const size_t n = (size_t)W * H;
    std::vector<T> ref(reff, reff + n), src(n), refmean, refstd, nccv(n), best(n), dep(n);
    std::vector<T> gsum(n, 0), gcount(n, 0);
    ref_stats(ref.data(), W, H, win, refmean, refstd);
    SweepScratch<T> s(n);
    for (int i = 0; i < N; ++i) {                         /* src/planesweep.cpp:125-127 */
        for (size_t j = 0; j < n; ++j) src[j] = (T)srcf[(size_t)i * n + j];
        std::fill(best.begin(), best.end(), (T)0);
        std::fill(dep.begin(), dep.end(), (T)0);
        for (int p = 0; p < P; ++p) {
            const T d = (T)planes[p];
            M3<T> Hm = homography<T>(K, Rref, tref, Rsrc + 9 * i, tsrc + 3 * i, alt != 0, d);
            plane_ncc(Hm, src.data(), ref.data(), refmean.data(), refstd.data(), W, H, win, (T)stdthresh, 1, s,
                      nccv.data());
            for (size_t j = 0; j < n; ++j)                 /* update_arrays_kernel, kernels.cu:66-82 */
                if (nccv[j] > best[j]) { best[j] = nccv[j]; dep[j] = d; }
        }
        for (size_t j = 0; j < n; ++j)                     /* sum_depthmap_NCC_kernel, kernels.cu:84-101 */
            if (best[j] > (T)nccthresh) { gsum[j] += dep[j]; gcount[j] += (T)1; }
        if (best_ncc_out) std::copy(best.begin(), best.end(), best_ncc_out + (size_t)i * n);
        if (best_depth_out) std::copy(dep.begin(), dep.end(), best_depth_out + (size_t)i * n);
    }
    for (size_t j = 0; j < n; ++j) {                        /* element_rdivide + set_QNAN_value */
        T v = gcount[j] != (T)0 ? gsum[j] / gcount[j] : std::numeric_limits<T>::quiet_NaN();
        if (v != v) v = (T)zfar;
        depth_out[j] = v;
    }
    return 0;
}

/* Running (argmin, minimum, second-smallest) of the aggregated costs in
 * plane order, strict '<' so that the first minimum wins. */
template <class T>
inline void reduce_best2(T c, int p, int32_t& arg, T& cmin, T& csec) {
    if (c < cmin) { csec = cmin; cmin = c; arg = p; }
    else if (c < csec) csec = c;
}

template <class T>
int sweep_agg(const float* reff, const float* srcf, int W, int H, int N, const float* K,
              const float* Rref, const float* tref, const float* Rsrc, const float* tsrc, int alt,
              const float* planes, int P, int win, float stdthresh, int off,
              int32_t* argmin, T* cmin, T* csec, T* volume = nullptr) {
    if (!window_ok(W, H, win) || N < 1 || P < 1) return 1;
    const size_t n = (size_t)W * H;
    std::vector<T> ref(reff, reff + n), refmean, refstd, nccv(n), acc(n);
    std::vector<std::vector<T>> src(N, std::vector<T>(n));
    for (int i = 0; i < N; ++i)
        for (size_t j = 0; j < n; ++j) src[i][j] = (T)srcf[(size_t)i * n + j];
    ref_stats(ref.data(), W, H, win, refmean, refstd);
    SweepScratch<T> s(n);
    const T inf = std::numeric_limits<T>::infinity();
    for (size_t j = 0; j < n; ++j) { argmin[j] = -1; cmin[j] = inf; csec[j] = inf; }
    for (int p = 0; p < P; ++p) {
        std::fill(acc.begin(), acc.end(), (T)0);
        for (int i = 0; i < N; ++i) {
            M3<double> Hd = homography<double>(K, Rref, tref, Rsrc + 9 * i, tsrc + 3 * i, alt != 0, (double)planes[p]);
            M3<T> Hm;
            for (int r = 0; r < 3; ++r) Hm.r[r] = mk<T>((T)Hd.r[r].x, (T)Hd.r[r].y, (T)Hd.r[r].z);
            plane_ncc(Hm, src[i].data(), ref.data(), refmean.data(), refstd.data(), W, H, win, (T)stdthresh, off, s,
                      nccv.data());
            for (size_t j = 0; j < n; ++j) acc[j] += nccv[j];
        }
        for (size_t j = 0; j < n; ++j) {
            T c = -acc[j] / (T)N;
            if (volume) volume[(size_t)p * n + j] = c;
            reduce_best2(c, p, argmin[j], cmin[j], csec[j]);
        }
    }
    return 0;
}

/* ---- TGV2 --------------------------------------------------------------- */

template <class T>
struct TgvState {
    int W, H, N;
    size_t n;
    std::vector<T> Ref, u, u0, u1x, u1y, ubar, u1xbar, u1ybar, Px, Py, Qx, Qy, Qz, Qw, prodsum;
    std::vector<T> T1, T2, T3, T4;
    std::vector<std::vector<T>> Src, It, Iu, r;
    TgvState(int W_, int H_, int N_) : W(W_), H(H_), N(N_), n((size_t)W_ * H_) {
        for (auto* v : {&Ref, &u, &u0, &u1x, &u1y, &ubar, &u1xbar, &u1ybar, &Px, &Py, &Qx, &Qy, &Qz, &Qw,
                        &prodsum, &T1, &T2, &T3, &T4})
            v->assign(n, (T)0);
        Src.assign(N, std::vector<T>(n)); It.assign(N, std::vector<T>(n, 0));
        Iu.assign(N, std::vector<T>(n, 0)); r.assign(N, std::vector<T>(n, 0));
    }
};

template <class T> inline T tpowf(T a, T b) { return std::pow(a, b); }
template <class T> inline T texpf(T a) { return std::exp(a); }

/* Anisotropic_diffusion_tensor_kernel, src/TGV2_kernels.cu:257-298 */
template <class T>
void tensor(const T* img, int W, int H, T beta, T gamma, T* T11, T* T12, T* T21, T* T22) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int xx = 0; xx < W; ++xx) {
            const size_t i = (size_t)y * W + xx;
            const int xn = std::min(xx + 1, W - 1), yn = std::min(y + 1, H - 1);
            T x = img[(size_t)y * W + xn] - img[i];
            T yv = img[(size_t)yn * W + xx] - img[i];
            T d = (T)std::sqrt(x * x + yv * yv);
            if (d > (T)0) {
                x = x / d;
                yv = yv / d;
                T k = texpf<T>(-beta * tpowf<T>(d, gamma));
                T11[i] = k * x * x + yv * yv;
                T12[i] = (k - (T)1) * x * yv;
                T21[i] = (k - (T)1) * x * yv;
                T22[i] = k * yv * yv + x * x;
            } else {
                T11[i] = 1; T12[i] = 0; T21[i] = 0; T22[i] = 1;
            }
        }
}

/* One warp's linearisation of source j at u0, src/planesweep.cpp:460-483,
 * kernels TGV2_kernels.cu:141-255 and kernels.cu:4-29. */
template <class T>
void linearise(TgvState<T>& S, int j, const M3<T>& K, const M3<T>& invK, const M3<T>& Rrel, const V3<T>& trel,
               T fx, T fy) {
    const int W = S.W, H = S.H;
    std::vector<T> xs(S.n), ys(S.n), X(S.n), Y(S.n), Z(S.n), dX(S.n), dY(S.n), dZ(S.n), dfx(S.n), dfy(S.n), Iw(S.n);
    const M3<T> RK = mm(Rrel, invK);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            V3<T> a = mk<T>((T)(x + 1), (T)(y + 1), (T)1);
            V3<T> x1 = S.u0[i] * mv(invK, a);                 /* transform_coordinates :156-177 */
            V3<T> x2 = mv(Rrel, x1) + trel;
            X[i] = x2.x; Y[i] = x2.y; Z[i] = x2.z;
            x1 = mv(K, x2);
            xs[i] = x1.x / x1.z - (T)1;
            ys[i] = x1.y / x1.z - (T)1;
            V3<T> dd = mv(RK, a);                             /* coordinate_derivatives :206-217 */
            dX[i] = dd.x; dY[i] = dd.y; dZ[i] = dd.z;
            dfx[i] = fx * (dX[i] * Z[i] - X[i] * dZ[i]) / (Z[i] * Z[i]);   /* derivativeF :233-234 */
            dfy[i] = fy * (dY[i] * Z[i] - Y[i] * dZ[i]) / (Z[i] * Z[i]);
        }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            Iw[i] = bilinear_at(S.Src[j].data(), W, H, xs[i], ys[i]);
        }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const int xn = std::min(x + 1, W - 1), yn = std::min(y + 1, H - 1);
            double ddx = (double)(Iw[(size_t)y * W + xn] - Iw[i]);    /* calculate_Iu :248-253 */
            double ddy = (double)(Iw[(size_t)yn * W + x] - Iw[i]);
            S.Iu[j][i] = (T)(ddx * (double)dfx[i] + ddy * (double)dfy[i]);
            S.It[j][i] = Iw[i] - S.Ref[i];                      /* subtract_kernel */
            S.r[j][i] = 0;
        }
}

/* TGV2_updateP_tensor_weighed_kernel, TGV2_kernels.cu:340-365: p from the
 * forward differences of ub minus (u1xb, u1yb), projected to the unit ball. */
template <class T>
void step_p_tensor(T* Px, T* Py, const T* T1, const T* T2, const T* T3, const T* T4, const T* ub, const T* u1xb,
                   const T* u1yb, T alpha1, T sigma, int W, int H) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int xx = 0; xx < W; ++xx) {
            const size_t i = (size_t)y * W + xx;
            const int xn = std::min(xx + 1, W - 1), yn = std::min(y + 1, H - 1);
            double x = (double)(ub[(size_t)y * W + xn] - ub[i] - u1xb[i]);
            double yv = (double)(ub[(size_t)yn * W + xx] - ub[i] - u1yb[i]);
            const T as = alpha1 * sigma;
            double dx = (double)Px[i] + (double)as * ((double)T1[i] * x + (double)T2[i] * yv);
            double dy = (double)Py[i] + (double)as * ((double)T3[i] * x + (double)T4[i] * yv);
            double d = (double)tmax<T>((T)1, (T)std::sqrt(dx * dx + dy * dy));
            Px[i] = (T)(dx / d);
            Py[i] = (T)(dy / d);
        }
}

/* TGV2_updateQ_kernel, TGV2_kernels.cu:54-85 */
template <class T>
void step_q(T* Qx, T* Qy, T* Qz, T* Qw, const T* u1xb, const T* u1yb, T alpha0, T sigma, int W, int H) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int xx = 0; xx < W; ++xx) {
            const size_t i = (size_t)y * W + xx;
            const int xn = std::min(xx + 1, W - 1), yn = std::min(y + 1, H - 1);
            T dx_u1x = u1xb[(size_t)y * W + xn] - u1xb[i];
            T dy_u1x = u1xb[(size_t)yn * W + xx] - u1xb[i];
            T dx_u1y = u1yb[(size_t)y * W + xn] - u1yb[i];
            T dy_u1y = u1yb[(size_t)yn * W + xx] - u1yb[i];
            const T as = alpha0 * sigma;
            double dx = (double)(Qx[i] + as * dx_u1x);
            double dy = (double)(Qy[i] + as * dy_u1y);
            double dz = (double)(Qz[i] + as * (dy_u1x + dx_u1y) / (T)2);
            double dw = (double)(Qw[i] + as * (dy_u1x + dx_u1y) / (T)2);
            double d = (double)tmax<T>((T)1, (T)std::sqrt(dx * dx + dy * dy + dz * dz + dw * dw));
            Qx[i] = (T)(dx / d);
            Qy[i] = (T)(dy / d);
            Qz[i] = (T)(dz / d);
            Qw[i] = (T)(dw / d);
        }
}

/* One primal-dual iteration, src/planesweep.cpp:485-509. */
template <class T>
void tgv_iteration(TgvState<T>& S, T lambda, T alpha0, T alpha1, T tau, T sigma) {
    const int W = S.W, H = S.H;
    step_p_tensor(S.Px.data(), S.Py.data(), S.T1.data(), S.T2.data(), S.T3.data(), S.T4.data(), S.ubar.data(),
                  S.u1xbar.data(), S.u1ybar.data(), alpha1, sigma, W, H);
    step_q(S.Qx.data(), S.Qy.data(), S.Qz.data(), S.Qw.data(), S.u1xbar.data(), S.u1ybar.data(), alpha0, sigma, W, H);
    /* prodsum := 0, then TGV2_updateR_kernel per source, TGV2_kernels.cu:87-105 */
    const T sl = sigma * lambda;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < S.n; ++i) {
        T ps = 0;
        for (int j = 0; j < S.N; ++j) {
            T rr = S.r[j][i] + sl * (S.It[j][i] + (S.u[i] - S.u0[i]) * S.Iu[j][i]);
            rr = rr / tmax<T>((T)1, (T)std::fabs(rr));
            S.r[j][i] = rr;
            ps += rr * S.Iu[j][i];
        }
        S.prodsum[i] = ps;
    }
    /* TGV2_updateU_tensor_weighed_kernel, TGV2_kernels.cu:300-338 */
    std::vector<T> un(S.n), u1xn(S.n), u1yn(S.n);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const int xp = std::max(x - 1, 0), yp = std::max(y - 1, 0);
            const size_t ixp = (size_t)y * W + xp, iyp = (size_t)yp * W + x;
            const T c_px = S.Px[i], c_py = S.Py[i], xp_px = S.Px[ixp], yp_px = S.Px[iyp], xp_py = S.Py[ixp],
                    yp_py = S.Py[iyp];
            T u = S.u[i] - tau * (-alpha1 * (S.T1[i] * (c_px - xp_px) + S.T2[i] * (c_py - xp_py) +
                                             S.T3[i] * (c_px - yp_px) + S.T4[i] * (c_py - yp_py)) +
                                  lambda * S.prodsum[i]);
            T u1x = S.u1x[i] - tau * (-alpha1 * (S.T1[i] * c_px + S.T2[i] * c_py) -
                                      alpha0 * (S.Qx[i] - S.Qx[ixp] + S.Qz[i] - S.Qz[iyp]));
            T u1y = S.u1y[i] - tau * (-alpha1 * (S.T3[i] * c_px + S.T4[i] * c_py) -
                                      alpha0 * (S.Qz[i] - S.Qz[ixp] + S.Qy[i] - S.Qy[iyp]));
            un[i] = u; u1xn[i] = u1x; u1yn[i] = u1y;
        }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < S.n; ++i) {
        S.ubar[i] = (T)2 * un[i] - S.u[i];
        S.u1xbar[i] = (T)2 * u1xn[i] - S.u1x[i];
        S.u1ybar[i] = (T)2 * u1yn[i] - S.u1y[i];
        S.u[i] = un[i]; S.u1x[i] = u1xn[i]; S.u1y[i] = u1yn[i];
    }
}

/* Iterations from a caller-given state, inside a given linearisation: the
 * loop body of src/planesweep.cpp:485-509 without the warp reset of
 * :442-483.  Planes of W*H in `state` (in and out): u, u1x, u1y, ubar,
 * u1xbar, u1ybar, px, py, qx, qy, qz, qw, r[N]; Tt = T11, T12, T21, T22. */
template <class T>
int tgv_iterate(int W, int H, int N, const T* Tt, const T* u0, const T* It, const T* Iu, T* state, int niters,
                double lambda, double alpha0, double alpha1, double tau, double sigma) {
    if (W < 2 || H < 2 || N < 1 || niters < 0) return 1;
    TgvState<T> S(W, H, N);
    const size_t n = S.n;
    std::vector<T>* planes[] = {&S.u, &S.u1x, &S.u1y, &S.ubar, &S.u1xbar, &S.u1ybar, &S.Px, &S.Py, &S.Qx, &S.Qy,
                                &S.Qz, &S.Qw};
    const T* in = state;
    for (auto* v : planes) { std::copy(in, in + n, v->begin()); in += n; }
    for (int j = 0; j < N; ++j) { std::copy(in, in + n, S.r[j].begin()); in += n; }
    std::copy(u0, u0 + n, S.u0.begin());
    for (auto* v : {&S.T1, &S.T2, &S.T3, &S.T4}) { std::copy(Tt, Tt + n, v->begin()); Tt += n; }
    for (int j = 0; j < N; ++j) {
        std::copy(It + (size_t)j * n, It + (size_t)(j + 1) * n, S.It[j].begin());
        std::copy(Iu + (size_t)j * n, Iu + (size_t)(j + 1) * n, S.Iu[j].begin());
    }
    for (int it = 0; it < niters; ++it) tgv_iteration(S, (T)lambda, (T)alpha0, (T)alpha1, (T)tau, (T)sigma);
    T* o = state;
    for (auto* v : planes) { std::copy(v->begin(), v->end(), o); o += n; }
    for (int j = 0; j < N; ++j) { std::copy(S.r[j].begin(), S.r[j].end(), o); o += n; }
    return 0;
}

template <class T>
void tgv_setup(TgvState<T>& S, const float* ref, const float* src, const float* Kf, const float* Rref,
               const float* tref, const float* Rsrc, const float* tsrc, int alt, T beta, T gamma,
               std::vector<M3<T>>& Rrel, std::vector<V3<T>>& trel) {
    const T scale = (T)1 / (T)255;                          /* element_scale(Ref, 1/255.f) */
    for (size_t i = 0; i < S.n; ++i) S.Ref[i] = (T)ref[i] * scale;
    tensor(S.Ref.data(), S.W, S.H, beta, gamma, S.T1.data(), S.T2.data(), S.T3.data(), S.T4.data());
    Rrel.resize(S.N); trel.resize(S.N);
    for (int j = 0; j < S.N; ++j) {
        for (size_t i = 0; i < S.n; ++i) S.Src[j][i] = (T)src[(size_t)j * S.n + i] * scale;
        relative(load_m<T>(Rref), load_v<T>(tref), load_m<T>(Rsrc + 9 * j), load_v<T>(tsrc + 3 * j), alt != 0,
                 Rrel[j], trel[j]);
    }
}

template <class T>
void tgv_warp_reset(TgvState<T>& S, const float* Kf, const std::vector<M3<T>>& Rrel, const std::vector<V3<T>>& trel) {
    /* src/planesweep.cpp:445-483 */
    S.u0 = S.u; S.ubar = S.u;
    for (auto* v : {&S.Px, &S.Py, &S.Qx, &S.Qy, &S.Qz, &S.Qw, &S.u1x, &S.u1y, &S.u1xbar, &S.u1ybar})
        std::fill(v->begin(), v->end(), (T)0);
    M3<T> K = load_m<T>(Kf), invK = inv(K);
    /* fx, fy: double in the host code, float at the launcher (:426, :469) */
    const T fx = (T)(double)Kf[0], fy = (T)(double)Kf[4];
    for (int j = 0; j < S.N; ++j) linearise(S, j, K, invK, Rrel[j], trel[j], fx, fy);
}

template <class T>
int tgv(const float* ref, const float* src, int W, int H, int N, const float* Kf, const float* Rref,
        const float* tref, const float* Rsrc, const float* tsrc, int alt, int niters, int warps, double lambda,
        double alpha0, double alpha1, double tau, double sigma, double beta, double gamma, const float* u_init,
        T* u_out, T* state_out) {
    if (W < 2 || H < 2 || N < 1 || niters < 0 || warps < 0) return 1;
    TgvState<T> S(W, H, N);
    std::vector<M3<T>> Rrel;
    std::vector<V3<T>> trel;
    tgv_setup(S, ref, src, Kf, Rref, tref, Rsrc, tsrc, alt, (T)beta, (T)gamma, Rrel, trel);
    for (size_t i = 0; i < S.n; ++i) S.u[i] = u_init ? (T)u_init[i] : (T)1;   /* set_value(u, 1) */
    S.ubar = S.u;
    for (int l = 0; l < warps; ++l) {
        tgv_warp_reset(S, Kf, Rrel, trel);
        for (int it = 0; it < niters; ++it)
            tgv_iteration(S, (T)lambda, (T)alpha0, (T)alpha1, (T)tau, (T)sigma);
    }
    std::copy(S.u.begin(), S.u.end(), u_out);
    if (state_out) {
        T* o = state_out;
        for (auto* v : {&S.u, &S.u1x, &S.u1y, &S.Px, &S.Py, &S.Qx, &S.Qy, &S.Qz}) {
            std::copy(v->begin(), v->end(), o);
            o += S.n;
        }
        for (auto* vv : {&S.It, &S.Iu, &S.r})
            for (int j = 0; j < N; ++j) {
                std::copy((*vv)[j].begin(), (*vv)[j].end(), o);
                o += S.n;
            }
    }
    return 0;
}

/* TGV2_updateU_sparseDepthTensor_kernel, TGV2_kernels.cu:399-432: the
 * primal step with the sparse-depth data term (prox of (w/2)(u - Ds)^2) and
 * the over-relaxation of (u, u1) into the bar arrays.  Per pixel in program
 * order, so that aliased u1 / u1bar pointers behave as in the reference's
 * driver (the bar write then reads the fresh u1). */
template <class T>
void step_u_sparse_tensor(T* u, T* u1x, T* u1y, T* ubar, T* u1xbar, T* u1ybar, const T* T1, const T* T2,
                          const T* T3, const T* T4, const T* Px, const T* Py, const T* Qx, const T* Qy, const T* Qz,
                          const T* w, const T* Ds, T alpha0, T alpha1, T tau, T theta, int W, int H) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const int xp = std::max(x - 1, 0), yp = std::max(y - 1, 0);
            const size_t ixp = (size_t)y * W + xp, iyp = (size_t)yp * W + x;
            const T c_px = Px[i], c_py = Py[i], xp_px = Px[ixp], yp_px = Px[iyp], xp_py = Py[ixp], yp_py = Py[iyp];
            u[i] = (u[i] + tau * (alpha1 * (T1[i] * (c_px - xp_px) + T2[i] * (c_py - xp_py) + T3[i] * (c_px - yp_px) +
                                            T4[i] * (c_py - yp_py)) +
                                  w[i] * Ds[i])) /
                   ((T)1 + tau * w[i]);
            u1x[i] = u1x[i] - tau * (-alpha1 * (T1[i] * c_px + T2[i] * c_py) - alpha0 * (Qx[i] - Qx[ixp] + Qz[i] - Qz[iyp]));
            u1y[i] = u1y[i] - tau * (-alpha1 * (T3[i] * c_px + T4[i] * c_py) - alpha0 * (Qz[i] - Qz[ixp] + Qy[i] - Qy[iyp]));
            ubar[i] = u[i] + theta * (u[i] - ubar[i]);
            u1xbar[i] = u1x[i] + theta * (u1x[i] - u1xbar[i]);
            u1ybar[i] = u1y[i] + theta * (u1y[i] - u1ybar[i]);
        }
}

/* PlaneSweep::TGVdenoiseFromSparse, src/planesweep.cpp:596-674.
 * weights = (sparse > 0) (calculateWeights_sparseDepth, TGV2_kernels.cu:
 * 434-445); Ds = sparse / zfar and u = ubar = dense / zfar (element_scale by
 * 1.f/zfar); the reference image / 255 gives the tensor; then per iteration
 * P (on u and u1), Q (on u1) and the sparse U step, whose u1 and u1bar
 * arguments are the same arrays (vxbar, vybar); output u * zfar.  p, q and
 * u1 start at zero.  ubar is written but never read, so theta does not
 * change the result. */
template <class T>
int tgv_sparse(const float* ref, const float* dense, const float* sparse, int W, int H, float zfar, int niters,
               double alpha0, double alpha1, double tau, double sigma, double theta, double beta, double gamma, T* out,
               T* u1_out) {
    if (W < 2 || H < 2 || niters < 0) return 1;
    const size_t n = (size_t)W * H;
    std::vector<T> rf(n), u(n), ubar(n), vx(n, 0), vy(n, 0), Px(n, 0), Py(n, 0), Qx(n, 0), Qy(n, 0), Qz(n, 0),
        Qw(n, 0), w(n), Ds(n), T1(n), T2(n), T3(n), T4(n);
    const T sz = (T)1 / (T)zfar, s255 = (T)1 / (T)255;
    for (size_t i = 0; i < n; ++i) {
        Ds[i] = (T)sparse[i];
        w[i] = Ds[i] > (T)0 ? (T)1 : (T)0;
        Ds[i] = Ds[i] * sz;
        ubar[i] = (T)dense[i] * sz;
        u[i] = ubar[i];
        rf[i] = (T)ref[i] * s255;
    }
    tensor(rf.data(), W, H, (T)beta, (T)gamma, T1.data(), T2.data(), T3.data(), T4.data());
    const T a0 = (T)alpha0, a1 = (T)alpha1, tf = (T)tau, sf = (T)sigma, th = (T)theta;
    for (int it = 0; it < niters; ++it) {
        step_p_tensor(Px.data(), Py.data(), T1.data(), T2.data(), T3.data(), T4.data(), u.data(), vx.data(), vy.data(),
                      a1, sf, W, H);
        step_q(Qx.data(), Qy.data(), Qz.data(), Qw.data(), vx.data(), vy.data(), a0, sf, W, H);
        step_u_sparse_tensor(u.data(), vx.data(), vy.data(), ubar.data(), vx.data(), vy.data(), T1.data(), T2.data(),
                             T3.data(), T4.data(), Px.data(), Py.data(), Qx.data(), Qy.data(), Qz.data(), w.data(),
                             Ds.data(), a0, a1, tf, th, W, H);
    }
    const T zf = (T)zfar;
    for (size_t i = 0; i < n; ++i) out[i] = u[i] * zf;
    if (u1_out) {
        std::copy(vx.begin(), vx.end(), u1_out);
        std::copy(vy.begin(), vy.end(), u1_out + n);
    }
    return 0;
}

/* compute3D_kernel, src/kernels.cu:405-429, as PlaneSweep::get3Dcoordinates
 * (src/planesweep.cpp:551-586) drives it: (Rr, t) = RelativeMatrices(Rref,
 * tref, I, 0), X = Rr * ((z * invK) * (x+1, y+1, 1)) + t for z < 5.5, else
 * z := -9 with x, y left as they were (zero here). */
template <class T>
void compute3d(const float* depth, int W, int H, const float* Kf, const float* Rref, const float* tref, int alt, T* X,
               T* Y, T* Z) {
    M3<T> Rr, I;
    V3<T> t;
    for (int r = 0; r < 3; ++r) I.r[r] = mk<T>(r == 0, r == 1, r == 2);
    relative(load_m<T>(Rref), load_v<T>(tref), I, mk<T>(0, 0, 0), alt != 0, Rr, t);
    const M3<T> invK = inv(load_m<T>(Kf));
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const T z = (T)depth[i];
            X[i] = 0; Y[i] = 0;
            if ((double)z < 5.5) {
                M3<T> zK;
                for (int r = 0; r < 3; ++r) zK.r[r] = z * invK.r[r];
                const V3<T> a = mv(Rr, mv(zK, mk<T>((T)(x + 1), (T)(y + 1), (T)1))) + t;
                X[i] = a.x; Y[i] = a.y; Z[i] = a.z;
            } else {
                Z[i] = (T)-9;
            }
        }
}

template <class S>
double energy(const S* u, const S* u1x, const S* u1y, const S* u0, const S* Tt, const S* It, const S* Iu, int W,
              int H, int N, double lambda, double alpha0, double alpha1, const float* weight = nullptr) {
    const size_t n = (size_t)W * H;
    const S *T11 = Tt, *T12 = Tt + n, *T21 = Tt + 2 * n, *T22 = Tt + 3 * n;
    double e1 = 0, e0 = 0, ed = 0;
#pragma omp parallel for reduction(+ : e1, e0, ed) schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            if (weight && weight[i] == 0.f) continue;     // pixel left out of the energy
            const size_t ixn = (size_t)y * W + std::min(x + 1, W - 1), iyn = (size_t)std::min(y + 1, H - 1) * W + x;
            double gx = ((double)u[ixn] - (double)u[i]) - (double)u1x[i];
            double gy = ((double)u[iyn] - (double)u[i]) - (double)u1y[i];
            double a = (double)T11[i] * gx + (double)T12[i] * gy;
            double b = (double)T21[i] * gx + (double)T22[i] * gy;
            e1 += std::sqrt(a * a + b * b);
            double ex = (double)u1x[ixn] - (double)u1x[i];
            double ey = (double)u1y[iyn] - (double)u1y[i];
            double ez = 0.5 * (((double)u1x[iyn] - (double)u1x[i]) + ((double)u1y[ixn] - (double)u1y[i]));
            e0 += std::sqrt(ex * ex + ey * ey + 2.0 * ez * ez);
            double du = (double)u[i] - (double)u0[i];
            for (int j = 0; j < N; ++j) ed += std::fabs((double)It[(size_t)j * n + i] + du * (double)Iu[(size_t)j * n + i]);
        }
    return alpha1 * e1 + alpha0 * e0 + lambda * ed;
}

/* ---- fusion -------------------------------------------------------------- */

constexpr int NB = 8;

inline void bin_centers(int bins, float* c) {       /* fusionData::binParams, fusion.h:1041-1054 */
    c[0] = -1.f;
    c[bins - 1] = 1.f;
    if (bins > 3) {
        for (int i = 1; i < bins - 1; ++i) c[i] = 2.f * float(i - 1) / float(bins - 3) - 1.f;
    } else if (bins == 3) {
        c[1] = 0.f;
    }
}

/* FusionUpdateHistogram_kernel (fusion.cu:5-43) + fusionData::updateHist
 * (fusion.h:883-907) + worldCoords (fusion.h:360-364) + bilinterp
 * (dev_functions.h:33-38). */
template <class T>
void fusion_hist(uint16_t* hist, int gw, int gh, int gd, const float* vol, const float* depth, int W, int H,
                 const float* Kf, const float* Rf, const float* Tf, float thresholdf) {
    const M3<T> K = load_m<T>(Kf), R = load_m<T>(Rf);
    const V3<T> Tv = load_v<T>(Tf);
    const V3<T> a = mk<T>((T)vol[0], (T)vol[1], (T)vol[2]), b = mk<T>((T)vol[3], (T)vol[4], (T)vol[5]);
    const T thresh = (T)thresholdf;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < gd; ++z)
        for (int y = 0; y < gh; ++y)
            for (int x = 0; x < gw; ++x) {
                const V3<T> size = b - a;
                V3<T> f = mk<T>((T)((x + .5) / (double)gw), (T)((y + .5) / (double)gh), (T)((z + .5) / (double)gd));
                V3<T> c = a + mk(size.x * f.x, size.y * f.y, size.z * f.z);
                c = mv(R, c) + Tv;
                c = mv(K, c);
                const T pxx = c.x / c.z, pxy = c.y / c.z;
                if ((pxx < (T)0) || (pxx > (T)(W - 1)) || (pxy < (T)0) || (pxy > (T)(H - 1))) continue;
                const int x0 = (int)tmax<T>(std::floor(pxx), (T)0), y0 = (int)tmax<T>(std::floor(pxy), (T)0);
                const int x1 = (int)tmin<T>((T)(x0 + 1), (T)(W - 1)), y1 = (int)tmin<T>((T)(y0 + 1), (T)(H - 1));
                const T fxr = pxx - std::floor(pxx), fyr = pxy - std::floor(pxy);
                const T a0 = (T)depth[x0 + (size_t)y0 * W], a1 = (T)depth[x1 + (size_t)y0 * W];
                const T b0 = (T)depth[x0 + (size_t)y1 * W], b1 = (T)depth[x1 + (size_t)y1 * W];
                const T l0 = a0 + fyr * (b0 - a0), l1 = a1 + fyr * (b1 - a1);   /* lerp(y0, y1, frac.y) */
                const T dep = l0 + fxr * (l1 - l0);                              /* lerp(x.x, x.y, frac.x) */
                uint16_t* h = hist + ((size_t)z * gh * gw + (size_t)y * gw + x) * NB;
                const T sd = dep - c.z;
                if (sd >= thresh) { h[NB - 1]++; continue; }
                if (sd <= -thresh) { h[0]++; continue; }
                const T q = tmin<T>((T)std::round((sd + thresh) / ((T)2 * thresh) * (T)(NB - 3) + (T)1), (T)(NB - 2));
                h[(unsigned char)q]++;
            }
}

/* sortedHist<8> (structs.h:212-284): insertion that drops values below the
 * current minimum, median = element[8]. */
template <class T>
struct SortedHist {
    T e[2 * NB + 1];
    int count;
    explicit SortedHist(const T* centers) : count(NB) { for (int i = 0; i < NB; ++i) e[i] = centers[i]; }
    void insert(T val) {
        const int size = 2 * NB + 1;
        if (count != 0) {
            for (int i = count - 1; i >= 0; --i) {
                int next = std::min(i + 1, size - 1);
                if (val < e[i]) e[next] = e[i];
                else { e[next] = val; break; }
            }
        } else {
            e[0] = val;
        }
        count = std::min(count + 1, size);
    }
    T median() const { return e[NB]; }
};

inline int Wi(const uint16_t* h, int i) {            /* fusion.h:813-820 */
    int r = 0;
    for (int j = 1; j <= i; ++j) r -= h[j - 1];
    for (int j = i + 1; j <= NB; ++j) r += h[j - 1];
    return r;
}

template <class T>
T prox_hist(T u, const uint16_t* h, T tau, T lambda, const T* centers) {   /* fusion.h:851-858 */
    SortedHist<T> s(centers);
    s.insert(u);
    for (int j = 1; j <= NB; ++j) s.insert(u + tau * lambda * (T)Wi(h, j));
    return s.median();
}

/* FusionUpdateU_kernel, fusion.cu:45-57 (divPBwd fusion.h:793-802).  The
 * reference computes the step in double and passes float to proxHist. */
template <class T>
void fusion_u(T* u, T* v, const T* p, const uint16_t* hist, int gw, int gh, int gd, double tau, double lambda) {
    float cf[NB];
    bin_centers(NB, cf);
    T centers[NB];
    for (int i = 0; i < NB; ++i) centers[i] = (T)cf[i];
    const size_t V = (size_t)gw * gh * gd;
    const T *px = p, *py = p + V, *pz = p + 2 * V;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < gd; ++z)
        for (int y = 0; y < gh; ++y)
            for (int x = 0; x < gw; ++x) {
                const size_t i = (size_t)z * gh * gw + (size_t)y * gw + x;
                T div = px[i] + py[i] + pz[i];
                if (x > 0) div -= px[i - 1];
                if (y > 0) div -= py[i - gw];
                if (z > 0) div -= pz[i - (size_t)gw * gh];
                const double un = (double)u[i];
                const double uu = un - tau * (double)(-div);
                const T unew = prox_hist<T>((T)uu, hist + i * NB, (T)tau, (T)lambda, centers);
                u[i] = unew;
                v[i] = (T)((double)((T)2 * unew) - un);
            }
}

/* FusionUpdateP_kernel, fusion.cu:59-70 (gradVFwd fusion.h:773-781,
 * projectUnitBall fusion.h:866-870). */
template <class T>
void fusion_p(T* p, const T* v, int gw, int gh, int gd, double sigma) {
    const size_t V = (size_t)gw * gh * gd;
    T *px = p, *py = p + V, *pz = p + 2 * V;
    const T s = (T)sigma;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < gd; ++z)
        for (int y = 0; y < gh; ++y)
            for (int x = 0; x < gw; ++x) {
                const size_t i = (size_t)z * gh * gw + (size_t)y * gw + x;
                const T vv = v[i];
                T gx = 0, gy = 0, gz = 0;
                if (x < gw - 1) gx = v[i + 1] - vv;
                if (y < gh - 1) gy = v[i + gw] - vv;
                if (z < gd - 1) gz = v[i + (size_t)gw * gh] - vv;
                T ax = px[i] + s * gx, ay = py[i] + s * gy, az = pz[i] + s * gz;
                T d = tmax<T>((T)1, (T)std::sqrt(ax * ax + ay * ay + az * az));
                px[i] = ax / d; py[i] = ay / d; pz[i] = az / d;
            }
}

/* ---- TVL1 (PlaneSweep::CudaDenoise, planesweep.cpp:297-379) -------------- */

template <class T>
int tvl1(const float* depth, const float* ref, int W, int H, float znear, float zfar, int niters, double lambda,
         double tau, double sigma, double theta, double beta, double gamma, T* out) {
    if (W < 2 || H < 2 || niters < 0) return 1;
    const size_t n = (size_t)W * H;
    std::vector<T> R(n, 0), Px(n, 0), Py(n, 0), raw(n), u(n), rf(n), T11(n), T12(n), T21(n), T22(n);
    const T s255 = (T)1 / (T)255;
    for (size_t i = 0; i < n; ++i) { rf[i] = (T)ref[i] * s255; u[i] = (T)depth[i]; raw[i] = (T)depth[i]; }
    tensor(rf.data(), W, H, (T)beta, (T)gamma, T11.data(), T12.data(), T21.data(), T22.data());
    const T mzn = -(T)znear;
    for (size_t i = 0; i < n; ++i) { u[i] += mzn; raw[i] += mzn; }          /* element_add */
    const double xscale = (double)((T)1 / ((T)zfar - (T)znear));
    const double inputscale = -sigma / (double)((T)zfar - (T)znear);
    for (size_t i = 0; i < n; ++i) { u[i] = u[i] * (T)xscale; raw[i] = raw[i] * (T)inputscale; }
    const T tauf = (T)tau, thetaf = (T)theta, lambdaf = (T)lambda, sigmaf = (T)sigma;
    for (int it = 0; it < niters; ++it) {
        const T cs = (T)(it == 0 ? 1 + sigma : sigma);
        /* denoising_TVL1_calculateP_tensor_weighed_kernel, kernels.cu:283-307 */
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int xx = 0; xx < W; ++xx) {
                const size_t i = (size_t)y * W + xx;
                const int xn = std::min(xx + 1, W - 1), yn = std::min(y + 1, H - 1);
                double x = (double)(u[(size_t)y * W + xn] - u[i]);
                double yv = (double)(u[(size_t)yn * W + xx] - u[i]);
                double dx = (double)Px[i] + (double)cs * ((double)T11[i] * x + (double)T12[i] * yv);
                double dy = (double)Py[i] + (double)cs * ((double)T21[i] * x + (double)T22[i] * yv);
                double d = (double)tmax<T>((T)1, (T)std::sqrt(dx * dx + dy * dy));
                Px[i] = (T)(dx / d);
                Py[i] = (T)(dy / d);
            }
        /* denoising_TVL1_update_kernel, kernels.cu:342-370 */
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t i = (size_t)y * W + x;
                const int yp = std::max(y - 1, 0);
                T r = R[i];
                r += raw[i];
                r += sigmaf * u[i];
                if (r > lambdaf) r = lambdaf;
                if (r < -lambdaf) r = -lambdaf;
                R[i] = r;
                double x_new;
                if (x == 0) x_new = (double)(u[i] + tauf * (Py[i] - Py[(size_t)yp * W + x]) - tauf * r);
                else x_new = (double)(u[i] + tauf * (Px[i] - Px[i - 1] + Py[i] - Py[(size_t)yp * W + x]) - tauf * r);
                u[i] = (T)(x_new + (double)thetaf * (x_new - (double)u[i]));
            }
    }
    const T sc = (T)zfar - (T)znear;
    for (size_t i = 0; i < n; ++i) { u[i] = u[i] * sc; u[i] += (T)znear; out[i] = u[i]; }
    return 0;
}

}  // namespace

/* ======================= C interface ==================================== */

extern "C" {

int orc_reference_planes(float znear, float zfar, unsigned nplanes, float* out, int maxn) {
    float dstep = (zfar - znear) / (float)(nplanes - 1);      /* planesweep.cpp:169 */
    int n = 0;
    for (float d = znear; d <= zfar; d += dstep) {             /* planesweep.cpp:200 */
        if (n >= maxn) break;
        out[n++] = d;
    }
    return n;
}

void orc_mat3_inv_f32(const float* m, float* out) {
    M3<float> r = inv(load_m<float>(m));
    for (int i = 0; i < 3; ++i) { out[3 * i] = r.r[i].x; out[3 * i + 1] = r.r[i].y; out[3 * i + 2] = r.r[i].z; }
}

void orc_relative_matrices_f32(const float* Rref, const float* tref, const float* Rsrc, const float* tsrc,
                               int alternative, float* Rrel, float* trel) {
    M3<float> R;
    V3<float> t;
    relative(load_m<float>(Rref), load_v<float>(tref), load_m<float>(Rsrc), load_v<float>(tsrc), alternative != 0, R, t);
    for (int i = 0; i < 3; ++i) { Rrel[3 * i] = R.r[i].x; Rrel[3 * i + 1] = R.r[i].y; Rrel[3 * i + 2] = R.r[i].z; }
    trel[0] = t.x; trel[1] = t.y; trel[2] = t.z;
}

void orc_homography_ref_f32(const float* K, const float* Rref, const float* tref, const float* Rsrc,
                            const float* tsrc, int alternative, float depth, float* H) {
    M3<float> h = homography<float>(K, Rref, tref, Rsrc, tsrc, alternative != 0, depth);
    for (int i = 0; i < 3; ++i) { H[3 * i] = h.r[i].x; H[3 * i + 1] = h.r[i].y; H[3 * i + 2] = h.r[i].z; }
}

void orc_homography_f64(const float* K, const float* Rref, const float* tref, const float* Rsrc,
                        const float* tsrc, int alternative, float depth, double* H) {
    M3<double> h = homography<double>(K, Rref, tref, Rsrc, tsrc, alternative != 0, (double)depth);
    for (int i = 0; i < 3; ++i) { H[3 * i] = h.r[i].x; H[3 * i + 1] = h.r[i].y; H[3 * i + 2] = h.r[i].z; }
}

void orc_windowed_mean_row_f32(float* out, const float* in, int winsize, int squared, int w, int h) {
    box_row<float>(out, in, winsize, squared != 0, w, h);
}
void orc_windowed_mean_col_f32(float* out, const float* in, int winsize, int squared, int w, int h) {
    box_col<float>(out, in, winsize, squared != 0, w, h);
}
void orc_bilinear_f32(float* out, const float* data, const float* xs, const float* ys, int M1, int M2, int N1,
                      int N2) {
    for (int k = 0; k < N2; ++k)
        for (int l = 0; l < N1; ++l)
            out[(size_t)k * N1 + l] = bilinear_at<float>(data, M1, M2, xs[(size_t)k * N1 + l], ys[(size_t)k * N1 + l]);
}
void orc_transform_indexes_f32(float* xs, float* ys, const float* H, int w, int h) {
    M3<float> Hm = load_m<float>(H);
    for (int k = 0; k < h; ++k)
        for (int l = 0; l < w; ++l) warp_coord(Hm, l, k, 1, xs[(size_t)k * w + l], ys[(size_t)k * w + l]);
}

int orc_sweep_ref_f32(const float* ref, const float* src, int W, int H, int N, const float* K, const float* Rref,
                      const float* tref, const float* Rsrc, const float* tsrc, int alternative, const float* planes,
                      int P, int winsize, float stdthresh, float nccthresh, float zfar, float* depth_out,
                      float* best_ncc, float* best_depth) {
    return sweep_ref<float>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, planes, P, winsize, stdthresh,
                            nccthresh, zfar, depth_out, best_ncc, best_depth);
}
int orc_sweep_ref_f64(const float* ref, const float* src, int W, int H, int N, const float* K, const float* Rref,
                      const float* tref, const float* Rsrc, const float* tsrc, int alternative, const float* planes,
                      int P, int winsize, float stdthresh, float nccthresh, float zfar, double* depth_out,
                      double* best_ncc, double* best_depth) {
    return sweep_ref<double>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, planes, P, winsize, stdthresh,
                             nccthresh, zfar, depth_out, best_ncc, best_depth);
}
int orc_sweep_agg_f32(const float* ref, const float* src, int W, int H, int N, const float* K, const float* Rref,
                      const float* tref, const float* Rsrc, const float* tsrc, int alternative, const float* planes,
                      int P, int winsize, float stdthresh, int coord_offset, int32_t* argmin, float* cost_min,
                      float* cost_second) {
    return sweep_agg<float>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, planes, P, winsize, stdthresh,
                            coord_offset, argmin, cost_min, cost_second);
}
int orc_sweep_agg_f64(const float* ref, const float* src, int W, int H, int N, const float* K, const float* Rref,
                      const float* tref, const float* Rsrc, const float* tsrc, int alternative, const float* planes,
                      int P, int winsize, float stdthresh, int coord_offset, int32_t* argmin, double* cost_min,
                      double* cost_second) {
    return sweep_agg<double>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, planes, P, winsize, stdthresh,
                             coord_offset, argmin, cost_min, cost_second);
}

int orc_sweep_agg_volume_f64(const float* ref, const float* src, int W, int H, int N, const float* K,
                             const float* Rref, const float* tref, const float* Rsrc, const float* tsrc,
                             int alternative, const float* planes, int P, int winsize, float stdthresh,
                             int coord_offset, int32_t* argmin, double* cost_min, double* cost_second,
                             double* volume) {
    return sweep_agg<double>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, planes, P, winsize, stdthresh,
                             coord_offset, argmin, cost_min, cost_second, volume);
}

int orc_sweep_agg_volume_f32(const float* ref, const float* src, int W, int H, int N, const float* K,
                             const float* Rref, const float* tref, const float* Rsrc, const float* tsrc,
                             int alternative, const float* planes, int P, int winsize, float stdthresh,
                             int coord_offset, int32_t* argmin, float* cost_min, float* cost_second, float* volume) {
    return sweep_agg<float>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, planes, P, winsize, stdthresh,
                            coord_offset, argmin, cost_min, cost_second, volume);
}

void orc_reduce_best2_f64(const double* costs, int P, int32_t* argmin, double* cost_min, double* cost_second) {
    const double inf = std::numeric_limits<double>::infinity();
    *argmin = -1; *cost_min = inf; *cost_second = inf;
    for (int p = 0; p < P; ++p) reduce_best2(costs[p], p, *argmin, *cost_min, *cost_second);
}

int orc_tgv_f32(const float* ref, const float* src, int W, int H, int N, const float* K, const float* Rref,
                const float* tref, const float* Rsrc, const float* tsrc, int alternative, int niters, int warps,
                double lambda, double alpha0, double alpha1, double tau, double sigma, double beta, double gamma,
                const float* u_init, float* u_out, float* state_out) {
    return tgv<float>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, niters, warps, lambda, alpha0, alpha1,
                      tau, sigma, beta, gamma, u_init, u_out, state_out);
}
int orc_tgv_f64(const float* ref, const float* src, int W, int H, int N, const float* K, const float* Rref,
                const float* tref, const float* Rsrc, const float* tsrc, int alternative, int niters, int warps,
                double lambda, double alpha0, double alpha1, double tau, double sigma, double beta, double gamma,
                const float* u_init, double* u_out, double* state_out) {
    return tgv<double>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, niters, warps, lambda, alpha0,
                       alpha1, tau, sigma, beta, gamma, u_init, u_out, state_out);
}

void orc_tensor_f32(const float* img, int W, int H, float beta, float gamma, float* T11, float* T12, float* T21,
                    float* T22) {
    tensor<float>(img, W, H, beta, gamma, T11, T12, T21, T22);
}

}  // extern "C"

template <class T>
static int tgv_linearise(const float* ref, const float* src, int W, int H, int N, const float* K, const float* Rref,
                         const float* tref, const float* Rsrc, const float* tsrc, int alternative, double beta,
                         double gamma, const float* u0, T* It, T* Iu, T* Tout) {
    if (W < 2 || H < 2 || N < 1) return 1;
    TgvState<T> S(W, H, N);
    std::vector<M3<T>> Rrel;
    std::vector<V3<T>> trel;
    tgv_setup(S, ref, src, K, Rref, tref, Rsrc, tsrc, alternative, (T)beta, (T)gamma, Rrel, trel);
    for (size_t i = 0; i < S.n; ++i) S.u[i] = (T)u0[i];
    tgv_warp_reset(S, K, Rrel, trel);
    for (int j = 0; j < N; ++j) {
        std::copy(S.It[j].begin(), S.It[j].end(), It + (size_t)j * S.n);
        std::copy(S.Iu[j].begin(), S.Iu[j].end(), Iu + (size_t)j * S.n);
    }
    std::copy(S.T1.begin(), S.T1.end(), Tout);
    std::copy(S.T2.begin(), S.T2.end(), Tout + S.n);
    std::copy(S.T3.begin(), S.T3.end(), Tout + 2 * S.n);
    std::copy(S.T4.begin(), S.T4.end(), Tout + 3 * S.n);
    return 0;
}

extern "C" {

int orc_tgv_linearise_f32(const float* ref, const float* src, int W, int H, int N, const float* K,
                          const float* Rref, const float* tref, const float* Rsrc, const float* tsrc, int alternative,
                          double beta, double gamma, const float* u0, float* It, float* Iu, float* Tout) {
    return tgv_linearise<float>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, beta, gamma, u0, It, Iu,
                                Tout);
}
int orc_tgv_linearise_f64(const float* ref, const float* src, int W, int H, int N, const float* K,
                          const float* Rref, const float* tref, const float* Rsrc, const float* tsrc, int alternative,
                          double beta, double gamma, const float* u0, double* It, double* Iu, double* Tout) {
    return tgv_linearise<double>(ref, src, W, H, N, K, Rref, tref, Rsrc, tsrc, alternative, beta, gamma, u0, It, Iu,
                                 Tout);
}

int orc_tgv_iterate_f32(int W, int H, int N, const float* Tt, const float* u0, const float* It, const float* Iu,
                        float* state, int niters, double lambda, double alpha0, double alpha1, double tau,
                        double sigma) {
    return tgv_iterate<float>(W, H, N, Tt, u0, It, Iu, state, niters, lambda, alpha0, alpha1, tau, sigma);
}
int orc_tgv_iterate_f64(int W, int H, int N, const double* Tt, const double* u0, const double* It, const double* Iu,
                        double* state, int niters, double lambda, double alpha0, double alpha1, double tau,
                        double sigma) {
    return tgv_iterate<double>(W, H, N, Tt, u0, It, Iu, state, niters, lambda, alpha0, alpha1, tau, sigma);
}

double orc_tgv_energy(const float* u, const float* u1x, const float* u1y, const float* u0, const float* Tt,
                      const float* It, const float* Iu, int W, int H, int N, double lambda, double alpha0,
                      double alpha1) {
    return energy<float>(u, u1x, u1y, u0, Tt, It, Iu, W, H, N, lambda, alpha0, alpha1);
}
double orc_tgv_energy_f64(const double* u, const double* u1x, const double* u1y, const double* u0, const double* Tt,
                          const double* It, const double* Iu, int W, int H, int N, double lambda, double alpha0,
                          double alpha1) {
    return energy<double>(u, u1x, u1y, u0, Tt, It, Iu, W, H, N, lambda, alpha0, alpha1);
}
double orc_tgv_energy_masked_f64(const double* u, const double* u1x, const double* u1y, const double* u0,
                                 const double* Tt, const double* It, const double* Iu, int W, int H, int N,
                                 double lambda, double alpha0, double alpha1, const float* weight) {
    return energy<double>(u, u1x, u1y, u0, Tt, It, Iu, W, H, N, lambda, alpha0, alpha1, weight);
}

void orc_fusion_bin_centers(int bins, float* out) { bin_centers(bins, out); }

void orc_fusion_hist_f32(uint16_t* hist, int gw, int gh, int gd, const float* vol, const float* depth, int W, int H,
                         const float* K, const float* R, const float* T, float threshold) {
    fusion_hist<float>(hist, gw, gh, gd, vol, depth, W, H, K, R, T, threshold);
}
void orc_fusion_hist_f64(uint16_t* hist, int gw, int gh, int gd, const float* vol, const float* depth, int W, int H,
                         const float* K, const float* R, const float* T, float threshold) {
    fusion_hist<double>(hist, gw, gh, gd, vol, depth, W, H, K, R, T, threshold);
}
void orc_fusion_u_f32(float* u, float* v, const float* p, const uint16_t* hist, int gw, int gh, int gd, double tau,
                      double lambda) {
    fusion_u<float>(u, v, p, hist, gw, gh, gd, tau, lambda);
}
void orc_fusion_u_f64(double* u, double* v, const double* p, const uint16_t* hist, int gw, int gh, int gd,
                      double tau, double lambda) {
    fusion_u<double>(u, v, p, hist, gw, gh, gd, tau, lambda);
}
void orc_fusion_p_f32(float* p, const float* v, int gw, int gh, int gd, double sigma) {
    fusion_p<float>(p, v, gw, gh, gd, sigma);
}
void orc_fusion_p_f64(double* p, const double* v, int gw, int gh, int gd, double sigma) {
    fusion_p<double>(p, v, gw, gh, gd, sigma);
}

// TV-L1 histogram energy of u (no counterpart in the reference, which has no
// energy): sum_i |grad u_i|_2 + lambda sum_i sum_k h_k(i) |u_i - c_k|, with the
// forward differences of FusionUpdateP_kernel (src/fusion.cu:59-70, zero at
// the far faces) and the bin centres of binParams (src/inc/fusion.h:1041-1054).
// Sequential double sum over voxels in x, y, z order.
double orc_fusion_energy_f64(const double* u, const uint16_t* hist, int gw, int gh, int gd, double lambda) {
    float cf[NB];
    bin_centers(NB, cf);
    const size_t sxy = (size_t)gw * gh;
    double e = 0.0;
    for (int z = 0; z < gd; ++z)
        for (int y = 0; y < gh; ++y)
            for (int x = 0; x < gw; ++x) {
                const size_t i = (size_t)z * sxy + (size_t)y * gw + x;
                const double gx = x < gw - 1 ? u[i + 1] - u[i] : 0.0;
                const double gy = y < gh - 1 ? u[i + gw] - u[i] : 0.0;
                const double gz = z < gd - 1 ? u[i + sxy] - u[i] : 0.0;
                double d = 0.0;
                for (int k = 0; k < NB; ++k) d += (double)hist[i * NB + k] * std::fabs(u[i] - (double)cf[k]);
                e += std::sqrt(gx * gx + gy * gy + gz * gz) + lambda * d;
            }
    return e;
}

float orc_prox_hist_insertion_f32(float u, const int* Wv, float tau, float lambda) {
    float c[NB];
    bin_centers(NB, c);
    SortedHist<float> s(c);
    s.insert(u);
    for (int j = 0; j < NB; ++j) s.insert(u + tau * lambda * (float)Wv[j]);
    return s.median();
}

int64_t orc_surface_count_f32(const float* u, int64_t n) {
    int64_t c = 0;
    for (int64_t i = 0; i < n; ++i) c += (u[i] < 0.f && u[i] > -1.f);
    return c;
}
int64_t orc_surface_count_f64(const double* u, int64_t n) {
    int64_t c = 0;
    for (int64_t i = 0; i < n; ++i) c += (u[i] < 0.0 && u[i] > -1.0);
    return c;
}

int orc_tvl1_f32(const float* depth, const float* ref, int W, int H, float znear, float zfar, int niters,
                 double lambda, double tau, double sigma, double theta, double beta, double gamma, float* out) {
    return tvl1<float>(depth, ref, W, H, znear, zfar, niters, lambda, tau, sigma, theta, beta, gamma, out);
}
int orc_tvl1_f64(const float* depth, const float* ref, int W, int H, float znear, float zfar, int niters,
                 double lambda, double tau, double sigma, double theta, double beta, double gamma, double* out) {
    return tvl1<double>(depth, ref, W, H, znear, zfar, niters, lambda, tau, sigma, theta, beta, gamma, out);
}

int orc_tgv_sparse_f32(const float* ref, const float* dense, const float* sparse, int W, int H, float zfar,
                       int niters, double alpha0, double alpha1, double tau, double sigma, double theta, double beta,
                       double gamma, float* out, float* u1_out) {
    return tgv_sparse<float>(ref, dense, sparse, W, H, zfar, niters, alpha0, alpha1, tau, sigma, theta, beta, gamma,
                             out, u1_out);
}
int orc_tgv_sparse_f64(const float* ref, const float* dense, const float* sparse, int W, int H, float zfar,
                       int niters, double alpha0, double alpha1, double tau, double sigma, double theta, double beta,
                       double gamma, double* out, double* u1_out) {
    return tgv_sparse<double>(ref, dense, sparse, W, H, zfar, niters, alpha0, alpha1, tau, sigma, theta, beta, gamma,
                              out, u1_out);
}

void orc_compute3d_f32(const float* depth, int W, int H, const float* K, const float* Rref, const float* tref,
                       int alt, float* X, float* Y, float* Z) {
    compute3d<float>(depth, W, H, K, Rref, tref, alt, X, Y, Z);
}
void orc_compute3d_f64(const float* depth, int W, int H, const float* K, const float* Rref, const float* tref,
                       int alt, double* X, double* Y, double* Z) {
    compute3d<double>(depth, W, H, K, Rref, tref, alt, X, Y, Z);
}

}  // extern "C"
