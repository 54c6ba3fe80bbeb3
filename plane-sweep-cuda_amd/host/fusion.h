// fusion.h -- the reference's fusion host API (src/inc/fusion.h:105-1055,
// src/inc/fusion.cu.h:40-96) over libpsw's C ABI: fusionData<8> owning a
// device voxel grid, value copies that alias it without owning it, geometry
// helpers, and the FusionUpdate{Histogram,U,P,Iteration}<8> calls.
//
// The device storage is SoA (u, v, p[3], uint16 hist[8]); copyTo/copyFrom
// convert to and from the reference's 28-byte AoS fusionvoxel<8> (bins wrap
// modulo 256 like the reference's unsigned char counters).
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/psw.h"
#include "structs.h"

namespace psw {

enum MemoryKind { Device, Host, Managed, Standard };  // src/inc/memory.h

template <unsigned char _nBins>
struct histogram {
    unsigned char bin[_nBins];
    histogram() { std::memset(bin, 0, sizeof(bin)); }
    unsigned char& first() { return bin[0]; }
    unsigned char& last() { return bin[_nBins - 1]; }
    unsigned char& operator()(unsigned char i) { return bin[i]; }
};

template <unsigned char _nBins>
struct fusionvoxel {
    float u;
    float v;
    float3 p;
    histogram<_nBins> h;
    fusionvoxel() : u(0), v(0), p(make_float3(0, 0, 0)), h() {}
};

template <unsigned char _histBins, MemoryKind memT = Device>
class fusionData {
    static_assert(_histBins == PSW_FUSION_BINS && memT == Device,
                  "libpsw provides device grids with 8 histogram bins (fusionData<8>)");
};

template <>
class fusionData<PSW_FUSION_BINS, Device> {
public:
    fusionData() { std::memset(&g_, 0, sizeof(g_)); binParams(); }
    fusionData(size_t w, size_t h, size_t d) { alloc(w, h, d, Rectangle3D()); }
    fusionData(size_t w, size_t h, size_t d, float3 x, float3 y) { alloc(w, h, d, Rectangle3D(x, y)); }
    fusionData(size_t w, size_t h, size_t d, const Rectangle3D& vol) { alloc(w, h, d, vol); }
    // A copy aliases the grid without owning it (fusion.h:176-182).
    fusionData(const fusionData& fd) : g_(fd.g_), own_(false), status_(fd.status_) { binParams(); }
    fusionData& operator=(const fusionData& fd) {
        if (this == &fd) return *this;
        release();
        g_ = fd.g_;
        own_ = false;
        status_ = fd.status_;
        return *this;
    }
    ~fusionData() { release(); }

    size_t width() const { return (size_t)g_.width; }
    size_t height() const { return (size_t)g_.height; }
    size_t depth() const { return (size_t)g_.depth; }
    size_t elements() const { return width() * height() * depth(); }
    unsigned char bins() const { return PSW_FUSION_BINS; }
    bool ManageData() const { return own_; }
    void setManageData(bool manage) { own_ = manage; }
    int status() const { return status_; }

    Rectangle3D volume() const {
        return Rectangle3D(make_float3(g_.volume[0], g_.volume[1], g_.volume[2]),
                           make_float3(g_.volume[3], g_.volume[4], g_.volume[5]));
    }
    void setVolume(const Rectangle3D& vol) {
        g_.volume[0] = vol.a.x; g_.volume[1] = vol.a.y; g_.volume[2] = vol.a.z;
        g_.volume[3] = vol.b.x; g_.volume[4] = vol.b.y; g_.volume[5] = vol.b.z;
    }
    void setVolume(float3 x, float3 y) { setVolume(Rectangle3D(x, y)); }

    int index(int nx = 0, int ny = 0, int nz = 0) const { return nx + ny * g_.width + nz * g_.width * g_.height; }
    int3 indexes(int ix) const {
        return int3{ix % g_.width, ix / g_.width % g_.height, ix / (g_.width * g_.height)};
    }
    // Voxel centre in world coordinates (fusion.h:360-364).
    float3 worldCoords(int x, int y, int z) const {
        const Rectangle3D v = volume();
        return v.a + v.size() * make_float3((float)((x + .5) / (double)width()), (float)((y + .5) / (double)height()),
                                            (float)((z + .5) / (double)depth()));
    }
    float binCenter(unsigned char i) const { return i < PSW_FUSION_BINS ? bin_[i] : 0.f; }
    const float* binCenters() const { return bin_; }
    float binStep() const { return binstep_; }

    void Resize(size_t w, size_t h, size_t d) {
        const Rectangle3D v = volume();
        release();
        alloc(w, h, d, v);
    }

    // Device -> host AoS copy with a row pitch in bytes (fusion.h:933-947).
    int copyTo(fusionvoxel<PSW_FUSION_BINS>* data, size_t npitch) const {
        const size_t V = elements();
        std::vector<float> u(V), v(V), p(3 * V);
        std::vector<uint16_t> h(V * PSW_FUSION_BINS);
        int st = psw_memcpy_dtoh(u.data(), g_.u, V * 4, nullptr);
        if (st == PSW_OK) st = psw_memcpy_dtoh(v.data(), g_.v, V * 4, nullptr);
        if (st == PSW_OK) st = psw_memcpy_dtoh(p.data(), g_.p, 3 * V * 4, nullptr);
        if (st == PSW_OK) st = psw_memcpy_dtoh(h.data(), g_.hist, V * 2 * PSW_FUSION_BINS, nullptr);
        if (st != PSW_OK) return st;
        for (size_t z = 0; z < depth(); ++z)
            for (size_t y = 0; y < height(); ++y) {
                auto* row = reinterpret_cast<fusionvoxel<PSW_FUSION_BINS>*>(reinterpret_cast<unsigned char*>(data) +
                                                                            (z * height() + y) * npitch);
                for (size_t x = 0; x < width(); ++x) {
                    const size_t i = (z * height() + y) * width() + x;
                    row[x].u = u[i];
                    row[x].v = v[i];
                    row[x].p = make_float3(p[i], p[V + i], p[2 * V + i]);
                    for (int b = 0; b < PSW_FUSION_BINS; ++b)
                        row[x].h.bin[b] = (unsigned char)(h[i * PSW_FUSION_BINS + b] & 0xff);
                }
            }
        return PSW_OK;
    }
    // Host AoS -> device (fusion.h:916-924).
    int copyFrom(const fusionvoxel<PSW_FUSION_BINS>* data, size_t npitch) {
        const size_t V = elements();
        std::vector<float> u(V), v(V), p(3 * V);
        std::vector<uint16_t> h(V * PSW_FUSION_BINS);
        for (size_t z = 0; z < depth(); ++z)
            for (size_t y = 0; y < height(); ++y) {
                auto* row = reinterpret_cast<const fusionvoxel<PSW_FUSION_BINS>*>(
                    reinterpret_cast<const unsigned char*>(data) + (z * height() + y) * npitch);
                for (size_t x = 0; x < width(); ++x) {
                    const size_t i = (z * height() + y) * width() + x;
                    u[i] = row[x].u;
                    v[i] = row[x].v;
                    p[i] = row[x].p.x; p[V + i] = row[x].p.y; p[2 * V + i] = row[x].p.z;
                    for (int b = 0; b < PSW_FUSION_BINS; ++b) h[i * PSW_FUSION_BINS + b] = row[x].h.bin[b];
                }
            }
        int st = psw_memcpy_htod(g_.u, u.data(), V * 4, nullptr);
        if (st == PSW_OK) st = psw_memcpy_htod(g_.v, v.data(), V * 4, nullptr);
        if (st == PSW_OK) st = psw_memcpy_htod(g_.p, p.data(), 3 * V * 4, nullptr);
        if (st == PSW_OK) st = psw_memcpy_htod(g_.hist, h.data(), V * 2 * PSW_FUSION_BINS, nullptr);
        return st;
    }

    const psw_fusion_grid* grid() const { return &g_; }

private:
    psw_fusion_grid g_;
    bool own_ = true;
    int status_ = PSW_OK;
    float bin_[PSW_FUSION_BINS];
    float binstep_ = 0.f;

    void alloc(size_t w, size_t h, size_t d, const Rectangle3D& vol) {
        binParams();
        const float v[6] = {vol.a.x, vol.a.y, vol.a.z, vol.b.x, vol.b.y, vol.b.z};
        status_ = psw_fusion_alloc(&g_, (int)w, (int)h, (int)d, v);
        own_ = status_ == PSW_OK;
    }
    void release() {
        if (own_) psw_fusion_free(&g_);
        own_ = false;
    }
    // fusionData::binParams (fusion.h:1041-1054)
    void binParams() {
        const int b = PSW_FUSION_BINS;
        bin_[0] = -1.f;
        bin_[b - 1] = 1.f;
        for (int i = 1; i < b - 1; i++) bin_[i] = 2.f * float(i - 1) / float(b - 3) - 1.f;
        binstep_ = 2 / float(b - 3);
    }
};

typedef fusionData<8> dfusionData8;

// src/inc/fusion.cu.h:40-96.  The launch shape is chosen by the library; the
// blocks/threads arguments are accepted for source compatibility.  Each call
// returns the libpsw status (0 = ok).
template <unsigned char _bins>
int FusionUpdateHistogram(fusionData<_bins> f, const float* depthmap, const Matrix3D K, const Matrix3D R,
                          const Vector3D t, const float threshold, const int width, const int height,
                          dim3 = dim3(), dim3 = dim3()) {
    float Kf[9], Rf[9], tf[3] = {t.x, t.y, t.z};
    K.toArray(Kf);
    R.toArray(Rf);
    return psw_fusion_update_histogram(f.grid(), depthmap, width, height, Kf, Rf, tf, threshold, nullptr);
}

template <unsigned char _bins>
int FusionUpdateU(fusionData<_bins> f, const double tau, const double lambda, dim3 = dim3(), dim3 = dim3()) {
    return psw_fusion_update_u(f.grid(), tau, lambda, nullptr);
}

template <unsigned char _bins>
int FusionUpdateP(fusionData<_bins> f, const double sigma, dim3 = dim3(), dim3 = dim3()) {
    return psw_fusion_update_p(f.grid(), sigma, nullptr);
}

template <unsigned char _bins>
int FusionUpdateIteration(fusionData<_bins> f, const float* depthmap, const Matrix3D K, const Matrix3D R,
                          const Vector3D t, const float threshold, const double tau, const double lambda,
                          const double sigma, const int width, const int height, dim3 = dim3(), dim3 = dim3()) {
    float Kf[9], Rf[9], tf[3] = {t.x, t.y, t.z};
    K.toArray(Kf);
    R.toArray(Rf);
    return psw_fusion_update_iteration(f.grid(), depthmap, width, height, Kf, Rf, tf, threshold, tau, lambda, sigma,
                                       nullptr);
}

}  // namespace psw
