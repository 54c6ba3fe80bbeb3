// planesweep.cpp -- PlaneSweep over libpsw's C ABI (see planesweep.h).
#include "planesweep.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>

namespace psw {

namespace {

struct Timer {
    std::chrono::high_resolution_clock::time_point t0 = std::chrono::high_resolution_clock::now();
    long long ms() const {
        return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::high_resolution_clock::now() - t0)
            .count();
    }
};

}  // namespace

PlaneSweep::PlaneSweep() : threads(DEFAULT_BLOCK_XDIM) {}

PlaneSweep::PlaneSweep(int argc, char** argv) {
    deviceInit(argc, argv);
    threads = dim3(DEFAULT_BLOCK_XDIM, maxThreadsPerBlock / DEFAULT_BLOCK_XDIM);
}

PlaneSweep::~PlaneSweep() { release(); }

void PlaneSweep::release() {
    if (strm) psw_stream_synchronize(strm);
    for (int i = 0; i < NSLOTS; ++i) {
        if (pool[i]) psw_free(pool[i]);
        pool[i] = nullptr;
        poolcap[i] = 0;
    }
    d_depthmap = nullptr;
    if (tgv) psw_tgv_destroy(tgv);
    tgv = nullptr;
    tgvw = tgvh = tgvn = 0;
    if (strm) psw_stream_destroy(strm);
    strm = nullptr;
}

psw_stream_t PlaneSweep::stream() {
    if (!strm && psw_stream_create(&strm) != PSW_OK) strm = nullptr;   // null stream as the fallback
    return strm;
}

size_t PlaneSweep::deviceBytes() const {
    size_t b = 0;
    for (int i = 0; i < NSLOTS; ++i) b += poolcap[i];
    return b + (tgv ? psw_tgv_state_bytes(tgvw, tgvh, tgvn) : 0);
}

// The slot's buffer with at least `bytes` bytes (grow-only; the old
// contents are not kept when it grows).
int PlaneSweep::buffer(Slot s, size_t bytes, float** out) {
    if (bytes > poolcap[s]) {
        if (strm) psw_stream_synchronize(strm);   // work queued on the old buffer is done
        if (pool[s]) psw_free(pool[s]);
        pool[s] = nullptr;
        poolcap[s] = 0;
        if (s == DEPTHMAP) d_depthmap = nullptr;
        const int st = psw_malloc(&pool[s], bytes);
        if (st != PSW_OK) return st;
        poolcap[s] = bytes;
    }
    *out = static_cast<float*>(pool[s]);
    if (s == DEPTHMAP) d_depthmap = *out;
    return PSW_OK;
}

// The TGV2 solver for this shape, created once per shape.
int PlaneSweep::solver(int w, int h, int nimages, psw_tgv** out) {
    if (!tgv || tgvw != w || tgvh != h || tgvn != nimages) {
        if (strm) psw_stream_synchronize(strm);
        if (tgv) psw_tgv_destroy(tgv);
        tgv = nullptr;
        const int st = psw_tgv_create(&tgv, w, h, nimages);
        if (st != PSW_OK) return st;
        tgvw = w; tgvh = h; tgvn = nimages;
    }
    *out = tgv;
    return PSW_OK;
}

// cudaDevInit (planesweep.cpp:30-58): a "device=N" argument selects the GPU
// (findCudaDevice's convention), otherwise device 0.
int PlaneSweep::deviceInit(int argc, char** argv) {
    int count = 0;
    if (psw_device_count(&count) != PSW_OK || count == 0) {
        std::cerr << "HIP error: no devices supporting HIP." << std::endl;
        return NO_CUDA_DEVICE;
    }
    int dev = 0;
    for (int i = 1; i < argc && argv; ++i) {
        const char* a = argv[i];
        while (*a == '-') ++a;
        if (std::strncmp(a, "device=", 7) == 0) dev = std::atoi(a + 7);
    }
    if (dev < 0 || dev >= count) {
        std::cerr << "Invalid device " << dev << " (" << count << " available)" << std::endl;
        return NO_CUDA_DEVICE;
    }
    if (device >= 0 && dev != device) release();   // the pool and stream belong to the old device
    if (psw_set_device(dev) != PSW_OK) return NO_CUDA_DEVICE;
    device = dev;
    return dev;
}

bool PlaneSweep::fail(const char* where, int status) {
    lasterror = std::string(where) + ": " + psw_status_string(status) + ": " + psw_last_error();
    std::cerr << "Exception caught: \n" << lasterror << std::endl;
    return false;
}

int PlaneSweep::nImages() const {
    return std::min(std::max((int)numberimages, 1), (int)HostSrc.size());
}

void PlaneSweep::setWindowSize(unsigned int sz) {
    if (sz % 2 == 0) std::cout << "Window size must be an odd number";
    else winsize = sz;
}

void PlaneSweep::setBlockXdim(int& threadsx) {
    if (threadsx * (int)threads.y > maxThreadsPerBlock) threadsx = maxThreadsPerBlock / threads.y;
    threads.x = threadsx;
}

void PlaneSweep::setBlockYdim(int& threadsy) {
    if (threadsy * (int)threads.x > maxThreadsPerBlock) threadsy = maxThreadsPerBlock / threads.x;
    threads.y = threadsy;
}

void PlaneSweep::RelativeMatrices(Matrix3D& Rrel, Vector3D& trel, const Matrix3D& Rref, const Vector3D& tref,
                                  const Matrix3D& Rsrc, const Vector3D& tsrc) const {
    float a[9], b[9], ta[3] = {tref.x, tref.y, tref.z}, tb[3] = {tsrc.x, tsrc.y, tsrc.z}, R[9], t[3];
    Rref.toArray(a);
    Rsrc.toArray(b);
    psw_relative_matrices(a, ta, b, tb, alternativemethod ? 1 : 0, R, t);
    Rrel = Matrix3D(R);
    trel = Vector3D(t);
}

std::vector<float> PlaneSweep::planeDepths() const {
    if (!planedepths.empty()) return planedepths;
    std::vector<float> d(numberplanes + 2);
    int n = 0;
    psw_reference_planes(znear, zfar, numberplanes, d.data(), (int)d.size(), &n);
    d.resize(n);
    return d;
}

// PlaneSweep::ConvertDepthtoUChar (planesweep.cpp:279-290)
void PlaneSweep::ConvertDepthtoUChar(const CamImage<float>& input, CamImage<uchar>& output) {
    output.reset(input.width(), input.height());
    for (size_t i = 0; i < input.area(); ++i) {
        const float v = input[i];
        if (v == v) output[i] = uchar(UCHAR_MAX * std::min(std::max((v - znear) / (zfar - znear), 0.f), 1.f));
        else output[i] = UCHAR_MAX;
    }
}

bool PlaneSweep::RunAlgorithm(int argc, char** argv) {
    Timer tm;
    depthmap.reset(HostRef.width(), HostRef.height());
    std::printf("Starting plane sweep algorithm...\n\n");
    if (deviceInit(argc, argv) == NO_CUDA_DEVICE) return false;
    const int w = (int)HostRef.width(), h = (int)HostRef.height();
    const int nimgs = nImages();
    if (nimgs < 1 || w == 0 || h == 0) return fail("RunAlgorithm", PSW_ERR_ARG);
    const std::vector<float> planes = planeDepths();
    const int P = (int)planes.size();
    if (P < 1) return fail("RunAlgorithm", PSW_ERR_ARG);
    const size_t n = (size_t)w * h;
    float Kf[9], Rr[9], tr[3] = {HostRef.t.x, HostRef.t.y, HostRef.t.z};
    K.toArray(Kf);
    HostRef.R.toArray(Rr);
    std::vector<float> Rs(9 * nimgs), ts(3 * nimgs), H((size_t)nimgs * P * 9);
    for (int i = 0; i < nimgs; ++i) {
        if (HostSrc[i].width() != (size_t)w || HostSrc[i].height() != (size_t)h) return fail("RunAlgorithm", PSW_ERR_ARG);
        HostSrc[i].R.toArray(&Rs[9 * i]);
        ts[3 * i] = HostSrc[i].t.x; ts[3 * i + 1] = HostSrc[i].t.y; ts[3 * i + 2] = HostSrc[i].t.z;
    }
    int st = psw_sweep_homographies_reference(H.data(), planes.data(), P, nimgs, Kf, Rr, tr, Rs.data(), ts.data(),
                                              alternativemethod ? 1 : 0);
    if (st != PSW_OK) return fail("RunAlgorithm", st);
    psw_stream_t sm = stream();
    float *dref, *dsrc, *dH, *ddep, *dout;
    if ((st = buffer(REF, n * 4, &dref)) != PSW_OK || (st = buffer(SRC, n * nimgs * 4, &dsrc)) != PSW_OK ||
        (st = buffer(HOM, H.size() * 4, &dH)) != PSW_OK || (st = buffer(PLANES, P * 4, &ddep)) != PSW_OK ||
        (st = buffer(OUT, n * 4, &dout)) != PSW_OK)
        return fail("RunAlgorithm", st);
    st = psw_memcpy_htod(dref, HostRef.data(), n * 4, sm);
    for (int i = 0; i < nimgs && st == PSW_OK; ++i) st = psw_memcpy_htod(dsrc + (size_t)i * n, HostSrc[i].data(), n * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_htod(dH, H.data(), H.size() * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_htod(ddep, planes.data(), P * 4, sm);
    if (st == PSW_OK)
        st = psw_sweep_fused_reference(dref, dsrc, w, h, nimgs, dH, ddep, P, (int)winsize, stdthresh, nccthresh, zfar,
                                       dout, sm);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmap.data(), dout, n * 4, sm);
    if (st != PSW_OK) return fail("RunAlgorithm", st);
    ConvertDepthtoUChar(depthmap, depthmap8u);
    depthavailable = true;
    std::cout << "Time taken for the algorithm to complete is " << tm.ms() << "ms\n\n";
    return true;
}

bool PlaneSweep::RunAggregated(bool zeroBasedK, bool fast) {
    const int w = (int)HostRef.width(), h = (int)HostRef.height();
    const int nimgs = nImages();
    const std::vector<float> planes = planeDepths();
    const int P = (int)planes.size();
    if (nimgs < 1 || nimgs > PSW_MAX_SOURCES || P < 1 || w == 0 || h == 0) return fail("RunAggregated", PSW_ERR_ARG);
    const size_t n = (size_t)w * h;
    float Kf[9], Rr[9], tr[3] = {HostRef.t.x, HostRef.t.y, HostRef.t.z};
    K.toArray(Kf);
    HostRef.R.toArray(Rr);
    std::vector<float> Rs(9 * nimgs), ts(3 * nimgs);
    for (int i = 0; i < nimgs; ++i) {
        HostSrc[i].R.toArray(&Rs[9 * i]);
        ts[3 * i] = HostSrc[i].t.x; ts[3 * i + 1] = HostSrc[i].t.y; ts[3 * i + 2] = HostSrc[i].t.z;
    }
    psw_stream_t sm = stream();
    float *dref, *dsrc, *dH, *ddep, *darg, *dmin, *dsec;
    int st;
    if ((st = buffer(REF, n * 4, &dref)) != PSW_OK || (st = buffer(SRC, n * nimgs * 4, &dsrc)) != PSW_OK ||
        (st = buffer(HOM, (size_t)nimgs * P * 36, &dH)) != PSW_OK || (st = buffer(PLANES, P * 4, &ddep)) != PSW_OK ||
        (st = buffer(ARG, n * 4, &darg)) != PSW_OK || (st = buffer(CMIN, n * 4, &dmin)) != PSW_OK ||
        (st = buffer(CSEC, n * 4, &dsec)) != PSW_OK)
        return fail("RunAggregated", st);
    st = psw_memcpy_htod(dref, HostRef.data(), n * 4, sm);
    for (int i = 0; i < nimgs && st == PSW_OK; ++i) st = psw_memcpy_htod(dsrc + (size_t)i * n, HostSrc[i].data(), n * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_htod(ddep, planes.data(), P * 4, sm);
    if (st == PSW_OK)
        st = psw_sweep_homographies(dH, ddep, P, nimgs, Kf, Rr, tr, Rs.data(), ts.data(), alternativemethod ? 1 : 0, sm);
    if (st == PSW_OK)
        st = psw_sweep_fused_aggregate(dref, dsrc, w, h, nimgs, dH, P, (int)winsize, stdthresh, zeroBasedK ? 0 : 1,
                                       fast ? 1 : 0, reinterpret_cast<int32_t*>(darg), dmin, dsec, sm);
    argmin.assign(n, -1);
    costmin.reset(w, h);
    costsecond.reset(w, h);
    depthmap.reset(w, h);
    if (st == PSW_OK) st = psw_memcpy_dtoh(argmin.data(), darg, n * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_dtoh(costmin.data(), dmin, n * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_dtoh(costsecond.data(), dsec, n * 4, sm);
    if (st != PSW_OK) return fail("RunAggregated", st);
    for (size_t i = 0; i < n; ++i) depthmap[i] = argmin[i] >= 0 ? planes[argmin[i]] : zfar;
    ConvertDepthtoUChar(depthmap, depthmap8u);
    depthavailable = true;
    return true;
}

bool PlaneSweep::Denoise(unsigned int, double) {
    std::cerr << "\nWarning: OpenCV was not found. Denoising aborted.\n\n";
    return false;
}

bool PlaneSweep::CudaDenoise(int argc, char** argv, const unsigned int niters, const double lambda, const double tau,
                             const double sigma, const double theta, const double beta, const double gamma) {
    Timer tm;
    std::printf("Starting TVL1 denoising...\n\n");
    if (!depthavailable) return false;
    if (deviceInit(argc, argv) == NO_CUDA_DEVICE) return false;
    const int w = (int)depthmap.width(), h = (int)depthmap.height();
    const size_t n = (size_t)w * h;
    if (HostRef.width() != (size_t)w || HostRef.height() != (size_t)h) return fail("CudaDenoise", PSW_ERR_ARG);
    psw_stream_t sm = stream();
    float *dout, *din, *dref, *work;
    int st;
    if ((st = buffer(DEPTHMAP, n * 4, &dout)) != PSW_OK || (st = buffer(DENSE, n * 4, &din)) != PSW_OK ||
        (st = buffer(REF, n * 4, &dref)) != PSW_OK || (st = buffer(WORK, 12 * n * 4, &work)) != PSW_OK)
        return fail("CudaDenoise", st);
    st = psw_memcpy_htod(din, depthmap.data(), n * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_htod(dref, HostRef.data(), n * 4, sm);
    if (st == PSW_OK)
        st = psw_tvl1_denoise(din, dref, w, h, znear, zfar, (int)niters, lambda, tau, sigma, theta, beta, gamma, dout,
                              work, sm);
    depthmapdenoised.reset(w, h);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmapdenoised.data(), dout, n * 4, sm);
    if (st != PSW_OK) return fail("CudaDenoise", st);
    ConvertDepthtoUChar(depthmapdenoised, depthmap8udenoised);
    std::cout << "Time taken for the TVL1 denoising to complete is " << tm.ms() << "ms\n\n";
    return true;
}

bool PlaneSweep::TGV(int argc, char** argv, const unsigned int niters, const unsigned int warps, const double lambda,
                     const double alpha0, const double alpha1, const double tau, const double sigma, const double beta,
                     const double gamma) {
    Timer tm;
    std::printf("\nStarting TGV...\n\n");
    if (deviceInit(argc, argv) == NO_CUDA_DEVICE) return false;
    const int w = (int)HostRef.width(), h = (int)HostRef.height();
    const int nimages = nImages();
    if (nimages < 1 || nimages > PSW_MAX_SOURCES || w < 2 || h < 2) return fail("TGV", PSW_ERR_ARG);
    const size_t n = (size_t)w * h;
    depthmapTGV.reset(w, h);
    depthmap8uTGV.reset(w, h);
    std::vector<float> Rrel(9 * nimages), trel(3 * nimages);
    for (int i = 0; i < nimages; ++i) {
        Matrix3D R;
        Vector3D t;
        RelativeMatrices(R, t, HostRef.R, HostRef.t, HostSrc[i].R, HostSrc[i].t);
        R.toArray(&Rrel[9 * i]);
        trel[3 * i] = t.x; trel[3 * i + 1] = t.y; trel[3 * i + 2] = t.z;
    }
    psw_stream_t sm = stream();
    psw_tgv* t = nullptr;
    float *dref, *dsrc, *dinit, *dout;
    int st;
    if ((st = solver(w, h, nimages, &t)) != PSW_OK || (st = buffer(REF, n * 4, &dref)) != PSW_OK ||
        (st = buffer(SRC, n * nimages * 4, &dsrc)) != PSW_OK || (st = buffer(INIT, n * 4, &dinit)) != PSW_OK ||
        (st = buffer(OUT, n * 4, &dout)) != PSW_OK)
        return fail("TGV", st);
    st = psw_memcpy_htod(dref, HostRef.data(), n * 4, sm);
    for (int i = 0; i < nimages && st == PSW_OK; ++i)
        st = psw_memcpy_htod(dsrc + (size_t)i * n, HostSrc[i].data(), n * 4, sm);
    const bool warm = tgvinit.width() == (size_t)w && tgvinit.height() == (size_t)h;
    if (st == PSW_OK && warm) st = psw_memcpy_htod(dinit, tgvinit.data(), n * 4, sm);
    if (st == PSW_OK) st = psw_tgv_set_images(t, dref, dsrc, beta, gamma, sm);
    psw_tgv_params prm = {lambda, alpha0, alpha1, tau, sigma, beta, gamma};
    float Kf[9];
    K.toArray(Kf);
    if (st == PSW_OK)
        st = psw_tgv_run(t, Kf, Rrel.data(), trel.data(), (int)niters, (int)warps, &prm, warm ? dinit : nullptr, dout,
                         sm);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmapTGV.data(), dout, n * 4, sm);
    if (st != PSW_OK) return fail("TGV", st);
    ConvertDepthtoUChar(depthmapTGV, depthmap8uTGV);
    std::cout << "Time taken for the TGV to complete is " << tm.ms() << "ms\n\n";
    return true;
}

bool PlaneSweep::TGVSolve(int max_iters, int check_every, double tol, double e_star, double lambda, double alpha0,
                          double alpha1, double tau, double sigma, int* iters, double* energy) {
    const int w = (int)HostRef.width(), h = (int)HostRef.height();
    const int nimages = nImages();
    if (nimages < 1 || nimages > PSW_MAX_SOURCES || w < 2 || h < 2) return fail("TGVSolve", PSW_ERR_ARG);
    const size_t n = (size_t)w * h;
    std::vector<float> Rrel(9 * nimages), trel(3 * nimages);
    for (int i = 0; i < nimages; ++i) {
        Matrix3D R;
        Vector3D t;
        RelativeMatrices(R, t, HostRef.R, HostRef.t, HostSrc[i].R, HostSrc[i].t);
        R.toArray(&Rrel[9 * i]);
        trel[3 * i] = t.x; trel[3 * i + 1] = t.y; trel[3 * i + 2] = t.z;
    }
    psw_stream_t sm = stream();
    psw_tgv* t = nullptr;
    float *dref, *dsrc, *dinit, *dout;
    int st;
    if ((st = solver(w, h, nimages, &t)) != PSW_OK || (st = buffer(REF, n * 4, &dref)) != PSW_OK ||
        (st = buffer(SRC, n * nimages * 4, &dsrc)) != PSW_OK || (st = buffer(INIT, n * 4, &dinit)) != PSW_OK ||
        (st = buffer(OUT, n * 4, &dout)) != PSW_OK)
        return fail("TGVSolve", st);
    st = psw_memcpy_htod(dref, HostRef.data(), n * 4, sm);
    for (int i = 0; i < nimages && st == PSW_OK; ++i)
        st = psw_memcpy_htod(dsrc + (size_t)i * n, HostSrc[i].data(), n * 4, sm);
    if (st == PSW_OK) st = psw_tgv_set_images(t, dref, dsrc, DEFAULT_TGV_BETA, DEFAULT_TGV_GAMMA, sm);
    const bool warm = tgvinit.width() == (size_t)w && tgvinit.height() == (size_t)h;
    if (st == PSW_OK && warm) st = psw_memcpy_htod(dinit, tgvinit.data(), n * 4, sm);
    if (st == PSW_OK) st = psw_tgv_set_u(t, warm ? dinit : nullptr, 1.f, sm);
    float Kf[9];
    K.toArray(Kf);
    if (st == PSW_OK) st = psw_tgv_linearise(t, Kf, Rrel.data(), trel.data(), sm);
    psw_tgv_params prm = {lambda, alpha0, alpha1, tau, sigma, DEFAULT_TGV_BETA, DEFAULT_TGV_GAMMA};
    int it = 0;
    double e = 0;
    if (st == PSW_OK) st = psw_tgv_solve(t, max_iters, check_every, tol, e_star, &prm, &it, &e, sm);
    depthmapTGV.reset(w, h);
    if (st == PSW_OK) st = psw_tgv_get_u(t, dout, sm);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmapTGV.data(), dout, n * 4, sm);
    if (st != PSW_OK) return fail("TGVSolve", st);
    if (iters) *iters = it;
    if (energy) *energy = e;
    ConvertDepthtoUChar(depthmapTGV, depthmap8uTGV);
    return true;
}

bool PlaneSweep::TGVdenoiseFromSparse(int argc, char** argv, const CamImage<float>& depth, const unsigned int niters,
                                      const double alpha0, const double alpha1, const double tau, const double sigma,
                                      const double theta, const double beta, const double gamma) {
    Timer tm;
    std::printf("\nStarting TGV denoising...\n\n");
    if (deviceInit(argc, argv) == NO_CUDA_DEVICE) return false;
    const int w = (int)HostRef.width(), h = (int)HostRef.height();
    const size_t n = (size_t)w * h;
    if (w < 1 || h < 1 || depth.width() != (size_t)w || depth.height() != (size_t)h ||
        depthmap.width() != (size_t)w || depthmap.height() != (size_t)h)
        return fail("TGVdenoiseFromSparse", PSW_ERR_ARG);
    depthmapTGV.reset(w, h);
    psw_stream_t sm = stream();
    float *dout, *dref, *ddense, *dsparse;   // dout: d_depthmap, which holds the result afterwards as in the reference
    int st;
    if ((st = buffer(DEPTHMAP, n * 4, &dout)) != PSW_OK || (st = buffer(REF, n * 4, &dref)) != PSW_OK ||
        (st = buffer(DENSE, n * 4, &ddense)) != PSW_OK || (st = buffer(SPARSE, n * 4, &dsparse)) != PSW_OK)
        return fail("TGVdenoiseFromSparse", st);
    float* ws = nullptr;
    const size_t wsb = psw_tgv_sparse_workspace_bytes(w, h);
    if ((st = buffer(WORK, wsb, &ws)) != PSW_OK) return fail("TGVdenoiseFromSparse", st);
    st = psw_memcpy_htod(dref, HostRef.data(), n * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_htod(ddense, depthmap.data(), n * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_htod(dsparse, depth.data(), n * 4, sm);
    if (st == PSW_OK)
        st = psw_tgv_sparse_denoise(dref, ddense, dsparse, w, h, zfar, (int)niters, alpha0, alpha1, tau, sigma, theta,
                                    beta, gamma, ws, wsb, dout, sm);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmapTGV.data(), dout, n * 4, sm);
    if (st != PSW_OK) return fail("TGVdenoiseFromSparse", st);
    ConvertDepthtoUChar(depthmapTGV, depthmap8uTGV);
    std::cout << "Time taken for the TGV to complete is " << tm.ms() << "ms\n\n";
    return true;
}

void PlaneSweep::get3Dcoordinates(CamImage<float>*& x, CamImage<float>*& y, CamImage<float>*& z) {
    Matrix3D Rr, I;
    Vector3D t, T(0, 0, 0);
    I.makeIdentity();
    RelativeMatrices(Rr, t, HostRef.R, HostRef.t, I, T);
    const int w = (int)HostRef.width(), h = (int)HostRef.height();
    const size_t n = (size_t)w * h;
    if (depthmapdenoised.width() != (size_t)w || depthmapdenoised.height() != (size_t)h) {
        fail("get3Dcoordinates", PSW_ERR_ARG);
        return;
    }
    float R9[9], t3[3] = {t.x, t.y, t.z}, iK[9];
    Rr.toArray(R9);
    invK.toArray(iK);
    psw_stream_t sm = stream();
    float *px, *py, *pz;
    int st = buffer(CX, n * 4, &px);
    if (st == PSW_OK) st = buffer(CY, n * 4, &py);
    if (st == PSW_OK) st = buffer(CZ, n * 4, &pz);
    if (st == PSW_OK) st = psw_memcpy_htod(pz, depthmapdenoised.data(), n * 4, sm);
    // x, y of pixels at or beyond the depth limit are not written by the
    // kernel (uninitialised in the reference); they read 0 here.
    if (st == PSW_OK) st = psw_fill_f32(px, 0.f, n, sm);
    if (st == PSW_OK) st = psw_fill_f32(py, 0.f, n, sm);
    const psw_dim3 zd = {0, 0, 0};
    if (st == PSW_OK) st = psw_compute3D(px, py, pz, R9, t3, iK, w, h, zd, zd, sm);
    coord_x.reset(w, h);
    coord_y.reset(w, h);
    coord_z.reset(w, h);
    if (st == PSW_OK) st = psw_memcpy_dtoh(coord_x.data(), px, n * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_dtoh(coord_y.data(), py, n * 4, sm);
    if (st == PSW_OK) st = psw_memcpy_dtoh(coord_z.data(), pz, n * 4, sm);
    if (st != PSW_OK) {
        fail("get3Dcoordinates", st);
        return;
    }
    x = &coord_x;
    y = &coord_y;
    z = &coord_z;
}

}  // namespace psw
