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

// Device allocation owned for the duration of one call.
struct DevBuf {
    void* p = nullptr;
    int status = PSW_OK;
    explicit DevBuf(size_t bytes) { status = psw_malloc(&p, bytes); }
    ~DevBuf() {
        if (p) psw_free(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    float* f() { return static_cast<float*>(p); }
};

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

PlaneSweep::~PlaneSweep() {
    if (d_depthmap) psw_free(d_depthmap);
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
    if (psw_set_device(dev) != PSW_OK) return NO_CUDA_DEVICE;
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
    DevBuf dref(n * 4), dsrc(n * nimgs * 4), dH(H.size() * 4), ddep(P * 4), dout(n * 4);
    for (DevBuf* b : {&dref, &dsrc, &dH, &ddep, &dout})
        if (b->status != PSW_OK) return fail("RunAlgorithm", b->status);
    st = psw_memcpy_htod(dref.p, HostRef.data(), n * 4, nullptr);
    for (int i = 0; i < nimgs && st == PSW_OK; ++i)
        st = psw_memcpy_htod(dsrc.f() + (size_t)i * n, HostSrc[i].data(), n * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_htod(dH.p, H.data(), H.size() * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_htod(ddep.p, planes.data(), P * 4, nullptr);
    if (st == PSW_OK)
        st = psw_sweep_fused_reference(dref.f(), dsrc.f(), w, h, nimgs, dH.f(), ddep.f(), P, (int)winsize, stdthresh,
                                       nccthresh, zfar, dout.f(), nullptr);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmap.data(), dout.p, n * 4, nullptr);
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
    DevBuf dref(n * 4), dsrc(n * nimgs * 4), dH((size_t)nimgs * P * 36), ddep(P * 4), darg(n * 4), dmin(n * 4),
        dsec(n * 4);
    for (DevBuf* b : {&dref, &dsrc, &dH, &ddep, &darg, &dmin, &dsec})
        if (b->status != PSW_OK) return fail("RunAggregated", b->status);
    int st = psw_memcpy_htod(dref.p, HostRef.data(), n * 4, nullptr);
    for (int i = 0; i < nimgs && st == PSW_OK; ++i)
        st = psw_memcpy_htod(dsrc.f() + (size_t)i * n, HostSrc[i].data(), n * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_htod(ddep.p, planes.data(), P * 4, nullptr);
    if (st == PSW_OK)
        st = psw_sweep_homographies(dH.f(), ddep.f(), P, nimgs, Kf, Rr, tr, Rs.data(), ts.data(),
                                    alternativemethod ? 1 : 0, nullptr);
    if (st == PSW_OK)
        st = psw_sweep_fused_aggregate(dref.f(), dsrc.f(), w, h, nimgs, dH.f(), P, (int)winsize, stdthresh,
                                       zeroBasedK ? 0 : 1, fast ? 1 : 0, static_cast<int32_t*>(darg.p), dmin.f(),
                                       dsec.f(), nullptr);
    argmin.assign(n, -1);
    costmin.reset(w, h);
    costsecond.reset(w, h);
    depthmap.reset(w, h);
    if (st == PSW_OK) st = psw_memcpy_dtoh(argmin.data(), darg.p, n * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_dtoh(costmin.data(), dmin.p, n * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_dtoh(costsecond.data(), dsec.p, n * 4, nullptr);
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
    if (d_depthmap) {
        psw_free(d_depthmap);
        d_depthmap = nullptr;
    }
    void* p = nullptr;
    int st = psw_malloc(&p, n * 4);
    if (st != PSW_OK) return fail("CudaDenoise", st);
    d_depthmap = static_cast<float*>(p);
    DevBuf din(n * 4), dref(n * 4), work(12 * n * 4);
    for (DevBuf* b : {&din, &dref, &work})
        if (b->status != PSW_OK) return fail("CudaDenoise", b->status);
    st = psw_memcpy_htod(din.p, depthmap.data(), n * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_htod(dref.p, HostRef.data(), n * 4, nullptr);
    if (st == PSW_OK)
        st = psw_tvl1_denoise(din.f(), dref.f(), w, h, znear, zfar, (int)niters, lambda, tau, sigma, theta, beta, gamma,
                              d_depthmap, work.f(), nullptr);
    depthmapdenoised.reset(w, h);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmapdenoised.data(), d_depthmap, n * 4, nullptr);
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
    psw_tgv* t = nullptr;
    int st = psw_tgv_create(&t, w, h, nimages);
    if (st != PSW_OK) return fail("TGV", st);
    DevBuf dref(n * 4), dsrc(n * nimages * 4), dinit(n * 4), dout(n * 4);
    for (DevBuf* b : {&dref, &dsrc, &dinit, &dout})
        if (b->status != PSW_OK) { psw_tgv_destroy(t); return fail("TGV", b->status); }
    st = psw_memcpy_htod(dref.p, HostRef.data(), n * 4, nullptr);
    for (int i = 0; i < nimages && st == PSW_OK; ++i)
        st = psw_memcpy_htod(dsrc.f() + (size_t)i * n, HostSrc[i].data(), n * 4, nullptr);
    const bool warm = tgvinit.width() == (size_t)w && tgvinit.height() == (size_t)h;
    if (st == PSW_OK && warm) st = psw_memcpy_htod(dinit.p, tgvinit.data(), n * 4, nullptr);
    if (st == PSW_OK) st = psw_tgv_set_images(t, dref.f(), dsrc.f(), beta, gamma, nullptr);
    psw_tgv_params prm = {lambda, alpha0, alpha1, tau, sigma, beta, gamma};
    float Kf[9];
    K.toArray(Kf);
    if (st == PSW_OK)
        st = psw_tgv_run(t, Kf, Rrel.data(), trel.data(), (int)niters, (int)warps, &prm, warm ? dinit.f() : nullptr,
                         dout.f(), nullptr);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmapTGV.data(), dout.p, n * 4, nullptr);
    psw_tgv_destroy(t);
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
    psw_tgv* t = nullptr;
    int st = psw_tgv_create(&t, w, h, nimages);
    if (st != PSW_OK) return fail("TGVSolve", st);
    DevBuf dref(n * 4), dsrc(n * nimages * 4), dinit(n * 4), dout(n * 4);
    st = dref.status | dsrc.status | dinit.status | dout.status;
    if (st == PSW_OK) st = psw_memcpy_htod(dref.p, HostRef.data(), n * 4, nullptr);
    for (int i = 0; i < nimages && st == PSW_OK; ++i)
        st = psw_memcpy_htod(dsrc.f() + (size_t)i * n, HostSrc[i].data(), n * 4, nullptr);
    if (st == PSW_OK) st = psw_tgv_set_images(t, dref.f(), dsrc.f(), DEFAULT_TGV_BETA, DEFAULT_TGV_GAMMA, nullptr);
    const bool warm = tgvinit.width() == (size_t)w && tgvinit.height() == (size_t)h;
    if (st == PSW_OK && warm) st = psw_memcpy_htod(dinit.p, tgvinit.data(), n * 4, nullptr);
    if (st == PSW_OK) st = psw_tgv_set_u(t, warm ? dinit.f() : nullptr, 1.f, nullptr);
    float Kf[9];
    K.toArray(Kf);
    if (st == PSW_OK) st = psw_tgv_linearise(t, Kf, Rrel.data(), trel.data(), nullptr);
    psw_tgv_params prm = {lambda, alpha0, alpha1, tau, sigma, DEFAULT_TGV_BETA, DEFAULT_TGV_GAMMA};
    int it = 0;
    double e = 0;
    if (st == PSW_OK) st = psw_tgv_solve(t, max_iters, check_every, tol, e_star, &prm, &it, &e, nullptr);
    depthmapTGV.reset(w, h);
    if (st == PSW_OK) st = psw_tgv_get_u(t, dout.f(), nullptr);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmapTGV.data(), dout.p, n * 4, nullptr);
    psw_tgv_destroy(t);
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
    if (d_depthmap) {   // holds the result afterwards, as in the reference
        psw_free(d_depthmap);
        d_depthmap = nullptr;
    }
    void* p = nullptr;
    int st = psw_malloc(&p, n * 4);
    if (st != PSW_OK) return fail("TGVdenoiseFromSparse", st);
    d_depthmap = static_cast<float*>(p);
    DevBuf dref(n * 4), ddense(n * 4), dsparse(n * 4);
    for (DevBuf* b : {&dref, &ddense, &dsparse})
        if (b->status != PSW_OK) return fail("TGVdenoiseFromSparse", b->status);
    st = psw_memcpy_htod(dref.p, HostRef.data(), n * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_htod(ddense.p, depthmap.data(), n * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_htod(dsparse.p, depth.data(), n * 4, nullptr);
    if (st == PSW_OK)
        st = psw_tgv_sparse_denoise(dref.f(), ddense.f(), dsparse.f(), w, h, zfar, (int)niters, alpha0, alpha1, tau,
                                    sigma, theta, beta, gamma, nullptr, 0, d_depthmap, nullptr);
    if (st == PSW_OK) st = psw_memcpy_dtoh(depthmapTGV.data(), d_depthmap, n * 4, nullptr);
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
    DevBuf Px(n * 4), Py(n * 4), X(n * 4);
    float R9[9], t3[3] = {t.x, t.y, t.z}, iK[9];
    Rr.toArray(R9);
    invK.toArray(iK);
    int st = Px.status | Py.status | X.status;
    if (st == PSW_OK) st = psw_memcpy_htod(X.p, depthmapdenoised.data(), n * 4, nullptr);
    // x, y of pixels at or beyond the depth limit are not written by the
    // kernel (uninitialised in the reference); they read 0 here.
    if (st == PSW_OK) st = psw_fill_f32(Px.f(), 0.f, n, nullptr);
    if (st == PSW_OK) st = psw_fill_f32(Py.f(), 0.f, n, nullptr);
    const psw_dim3 zd = {0, 0, 0};
    if (st == PSW_OK) st = psw_compute3D(Px.f(), Py.f(), X.f(), R9, t3, iK, w, h, zd, zd, nullptr);
    coord_x.reset(w, h);
    coord_y.reset(w, h);
    coord_z.reset(w, h);
    if (st == PSW_OK) st = psw_memcpy_dtoh(coord_x.data(), Px.p, n * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_dtoh(coord_y.data(), Py.p, n * 4, nullptr);
    if (st == PSW_OK) st = psw_memcpy_dtoh(coord_z.data(), X.p, n * 4, nullptr);
    if (st != PSW_OK) {
        fail("get3Dcoordinates", st);
        return;
    }
    x = &coord_x;
    y = &coord_y;
    z = &coord_z;
}

}  // namespace psw
