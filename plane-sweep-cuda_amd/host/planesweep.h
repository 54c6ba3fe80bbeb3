// planesweep.h -- the reference's PlaneSweep host class (src/inc/planesweep.h:
// 27-510) over libpsw's C ABI: same public members, setters, getters, run
// calls, defaults (src/inc/defines.h:23-82) and bool + message error
// convention.  Runs go to the fused gfx950 kernels; no CUDA runtime, no
// device reset on error (the reference's cudaDeviceReset would invalidate
// caller-owned fusion grids).
#pragma once

#include <string>
#include <vector>

#include "../../include/psw.h"
#include "cam_image.h"
#include "structs.h"

// Defaults of src/inc/defines.h
#define DEFAULT_Z_NEAR 0.1f
#define DEFAULT_Z_FAR 1.0f
#define DEFAULT_NUMBER_OF_PLANES 200
#define DEFAULT_NUMBER_OF_IMAGES 4
#define DEFAULT_WINDOW_SIZE 5
#define DEFAULT_STD_THRESHOLD 0.0001f
#define DEFAULT_NCC_THRESHOLD 0.5f
#define NO_CUDA_DEVICE -1
#define MAX_THREADS_PER_BLOCK 512
#define DEFAULT_BLOCK_XDIM 32
#define DEFAULT_TVL1_ITERATIONS 100
#define DEFAULT_TVL1_LAMBDA .3
#define DEFAULT_TVL1_TAU 0.02
#define DEFAULT_TVL1_SIGMA 6.f
#define DEFAULT_TVL1_THETA 1.f
#define DEFAULT_TVL1_BETA 0.f
#define DEFAULT_TVL1_GAMMA 1.f
#define DEFAULT_TGV_LAMBDA 0.5
#define DEFAULT_TGV_ALPHA0 2.0
#define DEFAULT_TGV_ALPHA1 1.5
#define DEFAULT_TGV_NITERS 30
#define DEFAULT_TGV_NWARPS 15
#define DEFAULT_TGV_SIGMA 1.f
#define DEFAULT_TGV_TAU 0.02
#define DEFAULT_TGV_BETA 0.f
#define DEFAULT_TGV_GAMMA 1.f
#define DEFAULT_FUSION_SD_THRESHOLD 0.05
#define DEFAULT_FUSION_TAU 0.1
#define DEFAULT_FUSION_LAMBDA 0.5
#define DEFAULT_FUSION_SIGMA 1.f

namespace psw {

class PlaneSweep {
public:
    CamImage<float> HostRef;
    std::vector<CamImage<float>> HostSrc;
    CamImage<uchar> HostRef8u;
    std::vector<CamImage<uchar>> HostSrc8u;

    PlaneSweep();
    PlaneSweep(int argc, char** argv);
    ~PlaneSweep();
    // The object owns device buffers and a stream: not copyable.
    PlaneSweep(const PlaneSweep&) = delete;
    PlaneSweep& operator=(const PlaneSweep&) = delete;

    // PlaneSweep::RunAlgorithm (planesweep.cpp:74-161): depthmap of HostRef
    // from HostSrc[0..numberimages) with the reference's per-source argmax
    // and thresholded average.
    bool RunAlgorithm(int argc, char** argv);
    // OpenCV's CPU denoiser in the reference (planesweep.cpp:261-277); OpenCV
    // is not part of this library: returns false with a message, as the
    // reference does when built without OpenCV.
    bool Denoise(unsigned int niter, double lambda);
    // PlaneSweep::CudaDenoise (planesweep.cpp:297-379).
    bool CudaDenoise(int argc, char** argv, const unsigned int niters = DEFAULT_TVL1_ITERATIONS,
                     const double lambda = DEFAULT_TVL1_LAMBDA, const double tau = DEFAULT_TVL1_TAU,
                     const double sigma = DEFAULT_TVL1_SIGMA, const double theta = DEFAULT_TVL1_THETA,
                     const double beta = DEFAULT_TVL1_BETA, const double gamma = DEFAULT_TVL1_GAMMA);
    // PlaneSweep::TGV (planesweep.cpp:381-536).
    bool TGV(int argc, char** argv, const unsigned int niters = DEFAULT_TGV_NITERS,
             const unsigned int warps = DEFAULT_TGV_NWARPS, const double lambda = DEFAULT_TGV_LAMBDA,
             const double alpha0 = DEFAULT_TGV_ALPHA0, const double alpha1 = DEFAULT_TGV_ALPHA1,
             const double tau = DEFAULT_TGV_TAU, const double sigma = DEFAULT_TGV_SIGMA,
             const double beta = DEFAULT_TGV_BETA, const double gamma = DEFAULT_TGV_GAMMA);
    // PlaneSweep::TGVdenoiseFromSparse (planesweep.cpp:596-674, work in
    // progress in the reference): densifies `depth` (> 0 where known) from
    // the current depthmap with the fused sparse-depth TGV2 kernel; result in
    // depthmapTGV and in the device buffer d_depthmap.
    bool TGVdenoiseFromSparse(int argc, char** argv, const CamImage<float>& depth, const unsigned int niters,
                              const double alpha0, const double alpha1, const double tau, const double sigma,
                              const double theta, const double beta, const double gamma);

    void RelativeMatrices(Matrix3D& Rrel, Vector3D& trel, const Matrix3D& Rref, const Vector3D& tref,
                          const Matrix3D& Rsrc, const Vector3D& tsrc) const;

    // Setters (planesweep.h:170-277)
    void setAlternativeRelativeMatrixMethod(bool method) { alternativemethod = method; }
    void setK(const Matrix3D& Km) { K = Km; invK = Km.inv(); }
    void setZ(float n, float f) { znear = n; zfar = f; }
    void setZnear(float n) { znear = n; }
    void setZfar(float f) { zfar = f; }
    void setNumberofPlanes(unsigned int n) { numberplanes = n; }
    void setNumberofImages(unsigned int n) { numberimages = n; }
    void setWindowSize(unsigned int sz);
    void setSTDthreshold(float th) { stdthresh = th; }
    void setNCCthreshold(float th) { nccthresh = th; }
    void setThreadsPerBlock(dim3 tpb) { threads = tpb; }
    void setBlockXdim(int& threadsx);
    void setBlockYdim(int& threadsy);

    // Getters
    bool getAlternativeRelativeMatrixMethod() const { return alternativemethod; }
    Matrix3D getK() const { return K; }
    Matrix3D getInverseK() const { return invK; }
    CamImage<float>* getDepthmap() { return &depthmap; }
    CamImage<float>* getDepthmapDenoised() { return &depthmapdenoised; }
    float* getDepthmapDenoisedPtr() { return d_depthmap; }   // device pointer owned by this object
    CamImage<uchar>* getDepthmap8u() { return &depthmap8u; }
    CamImage<uchar>* getDepthmap8uDenoised() { return &depthmap8udenoised; }
    CamImage<float>* getDepthmapTGV() { return &depthmapTGV; }
    CamImage<uchar>* getDepthmap8uTGV() { return &depthmap8uTGV; }
    // compute3D (kernels.cu:405-429): world coordinates of the denoised depthmap
    void get3Dcoordinates(CamImage<float>*& x, CamImage<float>*& y, CamImage<float>*& z);
    float getZnear() const { return znear; }
    float getZfar() const { return zfar; }
    unsigned int getNumberofPlanes() const { return numberplanes; }
    unsigned int getNumberofImages() const { return numberimages; }
    unsigned int getWindowSize() const { return winsize; }
    float getSTDthreshold() const { return stdthresh; }
    float getNCCthreshold() const { return nccthresh; }
    unsigned int getMaxThreadsPerBlock() const { return maxThreadsPerBlock; }
    dim3 getThreadsPerBlock() const { return threads; }

    // Additions (no reference counterpart).
    // Explicit plane depths instead of the znear..zfar loop (empty = the loop).
    void setPlaneDepths(const std::vector<float>& d) { planedepths = d; }
    // Aggregated sweep: cost = -(sum of NCC over sources)/N, argmin plane,
    // minimum and second-smallest cost; depth of the argmin plane in depthmap.
    // zeroBasedK: K uses 0-based pixel coordinates.
    bool RunAggregated(bool zeroBasedK = false, bool fast = true);
    const std::vector<int>& getArgmin() const { return argmin; }
    const CamImage<float>* getCostMin() const { return &costmin; }
    const CamImage<float>* getCostSecond() const { return &costsecond; }
    // Initial depth for TGV (empty = the reference's constant 1).
    void setTGVInit(const CamImage<float>& u) { tgvinit = u; }
    // Run TGV until the relative energy gap to e_star is <= tol (one warp).
    bool TGVSolve(int max_iters, int check_every, double tol, double e_star, double lambda = DEFAULT_TGV_LAMBDA,
                  double alpha0 = DEFAULT_TGV_ALPHA0, double alpha1 = DEFAULT_TGV_ALPHA1,
                  double tau = DEFAULT_TGV_TAU, double sigma = DEFAULT_TGV_SIGMA, int* iters = nullptr,
                  double* energy = nullptr);
    const std::string& lastError() const { return lasterror; }
    std::vector<float> planeDepths() const;
    // The object's stream: every run of this object is queued on it, and a
    // run returns once its results are copied back.  Created on first use.
    psw_stream_t stream();
    // Device bytes this object's buffer pool holds (grow-only: a run allocates
    // only when an image, source count or plane count is larger than before).
    size_t deviceBytes() const;

protected:
    Matrix3D K;
    Matrix3D invK;
    CamImage<float> depthmap, depthmapdenoised;
    CamImage<uchar> depthmap8u, depthmap8udenoised;
    CamImage<float> depthmapTGV;
    CamImage<uchar> depthmap8uTGV;
    float* d_depthmap = nullptr;   // the pool's DEPTHMAP buffer (getDepthmapDenoisedPtr)
    CamImage<float> coord_x, coord_y, coord_z;

    float znear = DEFAULT_Z_NEAR;
    float zfar = DEFAULT_Z_FAR;
    unsigned int numberplanes = DEFAULT_NUMBER_OF_PLANES;
    unsigned int numberimages = DEFAULT_NUMBER_OF_IMAGES;
    unsigned int winsize = DEFAULT_WINDOW_SIZE;
    float stdthresh = DEFAULT_STD_THRESHOLD;
    float nccthresh = DEFAULT_NCC_THRESHOLD;

    int maxThreadsPerBlock = MAX_THREADS_PER_BLOCK;
    dim3 blocks, threads;
    bool depthavailable = false;
    bool alternativemethod = false;

    std::vector<float> planedepths;
    std::vector<int> argmin;
    CamImage<float> costmin, costsecond, tgvinit;
    std::string lasterror;

    void ConvertDepthtoUChar(const CamImage<float>& input, CamImage<uchar>& output);
    int deviceInit(int argc, char** argv);
    bool fail(const char* where, int status);
    int nImages() const;

    // Per-object device buffers, one per role, grown on demand and freed
    // with the object (the reference allocates and frees its device images
    // on every call, src/planesweep.cpp:122-123,172-189,313-316).
    enum Slot { REF, SRC, HOM, PLANES, OUT, ARG, CMIN, CSEC, INIT, WORK, DENSE, SPARSE, CX, CY, CZ, DEPTHMAP, NSLOTS };
    void* pool[NSLOTS] = {};
    size_t poolcap[NSLOTS] = {};
    int device = -1;               // device the pool and the stream belong to
    psw_stream_t strm = nullptr;
    psw_tgv* tgv = nullptr;        // TGV2 solver, kept while the shape stays the same
    int tgvw = 0, tgvh = 0, tgvn = 0;
    int buffer(Slot s, size_t bytes, float** out);
    int solver(int w, int h, int nimages, psw_tgv** out);
    void release();
};

}  // namespace psw
