// host_api_test.cpp -- drives the reference-compatible C++ API (PlaneSweep,
// fusionData<8>, FusionUpdateIteration<8>, per-kernel launchers) the way the
// reference's GUI does (src/pclviewer.cpp:437-503, :633-704, :972-1083), on
// inputs written by tests/test_host_api.py, and writes the outputs back for
// comparison with the oracle.
//
// usage: host_api_test <dir>     (reads <dir>/scene.bin, <dir>/fusion.bin)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "fusion.h"
#include "kernels.h"
#include "planesweep.h"

using namespace psw;

static std::vector<float> readf(FILE* f, size_t n) {
    std::vector<float> v(n);
    if (fread(v.data(), 4, n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
    return v;
}
static void writef(const std::string& path, const void* p, size_t bytes) {
    FILE* f = fopen(path.c_str(), "wb");
    fwrite(p, 1, bytes, f);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    FILE* f = fopen((dir + "/scene.bin").c_str(), "rb");
    if (!f) return 2;
    int hdr[4];
    if (fread(hdr, 4, 4, f) != 4) return 2;
    const int W = hdr[0], H = hdr[1], N = hdr[2], P = hdr[3];
    std::vector<float> prm = readf(f, 8);  // znear zfar win stdthresh nccthresh tgv_iters tgv_warps lambda
    std::vector<float> K = readf(f, 9), Rr = readf(f, 9), tr = readf(f, 3);
    PlaneSweep ps(argc, argv);
    ps.HostRef.reset(W, H);
    ps.HostRef.R = Matrix3D(Rr.data());
    ps.HostRef.t = Vector3D(tr.data());
    ps.HostSrc.resize(N);
    for (int i = 0; i < N; ++i) {
        std::vector<float> R = readf(f, 9), t = readf(f, 3);
        ps.HostSrc[i].reset(W, H);
        ps.HostSrc[i].R = Matrix3D(R.data());
        ps.HostSrc[i].t = Vector3D(t.data());
    }
    std::vector<float> img = readf(f, (size_t)W * H);
    std::copy(img.begin(), img.end(), ps.HostRef.data());
    for (int i = 0; i < N; ++i) {
        img = readf(f, (size_t)W * H);
        std::copy(img.begin(), img.end(), ps.HostSrc[i].data());
    }
    fclose(f);

    ps.setK(Matrix3D(K.data()));
    ps.setZ(prm[0], prm[1]);
    ps.setNumberofPlanes(P);
    ps.setNumberofImages(N);
    ps.setWindowSize((unsigned)prm[2]);
    ps.setSTDthreshold(prm[3]);
    ps.setNCCthreshold(prm[4]);
    const size_t n = (size_t)W * H;
    if (!ps.RunAlgorithm(argc, argv)) return 3;
    writef(dir + "/out_sweep.bin", ps.getDepthmap()->data(), n * 4);
    writef(dir + "/out_sweep8u.bin", ps.getDepthmap8u()->data(), n);
    if (!ps.CudaDenoise(argc, argv, 30)) return 4;
    writef(dir + "/out_tvl1.bin", ps.getDepthmapDenoised()->data(), n * 4);
    if (!ps.TGV(argc, argv, (unsigned)prm[5], (unsigned)prm[6], prm[7])) return 5;
    writef(dir + "/out_tgv.bin", ps.getDepthmapTGV()->data(), n * 4);
    // world coordinates of the TVL1 depthmap, then sparse-depth TGV2 from
    // every 7th TVL1 depth over the sweep depthmap
    {
        CamImage<float>*cx = nullptr, *cy = nullptr, *cz = nullptr;
        ps.get3Dcoordinates(cx, cy, cz);
        if (!cx || !cy || !cz) return 12;
        writef(dir + "/out_x.bin", cx->data(), n * 4);
        writef(dir + "/out_y.bin", cy->data(), n * 4);
        writef(dir + "/out_z.bin", cz->data(), n * 4);
        CamImage<float> sparse(W, H);
        for (size_t i = 0; i < n; ++i) sparse.data()[i] = i % 7 == 0 ? ps.getDepthmapDenoised()->data()[i] : 0.f;
        if (!ps.TGVdenoiseFromSparse(argc, argv, sparse, 20, 2.0, 1.5, 0.02, 1.0, 1.0, 0.0, 1.0)) return 13;
        writef(dir + "/out_tgv_sparse.bin", ps.getDepthmapTGV()->data(), n * 4);
    }
    // error convention: an even window is rejected by the setter, a run with
    // too few planes fails with false
    ps.setWindowSize(4);
    if (ps.getWindowSize() != (unsigned)prm[2]) return 6;

    // per-kernel launchers: PlaneSweep::RunAlgorithm's reference statistics
    {
        void *dref = nullptr, *dtmp = nullptr, *dmean = nullptr;
        psw_malloc(&dref, n * 4); psw_malloc(&dtmp, n * 4); psw_malloc(&dmean, n * 4);
        psw_memcpy_htod(dref, ps.HostRef.data(), n * 4, nullptr);
        dim3 threads(32, 16), blocks((W + 31) / 32, (H + 15) / 16);
        int st = windowed_mean_column((float*)dtmp, (const float*)dref, (unsigned)prm[2], false, W, H, blocks, threads);
        st |= windowed_mean_row((float*)dmean, (const float*)dtmp, (unsigned)prm[2], false, W, H, blocks, threads);
        std::vector<float> mean(n);
        st |= psw_memcpy_dtoh(mean.data(), dmean, n * 4, nullptr);
        psw_free(dref); psw_free(dtmp); psw_free(dmean);
        if (st) return 7;
        writef(dir + "/out_refmean.bin", mean.data(), n * 4);
    }

    // fusion, as the GUI's reconstruct loop
    f = fopen((dir + "/fusion.bin").c_str(), "rb");
    if (!f) return 8;
    int fh[6];
    if (fread(fh, 4, 6, f) != 6) return 8;
    const int dW = fh[0], dH = fh[1], M = fh[2], gw = fh[3], gh = fh[4], gd = fh[5];
    std::vector<float> vol = readf(f, 6), fp = readf(f, 4), FK = readf(f, 9);
    Rectangle3D volm(make_float3(vol[0], vol[1], vol[2]), make_float3(vol[3], vol[4], vol[5]));
    fusionData<8> fd(gw, gh, gd, volm);
    if (fd.status() != PSW_OK) return 9;
    void* dmap = nullptr;
    psw_malloc(&dmap, (size_t)dW * dH * 4);
    std::vector<std::vector<float>> Rs(M), Ts(M);
    for (int k = 0; k < M; ++k) { Rs[k] = readf(f, 9); Ts[k] = readf(f, 3); }
    for (int k = 0; k < M; ++k) {
        std::vector<float> d = readf(f, (size_t)dW * dH);
        psw_memcpy_htod(dmap, d.data(), d.size() * 4, nullptr);
        if (FusionUpdateIteration<8>(fd, (const float*)dmap, Matrix3D(FK.data()), Matrix3D(Rs[k].data()),
                                     Vector3D(Ts[k].data()), fp[0], fp[1], fp[2], fp[3], dW, dH, dim3(16, 16, 2),
                                     dim3(1, 1, 1)) != PSW_OK)
            return 10;
    }
    fclose(f);
    psw_free(dmap);
    std::vector<fusionvoxel<8>> vox(fd.elements());
    if (fd.copyTo(vox.data(), gw * sizeof(fusionvoxel<8>)) != PSW_OK) return 11;
    std::vector<float> u(fd.elements());
    std::vector<unsigned char> hist(fd.elements() * 8);
    for (size_t i = 0; i < vox.size(); ++i) {
        u[i] = vox[i].u;
        for (int b = 0; b < 8; ++b) hist[i * 8 + b] = vox[i].h.bin[b];
    }
    writef(dir + "/out_fusion_u.bin", u.data(), u.size() * 4);
    writef(dir + "/out_fusion_hist.bin", hist.data(), hist.size());
    float3 c = fd.worldCoords(1, 2, 3);
    writef(dir + "/out_world.bin", &c, 12);
    printf("host api test ok\n");
    return 0;
}
