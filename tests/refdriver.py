"""The reference's host sequencing run over libpsw's per-kernel launchers.

PlaneSweep::RunAlgorithm / PlaneSweepThread (src/planesweep.cpp:74-259) and
PlaneSweep::TGV (src/planesweep.cpp:381-536) call the kernels of
src/inc/kernels.cu.h in a fixed order; these functions issue the same calls
through the C ABI (psw_<kernel>), so comparing their output with the oracle
checks every launcher in its real use.  Device buffers the reference leaves
uninitialised start at zero.
"""
import ctypes

import numpy as np

from psw_amd import stages
from psw_amd._lib import ZERO_DIM as Z
from psw_amd._lib import call
from psw_amd.device import DeviceArray


def _hp(a):
    return ctypes.c_void_p(np.ascontiguousarray(a).ctypes.data)


def _zeros(H, W):
    d = DeviceArray((H, W))
    d.fill(0.0)
    return d


def run_algorithm(ref, src, K, Rref, tref, Rsrc, tsrc, planes, winsize, stdthresh, nccthresh, zfar, alternative=0):
    ref = np.ascontiguousarray(ref, np.float32)
    H, W = ref.shape
    N = src.shape[0]
    dref = DeviceArray.from_host(ref)
    refmean, refstd, inter1 = _zeros(H, W), _zeros(H, W), _zeros(H, W)
    call("psw_windowed_mean_column", inter1.ptr, dref.ptr, winsize, 0, W, H, Z, Z, None)
    call("psw_windowed_mean_row", refmean.ptr, inter1.ptr, winsize, 0, W, H, Z, Z, None)
    call("psw_windowed_mean_column", inter1.ptr, dref.ptr, winsize, 1, W, H, Z, Z, None)
    call("psw_windowed_mean_row", refstd.ptr, inter1.ptr, winsize, 0, W, H, Z, Z, None)
    call("psw_calculate_STD", refstd.ptr, refmean.ptr, refstd.ptr, W, H, Z, Z, None)
    gdepth, gN = _zeros(H, W), _zeros(H, W)
    Hs = stages.homographies_reference(planes, K, Rref, tref, Rsrc, tsrc, alternative)
    for i in range(N):
        dsrc = DeviceArray.from_host(np.ascontiguousarray(src[i], np.float32))
        ncc, best, dep, it1 = _zeros(H, W), _zeros(H, W), _zeros(H, W), _zeros(H, W)
        dx, dy, warped = _zeros(H, W), _zeros(H, W), _zeros(H, W)
        for p, d in enumerate(planes):
            h = np.ascontiguousarray(Hs[i, p])
            call("psw_transform_indexes", dx.ptr, dy.ptr, _hp(h), W, H, Z, Z, None)
            call("psw_bilinear_interpolation", warped.ptr, dsrc.ptr, dx.ptr, dy.ptr, W, H, W, H, Z, Z, None)
            call("psw_windowed_mean_column", it1.ptr, warped.ptr, winsize, 0, W, H, Z, Z, None)
            call("psw_windowed_mean_row", dx.ptr, it1.ptr, winsize, 0, W, H, Z, Z, None)
            call("psw_windowed_mean_column", dy.ptr, warped.ptr, winsize, 1, W, H, Z, Z, None)
            call("psw_windowed_mean_row", it1.ptr, dy.ptr, winsize, 0, W, H, Z, Z, None)
            call("psw_calculate_STD", dy.ptr, dx.ptr, it1.ptr, W, H, Z, Z, None)
            call("psw_element_multiply", it1.ptr, dref.ptr, warped.ptr, W, H, Z, Z, None)
            call("psw_windowed_mean_column", warped.ptr, it1.ptr, winsize, 0, W, H, Z, Z, None)
            call("psw_windowed_mean_row", it1.ptr, warped.ptr, winsize, 0, W, H, Z, Z, None)
            call("psw_calcNCC", ncc.ptr, it1.ptr, refmean.ptr, dx.ptr, refstd.ptr, dy.ptr, float(stdthresh),
                 float(stdthresh), W, H, Z, Z, None)
            call("psw_update_arrays", dep.ptr, best.ptr, ncc.ptr, float(d), W, H, Z, Z, None)
        call("psw_sum_depthmap_NCC", gdepth.ptr, gN.ptr, dep.ptr, best.ptr, float(nccthresh), W, H, Z, Z, None)
    call("psw_element_rdivide", gdepth.ptr, gdepth.ptr, gN.ptr, W, H, Z, Z, None)
    call("psw_set_QNAN_value", gdepth.ptr, float(zfar), W, H, Z, Z, None)
    return gdepth.to_host()


def tgv(ref, src, K, Rref, tref, Rsrc, tsrc, niters, warps, lambda_=0.5, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0,
        beta=0.0, gamma=1.0, alternative=0):
    ref = np.ascontiguousarray(ref, np.float32)
    H, W = ref.shape
    N = src.shape[0]
    f = lambda: _zeros(H, W)  # noqa: E731
    Ref, Px, Py, u, u0, u1x, u1y, ubar, u1xbar, u1ybar, qx, qy, qz, qw, prodsum = [f() for _ in range(15)]
    x, y, X, Y, Z_, dX, dY, dZ, dfx, dfy, T1, T2, T3, T4 = [f() for _ in range(14)]
    call("psw_set_value", u.ptr, 1.0, W, H, Z, Z, None)
    call("psw_memcpy_dtod", ubar.ptr, u.ptr, u.nbytes, None)
    Ref.upload(ref)
    call("psw_element_scale", Ref.ptr, np.float32(1) / np.float32(255), W, H, Z, Z, None)
    call("psw_Anisotropic_diffusion_tensor", T1.ptr, T2.ptr, T3.ptr, T4.ptr, Ref.ptr, float(beta), float(gamma), W, H,
         Z, Z, None)
    Src, It, Iu, r = [], [], [], []
    Rrel, trel = stages.relative_all(Rref, tref, Rsrc, tsrc, alternative)
    for i in range(N):
        s = DeviceArray.from_host(np.ascontiguousarray(src[i], np.float32))
        call("psw_element_scale", s.ptr, np.float32(1) / np.float32(255), W, H, Z, Z, None)
        Src.append(s)
        It.append(f()); Iu.append(f()); r.append(f())
    K = np.ascontiguousarray(K, np.float32)
    invK = stages.mat3_inverse(K)
    fx, fy = float(K[0, 0]), float(K[1, 1])
    for _ in range(warps):
        call("psw_memcpy_dtod", u0.ptr, u.ptr, u.nbytes, None)
        call("psw_memcpy_dtod", ubar.ptr, u.ptr, u.nbytes, None)
        for b in (Px, Py, qx, qy, qz, qw, u1x, u1y, u1xbar, u1ybar):
            call("psw_set_value", b.ptr, 0.0, W, H, Z, Z, None)
        for i in range(N):
            call("psw_TGV2_transform_coordinates", x.ptr, y.ptr, X.ptr, Y.ptr, Z_.ptr, u0.ptr, _hp(K), _hp(Rrel[i]),
                 _hp(trel[i]), _hp(invK), W, H, Z, Z, None)
            call("psw_TGV2_calculate_coordinate_derivatives", dX.ptr, dY.ptr, dZ.ptr, _hp(invK), _hp(Rrel[i]), W, H,
                 Z, Z, None)
            call("psw_TGV2_calculate_derivativeF", dfx.ptr, dfy.ptr, X.ptr, dX.ptr, Y.ptr, dY.ptr, Z_.ptr, dZ.ptr, fx,
                 fy, W, H, Z, Z, None)
            call("psw_bilinear_interpolation", X.ptr, Src[i].ptr, x.ptr, y.ptr, W, H, W, H, Z, Z, None)
            call("psw_TGV2_calculate_Iu", Iu[i].ptr, X.ptr, dfx.ptr, dfy.ptr, W, H, Z, Z, None)
            call("psw_subtract", It[i].ptr, X.ptr, Ref.ptr, W, H, Z, Z, None)
            call("psw_set_value", r[i].ptr, 0.0, W, H, Z, Z, None)
        for _ in range(niters):
            call("psw_TGV2_updateP_tensor_weighed", Px.ptr, Py.ptr, T1.ptr, T2.ptr, T3.ptr, T4.ptr, ubar.ptr,
                 u1xbar.ptr, u1ybar.ptr, float(alpha1), float(sigma), W, H, Z, Z, None)
            call("psw_TGV2_updateQ", qx.ptr, qy.ptr, qz.ptr, qw.ptr, u1xbar.ptr, u1ybar.ptr, float(alpha0),
                 float(sigma), W, H, Z, Z, None)
            call("psw_set_value", prodsum.ptr, 0.0, W, H, Z, Z, None)
            for j in range(N):
                call("psw_TGV2_updateR", r[j].ptr, prodsum.ptr, u.ptr, u0.ptr, It[j].ptr, Iu[j].ptr, float(sigma),
                     float(lambda_), W, H, Z, Z, None)
            call("psw_TGV2_updateU_tensor_weighed", u.ptr, u1x.ptr, u1y.ptr, T1.ptr, T2.ptr, T3.ptr, T4.ptr, ubar.ptr,
                 u1xbar.ptr, u1ybar.ptr, Px.ptr, Py.ptr, qx.ptr, qy.ptr, qz.ptr, qw.ptr, prodsum.ptr, float(alpha0),
                 float(alpha1), float(tau), float(lambda_), W, H, Z, Z, None)
    return u.to_host()


def tgv_sparse(ref, dense, sparse, zfar, niters, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, theta=1.0, beta=0.0,
               gamma=1.0):
    """PlaneSweep::TGVdenoiseFromSparse (src/planesweep.cpp:596-674) over
    the launchers, with its aliasing: vxbar / vybar are both u1 and u1bar of
    the U step."""
    ref = np.ascontiguousarray(ref, np.float32)
    H, W = ref.shape
    f = lambda: _zeros(H, W)  # noqa: E731
    px, py, qx, qy, qz, qw, ubar, vx, vy, weights, T1, T2, T3, T4 = [f() for _ in range(14)]
    Ds = DeviceArray.from_host(np.ascontiguousarray(sparse, np.float32))
    Ref = DeviceArray.from_host(ref)
    ubar.upload(np.ascontiguousarray(dense, np.float32))
    sz = np.float32(1) / np.float32(zfar)
    call("psw_calculateWeights_sparseDepth", weights.ptr, Ds.ptr, W, H, Z, Z, None)
    call("psw_element_scale", Ds.ptr, sz, W, H, Z, Z, None)
    call("psw_element_scale", ubar.ptr, sz, W, H, Z, Z, None)
    u = f()
    call("psw_memcpy_dtod", u.ptr, ubar.ptr, u.nbytes, None)
    call("psw_element_scale", Ref.ptr, np.float32(1) / np.float32(255), W, H, Z, Z, None)
    call("psw_Anisotropic_diffusion_tensor", T1.ptr, T2.ptr, T3.ptr, T4.ptr, Ref.ptr, float(beta), float(gamma), W, H,
         Z, Z, None)
    for _ in range(niters):
        call("psw_TGV2_updateP_tensor_weighed", px.ptr, py.ptr, T1.ptr, T2.ptr, T3.ptr, T4.ptr, u.ptr, vx.ptr, vy.ptr,
             float(alpha1), float(sigma), W, H, Z, Z, None)
        call("psw_TGV2_updateQ", qx.ptr, qy.ptr, qz.ptr, qw.ptr, vx.ptr, vy.ptr, float(alpha0), float(sigma), W, H, Z,
             Z, None)
        call("psw_TGV2_updateU_sparseDepthTensor", u.ptr, vx.ptr, vy.ptr, ubar.ptr, vx.ptr, vy.ptr, T1.ptr, T2.ptr,
             T3.ptr, T4.ptr, px.ptr, py.ptr, qx.ptr, qy.ptr, qz.ptr, qw.ptr, weights.ptr, Ds.ptr, float(alpha0),
             float(alpha1), float(tau), float(theta), W, H, Z, Z, None)
    call("psw_element_scale", u.ptr, np.float32(zfar), W, H, Z, Z, None)
    return u.to_host()


def compute3d(depth, K, Rref, tref, alternative=0):
    """PlaneSweep::get3Dcoordinates (src/planesweep.cpp:551-586) over
    psw_compute3D; x, y of pixels beyond the depth limit read 0."""
    depth = np.ascontiguousarray(depth, np.float32)
    H, W = depth.shape
    Rr, t = stages.relative_matrices(Rref, tref, np.eye(3, dtype=np.float32), np.zeros(3, np.float32), alternative)
    invK = stages.mat3_inverse(np.ascontiguousarray(K, np.float32))
    px, py = _zeros(H, W), _zeros(H, W)
    X = DeviceArray.from_host(depth)
    call("psw_compute3D", px.ptr, py.ptr, X.ptr, _hp(np.ascontiguousarray(Rr, np.float32)),
         _hp(np.ascontiguousarray(t, np.float32)), _hp(invK), W, H, Z, Z, None)
    return px.to_host(), py.to_host(), X.to_host()
