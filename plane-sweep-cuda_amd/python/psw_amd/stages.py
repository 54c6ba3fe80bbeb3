"""Stage-level wrappers over libpsw for Python callers (tests, bench).

Every function moves numpy arrays to the device, calls the HIP entry points
of include/psw.h and returns numpy arrays.  Pose convention as the
reference: X_cam = R X_world + t (PlaneSweep::RelativeMatrices default).
"""
import ctypes

import numpy as np

from ._lib import FusionGrid, TgvParams, call, lib
from .device import DeviceArray


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _hp(a):
    return ctypes.c_void_p(a.ctypes.data)


# ---- host helpers -----------------------------------------------------------

def reference_planes(znear, zfar, nplanes):
    """The reference's float-accumulated plane depths (planesweep.cpp:169,200)."""
    out = np.zeros(int(nplanes) + 2, np.float32)
    n = ctypes.c_int(0)
    call("psw_reference_planes", float(znear), float(zfar), int(nplanes), _hp(out), out.size, ctypes.byref(n))
    return out[:n.value].copy()


def inverse_depth_planes(znear, zfar, nplanes):
    """Planes uniform in inverse depth over [znear, zfar]."""
    inv = np.linspace(1.0 / znear, 1.0 / zfar, int(nplanes))
    return (1.0 / inv).astype(np.float32)


def relative_matrices(Rref, tref, Rsrc, tsrc, alternative=0):
    Rrel = np.zeros(9, np.float32)
    trel = np.zeros(3, np.float32)
    a = [_f32(x).ravel() for x in (Rref, tref, Rsrc, tsrc)]
    call("psw_relative_matrices", *[_hp(x) for x in a], int(alternative), _hp(Rrel), _hp(trel))
    return Rrel.reshape(3, 3), trel


def mat3_inverse(m):
    m = _f32(m).ravel()
    out = np.zeros(9, np.float32)
    call("psw_mat3_inverse", _hp(m), _hp(out))
    return out.reshape(3, 3)


def homographies_reference(depths, K, Rref, tref, Rsrc, tsrc, alternative=0):
    depths = _f32(depths)
    Rsrc = _f32(Rsrc).reshape(-1, 9)
    N, P = Rsrc.shape[0], depths.size
    H = np.zeros((N, P, 9), np.float32)
    args = [_f32(x).ravel() for x in (K, Rref, tref, Rsrc, tsrc)]
    call("psw_sweep_homographies_reference", _hp(H), _hp(depths), P, N, *[_hp(x) for x in args], int(alternative))
    return H


# ---- stage 1: plane sweep -----------------------------------------------------

def sweep_reference(ref, src, K, Rref, tref, Rsrc, tsrc, planes, winsize, stdthresh, nccthresh, zfar,
                    alternative=0):
    """PlaneSweep::RunAlgorithm semantics on the fused kernel; K 1-based."""
    ref, src, planes = _f32(ref), _f32(src), _f32(planes)
    Hh, W = ref.shape
    N = src.shape[0]
    Hm = homographies_reference(planes, K, Rref, tref, Rsrc, tsrc, alternative)
    d_ref, d_src = DeviceArray.from_host(ref), DeviceArray.from_host(src)
    d_H, d_dep = DeviceArray.from_host(Hm), DeviceArray.from_host(planes)
    d_out = DeviceArray((Hh, W))
    call("psw_sweep_fused_reference", d_ref.ptr, d_src.ptr, W, Hh, N, d_H.ptr, d_dep.ptr, planes.size, int(winsize),
         float(stdthresh), float(nccthresh), float(zfar), d_out.ptr, None)
    return d_out.to_host()


class SweepAggregate:
    """Aggregated sweep with device-resident inputs (bench use)."""

    def __init__(self, ref, src, K, Rref, tref, Rsrc, tsrc, planes, winsize, stdthresh, coord_offset=0,
                 alternative=0):
        if isinstance(ref, DeviceArray):      # images already in device memory
            self.d_ref, self.d_src = ref, src
            self.H_img, self.W = ref.shape
            self.N = src.shape[0]
        else:
            ref, src = _f32(ref), _f32(src)
            self.H_img, self.W = ref.shape
            self.N = src.shape[0]
            self.d_ref, self.d_src = DeviceArray.from_host(ref), DeviceArray.from_host(src)
        self.planes = _f32(planes)
        self.P = self.planes.size
        self.winsize, self.stdthresh, self.coord_offset = int(winsize), float(stdthresh), int(coord_offset)
        self.d_dep = DeviceArray.from_host(self.planes)
        self.d_H = DeviceArray((self.N, self.P, 9))
        self.poses = [_f32(x).ravel() for x in (K, Rref, tref, Rsrc, tsrc)]
        self.alternative = int(alternative)
        n = self.W * self.H_img
        self.d_arg = DeviceArray((self.H_img, self.W), np.int32)
        self.d_cmin = DeviceArray((self.H_img, self.W))
        self.d_csec = DeviceArray((self.H_img, self.W))
        self.d_depth = DeviceArray((self.H_img, self.W))
        self.npix = n

    def build_homographies(self, stream=None):
        call("psw_sweep_homographies", self.d_H.ptr, self.d_dep.ptr, self.P, self.N,
             *[_hp(x) for x in self.poses], self.alternative, stream)

    def run(self, fast=False, stream=None):
        call("psw_sweep_fused_aggregate", self.d_ref.ptr, self.d_src.ptr, self.W, self.H_img, self.N, self.d_H.ptr,
             self.P, self.winsize, self.stdthresh, self.coord_offset, int(bool(fast)), self.d_arg.ptr,
             self.d_cmin.ptr, self.d_csec.ptr, stream)

    def depth(self, fallback, stream=None):
        call("psw_sweep_argmin_depth", self.d_depth.ptr, self.d_arg.ptr, self.d_dep.ptr, self.npix, float(fallback),
             stream)
        return self.d_depth

    def results(self, stream=None):
        """(argmin, min cost, second cost) on the host, after the work queued
        on `stream` (the stream the sweep ran on)."""
        return self.d_arg.to_host(stream), self.d_cmin.to_host(stream), self.d_csec.to_host(stream)


def sweep_aggregate(ref, src, K, Rref, tref, Rsrc, tsrc, planes, winsize, stdthresh, coord_offset=0, fast=False,
                    alternative=0):
    s = SweepAggregate(ref, src, K, Rref, tref, Rsrc, tsrc, planes, winsize, stdthresh, coord_offset, alternative)
    s.build_homographies()
    s.run(fast=fast)
    return s.results()


def homographies_device(depths, K, Rref, tref, Rsrc, tsrc, alternative=0):
    """FP64-on-GPU homographies (psw_sweep_homographies), returned to the host."""
    depths = _f32(depths)
    Rsrc = _f32(Rsrc).reshape(-1, 9)
    N, P = Rsrc.shape[0], depths.size
    d_dep = DeviceArray.from_host(depths)
    d_H = DeviceArray((N, P, 9))
    call("psw_sweep_homographies", d_H.ptr, d_dep.ptr, P, N, *[_hp(_f32(x).ravel()) for x in (K, Rref, tref, Rsrc, tsrc)],
         int(alternative), None)
    return d_H.to_host()


# ---- stage 2: TGV2 -------------------------------------------------------------

def tgv_params(lambda_=0.5, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, beta=0.0, gamma=1.0):
    """PlaneSweep::TGV defaults (defines.h:49-57)."""
    return TgvParams(lambda_, alpha0, alpha1, tau, sigma, beta, gamma)


class TgvSolver:
    """psw_tgv handle: device state of one TGV2 problem."""

    def __init__(self, W, H, N):
        h = ctypes.c_void_p()
        call("psw_tgv_create", ctypes.byref(h), int(W), int(H), int(N))
        self.h, self.W, self.H, self.N = h, int(W), int(H), int(N)

    def set_images(self, d_ref, d_src, beta=0.0, gamma=1.0, stream=None):
        call("psw_tgv_set_images", self.h, d_ref.ptr, d_src.ptr, float(beta), float(gamma), stream)

    def set_u(self, d_u=None, value=1.0, stream=None):
        call("psw_tgv_set_u", self.h, d_u.ptr if d_u is not None else None, float(value), stream)

    def linearise(self, K, Rrel, trel, stream=None):
        K, Rrel, trel = _f32(K).ravel(), _f32(Rrel).ravel(), _f32(trel).ravel()
        call("psw_tgv_linearise", self.h, _hp(K), _hp(Rrel), _hp(trel), stream)

    def set_primal(self, d_u=None, value=1.0, d_u1=None, stream=None):
        """Warm start inside the current linearisation; duals restart from 0."""
        call("psw_tgv_set_primal", self.h, d_u.ptr if d_u is not None else None, float(value),
             d_u1.ptr if d_u1 is not None else None, stream)

    def set_duals(self, d_p=None, d_q=None, d_r=None, stream=None):
        """Load p (2,H,W), q (3,H,W), r (N,H,W) device arrays; None means 0."""
        call("psw_tgv_set_duals", self.h, *[d.ptr if d is not None else None for d in (d_p, d_q, d_r)], stream)

    def set_energy_mask(self, qmax, prm, stream=None, count=True):
        """Leave out of the energy the pixels whose data step is unstable
        (q = tau sigma lambda^2 sum Iu^2 > qmax at them or a 4-neighbour);
        returns how many were left out (synchronous) when count is set."""
        c = ctypes.c_int64()
        call("psw_tgv_set_energy_mask", self.h, float(qmax), ctypes.byref(prm), ctypes.byref(c) if count else None,
             stream)
        return c.value if count else None

    def iterate(self, niters, prm, stream=None):
        call("psw_tgv_iterate", self.h, int(niters), ctypes.byref(prm), stream)

    def run(self, K, Rrel, trel, niters, warps, prm, d_out, d_init=None, stream=None):
        """PlaneSweep::TGV on the images already set: u from d_init (None:
        the constant 1), `warps` linearisations of `niters` iterations each;
        the depth into d_out (H, W)."""
        K, Rrel, trel = _f32(K).ravel(), _f32(Rrel).ravel(), _f32(trel).ravel()
        call("psw_tgv_run", self.h, _hp(K), _hp(Rrel), _hp(trel), int(niters), int(warps), ctypes.byref(prm),
             d_init.ptr if d_init is not None else None, d_out.ptr, stream)

    def energy(self, prm, stream=None):
        e = ctypes.c_double()
        call("psw_tgv_energy", self.h, ctypes.byref(prm), ctypes.byref(e), stream)
        return e.value

    def solve(self, max_iters, check_every, tol, e_star, prm, stream=None):
        it, e = ctypes.c_int(), ctypes.c_double()
        call("psw_tgv_solve", self.h, int(max_iters), int(check_every), float(tol), float(e_star), ctypes.byref(prm),
             ctypes.byref(it), ctypes.byref(e), stream)
        return it.value, e.value

    def get_u(self, stream=None):
        d = DeviceArray((self.H, self.W))
        call("psw_tgv_get_u", self.h, d.ptr, stream)
        return d.to_host(stream)

    def get_energy_mask(self, stream=None):
        """(H, W) weights of the energy: 1 counted, 0 left out."""
        d = DeviceArray((self.H, self.W))
        call("psw_tgv_get_energy_mask", self.h, d.ptr, stream)
        return d.to_host(stream)

    def state_device(self, stream=None):
        """Full state planes (as get_state) in a new device array."""
        d = DeviceArray((8 + 3 * self.N + 4, self.H, self.W))
        call("psw_tgv_get_state", self.h, d.ptr, stream)
        return d

    def get_state(self, stream=None):
        n = self.W * self.H
        planes = 8 + 3 * self.N + 3 + 1
        d = DeviceArray((planes, self.H, self.W))
        call("psw_tgv_get_state", self.h, d.ptr, stream)
        s = d.to_host(stream)
        N = self.N
        return {"u": s[0], "u1x": s[1], "u1y": s[2], "px": s[3], "py": s[4], "qx": s[5], "qy": s[6], "qz": s[7],
                "It": s[8:8 + N], "Iu": s[8 + N:8 + 2 * N], "r": s[8 + 2 * N:8 + 3 * N],
                "T11": s[8 + 3 * N], "T12": s[9 + 3 * N], "T22": s[10 + 3 * N], "u0": s[11 + 3 * N]}

    def __del__(self):
        try:
            lib.psw_tgv_destroy(self.h)
        except Exception:
            pass


def relative_all(Rref, tref, Rsrc, tsrc, alternative=0):
    Rsrc = _f32(Rsrc).reshape(-1, 3, 3)
    tsrc = _f32(tsrc).reshape(-1, 3)
    Rr, tr = zip(*[relative_matrices(Rref, tref, Rsrc[i], tsrc[i], alternative) for i in range(Rsrc.shape[0])])
    return np.stack(Rr), np.stack(tr)


def tgv_run(ref, src, K, Rref, tref, Rsrc, tsrc, niters, warps, prm, u_init=None, alternative=0, state=False):
    """PlaneSweep::TGV (planesweep.cpp:381-536) on the fused kernels; images 0..255, K 1-based."""
    ref, src = _f32(ref), _f32(src)
    Hh, W = ref.shape
    N = src.shape[0]
    Rrel, trel = relative_all(Rref, tref, Rsrc, tsrc, alternative)
    t = TgvSolver(W, Hh, N)
    d_ref, d_src = DeviceArray.from_host(ref), DeviceArray.from_host(src)
    t.set_images(d_ref, d_src, prm.beta, prm.gamma)
    d_init = DeviceArray.from_host(_f32(u_init)) if u_init is not None else None
    d_out = DeviceArray((Hh, W))
    call("psw_tgv_run", t.h, _hp(_f32(K).ravel()), _hp(_f32(Rrel).ravel()), _hp(_f32(trel).ravel()), int(niters),
         int(warps), ctypes.byref(prm), d_init.ptr if d_init is not None else None, d_out.ptr, None)
    u = d_out.to_host()
    if state:
        return u, t.get_state()
    return u


def tgv_sparse(ref, dense, sparse, zfar, niters, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, theta=1.0, beta=0.0,
               gamma=1.0, workspace=None, stream=None):
    """PlaneSweep::TGVdenoiseFromSparse (planesweep.cpp:596-674) on the fused
    kernel: reference image 0..255, dense initial depth, sparse depth (> 0
    where known).  Inputs may be host arrays or DeviceArrays; returns a
    DeviceArray when they are device-resident, else a host array.
    workspace: a DeviceArray of at least tgv_sparse_workspace_bytes() bytes,
    or None (allocated per call)."""
    dev = isinstance(ref, DeviceArray)
    d_ref, d_dense, d_sparse = [x if isinstance(x, DeviceArray) else DeviceArray.from_host(_f32(x))
                                for x in (ref, dense, sparse)]
    Hh, W = d_ref.shape
    d_out = DeviceArray((Hh, W))
    call("psw_tgv_sparse_denoise", d_ref.ptr, d_dense.ptr, d_sparse.ptr, W, Hh, float(zfar), int(niters),
         float(alpha0), float(alpha1), float(tau), float(sigma), float(theta), float(beta), float(gamma),
         workspace.ptr if workspace is not None else None, workspace.nbytes if workspace is not None else 0,
         d_out.ptr, stream)
    return d_out if dev else d_out.to_host(stream)


def tgv_sparse_workspace_bytes(W, H):
    return int(lib.psw_tgv_sparse_workspace_bytes(int(W), int(H)))


# ---- stage 3: fusion ---------------------------------------------------------------

class Fusion:
    """SoA voxel grid on the device (psw_fusion_grid)."""

    def __init__(self, w, h, d, volume):
        self.g = FusionGrid()
        vol = _f32(volume).ravel()
        call("psw_fusion_alloc", ctypes.byref(self.g), int(w), int(h), int(d), _hp(vol))
        self.shape = (int(d), int(h), int(w))
        self.V = int(w) * int(h) * int(d)
        self.volume = vol.copy()

    def clear(self, stream=None):
        call("psw_fusion_clear", ctypes.byref(self.g), stream)

    def update_histogram(self, d_depth, W, H, K, R, T, threshold, stream=None):
        call("psw_fusion_update_histogram", ctypes.byref(self.g), d_depth.ptr, int(W), int(H),
             *[_hp(_f32(x).ravel()) for x in (K, R, T)], float(threshold), stream)

    def update_histogram_batch(self, d_depths, m, W, H, K, R, T, threshold, stream=None):
        call("psw_fusion_update_histogram_batch", ctypes.byref(self.g), d_depths.ptr, int(m), int(W), int(H),
             *[_hp(_f32(x).ravel()) for x in (K, R, T)], float(threshold), stream)

    def update_u(self, tau, lambda_, stream=None):
        call("psw_fusion_update_u", ctypes.byref(self.g), float(tau), float(lambda_), stream)

    def update_p(self, sigma, stream=None):
        call("psw_fusion_update_p", ctypes.byref(self.g), float(sigma), stream)

    def update_iteration(self, d_depth, W, H, K, R, T, threshold, tau, lambda_, sigma, stream=None):
        call("psw_fusion_update_iteration", ctypes.byref(self.g), d_depth.ptr, int(W), int(H),
             *[_hp(_f32(x).ravel()) for x in (K, R, T)], float(threshold), float(tau), float(lambda_), float(sigma),
             stream)

    def solve(self, niters, tau, lambda_, sigma, stream=None):
        call("psw_fusion_solve", ctypes.byref(self.g), int(niters), float(tau), float(lambda_), float(sigma), stream)

    def energy(self, lambda_, stream=None):
        """FP64 TV-L1 histogram energy of the current u (synchronous)."""
        e = ctypes.c_double()
        call("psw_fusion_energy", ctypes.byref(self.g), float(lambda_), ctypes.byref(e), stream)
        return e.value

    def surface_points(self, stream=None):
        """(n, 3) float32 world coordinates of the voxels with -1 < u < 0,
        in z, y, x order."""
        c = ctypes.c_int64()
        call("psw_fusion_surface_points", ctypes.byref(self.g), None, 0, ctypes.byref(c), stream)
        n = c.value
        if n == 0:
            return np.zeros((0, 3), np.float32)
        d = DeviceArray((n, 3))
        call("psw_fusion_surface_points", ctypes.byref(self.g), d.ptr, n, ctypes.byref(c), stream)
        return d.to_host(stream)

    def surface_count(self, stream=None):
        c = ctypes.c_int64()
        call("psw_fusion_surface_count", ctypes.byref(self.g), ctypes.byref(c), stream)
        return c.value

    def _get(self, ptr, count, dtype):
        out = np.empty(count, dtype)
        call("psw_memcpy_dtoh", _hp(out), ctypes.c_void_p(ptr), out.nbytes, None)
        return out

    def _set(self, ptr, arr):
        arr = np.ascontiguousarray(arr)
        call("psw_memcpy_htod", ctypes.c_void_p(ptr), _hp(arr), arr.nbytes, None)

    def download(self):
        V = self.V
        d, h, w = self.shape
        return {"u": self._get(self.g.u, V, np.float32).reshape(d, h, w),
                "v": self._get(self.g.v, V, np.float32).reshape(d, h, w),
                "p": self._get(self.g.p, 3 * V, np.float32).reshape(3, d, h, w),
                "hist": self._get(self.g.hist, 8 * V, np.uint16).reshape(d, h, w, 8)}

    def upload(self, u=None, v=None, p=None, hist=None):
        if u is not None:
            self._set(self.g.u, _f32(u))
        if v is not None:
            self._set(self.g.v, _f32(v))
        if p is not None:
            self._set(self.g.p, _f32(p))
        if hist is not None:
            self._set(self.g.hist, np.ascontiguousarray(hist, np.uint16))

    def __del__(self):
        try:
            lib.psw_fusion_free(ctypes.byref(self.g))
        except Exception:
            pass


# ---- TVL1 ---------------------------------------------------------------------------

def tvl1_denoise(depth, ref, znear, zfar, niters=100, lambda_=0.3, tau=0.02, sigma=6.0, theta=1.0, beta=0.0,
                 gamma=1.0):
    """PlaneSweep::CudaDenoise (planesweep.cpp:297-379); defaults defines.h:40-46."""
    depth, ref = _f32(depth), _f32(ref)
    Hh, W = depth.shape
    d_dep, d_ref = DeviceArray.from_host(depth), DeviceArray.from_host(ref)
    d_out, d_work = DeviceArray((Hh, W)), DeviceArray((12, Hh, W))
    tvl1_denoise_device(d_dep, d_ref, W, Hh, znear, zfar, d_out, d_work, niters, lambda_, tau, sigma, theta, beta,
                        gamma)
    return d_out.to_host()


def tvl1_denoise_device(d_depth, d_ref, W, Hh, znear, zfar, d_out, d_work, niters=100, lambda_=0.3, tau=0.02,
                        sigma=6.0, theta=1.0, beta=0.0, gamma=1.0, stream=None):
    """psw_tvl1_denoise on device buffers, queued on `stream` (a run of the
    resident solver waits for its launch, whose error words it reads); d_work
    holds 12*W*H floats."""
    call("psw_tvl1_denoise", d_depth.ptr, d_ref.ptr, W, Hh, float(znear), float(zfar), int(niters), float(lambda_),
         float(tau), float(sigma), float(theta), float(beta), float(gamma), d_out.ptr, d_work.ptr, stream)
