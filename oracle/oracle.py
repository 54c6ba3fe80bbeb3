"""ctypes binding of the CPU oracle (oracle/_build/libpsw_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to compare the GPU results against.  The
oracle restates the reference's kernels and host sequencing on the CPU
(psw_oracle.cpp); the _f32 functions keep the reference's float arithmetic,
the _f64 functions run the same algorithm in double.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libpsw_oracle.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"oracle not built: {LIB_PATH} (run make -C oracle)")

_lib = ctypes.CDLL(LIB_PATH)
_vp = ctypes.c_void_p
_i, _f, _d, _u = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_uint
_i64 = ctypes.c_int64

_SIGS = {
    "orc_reference_planes": (_i, [_f, _f, _u, _vp, _i]),
    "orc_mat3_inv_f32": (None, [_vp, _vp]),
    "orc_relative_matrices_f32": (None, [_vp, _vp, _vp, _vp, _i, _vp, _vp]),
    "orc_homography_ref_f32": (None, [_vp, _vp, _vp, _vp, _vp, _i, _f, _vp]),
    "orc_homography_f64": (None, [_vp, _vp, _vp, _vp, _vp, _i, _f, _vp]),
    "orc_windowed_mean_row_f32": (None, [_vp, _vp, _i, _i, _i, _i]),
    "orc_windowed_mean_col_f32": (None, [_vp, _vp, _i, _i, _i, _i]),
    "orc_bilinear_f32": (None, [_vp, _vp, _vp, _vp, _i, _i, _i, _i]),
    "orc_transform_indexes_f32": (None, [_vp, _vp, _vp, _i, _i]),
    "orc_sweep_ref_f32": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _f, _f, _f, _vp, _vp,
                               _vp]),
    "orc_sweep_ref_f64": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _f, _f, _f, _vp, _vp,
                               _vp]),
    "orc_sweep_agg_f32": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _f, _i, _vp, _vp, _vp]),
    "orc_sweep_agg_f64": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _f, _i, _vp, _vp, _vp]),
    "orc_tgv_f32": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _d, _vp,
                         _vp, _vp]),
    "orc_tgv_f64": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _d, _vp,
                         _vp, _vp]),
    "orc_sweep_agg_volume_f64": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _f, _i, _vp,
                                      _vp, _vp, _vp]),
    "orc_sweep_agg_volume_f32": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _f, _i, _vp,
                                      _vp, _vp, _vp]),
    "orc_reduce_best2_f64": (None, [_vp, _i, _vp, _vp, _vp]),
    "orc_tgv_linearise_f64": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _vp, _vp, _vp, _vp]),
    "orc_tgv_iterate_f32": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _d]),
    "orc_tgv_iterate_f64": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _d]),
    "orc_tensor_f32": (None, [_vp, _i, _i, _f, _f, _vp, _vp, _vp, _vp]),
    "orc_tgv_linearise_f32": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _vp, _vp, _vp, _vp]),
    "orc_tgv_energy": (_d, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _d]),
    "orc_tgv_energy_f64": (_d, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _d]),
    "orc_tgv_energy_masked_f64": (_d, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _d, _vp]),
    "orc_fusion_bin_centers": (None, [_i, _vp]),
    "orc_fusion_hist_f32": (None, [_vp, _i, _i, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _f]),
    "orc_fusion_hist_f64": (None, [_vp, _i, _i, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _f]),
    "orc_fusion_u_f32": (None, [_vp, _vp, _vp, _vp, _i, _i, _i, _d, _d]),
    "orc_fusion_u_f64": (None, [_vp, _vp, _vp, _vp, _i, _i, _i, _d, _d]),
    "orc_fusion_p_f32": (None, [_vp, _vp, _i, _i, _i, _d]),
    "orc_fusion_p_f64": (None, [_vp, _vp, _i, _i, _i, _d]),
    "orc_fusion_energy_f64": (_d, [_vp, _vp, _i, _i, _i, _d]),
    "orc_prox_hist_insertion_f32": (_f, [_f, _vp, _f, _f]),
    "orc_surface_count_f32": (_i64, [_vp, _i64]),
    "orc_surface_count_f64": (_i64, [_vp, _i64]),
    "orc_tvl1_f32": (_i, [_vp, _vp, _i, _i, _f, _f, _i, _d, _d, _d, _d, _d, _d, _vp]),
    "orc_tvl1_f64": (_i, [_vp, _vp, _i, _i, _f, _f, _i, _d, _d, _d, _d, _d, _d, _vp]),
    "orc_tgv_sparse_f32": (_i, [_vp, _vp, _vp, _i, _i, _f, _i, _d, _d, _d, _d, _d, _d, _d, _vp, _vp]),
    "orc_tgv_sparse_f64": (_i, [_vp, _vp, _vp, _i, _i, _f, _i, _d, _d, _d, _d, _d, _d, _d, _vp, _vp]),
    "orc_compute3d_f32": (None, [_vp, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "orc_compute3d_f64": (None, [_vp, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp]),
}
for _n, (_r, _a) in _SIGS.items():
    _fn = getattr(_lib, _n)
    _fn.restype = _r
    _fn.argtypes = _a


def _c(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _ft(prec):
    return np.float32 if prec == "f32" else np.float64


def reference_planes(znear, zfar, nplanes):
    out = np.zeros(int(nplanes) + 2, np.float32)
    n = _lib.orc_reference_planes(znear, zfar, nplanes, _p(out), out.size)
    return out[:n].copy()


def mat3_inv(m):
    m = _c(m).ravel()
    o = np.zeros(9, np.float32)
    _lib.orc_mat3_inv_f32(_p(m), _p(o))
    return o.reshape(3, 3)


def relative_matrices(Rref, tref, Rsrc, tsrc, alternative=0):
    a = [_c(x).ravel() for x in (Rref, tref, Rsrc, tsrc)]
    R, t = np.zeros(9, np.float32), np.zeros(3, np.float32)
    _lib.orc_relative_matrices_f32(*[_p(x) for x in a], alternative, _p(R), _p(t))
    return R.reshape(3, 3), t


def homography_ref(K, Rref, tref, Rsrc, tsrc, depth, alternative=0):
    a = [_c(x).ravel() for x in (K, Rref, tref, Rsrc, tsrc)]
    H = np.zeros(9, np.float32)
    _lib.orc_homography_ref_f32(*[_p(x) for x in a], alternative, float(depth), _p(H))
    return H.reshape(3, 3)


def homography_f64(K, Rref, tref, Rsrc, tsrc, depth, alternative=0):
    a = [_c(x).ravel() for x in (K, Rref, tref, Rsrc, tsrc)]
    H = np.zeros(9, np.float64)
    _lib.orc_homography_f64(*[_p(x) for x in a], alternative, float(depth), _p(H))
    return H.reshape(3, 3)


def windowed_mean(img, winsize, squared, axis):
    img = _c(img)
    h, w = img.shape
    out = np.zeros_like(img)
    fn = _lib.orc_windowed_mean_row_f32 if axis == "row" else _lib.orc_windowed_mean_col_f32
    fn(_p(out), _p(img), int(winsize), int(bool(squared)), w, h)
    return out


def bilinear(data, xs, ys):
    data, xs, ys = _c(data), _c(xs), _c(ys)
    out = np.zeros_like(xs)
    _lib.orc_bilinear_f32(_p(out), _p(data), _p(xs), _p(ys), data.shape[1], data.shape[0], xs.shape[1], xs.shape[0])
    return out


def transform_indexes(H, w, h):
    H = _c(H).ravel()
    xs, ys = np.zeros((h, w), np.float32), np.zeros((h, w), np.float32)
    _lib.orc_transform_indexes_f32(_p(xs), _p(ys), _p(H), w, h)
    return xs, ys


def sweep_ref(ref, src, K, Rref, tref, Rsrc, tsrc, planes, winsize, stdthresh, nccthresh, zfar, alternative=0,
              prec="f32", per_source=False):
    ref, src, planes = _c(ref), _c(src), _c(planes)
    Hh, W = ref.shape
    N = src.shape[0]
    T = _ft(prec)
    out = np.zeros((Hh, W), T)
    bn = np.zeros((N, Hh, W), T) if per_source else None
    bd = np.zeros((N, Hh, W), T) if per_source else None
    fn = _lib.orc_sweep_ref_f32 if prec == "f32" else _lib.orc_sweep_ref_f64
    pose = [_c(x).ravel() for x in (K, Rref, tref, Rsrc, tsrc)]
    rc = fn(_p(ref), _p(src), W, Hh, N, *[_p(x) for x in pose], alternative, _p(planes), planes.size, winsize,
            stdthresh, nccthresh, zfar, _p(out), _p(bn), _p(bd))
    assert rc == 0, rc
    return (out, bn, bd) if per_source else out


def sweep_agg(ref, src, K, Rref, tref, Rsrc, tsrc, planes, winsize, stdthresh, coord_offset=0, alternative=0,
              prec="f32"):
    ref, src, planes = _c(ref), _c(src), _c(planes)
    Hh, W = ref.shape
    N = src.shape[0]
    T = _ft(prec)
    arg = np.zeros((Hh, W), np.int32)
    cmin, csec = np.zeros((Hh, W), T), np.zeros((Hh, W), T)
    fn = _lib.orc_sweep_agg_f32 if prec == "f32" else _lib.orc_sweep_agg_f64
    pose = [_c(x).ravel() for x in (K, Rref, tref, Rsrc, tsrc)]
    rc = fn(_p(ref), _p(src), W, Hh, N, *[_p(x) for x in pose], alternative, _p(planes), planes.size, winsize,
            stdthresh, coord_offset, _p(arg), _p(cmin), _p(csec))
    assert rc == 0, rc
    return arg, cmin, csec


def sweep_agg_volume(ref, src, K, Rref, tref, Rsrc, tsrc, planes, winsize, stdthresh, coord_offset=0,
                     alternative=0, prec="f64"):
    """Aggregated sweep that also returns the cost of every plane:
    (argmin, min, second, volume[P][H][W]), in FP64 or in the reference's
    float arithmetic (prec "f32")."""
    ref, src, planes = _c(ref), _c(src), _c(planes)
    Hh, W = ref.shape
    N = src.shape[0]
    T = _ft(prec)
    arg = np.zeros((Hh, W), np.int32)
    cmin, csec = np.zeros((Hh, W), T), np.zeros((Hh, W), T)
    vol = np.zeros((planes.size, Hh, W), T)
    pose = [_c(x).ravel() for x in (K, Rref, tref, Rsrc, tsrc)]
    fn = _lib.orc_sweep_agg_volume_f64 if prec == "f64" else _lib.orc_sweep_agg_volume_f32
    rc = fn(_p(ref), _p(src), W, Hh, N, *[_p(x) for x in pose], alternative, _p(planes), planes.size, winsize,
            stdthresh, coord_offset, _p(arg), _p(cmin), _p(csec), _p(vol))
    assert rc == 0, rc
    return arg, cmin, csec, vol


def reduce_best2(costs):
    """The aggregated sweep's reduction over one pixel's costs in plane order:
    (argmin, min, second-smallest)."""
    c = _c(costs, np.float64).ravel()
    a, m, s = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
    _lib.orc_reduce_best2_f64(_p(c), c.size, ctypes.byref(a), ctypes.byref(m), ctypes.byref(s))
    return a.value, m.value, s.value


TGV_STATE_NAMES = ["u", "u1x", "u1y", "ubar", "u1xbar", "u1ybar", "px", "py", "qx", "qy", "qz", "qw"]


def tgv_iterate(state, T4, u0, It, Iu, niters, lambda_=1.0, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0,
                prec="f64"):
    """niters TGV2 iterations from a given state inside a given linearisation.
    state: dict with the planes of TGV_STATE_NAMES and 'r' (N, H, W); missing
    planes start at 0, except ubar/u1xbar/u1ybar (default u, u1x, u1y) and qw
    (default qz).  T4 = (T11, T12, T21, T22).  Returns the new state dict."""
    T = _ft(prec)
    It, Iu = _c(It, T), _c(Iu, T)
    N, Hh, W = It.shape
    dflt = {"ubar": "u", "u1xbar": "u1x", "u1ybar": "u1y", "qw": "qz"}
    planes = []
    for k in TGV_STATE_NAMES:
        v = state.get(k)
        if v is None and k in dflt:
            v = state.get(dflt[k])
        planes.append(np.zeros((Hh, W), T) if v is None else _c(v, T).reshape(Hh, W))
    r = state.get("r")
    planes.append(np.zeros((N, Hh, W), T) if r is None else _c(r, T).reshape(N, Hh, W))
    st = np.ascontiguousarray(np.concatenate([p.reshape(-1) for p in planes]))
    Tt, u0 = _c(T4, T), _c(u0, T)
    fn = _lib.orc_tgv_iterate_f32 if prec == "f32" else _lib.orc_tgv_iterate_f64
    rc = fn(W, Hh, N, _p(Tt), _p(u0), _p(It), _p(Iu), _p(st), int(niters), lambda_, alpha0, alpha1, tau, sigma)
    assert rc == 0, rc
    n = W * Hh
    out = {k: st[i * n:(i + 1) * n].reshape(Hh, W) for i, k in enumerate(TGV_STATE_NAMES)}
    out["r"] = st[12 * n:].reshape(N, Hh, W)
    return out


def tgv(ref, src, K, Rref, tref, Rsrc, tsrc, niters, warps, lambda_=0.5, alpha0=2.0, alpha1=1.5, tau=0.02,
        sigma=1.0, beta=0.0, gamma=1.0, u_init=None, alternative=0, prec="f32", state=False):
    ref, src = _c(ref), _c(src)
    Hh, W = ref.shape
    N = src.shape[0]
    T = _ft(prec)
    u = np.zeros((Hh, W), T)
    st = np.zeros((8 + 3 * N, Hh, W), T) if state else None
    ui = _c(u_init) if u_init is not None else None
    fn = _lib.orc_tgv_f32 if prec == "f32" else _lib.orc_tgv_f64
    pose = [_c(x).ravel() for x in (K, Rref, tref, Rsrc, tsrc)]
    rc = fn(_p(ref), _p(src), W, Hh, N, *[_p(x) for x in pose], alternative, niters, warps, lambda_, alpha0, alpha1,
            tau, sigma, beta, gamma, _p(ui), _p(u), _p(st))
    assert rc == 0, rc
    if not state:
        return u
    names = ["u", "u1x", "u1y", "px", "py", "qx", "qy", "qz"]
    d = {k: st[i] for i, k in enumerate(names)}
    d["It"], d["Iu"], d["r"] = st[8:8 + N], st[8 + N:8 + 2 * N], st[8 + 2 * N:8 + 3 * N]
    return u, d


def tensor(img, beta, gamma):
    img = _c(img)
    Hh, W = img.shape
    T = [np.zeros_like(img) for _ in range(4)]
    _lib.orc_tensor_f32(_p(img), W, Hh, beta, gamma, *[_p(t) for t in T])
    return T


def tgv_linearise(ref, src, K, Rref, tref, Rsrc, tsrc, u0, beta=0.0, gamma=1.0, alternative=0, prec="f32"):
    """It (N,H,W), Iu (N,H,W) and T (T11, T12, T21, T22) of one warp at u0."""
    ref, src, u0 = _c(ref), _c(src), _c(u0)
    Hh, W = ref.shape
    N = src.shape[0]
    T = _ft(prec)
    It, Iu, Tt = np.zeros((N, Hh, W), T), np.zeros((N, Hh, W), T), np.zeros((4, Hh, W), T)
    pose = [_c(x).ravel() for x in (K, Rref, tref, Rsrc, tsrc)]
    fn = _lib.orc_tgv_linearise_f32 if prec == "f32" else _lib.orc_tgv_linearise_f64
    rc = fn(_p(ref), _p(src), W, Hh, N, *[_p(x) for x in pose], alternative, beta, gamma,
            _p(u0), _p(It), _p(Iu), _p(Tt))
    assert rc == 0
    return It, Iu, Tt


def tgv_energy(u, u1x, u1y, u0, T4, It, Iu, lambda_, alpha0, alpha1, weight=None):
    """T4 = [T11, T12, T21, T22] planes.  weight (H, W), optional: pixels
    where it is 0 are left out (FP64 sum)."""
    if weight is not None:
        a = [_c(x, np.float64) for x in (u, u1x, u1y, u0, T4, It, Iu)]
        w = _c(weight)
        Hh, W = a[0].shape
        return _lib.orc_tgv_energy_masked_f64(*[_p(x) for x in a], W, Hh, a[5].shape[0], lambda_, alpha0, alpha1,
                                              _p(w))
    if np.asarray(u).dtype == np.float64:
        a = [_c(x, np.float64) for x in (u, u1x, u1y, u0, T4, It, Iu)]
        fn = _lib.orc_tgv_energy_f64
    else:
        a = [_c(x) for x in (u, u1x, u1y, u0, T4, It, Iu)]
        fn = _lib.orc_tgv_energy
    Hh, W = a[0].shape
    N = a[5].shape[0]
    return fn(*[_p(x) for x in a], W, Hh, N, lambda_, alpha0, alpha1)


def bin_centers(bins=8):
    o = np.zeros(bins, np.float32)
    _lib.orc_fusion_bin_centers(bins, _p(o))
    return o


def fusion_hist(hist, vol, depth, K, R, T, threshold, prec="f32"):
    """hist: (d, h, w, 8) uint16, updated in place."""
    assert hist.dtype == np.uint16 and hist.flags.c_contiguous
    assert not threshold < 0, "negative threshold: the reference's bin index is undefined"
    gd, gh, gw, _ = hist.shape
    depth = _c(depth)
    fn = _lib.orc_fusion_hist_f32 if prec == "f32" else _lib.orc_fusion_hist_f64
    a = [_c(x).ravel() for x in (vol, K, R, T)]
    fn(_p(hist), gw, gh, gd, _p(a[0]), _p(depth), depth.shape[1], depth.shape[0], _p(a[1]), _p(a[2]), _p(a[3]),
       float(threshold))
    return hist


def fusion_u(u, v, p, hist, tau, lambda_, prec="f32"):
    """u, v: (d, h, w); p: (3, d, h, w); updated in place."""
    gd, gh, gw = u.shape
    fn = _lib.orc_fusion_u_f32 if prec == "f32" else _lib.orc_fusion_u_f64
    fn(_p(u), _p(v), _p(p), _p(hist), gw, gh, gd, tau, lambda_)


def fusion_p(p, v, sigma, prec="f32"):
    gd, gh, gw = v.shape
    fn = _lib.orc_fusion_p_f32 if prec == "f32" else _lib.orc_fusion_p_f64
    fn(_p(p), _p(v), gw, gh, gd, sigma)


def fusion_energy(u, hist, lambda_):
    """FP64 TV-L1 histogram energy of u (gd, gh, gw) with hist (gd, gh, gw, 8)."""
    u = np.ascontiguousarray(u, np.float64)
    hist = np.ascontiguousarray(hist, np.uint16)
    gd, gh, gw = u.shape
    return _lib.orc_fusion_energy_f64(_p(u), _p(hist), gw, gh, gd, lambda_)


def prox_hist_insertion(u, W, tau, lambda_):
    W = _c(W, np.int32)
    return _lib.orc_prox_hist_insertion_f32(u, _p(W), tau, lambda_)


def surface_count(u):
    u = np.ascontiguousarray(u)
    if u.dtype == np.float64:
        return _lib.orc_surface_count_f64(_p(u), u.size)
    u = _c(u)
    return _lib.orc_surface_count_f32(_p(u), u.size)


def tvl1(depth, ref, znear, zfar, niters=100, lambda_=0.3, tau=0.02, sigma=6.0, theta=1.0, beta=0.0, gamma=1.0,
         prec="f32"):
    depth, ref = _c(depth), _c(ref)
    Hh, W = depth.shape
    out = np.zeros((Hh, W), _ft(prec))
    fn = _lib.orc_tvl1_f32 if prec == "f32" else _lib.orc_tvl1_f64
    rc = fn(_p(depth), _p(ref), W, Hh, znear, zfar, niters, lambda_, tau, sigma, theta, beta, gamma, _p(out))
    assert rc == 0
    return out


def tgv_sparse(ref, dense, sparse, zfar, niters, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, theta=1.0, beta=0.0,
               gamma=1.0, prec="f32", u1=False):
    """PlaneSweep::TGVdenoiseFromSparse (src/planesweep.cpp:596-674)."""
    ref, dense, sparse = _c(ref), _c(dense), _c(sparse)
    Hh, W = ref.shape
    T = _ft(prec)
    out = np.zeros((Hh, W), T)
    v = np.zeros((2, Hh, W), T) if u1 else None
    fn = _lib.orc_tgv_sparse_f32 if prec == "f32" else _lib.orc_tgv_sparse_f64
    rc = fn(_p(ref), _p(dense), _p(sparse), W, Hh, zfar, niters, alpha0, alpha1, tau, sigma, theta, beta, gamma,
            _p(out), _p(v))
    assert rc == 0, rc
    return (out, v) if u1 else out


def compute3d(depth, K, Rref, tref, alternative=0, prec="f32"):
    """PlaneSweep::get3Dcoordinates (src/planesweep.cpp:551-586): X, Y, Z."""
    depth = _c(depth)
    Hh, W = depth.shape
    T = _ft(prec)
    out = [np.zeros((Hh, W), T) for _ in range(3)]
    fn = _lib.orc_compute3d_f32 if prec == "f32" else _lib.orc_compute3d_f64
    fn(_p(depth), W, Hh, *[_p(_c(x).ravel()) for x in (K, Rref, tref)], alternative, *[_p(o) for o in out])
    return out
