"""Sparse-depth TGV2 (PlaneSweep::TGVdenoiseFromSparse, src/planesweep.cpp:
596-674) and the 3-D coordinates of a depthmap (PlaneSweep::get3Dcoordinates,
src/planesweep.cpp:551-586, compute3D_kernel src/kernels.cu:405-429).

CPU: the oracle pinned by what the mathematics fixes -- a constant depth is a
fixed point; without regularisation the data step is the closed-form prox
(w = 1) or the identity (w = 0); without sparse data the first iteration is
the dense TGV2 step with lambda = 0 (pinned by the dense oracle's own tests);
the reference driver's u1/u1bar aliasing makes theta inert; back-projected
points reproject to their pixels.
GPU: the launcher sequence within 1e-6 of the FP32 oracle, the fused solver
within the TGV2 bar of the FP64 oracle, compute3D bit-exact."""
import numpy as np
import pytest

import oracle
from conftest import one_based

KW = dict(alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, theta=1.0, beta=0.0, gamma=1.0)


def _scene(seed=3, W=96, H=72, frac=0.05):
    from psw_amd import scene
    v = scene.make_views(seed=seed, width=W, height=H, nsrc=1)
    rng = np.random.default_rng(seed)
    gt = v["depth"].astype(np.float32)
    dense = (gt * (1 + 0.05 * rng.standard_normal(gt.shape))).astype(np.float32)
    sparse = np.where(rng.random(gt.shape) < frac, gt, 0).astype(np.float32)
    sparse[rng.random(gt.shape) < 0.01] = -1.0          # invalid samples (weight 0)
    return v, dense, sparse


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_constant_depth_is_fixed(prec):
    H, W = 20, 28
    ref = np.random.default_rng(0).uniform(0, 255, (H, W)).astype(np.float32)
    c = np.full((H, W), 2.5, np.float32)
    u = oracle.tgv_sparse(ref, c, c, 4.0, 200, prec=prec, **KW)
    np.testing.assert_array_equal(u, c)


def test_data_step_closed_form():
    """alpha0 = alpha1 = 0: u_k = Ds + (u_0 - Ds) / (1 + tau)^k where the
    sparse depth is > 0, u unchanged elsewhere (sparse 0, negative)."""
    H, W, zf, tau = 16, 20, 4.0, 0.1
    rng = np.random.default_rng(1)
    dense = rng.uniform(1, 3, (H, W)).astype(np.float32)
    sparse = rng.uniform(1, 3, (H, W)).astype(np.float32)
    sparse[rng.random((H, W)) < 0.3] = 0.0
    sparse[rng.random((H, W)) < 0.1] = -2.0
    w = sparse > 0
    ref = np.full((H, W), 50.0, np.float32)
    for k in (1, 7, 40):
        u = oracle.tgv_sparse(ref, dense, sparse, zf, k, alpha0=0.0, alpha1=0.0, tau=tau, prec="f64")
        ds, u0 = sparse.astype(np.float64) * (1 / zf), dense.astype(np.float64) * (1 / zf)
        exp = (ds + (u0 - ds) / (1 + tau) ** k) * zf
        np.testing.assert_allclose(u[w], exp[w], rtol=1e-12)
        np.testing.assert_array_equal(u[~w], (u0 * zf)[~w])


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_first_step_is_dense_step_without_data(prec):
    """No sparse samples and zfar = 1 (exact scaling): one iteration equals
    the dense TGV2 oracle's first iteration with lambda = 0 from the same u,
    bit for bit, in u and u1."""
    H, W = 18, 24
    rng = np.random.default_rng(2)
    ref = rng.uniform(0, 255, (H, W)).astype(np.float32)
    dense = rng.uniform(0.5, 3, (H, W)).astype(np.float32)
    kw = dict(KW, beta=1.5, gamma=0.8)
    u, u1 = oracle.tgv_sparse(ref, dense, np.zeros_like(dense), 1.0, 1, prec=prec, u1=True, **kw)
    K = np.array([[100, 0, 12], [0, 100, 9], [0, 0, 1]], np.float32)
    eye, z = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    dk = {k: v for k, v in kw.items() if k != "theta"}
    ud, st = oracle.tgv(ref, ref[None], K, eye, z, eye[None], z[None], 1, 1, lambda_=0.0, u_init=dense, prec=prec,
                        state=True, **dk)
    np.testing.assert_array_equal(u, ud)
    np.testing.assert_array_equal(u1[0], st["u1x"])
    np.testing.assert_array_equal(u1[1], st["u1y"])
    assert np.abs(u - dense).max() > 0 and np.abs(u1).max() > 0    # the step did something


def test_theta_has_no_effect():
    """The reference driver passes vxbar / vybar as both u1 and u1bar, and P
    reads u, so the over-relaxation parameter never reaches the result."""
    v, dense, sparse = _scene(W=48, H=36)
    outs = [oracle.tgv_sparse(v["ref"], dense, sparse, 8.0, 30, **dict(KW, theta=t)) for t in (0.0, 0.5, 1.0)]
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], outs[2])


def test_fp32_close_to_fp64():
    v, dense, sparse = _scene(W=64, H=48)
    u32 = oracle.tgv_sparse(v["ref"], dense, sparse, 8.0, 100, **KW)
    u64 = oracle.tgv_sparse(v["ref"], dense, sparse, 8.0, 100, prec="f64", **KW)
    assert np.max(np.abs(u32 - u64) / np.abs(u64)) <= 1e-5


def _pose(seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.3, 0.3, 3)
    cx, sx, cy, sy, cz, sz = np.cos(a[0]), np.sin(a[0]), np.cos(a[1]), np.sin(a[1]), np.cos(a[2]), np.sin(a[2])
    R = (np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]]) @ np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]]) @
         np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]]))
    return R.astype(np.float32), rng.uniform(-1, 1, 3).astype(np.float32)


@pytest.mark.parametrize("alternative", [0, 1])
def test_compute3d_reprojects(alternative):
    """Default method (X_cam = R X_w + t): K (R X + t) is z (x+1, y+1, 1);
    alternative method (R, t camera -> world): X = R z K^-1 p + t.  Depths
    from 5.5 up give z = -9 and x = y = 0."""
    H, W = 12, 16
    K = np.array([[120.0, 0.5, 8.5], [0, 118.0, 6.5], [0, 0, 1]], np.float32)
    R, t = _pose(4)
    rng = np.random.default_rng(5)
    depth = rng.uniform(0.5, 5.0, (H, W)).astype(np.float32)
    far = rng.random((H, W)) < 0.2
    depth[far] = rng.uniform(5.5, 9.0, far.sum())
    X, Y, Z = oracle.compute3d(depth, K, R, t, alternative, prec="f64")
    P = np.stack([X, Y, Z], -1)[~far]
    ys, xs = np.nonzero(~far)
    pix = np.stack([xs + 1.0, ys + 1.0, np.ones_like(xs, float)], -1)
    z = depth[~far].astype(np.float64)
    Kd, Rd, td = K.astype(np.float64), R.astype(np.float64), t.astype(np.float64)
    if alternative == 0:
        proj = (Kd @ (P @ Rd.T + td).T).T
        np.testing.assert_allclose(proj, pix * z[:, None], rtol=1e-10, atol=1e-10)
    else:
        exp = (z[:, None] * (np.linalg.inv(Kd) @ pix.T).T) @ Rd.T + td
        np.testing.assert_allclose(P, exp, rtol=1e-10, atol=1e-10)
    assert np.all(Z[far] == -9) and np.all(X[far] == 0) and np.all(Y[far] == 0)
    X32, Y32, Z32 = oracle.compute3d(depth, K, R, t, alternative)
    for a, b in ((X32, X), (Y32, Y), (Z32, Z)):
        assert np.max(np.abs(a - b)) <= 1e-5 * max(1.0, np.abs(b).max())


# ---- GPU -------------------------------------------------------------------------------------------------------

@pytest.mark.gpu
def test_sparse_launchers_vs_oracle(gpu):
    import refdriver
    v, dense, sparse = _scene()
    got = refdriver.tgv_sparse(v["ref"], dense, sparse, 8.0, 40, **KW)
    exp = oracle.tgv_sparse(v["ref"], dense, sparse, 8.0, 40, **KW)
    err = np.abs(got - exp) / np.abs(exp)
    print(f"sparse launchers vs f32 oracle: max rel {err.max():.3e}, identical {np.mean(got == exp):.4f}")
    assert err.max() <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("W,H,niters", [(96, 72, 100), (161, 83, 37), (640, 480, 50)])
def test_sparse_fused_vs_fp64(gpu, W, H, niters):
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    v, dense, sparse = _scene(W=W, H=H)
    kw = dict(KW, beta=1.0, gamma=1.0)
    u = stages.tgv_sparse(v["ref"], dense, sparse, 8.0, niters, **kw)
    u64 = oracle.tgv_sparse(v["ref"], dense, sparse, 8.0, niters, prec="f64", **kw)
    e = np.abs(u - u64) / np.abs(u64)
    print(f"sparse fused {W}x{H} x{niters} vs FP64: max {e.max():.3e} p99.9 {np.percentile(e, 99.9):.3e}")
    assert np.percentile(e, 99.9) <= 1e-4 and e.max() <= 1e-3
    ws = DeviceArray((stages.tgv_sparse_workspace_bytes(W, H) // 4,))
    ws.fill(np.nan)                                  # the set-up must overwrite every slot it reads
    np.testing.assert_array_equal(stages.tgv_sparse(v["ref"], dense, sparse, 8.0, niters, workspace=ws, **kw), u)


@pytest.mark.gpu
def test_sparse_edge_cases(gpu):
    from psw_amd import stages
    from psw_amd._lib import PswError
    from psw_amd.device import DeviceArray
    v, dense, sparse = _scene(W=40, H=30)
    u0 = stages.tgv_sparse(v["ref"], dense, sparse, 8.0, 0, **KW)
    exp0 = oracle.tgv_sparse(v["ref"], dense, sparse, 8.0, 0, **KW)
    np.testing.assert_array_equal(u0, exp0)
    u1 = stages.tgv_sparse(v["ref"][:1, :5], dense[:1, :5], sparse[:1, :5], 8.0, 3, **KW)   # one row
    assert np.all(np.isfinite(u1))
    with pytest.raises(PswError):
        stages.tgv_sparse(v["ref"], dense, sparse, 0.0, 3, **KW)
    small = DeviceArray((16,))
    with pytest.raises(PswError):
        stages.tgv_sparse(v["ref"], dense, sparse, 8.0, 3, workspace=small, **KW)


@pytest.mark.gpu
@pytest.mark.parametrize("alternative", [0, 1])
def test_compute3d_bit_exact(gpu, alternative):
    import refdriver
    H, W = 48, 64
    K = one_based(np.array([[60.0, 0, 31.5], [0, 60.0, 23.5], [0, 0, 1]], np.float32))
    R, t = _pose(7)
    rng = np.random.default_rng(8)
    depth = rng.uniform(0.3, 7.0, (H, W)).astype(np.float32)
    got = refdriver.compute3d(depth, K, R, t, alternative)
    exp = oracle.compute3d(depth, K, R, t, alternative)
    for g, e in zip(got, exp):
        np.testing.assert_array_equal(g, e)


@pytest.mark.gpu
@pytest.mark.parametrize("niters", [25, 26, 1, 2])
def test_sparse_launch_shapes_identical(gpu, monkeypatch, niters):
    """The tile shape (PSW_TGV_SPARSE_VARIANT, PSW_TGV_SPARSE_NARROW) and two
    to six iterations per launch (PSW_TGV_SPARSE_DOUBLE, PSW_TGV_SPARSE_M)
    never change results, for odd and even counts."""
    from psw_amd import stages
    v, dense, sparse = _scene(W=161, H=83)
    outs = []
    for dbl, var, narrow, m in (("0", "1", "1", "2"), ("0", "0", "1", "2"), ("0", "2", "1", "2"),
                                ("0", "3", "1", "2"), ("1", "1", "1", "2"), ("1", "0", "1", "2"),
                                ("1", "1", "0", "2"), ("1", "1", "1", "3"), ("1", "1", "1", "4"),
                                ("1", "1", "1", "5"), ("1", "1", "1", "6")):
        monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT", "0")
        monkeypatch.setenv("PSW_TGV_SPARSE_DOUBLE", dbl)
        monkeypatch.setenv("PSW_TGV_SPARSE_VARIANT", var)
        monkeypatch.setenv("PSW_TGV_SPARSE_NARROW", narrow)
        monkeypatch.setenv("PSW_TGV_SPARSE_M", m)
        outs.append(stages.tgv_sparse(v["ref"], dense, sparse, 8.0, niters, **KW))
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def _random_case(rng, H, W, frac=0.05):
    ref = rng.uniform(0, 255, (H, W)).astype(np.float32)
    dense = rng.uniform(0.5, 6.0, (H, W)).astype(np.float32)
    sparse = np.where(rng.random((H, W)) < frac, dense * (1 + 0.05 * rng.standard_normal((H, W))), 0)
    sparse = sparse.astype(np.float32)
    sparse[rng.random((H, W)) < 0.01] = -1.0
    return ref, dense, sparse


@pytest.mark.gpu
@pytest.mark.parametrize("niters", [2, 3, 25])
def test_sparse_resident_matches_fused(gpu, monkeypatch, niters):
    """The resident solver (all iterations in one cooperative launch of 64x20
    tiles; PSW_TGV_SPARSE_RESIDENT=2 makes it an error if it does not run)
    against the fused launches: bit-identical on the 96x72 scene and on sizes
    that cut the tiles (161x83, 65x9, 3x2, 200x41) and at 640x480."""
    from psw_amd import stages
    rng = np.random.default_rng(300 + niters)
    v, dense, sparse = _scene()
    cases = [(v["ref"], dense, sparse)]
    for H, W in [(83, 161), (9, 65), (2, 3), (41, 200), (480, 640)]:
        cases.append(_random_case(rng, H, W))
    for ref, dense, sparse in cases:
        print(f"[case] sparse resident {ref.shape[1]}x{ref.shape[0]} niters={niters}", flush=True)
        monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT", "0")
        fused = stages.tgv_sparse(ref, dense, sparse, 8.0, niters, **KW)
        monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT", "2")
        res = stages.tgv_sparse(ref, dense, sparse, 8.0, niters, **KW)
        np.testing.assert_array_equal(res.view(np.uint32), fused.view(np.uint32))


@pytest.mark.gpu
def test_sparse_resident_fallback_and_failure(gpu, monkeypatch):
    """Frames above 256 tiles (1242x375: 380) use the fused launches, with the
    same bits; a wait that runs out (a zero poll budget, forced from the host)
    is reported as an error; and the next run is whole again."""
    from psw_amd import stages
    from psw_amd._lib import PswError
    rng = np.random.default_rng(11)
    ref, dense, sparse = _random_case(rng, 375, 1242)
    print("[case] sparse 1242x375 resident on (falls back) vs fused", flush=True)
    monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT", "0")
    fused = stages.tgv_sparse(ref, dense, sparse, 8.0, 6, **KW)
    monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT", "1")
    got = stages.tgv_sparse(ref, dense, sparse, 8.0, 6, **KW)
    np.testing.assert_array_equal(got.view(np.uint32), fused.view(np.uint32))
    monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT", "2")
    with pytest.raises(PswError):
        stages.tgv_sparse(ref, dense, sparse, 8.0, 6, **KW)
    v, dense, sparse = _scene()
    print("[case] sparse 96x72 resident with a zero poll budget", flush=True)
    monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT_SPIN", "0")
    with pytest.raises(PswError) as e:
        stages.tgv_sparse(v["ref"], dense, sparse, 8.0, 5, **KW)
    assert "resident sparse-depth TGV2 solver" in str(e.value)
    monkeypatch.delenv("PSW_TGV_SPARSE_RESIDENT_SPIN")
    monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT", "0")
    fused = stages.tgv_sparse(v["ref"], dense, sparse, 8.0, 5, **KW)
    monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT", "2")
    got = stages.tgv_sparse(v["ref"], dense, sparse, 8.0, 5, **KW)
    np.testing.assert_array_equal(got.view(np.uint32), fused.view(np.uint32))


@pytest.mark.gpu
def test_sparse_resident_concurrent_calls(gpu, monkeypatch):
    """Two threads densifying different frames on their own streams at once
    (the resident runs of one device take turns on its exchange buffer): the
    same bits as each frame alone."""
    import threading
    from psw_amd import stages
    from psw_amd.device import DeviceArray, Stream
    monkeypatch.setenv("PSW_TGV_SPARSE_RESIDENT", "2")
    rng = np.random.default_rng(9)
    frames = [_random_case(rng, 83, 161) for _ in range(2)]
    alone = [stages.tgv_sparse(ref, dense, sparse, 8.0, 100, **KW) for ref, dense, sparse in frames]
    dev = [([DeviceArray.from_host(x) for x in f], DeviceArray((stages.tgv_sparse_workspace_bytes(161, 83) // 4,)),
            Stream()) for f in frames]
    errors = []

    def run(k):
        (d_ref, d_dense, d_sparse), ws, st = dev[k]
        try:
            for _ in range(5):
                out = stages.tgv_sparse(d_ref, d_dense, d_sparse, 8.0, 100, workspace=ws, stream=st.handle, **KW)
                st.synchronize()
                if not np.array_equal(out.to_host().view(np.uint32), alone[k].view(np.uint32)):
                    errors.append(k)
        except Exception as e:   # noqa: BLE001  (reported below)
            errors.append(repr(e))

    print("[case] sparse resident, two threads at once", flush=True)
    th = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th)
    assert not errors, errors

