"""Known-answer tests that pin the CPU oracle (no GPU).

The reference ships no tests or golden vectors, so the oracle is checked
against what the mathematics fixes: brute-force box filters, bilinear
interpolation at lattice points, identity-pose and fronto-parallel-plane
sweeps, TGV2 fixed points and invariants, the closed form of proxHist, hand
computed histogram bins and an independent float64 histogram, TVL1 on a
constant depth.  Each test is chosen so that a plausible slip in the oracle
(dropped term, wrong sign, transposed operand, off-by-one border) fails it.
"""
import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(1234)


# ---- sweep primitives ---------------------------------------------------------------

def _box_bruteforce(img, win, squared, axis):
    n = win // 2
    x = img.astype(np.float64) ** 2 if squared else img.astype(np.float64)
    pad = ((n, n), (0, 0)) if axis == "col" else ((0, 0), (n, n))
    p = np.pad(x, pad, mode="reflect")          # reflect-101, as kernels.cu:203-204
    out = np.zeros_like(x)
    for k in range(win):
        out += p[k:k + img.shape[0], :] if axis == "col" else p[:, k:k + img.shape[1]]
    return out


@pytest.mark.parametrize("win", [1, 3, 5, 9])
@pytest.mark.parametrize("axis", ["row", "col"])
@pytest.mark.parametrize("squared", [False, True])
def test_box_mean_exact_on_integers(win, axis, squared):
    img = RNG.integers(0, 256, size=(23, 31)).astype(np.float32)
    got = oracle.windowed_mean(img, win, squared, axis)
    # integer sums are exact in float; the single division rounds the same way
    exp = (_box_bruteforce(img, win, squared, axis).astype(np.float32) / np.float32(win))
    np.testing.assert_array_equal(got, exp)


def test_box_mean_reflect_is_101_not_edge():
    img = np.zeros((5, 6), np.float32)
    img[:, 1] = 10.0                               # reflect-101 mirrors column 1 into column -1
    got = oracle.windowed_mean(img, 3, False, "row")
    np.testing.assert_allclose(got[:, 0], (10 + 0 + 10) / 3, rtol=1e-6)


def test_bilinear_lattice_bounds_and_interior():
    d = RNG.uniform(0, 255, size=(12, 17)).astype(np.float32)
    ys, xs = np.mgrid[0:11, 0:16].astype(np.float32)
    np.testing.assert_array_equal(oracle.bilinear(d, xs, ys), d[:11, :16])     # lattice points
    edge = np.array([[16.0, -0.5, 3.0, np.nan, 1e30]], np.float32)
    yv = np.array([[2.0, 2.0, 11.0, 2.0, 2.0]], np.float32)
    np.testing.assert_array_equal(oracle.bilinear(d, edge, yv), np.zeros((1, 5), np.float32))
    x, y = np.float32(3.25), np.float32(7.75)
    exp = (0.25 * 0.25 * d[7, 3] * 0 + (1 - .25) * (1 - .75) * d[7, 3] + .25 * (1 - .75) * d[7, 4] +
           (1 - .25) * .75 * d[8, 3] + .25 * .75 * d[8, 4])
    got = oracle.bilinear(d, np.array([[x]]), np.array([[y]]))[0, 0]
    assert abs(got - exp) <= 1e-4


def test_transform_indexes_identity_and_translation():
    xs, ys = oracle.transform_indexes(np.eye(3), 7, 5)
    np.testing.assert_array_equal(xs, np.tile(np.arange(7, dtype=np.float32), (5, 1)))
    np.testing.assert_array_equal(ys, np.tile(np.arange(5, dtype=np.float32)[:, None], (1, 7)))
    H = np.array([[1, 0, 2.5], [0, 1, -1.0], [0, 0, 1]], np.float32)   # 1-based in, 0-based out
    xs, ys = oracle.transform_indexes(H, 4, 3)
    np.testing.assert_array_equal(xs[0], np.arange(4, dtype=np.float32) + 2.5)
    np.testing.assert_array_equal(ys[:, 0], np.arange(3, dtype=np.float32) - 1.0)


def test_matrix_inverse_and_relative_pose():
    from psw_amd import scene
    R1 = scene.rotation(RNG.normal(size=3), 0.4)
    R2 = scene.rotation(RNG.normal(size=3), 0.3)
    t1, t2 = RNG.normal(size=3), RNG.normal(size=3)
    np.testing.assert_allclose(oracle.mat3_inv(R1) @ R1.astype(np.float32), np.eye(3), atol=1e-6)
    Rrel, trel = oracle.relative_matrices(R1, t1, R2, t2, 0)
    # X_src = R2 X_w + t2 and X_w = R1^-1 (X_ref - t1)  =>  X_src = R2 R1^-1 X_ref + (t2 - R2 R1^-1 t1)
    np.testing.assert_allclose(Rrel, R2 @ R1.T, atol=1e-6)
    np.testing.assert_allclose(trel, t2 - R2 @ R1.T @ t1, atol=1e-5)
    # alternative method: R, t camera-to-world
    Rrel, trel = oracle.relative_matrices(R1, t1, R2, t2, 1)
    np.testing.assert_allclose(Rrel, R2.T @ R1, atol=1e-6)
    np.testing.assert_allclose(trel, R2.T @ (t1 - t2), atol=1e-5)


def test_homography_maps_plane_points():
    """A point on the plane z = d seen at pixel p in the reference projects to
    H p in the source (1-based pixels, as PlaneSweepThread)."""
    from psw_amd import scene
    K = scene.one_based(scene.intrinsics(64, 48, 60.0))
    R = scene.rotation([0.2, 1.0, 0.1], 0.03)
    C = np.array([0.04, -0.01, 0.02])
    t = -R @ C
    d = 2.5
    H = oracle.homography_ref(K, np.eye(3), np.zeros(3), R, t, d).astype(np.float64)
    Hd = oracle.homography_f64(K, np.eye(3), np.zeros(3), R, t, d)
    for (u, v) in [(1.0, 1.0), (30.0, 20.0), (64.0, 48.0)]:
        X = d * np.linalg.solve(K.astype(np.float64), [u, v, 1.0])
        x = K.astype(np.float64) @ (R.astype(np.float64) @ X + t.astype(np.float64))
        q = H @ [u, v, 1.0]
        assert np.allclose(q[:2] / q[2], x[:2] / x[2], atol=2e-3)
        q = Hd @ [u, v, 1.0]
        assert np.allclose(q[:2] / q[2], x[:2] / x[2], atol=1e-6)
    assert H[2, 2] == 1.0


def test_reference_plane_loop():
    for zn, zf, n in [(0.1, 1.0, 200), (0.5, 5.0, 64), (0.1, 5.0, 250), (0.3, 8.0, 256), (1.0, 2.0, 2)]:
        p = oracle.reference_planes(zn, zf, n)
        assert len(p) in (n, n - 1)
        assert p[0] == np.float32(zn) and p[-1] <= np.float32(zf) and np.all(np.diff(p) > 0)
        # float accumulation: d_k = fl(d_{k-1} + fl(dstep))
        dstep = (np.float32(zf) - np.float32(zn)) / np.float32(n - 1)
        acc = np.float32(zn)
        for k in range(len(p)):
            assert p[k] == acc
            acc = np.float32(acc + dstep)


# ---- sweep known answers ------------------------------------------------------------------

def _textured(h, w, seed=0):
    from psw_amd import scene
    yy, xx = np.mgrid[0:h, 0:w]
    pts = np.stack([xx.ravel() * 0.013, yy.ravel() * 0.017, np.full(xx.size, 0.5)], 1)
    return (scene.texture(pts, seed) * 255).reshape(h, w).astype(np.float32)


def test_sweep_identity_pose_picks_first_plane():
    """Src = Ref and Rsrc = Rref: every plane gives the same NCC, so the strict
    '>' keeps the first plane (kernels.cu:77) and the depth is planes[0]."""
    from psw_amd import scene
    img = _textured(40, 56)
    K = scene.one_based(scene.intrinsics(56, 40, 50.0))
    planes = oracle.reference_planes(0.5, 4.0, 10)
    I, z = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    d, bn, bd = oracle.sweep_ref(img, img[None], K, I, z, I[None], z[None], planes, 5, 1e-4, 0.5, 4.0,
                                 per_source=True)
    # the warp samples the last column/row as 0 (kernels.cu:21), so only the
    # interior sees the unchanged image
    assert np.all(bn[0, :-3, :-3] > 0.99)
    sel = bn[0] > 0.5
    np.testing.assert_array_equal(d[sel], np.full(sel.sum(), planes[0], np.float32))
    np.testing.assert_array_equal(d[~sel], np.full((~sel).sum(), 4.0, np.float32))
    np.testing.assert_array_equal(bd[0][bn[0] > 0], np.full((bn[0] > 0).sum(), planes[0], np.float32))
    arg, _, _ = oracle.sweep_agg(img, img[None], K, I, z, I[None], z[None], planes, 5, 1e-4, coord_offset=1)
    assert np.all(arg == 0)


def _plane_views(d_true, width=96, height=72, baseline=(0.08, 0.02, 0.0)):
    """Reference and one source image of a textured fronto-parallel plane z = d_true."""
    from psw_amd import scene
    K = scene.intrinsics(width, height, 80.0).astype(np.float64)
    Ki = np.linalg.inv(K)
    u, v = np.meshgrid(np.arange(width, dtype=np.float64), np.arange(height, dtype=np.float64))
    rays = np.stack([u.ravel(), v.ravel(), np.ones(u.size)], 1) @ Ki.T

    def render(C):
        X = np.asarray(C)[None] + d_true * rays           # rays have z = 1, camera looks along +z
        return (255 * scene.texture(np.stack([X[:, 0] * 3, X[:, 1] * 3, np.full(len(X), 0.3)], 1), 7)).reshape(
            height, width).astype(np.float32)

    C = np.array(baseline)
    ref, src = render([0.0, 0.0, 0.0]), render(C)
    R = np.eye(3, dtype=np.float32)
    return ref, src[None], K.astype(np.float32), R, np.zeros(3, np.float32), R[None], (-C).astype(np.float32)[None]


@pytest.mark.parametrize("mode", ["aggregate", "reference"])
def test_sweep_fronto_parallel_plane(mode):
    from psw_amd import scene
    d_true = 2.0
    ref, src, K, Rr, tr, Rs, ts = _plane_views(d_true)
    # disparities 6.4, 4.6, 3.2, 2.3, 1.6 px: neighbours at least 0.7 px apart
    planes = np.array([1.0, 1.4, 2.0, 2.8, 4.0], np.float32)
    inner = (slice(8, -8), slice(12, -12))
    if mode == "aggregate":
        arg, _, _ = oracle.sweep_agg(ref, src, K, Rr, tr, Rs, ts, planes, 7, 1e-4, coord_offset=0)
        assert np.mean(arg[inner] == 2) >= 0.98
    else:
        d = oracle.sweep_ref(ref, src, scene.one_based(K), Rr, tr, Rs, ts, planes, 7, 1e-4, 0.5, 4.0)
        assert np.mean(d[inner] == np.float32(d_true)) >= 0.98


def test_ncc_invariant_to_affine_intensity():
    v = _plane_views(2.4, 80, 60)
    ref, src, K, Rr, tr, Rs, ts = v
    planes = np.linspace(1.5, 3.5, 12).astype(np.float32)
    a1, c1, _ = oracle.sweep_agg(ref, src, K, Rr, tr, Rs, ts, planes, 5, 1e-4)
    a2, c2, _ = oracle.sweep_agg(ref, 0.5 * src + 20.0, K, Rr, tr, Rs, ts, planes, 5, 1e-4)
    # away from the border, where the warp samples no out-of-image zeros
    inner = (slice(6, -6), slice(12, -12))
    assert np.mean(a1[inner] == a2[inner]) >= 0.999
    assert np.max(np.abs(c1[inner] - c2[inner])) <= 1e-3


def test_sweep_fp32_fp64_agree(views_small):
    from psw_amd import stages
    v = views_small
    planes = stages.inverse_depth_planes(0.5, 6.0, 24)
    a32, _, _ = oracle.sweep_agg(v["ref"], v["src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, 7,
                                 1e-4)
    a64, _, _ = oracle.sweep_agg(v["ref"], v["src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, 7,
                                 1e-4, prec="f64")
    assert np.mean(a32 == a64) >= 0.999


# ---- TGV2 -----------------------------------------------------------------------------------

def test_tensor_is_identity_for_beta_zero():
    img = RNG.uniform(0, 1, size=(20, 30)).astype(np.float32)
    T11, T12, T21, T22 = oracle.tensor(img, 0.0, 1.0)
    assert np.all(T12 == 0) and np.all(T21 == 0)
    assert np.max(np.abs(T11 - 1)) <= 2.5e-7 and np.max(np.abs(T22 - 1)) <= 2.5e-7
    T11, T12, T21, T22 = oracle.tensor(img, 2.0, 1.0)          # anisotropic: T n = k n, T n_perp = n_perp
    gx = np.diff(img, axis=1, append=img[:, -1:])
    gy = np.diff(img, axis=0, append=img[-1:, :])
    g = np.hypot(gx, gy)
    m = g > 1e-3
    nx, ny = gx[m] / g[m], gy[m] / g[m]
    k = np.exp(-2.0 * g[m])
    np.testing.assert_allclose(T11[m] * nx + T12[m] * ny, k * nx, atol=1e-5)
    np.testing.assert_allclose(T11[m] * -ny + T12[m] * nx, -ny, atol=1e-5)


def test_tgv_fixed_point_constant_depth():
    """Identical images and poses: It = Iu = 0 (to rounding), so a constant
    u stays constant and p, q stay 0."""
    from psw_amd import scene
    img = _textured(36, 48)
    K = scene.one_based(scene.intrinsics(48, 36, 45.0))
    I, z = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    u0 = np.full((36, 48), 2.0, np.float32)
    u, st = oracle.tgv(img, img[None], K, I, z, I[None], z[None], 60, 2, lambda_=1.0, u_init=u0, state=True)
    assert np.max(np.abs(u - 2.0)) <= 1e-5
    assert np.max(np.abs(st["px"])) <= 1e-5 and np.max(np.abs(st["qz"])) <= 1e-5


def test_tgv_projections_and_energy_descent(views_small):
    from psw_amd import scene
    v = views_small
    K1 = scene.one_based(v["K"])
    kw = dict(lambda_=1.0, alpha0=2.0, alpha1=1.5)
    u, st = oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], 150, 1, state=True, **kw)
    assert np.max(np.hypot(st["px"], st["py"])) <= 1 + 1e-6
    assert np.max(np.sqrt(st["qx"] ** 2 + st["qy"] ** 2 + 2 * st["qz"] ** 2)) <= 1 + 1e-6
    assert np.max(np.abs(st["r"])) <= 1.0
    It, Iu, T4 = oracle.tgv_linearise(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"],
                                      np.ones_like(v["ref"]))
    e0 = oracle.tgv_energy(np.ones_like(u), np.zeros_like(u), np.zeros_like(u), np.ones_like(u), T4, It, Iu, 1.0, 2.0,
                           1.5)
    e1 = oracle.tgv_energy(u, st["u1x"], st["u1y"], np.ones_like(u), T4, It, Iu, 1.0, 2.0, 1.5)
    assert e1 < e0
    # data term at u = u0 is sum |It|: a closed form of the energy
    e_data = oracle.tgv_energy(np.ones_like(u), np.zeros_like(u), np.zeros_like(u), np.ones_like(u), T4, It, Iu, 1.0,
                               0.0, 0.0)
    assert abs(e_data - np.abs(It.astype(np.float64)).sum()) <= 1e-9 * e_data


def test_tgv_fp32_close_to_fp64(views_small):
    """One linearisation: the float algorithm stays within 1e-4 of double.
    Re-linearising (warps > 1) can move a sample across the image border,
    where the bilinear rule switches it to 0 (kernels.cu:21): a handful of
    pixels then differ more, so that case is bounded by a percentile."""
    from psw_amd import scene
    v = views_small
    K1 = scene.one_based(v["K"])
    args = (v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    a = oracle.tgv(*args, 200, 1, lambda_=1.0)
    b = oracle.tgv(*args, 200, 1, lambda_=1.0, prec="f64")
    assert np.max(np.abs(a - b) / np.abs(b)) <= 1e-4
    a = oracle.tgv(*args, 100, 2, lambda_=1.0)
    b = oracle.tgv(*args, 100, 2, lambda_=1.0, prec="f64")
    e = np.abs(a - b) / np.abs(b)
    assert np.percentile(e, 99.9) <= 1e-4 and e.max() <= 1e-3


# ---- fusion ---------------------------------------------------------------------------------

def _prox_closed_form(u, W, tau, lam):
    centres = oracle.bin_centers(8)
    tl = np.float32(tau) * np.float32(lam)
    vals = [np.float32(u)] + [np.float32(np.float32(u) + np.float32(tl * np.float32(w))) for w in W]
    allv = np.sort(np.concatenate([centres, np.array(vals, np.float32)]))
    return max(np.float32(-1.0), allv[8])


def test_bin_centres():
    np.testing.assert_allclose(oracle.bin_centers(8), [-1, -1, -0.6, -0.2, 0.2, 0.6, 1, 1], atol=1e-7)


def test_prox_hist_closed_form_matches_insertion():
    """sortedHist::insert drops values below the running minimum (-1), so
    proxHist = max(-1, median(centres U {u, p_1..p_8}))."""
    for _ in range(20000):
        h = RNG.integers(0, 40, size=8) * (RNG.uniform(size=8) < 0.6)
        W = [int(-h[:i].sum() + h[i:].sum()) for i in range(1, 9)]
        u = np.float32(RNG.uniform(-3, 3) if RNG.uniform() < 0.9 else RNG.choice([-1.0, 1.0, 0.2, -0.6]))
        tau, lam = np.float32(RNG.uniform(0.01, 0.3)), np.float32(RNG.uniform(0.1, 2.0))
        assert oracle.prox_hist_insertion(u, W, tau, lam) == _prox_closed_form(u, W, tau, lam)


def test_prox_hist_empty_histogram():
    for u in [-5.0, -1.0, -0.3, 0.0, 0.7, 1.0, 4.0]:
        assert oracle.prox_hist_insertion(u, [0] * 8, 0.1, 0.5) == np.float32(max(u, -1.0))


def test_histogram_hand_computed_bins():
    """Voxels on the optical axis of a camera at the origin looking along +z at
    a constant depthmap D = 2: sd = D - z, bins by fusion.h:883-907."""
    K = np.array([[50, 0, 20], [0, 50, 15], [0, 0, 1]], np.float32)
    depth = np.full((31, 41), 2.0, np.float32)
    vol = np.array([-0.5, -0.5, 0.0, 0.5, 0.5, 4.0], np.float32)      # 1 x 1 x 8 voxels, centres z = 0.25 .. 3.75
    hist = np.zeros((8, 1, 1, 8), np.uint16)
    oracle.fusion_hist(hist, vol, depth, K, np.eye(3), np.zeros(3), 0.3)
    bins = [int(np.argmax(hist[k, 0, 0])) for k in range(8)]
    assert bins == [7, 7, 7, 6, 1, 0, 0, 0]
    assert hist.sum() == 8


def _hist_float64(vol, depth, K, R, T, thr, shape):
    gd, gh, gw = shape
    a, b = vol[:3].astype(np.float64), vol[3:].astype(np.float64)
    zz, yy, xx = np.meshgrid(np.arange(gd), np.arange(gh), np.arange(gw), indexing="ij")
    f = np.stack([(xx + .5) / gw, (yy + .5) / gh, (zz + .5) / gd], -1)
    c = a + (b - a) * f
    c = c @ R.astype(np.float64).T + T.astype(np.float64)
    c = c @ K.astype(np.float64).T
    px, py, z = c[..., 0] / c[..., 2], c[..., 1] / c[..., 2], c[..., 2]
    H, W = depth.shape
    ok = (px >= 0) & (px <= W - 1) & (py >= 0) & (py <= H - 1)
    x0 = np.clip(np.floor(px), 0, W - 1).astype(int)
    y0 = np.clip(np.floor(py), 0, H - 1).astype(int)
    x1, y1 = np.minimum(x0 + 1, W - 1), np.minimum(y0 + 1, H - 1)
    fx, fy = px - np.floor(px), py - np.floor(py)
    D = depth.astype(np.float64)
    d = (D[y0, x0] * (1 - fy) + D[y1, x0] * fy) * (1 - fx) + (D[y0, x1] * (1 - fy) + D[y1, x1] * fy) * fx
    sd = d - z
    bin_ = np.where(sd >= thr, 7, np.where(sd <= -thr, 0, np.minimum(np.round((sd + thr) / (2 * thr) * 5 + 1), 6)))
    return ok, bin_.astype(int)


def test_histogram_matches_independent_float64():
    from psw_amd import scene
    o = scene.orbit_depthmaps(seed=9, n=2, width=64, height=48)
    vol = np.array([-3.0, -1.6, -4.0, 3.0, 1.6, 6.0], np.float32)
    shape = (20, 12, 16)
    for k in range(2):
        hist = np.zeros(shape + (8,), np.uint16)
        oracle.fusion_hist(hist, vol, o["depths"][k], o["K"], o["R"][k], o["t"][k], 0.3, prec="f64")
        ok, bins = _hist_float64(vol, o["depths"][k], o["K"], o["R"][k], o["t"][k], 0.3, shape)
        got_ok = hist.sum(-1) == 1
        assert np.mean(got_ok == ok) >= 0.999
        both = got_ok & ok
        assert np.mean(np.argmax(hist, -1)[both] == bins[both]) >= 0.999
        h32 = np.zeros(shape + (8,), np.uint16)
        oracle.fusion_hist(h32, vol, o["depths"][k], o["K"], o["R"][k], o["t"][k], 0.3)
        assert np.mean(np.all(h32 == hist, -1)) >= 0.999


def test_fusion_pd_invariants():
    from psw_amd import scene
    o = scene.orbit_depthmaps(seed=3, n=3, width=48, height=36)
    vol = np.array([-3.0, -1.6, -4.0, 3.0, 1.6, 6.0], np.float32)
    shape = (14, 10, 12)
    hist = np.zeros(shape + (8,), np.uint16)
    for k in range(3):
        oracle.fusion_hist(hist, vol, o["depths"][k], o["K"], o["R"][k], o["t"][k], 0.3)
    u, v, p = np.zeros(shape, np.float32), np.zeros(shape, np.float32), np.zeros((3,) + shape, np.float32)
    for _ in range(40):
        oracle.fusion_u(u, v, p, hist, 0.1, 0.5)
        oracle.fusion_p(p, v, 1.0)
    assert np.max(np.sqrt((p.astype(np.float64) ** 2).sum(0))) <= 1 + 1e-6
    assert u.min() >= -1.0
    # voxels never voted on keep u = 0 only while p around them is 0; empty grid stays 0
    ue, ve, pe = np.zeros(shape, np.float32), np.zeros(shape, np.float32), np.zeros((3,) + shape, np.float32)
    oracle.fusion_u(ue, ve, pe, np.zeros_like(hist), 0.1, 0.5)
    assert np.all(ue == 0)


def test_fusion_energy_closed_forms():
    """TV-L1 histogram energy: a unit spike in the interior has gradient
    norm sqrt(3) at itself and 1 at each of its three backward neighbours;
    a lone vote in bin k at u = 0 costs lambda |c_k|; constant u = c_k with
    all votes in bin k costs nothing."""
    shape = (4, 5, 6)
    u = np.zeros(shape)
    h = np.zeros(shape + (8,), np.uint16)
    assert oracle.fusion_energy(u, h, 0.5) == 0.0
    u[2, 2, 3] = 1.0
    np.testing.assert_allclose(oracle.fusion_energy(u, h, 0.5), 3.0 + np.sqrt(3.0), rtol=1e-15)
    # binParams (src/inc/fusion.h:1041-1054) in float: c_i = 2 (i - 1) / (8 - 3) - 1, ends at -1 and 1
    f32 = np.float32
    centres = np.array([-1.0] + [f32(2) * f32(i - 1) / f32(5) - f32(1) for i in range(1, 7)] + [1.0], np.float64)
    for k in range(8):
        h = np.zeros(shape + (8,), np.uint16)
        h[1, 3, 4, k] = 3
        np.testing.assert_allclose(oracle.fusion_energy(np.zeros(shape), h, 0.7), 0.7 * 3 * abs(centres[k]),
                                   rtol=1e-15)
        h[..., k] = 2
        assert oracle.fusion_energy(np.full(shape, centres[k]), h, 0.7) == 0.0


# ---- TVL1 -------------------------------------------------------------------------------------

def test_tvl1_constant_depth_is_fixed():
    ref = _textured(30, 40)
    d = np.full((30, 40), 1.7, np.float32)
    out = oracle.tvl1(d, ref, 0.5, 5.0, niters=40)
    assert np.max(np.abs(out - 1.7)) <= 1e-5
    out64 = oracle.tvl1(d, ref, 0.5, 5.0, niters=40, prec="f64")
    assert np.max(np.abs(out64 - np.float64(np.float32(1.7)))) <= 1e-12


def test_tvl1_reduces_total_variation():
    ref = _textured(30, 40)
    base = np.linspace(1.0, 3.0, 40, dtype=np.float32)[None].repeat(30, 0)
    noisy = (base + RNG.normal(0, 0.05, size=base.shape)).astype(np.float32)
    out = oracle.tvl1(noisy, ref, 0.5, 5.0, niters=100)
    tv = lambda a: np.abs(np.diff(a, axis=0)).sum() + np.abs(np.diff(a, axis=1)).sum()  # noqa: E731
    assert tv(out) < tv(noisy)
