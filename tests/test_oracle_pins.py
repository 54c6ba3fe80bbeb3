"""Pins of the oracle's TGV2 iteration, linearisation, energy and the
aggregated sweep's reduction against what the mathematics fixes (no GPU).

The reference has no tests, so every expectation here is computed by hand
from the reference's formulas (src/TGV2_kernels.cu:54-105, :238-255,
:300-365) on states chosen so the answer is a closed form:
  - an affine depth with u1 = grad u is a fixed point away from the borders,
    and the first iteration's border response is the clamped-difference
    stencil;
  - one iteration from a single nonzero dual (px, py, qx, qy, qz) moves u and
    u1 by the backward-difference stencil of the U step, with each tensor
    entry at its own place (T12 != T21 here, so a transposed tensor fails);
  - one iteration from a single nonzero primal (u, u1x, u1y) sets p and q by
    the forward-difference stencil of the P and Q steps (the /2 of the
    symmetrised gradient, and which neighbour q_z and q_y take, are pinned);
  - the data dual uses u, not ubar, with the sign of It + (u - u0) Iu;
  - the energy's regulariser has closed forms on affine and step fields;
  - Iu equals the derivative in depth of the warped intensity, and It the
    warped-minus-reference intensity, for a lateral translation of a linear
    image, where both are closed forms;
  - the sweep's (argmin, min, second) reduction on hand-built cost
    sequences, and against numpy's reduction of the oracle's cost volume.
Both precisions of the oracle are checked where the state is exactly
representable.
"""
import numpy as np
import pytest

import oracle

PRM = dict(lambda_=1.0, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0)
RNG = np.random.default_rng(20240607)


def _tensor(H, W, rng=RNG):
    """A per-pixel tensor with independent entries (T12 != T21)."""
    return np.stack([rng.uniform(0.5, 1.5, (H, W)), rng.uniform(-0.3, 0.3, (H, W)),
                     rng.uniform(-0.3, 0.3, (H, W)), rng.uniform(0.5, 1.5, (H, W))])


def _zeros(H, W, N=1):
    z = np.zeros((H, W))
    return {"u": z.copy(), "u1x": z.copy(), "u1y": z.copy(), "px": z.copy(), "py": z.copy(), "qx": z.copy(),
            "qy": z.copy(), "qz": z.copy(), "r": np.zeros((N, H, W))}


def _run(state, T4, u0, It, Iu, k, prec, **kw):
    p = dict(PRM)
    p.update(kw)
    return oracle.tgv_iterate(state, T4, u0, It, Iu, k, prec=prec, **p)


# ---- the second-order branch ------------------------------------------------------------

@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_affine_depth_with_matching_u1_is_a_fixed_point(prec):
    """u = a + b x + c y, u1 = (b, c), It = Iu = 0, zero duals: every
    gradient the P and Q steps see is exactly 0 away from the last row and
    column, so nothing moves there; the clamped differences at the border
    make p(W-1) = proj(sigma alpha1 T (-b, 0)), which feeds back into u and
    u1 one stencil step per iteration."""
    H, W, k = 40, 56, 6
    a, b, c = 1.5, 0.125, -0.0625              # dyadic: u and its differences are exact in float
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    u = a + b * x + c * y
    T4 = _tensor(H, W)
    if prec == "f32":
        T4 = T4.astype(np.float32).astype(np.float64)
    st = _zeros(H, W)
    st.update(u=u, u1x=np.full((H, W), b), u1y=np.full((H, W), c))
    z = np.zeros((1, H, W))
    out = _run(st, T4, u, z, z, k, prec)
    inner = (slice(0, H - 3 * k), slice(0, W - 3 * k))
    for name, exp in (("u", u), ("u1x", st["u1x"]), ("u1y", st["u1y"])):
        np.testing.assert_array_equal(out[name][inner], exp[inner])
    for name in ("px", "py", "qx", "qy", "qz"):
        assert np.all(out[name][inner] == 0), name
    # the first iteration's response at the right border column, rows 0..H-2:
    # p = sigma alpha1 (T11 (0 - b), T21 (0 - b)) there and 0 on its left; the
    # U step's backward differences (left neighbour 0, upper neighbour the
    # row above, clamped at row 0) then give
    # u += tau alpha1 (T11 px + T12 py + T21 (px - px_up) + T22 (py - py_up)),
    # u1x += tau alpha1 (T11 px + T12 py), u1y += tau alpha1 (T21 px + T22 py)
    one = _run(st, T4, u, z, z, 1, prec)
    s, t = PRM["sigma"] * PRM["alpha1"], PRM["tau"] * PRM["alpha1"]
    rows = slice(0, H - 1)
    T11, T12, T21, T22 = (T4[i][rows, W - 1] for i in range(4))
    px, py = s * T11 * (-b), s * T21 * (-b)
    px_up, py_up = np.concatenate([px[:1], px[:-1]]), np.concatenate([py[:1], py[:-1]])
    assert np.max(np.hypot(px, py)) < 1
    tol = dict(rtol=1e-6 if prec == "f32" else 1e-14, atol=0)
    np.testing.assert_allclose(one["px"][rows, W - 1], px, **tol)
    np.testing.assert_allclose(one["py"][rows, W - 1], py, **tol)
    np.testing.assert_allclose(one["u"][rows, W - 1] - u[rows, W - 1],
                               t * (T11 * px + T12 * py + T21 * (px - px_up) + T22 * (py - py_up)),
                               rtol=tol["rtol"], atol=2e-6 if prec == "f32" else 1e-15)   # |u| ~ 8: float ulp
    tol["atol"] = 1e-8 if prec == "f32" else 1e-17      # u1 ~ 0.1: float ulp 7e-9
    np.testing.assert_allclose(one["u1x"][rows, W - 1] - b, t * (T11 * px + T12 * py), **tol)
    np.testing.assert_allclose(one["u1y"][rows, W - 1] - c, t * (T21 * px + T22 * py), **tol)
    assert np.all(one["px"][:, :W - 1][rows] == 0)


def _single(H, W, name, val, y0, x0):
    st = _zeros(H, W)
    st[name][y0, x0] = val
    return st


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dual", ["px", "py", "qx", "qy", "qz"])
def test_one_step_from_a_single_dual(prec, dual):
    """u = u1 = 0 and It = Iu = 0: the P and Q steps see zero gradients, so
    the duals keep their values, and the U step moves u and u1 by the
    reference's backward-difference stencil of that one dual value."""
    H, W, y0, x0, a = 12, 14, 5, 6, 0.375
    T4 = _tensor(H, W)
    if prec == "f32":
        T4 = T4.astype(np.float32).astype(np.float64)
    z = np.zeros((1, H, W))
    out = _run(_single(H, W, dual, a, y0, x0), T4, np.zeros((H, W)), z, z, 1, prec)
    ta1, ta0 = PRM["tau"] * PRM["alpha1"], PRM["tau"] * PRM["alpha0"]
    T11, T12, T21, T22 = T4
    du, du1x, du1y = np.zeros((H, W)), np.zeros((H, W)), np.zeros((H, W))
    if dual == "px":           # u += tau alpha1 (T11 (px - px[x-1]) + T21 (px - px[y-1])); u1 += tau alpha1 T(:,0) px
        du[y0, x0] = ta1 * (T11[y0, x0] + T21[y0, x0]) * a
        du[y0, x0 + 1] = -ta1 * T11[y0, x0 + 1] * a
        du[y0 + 1, x0] = -ta1 * T21[y0 + 1, x0] * a
        du1x[y0, x0] = ta1 * T11[y0, x0] * a
        du1y[y0, x0] = ta1 * T21[y0, x0] * a
    elif dual == "py":
        du[y0, x0] = ta1 * (T12[y0, x0] + T22[y0, x0]) * a
        du[y0, x0 + 1] = -ta1 * T12[y0, x0 + 1] * a
        du[y0 + 1, x0] = -ta1 * T22[y0 + 1, x0] * a
        du1x[y0, x0] = ta1 * T12[y0, x0] * a
        du1y[y0, x0] = ta1 * T22[y0, x0] * a
    elif dual == "qx":         # u1x += tau alpha0 (qx - qx[x-1])
        du1x[y0, x0], du1x[y0, x0 + 1] = ta0 * a, -ta0 * a
    elif dual == "qy":         # u1y += tau alpha0 (qy - qy[y-1])
        du1y[y0, x0], du1y[y0 + 1, x0] = ta0 * a, -ta0 * a
    else:                      # u1x += tau alpha0 (qz - qz[y-1]); u1y += tau alpha0 (qz - qz[x-1])
        du1x[y0, x0], du1x[y0 + 1, x0] = ta0 * a, -ta0 * a
        du1y[y0, x0], du1y[y0, x0 + 1] = ta0 * a, -ta0 * a
    tol = dict(rtol=1e-6 if prec == "f32" else 1e-14, atol=1e-9 if prec == "f32" else 1e-17)
    np.testing.assert_allclose(out["u"], du, **tol)
    np.testing.assert_allclose(out["u1x"], du1x, **tol)
    np.testing.assert_allclose(out["u1y"], du1y, **tol)
    # ubar = 2 u_new - u_old with u_old = 0
    np.testing.assert_allclose(out["ubar"], 2 * du, **tol)
    # the duals are inside the unit ball and see zero gradients: unchanged
    assert out[dual][y0, x0] == a
    for other in ("px", "py", "qx", "qy", "qz"):
        if other != dual:
            assert np.all(out[other] == 0), other


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("primal", ["u", "u1x", "u1y"])
def test_one_step_from_a_single_primal(prec, primal):
    """A single nonzero primal value e (in u = ubar, or u1 = u1bar): the P
    step's forward differences give p = sigma alpha1 T (grad ubar - u1bar)
    and the Q step's q = sigma alpha0 (d_x u1x, d_y u1y, (d_y u1x + d_x u1y)/2)
    at the pixel and its left / upper neighbours."""
    H, W, y0, x0, e = 12, 14, 5, 6, 0.25
    T4 = _tensor(H, W)
    if prec == "f32":
        T4 = T4.astype(np.float32).astype(np.float64)
    z = np.zeros((1, H, W))
    out = _run(_single(H, W, primal, e, y0, x0), T4, np.zeros((H, W)), z, z, 1, prec)
    sa1, sa0 = PRM["sigma"] * PRM["alpha1"], PRM["sigma"] * PRM["alpha0"]
    T11, T12, T21, T22 = T4
    px, py = np.zeros((H, W)), np.zeros((H, W))
    qx, qy, qz = np.zeros((H, W)), np.zeros((H, W)), np.zeros((H, W))

    def p_at(yy, xx, gx, gy):
        px[yy, xx] = sa1 * (T11[yy, xx] * gx + T12[yy, xx] * gy)
        py[yy, xx] = sa1 * (T21[yy, xx] * gx + T22[yy, xx] * gy)

    if primal == "u":          # grad u at the pixel (-e, -e), at x-1 (e, 0), at y-1 (0, e)
        p_at(y0, x0, -e, -e)
        p_at(y0, x0 - 1, e, 0.0)
        p_at(y0 - 1, x0, 0.0, e)
    elif primal == "u1x":      # grad u - u1 = (-e, 0) at the pixel
        p_at(y0, x0, -e, 0.0)
        qx[y0, x0], qx[y0, x0 - 1] = -sa0 * e, sa0 * e
        qz[y0, x0], qz[y0 - 1, x0] = -sa0 * e / 2, sa0 * e / 2
    else:
        p_at(y0, x0, 0.0, -e)
        qy[y0, x0], qy[y0 - 1, x0] = -sa0 * e, sa0 * e
        qz[y0, x0], qz[y0, x0 - 1] = -sa0 * e / 2, sa0 * e / 2
    assert np.max(np.hypot(px, py)) < 1 and np.max(np.sqrt(qx ** 2 + qy ** 2 + 2 * qz ** 2)) < 1
    tol = dict(rtol=1e-6 if prec == "f32" else 1e-14, atol=1e-9 if prec == "f32" else 1e-17)
    for name, exp in (("px", px), ("py", py), ("qx", qx), ("qy", qy), ("qz", qz), ("qw", qz)):
        np.testing.assert_allclose(out[name], exp, err_msg=name, **tol)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_data_dual_uses_u_and_the_residual_sign(prec):
    """r = proj(r + sigma lambda (It + (u - u0) Iu)) with u, not ubar
    (src/planesweep.cpp:501 passes u), then u -= tau lambda sum_j r_j Iu_j.
    u is constant and ubar = u + 1, which leaves every gradient (and so p, q)
    at 0 but would change r if ubar were used there.  It and Iu vary per
    pixel, so some residuals saturate and some do not."""
    H, W, N = 6, 7, 3
    u0 = np.full((H, W), 2.0)
    u = np.full((H, W), 2.25)
    rng = np.random.default_rng(5)
    It = rng.uniform(-1.5, 1.5, (N, H, W))
    Iu = rng.uniform(-3.0, 3.0, (N, H, W))
    st = _zeros(H, W, N)
    st.update(u=u, ubar=u + 1.0)
    T4 = _tensor(H, W)
    out = _run(st, T4, u0, It, Iu, 1, prec)
    sl = PRM["sigma"] * PRM["lambda_"]
    rho = It + (u - u0)[None] * Iu
    r = np.clip(sl * rho, -1, 1)
    assert 0 < np.mean(np.abs(r) == 1) < 1
    tol = dict(rtol=1e-6 if prec == "f32" else 1e-14, atol=1e-12)
    np.testing.assert_allclose(out["r"], r, **tol)
    np.testing.assert_allclose(out["u"], u - PRM["tau"] * PRM["lambda_"] * np.sum(r * Iu, axis=0), **tol)
    assert np.all(out["px"] == 0) and np.all(out["qz"] == 0)


# ---- the energy's regulariser -------------------------------------------------------------

def _energy(u, u1x, u1y, T4, lam=0.0, a0=2.0, a1=1.5, N=1):
    H, W = u.shape
    z = np.zeros((N, H, W))
    return oracle.tgv_energy(np.asarray(u, np.float64), np.asarray(u1x, np.float64), np.asarray(u1y, np.float64),
                             np.zeros((H, W)), np.asarray(T4, np.float64), z, z, lam, a0, a1)


def test_energy_affine_depth_is_zero_inside():
    """u affine with u1 = grad u: only the clamped last row and column
    contribute, |T (-b, 0)| along the last column, |T (0, -c)| along the last
    row and |T (-b, -c)| at the corner; the second-order term is 0."""
    H, W, b, c = 17, 23, 0.3, -0.7
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    t11, t12, t21, t22 = 1.2, 0.4, -0.1, 0.9
    T4 = np.stack([np.full((H, W), t)for t in (t11, t12, t21, t22)])
    e = _energy(1.0 + b * x + c * y, np.full((H, W), b), np.full((H, W), c), T4)
    col = (H - 1) * abs(b) * np.hypot(t11, t21)
    row = (W - 1) * abs(c) * np.hypot(t12, t22)
    corner = np.hypot(t11 * -b + t12 * -c, t21 * -b + t22 * -c)
    assert abs(e - 1.5 * (col + row + corner)) <= 1e-12 * e


def test_energy_step_edges():
    """Closed forms with T = I: a unit step in u (first order), and steps in
    u1 along x and y (second order: |e(u1)|_F counts the off-diagonal twice,
    so a step s in u1x across y gives |s| / sqrt(2) per pixel of the edge)."""
    H, W, x0, y0, s = 9, 11, 4, 6, 0.5
    I4 = np.stack([np.ones((H, W)), np.zeros((H, W)), np.zeros((H, W)), np.ones((H, W))])
    z = np.zeros((H, W))
    u = np.where(np.arange(W)[None] >= x0, 1.0, 0.0) * np.ones((H, 1))
    assert abs(_energy(u, z, z, I4) - 1.5 * H) <= 1e-12
    # u1x steps across x: e_xx = s on column x0-1; T (grad u - u1) = (-u1x, 0) where u1x = s
    u1x = np.where(np.arange(W)[None] >= x0, s, 0.0) * np.ones((H, 1))
    exp = 2.0 * H * s + 1.5 * H * (W - x0) * s
    assert abs(_energy(z, u1x, z, I4) - exp) <= 1e-12 * exp
    # u1x steps across y: e_xy = s / 2 on row y0-1, counted twice
    u1x = np.where(np.arange(H)[:, None] >= y0, s, 0.0) * np.ones((1, W))
    exp = 2.0 * W * s / np.sqrt(2) + 1.5 * (H - y0) * W * s
    assert abs(_energy(z, u1x, z, I4) - exp) <= 1e-12 * exp
    # u1y steps across x: e_xy = s / 2 on column x0-1; across y: e_yy = s on row y0-1
    u1y = np.where(np.arange(W)[None] >= x0, s, 0.0) * np.ones((H, 1))
    exp = 2.0 * H * s / np.sqrt(2) + 1.5 * H * (W - x0) * s
    assert abs(_energy(z, z, u1y, I4) - exp) <= 1e-12 * exp
    u1y = np.where(np.arange(H)[:, None] >= y0, s, 0.0) * np.ones((1, W))
    exp = 2.0 * W * s + 1.5 * (H - y0) * W * s
    assert abs(_energy(z, z, u1y, I4) - exp) <= 1e-12 * exp


def test_energy_data_term_single_pixels():
    """lambda sum_j |It_j + (u - u0) Iu_j| at two pixels with hand values."""
    H, W = 4, 5
    u0 = np.full((H, W), 1.0)
    u = u0.copy()
    u[1, 2], u[3, 4] = 1.5, 0.75
    It, Iu = np.zeros((2, H, W)), np.zeros((2, H, W))
    It[0, 1, 2], Iu[0, 1, 2], It[1, 1, 2], Iu[1, 1, 2] = 0.1, -0.4, -0.3, 0.2      # |0.1-0.2| + |-0.3+0.1|
    It[1, 3, 4], Iu[1, 3, 4] = 0.05, 1.0                                            # |0.05 - 0.25|
    I4 = np.stack([np.ones((H, W)), np.zeros((H, W)), np.zeros((H, W)), np.ones((H, W))])
    e = oracle.tgv_energy(u, np.zeros((H, W)), np.zeros((H, W)), u0, I4, It, Iu, 0.5, 0.0, 0.0)
    assert abs(e - 0.5 * (0.1 + 0.2 + 0.2)) <= 1e-15


# ---- the linearisation (It, Iu) -------------------------------------------------------------

@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_linearisation_of_a_lateral_translation(prec):
    """R = I, t = (tx, ty, 0), constant depth u0, linear source image
    I = A x + B y + C: the warp is a shift by f t / u, bilinear sampling is
    exact, so
        It = (A fx tx + B fy ty) / (255 u0)            (reference = source image)
        Iu = d(warped I)/du = -(A fx tx + B fy ty) / (255 u0^2),
    and Iu matches the central difference of It in u0.  fx != fy, A != B and
    tx != ty, so a sign slip or swapped axes in dfx / dfy fails."""
    H, W = 40, 60
    A, B, C = 1.25, -0.75, 100.0
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    img = (A * x + B * y + C).astype(np.float32)
    fx, fy, cx, cy = 50.0, 40.0, 30.5, 20.5                 # 1-based principal point
    K = np.array([[fx, 0, cx], [0, fy, cy], [0, 0, 1]], np.float32)
    I3, z3 = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    tx, ty, u0v = 0.05, -0.03, 2.0
    tsrc = np.array([[tx, ty, 0.0]], np.float32)
    # the reference's relative pose: R_rel = R_src R_ref^-1, t_rel = t_src - R_rel t_ref
    args = (img, img[None], K, I3, z3, I3[None], tsrc)
    u0 = np.full((H, W), u0v, np.float32)
    It, Iu, _ = oracle.tgv_linearise(*args, u0, prec=prec)
    inner = (slice(1, H - 4), slice(0, W - 4))              # warp stays inside (y shifts by -0.6 px); last row/col clamp
    k = (A * fx * float(np.float32(tx)) + B * fy * float(np.float32(ty))) / 255.0
    tol = 2e-5 if prec == "f32" else 1e-9
    np.testing.assert_allclose(It[0][inner], k / u0v, rtol=tol, atol=tol * 1e-2)
    np.testing.assert_allclose(Iu[0][inner], -k / u0v ** 2, rtol=tol, atol=tol * 1e-2)
    if prec == "f64":
        h = 1e-4
        Itp, _, _ = oracle.tgv_linearise(*args, np.full((H, W), u0v + h, np.float32), prec=prec)
        Itm, _, _ = oracle.tgv_linearise(*args, np.full((H, W), u0v - h, np.float32), prec=prec)
        hp = float(np.float32(u0v + h)) - u0v
        hm = u0v - float(np.float32(u0v - h))
        fd = (Itp[0] - Itm[0]) / (hp + hm)
        np.testing.assert_allclose(Iu[0][inner], fd[inner], rtol=1e-6)


# ---- the aggregated sweep's reduction ------------------------------------------------------

@pytest.mark.parametrize("costs,exp", [
    ([3.0, 1.0, 2.0], (1, 1.0, 2.0)),
    ([2.0, 3.0, 1.0], (2, 1.0, 2.0)),
    ([1.0, 1.0, 4.0], (0, 1.0, 1.0)),            # a tied minimum: first plane, second = the tie
    ([5.0, 4.0, 3.0, 2.0], (3, 2.0, 3.0)),
    ([-0.5, 0.25, -0.5, -0.75], (3, -0.75, -0.5)),
    ([7.0], (0, 7.0, np.inf)),                   # one plane: no second
    ([np.nan, 2.0, 1.0], (2, 1.0, 2.0)),         # a NaN cost (zero-variance planes) never wins
])
def test_best_and_second_on_hand_sequences(costs, exp):
    a, m, s = oracle.reduce_best2(costs)
    assert (a, m, s) == exp


def test_sweep_aggregate_equals_reduction_of_its_cost_volume(views_tiny):
    """The aggregated sweep's (argmin, min, second) equal numpy's reduction
    of the FP64 cost volume (first minimum; second = the second-smallest of
    the multiset), and the volume is -(sum of NCC) / N in [-1, 1]."""
    from psw_amd import stages
    v = views_tiny
    planes = stages.inverse_depth_planes(0.5, 6.0, 20)
    a, m, s, vol = oracle.sweep_agg_volume(v["ref"], v["src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"],
                                           planes, 5, 1e-4)
    np.testing.assert_array_equal(a, np.argmin(vol, axis=0))
    srt = np.sort(vol, axis=0)
    np.testing.assert_array_equal(m, srt[0])
    np.testing.assert_array_equal(s, srt[1])
    assert np.all(vol >= -1 - 1e-12) and np.all(vol <= 1 + 1e-12)
    a64, m64, s64 = oracle.sweep_agg(v["ref"], v["src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes,
                                     5, 1e-4, prec="f64")
    np.testing.assert_array_equal(a64, a)
    np.testing.assert_array_equal(s64, s)
