"""GPU parity: the HIP kernels (called through the C ABI) against the CPU oracle.

Bars (stated per test):
  - reference-order kernels: bit-identical to the float oracle;
  - argmin plane: >= 99.9 % agreement with the FP64 oracle, disagreements
    only where the FP64 costs of the two planes are within a near-tie margin;
  - histogram counts: bit-identical to the float oracle;
  - refined depth / fused u: within 1e-4 relative of the FP64 oracle
    (relative to max(|ref|, 1) for u, whose range is about [-1, 1]).
Sizes are small so that the oracle finishes in seconds.
"""
import numpy as np
import pytest

import oracle
from conftest import one_based, rel_err

pytestmark = pytest.mark.gpu

STD_T = 1e-4      # DEFAULT_STD_THRESHOLD (defines.h:29)
NCC_T = 0.5       # DEFAULT_NCC_THRESHOLD (defines.h:30)


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


# ---- stage 1: sweep ------------------------------------------------------------------

@pytest.mark.parametrize("winsize", [5, 9])
def test_sweep_reference_launchers_bit_exact(gpu, views_tiny, winsize):
    """RunAlgorithm over the per-kernel launchers == float oracle, bit for bit."""
    import refdriver
    v = views_tiny
    K1 = one_based(v["K"])
    planes = oracle.reference_planes(0.5, 6.0, 12)
    _print_case("sweep_reference_launchers", win=winsize, P=planes.size)
    got = refdriver.run_algorithm(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, winsize,
                                  STD_T, NCC_T, 6.0)
    exp = oracle.sweep_ref(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, winsize,
                           STD_T, NCC_T, 6.0)
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("winsize,nplanes", [(5, 24), (9, 16), (3, 8)])
def test_sweep_fused_reference_bit_exact(gpu, views_small, winsize, nplanes):
    from psw_amd import stages
    v = views_small
    K1 = one_based(v["K"])
    planes = stages.reference_planes(0.5, 6.0, nplanes)
    np.testing.assert_array_equal(planes, oracle.reference_planes(0.5, 6.0, nplanes))
    _print_case("sweep_fused_reference", win=winsize, P=planes.size)
    got = stages.sweep_reference(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, winsize,
                                 STD_T, NCC_T, 6.0)
    exp = oracle.sweep_ref(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, winsize,
                           STD_T, NCC_T, 6.0)
    np.testing.assert_array_equal(got, exp)


def test_sweep_homographies_fp64_match_oracle(gpu, views_small):
    """GPU FP64 homography construction == oracle double arithmetic rounded to float."""
    from psw_amd import stages
    v = views_small
    planes = stages.inverse_depth_planes(0.3, 8.0, 32)
    got = stages.homographies_device(planes, v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    for i in range(v["Rsrc"].shape[0]):
        for p in range(planes.size):
            exp = oracle.homography_f64(v["K"], v["Rref"], v["tref"], v["Rsrc"][i], v["tsrc"][i], planes[p])
            np.testing.assert_array_equal(got[i, p].reshape(3, 3), exp.astype(np.float32))


@pytest.mark.parametrize("winsize,offset,ty", [(7, 0, "16"), (9, 0, "16"), (5, 1, "16"), (9, 0, "32"), (3, 1, "32"),
                                               (15, 0, "32")])
def test_sweep_fused_aggregate_bit_exact(gpu, views_small, monkeypatch, winsize, offset, ty):
    """Reference-order aggregated sweep == FP32 oracle, bit for bit, with
    32 x 16 and 32 x 32 tiles (PSW_SWEEP_TY)."""
    from psw_amd import stages
    monkeypatch.setenv("PSW_SWEEP_TY", ty)
    v = views_small
    K = v["K"] if offset == 0 else one_based(v["K"])
    planes = stages.inverse_depth_planes(0.5, 6.0, 24)
    _print_case("sweep_fused_aggregate", win=winsize, off=offset)
    arg, cmin, csec = stages.sweep_aggregate(v["ref"], v["src"], K, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes,
                                             winsize, STD_T, coord_offset=offset, fast=False)
    ea, ec, es = oracle.sweep_agg(v["ref"], v["src"], K, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, winsize,
                                  STD_T, coord_offset=offset)
    np.testing.assert_array_equal(arg, ea)
    np.testing.assert_array_equal(cmin, ec)
    np.testing.assert_array_equal(csec, es)


@pytest.mark.parametrize("fast,ty", [(False, "16"), (True, "16"), (True, "32")])
def test_sweep_plane_groups_identical(gpu, views_small, monkeypatch, fast, ty):
    """Splitting the plane loop into groups merged afterwards changes nothing:
    1, 3 and 24 groups give identical argmin / min / second-best; with the
    reference-order arithmetic 32-row tiles give the 16-row tiles' results
    too (the fast variant's running sums restart at other rows there)."""
    from psw_amd import stages
    v = views_small
    planes = stages.inverse_depth_planes(0.5, 6.0, 24)
    out = []
    cases = [(ty, "1"), (ty, "3"), (ty, "24")] + ([("32", "1"), ("32", "5")] if not fast else [])
    for t, g in cases:
        monkeypatch.setenv("PSW_SWEEP_TY", t)
        monkeypatch.setenv("PSW_SWEEP_GROUPS", g)
        out.append(stages.sweep_aggregate(v["ref"], v["src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"],
                                          planes, 7, STD_T, coord_offset=0, fast=fast))
    for o in out[1:]:
        for a, b in zip(out[0], o):
            np.testing.assert_array_equal(a, b)


def _near_tie_check(arg, a64, c64_planes_fn, margin):
    dis = np.argwhere(arg != a64)
    for (y, x) in dis:
        c = c64_planes_fn(y, x)
        assert abs(c[arg[y, x]] - c[a64[y, x]]) <= margin, (y, x, c[arg[y, x]], c[a64[y, x]])


@pytest.mark.parametrize("fast", [False, True])
def test_sweep_aggregate_vs_fp64(gpu, views_small, fast):
    """Argmin agreement >= 99.9 % with the FP64 oracle; disagreements are near-ties."""
    from psw_amd import stages
    v = views_small
    planes = stages.inverse_depth_planes(0.5, 6.0, 32)
    _print_case("sweep_aggregate_vs_fp64", fast=fast)
    arg, cmin, _ = stages.sweep_aggregate(v["ref"], v["src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"],
                                          planes, 7, STD_T, fast=fast)
    a64, c64, _ = oracle.sweep_agg(v["ref"], v["src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, 7,
                                   STD_T, prec="f64")
    agree = float(np.mean(arg == a64))
    print(f"argmin agreement vs FP64: {agree:.6f} ({int(np.sum(arg != a64))} of {arg.size})")
    assert agree >= 0.999
    # disagreeing pixels: the GPU's minimum cost equals the FP64 minimum to a near-tie margin
    mask = arg != a64
    if mask.any():
        assert np.max(np.abs(cmin[mask] - c64[mask])) <= (1e-3 if fast else 1e-5)


# ---- stage 2: TGV2 ----------------------------------------------------------------------

TGV_KW = dict(lambda_=1.0, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, beta=0.0, gamma=1.0)


def test_tgv_reference_launchers_vs_oracle(gpu, views_tiny):
    """PlaneSweep::TGV over the per-kernel launchers vs the float oracle."""
    import refdriver
    v = views_tiny
    K1 = one_based(v["K"])
    got = refdriver.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], 20, 2, **TGV_KW)
    exp = oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], 20, 2, **TGV_KW)
    err = rel_err(got, exp)
    print(f"launchers vs f32 oracle: max rel {err.max():.3e}, identical {np.mean(got == exp):.4f}")
    assert err.max() <= 1e-6


@pytest.mark.parametrize("niters,warps", [(30, 3), (200, 1)])
def test_tgv_fused_vs_fp64_oracle(gpu, views_small, niters, warps):
    """Fused TGV2 (FP32) within 1e-4 relative of the FP64 oracle."""
    from psw_amd import stages
    v = views_small
    K1 = one_based(v["K"])
    prm = stages.tgv_params(**TGV_KW)
    _print_case("tgv_fused", niters=niters, warps=warps)
    u = stages.tgv_run(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], niters, warps, prm)
    u64 = oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], niters, warps, prec="f64",
                     **TGV_KW)
    u32 = oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], niters, warps, **TGV_KW)
    e64 = np.abs(u - u64) / np.abs(u64)
    e32 = np.abs(u - u32) / np.abs(u32)
    print(f"vs FP64 oracle: max rel {e64.max():.3e} p99.9 {np.percentile(e64, 99.9):.3e} mean {e64.mean():.3e}; "
          f"vs FP32 oracle: max rel {e32.max():.3e}")
    if warps == 1:
        assert e64.max() <= 1e-4
    else:
        # re-linearisation can move a border sample across the image edge,
        # where the reference's bilinear rule switches it to 0 (kernels.cu:21)
        assert np.percentile(e64, 99.9) <= 1e-4 and e64.max() <= 1e-3


def test_tgv_launch_shapes_identical(gpu, views_small, monkeypatch):
    """The three tile/thread shapes of the one-iteration kernel and the
    two- and three-iterations-per-launch kernels give identical iterates (each
    pixel's arithmetic does not depend on the tiling); the full state (u, u1,
    p, q, r) is compared after 37 + 5 + 1 iterations, so the multi-iteration
    paths mix 3-, 2- and 1-iteration launches."""
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    v = views_small
    K1 = one_based(v["K"])
    Hh, W = v["ref"].shape
    N = v["src"].shape[0]
    prm = stages.tgv_params(**TGV_KW)
    Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    d_ref, d_src = DeviceArray.from_host(v["ref"]), DeviceArray.from_host(v["src"])
    d_init = DeviceArray.from_host(np.clip(v["depth"], 0.5, 6.0).astype(np.float32))
    states = []
    for double, triple, var, ty2 in (("0", "0", "0", "10"), ("0", "0", "1", "10"), ("0", "0", "2", "10"),
                                     ("1", "0", "1", "10"), ("1", "0", "0", "10"), ("1", "0", "1", "4"),
                                     ("1", "0", "1", "8"), ("1", "0", "1", "12"), ("1", "1", "1", "10"),
                                     ("1", "1", "2", "4")):
        monkeypatch.setenv("PSW_TGV_DOUBLE", double)
        monkeypatch.setenv("PSW_TGV_TRIPLE", triple)
        monkeypatch.setenv("PSW_TGV_VARIANT", var)
        monkeypatch.setenv("PSW_TGV2_TY", ty2)
        t = stages.TgvSolver(W, Hh, N)
        t.set_images(d_ref, d_src, prm.beta, prm.gamma)
        t.set_u(d_init)
        t.linearise(K1, Rrel, trel)
        t.iterate(37, prm)
        t.iterate(5, prm)
        t.iterate(1, prm)
        states.append(t.get_state())
        # and the reference driver (warps x iterations) through the same kernels
        states[-1]["run"] = stages.tgv_run(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], 25, 2,
                                           prm)
    for st in states[1:]:
        for k in ("u", "u1x", "u1y", "px", "py", "qx", "qy", "qz", "r", "run"):
            np.testing.assert_array_equal(states[0][k], st[k], err_msg=k)


def test_tgv_set_primal_and_duals(gpu, views_small):
    """Warm start inside a linearisation: set_primal/set_duals load exactly
    the given fields, leave u0/It/Iu alone, and iterating from a loaded state
    equals iterating on without the reload (state round trip)."""
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    v = views_small
    K1 = one_based(v["K"])
    Hh, W = v["ref"].shape
    N = v["src"].shape[0]
    prm = stages.tgv_params(**TGV_KW)
    Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    t = stages.TgvSolver(W, Hh, N)
    d_ref, d_src = DeviceArray.from_host(v["ref"]), DeviceArray.from_host(v["src"])
    t.set_images(d_ref, d_src, prm.beta, prm.gamma)
    t.set_u(DeviceArray.from_host(np.clip(v["depth"], 0.5, 6.0).astype(np.float32)))
    t.linearise(K1, Rrel, trel)
    lin = t.get_state()
    rng = np.random.default_rng(4)
    u = (2.0 + rng.random((Hh, W))).astype(np.float32)
    u1 = (0.01 * rng.standard_normal((2, Hh, W))).astype(np.float32)
    p = (0.3 * rng.random((2, Hh, W))).astype(np.float32)
    q = (0.2 * rng.random((3, Hh, W))).astype(np.float32)
    r = (rng.random((N, Hh, W)) - 0.5).astype(np.float32)
    t.set_primal(DeviceArray.from_host(u), d_u1=DeviceArray.from_host(u1))
    t.set_duals(DeviceArray.from_host(p), DeviceArray.from_host(q), DeviceArray.from_host(r))
    s = t.get_state()
    np.testing.assert_array_equal(s["u"], u)
    np.testing.assert_array_equal(np.stack([s["u1x"], s["u1y"]]), u1)
    np.testing.assert_array_equal(np.stack([s["px"], s["py"]]), p)
    np.testing.assert_array_equal(np.stack([s["qx"], s["qy"], s["qz"]]), q)
    np.testing.assert_array_equal(s["r"], r)
    for k in ("u0", "It", "Iu", "T11"):
        np.testing.assert_array_equal(s[k], lin[k])
    # constant u with NULL fields: u1 and every dual zero
    t.set_primal(None, 1.5)
    t.set_duals()
    s = t.get_state()
    assert np.all(s["u"] == 1.5) and not np.any(s["u1x"]) and not np.any(s["px"]) and not np.any(s["r"])
    # iterate 7, reload the state (u_prev := u differs from the running
    # iterate, so compare from a restart on both sides)
    t.set_primal(DeviceArray.from_host(u), d_u1=DeviceArray.from_host(u1))
    t.set_duals(DeviceArray.from_host(p), DeviceArray.from_host(q), DeviceArray.from_host(r))
    t.iterate(7, prm)
    a = t.get_state()
    t.set_primal(DeviceArray.from_host(u), d_u1=DeviceArray.from_host(u1))
    t.set_duals(DeviceArray.from_host(p), DeviceArray.from_host(q), DeviceArray.from_host(r))
    t.iterate(7, prm)
    b = t.get_state()
    for k in ("u", "u1x", "u1y", "px", "py", "qx", "qy", "qz", "r"):
        np.testing.assert_array_equal(a[k], b[k])


def test_tgv_warm_start_and_state(gpu, views_small):
    """u_init replaces the cold start; state invariants |p|,|q|,|r| <= 1."""
    from psw_amd import stages
    v = views_small
    K1 = one_based(v["K"])
    prm = stages.tgv_params(**TGV_KW)
    u_init = np.clip(v["depth"], 0.5, 6.0).astype(np.float32)
    u, st = stages.tgv_run(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], 40, 2, prm,
                           u_init=u_init, state=True)
    u64 = oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], 40, 2, prec="f64",
                     u_init=u_init, **TGV_KW)
    assert (np.abs(u - u64) / np.abs(u64)).max() <= 1e-4
    eps = 1e-6
    assert np.max(np.hypot(st["px"], st["py"])) <= 1 + eps
    assert np.max(np.sqrt(st["qx"] ** 2 + st["qy"] ** 2 + 2 * st["qz"] ** 2)) <= 1 + eps
    assert np.max(np.abs(st["r"])) <= 1 + eps


def test_tgv_energy_matches_oracle(gpu, views_small):
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    v = views_small
    K1 = one_based(v["K"])
    prm = stages.tgv_params(**TGV_KW)
    H, W = v["ref"].shape
    N = v["src"].shape[0]
    Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    t = stages.TgvSolver(W, H, N)
    d_ref, d_src = DeviceArray.from_host(v["ref"]), DeviceArray.from_host(v["src"])
    t.set_images(d_ref, d_src)
    t.set_u(value=1.0)
    t.linearise(K1, Rrel, trel)
    t.iterate(25, prm)
    e_gpu = t.energy(prm)
    s = t.get_state()
    T4 = np.stack([s["T11"], s["T12"], s["T12"], s["T22"]])
    e_orc = oracle.tgv_energy(s["u"], s["u1x"], s["u1y"], s["u0"], T4, s["It"], s["Iu"], 1.0, 2.0, 1.5)
    print(f"energy gpu {e_gpu:.12e} oracle {e_orc:.12e}")
    assert abs(e_gpu - e_orc) <= 1e-12 * abs(e_orc)


def test_tgv_linearisation_vs_oracle(gpu, views_small):
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    v = views_small
    K1 = one_based(v["K"])
    H, W = v["ref"].shape
    N = v["src"].shape[0]
    u0 = np.clip(v["depth"] * 1.03, 0.5, 6.0).astype(np.float32)
    Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    t = stages.TgvSolver(W, H, N)
    t.set_images(DeviceArray.from_host(v["ref"]), DeviceArray.from_host(v["src"]))
    t.set_u(DeviceArray.from_host(u0))
    t.linearise(K1, Rrel, trel)
    s = t.get_state()
    It, Iu, T4 = oracle.tgv_linearise(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], u0)
    np.testing.assert_array_equal(s["T11"], T4[0])
    np.testing.assert_array_equal(s["T12"], T4[1])
    np.testing.assert_array_equal(s["T22"], T4[3])
    np.testing.assert_array_equal(s["It"], It)
    # Iu: the oracle keeps the reference's double products, the kernel is FP32
    assert np.max(np.abs(s["Iu"] - Iu)) <= 1e-5 * max(1.0, np.abs(Iu).max())


def test_tgv_solve_stops_at_gap(gpu, views_small):
    """psw_tgv_solve stops once E - E* <= tol E* (E* from a long run); an
    energy below E* stops too (at the first check)."""
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    v = views_small
    K1 = one_based(v["K"])
    prm = stages.tgv_params(**TGV_KW)
    H, W = v["ref"].shape
    N = v["src"].shape[0]
    Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    d_ref, d_src = DeviceArray.from_host(v["ref"]), DeviceArray.from_host(v["src"])

    def fresh():
        t = stages.TgvSolver(W, H, N)
        t.set_images(d_ref, d_src)
        t.set_u(value=1.0)
        t.linearise(K1, Rrel, trel)
        return t

    long = fresh()
    long.iterate(3000, prm)
    e_star = long.energy(prm)
    t = fresh()
    iters, e = t.solve(3000, 10, 1e-3, e_star, prm)
    print(f"E* {e_star:.6e}; stopped after {iters} iterations at E {e:.6e}")
    assert iters < 3000 and iters % 10 == 0
    assert e - e_star <= 1e-3 * e_star
    assert abs(t.energy(prm) - e) <= 1e-12 * e   # the state is the one at the stop
    # the same stop as testing the energy after every chunk of 10 with the
    # stand-alone energy pass (the solve tests it inside the next launch)
    ref = fresh()
    k = 0
    while True:
        ref.iterate(10, prm)
        k += 10
        if ref.energy(prm) - e_star <= 1e-3 * e_star:
            break
    assert k == iters
    sa, sb = ref.get_state(), t.get_state()
    for name in ("u", "u1x", "u1y", "px", "py", "qx", "qy", "qz", "r"):
        np.testing.assert_array_equal(sa[name], sb[name], err_msg=name)
    # one-sided: a target far above the energy after 10 iterations stops there
    low = fresh()
    low.iterate(10, prm)
    e10 = low.energy(prm)
    t = fresh()
    iters, e = t.solve(3000, 10, 1e-3, 4.0 * e10, prm)
    assert iters == 10 and e == pytest.approx(e10, rel=1e-12)


@pytest.mark.parametrize("variant,double,ty2,triple,inflight", [
    ("0", "1", "10", "0", "2"), ("1", "1", "10", "0", "2"), ("2", "1", "10", "0", "2"), ("1", "0", "10", "0", "2"),
    ("1", "1", "4", "0", "2"), ("1", "1", "8", "0", "2"), ("1", "1", "12", "0", "2"), ("1", "1", "10", "1", "2"),
    ("2", "1", "8", "1", "2"), ("1", "1", "10", "0", "5"), ("1", "0", "10", "0", "8")])
def test_tgv_solve_max_iters_and_variants(gpu, views_small, monkeypatch, variant, double, ty2, triple, inflight):
    """Every launch shape, with and without two- and three-iteration
    launches, and with 2, 5 or 8 chunks in flight, stops at the same
    iteration with the same state; a solve that reaches max_iters reports
    max_iters and the energy of the final state."""
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    monkeypatch.setenv("PSW_TGV_INFLIGHT", inflight)
    monkeypatch.setenv("PSW_TGV_VARIANT", variant)
    monkeypatch.setenv("PSW_TGV_DOUBLE", double)
    monkeypatch.setenv("PSW_TGV2_TY", ty2)
    monkeypatch.setenv("PSW_TGV_TRIPLE", triple)
    v = views_small
    K1 = one_based(v["K"])
    prm = stages.tgv_params(**TGV_KW)
    H, W = v["ref"].shape
    N = v["src"].shape[0]
    Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    d_ref, d_src = DeviceArray.from_host(v["ref"]), DeviceArray.from_host(v["src"])
    t = stages.TgvSolver(W, H, N)
    t.set_images(d_ref, d_src)
    t.set_u(value=1.0)
    t.linearise(K1, Rrel, trel)
    iters, e = t.solve(95, 10, 1e-9, 1.0, prm)          # E* = 1: never within tolerance
    assert iters == 95
    assert abs(t.energy(prm) - e) <= 1e-12 * e
    t.set_u(value=1.0)
    t.linearise(K1, Rrel, trel)
    t.iterate(400, prm)
    e_star = t.energy(prm)
    t.set_u(value=1.0)
    t.linearise(K1, Rrel, trel)
    it2, e2 = t.solve(400, 7, 2e-2, e_star, prm)
    assert it2 < 400 and it2 % 7 == 0 and e2 - e_star <= 2e-2 * e_star
    assert abs(t.energy(prm) - e2) <= 1e-12 * e2
    s_stop = t.get_state()
    t.iterate(3, prm)                                    # resumes from the stop (needs u_{k-1} intact)
    ref = stages.TgvSolver(W, H, N)
    ref.set_images(d_ref, d_src)
    ref.set_u(value=1.0)
    ref.linearise(K1, Rrel, trel)
    ref.iterate(it2, prm)
    s_ref = ref.get_state()
    for name in ("u", "u1x", "u1y", "px", "py", "qx", "qy", "qz", "r"):
        np.testing.assert_array_equal(s_stop[name], s_ref[name], err_msg=name)
    ref.iterate(3, prm)
    np.testing.assert_array_equal(t.get_u(), ref.get_u())


# ---- stage 3: fusion ---------------------------------------------------------------------

def _fusion_inputs(n=3, W=96, H=72, grid=(24, 20, 28)):
    from psw_amd import scene
    o = scene.orbit_depthmaps(seed=11, n=n, width=W, height=H)
    vol = np.array([-3.0, -1.6, -4.0, 3.0, 1.6, 6.0], np.float32)
    return o, vol, grid


def test_fusion_histogram_bit_exact(gpu):
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    o, vol, (gw, gh, gd) = _fusion_inputs()
    thr = 0.25
    f = stages.Fusion(gw, gh, gd, vol)
    hist = np.zeros((gd, gh, gw, 8), np.uint16)
    H, W = o["depths"].shape[1:]
    for k in range(o["depths"].shape[0]):
        d = DeviceArray.from_host(o["depths"][k])
        f.update_histogram(d, W, H, o["K"], o["R"][k], o["t"][k], thr)
        oracle.fusion_hist(hist, vol, o["depths"][k], o["K"], o["R"][k], o["t"][k], thr)
    got = f.download()["hist"]
    print(f"votes: {int(hist.sum())}")
    assert hist.sum() > 0
    np.testing.assert_array_equal(got, hist)
    # batched launch == sequential updates
    g2 = stages.Fusion(gw, gh, gd, vol)
    dall = DeviceArray.from_host(o["depths"])
    g2.update_histogram_batch(dall, o["depths"].shape[0], W, H, np.repeat(o["K"][None], o["R"].shape[0], 0), o["R"],
                              o["t"], thr)
    np.testing.assert_array_equal(g2.download()["hist"], hist)


def test_fusion_pd_vs_oracle(gpu):
    """U/P steps, interleaved with histograms (FusionUpdateIteration) and the
    fused solve, against the FP64 oracle fed the same histograms."""
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    o, vol, (gw, gh, gd) = _fusion_inputs(n=4)
    thr, tau, lam, sig = 0.25, 0.1, 0.5, 1.0
    H, W = o["depths"].shape[1:]
    f = stages.Fusion(gw, gh, gd, vol)
    hist = np.zeros((gd, gh, gw, 8), np.uint16)
    u64, v64, p64 = np.zeros((gd, gh, gw)), np.zeros((gd, gh, gw)), np.zeros((3, gd, gh, gw))
    u32, v32, p32 = [np.zeros_like(a, np.float32) for a in (u64, v64, p64)]
    for k in range(o["depths"].shape[0]):
        d = DeviceArray.from_host(o["depths"][k])
        f.update_iteration(d, W, H, o["K"], o["R"][k], o["t"][k], thr, tau, lam, sig)
        oracle.fusion_hist(hist, vol, o["depths"][k], o["K"], o["R"][k], o["t"][k], thr)
        for (u, v, p, prec) in ((u64, v64, p64, "f64"), (u32, v32, p32, "f32")):
            oracle.fusion_u(u, v, p, hist, tau, lam, prec=prec)
            oracle.fusion_p(p, v, sig, prec=prec)
    g = f.download()
    np.testing.assert_array_equal(g["hist"], hist)
    e = rel_err(g["u"], u64)
    print(f"interleaved: u vs FP64 max {e.max():.3e}; vs FP32 {rel_err(g['u'], u32).max():.3e}")
    assert e.max() <= 1e-4
    # then the fused solve
    n = 60
    f.solve(n, tau, lam, sig)
    for _ in range(n):
        oracle.fusion_u(u64, v64, p64, hist, tau, lam, prec="f64")
        oracle.fusion_p(p64, v64, sig, prec="f64")
    g = f.download()
    e = rel_err(g["u"], u64)
    ep = np.abs(g["p"] - p64).max()
    print(f"solve {n}: u vs FP64 max {e.max():.3e} mean {e.mean():.3e}; p max abs {ep:.3e}")
    assert e.max() <= 1e-4
    assert ep <= 1e-4
    assert f.surface_count() == oracle.surface_count(g["u"])
    # FP64 energy of the solved u: GPU block partials vs the oracle's sequential sum
    e_gpu = f.energy(lam)
    e_orc = oracle.fusion_energy(g["u"].astype(np.float64), g["hist"], lam)
    print(f"fusion energy GPU {e_gpu:.12e} oracle {e_orc:.12e}")
    assert abs(e_gpu - e_orc) <= 1e-12 * abs(e_orc)


@pytest.mark.parametrize("niters", [41, 42, 2])
def test_fusion_solve_kernels_identical(gpu, monkeypatch, niters):
    """The z-marching PD kernels (registers only, with LDS-shared neighbour
    duals, with pipelined loads, two steps per launch) and the
    one-voxel-per-thread kernel give bit-identical grids (depth 28 leaves a
    partial z chunk; 20 rows a partial 8-row block; 42 iterations end the
    two-step launches with a one-step launch)."""
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    o, vol, (gw, gh, gd) = _fusion_inputs(n=3)
    H, W = o["depths"].shape[1:]
    out = []
    print(f"[case] fusion kernels identical, {niters} iterations", flush=True)
    for mode in ("1", "0", "2", "3", "4"):
        monkeypatch.setenv("PSW_FUSION_PD", mode)
        f = stages.Fusion(gw, gh, gd, vol)
        for k in range(o["depths"].shape[0]):
            f.update_histogram(DeviceArray.from_host(o["depths"][k]), W, H, o["K"], o["R"][k], o["t"][k], 0.25)
        f.solve(niters, 0.1, 0.5, 1.0)
        out.append(f.download())
    for o in out[1:]:
        for key in ("u", "v", "p", "hist"):
            np.testing.assert_array_equal(out[0][key], o[key])


@pytest.mark.parametrize("top", [255, 1000])
def test_fusion_solve_histogram_packing(gpu, monkeypatch, top):
    """The pipelined PD kernel reads the counts from a byte-packed copy when
    they all fit in a byte (top 255) and from the 16-bit histogram otherwise
    (top 1000); both give the one-voxel-per-thread kernel's grids bit for
    bit, and u stays within 1e-4 of the FP64 oracle."""
    from psw_amd import stages
    rng = np.random.default_rng(top)
    vol = np.array([-1.0, -1.0, -1.0, 1.0, 1.0, 1.0], np.float32)
    gw, gh, gd = 70, 21, 35
    hist = rng.integers(0, 6, (gd, gh, gw, 8)).astype(np.uint16)
    hot = rng.random((gd, gh, gw, 8)) < 0.01
    hist[hot] = rng.integers(200, top + 1, int(hot.sum()))
    print(f"[case] histogram packing, top count {top}, max {int(hist.max())}", flush=True)
    tau, lam, sig, n = 0.1, 0.5, 1.0, 9
    out = []
    for mode in ("0", "3", "4"):
        monkeypatch.setenv("PSW_FUSION_PD", mode)
        f = stages.Fusion(gw, gh, gd, vol)
        f.upload(hist=hist)
        f.solve(n, tau, lam, sig)
        out.append(f.download())
    for o in out[1:]:
        for key in ("u", "v", "p", "hist"):
            np.testing.assert_array_equal(out[0][key], o[key], err_msg=key)
    u64, v64, p64 = np.zeros((gd, gh, gw)), np.zeros((gd, gh, gw)), np.zeros((3, gd, gh, gw))
    for _ in range(n):
        oracle.fusion_u(u64, v64, p64, hist, tau, lam, prec="f64")
        oracle.fusion_p(p64, v64, sig, prec="f64")
    e = rel_err(out[2]["u"], u64)
    print(f"u vs FP64 max {e.max():.3e}")
    assert e.max() <= 1e-4


# ---- TVL1 ---------------------------------------------------------------------------------

def test_tvl1_vs_oracle(gpu, views_small):
    from psw_amd import stages
    v = views_small
    depth = np.clip(v["depth"] + 0.05 * np.sin(np.arange(v["depth"].size)).reshape(v["depth"].shape), 0.5, 6.0)
    depth = depth.astype(np.float32)
    got = stages.tvl1_denoise(depth, v["ref"], 0.5, 6.0, niters=50)
    e32 = oracle.tvl1(depth, v["ref"], 0.5, 6.0, niters=50)
    e64 = oracle.tvl1(depth, v["ref"], 0.5, 6.0, niters=50, prec="f64")
    print(f"TVL1 vs f32 oracle max abs {np.abs(got - e32).max():.3e}; vs f64 max rel {rel_err(got, e64).max():.3e}")
    assert rel_err(got, e64).max() <= 1e-4
    assert np.abs(got - e32).max() <= 1e-5


@pytest.mark.parametrize("niters,ty,merged,mmax,narrow", [
    (0, "8", "1", "2", "1"), (1, "8", "1", "2", "1"), (2, "8", "1", "2", "1"), (3, "8", "1", "2", "1"),
    (50, "8", "1", "2", "1"), (3, "4", "1", "2", "1"), (50, "4", "1", "2", "1"), (3, "16", "1", "2", "1"),
    (50, "16", "1", "2", "1"), (3, "8", "1", "3", "1"), (50, "8", "1", "3", "1"), (4, "8", "1", "4", "1"),
    (7, "8", "1", "4", "1"), (50, "8", "1", "4", "1"), (50, "16", "1", "4", "1"), (49, "4", "1", "4", "1"),
    (1, "8", "1", "2", "0"), (3, "8", "1", "2", "0"), (50, "8", "1", "2", "0"), (50, "4", "1", "2", "0"),
    (50, "16", "1", "2", "0"), (3, "8", "0", "2", "0"), (50, "8", "0", "2", "0"), (50, "8", "1", "3", "0"),
    (7, "8", "1", "4", "0"), (50, "16", "1", "4", "0"), (50, "8", "1", "5", "1"), (11, "8", "1", "6", "1"),
    (50, "4", "1", "6", "1"), (50, "8", "1", "6", "0"), (50, "16", "1", "6", "1")])
def test_tvl1_fused_matches_launch_sequence(gpu, monkeypatch, views_small, niters, ty, merged, mmax, narrow):
    """The fused TVL1 launches of up to mmax iterations (tvl1_fused.hip, tiles
    ty rows high and 64 - 2m (narrow) or 64 columns wide; with 64-wide tiles
    two iterations run in three phases, or in five with merged=0)
    against the reference's two launches per iteration (PSW_TVL1_FUSED=0):
    bit-identical on the 160x120 view and on sizes that cut the tiles (65x9,
    3x2, 130x17), and within the oracle tolerance of the FP32 oracle."""
    from psw_amd import stages
    v = views_small
    rng = np.random.default_rng(100 + niters)
    cases = [(v["depth"], v["ref"])]
    for h, w in [(9, 65), (2, 3), (17, 130)]:
        cases.append((rng.uniform(0.5, 6.0, (h, w)), rng.uniform(0, 255, (h, w))))
    for depth, ref in cases:
        depth = np.clip(depth + 0.05 * rng.standard_normal(depth.shape), 0.5, 6.0).astype(np.float32)
        ref = np.asarray(ref, np.float32)
        print(f"[case] tvl1 {depth.shape[1]}x{depth.shape[0]} niters={niters} ty={ty} merged={merged} m={mmax} "
              f"narrow={narrow}", flush=True)
        monkeypatch.setenv("PSW_TVL1_FUSED", "0")
        seq = stages.tvl1_denoise(depth, ref, 0.5, 6.0, niters=niters)
        monkeypatch.setenv("PSW_TVL1_FUSED", "1")
        monkeypatch.setenv("PSW_TVL1_RESIDENT", "0")
        monkeypatch.setenv("PSW_TVL1_TY", ty)
        monkeypatch.setenv("PSW_TVL1_MERGED", merged)
        monkeypatch.setenv("PSW_TVL1_M", mmax)
        monkeypatch.setenv("PSW_TVL1_NARROW", narrow)
        fused = stages.tvl1_denoise(depth, ref, 0.5, 6.0, niters=niters)
        np.testing.assert_array_equal(fused.view(np.uint32), seq.view(np.uint32))
        if depth.size < 5000:
            e32 = oracle.tvl1(depth, ref, 0.5, 6.0, niters=niters)
            assert np.abs(fused - e32).max() <= 1e-5


@pytest.mark.parametrize("niters", [2, 3, 50])
def test_tvl1_resident_matches_launch_sequence(gpu, monkeypatch, views_small, niters):
    """The resident TVL1 solver (all iterations in one cooperative launch of
    64x20 tiles, tvl1_fused.hip; PSW_TVL1_RESIDENT=2 makes it an error if it
    does not run) against the reference's two launches per iteration:
    bit-identical on the 160x120 view, on sizes that cut the tiles (65x9,
    3x2, 130x17, 200x41) and at 640x480 (240 tiles)."""
    from psw_amd import stages
    v = views_small
    rng = np.random.default_rng(200 + niters)
    cases = [(v["depth"], v["ref"])]
    for h, w in [(9, 65), (2, 3), (17, 130), (41, 200), (480, 640)]:
        cases.append((rng.uniform(0.5, 6.0, (h, w)), rng.uniform(0, 255, (h, w))))
    for depth, ref in cases:
        depth = np.clip(depth + 0.05 * rng.standard_normal(depth.shape), 0.5, 6.0).astype(np.float32)
        ref = np.asarray(ref, np.float32)
        print(f"[case] tvl1 resident {depth.shape[1]}x{depth.shape[0]} niters={niters}", flush=True)
        monkeypatch.setenv("PSW_TVL1_FUSED", "0")
        seq = stages.tvl1_denoise(depth, ref, 0.5, 6.0, niters=niters)
        monkeypatch.setenv("PSW_TVL1_FUSED", "1")
        monkeypatch.setenv("PSW_TVL1_RESIDENT", "2")
        res = stages.tvl1_denoise(depth, ref, 0.5, 6.0, niters=niters)
        np.testing.assert_array_equal(res.view(np.uint32), seq.view(np.uint32))


def test_tvl1_resident_fallback_and_failure(gpu, monkeypatch, views_small):
    """Frames above 256 tiles (1242x375: 380) use the fused launches, with the
    same bits; a wait that runs out (a zero poll budget, forced from the host)
    is reported as an error; and the next run is whole again."""
    from psw_amd import stages
    from psw_amd._lib import PswError
    rng = np.random.default_rng(7)
    depth = rng.uniform(0.5, 6.0, (375, 1242)).astype(np.float32)
    ref = rng.uniform(0, 255, (375, 1242)).astype(np.float32)
    print("[case] tvl1 1242x375 resident on (falls back) vs launch sequence", flush=True)
    monkeypatch.setenv("PSW_TVL1_FUSED", "0")
    seq = stages.tvl1_denoise(depth, ref, 0.5, 6.0, niters=6)
    monkeypatch.setenv("PSW_TVL1_FUSED", "1")
    monkeypatch.setenv("PSW_TVL1_RESIDENT", "1")
    got = stages.tvl1_denoise(depth, ref, 0.5, 6.0, niters=6)
    np.testing.assert_array_equal(got.view(np.uint32), seq.view(np.uint32))
    monkeypatch.setenv("PSW_TVL1_RESIDENT", "2")
    with pytest.raises(PswError):
        stages.tvl1_denoise(depth, ref, 0.5, 6.0, niters=6)
    v = views_small
    d = np.clip(v["depth"], 0.5, 6.0).astype(np.float32)
    print("[case] tvl1 160x120 resident with a zero poll budget", flush=True)
    monkeypatch.setenv("PSW_TVL1_RESIDENT_SPIN", "0")
    with pytest.raises(PswError) as e:
        stages.tvl1_denoise(d, v["ref"], 0.5, 6.0, niters=5)
    assert "resident TVL1 solver" in str(e.value)
    monkeypatch.delenv("PSW_TVL1_RESIDENT_SPIN")
    monkeypatch.setenv("PSW_TVL1_FUSED", "0")
    seq = stages.tvl1_denoise(d, v["ref"], 0.5, 6.0, niters=5)
    monkeypatch.setenv("PSW_TVL1_FUSED", "1")
    got = stages.tvl1_denoise(d, v["ref"], 0.5, 6.0, niters=5)
    np.testing.assert_array_equal(got.view(np.uint32), seq.view(np.uint32))


def test_tvl1_resident_concurrent_calls(gpu, monkeypatch):
    """Two threads denoising different frames on their own streams at once
    (the resident runs of one device take turns on its exchange buffer): the
    same bits as each frame alone."""
    import threading
    from psw_amd import stages
    from psw_amd.device import DeviceArray, Stream
    monkeypatch.setenv("PSW_TVL1_RESIDENT", "2")
    rng = np.random.default_rng(5)
    H, W = 120, 160
    frames = [(rng.uniform(0.5, 6.0, (H, W)).astype(np.float32), rng.uniform(0, 255, (H, W)).astype(np.float32))
              for _ in range(2)]
    alone = [stages.tvl1_denoise(d, r, 0.5, 6.0, niters=200) for d, r in frames]
    dev = [(DeviceArray.from_host(d), DeviceArray.from_host(r), DeviceArray((H, W)), DeviceArray((12, H, W)), Stream())
           for d, r in frames]
    errors = []

    def run(k):
        d_d, d_r, d_o, d_w, st = dev[k]
        try:
            for _ in range(5):
                stages.tvl1_denoise_device(d_d, d_r, W, H, 0.5, 6.0, d_o, d_w, 200, stream=st.handle)
                st.synchronize()
                if not np.array_equal(d_o.to_host().view(np.uint32), alone[k].view(np.uint32)):
                    errors.append(k)
        except Exception as e:   # noqa: BLE001  (reported below)
            errors.append(repr(e))

    print("[case] tvl1 resident, two threads at once", flush=True)
    th = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th)
    assert not errors, errors


# ---- error paths (forced from the host) -----------------------------------------------------

def test_error_paths(gpu, views_tiny):
    from psw_amd import stages
    from psw_amd._lib import PswError
    v = views_tiny
    K1 = one_based(v["K"])
    planes = stages.reference_planes(0.5, 6.0, 4)
    with pytest.raises(PswError) as e:   # even window (PlaneSweep::setWindowSize, planesweep.h:234)
        stages.sweep_reference(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, 4, STD_T,
                               NCC_T, 6.0)
    assert e.value.status == 1
    with pytest.raises(PswError):        # too many sources
        stages.TgvSolver(8, 8, 17)
    from psw_amd._lib import Dim3, call
    from psw_amd.device import DeviceArray
    d = DeviceArray((4, 4))
    with pytest.raises(PswError) as e:   # block above 1024 threads: launch error
        call("psw_set_value", d.ptr, 1.0, 4, 4, Dim3(1, 1, 1), Dim3(64, 32, 1), None)
    assert e.value.status == 2
