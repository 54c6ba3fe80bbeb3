"""The reference-compatible C++ API (plane-sweep-cuda_amd/host/) driven the
way the reference's GUI drives PlaneSweep and fusionData<8>, compared with the
oracle.  Runs tests/cpp/_build/host_api_test on files written here."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT, one_based, rel_err

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "_build", "host_api_test")


def test_host_api_end_to_end(gpu, views_tiny, tmp_path):
    from psw_amd import scene
    v = views_tiny
    H, W = v["ref"].shape
    N = v["src"].shape[0]
    K1 = one_based(v["K"])
    znear, zfar, P, win, stdt, ncct = 0.5, 6.0, 16, 5, 1e-4, 0.5
    tgv_iters, tgv_warps, lam = 20, 2, 0.5
    with open(tmp_path / "scene.bin", "wb") as f:
        np.array([W, H, N, P], np.int32).tofile(f)
        np.array([znear, zfar, win, stdt, ncct, tgv_iters, tgv_warps, lam], np.float32).tofile(f)
        for a in (K1, v["Rref"], v["tref"]):
            np.asarray(a, np.float32).tofile(f)
        for i in range(N):
            np.asarray(v["Rsrc"][i], np.float32).tofile(f)
            np.asarray(v["tsrc"][i], np.float32).tofile(f)
        v["ref"].astype(np.float32).tofile(f)
        v["src"].astype(np.float32).tofile(f)
    o = scene.orbit_depthmaps(seed=4, n=3, width=48, height=36)
    gw, gh, gd = 12, 10, 14
    vol = np.array([-3.0, -1.6, -4.0, 3.0, 1.6, 6.0], np.float32)
    thr, tau, fl, sig = 0.3, 0.1, 0.5, 1.0
    with open(tmp_path / "fusion.bin", "wb") as f:
        np.array([48, 36, 3, gw, gh, gd], np.int32).tofile(f)
        vol.tofile(f)
        np.array([thr, tau, fl, sig], np.float32).tofile(f)
        np.asarray(o["K"], np.float32).tofile(f)
        for k in range(3):
            np.asarray(o["R"][k], np.float32).tofile(f)
            np.asarray(o["t"][k], np.float32).tofile(f)
        o["depths"].astype(np.float32).tofile(f)
    print("[case] host_api_test", flush=True)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.returncode
    rd = lambda name, dt=np.float32: np.fromfile(tmp_path / name, dt)  # noqa: E731
    planes = oracle.reference_planes(znear, zfar, P)
    exp = oracle.sweep_ref(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, win, stdt,
                           ncct, zfar)
    np.testing.assert_array_equal(rd("out_sweep.bin").reshape(H, W), exp)
    # ConvertDepthtoUChar (planesweep.cpp:279-290)
    d8 = np.where(exp == exp, (255 * np.clip((exp - znear) / (zfar - znear), 0, 1)).astype(np.uint8), 255)
    np.testing.assert_array_equal(rd("out_sweep8u.bin", np.uint8).reshape(H, W), d8)
    tv = oracle.tvl1(exp, v["ref"], znear, zfar, niters=30)
    assert np.abs(rd("out_tvl1.bin").reshape(H, W) - tv).max() <= 1e-5
    tg = oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], tgv_iters, tgv_warps,
                    lambda_=lam, prec="f64")
    assert rel_err(rd("out_tgv.bin").reshape(H, W), tg).max() <= 1e-4
    tvd = rd("out_tvl1.bin").reshape(H, W)
    for name, e in zip(("out_x.bin", "out_y.bin", "out_z.bin"), oracle.compute3d(tvd, K1, v["Rref"], v["tref"])):
        np.testing.assert_array_equal(rd(name).reshape(H, W), e)
    sparse = np.where(np.arange(H * W).reshape(H, W) % 7 == 0, tvd, 0).astype(np.float32)
    ts = oracle.tgv_sparse(v["ref"], exp, sparse, zfar, 20, prec="f64")
    assert rel_err(rd("out_tgv_sparse.bin").reshape(H, W), ts).max() <= 1e-4
    refmean = oracle.windowed_mean(oracle.windowed_mean(v["ref"], win, False, "col"), win, False, "row")
    np.testing.assert_array_equal(rd("out_refmean.bin").reshape(H, W), refmean)
    # fusion through fusionData<8> / FusionUpdateIteration<8>
    hist = np.zeros((gd, gh, gw, 8), np.uint16)
    u, vv, p = np.zeros((gd, gh, gw)), np.zeros((gd, gh, gw)), np.zeros((3, gd, gh, gw))
    for k in range(3):
        oracle.fusion_hist(hist, vol, o["depths"][k], o["K"], o["R"][k], o["t"][k], thr)
        oracle.fusion_u(u, vv, p, hist, tau, fl, prec="f64")
        oracle.fusion_p(p, vv, sig, prec="f64")
    np.testing.assert_array_equal(rd("out_fusion_hist.bin", np.uint8).reshape(gd, gh, gw, 8), hist.astype(np.uint8))
    assert rel_err(rd("out_fusion_u.bin").reshape(gd, gh, gw), u).max() <= 1e-4
    c = rd("out_world.bin")
    a, b = vol[:3], vol[3:]
    frac = np.array([(1 + .5) / gw, (2 + .5) / gh, (3 + .5) / gd]).astype(np.float32)
    np.testing.assert_array_equal(c, a + (b - a) * frac)
