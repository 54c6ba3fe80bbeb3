"""Real frames: the reference's ICL-NUIM images and calibration
(tests/golden/real_icl_nuim.npz, written by tests/golden/make_real_fixture.py:
reference frame + 4 sources, 320x240 gray, the reference's K with its skew
term).  CPU: fixture sanity and the oracle's depths in the scene's range.
GPU: the sweep (both reductions) bit-identical to the oracle and TGV2 within
the parity bar on real images, with both relative-pose methods."""
import os

import numpy as np
import pytest

import oracle
from conftest import ROOT

FIX = os.path.join(ROOT, "tests", "golden", "real_icl_nuim.npz")
ZNEAR, ZFAR = 0.1, 1.0          # DEFAULT_Z_NEAR / DEFAULT_Z_FAR (defines.h:24-25)


def _load():
    with np.load(FIX) as z:
        return {k: z[k] for k in z.files}


def test_fixture_sanity():
    f = _load()
    assert f["images"].shape == (5, 240, 320) and f["images"].dtype == np.uint8
    for R in f["R"]:
        np.testing.assert_allclose(R.astype(np.float64) @ R.T, np.eye(3), atol=5e-5)
    assert f["K"][0, 1] > 0                       # the ATAN distortion parameter sits in the skew slot


def test_oracle_depths_in_scene_range():
    """Aggregated sweep over 48 planes with the default pose method: strong
    matches (mean cost below -0.8) and depths inside the reference's default
    range for most pixels."""
    f = _load()
    img = f["images"].astype(np.float32)
    planes = (1.0 / np.linspace(1 / 0.3, 1 / 2.0, 48)).astype(np.float32)
    a, c, _ = oracle.sweep_agg(img[0], img[1:], f["K"], f["R"][0], f["t"][0], f["R"][1:], f["t"][1:], planes, 7,
                               1e-4, coord_offset=0)
    d = planes[np.clip(a, 0, planes.size - 1)]
    assert c.mean() < -0.8
    assert np.mean((d > ZNEAR) & (d < ZFAR)) > 0.9


@pytest.mark.gpu
@pytest.mark.parametrize("alternative", [0, 1])
def test_real_sweep_bit_exact(gpu, alternative):
    from psw_amd import stages
    from conftest import one_based
    f = _load()
    img = f["images"].astype(np.float32)
    K1 = one_based(f["K"])
    pose = (f["R"][0], f["t"][0], f["R"][1:], f["t"][1:])
    planes = stages.reference_planes(ZNEAR, ZFAR, 40)
    got = stages.sweep_reference(img[0], img[1:], K1, *pose, planes, 7, 1e-4, 0.5, ZFAR, alternative=alternative)
    exp = oracle.sweep_ref(img[0], img[1:], K1, *pose, planes, 7, 1e-4, 0.5, ZFAR, alternative=alternative)
    np.testing.assert_array_equal(got, exp)
    ip = (1.0 / np.linspace(1 / 0.3, 1 / 2.0, 48)).astype(np.float32)
    ga = stages.sweep_aggregate(img[0], img[1:], f["K"], *pose, ip, 7, 1e-4, coord_offset=0, fast=False,
                                alternative=alternative)
    ea = oracle.sweep_agg(img[0], img[1:], f["K"], *pose, ip, 7, 1e-4, coord_offset=0, alternative=alternative)
    for g, e in zip(ga, ea):
        np.testing.assert_array_equal(g, e)


@pytest.mark.gpu
def test_real_tgv_vs_fp64(gpu):
    """TGV2 (2 linearisations x 30 iterations) from the oracle's aggregated
    sweep depth on the real frames, against the FP64 oracle.

    On these frames a few pixels within 16 px of the border, whose warps into
    the sources cross the image edge (a hard in/out test of the warp), end
    far apart in FP32 and FP64: the FP32 oracle itself is off by up to 2x
    there (p99.9 1.6e-3 over the frame).  So the synthetic-scene bar
    (p99.9 <= 1e-4, max <= 1e-3) applies to the interior, and over the full
    frame the GPU may be no further from FP64 than three times the FP32
    oracle's spread."""
    from psw_amd import stages
    from conftest import one_based
    f = _load()
    img = f["images"].astype(np.float32)
    K1 = one_based(f["K"])
    pose = (f["R"][0], f["t"][0], f["R"][1:], f["t"][1:])
    ip = (1.0 / np.linspace(1 / 0.3, 1 / 2.0, 48)).astype(np.float32)
    a, _, _ = oracle.sweep_agg(img[0], img[1:], f["K"], *pose, ip, 7, 1e-4, coord_offset=0)
    u0 = ip[np.clip(a, 0, ip.size - 1)]
    kw = dict(lambda_=1.0, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, beta=0.0, gamma=1.0)
    prm = stages.tgv_params(**kw)
    u = stages.tgv_run(img[0], img[1:], K1, *pose, 30, 2, prm, u_init=u0)
    u64 = oracle.tgv(img[0], img[1:], K1, *pose, 30, 2, prec="f64", u_init=u0, **kw)
    u32 = oracle.tgv(img[0], img[1:], K1, *pose, 30, 2, prec="f32", u_init=u0, **kw)
    e = np.abs(u - u64) / np.abs(u64)
    e32 = np.abs(u32 - u64) / np.abs(u64)
    B = 16
    ei = e[B:-B, B:-B]
    print(f"real frames TGV2 vs FP64: interior max {ei.max():.3e} p99.9 {np.percentile(ei, 99.9):.3e}; "
          f"frame p99.9 {np.percentile(e, 99.9):.3e} (FP32 oracle {np.percentile(e32, 99.9):.3e}), "
          f"pixels > 1e-3: {(e > 1e-3).sum()} (FP32 oracle {(e32 > 1e-3).sum()})")
    assert np.percentile(ei, 99.9) <= 1e-4 and ei.max() <= 1e-3
    assert np.percentile(e, 99.9) <= max(1e-4, 3 * np.percentile(e32, 99.9))
    assert (e > 1e-3).sum() <= max(10, 3 * (e32 > 1e-3).sum())
