"""Committed golden vectors (tests/golden/golden_small.npz, written by
tests/golden/make_golden.py from seeded inputs through the oracle).

CPU: the oracle and the scene generator reproduce the fixture exactly.
GPU: the HIP kernels reproduce the fixture (bit-exact where the kernels claim
bit-exactness, within the parity tolerance elsewhere)."""
import os

import numpy as np
import pytest

import oracle
from conftest import ROOT, rel_err

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden_small.npz")


def _load():
    z = np.load(GOLDEN)     # allow_pickle=False (default): data only
    inp = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return inp, out


def test_scene_generator_reproduces_fixture_inputs():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(ROOT, "tests", "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    inp, _ = _load()
    again = mg.make_inputs()
    for k, v in inp.items():
        np.testing.assert_array_equal(np.asarray(again[k]), v, err_msg=k)


def test_oracle_reproduces_fixture_outputs():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(ROOT, "tests", "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    inp, out = _load()
    again = mg.compute(inp)
    for k, v in out.items():
        np.testing.assert_array_equal(again[k], v, err_msg=k)


@pytest.mark.gpu
def test_gpu_matches_fixture(gpu):
    from psw_amd import scene, stages
    from psw_amd.device import DeviceArray
    inp, out = _load()
    K1 = scene.one_based(inp["K"])
    pose = (inp["Rref"], inp["tref"], inp["Rsrc"], inp["tsrc"])
    d = stages.sweep_reference(inp["ref"], inp["src"], K1, *pose, out["planes"], 5, 1e-4, 0.5, 6.0)
    np.testing.assert_array_equal(d, out["sweep_ref_f32"])
    a, c, s = stages.sweep_aggregate(inp["ref"], inp["src"], inp["K"], *pose, inp["agg_planes"], 7, 1e-4, fast=False)
    np.testing.assert_array_equal(a, out["agg_arg_f32"])
    np.testing.assert_array_equal(c, out["agg_cmin_f32"])
    np.testing.assert_array_equal(s, out["agg_csec_f32"])
    assert np.mean(a == out["agg_arg_f64"]) >= 0.999
    u = stages.tgv_run(inp["ref"], inp["src"], K1, *pose, 40, 2, stages.tgv_params(lambda_=1.0))
    e = rel_err(u, out["tgv_f64"])
    assert np.percentile(e, 99.9) <= 1e-4 and e.max() <= 1e-3
    gd, gh, gw = inp["grid"]
    f = stages.Fusion(gw, gh, gd, inp["fvol"])
    H, W = inp["fdepths"].shape[1:]
    for k in range(inp["fdepths"].shape[0]):
        f.update_histogram(DeviceArray.from_host(inp["fdepths"][k]), W, H, inp["fK"], inp["fR"][k], inp["ft"][k], 0.3)
    np.testing.assert_array_equal(f.download()["hist"], out["fusion_hist"])
    f.solve(5, 0.1, 0.5, 1.0)
    assert rel_err(f.download()["u"], out["fusion_u_f64"]).max() <= 1e-4
    t = stages.tvl1_denoise(out["sweep_ref_f32"], inp["ref"], 0.5, 6.0, niters=20)
    assert np.abs(t - out["tvl1_f32"]).max() <= 1e-5
