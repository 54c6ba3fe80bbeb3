"""Writes tests/golden/golden_small.npz: small seeded inputs and the oracle's
outputs for them (FP32 and FP64).  Run from the repository root:

    python tests/golden/make_golden.py

The fixture pins the oracle against regressions (tests/test_golden.py re-runs
it on the stored inputs) and gives GPU tests outputs to match without the
oracle.  Inputs come from the seeded scene generator; nothing here comes from
the reference (which has no test data for this path)."""
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path[:0] = [os.path.join(ROOT, "plane-sweep-cuda_amd", "python"), os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from psw_amd import scene  # noqa: E402

PARAMS = dict(znear=0.5, zfar=6.0, nplanes=12, win=5, stdthresh=1e-4, nccthresh=0.5, agg_planes=16, agg_win=7,
              tgv_iters=40, tgv_warps=2, tgv_lambda=1.0, fus_thr=0.3, fus_tau=0.1, fus_lambda=0.5, fus_sigma=1.0,
              fus_iters=5, tvl1_iters=20)


def compute(inp):
    p = PARAMS
    K1 = scene.one_based(inp["K"])
    pose = (inp["Rref"], inp["tref"], inp["Rsrc"], inp["tsrc"])
    out = {}
    planes = oracle.reference_planes(p["znear"], p["zfar"], p["nplanes"])
    out["planes"] = planes
    out["sweep_ref_f32"] = oracle.sweep_ref(inp["ref"], inp["src"], K1, *pose, planes, p["win"], p["stdthresh"],
                                            p["nccthresh"], p["zfar"])
    out["sweep_ref_f64"] = oracle.sweep_ref(inp["ref"], inp["src"], K1, *pose, planes, p["win"], p["stdthresh"],
                                            p["nccthresh"], p["zfar"], prec="f64")
    ap = inp["agg_planes"]
    a, c, s = oracle.sweep_agg(inp["ref"], inp["src"], inp["K"], *pose, ap, p["agg_win"], p["stdthresh"])
    out.update(agg_arg_f32=a, agg_cmin_f32=c, agg_csec_f32=s)
    a, c, s = oracle.sweep_agg(inp["ref"], inp["src"], inp["K"], *pose, ap, p["agg_win"], p["stdthresh"], prec="f64")
    out.update(agg_arg_f64=a, agg_cmin_f64=c, agg_csec_f64=s)
    for prec in ("f32", "f64"):
        out["tgv_" + prec] = oracle.tgv(inp["ref"], inp["src"], K1, *pose, p["tgv_iters"], p["tgv_warps"],
                                        lambda_=p["tgv_lambda"], prec=prec)
    hist = np.zeros(tuple(inp["grid"]) + (8,), np.uint16)
    for k in range(inp["fdepths"].shape[0]):
        oracle.fusion_hist(hist, inp["fvol"], inp["fdepths"][k], inp["fK"], inp["fR"][k], inp["ft"][k], p["fus_thr"])
    out["fusion_hist"] = hist
    for prec, dt in (("f32", np.float32), ("f64", np.float64)):
        u, v, pp = np.zeros(hist.shape[:3], dt), np.zeros(hist.shape[:3], dt), np.zeros((3,) + hist.shape[:3], dt)
        for _ in range(p["fus_iters"]):
            oracle.fusion_u(u, v, pp, hist, p["fus_tau"], p["fus_lambda"], prec=prec)
            oracle.fusion_p(pp, v, p["fus_sigma"], prec=prec)
        out["fusion_u_" + prec] = u
    out["tvl1_f32"] = oracle.tvl1(out["sweep_ref_f32"], inp["ref"], p["znear"], p["zfar"], niters=p["tvl1_iters"])
    return out


def make_inputs():
    v = scene.make_views(seed=21, width=48, height=36, nsrc=2)
    o = scene.orbit_depthmaps(seed=22, n=3, width=40, height=30)
    return {"ref": v["ref"], "src": v["src"], "K": v["K"], "Rref": v["Rref"], "tref": v["tref"], "Rsrc": v["Rsrc"],
            "tsrc": v["tsrc"], "agg_planes": (1.0 / np.linspace(1 / 0.5, 1 / 6.0, PARAMS["agg_planes"])).astype(
                np.float32), "fdepths": o["depths"], "fK": o["K"], "fR": o["R"], "ft": o["t"],
            "fvol": np.array([-3.0, -1.6, -4.0, 3.0, 1.6, 6.0], np.float32), "grid": np.array([14, 10, 12], np.int32)}


if __name__ == "__main__":
    inp = make_inputs()
    out = compute(inp)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_small.npz")
    np.savez_compressed(path, **{"in_" + k: v for k, v in inp.items()}, **{"out_" + k: v for k, v in out.items()})
    print("wrote", path, os.path.getsize(path), "bytes")
