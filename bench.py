#!/usr/bin/env python3
"""Throughput of the plain multiview-stereo pipeline on one MI355X.

Workload (one step = one depthmap): a seeded synthetic room, 640x480 float32
reference view and 8 source views (0..255 intensities), pinhole K fx=fy=480,
cx=319.5, cy=239.5; poses within 5 cm / 2 degrees of the reference.
  sweep : 256 fronto-parallel planes uniform in inverse depth over 0.3-8 m,
          9x9 NCC window, aggregated cost -(1/N) sum_s NCC, fused kernel with
          argmin / min / second-best (no cost volume in HBM);
  TGV2  : cold start u = 1, one linearisation, lambda=1, alpha0=2, alpha1=1.5,
          tau=0.02, sigma=1, iterated until the primal energy is within 1e-4
          (relative) of E*, the converged energy of the same problem
          (computed before the timed region), capped at 1000 iterations.
Inputs are resident in HBM when the timed region starts.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "plane-sweep-cuda_amd", "python"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

W, H, NSRC, NPLANES, WIN = 640, 480, 8, 256, 9
ZNEAR, ZFAR = 0.3, 8.0
STD_T = 1e-4
TGV = dict(lambda_=1.0, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, beta=0.0, gamma=1.0)
GAP_TOL = 1e-4
MAX_ITERS = 1000
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
SWEEP_FLOP_PER_EVAL = 50          # declared FLOP per (pixel, plane, source) NCC evaluation
TGV_FLOATS_PER_PX = 23 + 4 * NSRC  # algorithmic floats moved per pixel per fused iteration


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(v, planes, K0, K1, iters_per_map):
    """FP32 oracle (the CPU restatement, OpenMP) on a bounded sample of the
    same workload, scaled to one depthmap."""
    import oracle
    from psw_amd import scene  # noqa: F401
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nsp = 2
    t0 = time.perf_counter()
    oracle.sweep_agg(v["ref"], v["src"], K0, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes[:nsp], WIN, STD_T)
    t_sweep = (time.perf_counter() - t0) / nsp * NPLANES
    t0 = time.perf_counter()
    oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], 0, 1, **TGV)
    t_setup = time.perf_counter() - t0
    nit = 10
    t0 = time.perf_counter()
    oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], nit, 1, **TGV)
    t_it = (time.perf_counter() - t0 - t_setup) / nit
    t_map = t_sweep + t_setup + t_it * iters_per_map
    return {"value": 1.0 / t_map, "unit": "depthmaps/s", "cores": cores, "kind": "port",
            "sample": f"FP32 oracle: aggregated sweep over {nsp} of {NPLANES} planes x {NSRC} sources and "
                      f"{nit} TGV2 iterations at 640x480, scaled to {NPLANES} planes + {iters_per_map} iterations",
            "sweep_s_per_map": t_sweep, "tgv_s_per_iter": t_it, "tgv_setup_s": t_setup}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--exact-sweep", action="store_true", help="reference-order (bit-exact) sweep arithmetic")
    ap.add_argument("--max-iters", type=int, default=MAX_ITERS)
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--estar-iters", type=int, default=6000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described in DESIGN.md, section 'Sharding')", file=sys.stderr)
        sys.exit(2)

    import psw_amd
    from psw_amd import scene, stages
    from psw_amd.device import Event, Stream

    assert psw_amd.device_count() > 0, "no GPU"
    log("device:", psw_amd.device_name())
    t0 = time.perf_counter()
    v = scene.make_views(seed=args.seed, width=W, height=H, nsrc=NSRC)
    log(f"scene generated in {time.perf_counter() - t0:.1f} s")
    planes = stages.inverse_depth_planes(ZNEAR, ZFAR, NPLANES)
    K0 = v["K"]
    K1 = scene.one_based(K0)
    st = Stream()
    s = st.handle
    sweep = stages.SweepAggregate(v["ref"], v["src"], K0, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, WIN,
                                  STD_T, coord_offset=0)
    tgv = stages.TgvSolver(W, H, NSRC)
    tgv.set_images(sweep.d_ref, sweep.d_src, TGV["beta"], TGV["gamma"], s)
    Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    prm = stages.tgv_params(**TGV)
    fast = not args.exact_sweep

    # E*: converged energy of the cold-start problem (not timed)
    tgv.set_u(value=1.0, stream=s)
    tgv.linearise(K1, Rrel, trel, s)
    tgv.iterate(args.estar_iters - 1000, prm, s)
    e_prev = tgv.energy(prm, s)
    tgv.iterate(1000, prm, s)
    e_star = tgv.energy(prm, s)
    log(f"E* = {e_star:.9e} after {args.estar_iters} iterations (change over the last 1000: "
        f"{abs(e_prev - e_star) / e_star:.2e})")

    ev = [Event() for _ in range(3)]
    sweep_ms, tgv_ms, iters_used, gaps = [], [], [], []

    def step(timed):
        ev[0].record(st)
        sweep.build_homographies(s)
        sweep.run(fast=fast, stream=s)
        ev[1].record(st)
        tgv.set_u(value=1.0, stream=s)
        tgv.linearise(K1, Rrel, trel, s)
        iters, e = tgv.solve(args.max_iters, args.check_every, GAP_TOL, e_star, prm, s)
        ev[2].record(st)
        ev[2].synchronize()
        if timed:
            sweep_ms.append(ev[0].elapsed_ms(ev[1]))
            tgv_ms.append(ev[1].elapsed_ms(ev[2]))
            iters_used.append(iters)
            gaps.append(abs(e - e_star) / e_star)

    for _ in range(args.warmup):
        step(False)
    st.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    st.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step; sweep {np.mean(sweep_ms):.3f} ms, "
        f"TGV2 {np.mean(tgv_ms):.3f} ms ({np.mean(iters_used):.0f} iterations)")

    # Per-launch duration of the fused TGV2 iteration kernel (same stream and
    # buffers): events around back-to-back launches.
    tgv.set_u(value=1.0, stream=s)
    tgv.linearise(K1, Rrel, trel, s)
    nl = 200
    ev[0].record(st)
    tgv.iterate(nl, prm, s)
    ev[1].record(st)
    ev[1].synchronize()
    tgv_launch_ms = ev[0].elapsed_ms(ev[1]) / nl
    tgv_bytes = 4 * TGV_FLOATS_PER_PX * W * H
    achieved = tgv_bytes / (tgv_launch_ms * 1e-3) / 1e9
    # the sweep kernel alone
    ev[0].record(st)
    sweep.run(fast=fast, stream=s)
    ev[1].record(st)
    ev[1].synchronize()
    sweep_kernel_ms = ev[0].elapsed_ms(ev[1])
    sweep_tflops = SWEEP_FLOP_PER_EVAL * W * H * NPLANES * NSRC / (sweep_kernel_ms * 1e-3) / 1e12

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "tgv_iter_kernel_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    value = 1e3 / ms_per_step
    rec = {
        "metric": "640x480 depthmaps/s at equal TGV2 gap: learned warm-start vs plain TGV2",
        "value": value,
        "unit": "depthmaps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded room scene, generated in-process)",
        "config": {"workload": "plain pipeline (cold-start TGV2, no learned warm start): 640x480 reference + 8 "
                               "sources, 256 inverse-depth planes 0.3-8 m, 9x9 NCC, fused sweep+argmin, TGV2 to "
                               "relative energy gap 1e-4 (cap 1000 iterations)",
                   "sweep_arith": "fast" if fast else "reference-order",
                   "tgv_gap_tol": GAP_TOL, "check_every": args.check_every, "seed": args.seed},
        "roofline": {"kernel": "tgv_iter_kernel", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": tgv_bytes, "launch_ms": tgv_launch_ms},
        "stages": {"sweep_ms": float(np.mean(sweep_ms)), "tgv_ms": float(np.mean(tgv_ms)),
                   "tgv_iterations": float(np.mean(iters_used)), "tgv_rel_gap_at_stop": float(np.max(gaps)),
                   "e_star": e_star, "sweep_kernel_ms": sweep_kernel_ms, "sweep_tflops_declared": sweep_tflops,
                   "sweep_frac_fp32_peak": sweep_tflops / FP32_PEAK_TFLOPS},
    }
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        rec["cpu_baseline"] = cpu_baseline(v, planes, K0, K1, int(round(np.mean(iters_used))))
        log(f"cpu baseline sample took {time.perf_counter() - t0:.1f} s")
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
