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
    """FP32 oracle (the CPU restatement, OpenMP) on one complete depthmap of
    the same workload: the full aggregated sweep and as many TGV2 iterations
    as the GPU needed to reach the gap."""
    import oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    oracle.sweep_agg(v["ref"], v["src"], K0, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, WIN, STD_T)
    t_sweep = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], iters_per_map, 1, **TGV)
    t_tgv = time.perf_counter() - t0
    t_map = t_sweep + t_tgv
    return {"value": 1.0 / t_map, "unit": "depthmaps/s", "cores": cores, "kind": "port",
            "sample": f"one full depthmap on the FP32 oracle: aggregated sweep over all {NPLANES} planes x {NSRC} "
                      f"sources and {iters_per_map} TGV2 iterations at 640x480 (single run)",
            "sweep_s": t_sweep, "tgv_s": t_tgv}


# ---- fusion workload: 256^3 grid, 50 depthmaps, 300 primal-dual iterations ----

FUSION_GRID = 256
FUSION_MAPS = 50
FUSION_ITERS = 300
FUSION_VOLUME = np.array([-3.0, -1.6, -4.0, 3.0, 1.6, 6.0], np.float32)
FUSION_BYTES_PER_VOXEL_ITER = 56   # fused PD: read p 12, v 4, u 4, hist 16; write p 12, u 4, v 4


def fusion_workload(seed):
    from psw_amd import scene
    o = scene.orbit_depthmaps(seed=seed, n=FUSION_MAPS, width=W, height=H, noise_sigma=0.01)
    edge = float(np.mean((FUSION_VOLUME[3:] - FUSION_VOLUME[:3]) / FUSION_GRID))
    return o, 5.0 * edge   # truncation band of +-5 voxel edges


def run_fusion(o, thr, stream, ev):
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    g = FUSION_GRID
    f = stages.Fusion(g, g, g, FUSION_VOLUME)
    d_maps = DeviceArray.from_host(o["depths"])
    Ks = np.repeat(o["K"][None], FUSION_MAPS, 0)
    res = {}
    for rep in range(2):          # first repetition warms up
        f.clear(stream.handle)
        ev[0].record(stream)
        f.update_histogram_batch(d_maps, FUSION_MAPS, W, H, Ks, o["R"], o["t"], thr, stream.handle)
        ev[1].record(stream)
        f.solve(FUSION_ITERS, 0.1, 0.5, 1.0, stream.handle)
        ev[2].record(stream)
        ev[2].synchronize()
        res = {"hist_ms": ev[0].elapsed_ms(ev[1]), "pd_ms": ev[1].elapsed_ms(ev[2])}
    vox = g ** 3
    res["surface_voxels"] = f.surface_count(stream.handle)
    res["voxel_iters_per_s"] = vox * FUSION_ITERS / (res["pd_ms"] * 1e-3)
    res["pd_gbs_algorithmic"] = FUSION_BYTES_PER_VOXEL_ITER * vox * FUSION_ITERS / (res["pd_ms"] * 1e-3) / 1e9
    res["hist_voxel_maps_per_s"] = vox * FUSION_MAPS / (res["hist_ms"] * 1e-3)
    res["config"] = (f"{g}^3 voxels over the room box, {FUSION_MAPS} depthmaps 640x480 (ray-cast depth + 1 cm "
                     f"noise, orbiting cameras), 8 uint16 bins, band +-{thr:.3f} m (5 voxels), "
                     f"{FUSION_ITERS} PD iterations, tau 0.1 lambda 0.5 sigma 1")
    return res


def probe(names, seed):
    """Small fixed workloads for rocprofv3 runs (kernel stats and counters)."""
    import psw_amd
    from psw_amd import scene, stages
    from psw_amd.device import DeviceArray, Event, Stream
    assert psw_amd.device_count() > 0
    st = Stream()
    s = st.handle
    for name in names:
        log(f"probe {name}")
        if name == "copy":
            n = 64 * 1024 * 1024          # 256 MiB of floats
            a, b = DeviceArray((n,)), DeviceArray((n,))
            a.fill(1.0, s)
            for _ in range(5):
                psw_amd._lib.call("psw_copy_f32", b.ptr, a.ptr, n, s)
        elif name in ("tgv", "sweep", "sweep_exact"):
            v = scene.make_views(seed=seed, width=W, height=H, nsrc=NSRC)
            planes = stages.inverse_depth_planes(ZNEAR, ZFAR, NPLANES)
            sw = stages.SweepAggregate(v["ref"], v["src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes,
                                       WIN, STD_T)
            if name == "tgv":
                t = stages.TgvSolver(W, H, NSRC)
                t.set_images(sw.d_ref, sw.d_src, 0.0, 1.0, s)
                Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
                t.set_u(value=1.0, stream=s)
                t.linearise(scene.one_based(v["K"]), Rrel, trel, s)
                t.iterate(100, stages.tgv_params(**TGV), s)
            else:
                sw.build_homographies(s)
                for _ in range(3):
                    sw.run(fast=(name == "sweep"), stream=s)
        elif name == "fusion":
            o, thr = fusion_workload(seed)
            ev = [Event() for _ in range(3)]
            log(json.dumps(run_fusion(o, thr, st, ev)))
        st.synchronize()
    log("probe done")


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
    ap.add_argument("--no-fusion", action="store_true", help="skip the 256^3 fusion measurement")
    ap.add_argument("--probe", default="", help="comma list of tgv,sweep,sweep_exact,copy,fusion: run only these "
                                                "small workloads (for rocprofv3) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described in DESIGN.md, section 'Sharding')", file=sys.stderr)
        sys.exit(2)

    if args.probe:
        probe(args.probe.split(","), args.seed)
        return

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
    if not args.no_fusion:
        t0 = time.perf_counter()
        o, thr = fusion_workload(args.seed)
        log(f"fusion depthmaps rendered in {time.perf_counter() - t0:.1f} s")
        rec["fusion"] = run_fusion(o, thr, st, ev)
        log(f"fusion: histogram {rec['fusion']['hist_ms']:.2f} ms, {FUSION_ITERS} PD iterations "
            f"{rec['fusion']['pd_ms']:.2f} ms ({rec['fusion']['pd_gbs_algorithmic']:.0f} GB/s algorithmic)")
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        rec["cpu_baseline"] = cpu_baseline(v, planes, K0, K1, int(round(np.mean(iters_used))))
        log(f"cpu baseline sample took {time.perf_counter() - t0:.1f} s")
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
