#!/usr/bin/env python3
"""Throughput of the multiview-stereo pipeline on one MI355X, learned
warm-start TGV2 against plain TGV2 at equal energy gap.

Workload (one step = one refined 640x480 depthmap): a seeded synthetic room
rendered on the GPU, 640x480 float32 reference view and 8 source views
(0..255 intensities), pinhole K fx=fy=480, cx=319.5, cy=239.5; poses within
5 cm / 2 degrees of the reference.
  sweep : 256 fronto-parallel planes uniform in inverse depth over 0.3-8 m,
          9x9 NCC window, aggregated cost -(1/N) sum_s NCC, fused kernel with
          argmin / min / second-best (no cost volume in HBM);
  TGV2  : linearised at the sweep's argmin depth, lambda=1, alpha0=2,
          alpha1=1.5, tau=0.02, sigma=1, iterated until the primal energy is
          within 1e-4 (relative) of E*, the energy of the same problem after
          40000 iterations from the plain start (computed before the timed
          region), checked every 10 iterations, capped at 50000.  The energy
          leaves out the ~0.2 % of pixels where the reference's fixed steps
          make the data step unstable (tau sigma lambda^2 sum Iu^2 > 1 at the
          pixel or a neighbour): the iteration oscillates there forever;
  starts: learned = the BF16 U-Net's (u, u1, p, q) from the sweep results
          (the headline value, network time included); plain = the
          reference's warp start u = u0 (the sweep depth), u1 = p = q = r = 0;
          cold = the reference's constant u = 1.
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
MAX_ITERS = 50000
ESTAR_ITERS = 40000
QMAX = 1.0   # energy mask: pixels with tau sigma lambda^2 sum Iu^2 > QMAX (or at a 4-neighbour) left out
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
SWEEP_FLOP_PER_EVAL = 50          # declared FLOP per (pixel, plane, source) NCC evaluation
TGV_FLOATS_PER_PX = 23 + 4 * NSRC  # algorithmic floats moved per pixel per fused iteration


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(v, planes, K0, K1, ds, plain_iters, sample_iters=300):
    """FP32 oracle (the CPU restatement, OpenMP) on the same inputs: the full
    aggregated sweep of one depthmap, then `sample_iters` TGV2 iterations
    linearised at the GPU's sweep depth; the depthmap rate is the sweep time
    plus the sampled per-iteration time times the plain pipeline's iteration
    count (the CPU path has no learned start)."""
    import oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    oracle.sweep_agg(v["ref"], v["src"], K0, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes, WIN, STD_T)
    t_sweep = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], sample_iters, 1, u_init=ds, **TGV)
    t_it = (time.perf_counter() - t0) / sample_iters
    t_map = t_sweep + t_it * plain_iters
    return {"value": 1.0 / t_map, "unit": "depthmaps/s", "cores": cores, "kind": "port",
            "sample": f"FP32 oracle: one full aggregated sweep ({NPLANES} planes x {NSRC} sources, 640x480) and "
                      f"{sample_iters} TGV2 iterations, scaled to the plain pipeline's {plain_iters} iterations",
            "sweep_s": t_sweep, "tgv_iter_ms": t_it * 1e3}


# ---- fusion workload: 256^3 grid, 50 depthmaps, 300 primal-dual iterations ----

FUSION_GRID = 256
FUSION_MAPS = 50
FUSION_ITERS = 300
FUSION_VOLUME = np.array([-3.0, -1.6, -4.0, 3.0, 1.6, 6.0], np.float32)
# fused PD: read p 12, v 4, u 4, counts 8 (the byte-packed copy the solver
# reads when every count fits in a byte, as with 50 depthmaps); write p 12,
# u 4, v 4
FUSION_BYTES_PER_VOXEL_ITER = 48


def fusion_workload(seed):
    from psw_amd import scene
    o = scene.orbit_depthmaps_gpu(seed=seed, n=FUSION_MAPS, width=W, height=H, noise_sigma=0.01)
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
    res["pd_frac_hbm_peak"] = res["pd_gbs_algorithmic"] / HBM_PEAK_GBS
    res["pd_bytes_per_voxel_iter"] = FUSION_BYTES_PER_VOXEL_ITER
    res["hist_voxel_maps_per_s"] = vox * FUSION_MAPS / (res["hist_ms"] * 1e-3)
    res["config"] = (f"{g}^3 voxels over the room box, {FUSION_MAPS} depthmaps 640x480 (ray-cast depth + 1 cm "
                     f"noise, orbiting cameras), 8 uint16 bins, band +-{thr:.3f} m (5 voxels), "
                     f"{FUSION_ITERS} PD iterations, tau 0.1 lambda 0.5 sigma 1")
    return res


SPARSE_ITERS = 500
SPARSE_BYTES_PER_PIXEL_ITER = 80   # read u 12, p 8, q 12, T 12, sparse 4; write u 12, p 8, q 12


def run_sparse(pl, stream, ev, seed):
    """Sparse-depth TGV2 densification at 640x480: the sweep depthmap as the
    dense start, the scene's true depth at 5 % of the pixels as the sparse
    input.  Per-iteration time from a 500- and a 0-iteration call."""
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    s = stream.handle
    d_dense = pl.sweep_and_linearise()
    pl.st.synchronize()
    rng = np.random.default_rng(seed + 77)
    gt = pl.v["d_depth"].to_host()
    d_sparse = DeviceArray.from_host(np.where(rng.random(gt.shape) < 0.05, gt, 0).astype(np.float32))
    ws = DeviceArray((stages.tgv_sparse_workspace_bytes(W, H) // 4,))
    kw = dict(alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, theta=1.0, beta=0.0, gamma=1.0)

    def timed(k):
        ev[0].record(stream)
        stages.tgv_sparse(pl.v["d_ref"], d_dense, d_sparse, ZFAR, k, workspace=ws, stream=s, **kw)
        ev[1].record(stream)
        ev[1].synchronize()
        return ev[0].elapsed_ms(ev[1])

    timed(20)
    t_full, t_zero = timed(SPARSE_ITERS), timed(0)
    it_ms = (t_full - t_zero) / SPARSE_ITERS
    gbs = SPARSE_BYTES_PER_PIXEL_ITER * W * H / (it_ms * 1e-3) / 1e9
    multi = os.environ.get("PSW_TGV_SPARSE_DOUBLE", "1") != "0"   # several iterations per launch (default)
    m = int(os.environ.get("PSW_TGV_SPARSE_M", "4")) if multi else 1
    kernel = {1: "tgv_sparse_iter_kernel", 2: "tgv_sparse_iter2_kernel"}.get(m, f"tgv_sparse_iterm_kernel<{m}>")
    return {"kernel": kernel, "iterations_per_launch": m, "iter_us": it_ms * 1e3, "call_ms": t_full,
            "setup_output_ms": t_zero, "gbs_algorithmic": gbs, "frac_hbm_peak": gbs / HBM_PEAK_GBS,
            "bytes_per_pixel_iter": SPARSE_BYTES_PER_PIXEL_ITER,
            "config": f"{W}x{H}, sweep depth as the dense start, true depth at 5% of pixels, {SPARSE_ITERS} "
                      "iterations, alpha0 2 alpha1 1.5 tau 0.02 sigma 1"}


TVL1_ITERS = 1000
# per pixel and iteration, with nothing kept on chip between iterations: read
# depth, Px, Py, R, the scaled input and T11..T22 (36 B), write depth, Px, Py, R (16 B)
TVL1_BYTES_PER_PIXEL_ITER = 52


def run_tvl1(pl, stream, ev):
    """TVL1 denoising (PlaneSweep::CudaDenoise) of the 640x480 sweep depthmap
    with the reference's default parameters.  Per-iteration time from a
    1000- and a 0-iteration call, for the fused launches and for the
    reference's two launches per iteration (PSW_TVL1_FUSED=0)."""
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    d_depth = pl.sweep_and_linearise()
    pl.st.synchronize()
    d_out, d_work = DeviceArray((H, W)), DeviceArray((12, H, W))

    def timed(k):
        ev[0].record(stream)
        stages.tvl1_denoise_device(d_depth, pl.v["d_ref"], W, H, ZNEAR, ZFAR, d_out, d_work, k,
                                   stream=stream.handle)
        ev[1].record(stream)
        ev[1].synchronize()
        return ev[0].elapsed_ms(ev[1])

    res = {}
    prev = os.environ.get("PSW_TVL1_FUSED")
    try:
        for name, flag in (("launch_sequence", "0"), ("fused", "1")):
            os.environ["PSW_TVL1_FUSED"] = flag
            timed(20)
            t_full, t_zero = timed(TVL1_ITERS), timed(0)
            res[name] = (t_full - t_zero) / TVL1_ITERS
    finally:
        if prev is None:
            os.environ.pop("PSW_TVL1_FUSED", None)
        else:
            os.environ["PSW_TVL1_FUSED"] = prev
    gbs = TVL1_BYTES_PER_PIXEL_ITER * W * H / (res["fused"] * 1e-3) / 1e9
    m = int(os.environ.get("PSW_TVL1_M", "4"))                # iterations per launch (default 4, 56x8 tiles)
    return {"kernel": f"tvl1_iter_kernel<{m}>", "iterations_per_launch": m, "iter_us": res["fused"] * 1e3,
            "launch_sequence_iter_us": res["launch_sequence"] * 1e3,
            "speedup_vs_launch_sequence": res["launch_sequence"] / res["fused"], "gbs_algorithmic": gbs,
            "frac_hbm_peak": gbs / HBM_PEAK_GBS, "bytes_per_pixel_iter": TVL1_BYTES_PER_PIXEL_ITER,
            "config": f"{W}x{H} sweep depthmap, {TVL1_ITERS} iterations, lambda 0.3 tau 0.02 sigma 6 theta 1 "
                      "beta 0 gamma 1 (the reference's defaults)"}


class Pipeline:
    """Device state of the benchmarked pipeline for one seeded view."""

    def __init__(self, seed, weights):
        import psw_amd
        from psw_amd import scene, stages, warmstart
        from psw_amd.device import Event, Stream
        assert psw_amd.device_count() > 0, "no GPU"
        self.st = Stream()
        s = self.st.handle
        t0 = time.perf_counter()
        self.v = scene.make_views_gpu(seed, W, H, NSRC, stream=s)
        self.st.synchronize()
        log(f"scene rendered in {time.perf_counter() - t0:.2f} s")
        self.planes = stages.inverse_depth_planes(ZNEAR, ZFAR, NPLANES)
        v = self.v
        self.K0 = v["K"]
        self.K1 = scene.one_based(self.K0)
        self.sweep = stages.SweepAggregate(v["d_ref"], v["d_src"], self.K0, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"],
                                           self.planes, WIN, STD_T, coord_offset=0)
        self.tgv = stages.TgvSolver(W, H, NSRC)
        self.tgv.set_images(self.sweep.d_ref, self.sweep.d_src, TGV["beta"], TGV["gamma"], s)
        self.Rrel, self.trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
        self.prm = stages.tgv_params(**TGV)
        self.ws = warmstart.WarmStarter(W, H, weights) if weights else None
        self.ev = [Event() for _ in range(4)]

    def sweep_and_linearise(self, fast=True):
        s = self.st.handle
        self.sweep.build_homographies(s)
        self.sweep.run(fast=fast, stream=s)
        d_s = self.sweep.depth(ZFAR, s)
        return d_s

    def run(self, mode, e_star, max_iters, check_every, fast=True):
        """One depthmap; mode: learned | plain | cold.  Returns stage times
        (ms), iterations and the energy at the stop."""
        s, ev, st = self.st.handle, self.ev, self.st
        ev[0].record(st)
        d_s = self.sweep_and_linearise(fast)
        ev[1].record(st)
        if mode == "learned":
            self.ws.predict(self.sweep, ZFAR, s)
        ev[2].record(st)
        self.tgv.set_u(d_s, stream=s)
        self.tgv.linearise(self.K1, self.Rrel, self.trel, s)
        self.tgv.set_energy_mask(QMAX, self.prm, s, count=False)
        if mode == "learned":
            self.ws.start(self.tgv, s)
        elif mode == "cold":
            self.tgv.set_primal(None, 1.0, None, s)
        iters, e = self.tgv.solve(max_iters, check_every, GAP_TOL, e_star, self.prm, s)
        ev[3].record(st)
        ev[3].synchronize()
        return {"sweep_ms": ev[0].elapsed_ms(ev[1]), "net_ms": ev[1].elapsed_ms(ev[2]),
                "tgv_ms": ev[2].elapsed_ms(ev[3]), "iters": iters, "energy": e}


def probe(names, seed):
    """Small fixed workloads for rocprofv3 runs (kernel stats and counters)."""
    import psw_amd
    from psw_amd import warmstart
    from psw_amd.device import DeviceArray, Event, Stream
    assert psw_amd.device_count() > 0
    st = Stream()
    s = st.handle
    pl = None
    for name in names:
        log(f"probe {name}")
        if name == "copy":
            n = 64 * 1024 * 1024          # 256 MiB of floats
            a, b = DeviceArray((n,)), DeviceArray((n,))
            a.fill(1.0, s)
            for _ in range(5):
                psw_amd._lib.call("psw_copy_f32", b.ptr, a.ptr, n, s)
        elif name in ("tgv", "sweep", "sweep_exact", "unet"):
            if pl is None:
                pl = Pipeline(seed, warmstart.DEFAULT_WEIGHTS if name == "unet" else None)
            if name == "tgv":
                d_s = pl.sweep_and_linearise()
                pl.tgv.set_u(d_s, stream=pl.st.handle)
                pl.tgv.linearise(pl.K1, pl.Rrel, pl.trel, pl.st.handle)
                pl.tgv.iterate(100, pl.prm, pl.st.handle)
            elif name == "unet":
                if pl.ws is None:
                    pl.ws = warmstart.WarmStarter(W, H, warmstart.DEFAULT_WEIGHTS)
                pl.sweep_and_linearise()
                for _ in range(5):
                    pl.ws.predict(pl.sweep, ZFAR, pl.st.handle)
            else:
                for _ in range(3):
                    pl.sweep_and_linearise(fast=(name == "sweep"))
        elif name == "sparse":
            if pl is None:
                pl = Pipeline(seed, None)
            log(json.dumps(run_sparse(pl, st, [Event() for _ in range(2)], seed)))
        elif name == "tvl1":
            if pl is None:
                pl = Pipeline(seed, None)
            log(json.dumps(run_tvl1(pl, st, [Event() for _ in range(2)])))
        elif name == "fusion":
            o, thr = fusion_workload(seed)
            ev = [Event() for _ in range(3)]
            log(json.dumps(run_fusion(o, thr, st, ev)))
        st.synchronize()
        if pl is not None:
            pl.st.synchronize()
    log("probe done")


def run_batch(args):
    """The uneven batch: 32 seeded frames, resolution from
    {320x240, 640x480, 1242x375}, 4-8 sources, 64-256 planes per frame; sweep
    and TGV2 (plain start, linearised at the sweep depth, to the energy gap)
    for the whole batch in batched launches, against the same frames one at
    a time through the single-frame entry points."""
    import psw_amd
    from psw_amd import batch, stages
    from psw_amd.device import Stream
    from psw_amd.device import synchronize as device_synchronize
    log("device:", psw_amd.device_name())
    st = Stream()
    s = st.handle
    specs = batch.uneven_workload(args.seed, args.batch_frames)
    t0 = time.perf_counter()
    frames = [batch.Frame(sp, ZNEAR, ZFAR, WIN, STD_T) for sp in specs]
    device_synchronize()
    npx = sum(sp["W"] * sp["H"] for sp in specs)
    log(f"{len(frames)} frames ({npx / 1e6:.2f} Mpx) rendered in {time.perf_counter() - t0:.2f} s")
    prm = stages.tgv_params(**TGV)
    best, timings = batch.autotune(frames, prm, stream=st, log=log)
    log(f"autotuned: {best}")
    solvers = [f.tgv for f in frames]
    sweeps = [f.sweep for f in frames]

    def sweep_all():
        for sw in sweeps:
            sw.build_homographies(s)
        batch.sweep_batch(sweeps, True, best["sweep_groups"], best["sweep_workgroups"], s)

    def linearise_all():
        for f in frames:
            f.linearise(prm, QMAX, s)

    # E* per frame (not timed): batched iterations from the plain start
    sweep_all()
    linearise_all()
    batch.tgv_iterate_batch(solvers, args.batch_estar_iters, prm, best["tgv_variant"], s)
    e_star = [t.energy(prm, s) for t in solvers]
    st.synchronize()

    def batched():
        sweep_all()
        linearise_all()
        return batch.tgv_solve_batch(solvers, prm, e_star, args.max_iters, args.check_every, GAP_TOL,
                                     best["tgv_variant"], best["use_graph"], s)

    batched()                                   # warm-up
    st.synchronize()
    t0 = time.perf_counter()
    iters, energies = batched()
    st.synchronize()
    t_batch = time.perf_counter() - t0
    gaps = [abs(e - es) / es for e, es in zip(energies, e_star)]
    # the same frames one at a time through the single-frame entry points
    t0 = time.perf_counter()
    seq_ms = []
    for f, es in zip(frames, e_star):
        t1 = time.perf_counter()
        f.sweep.build_homographies(s)
        f.sweep.run(fast=True, stream=s)
        f.linearise(prm, QMAX, s)
        f.tgv.solve(args.max_iters, args.check_every, GAP_TOL, es, prm, s)
        st.synchronize()
        seq_ms.append((time.perf_counter() - t1) * 1e3)
    t_seq = time.perf_counter() - t0
    rec = {
        "metric": "uneven-batch depthmaps/s (32 frames, sweep + TGV2 to the energy gap)",
        "value": len(frames) / t_batch, "unit": "depthmaps/s", "n_gpus": 1, "higher_is_better": True,
        "batch_ms": t_batch * 1e3, "sequential_ms": t_seq * 1e3, "speedup_vs_sequential": t_seq / t_batch,
        "dtype": "f32", "data": "synthetic (seeded room scenes rendered on the GPU)",
        "config": {"workload": "32 frames, resolution in {320x240, 640x480, 1242x375}, 4-8 sources, 64-256 "
                               "inverse-depth planes 0.3-8 m, 9x9 NCC, TGV2 plain start to energy gap 1e-4 of E* "
                               f"({args.batch_estar_iters}-iteration run), energy mask qmax {QMAX}",
                   "frames": specs, "megapixels": npx / 1e6, "seed": args.seed},
        "autotune": {"chosen": best,
                     "sweep_ms": {f"groups{g}_wg{w}": v for (g, w), v in timings["sweep"].items()},
                     "tgv_us_per_iteration": {f"variant{v}": ms * 1e3 for v, ms in timings["tgv"].items()},
                     "solve_us_per_iteration": {("graph" if g else "direct"): ms * 1e3
                                                for g, ms in timings["graph"].items()}},
        "iterations": [int(i) for i in iters], "max_rel_gap_at_stop": float(max(gaps)),
        "sequential_per_frame_ms": seq_ms,
    }
    log(f"batch: {t_batch * 1e3:.1f} ms for {len(frames)} frames ({len(frames) / t_batch:.2f} depthmaps/s); "
        f"one at a time: {t_seq * 1e3:.1f} ms")
    print(json.dumps(rec), flush=True)


def main():
    from psw_amd import warmstart
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed learned warm-start depthmaps")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=2, help="timed plain-start depthmaps")
    ap.add_argument("--exact-sweep", action="store_true", help="reference-order (bit-exact) sweep arithmetic")
    ap.add_argument("--max-iters", type=int, default=MAX_ITERS)
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--estar-iters", type=int, default=ESTAR_ITERS)
    ap.add_argument("--weights", default=warmstart.DEFAULT_WEIGHTS)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cold", action="store_true", help="skip the constant-start (u = 1) run")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-fusion", action="store_true", help="skip the 256^3 fusion measurement")
    ap.add_argument("--batch", action="store_true", help="run the uneven 32-frame batch instead (its own JSON line)")
    ap.add_argument("--batch-frames", type=int, default=32)
    ap.add_argument("--batch-estar-iters", type=int, default=20000)
    ap.add_argument("--probe", default="", help="comma list of tgv,sweep,sweep_exact,unet,copy,fusion,sparse,tvl1: run only "
                                                "these small workloads (for rocprofv3) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (a depthmap's sweep, TGV2 and "
              "fusion each run on one GPU; frames would be spread over GPUs by the caller)", file=sys.stderr)
        sys.exit(2)
    if args.probe:
        probe(args.probe.split(","), args.seed)
        return
    if args.batch:
        run_batch(args)
        return
    if not os.path.exists(args.weights):
        sys.exit(f"bench.py: trained warm-start weights not found at {args.weights} "
                 "(python -m psw_amd.warmstart train --out <file>)")

    import psw_amd
    log("device:", psw_amd.device_name())
    pl = Pipeline(args.seed, args.weights)
    s, st, ev = pl.st.handle, pl.st, pl.ev
    fast = not args.exact_sweep

    # E*: the same problem (linearised at the sweep depth) iterated from the
    # plain start; not timed
    d_s = pl.sweep_and_linearise(fast)
    pl.tgv.set_u(d_s, stream=s)
    pl.tgv.linearise(pl.K1, pl.Rrel, pl.trel, s)
    excluded = pl.tgv.set_energy_mask(QMAX, pl.prm, s)
    pl.tgv.iterate(args.estar_iters - 5000, pl.prm, s)
    e_prev = pl.tgv.energy(pl.prm, s)
    pl.tgv.iterate(5000, pl.prm, s)
    e_star = pl.tgv.energy(pl.prm, s)
    u_star = pl.tgv.get_u(s)
    st.synchronize()
    ds_host = d_s.to_host(s)
    log(f"E* = {e_star:.9e} after {args.estar_iters} iterations (change over the last 5000: "
        f"{(e_prev - e_star) / e_star:.2e}); {excluded} pixels outside the energy mask")

    def timed(mode, k, warm):
        for _ in range(warm):
            pl.run(mode, e_star, args.max_iters, args.check_every, fast)
        st.synchronize()
        recs = []
        t0 = time.perf_counter()
        for _ in range(k):
            recs.append(pl.run(mode, e_star, args.max_iters, args.check_every, fast))
        st.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / k
        u = pl.tgv.get_u(s)
        st.synchronize()
        agg = {key: float(np.mean([r[key] for r in recs])) for key in ("sweep_ms", "net_ms", "tgv_ms", "iters")}
        agg.update(ms_per_depthmap=ms, depthmaps_per_s=1e3 / ms,
                   rel_gap_at_stop=float(max(abs(r["energy"] - e_star) / e_star for r in recs)),
                   reached_gap=bool(all(abs(r["energy"] - e_star) <= GAP_TOL * e_star for r in recs)),
                   rms_rel_diff_vs_estar_state=float(np.sqrt(np.mean(((u - u_star) / u_star) ** 2))),
                   median_rel_err_vs_true_depth=float(np.median(np.abs(u - pl.v["d_depth"].to_host())
                                                                / pl.v["d_depth"].to_host())))
        return agg

    learned = timed("learned", args.steps, args.warmup)
    log(f"learned: {learned['ms_per_depthmap']:.3f} ms/depthmap ({learned['iters']:.0f} TGV2 iterations, "
        f"network {learned['net_ms']:.3f} ms)")
    plain = timed("plain", args.plain_steps, 0)
    log(f"plain: {plain['ms_per_depthmap']:.3f} ms/depthmap ({plain['iters']:.0f} TGV2 iterations)")
    cold = None
    if not args.no_cold:
        cold = timed("cold", 1, 0)
        log(f"cold: {cold['ms_per_depthmap']:.3f} ms/depthmap ({cold['iters']:.0f} iterations, gap "
            f"{cold['rel_gap_at_stop']:.2e})")
    ms_per_step = learned["ms_per_depthmap"]

    # Per-launch duration of the fused TGV2 iteration kernel (same stream and
    # buffers): events around back-to-back launches.  By default a launch
    # does two iterations (tgv_iter2_kernel); PSW_TGV_TRIPLE=1 selects the
    # three-iteration kernel, PSW_TGV_DOUBLE=0 the one-iteration kernel.
    two = os.environ.get("PSW_TGV_DOUBLE", "1") != "0"
    three = two and os.environ.get("PSW_TGV_TRIPLE", "0") == "1"
    tgv_kernel, iters_per_launch = (("tgv_iter3_kernel", 3) if three else ("tgv_iter2_kernel", 2) if two
                                    else ("tgv_iter_kernel", 1))
    pl.tgv.set_u(d_s, stream=s)
    pl.tgv.linearise(pl.K1, pl.Rrel, pl.trel, s)
    nl = 198                     # whole launches of 1, 2 or 3 iterations
    ev[0].record(st)
    pl.tgv.iterate(nl, pl.prm, s)
    ev[1].record(st)
    ev[1].synchronize()
    tgv_launch_ms = ev[0].elapsed_ms(ev[1]) / nl * iters_per_launch
    tgv_bytes = 4 * TGV_FLOATS_PER_PX * W * H * iters_per_launch
    achieved = tgv_bytes / (tgv_launch_ms * 1e-3) / 1e9
    # the sweep kernel alone
    ev[0].record(st)
    pl.sweep.run(fast=fast, stream=s)
    ev[1].record(st)
    ev[1].synchronize()
    sweep_kernel_ms = ev[0].elapsed_ms(ev[1])
    sweep_tflops = SWEEP_FLOP_PER_EVAL * W * H * NPLANES * NSRC / (sweep_kernel_ms * 1e-3) / 1e12
    # the network alone (features, forward, warm-start fields)
    ev[0].record(st)
    for _ in range(10):
        pl.ws.predict(pl.sweep, ZFAR, s)
    ev[1].record(st)
    ev[1].synchronize()
    net_ms = ev[0].elapsed_ms(ev[1]) / 10
    net_tflops = warmstart.forward_flops(W, H) / (net_ms * 1e-3) / 1e12

    traffic = None
    tfile = os.path.join(ROOT, "profiles", tgv_kernel + "_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    rec = {
        "metric": "640x480 depthmaps/s at equal TGV2 gap: learned warm-start vs plain TGV2",
        "value": 1e3 / ms_per_step,
        "unit": "depthmaps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded room scene rendered on the GPU)",
        "config": {"workload": "learned warm-start pipeline: 640x480 reference + 8 sources, 256 inverse-depth planes "
                               "0.3-8 m, 9x9 NCC, fused sweep+argmin, BF16 U-Net warm start, TGV2 linearised at the "
                               "sweep depth to relative energy gap 1e-4 of E* (40000-iteration plain run)",
                   "sweep_arith": "fast" if fast else "reference-order", "tgv_gap_tol": GAP_TOL,
                   "check_every": args.check_every, "max_iters": args.max_iters, "estar_iters": args.estar_iters,
                   "seed": args.seed, "network_dtype": "bf16 (fp32 accumulation)"},
        "speedup_vs_plain": plain["ms_per_depthmap"] / ms_per_step,
        "learned": learned,
        "plain": plain,
        "cold": cold,
        "e_star": e_star,
        "energy_mask": {"qmax": QMAX, "excluded_pixels": excluded, "estar_change_last_5000": (e_prev - e_star) / e_star},
        "roofline": {"kernel": tgv_kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": tgv_bytes, "launch_ms": tgv_launch_ms,
                     "iterations_per_launch": iters_per_launch,
                     "bytes_note": f"algorithmic bytes of {iters_per_launch} TGV2 iteration(s): "
                                   f"{4 * TGV_FLOATS_PER_PX} B per pixel per iteration"},
        "stages": {"sweep_kernel_ms": sweep_kernel_ms, "sweep_tflops_declared": sweep_tflops,
                   "sweep_frac_fp32_peak": sweep_tflops / FP32_PEAK_TFLOPS, "unet_ms": net_ms,
                   "unet_tflops": net_tflops, "unet_frac_bf16_peak": net_tflops / BF16_PEAK_TFLOPS,
                   "unet_fwd_gflop": warmstart.forward_flops(W, H) / 1e9},
    }
    from psw_amd.device import Event, Stream
    rec["sparse_tgv"] = run_sparse(pl, Stream(), [Event() for _ in range(2)], args.seed)
    log(f"sparse TGV2: {rec['sparse_tgv']['iter_us']:.2f} us per iteration "
        f"({rec['sparse_tgv']['gbs_algorithmic']:.0f} GB/s algorithmic)")
    rec["tvl1"] = run_tvl1(pl, Stream(), [Event() for _ in range(2)])
    log(f"TVL1: {rec['tvl1']['iter_us']:.2f} us per iteration fused, "
        f"{rec['tvl1']['launch_sequence_iter_us']:.2f} us as two launches per iteration")
    if not args.no_fusion:
        t0 = time.perf_counter()
        o, thr = fusion_workload(args.seed)
        log(f"fusion depthmaps rendered in {time.perf_counter() - t0:.1f} s")
        from psw_amd.device import Event, Stream
        rec["fusion"] = run_fusion(o, thr, Stream(), [Event() for _ in range(3)])
        log(f"fusion: histogram {rec['fusion']['hist_ms']:.2f} ms, {FUSION_ITERS} PD iterations "
            f"{rec['fusion']['pd_ms']:.2f} ms ({rec['fusion']['pd_gbs_algorithmic']:.0f} GB/s algorithmic)")
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        vh = {"ref": pl.v["d_ref"].to_host(), "src": pl.v["d_src"].to_host(), "Rref": pl.v["Rref"],
              "tref": pl.v["tref"], "Rsrc": pl.v["Rsrc"], "tsrc": pl.v["tsrc"]}
        rec["cpu_baseline"] = cpu_baseline(vh, pl.planes, pl.K0, pl.K1, ds_host, int(round(plain["iters"])))
        log(f"cpu baseline sample took {time.perf_counter() - t0:.1f} s")
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
