#!/usr/bin/env python3
"""Throughput of the multiview-stereo pipeline on one MI355X, learned
warm-start TGV2 against plain TGV2 at equal energy gap.

Workload (one step = one refined 640x480 depthmap): seeded synthetic rooms
rendered on the GPU (eight scenes, seeds 2024-2031; the network was trained
on other seeds), 640x480 float32 reference view and 8 source views (0..255
intensities), pinhole K fx=fy=480, cx=319.5, cy=239.5; poses within 5 cm /
2 degrees of the reference.
  sweep : 256 fronto-parallel planes uniform in inverse depth over 0.3-8 m,
          9x9 NCC window, aggregated cost -(1/N) sum_s NCC, fused kernel with
          argmin / min / second-best (no cost volume in HBM);
  TGV2  : linearised at the sweep's argmin depth, lambda=1, alpha0=2,
          alpha1=1.5, tau=0.02, sigma=1, iterated until the primal energy is
          at most 1e-4 (relative) above E*, checked every 10 iterations,
          capped at 50000.  E* is the energy the reference's start settles
          at (computed per scene before the timed region): the energy never
          settles exactly but moves in a band a few 1e-5 wide (the pixels
          whose data step is unstable oscillate, and that reaches their
          neighbours), so E* is the mean of the plain start's energy every
          10000 iterations from 100000 to 200000 (the band is reported per
          scene).  The rule is one-sided because PDHG from another start can
          settle on a fixed point a few 1e-4 below E* (reported per scene:
          the learned start's mean over the same iterations).  The energy leaves out the ~0.2 % of pixels where the reference's
          fixed steps make the data step unstable (tau sigma lambda^2
          sum Iu^2 > 1 at the pixel or a neighbour): the iteration oscillates
          there forever;
  starts: learned = the BF16 U-Net's (u, u1, p, q) from the sweep results,
          kept only if its energy is below the plain start's, else the plain
          start, and restarted from the plain start if it has not reached the
          gap after 20000 iterations (the headline value, network time
          included); plain = the
          reference's warp start u = u0 (the sweep depth), u1 = p = q = r = 0;
          cold = the reference's constant u = 1.
Each start is timed over the same number of depthmaps, cycling over the
scenes; the speed-up is the median over scenes.  Inputs are resident in HBM
when the timed region starts.  Prints one JSON line.
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
ESTAR_ITERS = 100000     # the plain start's iterations before E* is sampled
ESTAR_END = 200000       # E*: the mean of its energy every 10000 iterations from ESTAR_ITERS to here
FALLBACK_AFTER = 20000   # a learned solve not at the gap by then restarts from the plain start
QMAX = 1.0   # energy mask: pixels with tau sigma lambda^2 sum Iu^2 > QMAX (or at a 4-neighbour) left out
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
SWEEP_FLOP_PER_EVAL = 50          # declared FLOP per (pixel, plane, source) NCC evaluation
SWEEP_HBM_BYTES = 4 * W * H * (1 + NSRC) + 16 * W * H   # images read once, outputs and keys once
TGV_UNFUSED_FLOATS_PER_PX = 27 + 4 * NSRC   # floats per pixel of one unfused iteration
SEEDS = tuple(range(2024, 2032))           # eight scenes; the network was trained on seeds from 100000


def tgv_compulsory_floats(m):
    """Floats per pixel one fused launch of m iterations must move: the state
    read once (u of k and k-1: 6, p 2, q 3, r N, T 3, u0 1, It N, Iu N) and
    written once (u of the last two iterations: 6, p 2, q 3, r N)."""
    return (15 + 3 * NSRC) + (11 + NSRC) if m > 1 else (15 + 3 * NSRC) + (8 + NSRC)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _median_run(fn, reps=3):
    """Seconds: one untimed warm-up run, then the median of `reps` timed runs."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(v, planes, K0, K1, ds, plain_iters, fusion_o=None, fusion_thr=None, sample_iters=300):
    """FP32 oracle (the CPU restatement, OpenMP) on the same inputs.  The
    headline: the full aggregated sweep of one depthmap, then `sample_iters`
    TGV2 iterations linearised at the GPU's sweep depth; the depthmap rate is
    the sweep time plus the sampled per-iteration time times the plain
    pipeline's iteration count (the CPU path has no learned start).  Beside
    it: the small plain pipeline (640x480, 4 sources, 64 planes over 0.5-5 m,
    7x7, 200 cold-start TGV2 iterations) and fusion primal-dual iterations on
    a 128^3 sub-grid of the fusion volume, per voxel-iteration.  Every time is
    the median of three runs after a warm-up run."""
    import oracle
    from psw_amd import stages
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t_sweep = _median_run(lambda: oracle.sweep_agg(v["ref"], v["src"], K0, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"],
                                                   planes, WIN, STD_T))
    t_it = _median_run(lambda: oracle.tgv(v["ref"], v["src"], K1, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"],
                                          sample_iters, 1, u_init=ds, **TGV)) / sample_iters
    t_map = t_sweep + t_it * plain_iters
    out = {"value": 1.0 / t_map, "unit": "depthmaps/s", "cores": cores, "kind": "port",
           "sample": f"FP32 oracle: one full aggregated sweep ({NPLANES} planes x {NSRC} sources, 640x480) and "
                     f"{sample_iters} TGV2 iterations, scaled to the plain pipeline's {plain_iters} iterations",
           "timing": "each time the median of 3 runs after 1 warm-up run",
           "sweep_s": t_sweep, "tgv_iter_ms": t_it * 1e3}
    # the small plain pipeline: 4 sources, 64 planes, 7x7, 200 cold-start iterations
    n4 = SMALL_NSRC
    src4 = np.ascontiguousarray(v["src"][:n4])
    p64 = stages.inverse_depth_planes(SMALL_ZNEAR, SMALL_ZFAR, SMALL_PLANES)
    t_s4 = _median_run(lambda: oracle.sweep_agg(v["ref"], src4, K0, v["Rref"], v["tref"], v["Rsrc"][:n4],
                                                v["tsrc"][:n4], p64, SMALL_WIN, STD_T))
    t_t4 = _median_run(lambda: oracle.tgv(v["ref"], src4, K1, v["Rref"], v["tref"], v["Rsrc"][:n4], v["tsrc"][:n4],
                                          SMALL_ITERS, 1, **TGV))
    out["small_pipeline"] = {"value": 1.0 / (t_s4 + t_t4), "unit": "depthmaps/s", "sweep_s": t_s4, "tgv_s": t_t4,
                             "sample": "640x480, 4 sources, 64 planes 0.5-5 m, 7x7, TGV2 cold start 200 iterations"}
    if fusion_o is not None:
        g, n_it = 128, 20
        hist = np.zeros((g, g, g, 8), np.uint16)
        for k in range(fusion_o["depths"].shape[0]):
            oracle.fusion_hist(hist, FUSION_VOLUME, fusion_o["depths"][k], fusion_o["K"], fusion_o["R"][k],
                               fusion_o["t"][k], fusion_thr)
        u, vv, pp = (np.zeros((g, g, g), np.float32), np.zeros((g, g, g), np.float32),
                     np.zeros((3, g, g, g), np.float32))
        def pd():
            for _ in range(n_it):
                oracle.fusion_u(u, vv, pp, hist, 0.1, 0.5)
                oracle.fusion_p(pp, vv, 1.0)
        t_pd = _median_run(pd)
        out["fusion_pd"] = {"value": g ** 3 * n_it / t_pd, "unit": "voxel-iterations/s",
                            "sample": f"{g}^3 sub-grid of the fusion volume, histograms of the same "
                                      f"{fusion_o['depths'].shape[0]} depthmaps, {n_it} U+P iterations"}
    return out


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
    input.  Per-iteration time from a 500- and a 0-iteration call, for the
    resident solver and for the fused launches."""
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

    # the resident solver (the default; =2 makes a run that falls back an
    # error) and the fused launches (=0)
    res = {}
    prev = os.environ.get("PSW_TGV_SPARSE_RESIDENT")
    try:
        for name, flag in (("fused", "0"), ("resident", "2")):
            os.environ["PSW_TGV_SPARSE_RESIDENT"] = flag
            timed(20)
            t_full, t_zero = timed(SPARSE_ITERS), timed(0)
            res[name] = ((t_full - t_zero) / SPARSE_ITERS, t_full, t_zero)
    finally:
        if prev is None:
            os.environ.pop("PSW_TGV_SPARSE_RESIDENT", None)
        else:
            os.environ["PSW_TGV_SPARSE_RESIDENT"] = prev
    it_ms, t_full, t_zero = res["resident"]
    f_ms = res["fused"][0]
    gbs = SPARSE_BYTES_PER_PIXEL_ITER * W * H / (f_ms * 1e-3) / 1e9
    multi = os.environ.get("PSW_TGV_SPARSE_DOUBLE", "1") != "0"   # several iterations per launch (default)
    m = int(os.environ.get("PSW_TGV_SPARSE_M", "4")) if multi else 1
    kernel = {1: "tgv_sparse_iter_kernel", 2: "tgv_sparse_iter2_kernel"}.get(m, f"tgv_sparse_iterm_kernel<{m}>")
    return {"kernel": "tgv_sparse_resident_kernel", "iter_us": it_ms * 1e3, "call_ms": t_full,
            "setup_output_ms": t_zero, "speedup_vs_fused": f_ms / it_ms,
            "note": "the default path: all iterations in one cooperative launch of 64x20 tiles that trade only their "
                    "edges' u, u1 each iteration (bit-identical to the fused launches); it reads the state once per "
                    "call, so its bound is the edge exchange and VALU issue, not HBM",
            "fused_kernel": kernel, "fused_iterations_per_launch": m, "fused_iter_us": f_ms * 1e3,
            "fused_gbs_per_step_bytes": gbs, "fused_frac_hbm_peak_per_step_bytes": gbs / HBM_PEAK_GBS,
            "bytes_per_pixel_iter": SPARSE_BYTES_PER_PIXEL_ITER,
            "config": f"{W}x{H}, sweep depth as the dense start, true depth at 5% of pixels, {SPARSE_ITERS} "
                      "iterations, alpha0 2 alpha1 1.5 tau 0.02 sigma 1"}


TVL1_ITERS = 1000
# per pixel and iteration, with nothing kept on chip between iterations: read
# depth, Px, Py, R, the scaled input and T11..T22 (36 B), write depth, Px, Py, R (16 B)
TVL1_BYTES_PER_PIXEL_ITER = 52


SMALL_NSRC, SMALL_PLANES, SMALL_WIN, SMALL_ZNEAR, SMALL_ZFAR, SMALL_ITERS = 4, 64, 7, 0.5, 5.0, 200


def run_small(pl, stream, ev, fast, reps=20):
    """The small plain pipeline on the first scene's views: the aggregated
    sweep over the first 4 sources (64 planes uniform in inverse depth over
    0.5-5 m, 7x7 window), then PlaneSweep::TGV from the constant start u = 1
    with one linearisation and 200 iterations (the reference's TGV call).
    Per depthmap over `reps` runs after two untimed ones, inputs resident."""
    from psw_amd import stages
    from psw_amd.device import DeviceArray
    v, s = pl.v, stream.handle
    d_src4 = DeviceArray.from_host(v["d_src"].to_host()[:SMALL_NSRC])
    sw = stages.SweepAggregate(v["d_ref"], d_src4, pl.K0, v["Rref"], v["tref"], v["Rsrc"][:SMALL_NSRC],
                               v["tsrc"][:SMALL_NSRC],
                               stages.inverse_depth_planes(SMALL_ZNEAR, SMALL_ZFAR, SMALL_PLANES), SMALL_WIN, STD_T,
                               coord_offset=0)
    tgv = stages.TgvSolver(W, H, SMALL_NSRC)
    tgv.set_images(v["d_ref"], d_src4, TGV["beta"], TGV["gamma"], s)
    Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"][:SMALL_NSRC], v["tsrc"][:SMALL_NSRC])
    d_u = DeviceArray((H, W))
    t_sw, t_tgv = [], []
    for i in range(reps + 2):
        ev[0].record(stream)
        sw.build_homographies(s)
        sw.run(fast=fast, stream=s)
        sw.depth(SMALL_ZFAR, s)
        ev[1].record(stream)
        tgv.run(pl.K1, Rrel, trel, SMALL_ITERS, 1, pl.prm, d_u, stream=s)
        ev[2].record(stream)
        ev[2].synchronize()
        if i >= 2:
            t_sw.append(ev[0].elapsed_ms(ev[1]))
            t_tgv.append(ev[1].elapsed_ms(ev[2]))
    ms = float(np.mean(t_sw) + np.mean(t_tgv))
    return {"value": 1e3 / ms, "unit": "depthmaps/s", "ms_per_depthmap": ms, "sweep_ms": float(np.mean(t_sw)),
            "tgv_ms": float(np.mean(t_tgv)), "tgv_us_per_iteration": float(np.mean(t_tgv)) * 1e3 / SMALL_ITERS,
            "runs": reps, "sweep_arith": "fast" if fast else "reference-order",
            "config": f"{W}x{H}, {SMALL_NSRC} sources, {SMALL_PLANES} planes {SMALL_ZNEAR}-{SMALL_ZFAR} m, "
                      f"{SMALL_WIN}x{SMALL_WIN}, TGV2 cold start {SMALL_ITERS} iterations (one linearisation)"}


def run_tvl1(pl, stream, ev):
    """TVL1 denoising (PlaneSweep::CudaDenoise) of the 640x480 sweep depthmap
    with the reference's default parameters.  Per-iteration time from a
    1000- and a 0-iteration call, for the resident solver (the default: one
    cooperative launch for all iterations; PSW_TVL1_RESIDENT=2 makes a run
    that falls back an error), the fused launches (PSW_TVL1_RESIDENT=0) and
    the reference's two launches per iteration (PSW_TVL1_FUSED=0)."""
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
    keys = ("PSW_TVL1_FUSED", "PSW_TVL1_RESIDENT")
    prev = {k: os.environ.get(k) for k in keys}
    try:
        for name, fused, resident in (("launch_sequence", "0", "0"), ("fused", "1", "0"), ("resident", "1", "2")):
            os.environ["PSW_TVL1_FUSED"], os.environ["PSW_TVL1_RESIDENT"] = fused, resident
            timed(20)
            t_full, t_zero = timed(TVL1_ITERS), timed(0)
            res[name] = (t_full - t_zero) / TVL1_ITERS
    finally:
        for k, val in prev.items():
            if val is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = val
    gbs = TVL1_BYTES_PER_PIXEL_ITER * W * H / (res["fused"] * 1e-3) / 1e9
    m = int(os.environ.get("PSW_TVL1_M", "4"))                # iterations per launch (default 4, 56x8 tiles)
    return {"kernel": "tvl1_resident_kernel", "iter_us": res["resident"] * 1e3,
            "speedup_vs_fused": res["fused"] / res["resident"],
            "speedup_vs_launch_sequence": res["launch_sequence"] / res["resident"],
            "note": "the default path: all iterations in one cooperative launch of 64x20 tiles that trade only their "
                    "edges' depth each iteration (bit-identical to the launch sequence); it reads the state once "
                    "per call, so its bound is the per-iteration edge exchange and VALU issue, not HBM",
            "fused_kernel": f"tvl1_iter_kernel<{m}>", "fused_iterations_per_launch": m,
            "fused_iter_us": res["fused"] * 1e3, "launch_sequence_iter_us": res["launch_sequence"] * 1e3,
            "fused_gbs_per_step_bytes": gbs, "fused_frac_hbm_peak_per_step_bytes": gbs / HBM_PEAK_GBS,
            "bytes_per_pixel_iter": TVL1_BYTES_PER_PIXEL_ITER,
            "config": f"{W}x{H} sweep depthmap, {TVL1_ITERS} iterations, lambda 0.3 tau 0.02 sigma 6 theta 1 "
                      "beta 0 gamma 1 (the reference's defaults)"}


class Pipeline:
    """Device state of the benchmarked pipeline for one seeded view."""

    def __init__(self, seed, weights, ws=None):
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
        self.ws = ws if ws is not None else (warmstart.WarmStarter(W, H, weights) if weights else None)
        self.ev = [Event() for _ in range(4)]
        self.seed = seed
        self.e_star = None
        self.last_start = None

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
        self.last_start = mode
        if mode == "learned":
            iters, e, self.last_start = self.ws.solve(self.tgv, d_s, self.prm, e_star, GAP_TOL, max_iters, check_every,
                                                      fallback_after=FALLBACK_AFTER, stream=s)
        else:
            if mode == "cold":
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
        elif name in ("tgv", "tgv_resident", "sweep", "sweep_exact", "unet"):
            if pl is None:
                pl = Pipeline(seed, warmstart.DEFAULT_WEIGHTS if name == "unet" else None)
            if name in ("tgv", "tgv_resident"):
                # tgv: 100 iterations on the launch sequence (50 two-iteration
                # launches); tgv_resident: 1000 in one resident launch
                d_s = pl.sweep_and_linearise()
                pl.tgv.set_u(d_s, stream=pl.st.handle)
                pl.tgv.linearise(pl.K1, pl.Rrel, pl.trel, pl.st.handle)
                saved = os.environ.get("PSW_TGV_RESIDENT")
                os.environ["PSW_TGV_RESIDENT"] = "1" if name == "tgv_resident" else "0"
                pl.tgv.iterate(1000 if name == "tgv_resident" else 100, pl.prm, pl.st.handle)
                if saved is None:
                    del os.environ["PSW_TGV_RESIDENT"]
                else:
                    os.environ["PSW_TGV_RESIDENT"] = saved
                pl.st.synchronize()
            elif name == "unet":
                if pl.ws is None:
                    pl.ws = warmstart.WarmStarter(W, H, warmstart.DEFAULT_WEIGHTS)
                pl.sweep_and_linearise()
                for _ in range(5):
                    pl.ws.predict(pl.sweep, ZFAR, pl.st.handle)
            else:
                for _ in range(3):
                    pl.sweep_and_linearise(fast=(name == "sweep"))
        elif name == "sweep_shapes":
            # the aggregated sweep at 640x480 x 8 sources x 256 planes x 9x9 for
            # each tile height and plane-group count, both arithmetic variants
            if pl is None:
                pl = Pipeline(seed, None)
            pl.sweep.build_homographies(pl.st.handle)
            ev = [Event(), Event()]
            times = {}
            saved = {k: os.environ.get(k) for k in ("PSW_SWEEP_TY", "PSW_SWEEP_GROUPS")}
            try:
                for ty in ("16", "32"):
                    for g in ("1", "2", "3", "5", "8", "12", "16", "24", "32", "48", "64"):
                        os.environ["PSW_SWEEP_TY"], os.environ["PSW_SWEEP_GROUPS"] = ty, g
                        for fast in (True, False):
                            pl.sweep.run(fast=fast, stream=pl.st.handle)
                            ev[0].record(pl.st)
                            for _ in range(5):
                                pl.sweep.run(fast=fast, stream=pl.st.handle)
                            ev[1].record(pl.st)
                            ev[1].synchronize()
                            times[f"tile32x{ty}_groups{g}_{'fast' if fast else 'reference-order'}_ms"] = \
                                ev[0].elapsed_ms(ev[1]) / 5
            finally:
                for k, val in saved.items():
                    if val is None:
                        os.environ.pop(k, None)
                    else:
                        os.environ[k] = val
            log(json.dumps(times))
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


def run_batch(args, emit=True):
    """The uneven batch: 32 seeded frames, resolution from
    {320x240, 640x480, 1242x375}, 4-8 sources, 64-256 planes per frame; sweep
    and TGV2 (plain start, linearised at the sweep depth, to the energy gap)
    for the whole batch in batched launches, against the same frames one at
    a time through the single-frame entry points.  Prints its own JSON line
    when `emit`; returns the record."""
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

    ty_env = os.environ.get("PSW_SWEEP_TY")
    os.environ["PSW_SWEEP_TY"] = str(best["sweep_tile_h"])

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
    gaps = [(e - es) / es for e, es in zip(energies, e_star)]
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
                     "sweep_ms": {f"tile32x{ty}_groups{g}_wg{w}": v for (ty, g, w), v in timings["sweep"].items()},
                     "tgv_us_per_iteration": {f"variant{v}": ms * 1e3 for v, ms in timings["tgv"].items()},
                     "solve_us_per_iteration": {("graph" if g else "direct"): ms * 1e3
                                                for g, ms in timings["graph"].items()}},
        "iterations": [int(i) for i in iters], "max_rel_gap_at_stop": float(max(gaps)),
        "sequential_per_frame_ms": seq_ms,
    }
    log(f"batch: {t_batch * 1e3:.1f} ms for {len(frames)} frames ({len(frames) / t_batch:.2f} depthmaps/s); "
        f"one at a time: {t_seq * 1e3:.1f} ms")
    if ty_env is None:
        os.environ.pop("PSW_SWEEP_TY", None)
    else:
        os.environ["PSW_SWEEP_TY"] = ty_env
    if emit:
        print(json.dumps(rec), flush=True)
    return rec


def main():
    from psw_amd import warmstart
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16,
                    help="timed depthmaps per start (learned, plain, cold), cycling over the scenes")
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--seeds", default=",".join(str(x) for x in SEEDS), help="comma list of scene seeds")
    ap.add_argument("--exact-sweep", action="store_true", help="reference-order (bit-exact) sweep arithmetic")
    ap.add_argument("--max-iters", type=int, default=MAX_ITERS)
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--estar-iters", type=int, default=ESTAR_ITERS)
    ap.add_argument("--weights", default=warmstart.DEFAULT_WEIGHTS)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cold", action="store_true", help="skip the constant-start (u = 1) run")
    ap.add_argument("--seed", type=int, default=SEEDS[0], help="scene of --probe and --batch")
    ap.add_argument("--no-fusion", action="store_true", help="skip the 256^3 fusion measurement")
    ap.add_argument("--batch", action="store_true", help="run the uneven 32-frame batch instead (its own JSON line)")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven 32-frame batch in the default run")
    ap.add_argument("--batch-frames", type=int, default=32)
    ap.add_argument("--batch-estar-iters", type=int, default=20000)
    ap.add_argument("--probe", default="", help="comma list of tgv,tgv_resident,sweep,sweep_exact,sweep_shapes,unet,copy,fusion,"
                                                "sparse,tvl1: run only these small workloads (for rocprofv3) and exit")
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
    from psw_amd.device import Event, Stream
    log("device:", psw_amd.device_name())
    fast = not args.exact_sweep
    seeds = [int(x) for x in args.seeds.split(",")]
    ws = warmstart.WarmStarter(W, H, args.weights)
    pipes = []
    for sd in seeds:
        pl = Pipeline(sd, args.weights, ws=ws)
        s, st = pl.st.handle, pl.st
        # E*: the same problem (linearised at the sweep depth) iterated from
        # the plain start; not timed
        d_s = pl.sweep_and_linearise(fast)
        pl.tgv.set_u(d_s, stream=s)
        pl.tgv.linearise(pl.K1, pl.Rrel, pl.trel, s)
        pl.excluded = pl.tgv.set_energy_mask(QMAX, pl.prm, s)
        # The energy does not settle: after some 10^4 iterations it moves in
        # a band a few 1e-5 wide for good (the pixels whose data step is
        # unstable oscillate, and that reaches their neighbours), so one
        # sample would be an arbitrary point of the band.  E* is the mean of
        # the samples every 10000 iterations from estar_iters to ESTAR_END.
        def band(n0):
            pl.tgv.iterate(n0, pl.prm, s)
            es, n = [pl.tgv.energy(pl.prm, s)], n0
            while n < max(ESTAR_END, n0):
                pl.tgv.iterate(10000, pl.prm, s)
                n += 10000
                es.append(pl.tgv.energy(pl.prm, s))
            return np.array(es), n
        es, n_it = band(args.estar_iters)
        pl.e_star_plain, pl.estar_iters = float(np.mean(es)), n_it
        pl.estar_band = [float((es.min() - es.mean()) / es.mean()), float((es.max() - es.mean()) / es.mean())]
        pl.estar_samples = len(es)
        pl.u_star = pl.tgv.get_u(s)
        pl.weight = pl.tgv.get_energy_mask(s)
        pl.d_s_host = d_s.to_host(s)
        pl.true_depth = pl.v["d_depth"].to_host(s)
        # the learned start iterated as long: does it settle at the same energy?
        pl.tgv.set_u(d_s, stream=s)
        pl.tgv.linearise(pl.K1, pl.Rrel, pl.trel, s)
        pl.tgv.set_energy_mask(QMAX, pl.prm, s, count=False)
        pl.ws.predict(pl.sweep, ZFAR, s)
        pl.ws.start(pl.tgv, s)
        es_l, _ = band(args.estar_iters)
        pl.e_star_learned = float(np.mean(es_l))
        pl.e_star = pl.e_star_plain
        st.synchronize()
        log(f"seed {sd}: E* = {pl.e_star:.9e}, the plain start's mean energy over {pl.estar_samples} samples "
            f"up to {n_it} iterations (band {pl.estar_band[0]:+.1e} .. {pl.estar_band[1]:+.1e}); the learned "
            f"start's mean over the same: {(pl.e_star_learned - pl.e_star) / pl.e_star:+.2e} relative; "
            f"{pl.excluded} pixels outside the energy mask")
        pipes.append(pl)

    def timed(mode, k, warm):
        """k depthmaps (cycling over the scenes) after `warm` untimed ones."""
        for i in range(warm):
            pipes[i % len(pipes)].run(mode, pipes[i % len(pipes)].e_star, args.max_iters, args.check_every, fast)
        for pl in pipes:
            pl.st.synchronize()
        recs = [[] for _ in pipes]
        t0 = time.perf_counter()
        for i in range(k):
            pl = pipes[i % len(pipes)]
            r = pl.run(mode, pl.e_star, args.max_iters, args.check_every, fast)
            r["start"] = pl.last_start
            recs[i % len(pipes)].append(r)
        for pl in pipes:
            pl.st.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / k
        return ms, recs

    res = {}
    for mode in ("learned", "plain", "cold"):
        if mode == "cold" and args.no_cold:
            continue
        ms, recs = timed(mode, args.steps, args.warmup if mode == "learned" else 0)
        res[mode] = (ms, recs)
        log(f"{mode}: {ms:.3f} ms/depthmap over {args.steps} depthmaps of {len(pipes)} scenes")
    ms_per_step = res["learned"][0]

    # per scene: times, iterations, stop states (one more untimed run of each start)
    per_seed = []
    for j, pl in enumerate(pipes):
        s = pl.st.handle
        rec = {"seed": pl.seed, "e_star": pl.e_star, "estar_iters": pl.estar_iters,
               "estar_samples": pl.estar_samples, "estar_band_rel": pl.estar_band,
               "learned_start_band_mean_rel": (pl.e_star_learned - pl.e_star) / pl.e_star,
               "excluded_pixels": pl.excluded}
        u_stop = {}
        for mode, (_, recs) in res.items():
            rs = recs[j]
            rec[mode] = {"runs": len(rs), "iters": float(np.mean([r["iters"] for r in rs])),
                         "ms_per_depthmap": float(np.mean([r["sweep_ms"] + r["net_ms"] + r["tgv_ms"] for r in rs])),
                         "sweep_ms": float(np.mean([r["sweep_ms"] for r in rs])),
                         "net_ms": float(np.mean([r["net_ms"] for r in rs])),
                         "tgv_ms": float(np.mean([r["tgv_ms"] for r in rs])),
                         "max_rel_gap_at_stop": float(max((r["energy"] - pl.e_star) / pl.e_star for r in rs)),
                         "reached_gap": bool(all(r["energy"] - pl.e_star <= GAP_TOL * pl.e_star for r in rs)),
                         "starts": sorted({r["start"] for r in rs})}
            pl.run(mode, pl.e_star, args.max_iters, args.check_every, fast)
            u_stop[mode] = pl.tgv.get_u(s)
            pl.st.synchronize()
            rec[mode]["stop_state_vs_estar_state"] = warmstart.judge(u_stop[mode], pl.u_star, pl.weight)
            rec[mode]["median_rel_err_vs_true_depth"] = float(np.median(np.abs(u_stop[mode] - pl.true_depth)
                                                                        / pl.true_depth))
        rec["speedup_vs_plain"] = rec["plain"]["ms_per_depthmap"] / rec["learned"]["ms_per_depthmap"]
        if "cold" in res:
            rec["speedup_vs_cold"] = rec["cold"]["ms_per_depthmap"] / rec["learned"]["ms_per_depthmap"]
        # the plain start's solve to the same gap judges the learned one
        rec["judge_learned_vs_plain_stop"] = warmstart.judge(u_stop["learned"], u_stop["plain"], pl.weight)
        per_seed.append(rec)
        log(f"seed {pl.seed}: iterations learned {rec['learned']['iters']:.0f} plain {rec['plain']['iters']:.0f}"
            + (f" cold {rec['cold']['iters']:.0f}" if "cold" in rec else "")
            + f"; speedup vs plain {rec['speedup_vs_plain']:.3f}")
    sp = np.array([r["speedup_vs_plain"] for r in per_seed])
    spc = np.array([r["speedup_vs_cold"] for r in per_seed]) if "cold" in res else None

    pl = pipes[0]
    s, st, ev = pl.st.handle, pl.st, pl.ev
    # Per-launch duration of the fused TGV2 iteration kernel (same stream and
    # buffers): events around back-to-back launches.  By default a launch
    # does two iterations (tgv_iter2_kernel); PSW_TGV_TRIPLE=1 selects the
    # three-iteration kernel, PSW_TGV_DOUBLE=0 the one-iteration kernel.
    two = os.environ.get("PSW_TGV_DOUBLE", "1") != "0"
    three = two and os.environ.get("PSW_TGV_TRIPLE", "0") == "1"
    tgv_kernel, m = (("tgv_iter3_kernel", 3) if three else ("tgv_iter2_kernel", 2) if two else ("tgv_iter_kernel", 1))
    d_s = pl.sweep_and_linearise(fast)
    pl.tgv.set_u(d_s, stream=s)
    pl.tgv.linearise(pl.K1, pl.Rrel, pl.trel, s)
    nl = 198                     # whole launches of 1, 2 or 3 iterations
    resident_env = os.environ.get("PSW_TGV_RESIDENT")
    os.environ["PSW_TGV_RESIDENT"] = "0"       # this measurement is of the launch sequence
    ev[0].record(st)
    pl.tgv.iterate(nl * m, pl.prm, s)
    ev[1].record(st)
    ev[1].synchronize()
    launch_ms = ev[0].elapsed_ms(ev[1]) / nl
    if resident_env is None:
        del os.environ["PSW_TGV_RESIDENT"]
    else:
        os.environ["PSW_TGV_RESIDENT"] = resident_env
    # the resident solver (tgv_resident_kernel: one cooperative launch, the
    # state on chip), the path the solves above took: per iteration over a
    # long run, the launch sequence beside it
    pl.tgv.set_u(d_s, stream=s)
    pl.tgv.linearise(pl.K1, pl.Rrel, pl.trel, s)
    n_res = 4000
    ev[0].record(st)
    pl.tgv.iterate(n_res, pl.prm, s)
    ev[1].record(st)
    ev[1].synchronize()
    resident_us = ev[0].elapsed_ms(ev[1]) * 1e3 / n_res
    compulsory = 4 * tgv_compulsory_floats(m) * W * H
    unfused_bytes = 4 * TGV_UNFUSED_FLOATS_PER_PX * W * H * m
    achieved = compulsory / (launch_ms * 1e-3) / 1e9
    traffic, tsrc = None, None
    tfile = os.path.join(ROOT, "profiles", tgv_kernel + "_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
        tsrc = f"constant: rocprofv3 FETCH_SIZE/WRITE_SIZE runs of {tgv_kernel} at this size, profiles/{tgv_kernel}_traffic.json"
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

    learned_ms, plain_ms = res["learned"][0], res["plain"][0]
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
        "data": f"synthetic (seeded room scenes rendered on the GPU, seeds {seeds[0]}-{seeds[-1]}, none used in training)",
        "config": {"workload": "learned warm-start pipeline: 640x480 reference + 8 sources, 256 inverse-depth planes "
                               "0.3-8 m, 9x9 NCC, fused sweep+argmin, BF16 U-Net warm start (plain start as "
                               "fallback), TGV2 linearised at the sweep depth until the energy is at most 1e-4 "
                               "(relative) above E* (the mean of the plain start's energy every 10000 iterations "
                               f"from {args.estar_iters} to {max(ESTAR_END, args.estar_iters)}: it moves in a band "
                               f"of a few 1e-5 for good); steps cycle over {len(seeds)} scenes",
                   "sweep_arith": "fast" if fast else "reference-order", "tgv_gap_tol": GAP_TOL,
                   "check_every": args.check_every, "max_iters": args.max_iters, "estar_iters": args.estar_iters,
                   "seeds": seeds, "network_dtype": "bf16 (fp32 accumulation)"},
        "speedup_vs_plain": float(np.median(sp)),
        "speedup_vs_plain_range": [float(sp.min()), float(sp.max())],
        "speedup_vs_plain_aggregate": plain_ms / learned_ms,
        "speedup_vs_cold": float(np.median(spc)) if spc is not None else None,
        "speedup_vs_cold_range": [float(spc.min()), float(spc.max())] if spc is not None else None,
        "speedup_vs_cold_aggregate": res["cold"][0] / learned_ms if "cold" in res else None,
        "target_3x_vs_cold_start": ({"comparator": "the reference's cold start (u = 1) on the same problem (linearised "
                                                   "at the sweep depth), the same gap",
                                     "median_per_scene": float(np.median(spc)), "met": bool(np.median(spc) >= 3.0)}
                                    if spc is not None else None),
        "plain": {"ms_per_depthmap": plain_ms, "depthmaps_per_s": 1e3 / plain_ms, "runs": args.steps},
        "cold": ({"ms_per_depthmap": res["cold"][0], "depthmaps_per_s": 1e3 / res["cold"][0], "runs": args.steps}
                 if "cold" in res else None),
        "tgv_only": {mode: {"ms_per_depthmap": float(np.mean([r["tgv_ms"] for rs in recs for r in rs])),
                            "depthmaps_per_s": 1e3 / float(np.mean([r["tgv_ms"] for rs in recs for r in rs]))}
                     for mode, (_, recs) in res.items()},
        "per_seed": per_seed,
        "roofline": {"kernel": tgv_kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": tsrc,
                     "bytes_per_launch": compulsory, "launch_ms": launch_ms, "iterations_per_launch": m,
                     "bytes_note": f"compulsory bytes of one launch of {m} iteration(s) ({4 * tgv_compulsory_floats(m)}"
                                   " B/px: the state read once and written once); the per-iteration count of an "
                                   f"unfused iteration, {4 * TGV_UNFUSED_FLOATS_PER_PX} B/px, gives "
                                   f"{unfused_bytes / 1e6:.1f} MB per launch "
                                   f"({unfused_bytes / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS:.2f} of peak), which "
                                   "counts bytes the fused launch never moves"},
        "tgv_resident": {"kernel": "tgv_resident_kernel", "us_per_iteration": resident_us,
                         "launch_sequence_us_per_iteration": launch_ms * 1e3 / m,
                         "speedup_vs_launch_sequence": launch_ms * 1e3 / m / resident_us,
                         "iterations_timed": n_res,
                         "note": "the solves above run this kernel (bit-identical to the launch sequence); it reads "
                                 "the state from HBM once per solve, so its bound is the per-iteration exchange of "
                                 "tile edges between workgroups and VALU issue, not HBM; the HBM roofline above is of "
                                 "the launch sequence's kernel (batched frames, frames above one tile per CU)"},
        "stages": {"sweep_kernel_ms": sweep_kernel_ms, "sweep_tflops_declared": sweep_tflops,
                   "sweep_frac_fp32_peak": sweep_tflops / FP32_PEAK_TFLOPS,
                   # algorithmic bytes: the images read once, outputs written once
                   "sweep_hbm_bytes": SWEEP_HBM_BYTES, "sweep_frac_hbm_peak":
                       SWEEP_HBM_BYTES / (sweep_kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "unet_ms": net_ms,
                   "unet_tflops": net_tflops, "unet_frac_bf16_peak": net_tflops / BF16_PEAK_TFLOPS,
                   "unet_fwd_gflop": warmstart.forward_flops(W, H) / 1e9},
    }
    rec["sparse_tgv"] = run_sparse(pl, Stream(), [Event() for _ in range(2)], seeds[0])
    log(f"TGV2 resident: {resident_us:.3f} us per iteration; launch sequence {launch_ms * 1e3 / m:.3f} us")
    log(f"sparse TGV2: {rec['sparse_tgv']['iter_us']:.2f} us per iteration resident, "
        f"{rec['sparse_tgv']['fused_iter_us']:.2f} fused")
    rec["tvl1"] = run_tvl1(pl, Stream(), [Event() for _ in range(2)])
    rec["small_pipeline"] = run_small(pl, Stream(), [Event() for _ in range(3)], fast)
    log(f"small pipeline: {rec['small_pipeline']['ms_per_depthmap']:.3f} ms per depthmap (sweep "
        f"{rec['small_pipeline']['sweep_ms']:.3f} ms, TGV2 {rec['small_pipeline']['tgv_ms']:.3f} ms)")
    log(f"TVL1: {rec['tvl1']['iter_us']:.2f} us per iteration resident, {rec['tvl1']['fused_iter_us']:.2f} fused, "
        f"{rec['tvl1']['launch_sequence_iter_us']:.2f} us as two launches per iteration")
    if not args.no_fusion:
        t0 = time.perf_counter()
        o, thr = fusion_workload(seeds[0])
        log(f"fusion depthmaps rendered in {time.perf_counter() - t0:.1f} s")
        rec["fusion"] = run_fusion(o, thr, Stream(), [Event() for _ in range(3)])
        log(f"fusion: histogram {rec['fusion']['hist_ms']:.2f} ms, {FUSION_ITERS} PD iterations "
            f"{rec['fusion']['pd_ms']:.2f} ms")
    if not args.no_batch:
        b = run_batch(args, emit=False)
        rec["uneven_batch"] = {k: b[k] for k in ("value", "unit", "batch_ms", "sequential_ms", "speedup_vs_sequential",
                                                 "max_rel_gap_at_stop")}
        lat = {}
        for fr, ms in zip(b["config"]["frames"], b["sequential_per_frame_ms"]):
            lat.setdefault(f"{fr['W']}x{fr['H']}", []).append(ms)
        rec["uneven_batch"].update(workload=b["config"]["workload"], megapixels=b["config"]["megapixels"],
                                   iterations_range=[min(b["iterations"]), max(b["iterations"])],
                                   autotuned=b["autotune"]["chosen"],
                                   sequential_latency_ms={k: {"frames": len(v), "min": float(np.min(v)),
                                                              "median": float(np.median(v)), "max": float(np.max(v))}
                                                          for k, v in lat.items()})
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        vh = {"ref": pl.v["d_ref"].to_host(), "src": pl.v["d_src"].to_host(), "Rref": pl.v["Rref"],
              "tref": pl.v["tref"], "Rsrc": pl.v["Rsrc"], "tsrc": pl.v["tsrc"]}
        plain_iters = int(round(per_seed[0]["plain"]["iters"]))
        rec["cpu_baseline"] = cpu_baseline(vh, pl.planes, pl.K0, pl.K1, pl.d_s_host, plain_iters,
                                           o if not args.no_fusion else None, thr if not args.no_fusion else None)
        log(f"cpu baseline samples took {time.perf_counter() - t0:.1f} s")
        cs = rec["cpu_baseline"]["small_pipeline"]
        rec["small_pipeline"]["vs_cpu_baseline"] = rec["small_pipeline"]["value"] / cs["value"]
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
