"""Many frames of uneven resolution, source and plane counts per launch:
the persistent batched sweep, the batched TGV2 solver (hipGraph-replayed
check chunks), a launch-shape autotuner, and the seeded uneven workload.

Every batched result is bit-identical, frame by frame, to the single-frame
entry points (tests/test_batch.py); the autotuner only picks launch shapes
(plane groups, resident workgroups, TGV2 tile shape), which never change
results.
"""
import ctypes
import os

import numpy as np

from . import scene, stages
from ._lib import SweepFrame, TgvParams, call
from .device import Event, Stream

RESOLUTIONS = ((320, 240), (640, 480), (1242, 375))


def _ptrs(solvers):
    arr = (ctypes.c_void_p * len(solvers))(*[t.h.value for t in solvers])
    return arr


def _prms(prms, n):
    if isinstance(prms, TgvParams):
        prms = [prms] * n
    return (TgvParams * n)(*prms)


def sweep_batch(sweeps, fast=True, groups=0, workgroups=0, stream=None):
    """psw_sweep_fused_aggregate_batch over SweepAggregate objects (same
    window, std threshold and coordinate offset); results land in each
    object's d_arg / d_cmin / d_csec.  Homographies must be built."""
    w = sweeps[0]
    fr = (SweepFrame * len(sweeps))()
    for i, sw in enumerate(sweeps):
        assert (sw.winsize, sw.stdthresh, sw.coord_offset) == (w.winsize, w.stdthresh, w.coord_offset)
        fr[i] = SweepFrame(sw.d_ref.ptr.value, sw.d_src.ptr.value, sw.W, sw.H_img, sw.N, sw.d_H.ptr.value, sw.P,
                           sw.d_arg.ptr.value, sw.d_cmin.ptr.value, sw.d_csec.ptr.value)
    call("psw_sweep_fused_aggregate_batch", fr, len(sweeps), w.winsize, w.stdthresh, w.coord_offset, int(bool(fast)),
         int(groups), int(workgroups), stream)


def tgv_iterate_batch(solvers, niters, prms, variant=-1, stream=None):
    call("psw_tgv_iterate_batch", _ptrs(solvers), len(solvers), int(niters), _prms(prms, len(solvers)), int(variant),
         stream)


def tgv_solve_batch(solvers, prms, e_stars, max_iters, check_every, tol, variant=-1, use_graph=True, stream=None):
    n = len(solvers)
    es = (ctypes.c_double * n)(*[float(e) for e in e_stars])
    it = (ctypes.c_int * n)()
    en = (ctypes.c_double * n)()
    call("psw_tgv_solve_batch", _ptrs(solvers), n, _prms(prms, n), es, int(max_iters), int(check_every), float(tol),
         int(variant), int(bool(use_graph)), it, en, stream)
    return np.array(it[:], np.int64), np.array(en[:], np.float64)


def uneven_workload(seed, n_frames=32, nsrc=(4, 8), planes=(64, 256)):
    """The seeded uneven batch: per frame a resolution from RESOLUTIONS, a
    source count in [nsrc] and a plane count in [planes] (inclusive)."""
    rng = np.random.Generator(np.random.PCG64(int(seed) + 4242))
    out = []
    for f in range(n_frames):
        W, H = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        out.append({"seed": int(seed) * 1000 + f, "W": W, "H": H,
                    "N": int(rng.integers(nsrc[0], nsrc[1] + 1)), "P": int(rng.integers(planes[0], planes[1] + 1))})
    return out


class Frame:
    """One frame of a batch: GPU-rendered views, its sweep and its solver."""

    def __init__(self, spec, znear=0.3, zfar=8.0, winsize=9, stdthresh=1e-4, stream=None):
        self.spec = spec
        W, H, N, P = spec["W"], spec["H"], spec["N"], spec["P"]
        K = scene.intrinsics(W, H, 480.0 * W / 640.0)
        self.v = scene.make_views_gpu(spec["seed"], W, H, N, K=K, stream=stream)
        v = self.v
        self.K0, self.K1 = K, scene.one_based(K)
        self.planes = stages.inverse_depth_planes(znear, zfar, P)
        self.zfar = zfar
        self.sweep = stages.SweepAggregate(v["d_ref"], v["d_src"], K, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"],
                                           self.planes, winsize, stdthresh, coord_offset=0)
        self.tgv = stages.TgvSolver(W, H, N)
        self.Rrel, self.trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])

    def linearise(self, prm, qmax=1.0, stream=None):
        d_s = self.sweep.depth(self.zfar, stream)
        self.tgv.set_images(self.sweep.d_ref, self.sweep.d_src, prm.beta, prm.gamma, stream)
        self.tgv.set_u(d_s, stream=stream)
        self.tgv.linearise(self.K1, self.Rrel, self.trel, stream)
        if qmax > 0:
            self.tgv.set_energy_mask(qmax, prm, stream, count=False)


def _time(fn, stream, reps=3):
    ev0, ev1 = Event(), Event()
    fn()
    ev0.record(stream)
    for _ in range(reps):
        fn()
    ev1.record(stream)
    ev1.synchronize()
    return ev0.elapsed_ms(ev1) / reps


def autotune(frames, prm, fast=True, solve_iters=200, stream=None, log=None):
    """Time the launch shapes of the batched sweep (NCC tile height x plane
    groups x resident workgroups; the tile height sets the LDS halo tile,
    32 + 2r by TY + 2r for window radius r, through PSW_SWEEP_TY), of the
    batched TGV2 iteration (tile shape of the one-iteration launches; the
    default shape lets the resident solver take the frames that fit it) and
    the batched solve's chunks replayed from a
    captured hipGraph or launched directly, on these frames; returns the
    fastest settings and the timings.  Results do not depend on the choice.
    Leaves the solvers iterated: linearise again before using them."""
    st = stream or Stream()
    s = st.handle
    sweeps = [f.sweep for f in frames]
    for sw in sweeps:
        sw.build_homographies(s)
    timings = {"sweep": {}, "tgv": {}}
    prev_ty = os.environ.get("PSW_SWEEP_TY")
    try:
        for ty in (16, 32):                       # NCC tile: 32 x 16 or 32 x 32 (halo = window radius)
            os.environ["PSW_SWEEP_TY"] = str(ty)
            for g in (1, 2, 4, 8):
                for wg in (1024, 2048, 4096):
                    ms = _time(lambda: sweep_batch(sweeps, fast, g, wg, s), st)
                    timings["sweep"][(ty, g, wg)] = ms
                    if log:
                        log(f"[autotune] sweep tile 32x{ty} groups {g} workgroups {wg}: {ms:.3f} ms")
    finally:
        if prev_ty is None:
            os.environ.pop("PSW_SWEEP_TY", None)
        else:
            os.environ["PSW_SWEEP_TY"] = prev_ty
    solvers = [f.tgv for f in frames]
    # per-iteration cost over solve_iters iterations per call: the default
    # shape (1) packs the frames the resident solver takes into resident
    # launches, whose set-up only pays off over many iterations
    for var in (0, 1, 2):
        ms = _time(lambda: tgv_iterate_batch(solvers, solve_iters, prm, var, s), st) / solve_iters
        timings["tgv"][var] = ms
        if log:
            log(f"[autotune] tgv variant {var}: {ms * 1e3:.1f} us per iteration")
    best_sweep = min(timings["sweep"], key=timings["sweep"].get)
    best_var = min(timings["tgv"], key=timings["tgv"].get)
    # the solve's chunks: hipGraph replay or direct launches (E* = 0: no stop)
    timings["graph"] = {}
    zero = [0.0] * len(solvers)
    for g in (False, True):
        ms = _time(lambda: tgv_solve_batch(solvers, prm, zero, solve_iters, 10, 1e-4, best_var, g, s), st) / solve_iters
        timings["graph"][g] = ms
        if log:
            log(f"[autotune] tgv solve chunks {'graph' if g else 'direct'}: {ms * 1e3:.1f} us per iteration")
    best_graph = min(timings["graph"], key=timings["graph"].get)
    return {"sweep_tile_h": best_sweep[0], "sweep_groups": best_sweep[1], "sweep_workgroups": best_sweep[2],
            "tgv_variant": best_var, "use_graph": bool(best_graph)}, timings
