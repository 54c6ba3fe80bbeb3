"""Learned TGV2 warm start: the BF16 U-Net of libpsw (psw_unet_*), its
training-data generation from the plain pipeline, training with Adam, and
inference into a TgvSolver.

Training pairs: a seeded room is rendered on the GPU (a window of a 640x480
camera, 8 source views), swept (aggregated NCC, 256 inverse-depth planes),
and TGV2 is linearised at the sweep depth and iterated to convergence from
the sweep start.  The network maps the sweep (reference image, argmin depth,
min and second-best cost) to the converged (u, u1, p, q).  At inference the
prediction becomes the solver's primal and dual start inside the same
linearisation; the solver then runs to the energy gap as usual, so a poor
prediction costs iterations, never accuracy.

    python -m psw_amd.warmstart train --out weights.npz      (on a GPU)
"""
import argparse
import ctypes
import os
import sys
import time

import numpy as np

from . import scene, stages
from ._lib import call, lib
from .device import DeviceArray, Stream

# loss weights of the 8 output channels: depth correction, u1 (2), p (2), q (3)
CHANNEL_WEIGHTS = np.array([4.0, 1.0, 1.0, 2.0, 2.0, 1.0, 1.0, 1.0], np.float32)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "weights",
                               "unet_warmstart.npz")


def param_count():
    return int(lib.psw_unet_param_count())


class UNet:
    """psw_unet handle sized for batch * width * height <= max_pixels."""

    def __init__(self, max_pixels, seed=1):
        h = ctypes.c_void_p()
        call("psw_unet_create", ctypes.byref(h), int(max_pixels))
        self.h = h
        self.max_pixels = int(max_pixels)
        self.nparams = param_count()
        if seed != 1:
            call("psw_unet_init", self.h, int(seed))

    # parameters ------------------------------------------------------------
    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        call("psw_unet_get_params", self.h, p.ctypes.data, p.size)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        call("psw_unet_set_params", self.h, p.ctypes.data, p.size)

    def get_grads(self):
        g = np.empty(self.nparams, np.float32)
        call("psw_unet_get_grads", self.h, g.ctypes.data, g.size)
        return g

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.get_params())

    def load(self, path):
        with np.load(path) as z:    # allow_pickle=False: plain arrays only
            self.set_params(z["params"])

    # data ------------------------------------------------------------------
    def features(self, slot, W, H, d_ref, d_argmin, d_cmin, d_csec, d_depths, nplanes, fallback, stream=None):
        call("psw_unet_features", self.h, int(slot), int(W), int(H), d_ref.ptr, d_argmin.ptr, d_cmin.ptr, d_csec.ptr,
             d_depths.ptr, int(nplanes), float(fallback), stream)

    def features_from_sweep(self, slot, sw, fallback, stream=None):
        self.features(slot, sw.W, sw.H_img, sw.d_ref, sw.d_arg, sw.d_cmin, sw.d_csec, sw.d_dep, sw.P, fallback, stream)

    def set_input(self, slot, W, H, d_x, d_depth, stream=None):
        call("psw_unet_set_input", self.h, int(slot), int(W), int(H), d_x.ptr, d_depth.ptr, stream)

    def targets(self, slot, W, H, d_state, stream=None):
        call("psw_unet_targets", self.h, int(slot), int(W), int(H), d_state.ptr, stream)

    def set_target(self, slot, W, H, d_t, stream=None):
        call("psw_unet_set_target", self.h, int(slot), int(W), int(H), d_t.ptr, stream)

    # compute ---------------------------------------------------------------
    def forward(self, batch, W, H, stream=None):
        call("psw_unet_forward", self.h, int(batch), int(W), int(H), stream)

    def output(self, slot, W, H, stream=None):
        d = DeviceArray((H, W, 8))
        call("psw_unet_output", self.h, int(slot), int(W), int(H), d.ptr, stream)
        return d

    def warm_start(self, slot, W, H, d_u, d_u1, d_p, d_q, stream=None):
        ptr = lambda d: d.ptr if d is not None else None   # noqa: E731
        call("psw_unet_warm_start", self.h, int(slot), int(W), int(H), ptr(d_u), ptr(d_u1), ptr(d_p), ptr(d_q),
             stream)

    def loss_grad(self, batch, W, H, weights=CHANNEL_WEIGHTS, want_loss=True, stream=None):
        w = np.ascontiguousarray(weights, np.float32)
        loss = ctypes.c_double()
        call("psw_unet_loss_grad", self.h, int(batch), int(W), int(H), w.ctypes.data,
             ctypes.byref(loss) if want_loss else None, stream)
        return loss.value if want_loss else None

    def adam_step(self, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        call("psw_unet_adam_step", self.h, float(lr), float(beta1), float(beta2), float(eps), int(step), stream)

    def __del__(self):
        try:
            lib.psw_unet_destroy(self.h)
        except Exception:
            pass


def forward_flops(W, H, batch=1):
    f = ctypes.c_double()
    call("psw_unet_flops", int(W), int(H), int(batch), ctypes.byref(f))
    return f.value


class WarmStarter:
    """Inference: sweep results of one view -> warm-started TgvSolver."""

    def __init__(self, W, H, weights=DEFAULT_WEIGHTS):
        self.W, self.H = int(W), int(H)
        self.net = UNet(self.W * self.H)
        if weights is not None:
            self.net.load(weights)
        n = self.W * self.H
        self.d_u = DeviceArray((self.H, self.W))
        self.d_u1 = DeviceArray((2, self.H, self.W))
        self.d_p = DeviceArray((2, self.H, self.W))
        self.d_q = DeviceArray((3, self.H, self.W))
        self.npix = n

    def predict(self, sw, fallback, stream=None):
        """Run the network on a SweepAggregate's results (after sw.run)."""
        self.net.features_from_sweep(0, sw, fallback, stream)
        self.net.forward(1, self.W, self.H, stream)
        self.net.warm_start(0, self.W, self.H, self.d_u, self.d_u1, self.d_p, self.d_q, stream)

    def start(self, tgv, stream=None):
        """Load the prediction into a solver already linearised at the sweep depth."""
        tgv.set_primal(self.d_u, 0.0, self.d_u1, stream)
        tgv.set_duals(self.d_p, self.d_q, None, stream)

    def solve(self, tgv, d_u0, prm, e_star, tol, max_iters, check_every, fallback_after=None, stream=None):
        """Solve from the prediction, with the plain start (u = u0, the rest 0)
        as fallback.  The prediction is kept only if its energy is below the
        plain start's; a solve from it that has not reached the gap after
        `fallback_after` iterations (default: max_iters) restarts from the
        plain start with the budget that is left.  tgv must be linearised at
        d_u0 (its state is the plain start).  Returns (iterations, energy,
        start) with start 'learned', 'plain' (rejected prediction) or
        'learned+plain' (restarted)."""
        e_plain = tgv.energy(prm, stream)
        self.start(tgv, stream)
        e_learned = tgv.energy(prm, stream)
        if not e_learned < e_plain:                  # NaN-safe: a NaN prediction is rejected
            tgv.set_primal(d_u0, 0.0, None, stream)
            tgv.set_duals(None, None, None, stream)
            it, e = tgv.solve(max_iters, check_every, tol, e_star, prm, stream)
            return it, e, "plain"
        first = max_iters if fallback_after is None else min(int(fallback_after), max_iters)
        it, e = tgv.solve(first, check_every, tol, e_star, prm, stream)
        if e_star <= 0 or e - e_star <= tol * e_star or first >= max_iters:
            return it, e, "learned"
        tgv.set_primal(d_u0, 0.0, None, stream)
        tgv.set_duals(None, None, None, stream)
        it2, e = tgv.solve(max_iters - it, check_every, tol, e_star, prm, stream)
        return it + it2, e, "learned+plain"


def judge(u, u_ref, weight):
    """Distance of a solve's stop state u from a reference stop state u_ref
    (the plain or cold start solved to the same gap), relative to |u_ref|:
    rms and max over the pixels the energy counts (weight 1) and over the
    ones it leaves out."""
    e = np.abs(np.asarray(u, np.float64) - u_ref) / np.maximum(np.abs(u_ref), 1e-12)
    keep = np.asarray(weight) > 0
    out = {}
    for name, sel in (("counted", keep), ("left_out", ~keep)):
        out[name] = {"pixels": int(sel.sum()),
                     "rms_rel": float(np.sqrt(np.mean(e[sel] ** 2))) if sel.any() else 0.0,
                     "max_rel": float(e[sel].max()) if sel.any() else 0.0,
                     "median_rel": float(np.median(e[sel])) if sel.any() else 0.0}
    return out


# ---- training data ------------------------------------------------------------------

FULL_W, FULL_H = 640, 480
NSRC, NPLANES, WIN, ZNEAR, ZFAR, STD_T = 8, 256, 9, 0.3, 8.0, 1e-4
TGV = dict(lambda_=1.0, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, beta=0.0, gamma=1.0)


class Dataset:
    """Sweep results and converged TGV2 states of many scene windows, kept in
    device memory (48 bytes per pixel)."""

    def __init__(self, crop_w, crop_h, tgv_iters=40000, stream=None, log=None):
        self.cw, self.ch = int(crop_w), int(crop_h)
        self.tgv_iters = int(tgv_iters)
        self.items = []
        self.stream = stream
        self.log = log or (lambda *a: None)
        self.planes = stages.inverse_depth_planes(ZNEAR, ZFAR, NPLANES)
        self.tgv = stages.TgvSolver(self.cw, self.ch, NSRC)
        self.prm = stages.tgv_params(**TGV)

    def add_scene(self, seed, x0, y0):
        s = self.stream
        K = scene.intrinsics(FULL_W, FULL_H)
        K[0, 2] -= x0
        K[1, 2] -= y0
        v = scene.make_views_gpu(seed, self.cw, self.ch, NSRC, K=K, stream=s)
        sw = stages.SweepAggregate(v["d_ref"], v["d_src"], K, v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], self.planes,
                                   WIN, STD_T, coord_offset=0)
        sw.build_homographies(s)
        sw.run(fast=True, stream=s)
        d_s = sw.depth(ZFAR, s)
        Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
        t = self.tgv
        t.set_images(sw.d_ref, sw.d_src, TGV["beta"], TGV["gamma"], s)
        t.set_u(d_s, stream=s)
        t.linearise(scene.one_based(K), Rrel, trel, s)
        t.iterate(self.tgv_iters, self.prm, s)
        st = t.state_device(s)
        d_state = DeviceArray((8, self.ch, self.cw))
        call("psw_memcpy_dtod", d_state.ptr, st.ptr, d_state.nbytes, s)
        self.items.append({"ref": sw.d_ref, "arg": sw.d_arg, "cmin": sw.d_cmin, "csec": sw.d_csec,
                           "state": d_state, "seed": seed})
        self._dep = sw.d_dep

    def load_slot(self, net, slot, i):
        it = self.items[i]
        s = self.stream
        net.features(slot, self.cw, self.ch, it["ref"], it["arg"], it["cmin"], it["csec"], self._dep, NPLANES, ZFAR, s)
        net.targets(slot, self.cw, self.ch, it["state"], s)


def train(n_scenes=1000, steps=20000, batch=8, crop=(256, 256), lr=1e-3, tgv_iters=20000, seed0=100000, log=print,
          out=None, init_seed=1):
    st = Stream()
    s = st.handle
    cw, ch = crop
    ds = Dataset(cw, ch, tgv_iters, s)
    rng = np.random.default_rng(seed0)
    t0 = time.time()
    for i in range(n_scenes):
        x0 = int(rng.integers(0, (FULL_W - cw) // 8 + 1)) * 8
        y0 = int(rng.integers(0, (FULL_H - ch) // 8 + 1)) * 8
        ds.add_scene(seed0 + i, x0, y0)
        if (i + 1) % 100 == 0:
            st.synchronize()
            log(f"[data] {i + 1}/{n_scenes} scenes in {time.time() - t0:.1f} s")
    st.synchronize()
    log(f"[data] {n_scenes} scenes, {cw}x{ch}, {tgv_iters} TGV2 iterations each: {time.time() - t0:.1f} s")
    net = UNet(batch * cw * ch, seed=init_seed)
    n_val = max(1, n_scenes // 10)
    n_tr = n_scenes - n_val
    t0 = time.time()
    hist = []
    for step in range(1, steps + 1):
        idx = rng.integers(0, n_tr, batch)
        for b, i in enumerate(idx):
            ds.load_slot(net, b, int(i))
        want = step % 1000 == 0 or step == 1
        loss = net.loss_grad(batch, cw, ch, want_loss=want, stream=s)
        cur_lr = lr * 0.5 * (1 + np.cos(np.pi * (step - 1) / steps))
        net.adam_step(cur_lr, step, stream=s)
        if want:
            hist.append((step, loss))
            log(f"[train] step {step} loss {loss:.5f} lr {cur_lr:.2e} ({time.time() - t0:.1f} s)")
    # validation loss on held-out scenes
    vl = []
    for j in range(0, n_val, batch):
        ids = list(range(n_tr + j, min(n_scenes, n_tr + j + batch)))
        for b, i in enumerate(ids):
            ds.load_slot(net, b, i)
        vl.append(net.loss_grad(len(ids), cw, ch, stream=s) * len(ids))
    val = float(np.sum(vl) / n_val)
    log(f"[train] validation loss {val:.5f} on {n_val} held-out scenes")
    if out:
        net.save(out)
        log(f"[train] weights written to {out}")
    return net, hist, val


def ceiling(seed=2024, weights=DEFAULT_WEIGHTS, estar_iters=40000, qmax=1.0, tol=1e-4, max_iters=50000,
            check_every=10, log=print):
    """How far a warm start can shorten the solve to the energy gap, on one
    640x480 scene: iterations from the plain start, from the converged state
    of the same problem (E*'s state, after estar_iters plain iterations) with
    all of it, without the data dual r, with r set to the sign of the
    residual, with relative noise on u, and from the network's prediction.
    Returns {start: iterations}."""
    from .device import Stream
    st = Stream()
    s = st.handle
    W, H = FULL_W, FULL_H
    v = scene.make_views_gpu(seed, W, H, NSRC, stream=s)
    planes = stages.inverse_depth_planes(ZNEAR, ZFAR, NPLANES)
    sw = stages.SweepAggregate(v["d_ref"], v["d_src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes,
                               WIN, STD_T, coord_offset=0)
    sw.build_homographies(s)
    sw.run(fast=True, stream=s)
    d_s = sw.depth(ZFAR, s)
    Rrel, trel = stages.relative_all(v["Rref"], v["tref"], v["Rsrc"], v["tsrc"])
    K1 = scene.one_based(v["K"])
    prm = stages.tgv_params(**TGV)
    t = stages.TgvSolver(W, H, NSRC)
    t.set_images(sw.d_ref, sw.d_src, TGV["beta"], TGV["gamma"], s)

    def relinearise():
        t.set_u(d_s, stream=s)
        t.linearise(K1, Rrel, trel, s)
        t.set_energy_mask(qmax, prm, s, count=False)

    relinearise()
    t.iterate(estar_iters, prm, s)
    e_star = t.energy(prm, s)
    S = t.get_state(s)
    st.synchronize()
    rho = S["It"] + (S["u"] - S["u0"])[None] * S["Iu"]

    def up(a):
        return DeviceArray.from_host(np.ascontiguousarray(a, np.float32))

    out = {}

    def run(label, u, u1, p, q, r):
        relinearise()
        keep = [up(u), up(u1), up(p), up(q), up(r) if r is not None else None]
        t.set_primal(keep[0], 0.0, keep[1], s)
        t.set_duals(keep[2], keep[3], keep[4], s)
        it, e = t.solve(max_iters, check_every, tol, e_star, prm, s)
        st.synchronize()
        out[label] = int(it)
        log(f"[ceiling] {label:44s} {it:6d} iterations (gap {(e - e_star) / e_star:+.2e})")

    zero2, zero3 = np.zeros((2, H, W), np.float32), np.zeros((3, H, W), np.float32)
    u1s, ps = np.stack([S["u1x"], S["u1y"]]), np.stack([S["px"], S["py"]])
    qs = np.stack([S["qx"], S["qy"], S["qz"]])
    run("plain start (u = u0, the rest 0)", S["u0"], zero2, zero2, zero3, None)
    run("converged state, r included", S["u"], u1s, ps, qs, S["r"])
    run("converged (u, u1, p, q), r = 0", S["u"], u1s, ps, qs, None)
    run("converged (u, u1, p, q), r = sign(residual)", S["u"], u1s, ps, qs, np.sign(rho))
    rng = np.random.default_rng(seed)
    for eps in (1e-4, 1e-3, 1e-2):
        un = (S["u"] * (1.0 + eps * rng.standard_normal(S["u"].shape))).astype(np.float32)
        run(f"converged, r included, u x (1 + {eps:g} noise)", un, u1s, ps, qs, S["r"])
        rn = S["It"] + (un - S["u0"])[None] * S["Iu"]
        run(f"converged, r = sign(residual), u x (1 + {eps:g} noise)", un, u1s, ps, qs, np.sign(rn))
    if weights is not None:
        ws = WarmStarter(W, H, weights)
        ws.predict(sw, ZFAR, s)
        uL, u1L, pL, qL = (ws.d_u.to_host(s), ws.d_u1.to_host(s), ws.d_p.to_host(s), ws.d_q.to_host(s))
        run("network (u, u1, p, q), r = 0", uL, u1L, pL, qL, None)
        rL = S["It"] + (uL - S["u0"])[None] * S["Iu"]
        run("network (u, u1, p, q), r = sign(residual)", uL, u1L, pL, qL, np.sign(rL))
        run("converged u, network (u1, p, q), r = 0", S["u"], u1L, pL, qL, None)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--scenes", type=int, default=1000)
    t.add_argument("--steps", type=int, default=20000)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--crop", type=int, default=256)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--tgv-iters", type=int, default=20000)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    c = sub.add_parser("ceiling", help="iterations to the gap from exact, perturbed and predicted starts")
    c.add_argument("--seed", type=int, default=2024)
    c.add_argument("--weights", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    if a.cmd == "ceiling":
        import json
        print(json.dumps(ceiling(a.seed, a.weights, log=lambda *x: print(*x, flush=True))), flush=True)
        return 0
    if a.cmd == "train":
        train(a.scenes, a.steps, a.batch, (a.crop, a.crop), a.lr, a.tgv_iters, out=a.out,
              log=lambda *x: print(*x, flush=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
