"""Batched frames of uneven size: every batched result equals, frame by
frame and bit for bit, the single-frame entry point on that frame (sweep
outputs, TGV2 states after n iterations, solve stops and states with and
without hipGraph replay, on the launch sequence and on the resident solver,
whose batched form packs frames of one source count into one launch)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [  # small uneven frames: size, sources, planes
    {"seed": 11, "W": 96, "H": 72, "N": 3, "P": 24},
    {"seed": 12, "W": 130, "H": 50, "N": 2, "P": 9},
    {"seed": 13, "W": 64, "H": 64, "N": 4, "P": 31},
    {"seed": 14, "W": 160, "H": 120, "N": 2, "P": 16},
]
TGV_KW = dict(lambda_=1.0, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, beta=0.0, gamma=1.0)


def _frames():
    from psw_amd import batch
    return [batch.Frame(sp, znear=0.5, zfar=6.0, winsize=7) for sp in SPECS]


@pytest.mark.parametrize("fast,groups,ty", [(True, 0, "16"), (False, 1, "16"), (True, 3, "16"), (True, 0, "32"),
                                             (False, 2, "32")])
def test_sweep_batch_matches_single(gpu, monkeypatch, fast, groups, ty):
    from psw_amd import batch
    monkeypatch.setenv("PSW_SWEEP_TY", ty)
    frames = _frames()
    single = []
    for f in frames:
        f.sweep.build_homographies()
        f.sweep.run(fast=fast)
        single.append(f.sweep.results())
    for f in frames:
        for d in (f.sweep.d_arg, f.sweep.d_cmin, f.sweep.d_csec):
            d.fill(0)
    batch.sweep_batch([f.sweep for f in frames], fast=fast, groups=groups, workgroups=97)
    for f, ref in zip(frames, single):
        for a, b in zip(f.sweep.results(), ref):
            np.testing.assert_array_equal(a, b)


def test_sweep_batch_many_frames(gpu):
    """19 uneven frames of 15 shapes in one batched sweep (the persistent
    kernel's frame lookup over many frames): every frame equals its
    single-frame sweep."""
    from psw_amd import batch
    specs = [{"seed": 40 + i, "W": 48 + 8 * (i % 5), "H": 32 + 8 * (i % 3), "N": 2 + i % 3, "P": 5 + i % 7}
             for i in range(19)]
    frames = [batch.Frame(sp, znear=0.5, zfar=6.0, winsize=5) for sp in specs]
    single = []
    for f in frames:
        f.sweep.build_homographies()
        f.sweep.run(fast=True)
        single.append(f.sweep.results())
    for f in frames:
        for d in (f.sweep.d_arg, f.sweep.d_cmin, f.sweep.d_csec):
            d.fill(0)
    print(f"[case] sweep batch of {len(frames)} frames", flush=True)
    batch.sweep_batch([f.sweep for f in frames], fast=True, groups=0, workgroups=0)
    for f, ref in zip(frames, single):
        for a, b in zip(f.sweep.results(), ref):
            np.testing.assert_array_equal(a, b)


def _linearised(frames, prm):
    for f in frames:
        f.sweep.build_homographies()
        f.sweep.run(fast=True)
        f.linearise(prm)


@pytest.mark.parametrize("batch_mb,resident", [("192", "0"), ("3", "0"), ("192", "1")])
def test_tgv_iterate_batch_matches_single(gpu, monkeypatch, batch_mb, resident):
    """Also with the frames split into several cache-sized groups (3 MB), and
    with the resident solver (frames of one source count packed into one
    launch)."""
    from psw_amd import batch, stages
    monkeypatch.setenv("PSW_TGV_BATCH_MB", batch_mb)
    monkeypatch.setenv("PSW_TGV_RESIDENT", resident)
    prm = stages.tgv_params(**TGV_KW)
    a, b = _frames(), _frames()
    _linearised(a, prm)
    _linearised(b, prm)
    for f in a:
        f.tgv.iterate(23, prm)
    batch.tgv_iterate_batch([f.tgv for f in b], 23, prm)
    for fa, fb in zip(a, b):
        sa, sb = fa.tgv.get_state(), fb.tgv.get_state()
        for k in ("u", "u1x", "u1y", "px", "py", "qx", "qy", "qz", "r"):
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


@pytest.mark.parametrize("use_graph,batch_mb,resident", [(False, "192", "0"), (True, "192", "0"), (True, "3", "0"),
                                                         (False, "192", "1")])
def test_tgv_solve_batch_matches_single(gpu, monkeypatch, use_graph, batch_mb, resident):
    """Per-frame E* from a long run; the batched solve stops every frame at
    the same iteration with the same state as the single solve; replaying
    the captured graph twice, with eager work in between, repeats it.  With
    batch_mb 3 the frames are solved in several cache-sized groups; with the
    resident solver the frames of one source count share a launch."""
    from psw_amd import batch, stages
    monkeypatch.setenv("PSW_TGV_BATCH_MB", batch_mb)
    monkeypatch.setenv("PSW_TGV_RESIDENT", resident)
    from psw_amd.device import Stream
    from psw_amd.device import synchronize as device_synchronize
    st = Stream()
    s = st.handle
    prm = stages.tgv_params(**TGV_KW)
    a, b = _frames(), _frames()
    _linearised(a, prm)
    e_star = []
    for f in a:
        f.tgv.iterate(3000, prm)
        e_star.append(f.tgv.energy(prm))
    _linearised(a, prm)
    ref_iters, ref_states = [], []
    for f, e in zip(a, e_star):
        it, _ = f.tgv.solve(3000, 10, 1e-3, e, prm)
        ref_iters.append(it)
        ref_states.append(f.tgv.get_state())
    print("single-frame stops:", ref_iters)
    assert all(it < 3000 for it in ref_iters)
    for rep in range(2):
        _linearised(b, prm)          # on the null stream
        device_synchronize()         # the solve runs on a non-blocking stream
        iters, _ = batch.tgv_solve_batch([f.tgv for f in b], prm, e_star, 3000, 10, 1e-3, use_graph=use_graph, stream=s)
        print(f"batched stops (graph={use_graph}, replay {rep}):", list(iters))
        assert list(iters) == ref_iters
        for f, ref in zip(b, ref_states):
            sb = f.tgv.get_state()
            for k in ("u", "u1x", "u1y", "px", "py", "qx", "qy", "qz", "r"):
                np.testing.assert_array_equal(sb[k], ref[k], err_msg=k)
