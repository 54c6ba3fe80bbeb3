"""psw_sweep_fused_aggregate_noalloc (the plane groups' per-pixel keys in a
static device array, no allocation) against psw_sweep_fused_aggregate,
through the C ABI: bit-identical argmin / min / second-best on two calls in
a row, the second on a smaller input, in both arithmetic variants; plane
groups for a frame whose keys exceed the static array are rejected with
PSW_ERR_ARG before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PSW_ERR_ARG = 1


def _noalloc(sw, fast, groups, stream=None, size=None):
    from psw_amd._lib import lib
    out = [type(d)(d.shape, d.dtype) for d in (sw.d_arg, sw.d_cmin, sw.d_csec)]
    W, H = size if size else (sw.W, sw.H_img)
    rc = lib.psw_sweep_fused_aggregate_noalloc(sw.d_ref.ptr, sw.d_src.ptr, W, H, sw.N, sw.d_H.ptr, sw.P,
                                               sw.winsize, ctypes.c_float(sw.stdthresh), sw.coord_offset,
                                               int(bool(fast)), int(groups), out[0].ptr, out[1].ptr, out[2].ptr,
                                               stream)
    return rc, out


@pytest.mark.parametrize("fast", [True, False])
def test_noalloc_sweep_matches_allocating_entry(gpu, fast):
    from psw_amd import scene, stages
    from psw_amd._lib import lib
    cases = [(640, 480, 4, 64, 7, 0), (320, 240, 3, 40, 5, 3)]    # W, H, sources, planes, window, groups
    for (W, H, N, P, win, groups) in cases:
        print(f"[case] noalloc sweep {W}x{H} N={N} P={P} win={win} groups={groups} fast={fast}", flush=True)
        v = scene.make_views_gpu(77 + W, W, H, N)
        planes = stages.inverse_depth_planes(0.5, 5.0, P)
        sw = stages.SweepAggregate(v["d_ref"], v["d_src"], v["K"], v["Rref"], v["tref"], v["Rsrc"], v["tsrc"], planes,
                                   win, 1e-4, coord_offset=0)
        sw.build_homographies()
        sw.run(fast=fast)
        exp = sw.results()
        rc, out = _noalloc(sw, fast, groups)
        assert rc == 0, lib.psw_last_error()
        for a, b in zip([d.to_host() for d in out], exp):
            np.testing.assert_array_equal(a, b)
    # plane groups of a frame above 2^21 pixels (32 MiB of keys) are refused
    # before any launch: the size check runs first, so the small buffers
    # passed are never touched
    rc, _ = _noalloc(sw, True, 8, size=(2048, 1100))
    assert rc == PSW_ERR_ARG
    assert b"static" in lib.psw_last_error()
