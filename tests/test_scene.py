"""Seeded synthetic scene generator (numpy, CPU) and its GPU renderer.

CPU: per-seed digests of the generated arrays are committed below, so a
change to the generator is caught; ray-cast depth is checked against the
analytic distance to the room's walls.  GPU: psw_scene_render reproduces the
numpy renderer's depth and noise-free image bit for bit, and its pixel noise
has the requested mean and standard deviation."""
import hashlib

import numpy as np
import pytest

from psw_amd import scene

# sha256 of (ref, src, depth) bytes for make_views(seed, 32, 24, nsrc=2)
DIGESTS = {
    1: "e01f746e1882d83b10f82070f16316159fbb155c7909b3aeb209f4f0d33cc0da",
    7: "46abe2ad141762280dcc75561a9b7d7300d2ed8ff1839200ff364ee99c368368",
}


def _digest(v):
    h = hashlib.sha256()
    for k in ("ref", "src", "depth"):
        h.update(np.ascontiguousarray(v[k]).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("seed", sorted(DIGESTS))
def test_scene_digest(seed):
    v = scene.make_views(seed=seed, width=32, height=24, nsrc=2)
    assert _digest(v) == DIGESTS[seed]


def test_depth_matches_analytic_walls():
    """Empty room seen from the origin with the identity pose: the depth of
    pixel (u, v) is the distance along the ray (x, y, 1) to the nearest wall,
    min over axes of (wall - 0) / ray component."""
    room = scene.Room(seed=3, n_boxes=0)
    W, H = 40, 30
    K = scene.intrinsics(W, H, 30.0)
    _, depth = room.render(K, np.eye(3), np.zeros(3), W, H, noise_sigma=0.0)
    u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    rx, ry = (u - K[0, 2]) / K[0, 0], (v - K[1, 2]) / K[1, 1]
    tx = np.where(rx > 0, room.hi[0] / np.maximum(rx, 1e-300), np.where(rx < 0, room.lo[0] / np.minimum(rx, -1e-300),
                                                                         np.inf))
    ty = np.where(ry > 0, room.hi[1] / np.maximum(ry, 1e-300), np.where(ry < 0, room.lo[1] / np.minimum(ry, -1e-300),
                                                                         np.inf))
    expect = np.minimum(np.minimum(tx, ty), room.hi[2])
    np.testing.assert_allclose(depth, expect, rtol=1e-6)
    assert np.isclose(depth[H // 2, W // 2], room.hi[2])


@pytest.mark.gpu
def test_gpu_renderer_matches_numpy(gpu):
    v = scene.make_views(seed=12, width=96, height=72, nsrc=2, noise_sigma=0.0)
    g = scene.make_views_gpu(seed=12, width=96, height=72, nsrc=2, noise_sigma=0.0)
    np.testing.assert_array_equal(g["d_depth"].to_host(), v["depth"])
    np.testing.assert_array_equal(g["d_ref"].to_host(), v["ref"])
    np.testing.assert_array_equal(g["d_src"].to_host(), v["src"])


@pytest.mark.gpu
def test_gpu_renderer_noise(gpu):
    room = scene.Room(seed=4)
    K = scene.intrinsics(256, 192, 192.0)
    R, t = np.eye(3), np.zeros(3)
    clean, _ = scene.render_gpu(room, K, R, t, 256, 192, 0.0, 0)
    a, _ = scene.render_gpu(room, K, R, t, 256, 192, 0.05, 99)
    b, _ = scene.render_gpu(room, K, R, t, 256, 192, 0.05, 99)
    c, _ = scene.render_gpu(room, K, R, t, 256, 192, 0.05, 100)
    n = a.to_host() - clean.to_host()
    assert abs(n.mean()) < 2e-3 and abs(n.std() - 0.05) < 2e-3
    np.testing.assert_array_equal(a.to_host(), b.to_host())
    assert not np.array_equal(a.to_host(), c.to_host())


@pytest.mark.gpu
def test_gpu_orbit_depthmaps_identical(gpu):
    a = scene.orbit_depthmaps(seed=5, n=3, width=48, height=36, noise_sigma=0.01)
    b = scene.orbit_depthmaps_gpu(seed=5, n=3, width=48, height=36, noise_sigma=0.01)
    for k in ("K", "R", "t", "depths"):
        np.testing.assert_array_equal(a[k], b[k])
