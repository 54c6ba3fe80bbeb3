"""Outputs: the fused-surface point extraction (GPU, ordered compaction)
against numpy on the same grid, the GUI's depth colour map, and PLY files."""
import numpy as np
import pytest

from psw_amd import ply


def test_depth_color_bands():
    """RGBdepthmapColor (src/pclviewer.cpp:1085-1116) at the band edges."""
    cases = {0: (0, 0, 128), 31: (0, 0, 252), 32: (0, 0, 255), 95: (0, 252, 255), 96: (0, 255, 255),
             159: (252, 255, 3), 160: (255, 255, 0), 223: (255, 3, 0), 224: (255, 0, 0), 255: (131, 0, 0)}
    got = ply.depth_color(np.array(list(cases), np.uint8))
    np.testing.assert_array_equal(got, np.array(list(cases.values()), np.uint8))


def test_ply_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    xyz = rng.standard_normal((50, 3)).astype(np.float32)
    rgb = rng.integers(0, 256, (50, 3)).astype(np.uint8)
    p = tmp_path / "c.ply"
    ply.write_ply(p, xyz, rgb)
    x2, c2 = ply.read_ply(p)
    np.testing.assert_array_equal(x2, xyz)
    np.testing.assert_array_equal(c2, rgb)
    ply.write_ply(p, xyz)
    x3, c3 = ply.read_ply(p)
    np.testing.assert_array_equal(x3, xyz)
    assert c3 is None


@pytest.mark.gpu
def test_surface_points_match_numpy(gpu):
    from psw_amd import stages
    gw, gh, gd = 40, 30, 23
    vol = np.array([-3.0, -1.6, -4.0, 3.0, 1.6, 6.0], np.float32)
    f = stages.Fusion(gw, gh, gd, vol)
    rng = np.random.default_rng(3)
    u = rng.uniform(-1.5, 0.5, (gd, gh, gw)).astype(np.float32)
    u[u > 0.3] = -1.0          # exactly -1: not on the surface
    f.upload(u=u)
    pts = f.surface_points()
    z, y, x = np.nonzero((u < 0) & (u > -1))      # C order = the reference's z, y, x loops
    fr = [np.array([(c + 0.5) / n for c in idx], np.float64).astype(np.float32) for idx, n in
          ((x, gw), (y, gh), (z, gd))]
    a, size = vol[:3], vol[3:] - vol[:3]
    exp = np.stack([a[k] + size[k] * fr[k] for k in range(3)], 1).astype(np.float32)
    assert pts.shape == exp.shape and f.surface_count() == exp.shape[0]
    np.testing.assert_array_equal(pts, exp)
    cloud, rgb = ply.fused_surface_cloud(f)
    np.testing.assert_array_equal(cloud[:, 0], -exp[:, 0])
    assert rgb.shape == exp.shape
