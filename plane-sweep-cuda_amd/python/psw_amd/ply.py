"""Point-cloud outputs of the pipeline, as the reference's GUI writes them:
ASCII PLY files of the depthmap clouds and of the fused surface
(src/pclviewer.cpp:924-929, pcl::io::savePLYFileASCII), with the GUI's
depth colour map (RGBdepthmapColor, src/pclviewer.cpp:1085-1116) and its
display convention for the fused surface (x negated, colour from the x
position in the volume, src/pclviewer.cpp:1058-1074)."""
import numpy as np


def depth_color(gray):
    """RGBdepthmapColor for uint8 values (vectorised): blue, light blue,
    yellow, red, dark red bands of 32/64/64/64/32 values."""
    d = np.asarray(gray).astype(np.int32)
    r, g, b = np.zeros_like(d), np.zeros_like(d), np.zeros_like(d)
    m = d < 32
    b = np.where(m, 128 + d * 4, b)
    m2 = (d >= 32) & (d < 96)
    g = np.where(m2, (d - 32) * 4, g)
    b = np.where(m2, 255, b)
    m3 = (d >= 96) & (d < 160)
    r = np.where(m3, (d - 96) * 4, r)
    g = np.where(m3, 255, g)
    b = np.where(m3, 255 - (d - 96) * 4, b)
    m4 = (d >= 160) & (d < 224)
    r = np.where(m4, 255, r)
    g = np.where(m4, 255 - (d - 160) * 4, g)
    m5 = d >= 224
    r = np.where(m5, 255 - (d - 224) * 4, r)
    return np.stack([r, g, b], -1).astype(np.uint8)


def write_ply(path, xyz, rgb=None):
    """ASCII PLY with float x, y, z and optional uchar r, g, b per vertex."""
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    with open(path, "w") as f:
        f.write("ply\nformat ascii 1.0\n")
        f.write(f"element vertex {xyz.shape[0]}\n")
        f.write("property float x\nproperty float y\nproperty float z\n")
        if rgb is not None:
            f.write("property uchar red\nproperty uchar green\nproperty uchar blue\n")
        f.write("end_header\n")
        if rgb is None:
            for p in xyz:
                f.write(f"{p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
        else:
            rgb = np.asarray(rgb, np.uint8).reshape(-1, 3)
            for p, c in zip(xyz, rgb):
                f.write(f"{p[0]:.9g} {p[1]:.9g} {p[2]:.9g} {c[0]} {c[1]} {c[2]}\n")


def read_ply(path):
    """The vertices (and colours, if any) of an ASCII PLY written by write_ply."""
    with open(path) as f:
        assert f.readline().strip() == "ply"
        n, props = 0, []
        for line in f:
            line = line.strip()
            if line.startswith("element vertex"):
                n = int(line.split()[2])
            elif line.startswith("property"):
                props.append(line.split()[2])
            elif line == "end_header":
                break
        rows = np.array([f.readline().split() for _ in range(n)], dtype=np.float64).reshape(n, len(props))
    xyz = rows[:, :3].astype(np.float32)
    rgb = rows[:, 3:6].astype(np.uint8) if len(props) >= 6 else None
    return xyz, rgb


def fused_surface_cloud(fusion):
    """The GUI's fused-surface cloud: voxel centres with -1 < u < 0, x negated,
    coloured by the x position within the volume."""
    pts = fusion.surface_points()
    a = np.asarray(fusion.volume[:3], np.float32)
    size = np.asarray(fusion.volume[3:], np.float32) - a
    gray = np.clip((255 * (pts[:, 0] - a[0]) / size[0]), 0, 255).astype(np.uint8)
    out = pts.copy()
    out[:, 0] = -out[:, 0]
    return out, depth_color(gray)
