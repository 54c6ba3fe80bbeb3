"""Writes tests/golden/real_icl_nuim.npz, a small real-image fixture for the
sweep and TGV2 parity tests, from the data files the reference ships:

  src/PlaneSweep/im0.png .. im9.png   640x480 RGB8 frames (ICL-NUIM)
  src/PlaneSweep/Calibration.txt      intrinsics and T_world_frame poses

Processing (this script's own, stated here):
  - PNG decoding (PIL);
  - grayscale with the reference's weights 0.2989 R + 0.5870 G + 0.1140 B
    (src/inc/defines.h:19-21, rgb2gray at src/pclviewer.cpp:870-885);
  - 2x2 box downsampling to 320x240 and rounding to integers 0..255 (uint8),
    so that the fixture stays small;
  - K as the reference builds it from the calibration (src/pclviewer.cpp:
    347-349, the ATAN distortion parameter in the skew slot), scaled by 1/2
    for the downsampled frames (pixel centres: x' = (x + 0.5) / 2 - 0.5);
  - poses as given (the "T_world_frame" rows), which the reference's GUI
    feeds to its default relative-pose method (src/pclviewer.cpp:355-405,
    src/planesweep.cpp:541-543): with it the aggregated NCC cost over the
    four sources averages -0.92, with the alternative method -0.36.
Only the reference frame im0 and sources im1..im4 are kept.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_real_fixture.py
"""
import os
import numpy as np

SRC = "/root/reference/src/PlaneSweep"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "real_icl_nuim.npz")


def read_png_rgb8(path):
    from PIL import Image
    return np.asarray(Image.open(path).convert("RGB"))


def read_calibration(path):
    lines = [ln.strip() for ln in open(path) if ln.strip()]
    fx, fy, cx, cy, omega = map(float, lines[0].split())
    poses = {}
    for i, ln in enumerate(lines):
        if ln.endswith(".png") and ln.startswith("im"):
            j = lines.index("Transform T_world_frame is", i)
            M = np.array([[float(v) for v in lines[j + 1 + r].split()] for r in range(3)])
            poses[ln] = (M[:, :3], M[:, 3])
    return (fx, fy, cx, cy, omega), poses


def main():
    (fx, fy, cx, cy, omega), poses = read_calibration(os.path.join(SRC, "Calibration.txt"))
    K = np.array([[fx * 640, (1 - omega) * 640, cx * 640], [0, fy * 480, cy * 480], [0, 0, 1]], np.float64)
    S = np.array([[0.5, 0, -0.25], [0, 0.5, -0.25], [0, 0, 1]])       # 2x downsampling of pixel coordinates
    Kd = (S @ K).astype(np.float32)
    grays, Rs, ts = [], [], []
    for k in range(5):
        name = f"im{k}.png"
        rgb = read_png_rgb8(os.path.join(SRC, name)).astype(np.float64)
        g = 0.2989 * rgb[..., 0] + 0.5870 * rgb[..., 1] + 0.1140 * rgb[..., 2]
        g = 0.25 * (g[0::2, 0::2] + g[0::2, 1::2] + g[1::2, 0::2] + g[1::2, 1::2])
        grays.append(np.clip(np.rint(g), 0, 255).astype(np.uint8))
        R, t = poses[name]
        Rs.append(R)
        ts.append(t)
    np.savez_compressed(OUT, images=np.stack(grays), K=Kd, R=np.array(Rs, np.float32), t=np.array(ts, np.float32))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
