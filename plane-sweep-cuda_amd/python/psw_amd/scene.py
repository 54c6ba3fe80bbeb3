"""Seeded synthetic room scenes for the multiview-stereo path.

A room is an axis-aligned interior box (walls, floor, ceiling) holding random
axis-aligned boxes.  Every surface carries a solid value-noise texture
(intensity in [0, 1]); images are ray cast through a pinhole camera and get
Gaussian pixel noise.  Randomness comes from integer hashing (texture) and
numpy's PCG64 generator (layout, poses, noise), so a seed gives the same scene
on every machine.

Pose convention: X_cam = R X_world + t (the reference's default).  Pixel
coordinates of K are 0-based; one_based(K) gives the reference's 1-based K.
"""
import numpy as np

FOCAL = 480.0


def intrinsics(width=640, height=480, focal=FOCAL):
    """0-based pinhole K with the principal point at the image centre."""
    return np.array([[focal, 0.0, (width - 1) / 2.0], [0.0, focal, (height - 1) / 2.0], [0.0, 0.0, 1.0]],
                    np.float32)


def one_based(K):
    """K for the reference's 1-based pixel coordinates (reader.cpp:201-206)."""
    K1 = np.array(K, np.float32).copy()
    K1[0, 2] += 1.0
    K1[1, 2] += 1.0
    return K1


def rotation(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = np.cos(angle), np.sin(angle)
    C = 1 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


# ---- solid value-noise texture ------------------------------------------------

def _hash(ix, iy, iz, seed):
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (iy.astype(np.uint32) * np.uint32(0xD8163841)) \
        ^ (iz.astype(np.uint32) * np.uint32(0xCB1AB31F)) ^ np.uint32(seed * 0x9E3779B1 & 0xFFFFFFFF)
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return (h >> np.uint32(8)).astype(np.float64) / float(1 << 24)


def value_noise(p, seed):
    """Trilinear value noise with smoothstep weights at integer lattice points."""
    f = np.floor(p)
    i = f.astype(np.int64)
    t = p - f
    w = t * t * (3.0 - 2.0 * t)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        wx = w[:, 0] if dx else 1.0 - w[:, 0]
        for dy in (0, 1):
            wy = w[:, 1] if dy else 1.0 - w[:, 1]
            for dz in (0, 1):
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += wx * wy * wz * _hash(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return out


def texture(points, seed):
    """Four octaves of value noise, wavelengths ~25 cm down to ~2 cm, in [0, 1]."""
    acc = np.zeros(points.shape[0])
    total = 0.0
    for k, (freq, amp) in enumerate(((4.0, 0.40), (9.0, 0.30), (21.0, 0.20), (47.0, 0.10))):
        acc += amp * value_noise(points * freq + 17.0 * k, seed + 101 * k)
        total += amp
    v = acc / total
    return np.clip((v - 0.5) * 1.8 + 0.5, 0.0, 1.0)


# ---- room ----------------------------------------------------------------------

class Room:
    """Interior box [lo, hi] plus random boxes; y points down (camera convention)."""

    def __init__(self, seed=0, n_boxes=10, lo=(-3.0, -1.6, -4.0), hi=(3.0, 1.6, 6.0)):
        self.seed = int(seed)
        self.lo = np.array(lo, np.float64)
        self.hi = np.array(hi, np.float64)
        rng = np.random.Generator(np.random.PCG64(self.seed))
        boxes = []
        for _ in range(n_boxes):
            size = rng.uniform([0.3, 0.3, 0.3], [1.2, 1.4, 1.2])
            cx = rng.uniform(self.lo[0] + 0.4, self.hi[0] - 0.4)
            cz = rng.uniform(1.6, self.hi[2] - 0.8)
            # stand on the floor (floor at y = hi[1]) or hang from nowhere (floating)
            if rng.uniform() < 0.7:
                cy = self.hi[1] - size[1] / 2
            else:
                cy = rng.uniform(self.lo[1] + size[1] / 2, self.hi[1] - size[1] / 2)
            c = np.array([cx, cy, cz])
            boxes.append((c - size / 2, c + size / 2))
        self.boxes = boxes

    def raycast(self, origin, dirs):
        """Distance along each ray (dirs: n x 3, not normalised) to the first surface."""
        o = np.asarray(origin, np.float64)
        d = np.where(np.abs(dirs) < 1e-12, 1e-12, dirs)
        inv = 1.0 / d
        # exit from the interior box
        t1 = (self.lo - o) * inv
        t2 = (self.hi - o) * inv
        t_room = np.min(np.maximum(t1, t2), axis=1)
        t = t_room
        for lo, hi in self.boxes:
            a = (lo - o) * inv
            b = (hi - o) * inv
            tn = np.max(np.minimum(a, b), axis=1)
            tf = np.min(np.maximum(a, b), axis=1)
            hit = (tn <= tf) & (tn > 1e-6)
            t = np.where(hit & (tn < t), tn, t)
        return t

    def render(self, K, R, t, width, height, noise_sigma=0.01, noise_seed=None):
        """Image (float32 in [0, 1] plus Gaussian noise) and depth (z in the camera frame)."""
        K = np.asarray(K, np.float64)
        R = np.asarray(R, np.float64)
        t = np.asarray(t, np.float64).reshape(3)
        u, v = np.meshgrid(np.arange(width, dtype=np.float64), np.arange(height, dtype=np.float64))
        u, v = u.ravel(), v.ravel()
        # Element-wise products in a fixed order (no BLAS), so that the GPU
        # renderer (psw_scene_render) reproduces every ray bit for bit.
        iK = np.linalg.inv(K)
        rc = [iK[j, 0] * u + iK[j, 1] * v + iK[j, 2] for j in range(3)]    # z component = 1
        dirs = np.stack([rc[0] * R[0, j] + rc[1] * R[1, j] + rc[2] * R[2, j] for j in range(3)], axis=1)
        C = np.array([-(R[0, j] * t[0] + R[1, j] * t[1] + R[2, j] * t[2]) for j in range(3)])
        s = self.raycast(C, dirs)
        pts = C + s[:, None] * dirs
        img = texture(pts, self.seed).reshape(height, width)
        if noise_sigma > 0:
            rng = np.random.Generator(np.random.PCG64(self.seed * 7919 + 1 if noise_seed is None else noise_seed))
            img = img + rng.normal(0.0, noise_sigma, img.shape)
        depth = s.reshape(height, width)               # z of rays_cam is 1
        return img.astype(np.float32), depth.astype(np.float32)


def random_poses(n, seed, max_translation=0.05, max_angle_deg=2.0):
    """n poses near the identity: |t| <= max_translation (uniform in the ball),
    rotation about a uniformly drawn axis by an angle uniform in [0, max_angle]."""
    rng = np.random.Generator(np.random.PCG64(int(seed) + 1000003))
    Rs, ts = [], []
    for _ in range(n):
        axis = rng.normal(size=3)
        ang = np.deg2rad(rng.uniform(0.0, max_angle_deg))
        R = rotation(axis, ang)
        dvec = rng.normal(size=3)
        dvec /= np.linalg.norm(dvec)
        C = dvec * max_translation * rng.uniform() ** (1.0 / 3.0)   # camera centre
        Rs.append(R)
        ts.append(-R @ C)
    return np.array(Rs, np.float32), np.array(ts, np.float32)


def make_views(seed, width=640, height=480, nsrc=4, noise_sigma=0.01, n_boxes=10, scale=255.0, focal=None):
    """Reference view at the world origin plus nsrc nearby source views.

    Returns a dict with ref (H, W), src (N, H, W) in units of `scale` (the
    reference works in 0..255), depth of the reference view, K (0-based) and
    the poses.  The focal length defaults to 480 px scaled with the width, so
    smaller images see the same field of view."""
    room = Room(seed, n_boxes=n_boxes)
    K = intrinsics(width, height, FOCAL * width / 640.0 if focal is None else focal)
    Rref, tref = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    Rs, ts = random_poses(nsrc, seed)
    ref, depth = room.render(K, Rref, tref, width, height, noise_sigma, noise_seed=seed * 31 + 7)
    src = np.stack([room.render(K, Rs[i], ts[i], width, height, noise_sigma, noise_seed=seed * 31 + 8 + i)[0]
                    for i in range(nsrc)])
    return {"ref": ref * np.float32(scale), "src": src * np.float32(scale), "depth": depth, "K": K,
            "Rref": Rref, "tref": tref, "Rsrc": Rs, "tsrc": ts, "room": room}


def look_at(C, target, up=(0.0, -1.0, 0.0)):
    """Pose (R, t) of a camera at C looking at target, y down."""
    C = np.asarray(C, np.float64)
    f = np.asarray(target, np.float64) - C
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    R = np.stack([r, d, f])
    return R.astype(np.float32), (-R @ C).astype(np.float32)


def orbit_depthmaps(seed, n=50, width=640, height=480, radius=1.2, centre=(0.0, 0.2, 1.0), noise_sigma=0.0):
    """n depthmaps from cameras on a horizontal circle around `centre`, each
    looking outwards through the centre's opposite side (fusion input)."""
    room = Room(seed)
    K = intrinsics(width, height)
    c = np.asarray(centre, np.float64)
    Rs, ts, depths = [], [], []
    rng = np.random.Generator(np.random.PCG64(int(seed) + 77))
    for k in range(n):
        a = 2.0 * np.pi * k / n
        C = c + radius * np.array([np.cos(a), 0.0, np.sin(a)])
        target = c + 3.0 * np.array([np.cos(a), 0.15, np.sin(a)])
        R, t = look_at(C, target)
        _, dep = room.render(K, R, t, width, height, noise_sigma=0.0)
        if noise_sigma > 0:
            dep = dep + rng.normal(0.0, noise_sigma, dep.shape).astype(np.float32)
        Rs.append(R)
        ts.append(t)
        depths.append(dep)
    return {"K": K, "R": np.array(Rs), "t": np.array(ts), "depths": np.array(depths, np.float32), "room": room}


# ---- GPU renderer (psw_scene_render) -----------------------------------------------

def room_struct(room):
    """The psw_room C struct of a Room (layout, and the per-octave texture
    hash salts that _hash derives from the seed)."""
    from ._lib import PSW_SCENE_MAX_BOXES, PswRoom
    assert len(room.boxes) <= PSW_SCENE_MAX_BOXES
    r = PswRoom()
    for j in range(3):
        r.lo[j], r.hi[j] = float(room.lo[j]), float(room.hi[j])
    r.nboxes = len(room.boxes)
    for b, (lo, hi) in enumerate(room.boxes):
        for j in range(3):
            r.box_lo[b][j], r.box_hi[b][j] = float(lo[j]), float(hi[j])
    for k in range(4):
        r.octave_salt[k] = ((room.seed + 101 * k) * 0x9E3779B1) & 0xFFFFFFFF
    return r


def render_gpu(room, K, R, t, width, height, noise_sigma=0.01, noise_seed=0, scale=1.0, d_img=None, d_depth=None,
               stream=None):
    """Room.render on the GPU into device arrays (allocated when not given).
    Depth and the noise-free image equal Room.render's; the pixel noise is a
    counter-based Gaussian keyed by noise_seed (not numpy's PCG64 stream)."""
    import ctypes
    from ._lib import call
    from .device import DeviceArray
    iK = np.ascontiguousarray(np.linalg.inv(np.asarray(K, np.float64)))
    Rd = np.ascontiguousarray(np.asarray(R, np.float64))
    td = np.ascontiguousarray(np.asarray(t, np.float64).reshape(3))
    if d_img is None:
        d_img = DeviceArray((height, width))
    if d_depth is None:
        d_depth = DeviceArray((height, width))
    rs = room_struct(room)
    call("psw_scene_render", ctypes.byref(rs), iK.ctypes.data, Rd.ctypes.data, td.ctypes.data, int(width),
         int(height), float(noise_sigma), int(noise_seed) & 0xFFFFFFFFFFFFFFFF, float(scale),
         d_img.ptr if d_img is not False else None, d_depth.ptr if d_depth is not False else None, stream)
    return d_img, d_depth


def make_views_gpu(seed, width=640, height=480, nsrc=4, noise_sigma=0.01, n_boxes=10, scale=255.0, focal=None,
                   K=None, stream=None):
    """make_views rendered on the GPU: same room, K and poses; images stay in
    device memory (d_ref (H, W), d_src (N, H, W), d_depth (H, W)).  K may be
    given (0-based), e.g. to render a window of a larger camera."""
    from .device import DeviceArray
    room = Room(seed, n_boxes=n_boxes)
    if K is None:
        K = intrinsics(width, height, FOCAL * width / 640.0 if focal is None else focal)
    K = np.asarray(K, np.float32)
    Rref, tref = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    Rs, ts = random_poses(nsrc, seed)
    d_ref, d_depth = render_gpu(room, K, Rref, tref, width, height, noise_sigma, seed * 31 + 7, scale, stream=stream)
    d_src = DeviceArray((nsrc, height, width))
    for i in range(nsrc):
        render_gpu(room, K, Rs[i], ts[i], width, height, noise_sigma, seed * 31 + 8 + i, scale,
                   d_img=d_src.view(i), d_depth=False, stream=stream)
    return {"d_ref": d_ref, "d_src": d_src, "d_depth": d_depth, "K": K, "Rref": Rref, "tref": tref, "Rsrc": Rs,
            "tsrc": ts, "room": room}


def orbit_depthmaps_gpu(seed, n=50, width=640, height=480, radius=1.2, centre=(0.0, 0.2, 1.0), noise_sigma=0.0):
    """orbit_depthmaps with the depth ray cast on the GPU (bit-identical to the
    numpy renderer) and the same numpy depth noise: identical output, faster."""
    room = Room(seed)
    K = intrinsics(width, height)
    c = np.asarray(centre, np.float64)
    Rs, ts, depths = [], [], []
    rng = np.random.Generator(np.random.PCG64(int(seed) + 77))
    for k in range(n):
        a = 2.0 * np.pi * k / n
        C = c + radius * np.array([np.cos(a), 0.0, np.sin(a)])
        target = c + 3.0 * np.array([np.cos(a), 0.15, np.sin(a)])
        R, t = look_at(C, target)
        _, d_dep = render_gpu(room, K, R, t, width, height, 0.0, 0, d_img=False)
        dep = d_dep.to_host()
        if noise_sigma > 0:
            dep = dep + rng.normal(0.0, noise_sigma, dep.shape).astype(np.float32)
        Rs.append(R)
        ts.append(t)
        depths.append(dep)
    return {"K": K, "R": np.array(Rs), "t": np.array(ts), "depths": np.array(depths, np.float32), "room": room}
