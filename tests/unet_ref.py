"""float64 numpy references for the U-Net layers (NHWC), used by the GPU
tests; pinned on the CPU by tests/test_unet_ref.py (brute force and adjoint
identities)."""
import numpy as np


def bf16_bits(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_val(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def conv_ref(x, w, b=None):
    """x (B,H,W,C) float64, w (9,C,Co): y[p] = sum_tap x[p + off(tap)] w[tap] (zero padding)."""
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2, W + 2, C))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((B, H, W, w.shape[2]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        y += xp[:, ky:ky + H, kx:kx + W] @ w[tap]
    return y if b is None else y + b


def dgrad_ref(dz, w):
    B, H, W, _ = dz.shape
    dp = np.zeros((B, H + 2, W + 2, dz.shape[3]))
    dp[:, 1:-1, 1:-1] = dz
    dx = np.zeros((B, H, W, w.shape[1]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        dx += dp[:, 2 - ky:2 - ky + H, 2 - kx:2 - kx + W] @ w[tap].T
    return dx


def wgrad_ref(x, dz):
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2, W + 2, C))
    xp[:, 1:-1, 1:-1] = x
    dw = np.zeros((9, C, dz.shape[3]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        dw[tap] = np.einsum("bhwc,bhwo->co", xp[:, ky:ky + H, kx:kx + W], dz)
    return dw, dz.sum(axis=(0, 1, 2))


# ---- the whole network (same layer table and parameter layout as libpsw) -----------

LAYERS = [  # ca, cb, cout, level
    (8, 0, 32, 0), (32, 0, 32, 0), (32, 0, 48, 1), (48, 0, 48, 1), (48, 0, 64, 2), (64, 0, 64, 2), (64, 0, 96, 3),
    (96, 0, 96, 3), (64, 96, 64, 2), (64, 0, 64, 2), (48, 64, 48, 1), (48, 0, 48, 1), (32, 48, 32, 0), (32, 0, 32, 0),
]


def param_blocks():
    """(start, end) of every W and bias block in the flat parameter vector."""
    out, o = [], 0
    for ca, cb, co, _ in LAYERS:
        out.append((o, o + 9 * (ca + cb) * co))
        o += 9 * (ca + cb) * co
        out.append((o, o + co))
        o += co
    out.append((o, o + 32 * 8))
    o += 32 * 8
    out.append((o, o + 8))
    return out


def unet_forward(theta, x, round_bf16=True):
    """Forward pass in float64.  round_bf16: round weights, the input and
    every stored activation to bf16 as the GPU does (head in full precision)."""
    rb = (lambda a: bf16_val(bf16_bits(a)).astype(np.float64)) if round_bf16 else (lambda a: a)
    blocks = param_blocks()
    act = [rb(np.asarray(x, np.float64))]
    pools = []

    def pool(a):
        return rb(0.25 * ((a[:, 0::2, 0::2] + a[:, 0::2, 1::2]) + (a[:, 1::2, 0::2] + a[:, 1::2, 1::2])))

    src_a = [0, 1, -1, 3, -2, 5, -3, 7, 6, 9, 4, 11, 2, 13]
    src_b = [None] * 8 + [8, None, 10, None, 12, None]
    for L, (ca, cb, co, _) in enumerate(LAYERS):
        a = act[src_a[L]] if src_a[L] >= 0 else pools[-src_a[L] - 1]
        inp = a if src_b[L] is None else np.concatenate([a, up2(act[src_b[L]])], axis=3)
        (w0, w1), (b0, b1) = blocks[2 * L], blocks[2 * L + 1]
        w = rb(theta[w0:w1].astype(np.float64)).reshape(9, ca + cb, co)
        act.append(rb(np.maximum(conv_ref(inp, w, theta[b0:b1].astype(np.float64)), 0.0)))
        if L in (1, 3, 5):
            pools.append(pool(act[-1]))
    (h0, h1), (c0, c1) = blocks[-2], blocks[-1]
    return act[-1] @ theta[h0:h1].astype(np.float64).reshape(32, 8) + theta[c0:c1]


def huber_loss(out, tgt, wc):
    d = out - tgt
    ad = np.abs(d)
    return float(np.sum(np.where(ad <= 1, 0.5 * d * d, ad - 0.5) * np.asarray(wc, np.float64)) / (d.size // 8))


def unet_grad(theta, x, tgt, wc, round_bf16=True):
    """Gradient of huber_loss(unet_forward(theta, x), tgt) with the GPU's
    rounding points: bf16 weights, activations and ReLU-masked output
    gradients dZ (round_bf16=False: none); float64 arithmetic elsewhere.
    Returns the flat gradient."""
    rb = (lambda a: bf16_val(bf16_bits(a)).astype(np.float64)) if round_bf16 else (lambda a: np.asarray(a, np.float64))
    blocks = param_blocks()
    th = np.asarray(theta, np.float64)
    act = [rb(np.asarray(x, np.float64))]
    pools = []
    inputs = []
    src_a = [0, 1, -1, 3, -2, 5, -3, 7, 6, 9, 4, 11, 2, 13]
    src_b = [None] * 8 + [8, None, 10, None, 12, None]
    for L, (ca, cb, co, _) in enumerate(LAYERS):
        a = act[src_a[L]] if src_a[L] >= 0 else pools[-src_a[L] - 1]
        inp = a if src_b[L] is None else np.concatenate([a, up2(act[src_b[L]])], axis=3)
        inputs.append(inp)
        (w0, w1), (b0, b1) = blocks[2 * L], blocks[2 * L + 1]
        act.append(rb(np.maximum(conv_ref(inp, rb(th[w0:w1]).reshape(9, ca + cb, co), th[b0:b1]), 0.0)))
        if L in (1, 3, 5):
            y = act[-1]
            pools.append(rb(0.25 * ((y[:, 0::2, 0::2] + y[:, 0::2, 1::2]) + (y[:, 1::2, 0::2] + y[:, 1::2, 1::2]))))
    (h0, h1), (c0, c1) = blocks[-2], blocks[-1]
    Wh = th[h0:h1].reshape(32, 8)
    out = act[-1] @ Wh + th[c0:c1]
    n = out.size // 8
    gout = np.clip(out - tgt, -1, 1) * np.asarray(wc, np.float64) / n
    grad = np.zeros_like(th)
    grad[h0:h1] = np.einsum("bhwc,bhwo->co", act[-1], gout).ravel()
    grad[c0:c1] = gout.sum(axis=(0, 1, 2))
    gact = [None] * 15
    gpool = [None] * 3
    gact[14] = gout @ Wh.T
    for L in range(13, -1, -1):
        ca, cb, co, _ = LAYERS[L]
        dz = rb(np.where(act[L + 1] > 0, gact[L + 1], 0.0))
        dw, db = wgrad_ref(inputs[L], dz)
        (w0, w1), (b0, b1) = blocks[2 * L], blocks[2 * L + 1]
        grad[w0:w1] = dw.ravel()
        grad[b0:b1] = db
        if L == 0:
            break
        dx = dgrad_ref(dz, rb(th[w0:w1]).reshape(9, ca + cb, co))
        if cb:
            gact[src_a[L]] = dx[..., :ca]
            g = dx[..., ca:]
            gact[src_b[L]] = (g[:, 0::2, 0::2] + g[:, 0::2, 1::2]) + (g[:, 1::2, 0::2] + g[:, 1::2, 1::2])
        elif src_a[L] >= 0:
            gact[src_a[L]] = dx
        else:
            pl = -src_a[L] - 1
            gpool[pl] = dx
            gact[2 * pl + 2] = gact[2 * pl + 2] + 0.25 * up2(dx)
    return grad
