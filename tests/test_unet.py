"""Learned warm start: the BF16 MFMA U-Net kernels against float64 numpy
references computed from the same bf16-rounded operands, finite-difference
checks of the network gradient, and a training run whose loss must fall.

Bars: convolutions (forward, data gradient, weight gradient) within
1e-5 of sum |terms| (FP32 accumulation of exact bf16 products); ReLU bf16
outputs within half a bf16 ulp plus that; network gradient within 15 % of
central differences along random directions (weights are rounded to bf16, so
the loss is only piecewise smooth at the 2^-9 scale)."""
import numpy as np
import pytest

from unet_ref import bf16_bits, bf16_val, conv_ref, dgrad_ref, up2, wgrad_ref

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    from psw_amd.device import DeviceArray
    return DeviceArray.from_host(a if dtype is None else np.ascontiguousarray(a, dtype))


CONV_CASES = [  # ca, cb, cout, B, H, W
    (8, 0, 32, 1, 12, 20),
    (32, 48, 32, 2, 8, 16),
    (64, 96, 64, 1, 8, 8),
    (48, 0, 96, 1, 6, 10),
    (32, 0, 48, 3, 5, 7),
]


@pytest.mark.parametrize("ca,cb,cout,B,H,W", CONV_CASES)
def test_conv3x3_forward(gpu, ca, cb, cout, B, H, W):
    from psw_amd._lib import call
    from psw_amd.device import DeviceArray
    rng = np.random.default_rng(ca * 1000 + cout)
    xa = bf16_bits(rng.standard_normal((B, H, W, ca)))
    xb = bf16_bits(rng.standard_normal((B, H // 2, W // 2, cb))) if cb else None
    w = rng.standard_normal((9, ca + cb, cout)).astype(np.float32) * 0.2
    b = rng.standard_normal(cout).astype(np.float32)
    x = bf16_val(xa).astype(np.float64)
    if cb:
        x = np.concatenate([x, up2(bf16_val(xb).astype(np.float64))], axis=3)
    wr = bf16_val(bf16_bits(w)).astype(np.float64)
    ref = conv_ref(x, wr, b.astype(np.float64))
    mag = conv_ref(np.abs(x), np.abs(wr)) + np.abs(b)
    d_xa = _dev(xa)
    d_xb = _dev(xb) if cb else None
    d_w, d_b = _dev(w), _dev(b)
    yf = DeviceArray((B, H, W, cout))
    call("psw_conv3x3_bf16", d_xa.ptr, ca, d_xb.ptr if cb else None, cb, d_w.ptr, d_b.ptr, cout, B, W, H, 0, yf.ptr,
         None)
    got = yf.to_host()
    err = np.abs(got - ref)
    print(f"fp32 out: max err {err.max():.3e}, max err / sum|terms| {(err / mag).max():.3e}")
    assert np.all(err <= 1e-5 * mag + 1e-6)
    yb = DeviceArray((B, H, W, cout), np.uint16)
    call("psw_conv3x3_bf16", d_xa.ptr, ca, d_xb.ptr if cb else None, cb, d_w.ptr, d_b.ptr, cout, B, W, H, 1, yb.ptr,
         None)
    gb = bf16_val(yb.to_host()).astype(np.float64)
    r = np.maximum(ref, 0.0)
    assert np.all(np.abs(gb - r) <= 2.0 ** -8 * r + 1e-5 * mag + 1e-6)


# One 8 x 32 image is one tile, so the batch size picks the block width of the
# launch: 16 output channels per block at small batches (unrolled full-slab
# loop, weights staged in LDS), 32 in the middle, all of them at 1024 tiles
# (the general K loop).  The same image must come out bit-identical.
@pytest.mark.parametrize("ca,cb,cout,batches", [
    (64, 96, 96, (1, 400, 1024)),   # five full slabs; one, two and six output blocks per block
    (48, 64, 96, (1, 400, 1024)),   # three full slabs and a 16-channel one
    (8, 0, 32, (1, 1024)),          # one 8-channel slab; one and two output blocks per block
])
def test_conv3x3_block_widths_agree(gpu, ca, cb, cout, batches):
    from psw_amd._lib import call
    from psw_amd.device import DeviceArray
    H, W = 8, 32
    rng = np.random.default_rng(ca + cb + cout)
    xa0 = bf16_bits(rng.standard_normal((1, H, W, ca)))
    xb0 = bf16_bits(rng.standard_normal((1, H // 2, W // 2, cb))) if cb else None
    w = rng.standard_normal((9, ca + cb, cout)).astype(np.float32) * 0.2
    b = rng.standard_normal(cout).astype(np.float32)
    d_w, d_b = _dev(w), _dev(b)
    outs = []
    for B in batches:
        xa = np.concatenate([xa0, bf16_bits(rng.standard_normal((B - 1, H, W, ca)))]) if B > 1 else xa0
        xb = None
        if cb:
            xb = np.concatenate([xb0, bf16_bits(rng.standard_normal((B - 1, H // 2, W // 2, cb)))]) if B > 1 else xb0
        d_xa = _dev(xa)
        d_xb = _dev(xb) if cb else None
        yf = DeviceArray((B, H, W, cout))
        call("psw_conv3x3_bf16", d_xa.ptr, ca, d_xb.ptr if cb else None, cb, d_w.ptr, d_b.ptr, cout, B, W, H, 0,
             yf.ptr, None)
        yb = DeviceArray((B, H, W, cout), np.uint16)
        call("psw_conv3x3_bf16", d_xa.ptr, ca, d_xb.ptr if cb else None, cb, d_w.ptr, d_b.ptr, cout, B, W, H, 1,
             yb.ptr, None)
        outs.append((yf.to_host()[0], yb.to_host()[0]))
    x = bf16_val(xa0).astype(np.float64)
    if cb:
        x = np.concatenate([x, up2(bf16_val(xb0).astype(np.float64))], axis=3)
    wr = bf16_val(bf16_bits(w)).astype(np.float64)
    ref = conv_ref(x, wr, b.astype(np.float64))[0]
    mag = conv_ref(np.abs(x), np.abs(wr))[0] + np.abs(b)
    assert np.all(np.abs(outs[0][0] - ref) <= 1e-5 * mag + 1e-6)
    for B, (f, h) in zip(batches[1:], outs[1:]):
        print(f"batch {B}: fp32 max |diff| to batch 1 {np.abs(f - outs[0][0]).max():.3e}")
        assert np.array_equal(f.view(np.uint32), outs[0][0].view(np.uint32))
        assert np.array_equal(h, outs[0][1])


@pytest.mark.parametrize("cout,cin,B,H,W", [(32, 32, 1, 12, 20), (96, 160, 1, 4, 8), (48, 112, 2, 6, 8),
                                            (32, 80, 1, 8, 8), (64, 64, 2, 5, 9)])
def test_conv3x3_dgrad(gpu, cout, cin, B, H, W):
    from psw_amd._lib import call
    from psw_amd.device import DeviceArray
    rng = np.random.default_rng(cout * 7 + cin)
    dz = bf16_bits(rng.standard_normal((B, H, W, cout)))
    w = rng.standard_normal((9, cin, cout)).astype(np.float32) * 0.2
    wr = bf16_val(bf16_bits(w)).astype(np.float64)
    dzr = bf16_val(dz).astype(np.float64)
    ref = dgrad_ref(dzr, wr)
    mag = dgrad_ref(np.abs(dzr), np.abs(wr))
    d_dz, d_w = _dev(dz), _dev(w)
    dx = DeviceArray((B, H, W, cin))
    call("psw_conv3x3_bf16_dgrad", d_dz.ptr, cout, d_w.ptr, cin, B, W, H, dx.ptr, None)
    err = np.abs(dx.to_host() - ref)
    print(f"dgrad: max err / sum|terms| {(err / (mag + 1e-30)).max():.3e}")
    assert np.all(err <= 1e-5 * mag + 1e-6)


@pytest.mark.parametrize("ca,cb,cout,B,H,W", [(8, 0, 32, 1, 12, 20), (32, 48, 32, 2, 8, 16), (64, 96, 64, 1, 8, 8),
                                              (48, 0, 96, 2, 9, 11), (48, 64, 48, 1, 16, 16)])
def test_conv3x3_wgrad(gpu, ca, cb, cout, B, H, W):
    from psw_amd._lib import call
    from psw_amd.device import DeviceArray
    rng = np.random.default_rng(ca + cb * 3 + cout)
    xa = bf16_bits(rng.standard_normal((B, H, W, ca)))
    xb = bf16_bits(rng.standard_normal((B, H // 2, W // 2, cb))) if cb else None
    dz = bf16_bits(rng.standard_normal((B, H, W, cout)))
    x = bf16_val(xa).astype(np.float64)
    if cb:
        x = np.concatenate([x, up2(bf16_val(xb).astype(np.float64))], axis=3)
    dzr = bf16_val(dz).astype(np.float64)
    dw, db = wgrad_ref(x, dzr)
    mw, mb = wgrad_ref(np.abs(x), np.abs(dzr))
    out = DeviceArray(((9 * (ca + cb) + 1) * cout,))
    d_xa, d_dz = _dev(xa), _dev(dz)
    d_xb = _dev(xb) if cb else None
    call("psw_conv3x3_bf16_wgrad", d_xa.ptr, ca, d_xb.ptr if cb else None, cb, d_dz.ptr, cout, B, W, H, out.ptr, None)
    g = out.to_host()
    gw = g[:9 * (ca + cb) * cout].reshape(9, ca + cb, cout)
    gb = g[9 * (ca + cb) * cout:]
    assert np.all(np.abs(gw - dw) <= 1e-5 * mw + 1e-6)
    assert np.all(np.abs(gb - db) <= 1e-5 * mb + 1e-6)


def _tiny_inputs(seed, B, W, H):
    rng = np.random.default_rng(seed)
    xs, ds, ts = [], [], []
    for b in range(B):
        x = rng.uniform(-1, 1, (H, W, 8)).astype(np.float32)
        x[..., 5:] = 0
        xs.append(x)
        ds.append(rng.uniform(1.0, 4.0, (H, W)).astype(np.float32))
        ts.append(rng.uniform(-0.8, 0.8, (H, W, 8)).astype(np.float32))
    return np.stack(xs), np.stack(ds), np.stack(ts)


def _tiny_net(seed=0, B=1, W=16, H=16):
    from psw_amd import warmstart
    net = warmstart.UNet(B * W * H, seed=seed + 2)
    x, ds, t = _tiny_inputs(seed, B, W, H)
    for b in range(B):
        net.set_input(b, W, H, _dev(x[b]), _dev(ds[b]))
        net.set_target(b, W, H, _dev(t[b]))
    # a head large enough that deeper layers matter for the loss
    p = net.get_params()
    nh = 32 * 8 + 8
    p[-nh:-8] = np.random.default_rng(seed + 100).standard_normal(32 * 8).astype(np.float32) * 0.3
    net.set_params(p)
    return net, x, t


def test_network_forward_matches_numpy(gpu):
    """Whole forward pass against the float64 model with the same bf16
    rounding of weights and activations (accumulation order differs, so
    activations may round differently: 1 % of the output scale)."""
    from unet_ref import unet_forward
    B, W, H = 2, 16, 24
    net, x, _ = _tiny_net(3, B, W, H)
    net.forward(B, W, H)
    theta = net.get_params()
    ref = unet_forward(theta, x)
    got = np.stack([net.output(b, W, H).to_host() for b in range(B)])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"forward vs float64 model: max err / max |out| = {err:.3e}")
    assert err <= 1e-2


def test_network_gradient_matches_numpy_backprop(gpu):
    """Every parameter block's GPU gradient against float64 backpropagation
    with the same bf16 rounding points (weights, activations, dZ).  FP32
    versus float64 accumulation makes some activations and dZ values round
    to the neighbouring bf16 number and a few ReLU masks flip; the effect
    grows towards the input layer.  Bars: every block within 3 % in norm,
    the median block within 0.5 %."""
    from psw_amd import warmstart
    from unet_ref import param_blocks, unet_grad
    B, W, H = 2, 16, 32
    net, x, t = _tiny_net(4, B, W, H)
    theta = net.get_params()
    net.loss_grad(B, W, H)
    g = net.get_grads().astype(np.float64)
    ref = unet_grad(theta, x, t, warmstart.CHANNEL_WEIGHTS)
    errs = []
    for blk, (lo, hi) in enumerate(param_blocks()):
        e = np.linalg.norm(g[lo:hi] - ref[lo:hi]) / max(np.linalg.norm(ref[lo:hi]), 1e-12)
        print(f"block {blk}: |g| {np.linalg.norm(ref[lo:hi]):.4e}, rel diff {e:.2e}")
        errs.append(e)
    assert max(errs) <= 3e-2 and float(np.median(errs)) <= 5e-3, errs


def test_network_gradient_finite_differences(gpu):
    """The GPU gradient (bf16 network) against central differences of the
    float64 model without any rounding, along random directions in the
    blocks nearest the output (head, d0, d2a): 5 %.  Deeper blocks see the
    bf16 rounding of every activation on the way (see the rounding-matched
    backprop test above)."""
    from psw_amd import warmstart
    from unet_ref import huber_loss, param_blocks, unet_forward
    B, W, H = 1, 16, 16
    net, x, t = _tiny_net(0, B, W, H)
    theta = net.get_params().astype(np.float64)
    net.loss_grad(B, W, H)
    g = net.get_grads().astype(np.float64)
    blocks = param_blocks()
    assert blocks[-1][1] == warmstart.param_count() == theta.size
    wc = warmstart.CHANNEL_WEIGHTS
    rng = np.random.default_rng(5)
    f = lambda th: huber_loss(unet_forward(th, x, round_bf16=False), t, wc)   # noqa: E731
    for blk in [len(blocks) - 2, len(blocks) - 1, 26, 27, 16]:
        lo, hi = blocks[blk]
        v = np.zeros_like(theta)
        v[lo:hi] = rng.standard_normal(hi - lo)
        eps = 1e-4 * max(float(np.abs(theta[lo:hi]).mean()), 1e-2) / float(np.abs(v[lo:hi]).mean())
        fd = (f(theta + eps * v) - f(theta - eps * v)) / (2 * eps)
        an = float(np.dot(g, v))
        print(f"block {blk}: GPU gradient {an:.6e}, float64 central difference {fd:.6e}")
        assert abs(fd - an) <= 0.05 * abs(fd) + 1e-7, (blk, fd, an)


def test_training_loss_decreases(gpu):
    """A fixed synthetic regression (targets are smooth functions of the
    input channels): 80 Adam steps must cut the loss by more than half."""
    from psw_amd import warmstart
    B, W, H = 2, 32, 32
    net = warmstart.UNet(B * W * H, seed=3)
    rng = np.random.default_rng(11)
    for b in range(B):
        x = np.zeros((H, W, 8), np.float32)
        yy, xx = np.mgrid[0:H, 0:W] / 32.0
        x[..., 0] = np.sin(6 * xx + b) * 0.5
        x[..., 1] = np.cos(5 * yy) * 0.5
        x[..., 2] = rng.uniform(-0.5, 0.5, (H, W))
        t = np.zeros((H, W, 8), np.float32)
        t[..., 0] = x[..., 0] + x[..., 1]
        t[..., 3] = 0.5 * x[..., 0] * x[..., 1]
        t[..., 5] = np.roll(x[..., 0], 1, axis=1)
        net.set_input(b, W, H, _dev(x), _dev(np.ones((H, W), np.float32)))
        net.set_target(b, W, H, _dev(t))
    first = net.loss_grad(B, W, H)
    for step in range(1, 81):
        if step > 1:
            net.loss_grad(B, W, H, want_loss=False)
        net.adam_step(1e-3, step)
    last = net.loss_grad(B, W, H)
    print(f"loss {first:.5f} -> {last:.5f}")
    assert last < 0.5 * first


def test_warm_start_fields(gpu):
    """psw_unet_warm_start maps the network output to TGV2 fields:
    u = ds (1 + 0.1 o0), u1 = 0.01 ds (o1, o2), p and q projected."""
    from psw_amd.device import DeviceArray
    B, W, H = 1, 16, 8
    net, _, _ = _tiny_net(1, B, W, H)
    net.forward(B, W, H)
    out = net.output(0, W, H).to_host().astype(np.float64)
    d = [DeviceArray(s) for s in ((H, W), (2, H, W), (2, H, W), (3, H, W))]
    net.warm_start(0, W, H, *d)
    u, u1, p, q = [a.to_host() for a in d]
    o = np.clip(out, -10, 10)
    # ds is recovered from u and o0
    dsr = u / (1 + 0.1 * o[..., 0])
    np.testing.assert_allclose(u1[0], 0.01 * dsr * o[..., 1], rtol=1e-5, atol=1e-7)
    n2 = np.maximum(1, np.hypot(out[..., 3], out[..., 4]))
    np.testing.assert_allclose(p[0], out[..., 3] / n2, rtol=1e-5, atol=1e-7)
    n3 = np.maximum(1, np.sqrt(out[..., 5] ** 2 + out[..., 6] ** 2 + 2 * out[..., 7] ** 2))
    np.testing.assert_allclose(q[2], out[..., 7] / n3, rtol=1e-5, atol=1e-7)
