"""CPU checks of the numpy references the U-Net kernels are tested against:
the convolution against a brute-force loop, the data and weight gradients
through the adjoint identities <conv(x, w), dz> = <x, dgrad(dz, w)> =
<w, wgrad(x, dz)>, and bf16 rounding against hand-computed bit patterns."""
import numpy as np

from unet_ref import bf16_bits, bf16_val, conv_ref, dgrad_ref, up2, wgrad_ref


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0, 65504.0], np.float32)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 ties up to 1+2^-6
    np.testing.assert_array_equal(bf16_bits(x), np.array([0x3F80, 0x3F80, 0x3F82, 0xC020, 0x0000, 0x4780], np.uint16))
    np.testing.assert_array_equal(bf16_val(bf16_bits(x))[:5], np.array([1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, 0.0],
                                                                      np.float32))


def test_conv_ref_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 5, 3))
    w = rng.standard_normal((9, 3, 2))
    y = conv_ref(x, w)
    for b in range(2):
        for i in range(4):
            for j in range(5):
                s = np.zeros(2)
                for tap in range(9):
                    yy, xx = i + tap // 3 - 1, j + tap % 3 - 1
                    if 0 <= yy < 4 and 0 <= xx < 5:
                        s += x[b, yy, xx] @ w[tap]
                np.testing.assert_allclose(y[b, i, j], s, rtol=1e-12, atol=1e-12)


def test_gradient_refs_are_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 6, 7, 4))
    w = rng.standard_normal((9, 4, 5))
    dz = rng.standard_normal((2, 6, 7, 5))
    lhs = np.sum(conv_ref(x, w) * dz)
    np.testing.assert_allclose(np.sum(x * dgrad_ref(dz, w)), lhs, rtol=1e-10)
    dw, db = wgrad_ref(x, dz)
    np.testing.assert_allclose(np.sum(w * dw), lhs, rtol=1e-10)
    np.testing.assert_allclose(db, dz.sum(axis=(0, 1, 2)))


def test_upsample_nearest():
    x = np.arange(2 * 2 * 3 * 1, dtype=float).reshape(2, 2, 3, 1)
    u = up2(x)
    assert u.shape == (2, 4, 6, 1)
    assert u[1, 3, 5, 0] == x[1, 1, 2, 0] and u[0, 2, 1, 0] == x[0, 1, 0, 0]


def test_network_model_shapes_and_zero_head():
    """The float64 network model: parameter layout size, output shape, and a
    zero head gives exactly the head bias."""
    from unet_ref import param_blocks, unet_forward
    n = param_blocks()[-1][1]
    assert n == 480264
    rng = np.random.default_rng(0)
    th = rng.standard_normal(n) * 0.05
    lo, hi = param_blocks()[-2]
    th[lo:hi] = 0.0
    b0, b1 = param_blocks()[-1]
    out = unet_forward(th, rng.standard_normal((1, 16, 8, 8)), round_bf16=False)
    assert out.shape == (1, 16, 8, 8)
    np.testing.assert_array_equal(out, np.broadcast_to(th[b0:b1], out.shape))


def test_network_backprop_matches_finite_differences():
    """unet_grad on the float64 model without rounding against central
    differences along a random direction in every parameter block (pins the
    routing of concatenation, pooling and upsampling gradients)."""
    from unet_ref import huber_loss, param_blocks, unet_forward, unet_grad
    rng = np.random.default_rng(2)
    blocks = param_blocks()
    n = blocks[-1][1]
    th = np.zeros(n)
    for k in range(0, len(blocks) - 2, 2):
        lo, hi = blocks[k]
        fan = (hi - lo) / (blocks[k + 1][1] - blocks[k + 1][0]) / 1.0
        th[lo:hi] = rng.standard_normal(hi - lo) * np.sqrt(2.0 / fan)
        th[blocks[k + 1][0]:blocks[k + 1][1]] = rng.standard_normal(blocks[k + 1][1] - blocks[k + 1][0]) * 0.05
    lo, hi = blocks[-2]
    th[lo:hi] = rng.standard_normal(hi - lo) * 0.3
    x = rng.uniform(-1, 1, (1, 16, 16, 8))
    t = rng.uniform(-0.8, 0.8, (1, 16, 16, 8))
    wc = np.array([1.0, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5])
    g = unet_grad(th, x, t, wc, round_bf16=False)
    f = lambda p: huber_loss(unet_forward(p, x, round_bf16=False), t, wc)   # noqa: E731
    for k, (lo, hi) in enumerate(blocks):
        v = np.zeros(n)
        v[lo:hi] = rng.standard_normal(hi - lo)
        eps = 1e-6
        fd = (f(th + eps * v) - f(th - eps * v)) / (2 * eps)
        np.testing.assert_allclose(np.dot(g, v), fd, rtol=1e-4, atol=1e-9, err_msg=f"block {k}")
