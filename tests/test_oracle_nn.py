"""Pins of the CPU network oracle (oracle/nn_ref.py) that the GPU network
kernels are checked against: direct sums, adjoint identities and central
finite differences of the float64 loss."""
import numpy as np
import pytest

from oracle import nn_ref as R


def brute_conv(x, W, stride, out_shape):
    B, nz, ny, nx, cin = x.shape
    oz, oy, ox = out_shape
    out = np.zeros((B, oz, oy, ox, W.shape[2]))
    for b in range(B):
        for z in range(oz):
            for y in range(oy):
                for xx in range(ox):
                    for t, tz, ty, tx in R.taps():
                        sz, sy, sx = stride * z + tz, stride * y + ty, stride * xx + tx
                        if 0 <= sz < nz and 0 <= sy < ny and 0 <= sx < nx:
                            out[b, z, y, xx] += x[b, sz, sy, sx] @ W[t]
    return out


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_matches_direct_sum(stride):
    rng = np.random.default_rng(1)
    shape = (5, 3, 5)
    out_shape = shape if stride == 1 else tuple((s + 1) // 2 for s in shape)
    x = rng.standard_normal((2,) + shape + (3,))
    W = rng.standard_normal((27, 3, 4))
    ref = brute_conv(x, W, stride, out_shape)
    got = R.conv_fwd(x, W, stride, out_shape, np.float64)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_adjoints(stride):
    rng = np.random.default_rng(2)
    shape = (9, 5, 9)
    out_shape = shape if stride == 1 else tuple((s + 1) // 2 for s in shape)
    x = rng.standard_normal((2,) + shape + (5,))
    z = rng.standard_normal((2,) + out_shape + (6,))
    W = rng.standard_normal((27, 5, 6))
    lhs = np.sum(R.conv_fwd(x, W, stride, out_shape, np.float64) * z)
    rhs_x = np.sum(x * R.conv_bwd_data(z, W, stride, shape, np.float64))
    assert abs(lhs - rhs_x) <= 1e-10 * abs(lhs)
    # weight adjoint: d<conv(x, W), z>/dW = conv_bwd_weight(x, z)
    dW = R.conv_bwd_weight(x, z, stride, np.float64)
    assert abs(np.sum(dW * W) - lhs) <= 1e-10 * abs(lhs)


def test_prolongation_is_trilinear_and_adjoint():
    rng = np.random.default_rng(3)
    c = rng.standard_normal((2, 3, 5, 9, 4))
    f = R.prolong(c)
    assert f.shape == (2, 5, 9, 17, 4)
    # coarse nodes are injected, midpoints average their neighbours
    assert np.array_equal(f[:, ::2, ::2, ::2], c)
    assert np.allclose(f[:, 1, 2, 4], 0.5 * (c[:, 0, 1, 2] + c[:, 1, 1, 2]))
    assert np.allclose(f[:, 1, 1, 1], c[:, 0:2, 0:2, 0:2].mean(axis=(1, 2, 3)))
    # a linear field is reproduced exactly
    zz, yy, xx = np.meshgrid(np.arange(3), np.arange(5), np.arange(9), indexing="ij")
    lin = (0.3 * zz - 1.2 * yy + 0.7 * xx)[None, ..., None]
    zf, yf, xf = np.meshgrid(np.arange(5), np.arange(9), np.arange(17), indexing="ij")
    assert np.allclose(R.prolong(lin)[0, ..., 0], 0.15 * zf - 0.6 * yf + 0.35 * xf)
    g = rng.standard_normal(f.shape)
    assert abs(np.sum(f * g) - np.sum(c * R.prolong_adj(g))) <= 1e-10 * abs(np.sum(f * g))


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0],
                    dtype=np.float32)
    got = R.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                       # tie -> even (mantissa 0)
    assert got[2] == np.float32(1.0 + 4 * 2 ** -8)  # tie -> even (rounds up)
    assert got[3] == 1.0
    assert got[4] == -2.5 and got[5] == 0.0
    bits = R.to_bf16_bits(vals)
    assert np.array_equal(R.from_bf16_bits(bits), got)


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    nx, ny, nz = 17, 9, 9
    layout, npar = R.param_layout()
    p = np.zeros(npar)
    for i, (_, _, _, cin, cout, _, _) in enumerate(R.LAYERS):
        w_off, b_off = layout[i]
        p[w_off:w_off + 27 * cin * cout] = rng.standard_normal(27 * cin * cout) * np.sqrt(2 / (27 * cin))
        p[b_off:b_off + cout] = 0.1 * rng.standard_normal(cout)
    x = np.zeros((1, nz, ny, nx, 16))
    x[..., :8] = rng.standard_normal((1, nz, ny, nx, 8))
    tgt = rng.standard_normal((1, nz, ny, nx, 3))
    net = R.Net(nx, ny, nz, p, bf16_mode=False)
    loss, g = net.loss_and_grad(x, tgt)
    assert np.isfinite(loss) and g.shape == (npar,)
    # small steps: a larger one lets some leaky-ReLU inputs cross zero
    for k in range(3):
        d = rng.standard_normal(npar)
        eps = 1e-7
        lp, _ = R.Net(nx, ny, nz, p + eps * d, bf16_mode=False).loss_and_grad(x, tgt)
        lm, _ = R.Net(nx, ny, nz, p - eps * d, bf16_mode=False).loss_and_grad(x, tgt)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-7 * abs(fd) + 1e-12, (k, fd, g @ d)
    # per-layer: a perturbation of one layer's weights only
    for i in (0, 6, 7, 13):
        w_off, b_off = layout[i]
        d = np.zeros(npar)
        d[w_off:b_off] = rng.standard_normal(b_off - w_off)
        lp, _ = R.Net(nx, ny, nz, p + 1e-7 * d, bf16_mode=False).loss_and_grad(x, tgt)
        lm, _ = R.Net(nx, ny, nz, p - 1e-7 * d, bf16_mode=False).loss_and_grad(x, tgt)
        fd = (lp - lm) / 2e-7
        assert abs(fd - g @ d) <= 1e-6 * abs(fd) + 1e-12, (R.LAYERS[i][0], fd, g @ d)


def test_param_count_is_about_one_million():
    _, npar = R.param_layout()
    assert 1_000_000 <= npar <= 1_100_000


def test_adam_first_step_is_lr_times_sign():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    m, v = np.zeros(3), np.zeros(3)
    out = R.adam_step(p.copy(), g, m, v, 1, 0.01, eps=0.0)
    assert np.allclose(out, p - 0.01 * np.sign(g))
