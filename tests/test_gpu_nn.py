"""GPU parity of the warm-start network kernels (MFMA convolutions, their
data and weight gradients, prolongation, loss, Adam) against the CPU network
oracle (oracle/nn_ref.py), layer by layer and for the whole network."""
import ctypes as C

import numpy as np
import pytest

import gpufem_amd as gf
from oracle import nn_ref as R

pytestmark = pytest.mark.gpu


def _i3(v):
    return np.ascontiguousarray(v, dtype=np.int32)


def bf16_data(rng, shape):
    bits = R.to_bf16_bits(rng.standard_normal(shape).astype(np.float32))
    return bits, R.from_bf16_bits(bits).reshape(shape).astype(np.float64)


def grid_out(vin, stride):
    return vin if stride == 1 else [(n + 1) // 2 for n in vin]


CONV_CASES = [  # cin, cout, stride
    (16, 16, 1), (48, 16, 1), (16, 32, 2), (32, 64, 2), (96, 32, 1), (192, 64, 1), (64, 128, 2),
]


@pytest.mark.parametrize("cin,cout,stride", CONV_CASES)
def test_conv_forward_matches_oracle(gpu_ctx, cin, cout, stride):
    ctx, lib = gpu_ctx, gpu_ctx.lib
    rng = np.random.default_rng(cin * 1000 + cout + stride)
    vin = [17, 9, 9]
    vout = grid_out(vin, stride)
    B = 2
    xb, x = bf16_data(rng, (B, vin[2], vin[1], vin[0], cin))
    W = (rng.standard_normal((27, cin, cout)) / np.sqrt(27 * cin)).astype(np.float32)
    bias = rng.standard_normal(cout).astype(np.float32)
    ref = R.conv_fwd(x, R.bf16(W).astype(np.float64), stride, (vout[2], vout[1], vout[0]),
                     np.float64) + bias
    d_x, d_w, d_b = (gf.DevBuf.from_array(ctx, a) for a in (xb, W, bias))
    d_o = gf.DevBuf(ctx, ref.size * 4)
    ctx._check(lib.gf_nn_conv3d(ctx.handle, d_x.ptr, B, gf._ptr(_i3(vin)), cin, d_w.ptr, d_b.ptr,
                                cout, stride, 0, 0, 0.0, gf._ptr(_i3(vout)), None, d_o.ptr, 0))
    got = d_o.download(np.float32).reshape(ref.shape)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    # bf16 output with bias + leaky ReLU fused
    d_o16 = gf.DevBuf(ctx, ref.size * 2)
    ctx._check(lib.gf_nn_conv3d(ctx.handle, d_x.ptr, B, gf._ptr(_i3(vin)), cin, d_w.ptr, d_b.ptr,
                                cout, stride, 0, 1, 0.1, gf._ptr(_i3(vout)), d_o16.ptr, None, 0))
    got16 = R.from_bf16_bits(d_o16.download(np.uint16)).reshape(ref.shape)
    act = np.where(ref > 0, ref, 0.1 * ref)
    assert np.all(np.abs(got16 - act) <= 2.0 ** -7 * np.abs(act) + 1e-6)


@pytest.mark.parametrize("cin,cout,stride", CONV_CASES)
def test_conv_data_gradient_matches_oracle(gpu_ctx, cin, cout, stride):
    ctx, lib = gpu_ctx, gpu_ctx.lib
    rng = np.random.default_rng(7 + cin + cout * 3 + stride)
    vin = [17, 9, 9]
    vout = grid_out(vin, stride)
    B = 2
    # layer cin -> cout; data gradient maps dz [.., cout] on vout to dx [.., cin] on vin
    dzb, dz = bf16_data(rng, (B, vout[2], vout[1], vout[0], cout))
    W = (rng.standard_normal((27, cin, cout)) / np.sqrt(27 * cout)).astype(np.float32)
    ref = R.conv_bwd_data(dz, R.bf16(W).astype(np.float64), stride, (vin[2], vin[1], vin[0]),
                          np.float64)
    d_dz, d_w = gf.DevBuf.from_array(ctx, dzb), gf.DevBuf.from_array(ctx, W)
    d_o = gf.DevBuf(ctx, ref.size * 4)
    ctx._check(lib.gf_nn_conv3d(ctx.handle, d_dz.ptr, B, gf._ptr(_i3(vout)), cout, d_w.ptr, None,
                                cin, stride, 1, 0, 0.0, gf._ptr(_i3(vin)), None, d_o.ptr, 0))
    got = d_o.download(np.float32).reshape(ref.shape)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    # accumulate into an existing fp32 gradient
    ctx._check(lib.gf_nn_conv3d(ctx.handle, d_dz.ptr, B, gf._ptr(_i3(vout)), cout, d_w.ptr, None,
                                cin, stride, 1, 0, 0.0, gf._ptr(_i3(vin)), None, d_o.ptr, 1))
    got2 = d_o.download(np.float32).reshape(ref.shape)
    assert np.max(np.abs(got2 - 2 * ref)) <= 2e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("cin,cout,stride", CONV_CASES)
def test_conv_weight_gradient_matches_oracle(gpu_ctx, cin, cout, stride):
    ctx, lib = gpu_ctx, gpu_ctx.lib
    rng = np.random.default_rng(11 + cin * 7 + cout + stride)
    vin = [17, 9, 9]
    vout = grid_out(vin, stride)
    B = 3
    xb, x = bf16_data(rng, (B, vin[2], vin[1], vin[0], cin))
    dzb, dz = bf16_data(rng, (B, vout[2], vout[1], vout[0], cout))
    ref = R.conv_bwd_weight(x, dz, stride, np.float64)
    refb = dz.reshape(-1, cout).sum(axis=0)
    d_x, d_dz = gf.DevBuf.from_array(ctx, xb), gf.DevBuf.from_array(ctx, dzb)
    d_w, d_b = gf.DevBuf(ctx, ref.size * 4), gf.DevBuf(ctx, cout * 4)
    ctx._check(lib.gf_nn_conv3d_wgrad(ctx.handle, d_x.ptr, B, gf._ptr(_i3(vin)), cin, d_dz.ptr,
                                      gf._ptr(_i3(vout)), cout, stride, d_w.ptr, d_b.ptr))
    got = d_w.download(np.float32).reshape(ref.shape)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    gotb = d_b.download(np.float32)
    assert np.max(np.abs(gotb - refb)) <= 1e-5 * np.max(np.abs(refb))


def _random_params(rng):
    layout, npar = R.param_layout()
    p = np.zeros(npar, dtype=np.float32)
    for i, (_, _, _, cin, cout, _, _) in enumerate(R.LAYERS):
        w_off, b_off = layout[i]
        p[w_off:w_off + 27 * cin * cout] = rng.standard_normal(27 * cin * cout) * np.sqrt(2 / (27 * cin))
        p[b_off:b_off + cout] = 0.05 * rng.standard_normal(cout)
    return p


def test_network_layout_matches_oracle(gpu_ctx):
    net = gf.Net(gpu_ctx, 17, 9, 9, 2)
    layout, npar = R.param_layout()
    assert net.nparams == npar
    for (cin, cout, stride, act, lin, lout, w_off, b_off), spec, lay in zip(net.layers(), R.LAYERS,
                                                                              layout):
        assert (cin, cout, stride, act) == (spec[3], spec[4], spec[5], spec[6])
        assert (w_off, b_off) == lay


def test_network_forward_and_gradient_match_oracle(gpu_ctx):
    ctx = gpu_ctx
    rng = np.random.default_rng(21)
    nx, ny, nz, B = 17, 9, 9, 2
    net = gf.Net(ctx, nx, ny, nz, B)
    p = _random_params(rng)
    net.set_params(p)
    x = np.zeros((B, nz, ny, nx, 16), dtype=np.float32)
    x[..., :8] = rng.standard_normal((B, nz, ny, nx, 8))
    xb = R.to_bf16_bits(x)
    tgt = np.zeros((B, nz, ny, nx, 4), dtype=np.float32)
    tgt[..., :3] = rng.standard_normal((B, nz, ny, nx, 3))
    d_in, d_t = gf.DevBuf.from_array(ctx, xb), gf.DevBuf.from_array(ctx, tgt)
    d_out = gf.DevBuf(ctx, B * nx * ny * nz * 16 * 4)
    net.forward(d_in.ptr, B, d_out.ptr)
    pred = d_out.download(np.float32).reshape(B, nz, ny, nx, 16)
    onet = R.Net(nx, ny, nz, p, bf16_mode=True)
    ref = onet.forward(R.from_bf16_bits(xb).reshape(x.shape))
    err = np.linalg.norm(pred - ref) / np.linalg.norm(ref)
    assert err <= 1e-2, err
    loss = net.backward(d_in.ptr, d_t.ptr, B)
    g = net.grads()
    oloss, og = onet.loss_and_grad(R.from_bf16_bits(xb).reshape(x.shape), tgt[..., :3])
    assert abs(loss - oloss) <= 1e-2 * oloss
    # bf16 storage of every activation and gradient: the first layer, at the
    # end of the backward chain, collects the most rounding (a few percent);
    # a wrong index or operand would give O(1) differences
    layout, _ = R.param_layout()
    for i, spec in enumerate(R.LAYERS):
        w_off, b_off = layout[i]
        end = b_off + spec[4]
        a, b = g[w_off:end].astype(np.float64), og[w_off:end].astype(np.float64)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        assert rel <= 6e-2 and cos >= 0.998, (spec[0], rel, cos)
    # Adam on the GPU = the oracle's Adam applied to the GPU's gradient
    net.adam(1e-3)
    m, v = np.zeros_like(p), np.zeros_like(p)
    p_ref = R.adam_step(p.astype(np.float64), g.astype(np.float64), m.astype(np.float64),
                        v.astype(np.float64), 1, 1e-3)
    assert np.max(np.abs(net.params() - p_ref)) <= 1e-6


def test_training_reduces_loss(gpu_ctx):
    ctx = gpu_ctx
    rng = np.random.default_rng(5)
    nx, ny, nz, B = 17, 9, 9, 4
    net = gf.Net(ctx, nx, ny, nz, B, seed=3)
    zz, yy, xx = np.meshgrid(np.linspace(0, 1, nz), np.linspace(0, 1, ny), np.linspace(0, 1, nx),
                             indexing="ij")
    x = np.zeros((B, nz, ny, nx, 16), dtype=np.float32)
    tgt = np.zeros((B, nz, ny, nx, 4), dtype=np.float32)
    for b in range(B):
        a = rng.standard_normal(3)
        x[b, ..., 0] = a[0]
        x[b, ..., 1] = a[1]
        x[b, ..., 5], x[b, ..., 6], x[b, ..., 7] = xx, yy, zz
        tgt[b, ..., 0] = a[0] * xx ** 2
        tgt[b, ..., 1] = a[1] * xx * yy
        tgt[b, ..., 2] = (a[0] - a[1]) * zz
    d_in = gf.DevBuf.from_array(ctx, R.to_bf16_bits(x))
    d_t = gf.DevBuf.from_array(ctx, tgt)
    losses = [net.train_step(d_in.ptr, d_t.ptr, B, 2e-3) for _ in range(60)]
    assert np.all(np.isfinite(losses))
    assert min(losses[-5:]) < 0.2 * losses[0], losses[::10]
