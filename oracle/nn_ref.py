"""CPU restatement of the learned warm-start network (numpy).

TEST INFRASTRUCTURE ONLY: the oracle the HIP network kernels of gpufem_amd
(csrc/nn.hip, csrc/nn.cpp) are tested against.  Only tests/ may import it.

The reference has no learned component; this module restates textbook
definitions (3x3x3 zero-padded convolution, strided and transposed; trilinear
vertex-centred prolongation and its adjoint; leaky ReLU; mean squared error;
reverse-mode differentiation of their composition; Adam) for the same layer
table as the HIP network.  It runs in float64 (for finite-difference pins) or
float32 with the HIP path's bfloat16 roundings emulated (`bf16=True`):
activations, activation gradients and the weights the kernels read are
rounded to bfloat16, products accumulate in float32, biases, weight
gradients and the final layer stay float32.

Pins (tests/test_oracle_nn.py): central finite differences of the float64
loss against the float64 gradient; adjoint identities <A x, y> = <x, A^T y>
of the convolution (strides 1 and 2) and of the prolongation; the
convolution against a direct 7-loop sum on tiny tensors.
"""
import numpy as np

SLOPE = 0.1
CIN, COUT, OUT_USED = 16, 16, 3

# name, (in buffer, channel offset), (out buffer, channel offset), cin, cout,
# stride, activation -- the same table as csrc/nn.cpp
LAYERS = [
    ("e0a", ("in", 0), ("a0", 0), 16, 16, 1, 1),
    ("e0b", ("a0", 0), ("cat0", 32), 16, 16, 1, 1),
    ("e1a", ("cat0", 32), ("a1", 0), 16, 32, 2, 1),
    ("e1b", ("a1", 0), ("cat1", 64), 32, 32, 1, 1),
    ("e2a", ("cat1", 64), ("a2", 0), 32, 64, 2, 1),
    ("e2b", ("a2", 0), ("cat2", 128), 64, 64, 1, 1),
    ("e3a", ("cat2", 128), ("a3", 0), 64, 128, 2, 1),
    ("d2a", ("cat2", 0), ("b2", 0), 192, 64, 1, 1),
    ("d2b", ("b2", 0), ("c2", 0), 64, 64, 1, 1),
    ("d1a", ("cat1", 0), ("b1", 0), 96, 32, 1, 1),
    ("d1b", ("b1", 0), ("c1", 0), 32, 32, 1, 1),
    ("d0a", ("cat0", 0), ("b0", 0), 48, 16, 1, 1),
    ("d0b", ("b0", 0), ("c0", 0), 16, 16, 1, 1),
    ("out", ("c0", 0), ("pred", 0), 16, COUT, 1, 0),
]
# prolongation (coarse buffer, fine buffer, channels) after layer index
UPS = {6: ("a3", "cat2", 128), 8: ("c2", "cat1", 64), 10: ("c1", "cat0", 32)}
BUFS = {"in": (0, 16), "a0": (0, 16), "cat0": (0, 48), "a1": (1, 32), "cat1": (1, 96),
        "a2": (2, 64), "cat2": (2, 192), "a3": (3, 128), "b2": (2, 64), "c2": (2, 64),
        "b1": (1, 32), "c1": (1, 32), "b0": (0, 16), "c0": (0, 16), "pred": (0, 16)}


def bf16(a):
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def to_bf16_bits(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)


def levels(nx, ny, nz):
    lv = [(nz, ny, nx)]
    for _ in range(3):
        z, y, x = lv[-1]
        lv.append(((z + 1) // 2, (y + 1) // 2, (x + 1) // 2))
    return lv


def param_layout():
    out, off = [], 0
    for name, _, _, cin, cout, _, _ in LAYERS:
        w_off = off
        off += 27 * cin * cout
        b_off = off
        off += cout
        out.append((w_off, b_off))
    return out, off


def taps():
    for t in range(27):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def _window(xp, t, stride, out_shape):
    """Slice of the 1-padded input holding x[s v + t] for every output v."""
    _, dz, dy, dx = t
    oz, oy, ox = out_shape
    s = stride
    return xp[:, 1 + dz:2 + dz + s * (oz - 1):s, 1 + dy:2 + dy + s * (oy - 1):s,
              1 + dx:2 + dx + s * (ox - 1):s, :]


def conv_fwd(x, W, stride, out_shape, dtype):
    """sum_t x[s v + t] W[t]: x [B, nz, ny, nx, cin], W [27, cin, cout]."""
    xp = np.pad(x.astype(dtype), ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((x.shape[0],) + tuple(out_shape) + (W.shape[2],), dtype=dtype)
    for t in taps():
        out += _window(xp, t, stride, out_shape) @ W[t[0]].astype(dtype)
    return out


def conv_bwd_data(dz, W, stride, in_shape, dtype):
    """Adjoint of conv_fwd with respect to x."""
    B = dz.shape[0]
    dxp = np.zeros((B,) + tuple(s + 2 for s in in_shape) + (W.shape[1],), dtype=dtype)
    out_shape = dz.shape[1:4]
    for t in taps():
        _, tz, ty, tx = t
        s = stride
        oz, oy, ox = out_shape
        dxp[:, 1 + tz:2 + tz + s * (oz - 1):s, 1 + ty:2 + ty + s * (oy - 1):s,
            1 + tx:2 + tx + s * (ox - 1):s, :] += dz.astype(dtype) @ W[t[0]].T.astype(dtype)
    return dxp[:, 1:-1, 1:-1, 1:-1, :]


def conv_bwd_weight(x, dz, stride, dtype):
    out_shape = dz.shape[1:4]
    xp = np.pad(x.astype(dtype), ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dW = np.zeros((27, x.shape[-1], dz.shape[-1]), dtype=dtype)
    d2 = dz.reshape(-1, dz.shape[-1]).astype(dtype)
    for t in taps():
        dW[t[0]] = _window(xp, t, stride, out_shape).reshape(-1, x.shape[-1]).T @ d2
    return dW


def _up1d(a, axis):
    n = a.shape[axis]
    shape = list(a.shape)
    shape[axis] = 2 * n - 1
    out = np.zeros(shape, dtype=a.dtype)
    ev = [slice(None)] * a.ndim
    od = [slice(None)] * a.ndim
    ev[axis] = slice(0, None, 2)
    od[axis] = slice(1, None, 2)
    lo = [slice(None)] * a.ndim
    hi = [slice(None)] * a.ndim
    lo[axis] = slice(0, n - 1)
    hi[axis] = slice(1, n)
    out[tuple(ev)] = a
    out[tuple(od)] = 0.5 * (a[tuple(lo)] + a[tuple(hi)])
    return out


def _up1d_adj(f, axis):
    m = f.shape[axis]
    n = (m + 1) // 2
    shape = list(f.shape)
    shape[axis] = n
    out = np.zeros(shape, dtype=f.dtype)
    ev = [slice(None)] * f.ndim
    ev[axis] = slice(0, None, 2)
    out += f[tuple(ev)]
    od = [slice(None)] * f.ndim
    od[axis] = slice(1, None, 2)
    half = 0.5 * f[tuple(od)]
    lo = [slice(None)] * f.ndim
    hi = [slice(None)] * f.ndim
    lo[axis] = slice(0, n - 1)
    hi[axis] = slice(1, n)
    out[tuple(lo)] += half
    out[tuple(hi)] += half
    return out


def prolong(c):
    """Trilinear vertex-centred prolongation of [B, z, y, x, C]."""
    return _up1d(_up1d(_up1d(c, 1), 2), 3)


def prolong_adj(f):
    return _up1d_adj(_up1d_adj(_up1d_adj(f, 3), 2), 1)


class Net:
    """The HIP network's forward and backward on the CPU."""

    def __init__(self, nx, ny, nz, params, bf16_mode=True):
        self.lv = levels(nx, ny, nz)
        self.layout, self.nparams = param_layout()
        assert params.size == self.nparams
        self.bf16 = bf16_mode
        self.dtype = np.float32 if bf16_mode else np.float64
        self.params = params.astype(self.dtype)

    def _r(self, a):
        return bf16(a) if self.bf16 else a

    def _w(self, i):
        _, _, _, cin, cout, _, _ = LAYERS[i]
        w_off, b_off = self.layout[i]
        W = self.params[w_off:w_off + 27 * cin * cout].reshape(27, cin, cout)
        b = self.params[b_off:b_off + cout]
        return (bf16(W) if self.bf16 else W), b

    def forward(self, x):
        """x: [B, nz, ny, nx, 16]; returns predictions [B, nz, ny, nx, 16]."""
        B = x.shape[0]
        buf = {}
        for name, (lvl, ch) in BUFS.items():
            if name in ("in", "pred"):
                continue
            buf[name] = np.zeros((B,) + self.lv[lvl] + (ch,), dtype=self.dtype)
        buf["in"] = self._r(x.astype(self.dtype))
        for i, (name, (ib, io), (ob, oo), cin, cout, s, act) in enumerate(LAYERS):
            W, b = self._w(i)
            xin = buf[ib][..., io:io + cin]
            z = conv_fwd(xin, W, s, self.lv[BUFS[ob][0]], self.dtype) + b
            if act:
                z = np.where(z > 0, z, SLOPE * z)
            if ob == "pred":
                buf["pred"] = z
            else:
                buf[ob][..., oo:oo + cout] = self._r(z)
            if i in UPS:
                cb, fb, c = UPS[i]
                buf[fb][..., 0:c] = self._r(prolong(buf[cb][..., 0:c]))
        self.buf = buf
        return buf["pred"]

    def loss_and_grad(self, x, target):
        """target: [B, nz, ny, nx, 3]; returns (loss, flat gradient)."""
        pred = self.forward(x)
        buf = self.buf
        d = pred[..., :OUT_USED] - target.astype(self.dtype)
        count = d.size
        loss = float(np.sum(d.astype(np.float64) ** 2) / count)
        grads = np.zeros(self.nparams, dtype=self.dtype)
        gbuf = {k: np.zeros_like(v) for k, v in buf.items() if k != "in"}
        dp = np.zeros_like(pred)
        dp[..., :OUT_USED] = 2.0 * d / count
        gbuf["pred"] = self._r(dp)
        for i in range(len(LAYERS) - 1, -1, -1):
            if i in UPS:
                cb, fb, c = UPS[i]
                gbuf[cb][..., 0:c] = self._r(prolong_adj(gbuf[fb][..., 0:c]))
            name, (ib, io), (ob, oo), cin, cout, s, act = LAYERS[i]
            W, _ = self._w(i)
            dz = gbuf[ob][..., oo:oo + cout]
            if act:
                y = buf[ob][..., oo:oo + cout]
                dz = np.where(y > 0, dz, self._r(SLOPE * dz))
                gbuf[ob][..., oo:oo + cout] = dz
            xin = buf[ib][..., io:io + cin]
            w_off, b_off = self.layout[i]
            grads[w_off:w_off + 27 * cin * cout] = conv_bwd_weight(xin, dz, s, self.dtype).reshape(-1)
            grads[b_off:b_off + cout] = dz.reshape(-1, cout).sum(axis=0)
            if ib == "in":
                continue
            dx = conv_bwd_data(dz, W, s, self.lv[BUFS[ib][0]], self.dtype)
            if name in ("e1a", "e2a", "e3a"):
                gbuf[ib][..., io:io + cin] = self._r(gbuf[ib][..., io:io + cin] + dx)
            else:
                gbuf[ib][..., io:io + cin] = self._r(dx)
        return loss, grads


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (bias correction folded into the step size)."""
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    lr_t = lr * np.sqrt(1 - b2 ** t) / (1 - b1 ** t)
    p -= lr_t * m / (np.sqrt(v) + eps)
    return p
