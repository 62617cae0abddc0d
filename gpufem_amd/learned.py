"""Learned warm start for the Brick8 cantilever family.

The solver generates its own training data: for members k of the family
(gpufem_amd.problems.cantilever_problem: random modulus field, random top
traction, x = 0 clamped) the grouped FP64 Jacobi-PCG computes displacement
fields; the network of libgpufem (gf_nn_*, a BF16 3D U-Net on the node grid
with hand-written HIP forward, backward and Adam kernels) learns
features -> displacement; its prediction becomes the PCG initial guess.
Zero-start PCG is the fallback and the yardstick: every warm-started solve
must reach the same true residual ||b - K u|| <= rtol ||b||.

Features per node (16 channels, bf16):
  0      log(E / 210000) / sigma, averaged over the node's elements
  1..3   nodal force / s_f  (s_f = RMS of the force on the top-face nodes)
  4      1 on the clamped face, 0 elsewhere (the prediction is zeroed there)
  5..7   x / nx, y / ny, z / nz
  8..10  beam-theory displacement estimate (x, y, z) / (s_f c_beam): the
         member's loads reduced to shear forces and bending moments along x,
         each cross-section's bending and shear stiffness from its moduli,
         integrated from the clamp (Timoshenko beam, plane sections); this
         carries the non-local clamp-to-tip dependence the convolutions'
         receptive field cannot reach
  11..13 1 on the faces y = 0 or ny, z = 0 or nz, x = nx
  14..15 zero
Target: the correction to the beam estimate, (u - beam) / (s_f c_u), with
beam the estimate as carried (bf16) in channels 8..10 and c_u the RMS of
(u - beam) / s_f over the training set; the output layer starts at zero, so
an untrained network reproduces the beam estimate.  Prediction:
x0 = s_f c_u (pred + (c_beam / c_u) feat[8..10]), zero on the clamped face.

Usage (trains, evaluates, writes the weights file):
  python -m gpufem_amd.learned --train 2000 --holdout 256 --epochs 20 --out W.npz
"""
import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import DevBuf, Net, _ptr, problems
from ._lib import GF_OK  # noqa: F401

MASK_CHANNEL = 4
NCH = 16
FEATURES = "v3-beam-correction"  # feature set and target the weights were trained on
BASE_CHANNEL = 8  # first of the three beam-estimate channels
HOLDOUT_BASE = 1_000_000  # member index of the first held-out sample


def bf16_bits(a):
    """float32 -> bfloat16 bit patterns (round to nearest even)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


class Family:
    """Geometry shared by the members of one cantilever family."""

    def __init__(self, nx=64, ny=32, nz=32):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.grid = (nx + 1, ny + 1, nz + 1)  # node grid (x, y, z)
        self.nodes = (nx + 1) * (ny + 1) * (nz + 1)
        self.template = problems.cantilever_problem(0, nx, ny, nz)
        z, y, x = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1),
                              indexing="ij")
        base = np.zeros((nz + 1, ny + 1, nx + 1, NCH), dtype=np.float32)
        base[..., MASK_CHANNEL] = (x == 0)
        base[..., 5] = x / nx
        base[..., 6] = y / ny
        base[..., 7] = z / nz
        self.base = base
        cnt = np.zeros((nz + 1, ny + 1, nx + 1))
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    cnt[dz:dz + nz, dy:dy + ny, dx:dx + nx] += 1.0
        self.node_count = cnt
        self.base_flat = base.reshape(-1, NCH)  # the constant channels, [nodes, 16]
        base[..., 11] = (y == 0) | (y == ny)
        base[..., 12] = (z == 0) | (z == nz)
        base[..., 13] = (x == nx)
        top = np.zeros((nz + 1, ny + 1, nx + 1), dtype=bool)
        top[nz] = True
        self.top = top.reshape(-1)
        self._zc = np.arange(nz + 1) - 0.5 * nz  # node offsets from the section centre
        self._yc = np.arange(ny + 1) - 0.5 * ny
        t = self.template
        w = self.beam_estimate(t.elem_E, t.f)
        self.c_beam = float(np.sqrt(np.mean(w ** 2))) / self.force_scale(t.f)

    def problem(self, k):
        return problems.cantilever_problem(k, self.nx, self.ny, self.nz, template=self.template)

    def node_log_modulus(self, elem_E):
        g = np.log(np.asarray(elem_E).reshape(self.nz, self.ny, self.nx) / problems.E_DEFAULT)
        s = np.zeros((self.nz + 1, self.ny + 1, self.nx + 1))
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    s[dz:dz + self.nz, dy:dy + self.ny, dx:dx + self.nx] += g
        return s / self.node_count / problems.CANTILEVER_SIGMA

    def force_scale(self, f):
        F = np.asarray(f).reshape(-1, 3)
        return float(np.sqrt(np.mean(F[self.top] ** 2)))

    def beam_estimate(self, elem_E, f):
        """Timoshenko-beam displacement [nz+1, ny+1, nx+1, 3] of the member
        with moduli elem_E and nodal forces f, clamped at x = 0: per slab
        (element column i) the shear force V and bending moment M of the
        forces beyond the slab centre, curvature M / (E_b I) with E_b the
        slab's moduli weighted by the squared distance from the section's
        centre, shear strain V / (k G A) with the slab's mean modulus; slopes
        and deflections summed from the clamp; u_x from the axial force and
        from the rotated plane sections."""
        nx, ny, nz = self.nx, self.ny, self.nz
        F = np.asarray(f).reshape(nz + 1, ny + 1, nx + 1, 3)
        q = F.sum(axis=(0, 1))  # nodal force resultant per x station
        E = np.asarray(elem_E).reshape(nz, ny, nx)
        zc = np.arange(nz) + 0.5 - 0.5 * nz
        yc = np.arange(ny) + 0.5 - 0.5 * ny
        Ebz = (E * (zc ** 2)[:, None, None]).sum(axis=(0, 1)) / (ny * (zc ** 2).sum())
        Eby = (E * (yc ** 2)[None, :, None]).sum(axis=(0, 1)) / (nz * (yc ** 2).sum())
        Em = E.mean(axis=(0, 1))
        A, Iy, Iz = ny * nz, ny * nz ** 3 / 12.0, nz * ny ** 3 / 12.0
        G = Em / (2.0 * (1.0 + problems.NU_DEFAULT))
        kap = 5.0 / 6.0
        xs = np.arange(nx + 1, dtype=float)
        mid = np.arange(nx) + 0.5
        beyond = xs[None, :] > mid[:, None]                       # [slab, station]
        V = beyond.astype(float) @ q                              # [slab, 3]
        M = (beyond * (xs[None, :] - mid[:, None])) @ q           # moment arms
        kz, ky = M[:, 2] / (Ebz * Iy), M[:, 1] / (Eby * Iz)        # curvatures
        gz, gy = V[:, 2] / (kap * G * A), V[:, 1] / (kap * G * A)  # shear strains
        ex = V[:, 0] / (Em * A)                                   # axial strain
        thz = np.concatenate([[0.0], np.cumsum(kz)])              # slopes at stations
        thy = np.concatenate([[0.0], np.cumsum(ky)])
        wz = np.concatenate([[0.0], np.cumsum(0.5 * (thz[:-1] + thz[1:]) + gz)])
        wy = np.concatenate([[0.0], np.cumsum(0.5 * (thy[:-1] + thy[1:]) + gy)])
        ua = np.concatenate([[0.0], np.cumsum(ex)])
        out = np.zeros((nz + 1, ny + 1, nx + 1, 3))
        out[..., 0] = (ua[None, None, :] - self._zc[:, None, None] * thz[None, None, :]
                       - self._yc[None, :, None] * thy[None, None, :])
        out[..., 1] = wy[None, None, :]
        out[..., 2] = wz[None, None, :]
        return out

    def features(self, p):
        """(bf16 bits [nodes * 16], s_f) of problem p."""
        x = self.base.copy()
        x[..., 0] = self.node_log_modulus(p.elem_E)
        s_f = self.force_scale(p.f)
        x[..., 1:4] = (p.f / s_f).reshape(self.nz + 1, self.ny + 1, self.nx + 1, 3)
        x[..., 8:11] = self.beam_estimate(p.elem_E, p.f) / (s_f * self.c_beam)
        return bf16_bits(x).reshape(-1), s_f

    def beam_dofs(self, feats, s_f):
        """The beam estimate as the features carry it (bf16), in DOF units
        [nodes, 3]."""
        f = np.asarray(feats, dtype=np.uint16).reshape(self.nodes, NCH)[:, BASE_CHANNEL:BASE_CHANNEL + 3]
        return (f.astype(np.uint32) << 16).view(np.float32).astype(np.float64) * (s_f * self.c_beam)

    def target(self, u, feats, s_f, c_u):
        t = np.zeros((self.nodes, 4), dtype=np.float32)
        t[:, :3] = (np.asarray(u).reshape(-1, 3) - self.beam_dofs(feats, s_f)) / (s_f * c_u)
        return t.reshape(-1)

    def base_weight(self, c_u):
        return self.c_beam / c_u

    def x0(self, pred, feats, s_f, c_u):
        """Host version of gf_nn_to_dofs: the DOF vector from one sample's
        prediction rows pred [nodes, >= 3] and its features."""
        x = s_f * c_u * np.asarray(pred)[:, :3].astype(np.float64) + self.beam_dofs(feats, s_f)
        x[self.base_flat[:, MASK_CHANNEL] > 0.5] = 0.0
        return x.reshape(-1)


def log_default(msg):
    print("[learned] " + msg, file=sys.stderr, flush=True)


def solve_members(ctx, fam, ks, group_size=16, rtol=1e-6, log=log_default, workers=8):
    """Displacements of members ks from the grouped FP64 Jacobi-PCG, group_size
    members per launch sequence.  The group's structure is built once and
    only the moduli change between groups; the next group's problems are
    generated on host threads while the GPU solves.  Yields (k, problem, u,
    iterations)."""
    from concurrent.futures import ThreadPoolExecutor
    ks = list(ks)
    chunks = [ks[i:i + group_size] for i in range(0, len(ks), group_size)]
    if not chunks:
        return
    pool = ThreadPoolExecutor(max_workers=workers)
    futs = [pool.submit(fam.problem, k) for k in chunks[0]]
    G = None
    try:
        for ci, chunk in enumerate(chunks):
            probs = [f.result() for f in futs]
            if ci + 1 < len(chunks):
                futs = [pool.submit(fam.problem, k) for k in chunks[ci + 1]]
            t0 = time.perf_counter()
            if G is None or len(G.problems) != len(probs):
                if G is not None:
                    G.free()
                G = ctx.group(probs)
            else:
                G.problems = probs
                G.set_E([p.elem_E for p in probs])
            G.assemble()
            G.apply_penalty()
            x, st, its = G.pcg(np.concatenate([p.f for p in probs]), rtol=rtol)
            log("solved members %d..%d: %d its max, %.2f s" % (
                chunk[0], chunk[-1], int(its.max()), time.perf_counter() - t0))
            for k, p, u, it in zip(chunk, probs, G.split(x), its):
                yield k, p, u, int(it)
    finally:
        pool.shutdown(wait=True)
        if G is not None:
            G.free()


def cosine_lr(step, total, lr0, lr1, warmup=0):
    """Linear warm-up over `warmup` steps, then cosine decay from lr0 to lr1.
    (Without normalization layers, full-size Adam steps on a fresh network
    move whole layers coherently and blow the activations up.)"""
    if step < warmup:
        return lr0 * (step + 1) / warmup
    s, T = step - warmup, max(total - warmup, 1)
    return lr1 + 0.5 * (lr0 - lr1) * (1.0 + np.cos(np.pi * min(s, T) / T))


def train(ctx, fam, feats, targets, epochs, batch, lr0=5e-4, lr1=5e-6, seed=0x5EED0008,
          net=None, log=log_default, warmup=1000):
    """Adam on the GPU over the dataset (feats [N, nodes*16] uint16, targets
    [N, nodes*4] float32), batch order reshuffled every epoch.  Returns the
    network and the per-epoch mean losses."""
    N = feats.shape[0]
    nb = N // batch
    if nb < 1:
        raise ValueError("fewer samples than one batch")
    if net is None:
        net = Net(ctx, *fam.grid, batch)
        # output layer from zero: the untrained prediction is the base estimate
        p = net.params()
        w_off = net.layers()[-1][6]
        p[w_off:] = 0.0
        net.set_params(p)
    in_bytes, t_bytes = fam.nodes * NCH * 2, fam.nodes * 4 * 4
    d_in = DevBuf(ctx, nb * batch * in_bytes)
    d_t = DevBuf(ctx, nb * batch * t_bytes)
    for i in range(nb * batch):
        d_in.upload(feats[i], i * in_bytes)
        d_t.upload(targets[i], i * t_bytes)
    rng = np.random.default_rng(seed)
    total = epochs * nb
    step, history = 0, []
    for ep in range(epochs):
        t0 = time.perf_counter()
        losses = []
        for s in rng.permutation(nb):
            lr = cosine_lr(step, total, lr0, lr1, warmup)
            losses.append(net.train_step(d_in.at(s * batch * in_bytes), d_t.at(s * batch * t_bytes),
                                         batch, lr))
            step += 1
        history.append(float(np.mean(losses)))
        log("epoch %d: loss %.4e (%.1f s)" % (ep, history[-1], time.perf_counter() - t0))
    d_in.free()
    d_t.free()
    return net, history


def evaluate(ctx, fam, net, c_u, ks, rtol=1e-8, log=log_default, log_every=1):
    """Zero-start vs warm-start FP64 Jacobi-PCG on members ks (device-resident
    vectors).  Returns per-member records (every log_every-th one logged)."""
    lib = ctx.lib
    p0 = fam.problem(ks[0])
    mesh = ctx.upload_mesh(p0.elem_type, p0.ngauss, p0.conn, p0.coords, p0.E, p0.nu,
                           elem_E=p0.elem_E)
    K = ctx.symbolic(mesh)
    n = p0.ndof
    d_b, d_x = DevBuf(ctx, n * 8), DevBuf(ctx, n * 8)
    d_in = DevBuf(ctx, fam.nodes * NCH * 2)
    d_pred = DevBuf(ctx, net.batch_max * fam.nodes * NCH * 4)
    out = []
    for k in ks:
        p = fam.problem(k)
        eE = np.ascontiguousarray(p.elem_E, dtype=np.float64)
        ctx._check(lib.gf_mesh_set_E(ctx.handle, mesh.handle, _ptr(eE)))
        ctx.assemble(mesh, K)
        ctx.apply_penalty(K, p.fixed_dofs)
        d_b.upload(np.ascontiguousarray(p.f, dtype=np.float64))
        st0 = _pcg_dev(ctx, K, d_b, d_x, 0, rtol)
        u0 = d_x.download(np.float64)
        feats, s_f = fam.features(p)
        d_in.upload(feats)
        t0 = time.perf_counter()
        net.forward(d_in.ptr, 1, d_pred.ptr)
        net.to_dofs(d_pred.ptr, d_in.ptr, 0, MASK_CHANNEL, s_f * c_u, d_x.ptr,
                    base_channel=BASE_CHANNEL, base_weight=fam.base_weight(c_u))
        x0 = d_x.download(np.float64)  # also synchronises
        t_inf = time.perf_counter() - t0
        # the guarded guess: kept only when its residual is below ||b||
        st1 = _pcg_dev(ctx, K, d_b, d_x, 2, rtol)
        u1 = d_x.download(np.float64)
        rec = {
            "k": int(k),
            "zero_it": int(st0.iterations), "zero_ms": float(st0.ms_device),
            "zero_true_relres": float(st0.true_relres),
            "warm_it": int(st1.iterations), "warm_ms": float(st1.ms_device),
            "warm_true_relres": float(st1.true_relres),
            "warm_x0_rejected": int(st1.x0_rejected),
            "inference_ms": 1e3 * t_inf,
            "x0_rel_err": float(np.linalg.norm(x0 - u0) / np.linalg.norm(u0)),
            # the untrained network's guess (the beam estimate) for comparison
            "base_rel_err": float(np.linalg.norm(fam.x0(np.zeros((fam.nodes, 3)), feats, s_f, c_u)
                                                 - u0) / np.linalg.norm(u0)),
            "solution_rel_diff": float(np.linalg.norm(u1 - u0) / np.linalg.norm(u0)),
        }
        out.append(rec)
        if (len(out) - 1) % log_every == 0 or len(out) == len(ks):
            log("member %d: zero %d it %.1f ms | warm %d it %.1f ms + inference %.2f ms | "
                "x0 err %.2e" % (k, rec["zero_it"], rec["zero_ms"], rec["warm_it"],
                                 rec["warm_ms"], rec["inference_ms"], rec["x0_rel_err"]))
    return out


def _pcg_dev(ctx, K, d_b, d_x, use_x0, rtol):
    from . import _lib, solve_opts
    o = solve_opts(device_ptrs=1, use_x0=use_x0, rtol=rtol)
    st = _lib.SolveStats()
    ctx._check(ctx.lib.gf_pcg(ctx.handle, K.handle, d_b.ptr, d_x.ptr, 1, C.byref(o), C.byref(st)))
    return st


def summarize(recs):
    z_it = np.array([r["zero_it"] for r in recs], dtype=float)
    w_it = np.array([r["warm_it"] for r in recs], dtype=float)
    z_ms = np.array([r["zero_ms"] for r in recs])
    w_ms = np.array([r["warm_ms"] + r["inference_ms"] for r in recs])
    return {
        "members": len(recs),
        "zero_iterations_mean": float(z_it.mean()), "warm_iterations_mean": float(w_it.mean()),
        "iteration_ratio": float(w_it.sum() / z_it.sum()),
        "zero_ms_mean": float(z_ms.mean()), "warm_plus_inference_ms_mean": float(w_ms.mean()),
        "time_ratio": float(w_ms.sum() / z_ms.sum()),
        "max_true_relres_zero": float(max(r["zero_true_relres"] for r in recs)),
        "max_true_relres_warm": float(max(r["warm_true_relres"] for r in recs)),
        "x0_rel_err_mean": float(np.mean([r["x0_rel_err"] for r in recs])),
        "base_rel_err_mean": float(np.mean([r.get("base_rel_err", np.nan) for r in recs])),
        "max_solution_rel_diff": float(max(r["solution_rel_diff"] for r in recs)),
        "warm_x0_rejected": int(sum(r.get("warm_x0_rejected", 0) for r in recs)),
    }


def save_weights(path, net, c_u, meta):
    np.savez(path, params=net.params(), c_u=np.float64(c_u),
             grid=np.array(net_grid(net), dtype=np.int32),
             meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))


def net_grid(net):
    return (net.nx, net.ny, net.nz)


def load_weights(ctx, path, batch_max=1):
    z = np.load(path, allow_pickle=False)
    nx, ny, nz = (int(v) for v in z["grid"])
    net = Net(ctx, nx, ny, nz, batch_max)
    net.set_params(z["params"])
    meta = json.loads(bytes(z["meta"]).decode())
    if meta.get("features") != FEATURES:
        raise ValueError("%s was trained on features %r, this code builds %r"
                         % (path, meta.get("features"), FEATURES))
    return net, float(z["c_u"]), meta


def main(argv=None):
    import gpufem_amd as gf
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--holdout", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--group", type=int, default=16)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--out", default="learned_brick8_cantilever.npz")
    ap.add_argument("--report", default="learned_report.json")
    args = ap.parse_args(argv)
    ctx = gf.Context(0)
    fam = Family()
    t0 = time.perf_counter()
    feats = np.empty((args.train, fam.nodes * NCH), dtype=np.uint16)
    rs, scales, its = [], [], []
    for i, (k, p, u, it) in enumerate(solve_members(ctx, fam, range(args.train), args.group)):
        feats[i], s_f = fam.features(p)
        rs.append(u.astype(np.float32))
        scales.append(s_f)
        its.append(it)
    t_data = time.perf_counter() - t0
    # output scale: RMS of the correction to the beam estimate, per unit force
    c_u = float(np.sqrt(np.mean([np.mean(((u.reshape(-1, 3) - fam.beam_dofs(f, s)) / s) ** 2)
                                 for u, f, s in zip(rs, feats, scales)])))
    targets = np.stack([fam.target(u, f, s, c_u) for u, f, s in zip(rs, feats, scales)])
    del rs
    log_default("dataset: %d members in %.1f s, c_u %.4e, mean its %.0f" % (
        args.train, t_data, c_u, np.mean(its)))
    t0 = time.perf_counter()
    net, hist = train(ctx, fam, feats, targets, args.epochs, args.batch, lr0=args.lr)
    t_train = time.perf_counter() - t0
    meta = {"family": "brick8_cantilever_64x32x32", "features": FEATURES,
            "train_members": [0, args.train],
            "epochs": args.epochs, "batch": args.batch, "loss_history": hist,
            "train_seconds": t_train, "data_seconds": t_data}
    save_weights(args.out, net, c_u, meta)
    eval_net = Net(ctx, *fam.grid, 1)
    eval_net.set_params(net.params())
    net.free()
    ks = list(range(HOLDOUT_BASE, HOLDOUT_BASE + args.holdout))
    recs = evaluate(ctx, fam, eval_net, c_u, ks)
    rep = {"summary": summarize(recs), "meta": meta, "members": recs}
    with open(args.report, "w") as fh:
        json.dump(rep, fh, indent=1)
    print(json.dumps(rep["summary"]), flush=True)


if __name__ == "__main__":
    main()
