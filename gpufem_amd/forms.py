"""The three forms beside plain FP64 PCG, each timed against plain FP64 PCG
on the same meshes (one GPU, vectors resident in HBM):

  mixed    Brick8 128^3 cube, lognormal E: FP32 inner PCG in FP64 iterative
           refinement vs FP64 Jacobi-PCG
  mixed_plate  the same on the Q8 1024x512 plate (a bending-dominated
           cantilever whose operator is too ill-conditioned for FP32)
  batched  Brick20 48^3 cube, 32 random top-face load cases: one block PCG
           (ELL SpMM per iteration) vs the same load cases one after another
  grouped  256 Q4 plates of uneven sizes (16x8 .. 256x128): one grouped
           assembly + segmented PCG vs plate after plate
  learned  Brick8 64x32x32 cantilevers: PCG from the network's prediction vs
           from zero, on held-out members
  combined the same held-out cantilevers: grouped batch + warm start, and
           mixed precision + warm start (batched inference), each vs plain
           FP64 PCG from zero member after member

Every solve stops at ||b - K u|| <= 1e-8 ||b|| (true residual, checked) or
tighter.  Accuracy bar: a form's displacements must lie within 1e-9
(relative 2-norm) of the FP64 solution converged to a true residual of
1e-12.  Plain FP64 PCG to 1e-8 is the time baseline.  A form whose answer at
1e-8 misses the bar is run again at 1e-9, 1e-10, ... and its speed is quoted
at the loosest tolerance that meets the bar (its operating point); the
answers at 1e-8 are reported beside it.
"""
import ctypes as C
import os
import sys
import time

import numpy as np

from . import GF_PREC_MIXED, DevBuf, _lib, solve_opts
from . import problems as P

RTOL = 1e-8
REF_RTOL = 1e-12   # the converged FP64 reference
ERR_BAR = 1e-9     # displacement accuracy required of every form
# operating-point search: half-decade steps from the reference's 1e-8
OP_RTOLS = (1e-8, 3e-9, 1e-9, 3e-10, 1e-10, 3e-11, 1e-11, 3e-12)


def _log(msg):
    print("[forms] " + msg, file=sys.stderr, flush=True)


def _setup(ctx, p):
    mesh = ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E)
    K = ctx.symbolic(mesh)
    ta = ctx.assemble(mesh, K)
    ctx.apply_penalty(K, p.fixed_dofs)
    ctx.tune(K)
    return mesh, K, ta


SCHEDULE = None  # None: the library's choice; main() can force one


def _solve(ctx, K, d_b, d_x, nrhs=1, **kw):
    if SCHEDULE is not None:
        kw.setdefault("schedule", SCHEDULE)
    kw.setdefault("rtol", RTOL)
    o = solve_opts(device_ptrs=1, **kw)
    st = _lib.SolveStats()
    ctx._check(ctx.lib.gf_pcg(ctx.handle, K.handle, d_b.ptr, d_x.ptr, nrhs, C.byref(o),
                              C.byref(st)))
    return st


def _assembly(p, K, ta, ctx, mesh):
    """Numeric assembly time of each kernel variant (3D: also the FP64
    matrix-core one); the matrix is left assembled by the fastest of the
    tuned variants (0, 1 and, where it fits, the node gather 3), with the
    penalty applied."""
    chosen, ms01 = ctx.assemble_tune(mesh, K)
    variants = {"element_major": float(ms01[0]), "pair_major": float(ms01[1])}
    if ms01[3] > 0:
        variants["node_gather"] = float(ms01[3])
    if p.dim == 3:
        ctx.assemble_tune(mesh, K, 2)
        variants["matrix_core"] = min(ctx.assemble(mesh, K).ms_device for _ in range(2))
        ctx.assemble_tune(mesh, K, chosen)
    t2 = ctx.assemble(mesh, K)
    ctx.apply_penalty(K, p.fixed_dofs)
    ms = min(ta.ms_device, t2.ms_device, *[v for v in ms01 if v > 0])
    return {"ms": ms, "dof_per_s": p.ndof / (ms * 1e-3), "elements": int(p.conn.shape[0]),
            "variant": chosen, "variant_ms": variants}


def _stats(st):
    return {"iterations": int(st.iterations), "seconds": st.ms_device / 1e3,
            "true_relres": float(st.true_relres)}


def _rel_err(U, U_ref):
    """max over load cases of ||u - u_ref|| / ||u_ref|| (rows = load cases)."""
    U = np.atleast_2d(U)
    U_ref = np.atleast_2d(U_ref)
    return max(float(np.linalg.norm(U[k] - U_ref[k]) / max(np.linalg.norm(U_ref[k]), 1e-300))
               for k in range(U.shape[0]))


def operating_point(solve, U_ref, log_name):
    """Runs solve(rtol) for rtol in OP_RTOLS until the displacements meet
    ERR_BAR against U_ref.  Returns (record of the passing run or the last,
    list of every run)."""
    runs = []
    for rtol in OP_RTOLS:
        st, U = solve(rtol)
        err = _rel_err(U, U_ref)
        rec = dict(_stats(st), rtol=rtol, refinements=int(st.refinements),
                   fallbacks=int(st.fallbacks), displacement_rel_err=err, meets_bar=err <= ERR_BAR)
        runs.append(rec)
        _log("%s: rtol %.0e: %d it %.3f s, error %.2e vs the converged FP64 solution%s" % (
            log_name, rtol, st.iterations, st.ms_device / 1e3, err,
            "" if err <= ERR_BAR else " (misses %.0e)" % ERR_BAR))
        if err <= ERR_BAR:
            break
    return runs[-1], runs


def _mixed_on(ctx, p, K, name):
    """Plain FP64 PCG to 1e-8 and the mixed form at its operating point, on
    an assembled and constrained matrix K of problem p."""
    d_b, d_x = DevBuf.from_array(ctx, p.f), DevBuf(ctx, p.ndof * 8)
    st_ref = _solve(ctx, K, d_b, d_x, rtol=REF_RTOL, maxit=1000000)
    u_ref = d_x.download(np.float64)
    st64 = _solve(ctx, K, d_b, d_x)
    u64 = d_x.download(np.float64)
    _log("%s: FP64 %d it %.3f s (reference %d it to %.0e)" % (
        name, st64.iterations, st64.ms_device / 1e3, st_ref.iterations, REF_RTOL))

    def mixed(rtol):
        st = _solve(ctx, K, d_b, d_x, precision=_lib.GF_PREC_MIXED, rtol=rtol)
        return st, d_x.download(np.float64)

    op, runs = operating_point(mixed, u_ref, name)
    r = {"workload": p.name, "dof": p.ndof, "nnz": K.nnz,
         "reference": dict(_stats(st_ref), rtol=REF_RTOL),
         "fp64": dict(_stats(st64), displacement_rel_err=_rel_err(u64, u_ref)),
         "mixed": op, "mixed_runs": runs,
         "speedup": st64.ms_device / max(op["seconds"] * 1e3, 1e-9)}
    d_b.free()
    d_x.free()
    return r


def form_mixed(ctx, n=128, prepared=None):
    """Brick8 n^3 cube with lognormal moduli; `prepared` = (problem, mesh,
    matrix) already assembled with the penalty applied (the benchmark's own
    workload), else built here."""
    if prepared is None:
        p = P.config_brick8_cube(n)
        _log("mixed: %s, %d DOF" % (p.name, p.ndof))
        mesh, K, ta = _setup(ctx, p)
        asm = _assembly(p, K, ta, ctx, mesh)
    else:
        p, mesh, K = prepared
        asm = None
    r = _mixed_on(ctx, p, K, "mixed")
    if asm is not None:
        r["assembly"] = asm
    return r


def form_mixed_plate(ctx, nx=1024, ny=512):
    p = P.config_q8_plate(nx, ny)
    _log("mixed_plate: %s, %d DOF" % (p.name, p.ndof))
    mesh, K, ta = _setup(ctx, p)
    return _mixed_on(ctx, p, K, "mixed_plate")


def form_batched(ctx, n=48, nrhs=32):
    p = P.config_brick20_cube(n)
    F = P.random_top_loads(p, nrhs, P.SEED_BRICK20_RHS)
    _log("batched: %s, %d DOF x %d load cases" % (p.name, p.ndof, nrhs))
    mesh, K, ta = _setup(ctx, p)
    asm = _assembly(p, K, ta, ctx, mesh)
    d_b = DevBuf.from_array(ctx, np.ascontiguousarray(F))
    d_x = DevBuf(ctx, F.nbytes)
    st_ref = _solve(ctx, K, d_b, d_x, nrhs=nrhs, batched=1, rtol=REF_RTOL, maxit=1000000)
    X_ref = d_x.download(np.float64).reshape(nrhs, -1)
    stb = _solve(ctx, K, d_b, d_x, nrhs=nrhs, batched=1)
    Xb = d_x.download(np.float64).reshape(nrhs, -1)
    _log("batched: block PCG %d it %.3f s (reference %d it to %.0e)" % (
        stb.iterations, stb.ms_device / 1e3, st_ref.iterations, REF_RTOL))
    # the baseline: every load case solved on its own by plain FP64 PCG
    seq_s, seq_it, worst = 0.0, [], 0.0
    d_b1, d_x1 = DevBuf(ctx, p.ndof * 8), DevBuf(ctx, p.ndof * 8)
    for k in range(nrhs):
        d_b1.upload(F[k])
        st = _solve(ctx, K, d_b1, d_x1)
        seq_s += st.ms_device / 1e3
        seq_it.append(int(st.iterations))
        xk = d_x1.download(np.float64)
        worst = max(worst, float(np.linalg.norm(Xb[k] - xk) / np.linalg.norm(xk)))
    _log("batched: %d load cases one after another: %.3f s" % (nrhs, seq_s))
    d_b1.free()
    d_x1.free()

    def block(rtol, **kw):
        st = _solve(ctx, K, d_b, d_x, nrhs=nrhs, batched=1, rtol=rtol, **kw)
        return st, d_x.download(np.float64).reshape(nrhs, -1)

    op_b, runs_b = operating_point(block, X_ref, "batched")
    op_m, runs_m = operating_point(lambda rtol: block(rtol, precision=GF_PREC_MIXED), X_ref,
                                   "batched mixed")
    r = {"workload": p.name + "_%drhs" % nrhs, "dof": p.ndof, "nnz": K.nnz, "nrhs": nrhs,
         "assembly": asm,
         "reference": dict(_stats(st_ref), rtol=REF_RTOL),
         "block": dict(_stats(stb), displacement_rel_err=_rel_err(Xb, X_ref)),
         "block_operating_point": op_b, "block_runs": runs_b,
         "block_mixed": op_m, "block_mixed_runs": runs_m,
         "sequential": {"seconds": seq_s, "iterations": seq_it},
         "block_vs_single_max_rel_diff": worst}
    r["speedup_block"] = seq_s / op_b["seconds"]
    r["speedup_block_mixed"] = seq_s / op_m["seconds"]
    r["speedup"] = max(r["speedup_block"], r["speedup_block_mixed"])
    return r


def form_grouped(ctx, count=256):
    sizes = P.q4_batch_sizes(count)
    probs = [P.structured(P.GF_Q4, nx, ny, ngauss=2, name="q4_%dx%d" % (nx, ny))
             for nx, ny in sizes]
    ndof = sum(q.ndof for q in probs)
    _log("grouped: %d Q4 plates, %d DOF in all" % (count, ndof))
    G = ctx.group(probs)
    ta = G.assemble()
    G.apply_penalty()
    x, st, its = G.pcg(rtol=RTOL)
    _log("grouped: one launch sequence: assembly %.3f ms, PCG %d it max %.3f s" % (
        ta.ms_device, st.iterations, st.ms_device / 1e3))
    seq_asm, seq_pcg, worst = 0.0, 0.0, 0.0
    xs = G.split(x)
    for k, q in enumerate(probs):
        mesh = ctx.upload_mesh(q.elem_type, q.ngauss, q.conn, q.coords, q.E, q.nu)
        K = ctx.symbolic(mesh)
        t = ctx.assemble(mesh, K)
        ctx.apply_penalty(K, q.fixed_dofs)
        d_b, d_x = DevBuf.from_array(ctx, q.f), DevBuf(ctx, q.ndof * 8)
        s1 = _solve(ctx, K, d_b, d_x)
        seq_asm += t.ms_device / 1e3
        seq_pcg += s1.ms_device / 1e3
        xk = d_x.download(np.float64)
        worst = max(worst, float(np.linalg.norm(xs[k] - xk) / np.linalg.norm(xk)))
    _log("grouped: plate after plate: assembly %.3f s, PCG %.3f s" % (seq_asm, seq_pcg))
    r = {"workload": "q4_plates_%d_16x8_to_256x128" % count, "dof": ndof, "meshes": count,
         "grouped": {"assembly_s": ta.ms_device / 1e3, "pcg_s": st.ms_device / 1e3,
                     "iterations_max": int(its.max()), "true_relres": float(st.true_relres)},
         "sequential": {"assembly_s": seq_asm, "pcg_s": seq_pcg},
         "grouped_vs_single_max_rel_diff": worst}
    r["speedup"] = (seq_asm + seq_pcg) / (r["grouped"]["assembly_s"] + r["grouped"]["pcg_s"])
    return r


def form_learned(ctx, weights, members=256):
    from . import learned as L
    if not os.path.exists(weights):
        return {"skipped": "no trained weights at %s" % weights}
    net, c_u, meta = L.load_weights(ctx, weights, batch_max=1)
    fam = L.Family()
    ks = list(range(L.HOLDOUT_BASE, L.HOLDOUT_BASE + members))
    _log("learned: %d held-out cantilevers" % members)
    recs = L.evaluate(ctx, fam, net, c_u, ks, rtol=RTOL, log=_log, log_every=32)
    s = L.summarize(recs)
    s["workload"] = "brick8_cantilever_64x32x32"
    s["trained_on"] = meta.get("train_members")
    s["speedup"] = 1.0 / s["time_ratio"]
    return s


def form_combined(ctx, weights, members=64, group=16):
    """Learned warm start + grouped batch, and learned warm start + mixed
    precision, on held-out cantilevers, against plain FP64 PCG from zero to
    1e-8 member after member (same meshes, same true-residual stopping test;
    assembly and inference counted).  Each combination is quoted at the
    loosest tolerance of OP_RTOLS where every member's displacements are
    within ERR_BAR of its FP64 solution converged to REF_RTOL."""
    from . import learned as L
    if not os.path.exists(weights):
        return {"skipped": "no trained weights at %s" % weights}
    net, c_u, meta = L.load_weights(ctx, weights, batch_max=group)
    fam = L.Family()
    ks = list(range(L.HOLDOUT_BASE, L.HOLDOUT_BASE + members))
    probs = [fam.problem(k) for k in ks]
    _log("combined: %d held-out cantilevers, groups of %d" % (members, group))
    p0 = probs[0]
    mesh = ctx.upload_mesh(p0.elem_type, p0.ngauss, p0.conn, p0.coords, p0.E, p0.nu,
                           elem_E=p0.elem_E)
    K = ctx.symbolic(mesh)
    d_b, d_x = DevBuf(ctx, p0.ndof * 8), DevBuf(ctx, p0.ndof * 8)

    def member(p):
        ctx._check(ctx.lib.gf_mesh_set_E(ctx.handle, mesh.handle,
                                         np.ascontiguousarray(p.elem_E).ctypes.data_as(C.c_void_p)))
        ta = ctx.assemble(mesh, K)
        ctx.apply_penalty(K, p.fixed_dofs)
        d_b.upload(np.ascontiguousarray(p.f))
        return ta

    # plain: one member after another to 1e-8, and each member's reference
    plain_s, plain_it, u_ref = 0.0, [], []
    for p in probs:
        ta = member(p)
        st = _solve(ctx, K, d_b, d_x)
        plain_s += (ta.ms_device + st.ms_device) / 1e3
        plain_it.append(int(st.iterations))
        _solve(ctx, K, d_b, d_x, rtol=REF_RTOL, maxit=1000000)
        u_ref.append(d_x.download(np.float64))
    _log("combined: plain %.3f s, %.0f it mean" % (plain_s, np.mean(plain_it)))
    nodes = fam.nodes
    d_in = DevBuf(ctx, group * nodes * L.NCH * 2)
    d_pred = DevBuf(ctx, group * nodes * L.NCH * 4)

    def guesses(chunk):
        fs = []
        for i, p in enumerate(chunk):
            f, s_f = fam.features(p)
            d_in.upload(f, i * nodes * L.NCH * 2)
            fs.append((f, s_f))
        t0 = time.perf_counter()
        net.forward(d_in.ptr, len(chunk), d_pred.ptr)
        pred = d_pred.download(np.float32, len(chunk) * nodes * L.NCH).reshape(len(chunk), nodes,
                                                                             L.NCH)
        x0 = [fam.x0(pred[i], f, s_f, c_u) for i, (f, s_f) in enumerate(fs)]
        return x0, time.perf_counter() - t0

    state = {"G": None}

    def grouped(rtol):
        secs, its, res, err = 0.0, [], 0.0, 0.0
        for g0 in range(0, members, group):
            chunk = probs[g0:g0 + group]
            G = state["G"]
            if G is None or len(G.problems) != len(chunk):
                G = state["G"] = ctx.group(chunk)
            else:
                G.problems = chunk
                G.set_E([p.elem_E for p in chunk])
            ta = G.assemble()
            G.apply_penalty()
            x0, t_inf = guesses(chunk)
            x, st, itv = G.pcg(np.concatenate([p.f for p in chunk]), x0=np.concatenate(x0),
                               rtol=rtol)
            secs += (ta.ms_device + st.ms_device) / 1e3 + t_inf
            its += [int(v) for v in itv]
            res = max(res, float(st.true_relres))
            for i, xi in enumerate(G.split(x)):
                err = max(err, _rel_err(xi, u_ref[g0 + i]))
        return secs, its, res, err

    def mixed(rtol):
        secs, its, res, err = 0.0, [], 0.0, 0.0
        for g0 in range(0, members, group):
            chunk = probs[g0:g0 + group]
            x0, t_inf = guesses(chunk)
            secs += t_inf
            for i, p in enumerate(chunk):
                ta = member(p)
                d_x.upload(np.ascontiguousarray(x0[i]))
                st = _solve(ctx, K, d_b, d_x, use_x0=1, precision=GF_PREC_MIXED, rtol=rtol)
                secs += (ta.ms_device + st.ms_device) / 1e3
                its.append(int(st.iterations))
                res = max(res, float(st.true_relres))
                err = max(err, _rel_err(d_x.download(np.float64), u_ref[g0 + i]))
        return secs, its, res, err

    def at_operating_point(run, name):
        runs = []
        for rtol in OP_RTOLS:
            secs, its, res, err = run(rtol)
            runs.append({"rtol": rtol, "seconds": secs, "iterations_mean": float(np.mean(its)),
                         "max_true_relres": res, "displacement_rel_err": err,
                         "meets_bar": err <= ERR_BAR, "speedup": plain_s / secs})
            _log("combined: %s rtol %.0e: %.3f s, %.0f it mean, worst error %.2e%s" % (
                name, rtol, secs, np.mean(its), err, "" if err <= ERR_BAR else " (misses)"))
            if err <= ERR_BAR:
                break
        return dict(runs[-1], runs=runs)

    g = at_operating_point(grouped, "grouped + warm start")
    m = at_operating_point(mixed, "mixed + warm start")
    return {"workload": "brick8_cantilever_64x32x32_x%d" % members, "members": members,
            "group": group,
            "plain_fp64": {"seconds": plain_s, "iterations_mean": float(np.mean(plain_it)),
                           "rtol": RTOL},
            "grouped_learned": g, "mixed_learned": m,
            "speedup": max(g["speedup"] if g["meets_bar"] else 0.0,
                           m["speedup"] if m["meets_bar"] else 0.0)}


FORMS = {"mixed": form_mixed, "mixed_plate": form_mixed_plate, "batched": form_batched,
         "grouped": form_grouped}


def run(ctx, names, weights, prepared=None):
    """Runs the named forms; prepared = {"mixed": (problem, mesh, matrix)}
    hands a form a workload the caller already assembled."""
    out = {}
    prepared = prepared or {}
    for name in names:
        t0 = time.perf_counter()
        try:
            if name == "learned":
                out[name] = form_learned(ctx, weights)
            elif name == "combined":
                out[name] = form_combined(ctx, weights)
            elif name == "mixed" and "mixed" in prepared:
                out[name] = form_mixed(ctx, prepared=prepared["mixed"])
            else:
                out[name] = FORMS[name](ctx)
        except Exception as e:  # report, keep the other forms
            out[name] = {"error": "%s: %s" % (type(e).__name__, e)}
        out[name]["wall_s"] = time.perf_counter() - t0
        _log("%s done in %.1f s" % (name, out[name]["wall_s"]))
    return out


def main(argv=None):
    """python -m gpufem_amd.forms [names] [--weights W]: time the named forms
    (comma list, default all) on GPU 0 and print one JSON line."""
    import argparse
    import json
    from . import Context
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="?",
                    default="mixed,mixed_plate,batched,grouped,learned,combined")
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                      "weights", "brick8_cantilever_64x32x32.npz"))
    ap.add_argument("--schedule", choices=["auto", "kernels", "graph"], default="auto",
                    help="FP64 solve schedule (kernels: no hipGraph, e.g. under a tracer)")
    a = ap.parse_args(argv)
    global SCHEDULE
    from . import GF_SCHED_GRAPH, GF_SCHED_KERNELS
    SCHEDULE = {"auto": None, "kernels": GF_SCHED_KERNELS, "graph": GF_SCHED_GRAPH}[a.schedule]
    ctx = Context(0)
    print(json.dumps(run(ctx, [n for n in a.names.split(",") if n], a.weights)), flush=True)


if __name__ == "__main__":
    main()
