"""Benchmark of the assembly-and-solve path on one MI355X.

A step is one pass of the hot path over the workload with inputs resident in
HBM: numeric assembly (Ke + colored scatter into ELL), the penalty constraint,
and a Jacobi-PCG solve from zero to ||b - K u|| <= 1e-8 ||b||.

The default workload is the largest single-GPU configuration, the Brick8
cube of 128^3 elements (6,440,067 DOF, 2x2x2 Gauss points, lognormal
per-element moduli).  `value` is assembled DOF/s: DOF / numeric assembly
time, the time covering every copy of the values the solver reads (the node
gather writes the block-sliced copy beside the ELL values).  The `solve`
object reports the PCG to 1e-8; `roofline` reports the SpMV kernel the
solver runs every iteration; `cpu_baseline` times the FP64 CPU restatement of
the reference OpenMP path on the host cores; `forms` times the mixed,
batched, grouped and learned forms against plain FP64 PCG.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "assembled DOF/s + PCG time-to-1e-8 resid (mixed/batched/learned vs plain FP64)"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def build_problem(name):
    from gpufem_amd import problems as P
    try:
        return P.by_name(name)
    except ValueError as e:
        raise SystemExit(str(e))


def ke_flops_per_element(dim, nen, ngp):
    """Dense-convention Ke FLOPs: ngp (2 s^2 ne + 2 s ne^2)."""
    s = 3 if dim == 2 else 6
    ne = dim * nen
    return ngp * (2 * s * s * ne + 2 * s * ne * ne)


def cpu_baseline(p, gpu_iters, budget_s, full=False):
    """FP64 restatement of the reference OpenMP path on the host cores.  The
    PCG is a bounded sample extrapolated to the GPU's iteration count, or
    (full) the whole solve to 1e-8, reporting its progress as it goes."""
    import oracle as O
    threads = O.default_threads()
    t0 = time.perf_counter()
    K, t_full = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E,
                           skip_zero=False, nthreads=threads)
    # numeric re-assembly: structure already present
    O.load().orc_ell_zero_values(K.m)
    _, t_num = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E,
                          skip_zero=False, nthreads=threads, into=K)
    O.apply_penalty(K, p.dim, p.supports)
    if full:
        _, it, rel, t_pcg = O.pcg_jacobi(K, p.f, rtol=1e-8, maxit=1000000, nthreads=threads,
                                         progress_every=100)
        return {
            "value": p.ndof / t_num,
            "unit": "DOF/s",
            "cores": threads,
            "kind": "port",
            "sample": ("full %s: colored OpenMP assembly (numeric pass timed, %.3f s; with "
                       "insertion %.3f s) + Jacobi-PCG from zero to 1e-8 measured whole: %d "
                       "iterations, %.1f s" % (p.name, t_num, t_full, it, t_pcg)),
            "assembly_s": t_num,
            "pcg_s_per_iter": t_pcg / max(it, 1),
            "pcg_to_1e8_s": t_pcg,
            "pcg_iterations": it,
            "pcg_true_relres": rel,
        }
    # bounded PCG sample: iterations until about half the remaining budget is spent
    spent = time.perf_counter() - t0
    probe_it = 5
    _, _, _, t_probe = O.pcg_jacobi(K, p.f, rtol=0.0, maxit=probe_it, nthreads=threads)
    per_it = t_probe / probe_it
    n_it = int(max(5, min(2000, (budget_s - spent) / 2 / max(per_it, 1e-6))))
    _, it, _, t_pcg = O.pcg_jacobi(K, p.f, rtol=0.0, maxit=n_it, nthreads=threads)
    per_it = t_pcg / max(it, 1)
    return {
        "value": p.ndof / t_num,
        "unit": "DOF/s",
        "cores": threads,
        "kind": "port",
        "sample": ("full %s: colored OpenMP assembly (numeric pass timed, %.3f s; with "
                   "insertion %.3f s) + %d Jacobi-PCG iterations (%.4f s/iter); PCG to 1e-8 "
                   "extrapolated to the GPU iteration count %d" % (p.name, t_num, t_full, it,
                                                                   per_it, gpu_iters)),
        "assembly_s": t_num,
        "pcg_s_per_iter": per_it,
        "pcg_to_1e8_s_extrapolated": per_it * gpu_iters,
    }


COPY_BYTES = 1 << 30


def probe(ctx, lib, p, K, variant, nrep=20):
    """nrep launches of the solver's SpMV kernel, then nrep of a plain copy of
    COPY_BYTES (read COPY_BYTES, write COPY_BYTES): the copy calibrates the
    per-dispatch byte counters of the same run."""
    import gpufem_amd as gf
    ms_spmv = C.c_double()
    gf.check(lib.gf_profile_spmv(ctx.handle, K.handle, nrep, C.byref(ms_spmv)), ctx.handle)
    src, dst = gf.DevBuf(ctx, COPY_BYTES), gf.DevBuf(ctx, COPY_BYTES)
    ms_copy = C.c_double()
    gf.check(lib.gf_stream_copy(ctx.handle, dst.ptr, src.ptr, COPY_BYTES, nrep,
                                C.byref(ms_copy)), ctx.handle)
    print(json.dumps({"workload": p.name, "spmv_variant": variant, "nnz": K.nnz, "n": p.ndof,
                      "spmv_ms": ms_spmv.value, "spmv_launches": nrep + 1,
                      "copy_bytes_each_way": COPY_BYTES, "copy_ms": ms_copy.value,
                      "copy_launches": nrep + 1,
                      "copy_gbs": 2 * COPY_BYTES / (ms_copy.value * 1e-3) / 1e9}), flush=True)


def measured_traffic(workload, variant):
    """HBM bytes per SpMV launch from the committed hardware-counter summary
    (two rocprofv3 --pmc runs of `bench.py --probe`, FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a copy kernel of
    known size), when it covers this workload and SpMV variant; else None."""
    path = os.path.join(ROOT, "profiles", "spmv_%s_hbm_bytes.md" % workload)
    if not os.path.exists(path):
        return None
    kv = {}
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_") and ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
    if kv.get("traffic_workload") != workload or kv.get("traffic_spmv_variant") != str(variant):
        return None
    return int(kv["traffic_bytes_per_launch"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="brick8_cube_128")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the host PCG to 1e-8 whole instead of a bounded sample")
    ap.add_argument("--schedule", default="graph",
                    choices=["kernels", "graph", "persistent"])
    ap.add_argument("--compare", default="all", choices=["all", "none"],
                    help="also time the other FP64 schedules and the mixed solve")
    ap.add_argument("--forms", default="mixed,mixed_plate,batched,grouped,learned,combined",
                    help="comma list of forms timed against plain FP64 on their own "
                         "workloads (mixed, batched, grouped, learned) or 'none'")
    ap.add_argument("--weights", default=os.path.join(ROOT, "gpufem_amd", "weights",
                                                      "brick8_cantilever_64x32x32.npz"))
    ap.add_argument("--probe", action="store_true",
                    help="only launch the solver's SpMV kernel and a plain copy kernel of known "
                         "size (for hardware-counter runs); prints one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: --gpus %d is not supported (the path would "
              "shard by mesh slabs with a halo exchange per SpMV; not built)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import gpufem_amd as gf
    lib = gf.load()
    log("building %s" % args.workload)
    p = build_problem(args.workload)
    n = p.ndof
    ctx = gf.Context(0)
    ncu, arch = ctx.device_info()
    log("device %s, %d CUs; %d DOF, %d elements" % (arch, ncu, n, p.grid.nelem))
    mesh = ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E)
    t0 = time.perf_counter()
    K = ctx.symbolic(mesh)
    t_sym = time.perf_counter() - t0
    log("symbolic: width %d, nnz %d, %d colors, %.3f s" % (K.width, K.nnz, K.ncolors, t_sym))
    # the same structure built on the device (timed after a warm-up build)
    try:
        ctx.symbolic(mesh, device=True).free()
        t0 = time.perf_counter()
        Kd = ctx.symbolic(mesh, device=True)
        t_sym_dev = time.perf_counter() - t0
        sym_dev_rounds = (Kd.symbolic_timing.launches - 12) // 2
        Kd.free()
        log("symbolic on the device: %.3f s (%d coloring rounds)" % (t_sym_dev, sym_dev_rounds))
    except gf.GpufemError as e:
        log("symbolic on the device: %s" % e)
        t_sym_dev, sym_dev_rounds = float("nan"), -1
    asm_variant, asm_variant_ms = ctx.assemble_tune(mesh, K)
    log("assembly kernels (ms): %s -> %d" % (" ".join("%.3f" % v for v in asm_variant_ms),
                                             asm_variant))
    ctx.apply_penalty(K, p.fixed_dofs)
    variant, variant_ms = ctx.tune(K)
    log("spmv variants (ms): %s -> %d" % (" ".join("%.4f" % v for v in variant_ms), variant))
    sched = {"kernels": gf.GF_SCHED_KERNELS, "graph": gf.GF_SCHED_GRAPH,
             "persistent": gf.GF_SCHED_PERSISTENT}[args.schedule]
    if args.probe:
        probe(ctx, lib, p, K, variant)
        return

    # inputs resident in HBM
    vb = n * 8
    d_b, d_x = C.c_void_p(), C.c_void_p()
    gf.check(lib.gf_dev_alloc(ctx.handle, vb, C.byref(d_b)), ctx.handle)
    gf.check(lib.gf_dev_alloc(ctx.handle, vb, C.byref(d_x)), ctx.handle)
    f = np.ascontiguousarray(p.f)
    gf.check(lib.gf_copy_h2d(ctx.handle, d_b, f.ctypes.data_as(C.c_void_p), vb), ctx.handle)
    opts = gf.solve_opts(device_ptrs=1, rtol=1e-8, schedule=sched)
    dofs = np.ascontiguousarray(p.fixed_dofs)

    def solve(o):
        st = gf._lib.SolveStats()
        gf.check(lib.gf_pcg(ctx.handle, K.handle, d_b, d_x, 1, C.byref(o), C.byref(st)),
                 ctx.handle)
        return st

    def step():
        ta = ctx.assemble(mesh, K)
        ctx.apply_penalty(K, dofs)
        return ta, solve(opts)

    for i in range(args.warmup):
        ta, st = step()
        log("warmup %d: assembly %.3f ms, pcg %d it %.1f ms" % (i, ta.ms_device, st.iterations,
                                                                  st.ms_device))
    asm_ms, pcg_ms, iters, relres = [], [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ta, st = step()
        asm_ms.append(ta.ms_device)
        pcg_ms.append(st.ms_device)
        iters.append(st.iterations)
        relres.append(st.true_relres)
        log("step %d: assembly %.3f ms, pcg %d it %.1f ms, true relres %.2e" % (
            i, ta.ms_device, st.iterations, st.ms_device, st.true_relres))
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    t_asm = float(np.mean(asm_ms)) / 1e3
    value = n / t_asm

    solves = {"fp64_" + args.schedule: {
        "iterations": float(np.mean(iters)), "seconds": float(np.mean(pcg_ms)) / 1e3,
        "true_relres": float(np.max(relres))}}
    if args.compare == "all":
        # each comparison is capped at a few times the FP64 iteration count
        cap = int(4 * max(iters)) + 100
        for name, o in [
                ("fp64_graph", gf.solve_opts(device_ptrs=1, rtol=1e-8, maxit=cap,
                                             schedule=gf.GF_SCHED_GRAPH)),
                ("fp64_kernels", gf.solve_opts(device_ptrs=1, rtol=1e-8, maxit=cap,
                                               schedule=gf.GF_SCHED_KERNELS)),
                ("fp64_persistent", gf.solve_opts(device_ptrs=1, rtol=1e-8, maxit=cap,
                                                  schedule=gf.GF_SCHED_PERSISTENT)),
                ("mixed", gf.solve_opts(device_ptrs=1, rtol=1e-8, maxit=cap,
                                        precision=gf.GF_PREC_MIXED))]:
            if name in solves:
                continue
            log("%s: solving ..." % name)
            try:
                st = solve(o)
            except gf.GpufemError as e:
                log("%s: %s" % (name, e))
                solves[name] = {"error": str(e)}
                continue
            solves[name] = {"iterations": st.iterations, "seconds": st.ms_device / 1e3,
                            "true_relres": st.true_relres, "refinements": st.refinements,
                            "fallbacks": st.fallbacks}
            log("%s: %d it, %.3f s, true relres %.2e" % (name, st.iterations,
                                                         st.ms_device / 1e3, st.true_relres))
    if args.compare == "all" and "error" not in solves.get("fp64_persistent", {"error": 1}):
        # what the persistent schedule's three grid barriers per iteration cost
        us, nbp = ctx.profile_barrier(K)
        solves["persistent_barrier"] = {"us_per_barrier": us, "workgroups": nbp,
                                        "us_per_iteration": 3 * us}
        log("persistent schedule: %.2f us per grid barrier (%d workgroups)" % (us, nbp))
    if args.compare == "all":
        # the elimination alternative to the penalty, same mesh and load
        ctx.assemble(mesh, K)
        ctx.apply_elimination(K, dofs)
        ctx.zero_dofs(d_b, n, dofs)
        for name, o in [
                ("fp64_elimination", gf.solve_opts(device_ptrs=1, rtol=1e-8, maxit=cap,
                                                   schedule=gf.GF_SCHED_GRAPH)),
                ("mixed_elimination", gf.solve_opts(device_ptrs=1, rtol=1e-8, maxit=cap,
                                                    precision=gf.GF_PREC_MIXED))]:
            log("%s: solving ..." % name)
            try:
                st = solve(o)
            except gf.GpufemError as e:
                log("%s: %s" % (name, e))
                solves[name] = {"error": str(e)}
                continue
            solves[name] = {"iterations": st.iterations, "seconds": st.ms_device / 1e3,
                            "true_relres": st.true_relres, "refinements": st.refinements,
                            "fallbacks": st.fallbacks}
            log("%s: %d it, %.3f s, true relres %.2e" % (name, st.iterations,
                                                         st.ms_device / 1e3, st.true_relres))
    # roofline: the solver's SpMV kernel (q = K p with the p.q partials),
    # timed with HIP events around each SpMV of 300 PCG iterations (the
    # kernel as the solve runs it, p rewritten in between); the back-to-back
    # figure (same p every launch) is reported beside it
    ms_spmv_b2b = C.c_double()
    gf.check(lib.gf_profile_spmv(ctx.handle, K.handle, 50, C.byref(ms_spmv_b2b)), ctx.handle)
    ms_spmv_pcg, ms_iter_pcg = ctx.profile_pcg(K, f, 300)
    # algorithmic bytes of the ELL format (values, int32 columns, x, y)
    spmv_bytes = 12 * K.nnz + 16 * n
    achieved = spmv_bytes / (ms_spmv_pcg * 1e-3) / 1e9
    # the least the block-sliced format must move: values, one 16-bit column
    # offset per d x d block, x once, y once
    d = p.dim
    fmt_bytes = 8 * K.nnz + 2 * K.nnz // (d * d) + 16 * n
    achieved_fmt = fmt_bytes / (ms_spmv_pcg * 1e-3) / 1e9
    # attainable streaming rate on this GPU: a plain copy kernel
    cs, cd = gf.DevBuf(ctx, COPY_BYTES), gf.DevBuf(ctx, COPY_BYTES)
    ms_copy = C.c_double()
    gf.check(lib.gf_stream_copy(ctx.handle, cd.ptr, cs.ptr, COPY_BYTES, 20, C.byref(ms_copy)),
             ctx.handle)
    copy_gbs = 2 * COPY_BYTES / (ms_copy.value * 1e-3) / 1e9
    cs.free()
    cd.free()
    log("spmv %.4f ms in the solve, %.4f ms back to back (%.0f GB/s algorithmic); "
        "copy kernel %.0f GB/s" % (ms_spmv_pcg, ms_spmv_b2b.value, achieved, copy_gbs))
    ne = p.grid.nelem
    ngp = p.ngauss ** p.dim
    asm_flops = ke_flops_per_element(p.dim, p.grid.nen, ngp) * ne
    pcg_iter_bytes = 12 * K.nnz + 104 * n
    it_mean = float(np.mean(iters))
    pcg_s = float(np.mean(pcg_ms)) / 1e3
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": p.name, "elements": ne, "dof": n, "nnz": K.nnz, "ell_width": K.width,
            "ngauss": p.ngauss, "colors": K.ncolors,
            "step": "numeric assembly + penalty + Jacobi-PCG from zero to 1e-8 (FP64)",
        },
        "assembly": {
            "ms_numeric": float(np.mean(asm_ms)),
            "s_symbolic_host": t_sym,
            "dof_per_s_full": n / (t_asm + t_sym),
            "s_symbolic_device": t_sym_dev,
            "coloring_rounds_device": sym_dev_rounds,
            "dof_per_s_full_device_symbolic": n / (t_asm + t_sym_dev),
            "gflops_dense_convention": asm_flops / t_asm / 1e9,
            "fp64_peak_frac": asm_flops / t_asm / 1e12 / FP64_PEAK_TFLOPS,
            "kernel_variant": asm_variant,
            "kernel_variant_ms": [float(v) for v in asm_variant_ms],
        },
        "solve": dict(solves, **{
            "spmv_variant": variant, "spmv_variant_ms": [float(v) for v in variant_ms],
            "fp64_gbs_per_iteration": pcg_iter_bytes * it_mean / pcg_s / 1e9}),
        "roofline": {
            "kernel": "k_spmv_pq",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(p.name, variant),
            "algorithmic_bytes_per_launch": spmv_bytes,
            "format_bytes_per_launch": fmt_bytes,
            "achieved_format_bytes": achieved_fmt,
            "frac_format_bytes": achieved_fmt / HBM_PEAK_GBS,
            "ms_per_launch": ms_spmv_pcg,
            "ms_per_launch_back_to_back": ms_spmv_b2b.value,
            "ms_per_pcg_iteration_timed": ms_iter_pcg,
            "copy_kernel_gbs": copy_gbs,
            "frac_of_copy": achieved / copy_gbs,
        },
        "cpu_baseline": None,
    }
    lib.gf_dev_free(ctx.handle, d_b)
    lib.gf_dev_free(ctx.handle, d_x)
    if args.forms != "none":
        from gpufem_amd import forms
        # the mixed form runs on this workload when it is the Brick8 cube:
        # the matrix is reassembled with the penalty, as a step leaves it
        prepared = {}
        if p.name.startswith("brick8_cube_"):
            ctx.assemble(mesh, K)
            ctx.apply_penalty(K, dofs)
            prepared["mixed"] = (p, mesh, K)
        out["forms"] = forms.run(ctx, [f for f in args.forms.split(",") if f], args.weights,
                                 prepared)
    del K, mesh
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(p, int(round(it_mean)), args.cpu_budget,
                                           full=args.cpu_full)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
