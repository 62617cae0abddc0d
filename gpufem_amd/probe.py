"""Short launch sequences of one kernel family, for hardware-counter runs.

Each mode builds its workload, launches the kernel of interest a few times,
then launches a plain streaming copy of known size (COPY_BYTES read and
COPY_BYTES written per launch).  The copy's counter readings in the same run
calibrate the per-dispatch byte counters (FETCH_SIZE, WRITE_SIZE), so the
kernel's HBM bytes are counter value x (known bytes / copy counter value).
Prints one JSON line naming what was launched.

  python -m gpufem_amd.probe assembly --workload brick8_cube_128 [--variant V]
  python -m gpufem_amd.probe spmv --workload q8_plate_1024x512
  python -m gpufem_amd.probe spmm [--workload brick20_cube_48 --nrhs 32 --iters 8]
  python -m gpufem_amd.probe summarize RESULTS_DB [RESULTS_DB ...]

`summarize` reads rocprofv3 result databases (SQLite) and prints, per kernel
and counter, the dispatch count and the mean value per dispatch; with
--kernel-trace databases it prints per-kernel call counts and durations.

The product path never imports this module.
"""
import argparse
import ctypes as C
import json
import sys

import numpy as np

from . import Context, DevBuf, GpufemError, check, load, solve_opts
from . import _lib
from . import problems as P

COPY_BYTES = 1 << 30


def _log(msg):
    print("[probe] " + msg, file=sys.stderr, flush=True)


def _copy(ctx, lib, nrep):
    src, dst = DevBuf(ctx, COPY_BYTES), DevBuf(ctx, COPY_BYTES)
    ms = C.c_double()
    check(lib.gf_stream_copy(ctx.handle, dst.ptr, src.ptr, COPY_BYTES, nrep, C.byref(ms)),
          ctx.handle)
    src.free()
    dst.free()
    return {"copy_bytes_each_way": COPY_BYTES, "copy_ms": ms.value, "copy_launches": nrep + 1,
            "copy_gbs": 2 * COPY_BYTES / (ms.value * 1e-3) / 1e9}


def _setup(ctx, p):
    mesh = ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E)
    K = ctx.symbolic(mesh)
    _log("%s: %d DOF, %d elements, nnz %d" % (p.name, p.ndof, p.grid.nelem, K.nnz))
    return mesh, K


def probe_assembly(ctx, lib, a):
    p = P.by_name(a.workload)
    mesh, K = _setup(ctx, p)
    chosen, ms = ctx.assemble_tune(mesh, K, a.variant)
    _log("assembly variant %d (tuner ms %s)" % (chosen, ms.tolist()))
    t = [ctx.assemble(mesh, K).ms_device for _ in range(a.reps)]
    out = {"mode": "assembly", "workload": p.name, "variant": chosen, "elements": p.grid.nelem,
           "dof": p.ndof, "nnz": K.nnz, "assembly_launches": a.reps, "assembly_ms": t}
    out.update(_copy(ctx, lib, a.reps))
    return out


def probe_spmv(ctx, lib, a):
    p = P.by_name(a.workload)
    mesh, K = _setup(ctx, p)
    ctx.assemble(mesh, K)
    ctx.apply_penalty(K, p.fixed_dofs)
    variant, _ = ctx.tune(K)
    ms = ctx.profile_spmv(K, a.reps)
    out = {"mode": "spmv", "workload": p.name, "spmv_variant": variant, "nnz": K.nnz,
           "n": p.ndof, "spmv_ms": ms, "spmv_launches": a.reps + 1}
    out.update(_copy(ctx, lib, a.reps))
    return out


def probe_spmm(ctx, lib, a):
    p = P.by_name(a.workload)
    F = P.random_top_loads(p, a.nrhs, P.SEED_BRICK20_RHS)
    mesh, K = _setup(ctx, p)
    ctx.assemble(mesh, K)
    ctx.apply_penalty(K, p.fixed_dofs)
    variant, _ = ctx.tune(K)
    d_b = DevBuf.from_array(ctx, np.ascontiguousarray(F))
    d_x = DevBuf(ctx, F.nbytes)
    o = solve_opts(device_ptrs=1, rtol=1e-30, maxit=a.iters, batched=1, verify=0)
    st = _lib.SolveStats()
    code = lib.gf_pcg(ctx.handle, K.handle, d_b.ptr, d_x.ptr, a.nrhs, C.byref(o), C.byref(st))
    if code not in (0, _lib.GF_ERR_NOCONV):
        check(code, ctx.handle)
    out = {"mode": "spmm", "workload": p.name, "nrhs": a.nrhs, "spmv_variant": variant,
           "nnz": K.nnz, "n": p.ndof, "block_iterations": int(st.iterations),
           "block_ms": st.ms_device}
    out.update(_copy(ctx, lib, a.reps))
    return out


def summarize(paths):
    """Per-kernel counter means (or durations) of rocprofv3 result databases."""
    import collections
    import sqlite3
    for path in paths:
        db = sqlite3.connect(path)
        print("== %s" % path)
        rows = db.execute("select name, counter_name, dispatch_id, counter_value "
                          "from pmc_events").fetchall()
        if rows:
            # a counter with several hardware instances (shader engines, XCDs)
            # has one row per instance: summed per dispatch, then averaged
            per = collections.defaultdict(float)
            for name, cn, d, v in rows:
                per[(name, cn, d)] += v
            agg = collections.defaultdict(list)
            for (name, cn, d), v in per.items():
                agg[(name, cn)].append(v)
            print("| kernel | counter | dispatches | mean per dispatch |")
            print("|---|---|---|---|")
            for (name, cn), vs in sorted(agg.items(), key=lambda kv: (kv[0][0], kv[0][1])):
                print("| `%s` | %s | %d | %.1f |" % (name[:100], cn, len(vs), sum(vs) / len(vs)))
        else:
            ks = db.execute("select name, count(*), sum(duration), avg(duration), min(duration), "
                            "max(duration) from kernels group by name order by sum(duration) "
                            "desc").fetchall()
            print("| kernel | calls | total ms | avg us | min us | max us |")
            print("|---|---|---|---|---|---|")
            for name, c, tot, avg, mn, mx in ks:
                print("| `%s` | %d | %.3f | %.2f | %.2f | %.2f |" % (
                    name[:100], c, tot / 1e6, avg / 1e3, mn / 1e3, mx / 1e3))


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "summarize":
        summarize(argv[1:])
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["assembly", "spmv", "spmm"])
    ap.add_argument("--workload", default=None)
    ap.add_argument("--variant", type=int, default=-1, help="assembly kernel variant (-1: tuner)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--nrhs", type=int, default=32)
    ap.add_argument("--iters", type=int, default=8, help="block PCG iterations (spmm)")
    a = ap.parse_args(argv)
    if a.workload is None:
        a.workload = {"assembly": "brick8_cube_128", "spmv": "q8_plate_1024x512",
                      "spmm": "brick20_cube_48"}[a.mode]
    lib = load()
    ctx = Context(0)
    fn = {"assembly": probe_assembly, "spmv": probe_spmv, "spmm": probe_spmm}[a.mode]
    try:
        out = fn(ctx, lib, a)
    except GpufemError as e:
        _log("error: %s" % e)
        return 1
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
