"""Result tables in the reference's "FEMGPU Output File" format, and the
strategy sweep that fills them from this library's tuners.

Format (FileIO::OpenOutputFile, writeAsmOutputHeader, writeSolOutputHeader,
writeNumTab, writeMatFormat, src/Utils/fileIO.cc:408-436, 457-488,
538-600; outfileInit / outfileWriteGptsColoring / outfileWriteHeader,
src/interface.cpp:195-234): a new file opens with "FEMGPU Output File" and
"Date: <asctime>", then "Gps:<n>\\tColoring:<0|1>\\t"; an appended run opens
with a line of 59 '#'.  Each run then writes "Matrix Format: ELL" and, per
block, "Threads:<n>\\t", a tab-terminated header, one tab-terminated row per
mesh (integers as %i, times in seconds and the matrix size in MB as %.3f)
and an empty line.  Older files of the reference omit the trailing tabs;
read_results accepts both.

The sweep (the role of RunBatches, src/interface.cpp:503-605) reads
neutral files, assembles each mesh on the GPU and writes one row per mesh:
AsmGPU is the numeric assembly time (the assembly tuner's kernel), and the
five solver columns are the PCG time with the SpMV variant that stands in
for each of the reference's OpenCL strategies forced (NAIVE -> thread per
row, NAIVEUR -> thread per row with four slots in flight, SHARE -> two
waves per row tile, BLOCK -> node-block columns, BLOCKUR -> block-sliced
copy).  The CPU columns are filled only when the caller passes CPU timings
(this library has no CPU path); otherwise they hold -1.  An extended
table with every SpMV variant's time, as the SpMV tuner measured it, is
returned alongside.
"""
import os
import time

import numpy as np

ASM_COLUMNS = ("DOF", "NODES", "ELEM", "SIZE", "AsmCPU", "AsmGPU")
SOL_COLUMNS = ("DOF", "ELEM", "SIZE", "AsmCPU", "AsmGPU", "SolCPU", "SolNaive", "SolNaiveUR",
               "SolShr", "SolBlk", "SolBlkUR")
SEPARATOR = "#" * 59
# SpMV variant per reference strategy column, by model dimension
STRATEGY_VARIANTS = {
    2: {"SolNaive": 0, "SolNaiveUR": 1, "SolShr": 3, "SolBlk": 5, "SolBlkUR": 7},
    3: {"SolNaive": 0, "SolNaiveUR": 1, "SolShr": 3, "SolBlk": 6, "SolBlkUR": 8},
}


def _num(v):
    if isinstance(v, (int, np.integer)):
        return "%i\t" % v
    return "%.3f\t" % v


class ResultsFile:
    """Writer of one run (a new file, or appended to an existing one)."""

    def __init__(self, path, append=False, gauss=2, coloring=True, fmt="ELL", date=None):
        self.fh = open(path, "a" if append else "w")
        if append:
            self.fh.write(SEPARATOR + "\n")
        else:
            self.fh.write("FEMGPU Output File\n")
            self.fh.write("Date: " + (date or time.asctime()) + "\n")
            self.fh.write("Gps:%i\tColoring:%i\t\n" % (gauss, int(bool(coloring))))
        self.fh.write("Matrix Format: %s\n" % fmt)

    def block(self, threads, solve):
        self.fh.write("Threads:%i\t\n" % threads)
        self.fh.write("".join(c + "\t" for c in (SOL_COLUMNS if solve else ASM_COLUMNS)) + "\n")

    def row(self, values):
        self.fh.write("".join(_num(v) for v in values) + "\n")

    def end_block(self):
        self.fh.write("\n")

    def close(self):
        self.fh.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _parse_num(s):
    try:
        return int(s)
    except ValueError:
        return float(s)


def read_results(path):
    """Blocks of a results file: dicts with format, gauss, coloring, threads,
    columns and rows (numbers)."""
    blocks, fmt, gauss, coloring, cur = [], None, None, None, None
    with open(path) as fh:
        lines = [ln.rstrip("\n") for ln in fh]
    i = 0
    while i < len(lines):
        s = lines[i]
        t = s.strip()
        if t.startswith("Gps:"):
            parts = [x for x in t.replace("\t", " ").replace(":", " ").split() if x]
            gauss, coloring = int(parts[1]), int(parts[3])
        elif t.startswith("Matrix Format:"):
            fmt = t.split(":", 1)[1].strip()
        elif t.startswith("Threads:"):
            cur = {"format": fmt, "gauss": gauss, "coloring": coloring,
                   "threads": int(t.split(":", 1)[1].strip()), "columns": [], "rows": []}
            blocks.append(cur)
            i += 1
            cur["columns"] = [c for c in lines[i].split("\t") if c]
        elif t and cur is not None and t[0].isdigit():
            cur["rows"].append([_parse_num(c) for c in s.split("\t") if c])
        i += 1
    return blocks


def sweep(ctx, paths, out, gauss=None, solve=True, append=False, reference_solve=True,
          cpu=None, log=None):
    """One results row per neutral file in paths, written to out.

    reference_solve: the solver columns time the reference's solve
    (unpreconditioned CG, at most 3000 iterations, ||r|| <= 1e-3 ||b||,
    src/interface.cpp:304-328); otherwise Jacobi-PCG to 1e-8 on the true
    residual.  cpu: optional callable(problem) -> (asm_seconds,
    solve_seconds, threads) for the CPU columns.  Returns the per-mesh
    records (with every SpMV variant's time from the SpMV tuner)."""
    from . import GF_PRECOND_JACOBI, GF_PRECOND_NONE, nf
    recs, rows = [], []
    threads, gps = 1, gauss or 2
    for path in paths:
        p = nf.read_nf(path, ngauss=gauss)
        gps = p.ngauss
        mesh = ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
        K = ctx.symbolic(mesh)
        variant, asm_ms = ctx.assemble_tune(mesh, K)
        ta = ctx.assemble(mesh, K)
        ctx.apply_penalty(K, p.fixed_dofs)
        size_mb = float(K.n) * K.width * (8 + 4) / (1024 * 1024)
        spmv_best, spmv_ms = ctx.tune(K)
        if reference_solve:
            opts = dict(precond=GF_PRECOND_NONE, maxit=3000, rtol=1e-3, verify=0)
        else:
            opts = dict(precond=GF_PRECOND_JACOBI, rtol=1e-8)
        sol = {}
        if solve:
            for name, v in STRATEGY_VARIANTS[p.dim].items():
                ctx.tune(K, v)
                _, st = ctx.pcg(K, p.f, raise_on_fail=False, **opts)
                sol[name] = st.ms_device / 1e3
            ctx.tune(K, spmv_best)
        asm_cpu, sol_cpu = -1.0, -1.0
        if cpu is not None:
            asm_cpu, sol_cpu, threads = cpu(p)
        t_asm = ta.ms_device / 1e3
        nnodes, nelem = p.coords.shape[0], p.conn.shape[0]
        if solve:
            rows.append([int(p.ndof), int(nelem), size_mb, asm_cpu, t_asm, sol_cpu] +
                        [sol[c] for c in SOL_COLUMNS[6:]])
        else:
            rows.append([int(p.ndof), int(nnodes), int(nelem), size_mb, asm_cpu, t_asm])
        recs.append({"file": os.path.basename(path), "dof": int(p.ndof), "elements": int(nelem),
                     "assembly_variant": int(variant),
                     "assembly_variant_ms": [float(v) for v in asm_ms],
                     "spmv_variant": int(spmv_best), "spmv_variant_ms": [float(v) for v in spmv_ms],
                     "solve_s": sol})
        if log:
            log("%s: %d DOF, assembly %.3f ms, solves %s" % (
                os.path.basename(path), p.ndof, ta.ms_device,
                " ".join("%s %.3f s" % kv for kv in sol.items())))
        K.free()
        mesh.free()
    with ResultsFile(out, append=append, gauss=gps, coloring=True) as rf:
        rf.block(threads, solve)
        for r in rows:
            rf.row(r)
        rf.end_block()
    return recs


# The meshes of the reference's published assembly tables
# (Report/test_models_home/*/_*_col.txt) and the table each appears in.
REFERENCE_MESHES = (("placa128_Q4.nf", "Q4_col.txt"), ("placa128_Q8.nf", "Q8_col.txt"),
                    ("rubik18_brk8.nf", "brk8_col.txt"), ("rubik20_brk20.nf", "brk20_col.txt"))


def reference_asm_times(table, dof, gauss):
    """{threads: (AsmCPU, AsmGPU)} of the ELL-format rows of a reference
    result table for the mesh with `dof` DOF and `gauss` points per axis."""
    out = {}
    for blk in read_results(table):
        if blk["format"] != "ELL" or blk["gauss"] != gauss or "AsmGPU" not in blk["columns"]:
            continue
        cols = blk["columns"]
        for row in blk["rows"]:
            if row and int(row[0]) == dof:
                out[blk["threads"]] = (row[cols.index("AsmCPU")], row[cols.index("AsmGPU")])
    return out


def reference_meshes(ctx, nf_dir, table_dir, log=None):
    """Assembles the reference's benchmark meshes with 2 and 3 Gauss points
    per axis and returns one record per (mesh, rule) with this library's
    numeric and full (symbolic + numeric) assembly times beside the
    reference's own AsmCPU / AsmGPU seconds at 1..4 threads.  The
    reference's AsmGPU covers allocation, copies, the OpenCL Ke kernel, the
    read-back and the host scatter (src/FEM/StiffAlgoGpuOmp.cc:86-172)."""
    from . import nf
    recs = []
    for name, table in REFERENCE_MESHES:
        for gauss in (2, 3):
            p = nf.read_nf(os.path.join(nf_dir, name), ngauss=gauss)
            mesh = ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
            t0 = time.perf_counter()
            K = ctx.symbolic(mesh)
            t_sym = time.perf_counter() - t0
            variant, ms = ctx.assemble_tune(mesh, K)
            t = min(ctx.assemble(mesh, K).ms_device for _ in range(5)) / 1e3
            ref = reference_asm_times(os.path.join(table_dir, table), p.ndof, gauss)
            best_gpu = min((v[1] for v in ref.values()), default=None)
            rec = {"mesh": name, "dof": int(p.ndof), "elements": int(p.conn.shape[0]),
                   "gauss": gauss, "assembly_variant": int(variant), "asm_s": t,
                   "asm_dof_per_s": p.ndof / t, "symbolic_s": t_sym,
                   "full_dof_per_s": p.ndof / (t + t_sym),
                   "reference": {str(k): {"AsmCPU": v[0], "AsmGPU": v[1]} for k, v in ref.items()},
                   "reference_best_AsmGPU_s": best_gpu,
                   "speedup_vs_reference_AsmGPU": (best_gpu / t) if best_gpu else None}
            recs.append(rec)
            if log:
                log("%s %d GP: %d DOF, assembly %.4f ms (%.3g DOF/s), full %.3f s; "
                    "reference AsmGPU best %s s" % (name, gauss, p.ndof, t * 1e3, p.ndof / t,
                                                    t + t_sym, best_gpu))
            K.free()
            mesh.free()
    return recs


def main(argv=None):
    """python -m gpufem_amd.results OUT file.nf ... [--append] [--asm-only]
    [--pcg] [--gauss N]
    python -m gpufem_amd.results --reference-meshes NF_DIR TABLE_DIR"""
    import argparse
    import json
    from . import Context
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("files", nargs="+")
    ap.add_argument("--append", action="store_true")
    ap.add_argument("--asm-only", action="store_true")
    ap.add_argument("--pcg", action="store_true",
                    help="solver columns: Jacobi-PCG to 1e-8 instead of the reference's CG")
    ap.add_argument("--gauss", type=int, default=None)
    if argv is None:
        import sys
        argv = sys.argv[1:]
    if argv and argv[0] == "--reference-meshes":
        ctx = Context(0)
        recs = reference_meshes(ctx, argv[1], argv[2],
                                log=lambda m: print("[results] " + m, flush=True))
        print(json.dumps(recs))
        return
    a = ap.parse_args(argv)
    ctx = Context(0)
    recs = sweep(ctx, a.files, a.out, gauss=a.gauss, solve=not a.asm_only, append=a.append,
                 reference_solve=not a.pcg,
                 log=lambda m: print("[results] " + m, flush=True))
    print(json.dumps(recs))


if __name__ == "__main__":
    main()
