"""The strategy sweep on the reference's own neutral files: one row per
mesh in the reference's result-table layout, filled from the GPU
assembly and the solver with each strategy's SpMV variant forced."""
import os

import numpy as np
import pytest

from gpufem_amd import nf
from gpufem_amd import results as Rz

pytestmark = pytest.mark.gpu

NF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nf")


@pytest.mark.parametrize("reference_solve", [True, False])
def test_sweep_writes_reference_layout(gpu_ctx, tmp_path, reference_solve):
    files = [os.path.join(NF, "placa2_Q4.nf"), os.path.join(NF, "rubik10_brk8.nf")]
    out = str(tmp_path / "sweep.txt")
    recs = Rz.sweep(gpu_ctx, files, out, reference_solve=reference_solve)
    blocks = Rz.read_results(out)
    assert len(blocks) == 1 and tuple(blocks[0]["columns"]) == Rz.SOL_COLUMNS
    rows = blocks[0]["rows"]
    assert len(rows) == 2
    for row, path, rec in zip(rows, files, recs):
        p = nf.read_nf(path)
        assert row[0] == p.ndof and row[1] == p.conn.shape[0]
        assert row[3] == -1.0 and row[5] == -1.0          # no CPU timings passed
        # seconds at %.3f: small meshes print 0.000, the records hold the times
        assert row[4] >= 0 and all(v >= 0 for v in row[6:])
        ams = rec["assembly_variant_ms"]   # -1 for variants not timed (matrix cores)
        assert ams[0] > 0 and ams[1] > 0 and ams[rec["assembly_variant"]] > 0
        assert all(v > 0 for v in rec["solve_s"].values())
        assert [round(rec["solve_s"][c], 3) for c in Rz.SOL_COLUMNS[6:]] == row[6:]
        ms = np.array(rec["spmv_variant_ms"])
        assert ms[rec["spmv_variant"]] == ms[ms > 0].min()
    # appended asm-only run
    Rz.sweep(gpu_ctx, files[:1], out, solve=False, append=True)
    blocks = Rz.read_results(out)
    assert tuple(blocks[1]["columns"]) == Rz.ASM_COLUMNS and blocks[1]["rows"][0][1] == \
        nf.read_nf(files[0]).coords.shape[0]
