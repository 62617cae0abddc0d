"""The reference's own meshes (neutral files in tests/golden/nf/) through the
GPU path, compared with the CPU oracle: ELL structure bit-exact, values
within 1e-12, displacements within 1e-9."""
import os

import numpy as np
import pytest

import oracle as O
from gpufem_amd import nf

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nf")


@pytest.mark.parametrize("name", ["rubik10_brk8.nf", "brick_brk20v2.nf", "placa2_Q4.nf",
                                  "placa2_Q8.nf", "brick_brk8v2.nf",
                                  # the meshes of the reference's published tables
                                  "placa128_Q4.nf", "placa128_Q8.nf", "rubik18_brk8.nf",
                                  "rubik20_brk20.nf"])
def test_reference_mesh_matches_oracle(gpu_ctx, name):
    p = nf.read_nf(os.path.join(HERE, name))
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    val, col, rownnz = gpu_ctx.download(K)
    ov, oc, orn = Ko.export(K.width)
    assert np.array_equal(rownnz, orn)
    mask = np.arange(K.width)[None, :] < rownnz[:, None]
    assert np.array_equal(col.reshape(K.width, -1).T[mask], oc.reshape(K.width, -1).T[mask])
    scale = np.abs(ov).max()
    assert np.abs(val - ov).max() <= 1e-12 * scale
    gpu_ctx.apply_penalty(K, p.fixed_dofs)
    O.apply_penalty(Ko, p.dim, p.supports)
    if not np.any(p.f):
        return
    x, st = gpu_ctx.pcg(K, p.f)
    xo, ito, _, _ = O.pcg_jacobi(Ko, p.f, rtol=1e-8)
    assert st.converged and st.true_relres <= 1e-8
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


def test_reference_mesh_sweep_reads_reference_tables(gpu_ctx):
    """The reference-mesh sweep assembles each mesh with both Gauss rules and
    finds the reference's own AsmGPU rows for it (1..4 threads)."""
    from gpufem_amd import results
    recs = results.reference_meshes(gpu_ctx, HERE, os.path.join(os.path.dirname(HERE), "results"))
    assert len(recs) == 8
    for r in recs:
        assert set(r["reference"]) == {"1", "2", "3", "4"}, (r["mesh"], r["gauss"])
        assert r["asm_s"] > 0 and r["reference_best_AsmGPU_s"] > 0
