"""GPU tests of the block-sliced copy the solver reads: written by the node
gather beside the ELL values, kept current by the penalty, and read through
32-bit node columns when some node-column offset does not fit 16 bits; and
the grouped batch over uneven 3D meshes."""
import numpy as np
import pytest

import gpufem_amd as gf
import oracle as O
from gpufem_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("t,dims,bsell,asm", [(gf.GF_Q8, (20, 10), 9, 3), (gf.GF_Q4, (33, 17), 7, 3),
                                              (gf.GF_BRICK8, (7, 6, 5), 10, 3),
                                              (gf.GF_BRICK8, (5, 4, 3), 8, 3),
                                              (gf.GF_BRICK20, (4, 3, 3), 10, 3),
                                              (gf.GF_BRICK20, (4, 3, 3), 10, 2)])
def test_gather_writes_block_sliced_copy(gpu_ctx, t, dims, bsell, asm):
    """With the copy built, the node gather (asm 3; asm 2 over the element
    tensors of the matrix cores) writes it and the penalty updates its
    diagonal: its products stay bitwise equal to the thread-per-row ELL
    kernel's, with no re-layout in between."""
    p = P.structured(t, *dims)
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    gpu_ctx.tune(K, bsell)                       # builds the copy from the ELL values
    gpu_ctx.assemble_tune(mesh, K, asm)          # node gather from here on
    x = np.random.default_rng(4).normal(size=p.ndof)
    rng = np.random.default_rng(7)
    for step in range(2):
        mesh.set_E(210000.0 * np.exp(0.5 * rng.normal(size=p.grid.nelem)))
        gpu_ctx.assemble(mesh, K)
        for penalty in (False, True):
            if penalty:
                gpu_ctx.apply_penalty(K, p.fixed_dofs)
            gpu_ctx.tune(K, bsell)
            yb, _ = gpu_ctx.spmv(K, x)
            gpu_ctx.tune(K, 0)
            y0, _ = gpu_ctx.spmv(K, x)
            assert np.array_equal(yb, y0), (step, penalty)
    # and the assembled values are the oracle's
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu,
                       elem_E=mesh._arrays[2], skip_zero=False)
    O.apply_penalty(Ko, p.dim, p.supports)
    yo = Ko.axy(x)
    assert np.abs(y0 - yo).max() <= 1e-12 * np.abs(yo).max()


@pytest.mark.parametrize("t,dims,bsell,asm", [(gf.GF_Q8, (20, 10), 9, 3),
                                              (gf.GF_BRICK8, (7, 6, 5), 10, 3),
                                              (gf.GF_BRICK20, (4, 3, 3), 10, 3),
                                              (gf.GF_BRICK20, (4, 3, 3), 10, 2)])
def test_stale_ell_values_materialise(gpu_ctx, t, dims, bsell, asm):
    """Once the copy exists the node gather writes it and only the diagonal
    of the ELL values; every reader of the other slots (download, the
    elimination, the ELL SpMV kernels) gets them materialised from the copy:
    the results are bitwise those of a matrix whose gather wrote every ELL
    value (no copy built), penalty and elimination included."""
    p = P.structured(t, *dims)
    E = 210000.0 * np.exp(0.4 * np.random.default_rng(3).normal(size=p.grid.nelem))
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=E)
    K = gpu_ctx.symbolic(mesh)      # with the copy: ELL values left stale
    K2 = gpu_ctx.symbolic(mesh)     # without a copy: every ELL value written
    K3 = gpu_ctx.symbolic(mesh)     # copy built from fully written ELL values
    for M in (K, K2, K3):
        gpu_ctx.assemble_tune(mesh, M, asm)
    gpu_ctx.tune(K, bsell)
    gpu_ctx.tune(K2, 0)
    for M in (K, K2, K3):
        gpu_ctx.assemble(mesh, M)
        gpu_ctx.apply_penalty(M, p.fixed_dofs)
    gpu_ctx.tune(K3, bsell)
    x1, s1 = gpu_ctx.pcg(K, p.f)                # reads the copy; ELL stays stale
    x3, s3 = gpu_ctx.pcg(K3, p.f)
    assert s1.iterations == s3.iterations and np.array_equal(x1, x3)
    v1, c1, _ = gpu_ctx.download(K)             # materialised
    v2, c2, _ = gpu_ctx.download(K2)
    assert np.array_equal(v1, v2) and np.array_equal(c1, c2)
    # elimination after a stale assembly
    dofs = np.asarray(p.fixed_dofs, dtype=np.int32)
    for M in (K, K2):
        gpu_ctx.assemble(mesh, M)
        gpu_ctx.apply_elimination(M, dofs)
    v1, _, _ = gpu_ctx.download(K)
    v2, _, _ = gpu_ctx.download(K2)
    assert np.array_equal(v1, v2)


def test_wide_node_columns_match_thread_per_row(gpu_ctx):
    """A Q4 plate of 33,153 nodes numbered at random: neighbouring nodes lie
    more than 32,767 apart, so the block-sliced kernels read 32-bit node
    columns instead of 16-bit offsets; the products and the solve stay
    bitwise those of the thread-per-row kernel."""
    g = gf.Grid(gf.GF_Q4, 256, 128)
    conn, coords = g.fill()
    perm = np.random.default_rng(12).permutation(g.nnodes)   # old node -> new node
    conn2 = (perm[conn - 1] + 1).astype(np.int32)
    coords2 = np.empty_like(coords)
    coords2[perm] = coords
    a, b = conn2[:, 0] - 1, conn2[:, 2] - 1
    assert np.abs(a - b).max() > 32767
    mesh = gpu_ctx.upload_mesh(gf.GF_Q4, 2, conn2, coords2, 210000.0, 0.3)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    left = np.nonzero(coords2[:, 0] == 0.0)[0]
    dofs = np.sort(np.concatenate([2 * left, 2 * left + 1])).astype(np.int32)
    gpu_ctx.apply_penalty(K, dofs)
    x = np.random.default_rng(2).normal(size=K.n)
    gpu_ctx.tune(K, 0)
    y0, _ = gpu_ctx.spmv(K, x)
    b = np.zeros(K.n)
    b[2 * np.nonzero(coords2[:, 0] == 256.0)[0] + 1] = -1.0
    x0, s0 = gpu_ctx.pcg(K, b, schedule=gf.GF_SCHED_KERNELS)
    for v in (7, 9):
        gpu_ctx.tune(K, v)
        y1, _ = gpu_ctx.spmv(K, x)
        assert np.array_equal(y0, y1), v
    x1, s1 = gpu_ctx.pcg(K, b, schedule=gf.GF_SCHED_KERNELS)
    assert s1.converged and abs(s1.iterations - s0.iterations) <= 1 + s0.iterations // 100
    assert np.linalg.norm(x1 - x0) <= 1e-7 * np.linalg.norm(x0)


def oracle_solution(p, rtol=1e-10):
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E,
                       skip_zero=False)
    O.apply_penalty(Ko, p.dim, p.supports)
    return O.pcg_jacobi(Ko, p.f, rtol=rtol)


@pytest.mark.parametrize("t,sizes", [
    (gf.GF_BRICK8, [(4, 3, 2), (14, 13, 12), (2, 2, 1), (9, 5, 3)]),   # 2,730 nodes in the second
    (gf.GF_BRICK20, [(3, 2, 2), (9, 8, 7), (1, 1, 1), (4, 3, 3)])])    # 2,638 nodes in the second
def test_group_uneven_3d_meshes(gpu_ctx, t, sizes):
    """Grouped 3D meshes of uneven sizes, one spanning several 2,048-node
    windows of the block-sliced copy: each mesh matches the oracle and the
    single-mesh solver."""
    rng = np.random.default_rng(5)
    probs = []
    for k, dims in enumerate(sizes):
        q = P.structured(t, *dims)
        q.elem_E = 210000.0 * np.exp(0.3 * rng.normal(size=q.grid.nelem))
        probs.append(q)
    assert max(q.grid.nnodes for q in probs) > 2048
    G = gpu_ctx.group(probs)
    G.assemble()
    G.apply_penalty()
    x, st, its = G.pcg(rtol=1e-10)
    assert st.converged and st.true_relres <= 1.05e-10
    for q, xk, it in zip(probs, G.split(x), its):
        xo, ito, _, _ = oracle_solution(q)
        assert np.linalg.norm(xk - xo) <= 1e-9 * np.linalg.norm(xo)
        mesh = gpu_ctx.upload_mesh(q.elem_type, q.ngauss, q.conn, q.coords, q.E, q.nu,
                                   elem_E=q.elem_E)
        K = gpu_ctx.symbolic(mesh)
        gpu_ctx.assemble(mesh, K)
        gpu_ctx.apply_penalty(K, q.fixed_dofs)
        xs, ss = gpu_ctx.pcg(K, q.f, rtol=1e-10)
        assert abs(int(it) - ss.iterations) <= 2 + ss.iterations // 100
        assert np.linalg.norm(xk - xs) <= 1e-9 * np.linalg.norm(xs)
