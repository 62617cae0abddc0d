"""GPU parity of the device-side constraints and load vector against the
FP64 oracle: zero substitution (the elimination alternative to the penalty,
src/FEM/fem.cc:146-163), clearing right-hand sides at fixed DOFs, and the
force vector built by assignment (src/FEM/FemData.cc:370-385)."""
import numpy as np
import pytest

import gpufem_amd as gf
import oracle as O
from gpufem_amd import problems as P

pytestmark = pytest.mark.gpu

CASES = [(gf.GF_Q4, 12, 6, 0), (gf.GF_Q8, 9, 4, 0), (gf.GF_BRICK8, 4, 3, 3),
         (gf.GF_BRICK20, 3, 2, 2)]


def _assemble_both(ctx, p):
    mesh = ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = ctx.symbolic(mesh)
    ctx.assemble(mesh, K)
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    return mesh, K, Ko


@pytest.mark.parametrize("t,nx,ny,nz", CASES)
def test_elimination_matches_oracle(gpu_ctx, t, nx, ny, nz):
    p = P.structured(t, nx, ny, nz)
    mesh, K, Ko = _assemble_both(gpu_ctx, p)
    gpu_ctx.apply_elimination(K, p.fixed_dofs)
    O.apply_elimination(Ko, p.dim, p.supports)
    val, col, rn = gpu_ctx.download(K)
    ov, oc, orn = Ko.export(K.width)
    assert np.array_equal(rn, orn) and np.array_equal(col, oc)
    assert np.abs(val - ov).max() <= 1e-12 * np.abs(ov).max()
    # the eliminated entries are exact zeros and ones on both sides
    fixed = np.zeros(p.ndof, bool)
    fixed[p.fixed_dofs] = True
    rows = np.broadcast_to(np.arange(p.ndof), col.shape)
    live = np.arange(col.shape[0])[:, None] < rn[None, :]
    hit = live & (fixed[rows] | fixed[col])
    assert np.array_equal(val[hit], ov[hit])
    assert np.all(val[hit & (rows == col)] == 1.0) and np.all(val[hit & (rows != col)] == 0.0)


@pytest.mark.parametrize("t,nx,ny,nz", CASES)
def test_elimination_solve_matches_oracle_and_penalty(gpu_ctx, t, nx, ny, nz):
    p = P.structured(t, nx, ny, nz)
    mesh, K, Ko = _assemble_both(gpu_ctx, p)
    gpu_ctx.apply_elimination(K, p.fixed_dofs)
    d_f = gf.DevBuf.from_array(gpu_ctx, np.ascontiguousarray(p.f))
    gpu_ctx.zero_dofs(d_f, p.ndof, p.fixed_dofs)
    f = d_f.download(np.float64)
    assert np.all(f[p.fixed_dofs] == 0.0)
    free = np.setdiff1d(np.arange(p.ndof), p.fixed_dofs)
    assert np.array_equal(f[free], p.f[free])
    x, st = gpu_ctx.pcg(K, f)
    assert st.converged and st.true_relres <= 1e-8
    assert np.all(x[p.fixed_dofs] == 0.0)
    # against the oracle both solve to 1e-10, so that the two iterations'
    # different rounding stays below the 1e-9 displacement bar
    xt, st = gpu_ctx.pcg(K, f, rtol=1e-10)
    O.apply_elimination(Ko, p.dim, p.supports)
    xo, _, _, _ = O.pcg_jacobi(Ko, f, rtol=1e-10)
    assert np.linalg.norm(xt - xo) <= 1e-9 * np.linalg.norm(xo)
    # the penalty solution differs by the penalty's O(1e-9) perturbation
    K2 = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K2)
    gpu_ctx.apply_penalty(K2, p.fixed_dofs)
    xp, _ = gpu_ctx.pcg(K2, p.f)
    assert np.linalg.norm(xp - x) <= 1e-6 * np.linalg.norm(x)


def test_zero_dofs_several_right_hand_sides(gpu_ctx):
    rng = np.random.default_rng(3)
    n, nrhs = 1000, 3
    V = rng.standard_normal((nrhs, n))
    dofs = np.array([0, 999, 17, 17, 500], dtype=np.int32)
    d = gf.DevBuf.from_array(gpu_ctx, V)
    gpu_ctx.zero_dofs(d, n, dofs, nrhs=nrhs)
    W = d.download(np.float64).reshape(nrhs, n)
    ref = V.copy()
    ref[:, dofs] = 0.0
    assert np.array_equal(W, ref)
    with pytest.raises(gf.GpufemError):
        gpu_ctx.zero_dofs(d, n, np.array([n], dtype=np.int32))


@pytest.mark.parametrize("dim", [2, 3])
def test_load_vector_matches_oracle(gpu_ctx, dim):
    rng = np.random.default_rng(10 + dim)
    nnodes, nload = 700, 2000
    loads = np.zeros((nload, dim + 1))
    loads[:, 0] = rng.integers(1, nnodes + 1, nload)       # many repeated nodes
    comp = rng.standard_normal((nload, dim))
    comp[rng.random((nload, dim)) < 0.4] = 0.0             # zeros overwrite nothing
    loads[:, 1:] = comp
    f = gpu_ctx.load_vector(dim, nnodes, loads).download(np.float64)
    assert np.array_equal(f, O.force_vector(dim, nnodes, loads))
    # the reference's cantilever loads and an empty list
    p = P.structured(gf.GF_Q4 if dim == 2 else gf.GF_BRICK8, 6, 4, 0 if dim == 2 else 3)
    g = gpu_ctx.load_vector(dim, p.grid.nnodes, p.loads).download(np.float64)
    assert np.array_equal(g, O.force_vector(dim, p.grid.nnodes, p.loads))
    e = gpu_ctx.load_vector(dim, 5, np.zeros((0, dim + 1))).download(np.float64)
    assert not e.any()
    with pytest.raises(gf.GpufemError):
        gpu_ctx.load_vector(dim, nnodes, np.array([[nnodes + 1] + [1.0] * dim]))
