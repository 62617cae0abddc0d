"""GPU parity of the grouped batch (many uneven meshes assembled and solved
as one block-diagonal system) against the FP64 oracle, mesh by mesh."""
import numpy as np
import pytest

import gpufem_amd as gf
import oracle as O
from gpufem_amd import problems as P

pytestmark = pytest.mark.gpu


def oracle_solution(p, rtol=1e-8):
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E,
                       skip_zero=False)
    O.apply_penalty(Ko, p.dim, p.supports)
    return Ko, O.pcg_jacobi(Ko, p.f, rtol=rtol)


def test_group_assembly_and_solve_match_each_mesh(gpu_ctx):
    sizes = [(16, 8), (5, 3), (24, 12), (9, 4), (16, 8)]
    probs = [P.structured(gf.GF_Q4, nx, ny, E=210000.0 * (1 + 0.1 * k))
             for k, (nx, ny) in enumerate(sizes)]
    G = gpu_ctx.group(probs)
    assert G.n == sum(p.ndof for p in probs)
    G.assemble()
    G.apply_penalty()
    x, st, its = G.pcg()
    assert st.converged and st.true_relres <= 1.05e-8
    for k, (p, xk) in enumerate(zip(probs, G.split(x))):
        Ko, (xo, ito, _, _) = oracle_solution(p)
        assert abs(int(its[k]) - ito) <= max(2, ito // 100), (k, its[k], ito)
        assert np.linalg.norm(xk - xo) <= 1e-9 * np.linalg.norm(xo)
    # identical meshes 0 and 4 take identical iterations (the E scale does not
    # change Jacobi-PCG) and converge independently of their neighbours
    assert its[0] == its[4]


def test_group_matches_single_mesh_solver(gpu_ctx):
    probs = [P.structured(gf.GF_Q8, 8, 4), P.structured(gf.GF_Q8, 12, 6)]
    G = gpu_ctx.group(probs)
    G.assemble()
    G.apply_penalty()
    x, st, its = G.pcg()
    for p, xk, it in zip(probs, G.split(x), its):
        mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
        K = gpu_ctx.symbolic(mesh)
        gpu_ctx.assemble(mesh, K)
        gpu_ctx.apply_penalty(K, p.fixed_dofs)
        xs, ss = gpu_ctx.pcg(K, p.f)
        assert abs(int(it) - ss.iterations) <= 2
        assert np.linalg.norm(xk - xs) <= 1e-9 * np.linalg.norm(xs)


def test_group_solve_is_deterministic(gpu_ctx):
    """The per-mesh reductions (one workgroup per mesh, the last to finish
    setting the group's state) add in a fixed order: the same solve twice
    gives the same bits and per-mesh iteration counts."""
    probs = [P.structured(gf.GF_Q4, nx, ny) for nx, ny in [(16, 8), (5, 3), (24, 12)]]
    G = gpu_ctx.group(probs)
    G.assemble()
    G.apply_penalty()
    x1, s1, its1 = G.pcg()
    x2, s2, its2 = G.pcg()
    assert s1.converged
    assert np.array_equal(x1, x2) and np.array_equal(its1, its2)


def test_group_rejects_mixed_element_types(gpu_ctx):
    with pytest.raises(gf.GpufemError):
        gpu_ctx.group([P.structured(gf.GF_Q4, 4, 2), P.structured(gf.GF_Q8, 4, 2)])
