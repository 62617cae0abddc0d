"""GPU parity of the assembly path (Ke + colored scatter + penalty) against the
FP64 oracle: ELL structure bit-exact, values within 1e-12 relative (to the
largest stored magnitude), colors identical, repeated runs bitwise identical."""
import numpy as np
import pytest

import gpufem_amd as gf
import oracle as O
from gpufem_amd import problems as P

pytestmark = pytest.mark.gpu

VAL_RTOL = 1e-12

CASES = [
    (gf.GF_Q4, 64, 32, 0, 2),      # the reference-sized plate (4,290 DOF)
    (gf.GF_Q4, 7, 5, 0, 3),
    (gf.GF_Q8, 24, 12, 0, 3),
    (gf.GF_Q8, 9, 4, 0, 2),
    (gf.GF_BRICK8, 8, 6, 5, 2),
    (gf.GF_BRICK8, 4, 4, 3, 3),
    (gf.GF_BRICK20, 4, 3, 3, 3),
    (gf.GF_BRICK20, 3, 3, 2, 2),
]


def build_both(ctx, t, nx, ny, nz, ng, distort=0.0, elem_E=None, seed=0):
    g = gf.Grid(t, nx, ny, nz)
    conn, coords = g.fill()
    if distort:
        rng = np.random.default_rng(seed)
        coords = coords + distort * rng.uniform(-1, 1, coords.shape)
    mesh = ctx.upload_mesh(t, ng, conn, coords, 210000.0, 0.3, elem_E=elem_E)
    K = ctx.symbolic(mesh)
    ctx.assemble(mesh, K)
    Ko, _ = O.assemble(g.dim, ng, conn, coords, 210000.0, 0.3, elem_E=elem_E, skip_zero=False)
    return g, conn, coords, mesh, K, Ko


def compare(ctx, K, Ko):
    val, col, rn = ctx.download(K)
    width = Ko.max_rownnz()
    assert K.width == width
    ov, oc, orn = Ko.export(width)
    assert np.array_equal(rn, orn)
    assert np.array_equal(col, oc)                   # bit-exact structure (incl. padding)
    scale = np.abs(ov).max()
    err = np.abs(val - ov).max()
    assert err <= VAL_RTOL * scale, "max |dK| = %g (scale %g)" % (err, scale)
    return val, col, rn


@pytest.mark.parametrize("t,nx,ny,nz,ng", CASES)
def test_assembly_parity(gpu_ctx, t, nx, ny, nz, ng):
    g, conn, coords, mesh, K, Ko = build_both(gpu_ctx, t, nx, ny, nz, ng)
    compare(gpu_ctx, K, Ko)
    col_o, nc = O.greedy_coloring(conn, g.nnodes)
    assert K.ncolors == nc
    assert np.array_equal(gpu_ctx.colors(mesh, K), col_o)


@pytest.mark.parametrize("t,nx,ny,nz,ng", [(gf.GF_Q4, 6, 5, 0, 2), (gf.GF_Q8, 5, 4, 0, 3),
                                           (gf.GF_BRICK8, 3, 3, 3, 2),
                                           (gf.GF_BRICK20, 2, 3, 2, 3)])
def test_assembly_parity_distorted_variable_E(gpu_ctx, t, nx, ny, nz, ng):
    nel = gf.Grid(t, nx, ny, nz).nelem
    E = P.lognormal_E(nel, 1234)
    g, conn, coords, mesh, K, Ko = build_both(gpu_ctx, t, nx, ny, nz, ng, distort=0.08,
                                              elem_E=E, seed=3)
    compare(gpu_ctx, K, Ko)


def test_assembly_deterministic(gpu_ctx):
    g, conn, coords, mesh, K, Ko = build_both(gpu_ctx, gf.GF_BRICK8, 6, 5, 4, 2)
    v1, _, _ = gpu_ctx.download(K)
    gpu_ctx.assemble(mesh, K)
    v2, _, _ = gpu_ctx.download(K)
    assert np.array_equal(v1, v2)


@pytest.mark.parametrize("t,nx,ny,nz,ng", [(gf.GF_Q4, 20, 9, 0, 2), (gf.GF_Q8, 11, 6, 0, 3),
                                           (gf.GF_BRICK8, 7, 5, 4, 2), (gf.GF_BRICK8, 4, 3, 3, 3),
                                           (gf.GF_BRICK20, 4, 3, 3, 3),
                                           (gf.GF_BRICK20, 3, 3, 2, 2)])
def test_pair_major_assembly_is_bitwise_element_major(gpu_ctx, t, nx, ny, nz, ng):
    g, conn, coords, mesh, K, Ko = build_both(gpu_ctx, t, nx, ny, nz, ng)
    gpu_ctx.assemble_tune(mesh, K, 0)
    gpu_ctx.assemble(mesh, K)
    v0, c0, r0 = gpu_ctx.download(K)
    gpu_ctx.assemble_tune(mesh, K, 1)
    gpu_ctx.assemble(mesh, K)
    v1, c1, r1 = gpu_ctx.download(K)
    assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
    chosen, ms = gpu_ctx.assemble_tune(mesh, K)
    b20 = t == gf.GF_BRICK20
    assert chosen in ((0, 1, 2, 3) if b20 else (0, 1, 3))
    assert ms[0] > 0 and ms[1] > 0 and (ms[2] > 0) == b20 and ms[3] > 0 and ms[4] == -1.0
    with pytest.raises(gf.GpufemError):
        gpu_ctx.assemble_tune(mesh, K, 5)


@pytest.mark.parametrize("t,nx,ny,nz,ng", [(gf.GF_BRICK8, 7, 5, 4, 2), (gf.GF_BRICK8, 4, 3, 3, 3),
                                           (gf.GF_BRICK20, 4, 3, 3, 3),
                                           (gf.GF_BRICK20, 3, 3, 2, 2)])
def test_matrix_core_assembly_matches_oracle(gpu_ctx, t, nx, ny, nz, ng):
    """Element tensors T_pq = sum_g w detJ E G_p^T G_q as FP64 MFMA tiles, then
    the node gather over them (another summation order): within 1e-12 of the
    oracle on distorted elements with per-element moduli, deterministic,
    and a re-assembly overwrites earlier values (penalty included)."""
    g = gf.Grid(t, nx, ny, nz)
    E = np.random.default_rng(41).uniform(1e5, 3e5, g.nelem)
    g_, conn, coords, mesh, K, Ko = build_both(gpu_ctx, t, nx, ny, nz, ng, distort=0.15,
                                               elem_E=E, seed=4)
    gpu_ctx.assemble_tune(mesh, K, 2)
    gpu_ctx.assemble(mesh, K)
    v1, _, _ = compare(gpu_ctx, K, Ko)
    gpu_ctx.apply_penalty(K, np.arange(0, K.n, 7, dtype=np.int32))
    gpu_ctx.assemble(mesh, K)
    v2, _, _ = compare(gpu_ctx, K, Ko)
    assert np.array_equal(v1, v2)


GATHER_CASES = [(gf.GF_Q4, 64, 32, 0, 2), (gf.GF_Q4, 7, 5, 0, 3), (gf.GF_Q8, 24, 12, 0, 3),
                (gf.GF_Q8, 9, 4, 0, 2), (gf.GF_BRICK8, 8, 6, 5, 2), (gf.GF_BRICK8, 4, 4, 3, 3)]


@pytest.mark.parametrize("t,nx,ny,nz,ng", GATHER_CASES + [(gf.GF_BRICK20, 4, 3, 3, 3),
                                                          (gf.GF_BRICK20, 5, 3, 2, 2)])
def test_gather_assembly_matches_oracle(gpu_ctx, t, nx, ny, nz, ng):
    """Node gather (lanes of a node or node row share its elements, element
    contributions summed in element order, every live slot written once):
    within 1e-12 of the oracle on distorted elements with per-element moduli,
    deterministic, and a re-assembly overwrites earlier values (penalty
    included)."""
    g = gf.Grid(t, nx, ny, nz)
    rng = np.random.default_rng(31)
    E = rng.uniform(1e5, 3e5, g.nelem)
    g_, conn, coords, mesh, K, Ko = build_both(gpu_ctx, t, nx, ny, nz, ng, distort=0.15,
                                               elem_E=E, seed=8)
    gpu_ctx.assemble_tune(mesh, K, 3)
    gpu_ctx.assemble(mesh, K)
    v1, _, _ = compare(gpu_ctx, K, Ko)
    gpu_ctx.apply_penalty(K, np.arange(0, K.n, 7, dtype=np.int32))
    gpu_ctx.assemble(mesh, K)
    v2, _, _ = compare(gpu_ctx, K, Ko)
    assert np.array_equal(v1, v2)


@pytest.mark.parametrize("t,nx,ny,nz,ng", GATHER_CASES + [(gf.GF_BRICK20, 4, 3, 3, 3),
                                                          (gf.GF_BRICK20, 3, 3, 2, 2)])
def test_two_phase_assembly_matches_oracle(gpu_ctx, t, nx, ny, nz, ng):
    """Two phases (J^-1 and w detJ E of every element at its Gauss points,
    then the node gather reading them): within 1e-12 of the oracle on
    distorted elements with per-element moduli, bitwise equal to the node
    gather that forms the geometry itself where that gather also takes every
    row of a node in one lane (not Brick20: one row per lane, the sums in
    another order), deterministic, and a re-assembly
    overwrites earlier values (penalty included); with the block-sliced copy
    built it is written too (the products stay bitwise those of the ELL
    values)."""
    g = gf.Grid(t, nx, ny, nz)
    rng = np.random.default_rng(37)
    E = rng.uniform(1e5, 3e5, g.nelem)
    g_, conn, coords, mesh, K, Ko = build_both(gpu_ctx, t, nx, ny, nz, ng, distort=0.15,
                                               elem_E=E, seed=8)
    gpu_ctx.assemble_tune(mesh, K, 3)
    gpu_ctx.assemble(mesh, K)
    v3, _, _ = gpu_ctx.download(K)
    gpu_ctx.assemble_tune(mesh, K, 4)
    gpu_ctx.assemble(mesh, K)
    v1, _, _ = compare(gpu_ctx, K, Ko)
    if t != gf.GF_BRICK20:
        assert np.array_equal(v1, v3)
    gpu_ctx.apply_penalty(K, np.arange(0, K.n, 7, dtype=np.int32))
    gpu_ctx.assemble(mesh, K)
    v2, _, _ = compare(gpu_ctx, K, Ko)
    assert np.array_equal(v1, v2)
    bsell = 7 if g.dim == 2 else 8
    gpu_ctx.tune(K, bsell)
    gpu_ctx.assemble(mesh, K)
    x = rng.normal(size=K.n)
    yb, _ = gpu_ctx.spmv(K, x)
    gpu_ctx.tune(K, 0)
    y0, _ = gpu_ctx.spmv(K, x)
    assert np.array_equal(yb, y0)


def test_matrix_core_assembly_rejects_2d(gpu_ctx):
    g, conn, coords, mesh, K, Ko = build_both(gpu_ctx, gf.GF_Q8, 4, 3, 0, 3)
    with pytest.raises(gf.GpufemError):
        gpu_ctx.assemble_tune(mesh, K, 2)


def test_penalty_parity(gpu_ctx):
    p = P.config_q4_plate()
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    pen = gpu_ctx.apply_penalty(K, p.fixed_dofs, 1e4)
    Ko, _ = O.assemble(2, 2, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    pen_o = O.apply_penalty(Ko, 2, p.supports, 1e4)
    assert abs(pen - pen_o) <= 1e-12 * pen_o
    compare(gpu_ctx, K, Ko)


def test_penalty_repeated_dof_multiplies_twice(gpu_ctx):
    n = 8
    val = np.zeros((1, n))
    val[0] = np.arange(1, n + 1)
    col = np.arange(n, dtype=np.int32)[None, :]
    K = gpu_ctx.upload_ell(val, col, np.ones(n, dtype=np.int32))
    pen = gpu_ctx.apply_penalty(K, np.array([3, 3, 5], dtype=np.int32), 10.0)
    assert pen == 80.0
    v, _, _ = gpu_ctx.download(K)
    assert v[0, 3] == 4.0 * 80.0 * 80.0 and v[0, 5] == 6.0 * 80.0 and v[0, 0] == 1.0


def test_degenerate_element_reported(gpu_ctx):
    g = gf.Grid(gf.GF_Q4, 2, 1)
    conn, coords = g.fill()
    conn[0] = conn[0][[1, 0, 3, 2]]              # mirrored element: detJ < 0
    mesh = gpu_ctx.upload_mesh(gf.GF_Q4, 2, conn, coords)
    K = gpu_ctx.symbolic(mesh)
    with pytest.raises(gf.GpufemError) as e:
        gpu_ctx.assemble(mesh, K)
    assert e.value.code == gf.GF_ERR_DEGENERATE
    for v in (3, 4):
        gpu_ctx.assemble_tune(mesh, K, v)
        with pytest.raises(gf.GpufemError) as e:
            gpu_ctx.assemble(mesh, K)
        assert e.value.code == gf.GF_ERR_DEGENERATE
