"""GPU parity of SpMV and the (Jacobi-preconditioned) CG against the
reference's known answers and the FP64 oracle."""
import json
import os

import numpy as np
import pytest

import gpufem_amd as gf
import oracle as O
from gpufem_amd import problems as P

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def upload_oracle(ctx, m):
    val, col, rn = m.export()
    return ctx.upload_ell(val, col, rn)


def diag_corner(n):
    m = O.EllMatrix(n)
    for i in range(n):
        m.add(i, i, float((i + 1) ** 2))
    m.set(0, 0, 2)
    m.set(0, 1, 2)
    m.set(1, 0, 2)
    return m


# ---- SpMV known answers (src/Test/AxyGPUTest.cpp:44-197) ------------------------------
@pytest.mark.parametrize("n", GOLD["spmv_diagonal"]["sizes"])
def test_spmv_diagonal(gpu_ctx, n):
    m = O.EllMatrix(n)
    for i in range(n):
        m.add(i, i, i + 1.0)
    K = upload_oracle(gpu_ctx, m)
    y, _ = gpu_ctx.spmv(K, np.arange(1, n + 1, dtype=float))
    assert np.array_equal(y, np.arange(1, n + 1, dtype=float) ** 2)


@pytest.mark.parametrize("n", GOLD["spmv_band_ones"]["sizes"])
def test_spmv_band(gpu_ctx, n):
    val = np.ones((8, n))
    col = np.tile(np.arange(8, dtype=np.int32)[:, None], (1, n))
    K = gpu_ctx.upload_ell(val, col, np.full(n, 8, dtype=np.int32))
    y, _ = gpu_ctx.spmv(K, np.full(n, 2.0))
    assert np.all(y == 16.0)


def test_spmv_shifted_band(gpu_ctx):
    m = O.EllMatrix(64)
    for i in range(64):
        for j in range(4):
            m.add(i, i + j, 1)
    K = upload_oracle(gpu_ctx, m)
    y, _ = gpu_ctx.spmv(K, np.full(64, 2.0))
    assert np.array_equal(y, m.axy(np.full(64, 2.0)))


def test_spmv_random_matches_oracle_and_multi_rhs(gpu_ctx):
    g = gf.Grid(gf.GF_Q8, 20, 10)
    conn, coords = g.fill()
    Ko, _ = O.assemble(2, 3, conn, coords, 210000.0, 0.3, skip_zero=False)
    K = upload_oracle(gpu_ctx, Ko)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(3, g.ndof))
    Y, _ = gpu_ctx.spmv(K, X)
    for k in range(3):
        # same slot order as the oracle's Axy, so FP64 results agree to rounding
        yo = Ko.axy(X[k])
        assert np.abs(Y[k] - yo).max() <= 1e-13 * np.abs(yo).max()


# ---- CG known answers (src/Test/AxyGPUTest.cpp:201-367) -------------------------------
@pytest.mark.parametrize("n", GOLD["cg_diag_corner"]["sizes"])
@pytest.mark.parametrize("sched", [gf.GF_SCHED_KERNELS, gf.GF_SCHED_GRAPH])
def test_cg_kat_diag_corner(gpu_ctx, n, sched):
    K = upload_oracle(gpu_ctx, diag_corner(n))
    b = np.arange(1, n + 1, dtype=float) ** 2
    x, st = gpu_ctx.pcg(K, b, precond=gf.GF_PRECOND_NONE, rtol=1e-6, maxit=500, schedule=sched)
    assert abs(x[0] + 1) < 1e-3 and abs(x[1] - 1.5) < 1e-3 and np.all(np.abs(x[2:] - 1) < 1e-3)
    x, st = gpu_ctx.pcg(K, b, rtol=1e-12, schedule=sched)
    assert abs(x[0] + 1) < 1e-9 and abs(x[1] - 1.5) < 1e-9 and st.converged


def test_cg_kat_tridiag(gpu_ctx):
    m = O.EllMatrix(4)
    for i in range(4):
        m.add(i, i, 2)
        m.add(i, i + 1, -1)
        m.add(i, i - 1, -1)
    K = upload_oracle(gpu_ctx, m)
    x, st = gpu_ctx.pcg(K, np.ones(4), precond=gf.GF_PRECOND_NONE, rtol=1e-5, maxit=200)
    assert np.allclose(x, GOLD["cg_tridiag_4"]["x"], atol=1e-3)


def test_cg_kat_ones_plus_identity(gpu_ctx):
    n = 32
    val = np.ones((n, n))
    col = np.tile(np.arange(n, dtype=np.int32)[:, None], (1, n))
    for i in range(n):
        val[i, i] = 2.0
    K = gpu_ctx.upload_ell(val, col, np.full(n, n, dtype=np.int32))
    x, st = gpu_ctx.pcg(K, np.ones(n), precond=gf.GF_PRECOND_NONE, rtol=1e-5, maxit=500)
    assert np.allclose(x, GOLD["cg_ones_plus_identity_32"]["x_each"], atol=1e-3)


def test_reference_cg_iterations_match_oracle(gpu_ctx):
    """Unpreconditioned CG with the reference's stopping rule (r.r <= eps^2 r0.r0)."""
    p = P.config_q4_plate()
    Ko, _ = O.assemble(2, 2, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    O.apply_penalty(Ko, 2, p.supports)
    xo, ito, _ = O.cg(Ko, p.f, 100000, 1e-8)
    K = upload_oracle(gpu_ctx, Ko)
    x, st = gpu_ctx.pcg(K, p.f, precond=gf.GF_PRECOND_NONE, rtol=1e-8)
    assert abs(st.iterations - ito) <= max(3, ito // 50)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


# ---- the assembled problem: plate 64x32, Jacobi-PCG to 1e-8 ----------------------------
@pytest.mark.parametrize("t,nx,ny,nz", [(gf.GF_Q4, 64, 32, 0), (gf.GF_Q8, 32, 16, 0),
                                        (gf.GF_BRICK8, 8, 8, 8), (gf.GF_BRICK20, 4, 4, 4)])
def test_pcg_parity_assembled(gpu_ctx, t, nx, ny, nz):
    p = P.structured(t, nx, ny, nz)
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    gpu_ctx.apply_penalty(K, p.fixed_dofs, 1e4)
    x, st = gpu_ctx.pcg(K, p.f, rtol=1e-8)
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    O.apply_penalty(Ko, p.dim, p.supports)
    xo, ito, relo, _ = O.pcg_jacobi(Ko, p.f, rtol=1e-8)
    assert st.converged and st.true_relres <= 1.2e-8
    assert abs(st.iterations - ito) <= max(2, ito // 100)
    # same algorithm, same stopping rule: displacements within 1e-9 relative
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


def test_pcg_schedules_bitwise_identical(gpu_ctx):
    p = P.config_q4_plate()
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    gpu_ctx.apply_penalty(K, p.fixed_dofs)
    x1, s1 = gpu_ctx.pcg(K, p.f, schedule=gf.GF_SCHED_KERNELS)
    x2, s2 = gpu_ctx.pcg(K, p.f, schedule=gf.GF_SCHED_GRAPH)
    x3, s3 = gpu_ctx.pcg(K, p.f, schedule=gf.GF_SCHED_GRAPH)
    assert s1.iterations == s2.iterations == s3.iterations
    assert np.array_equal(x1, x2) and np.array_equal(x2, x3)


def test_pcg_multi_rhs_and_warm_start(gpu_ctx):
    p = P.structured(gf.GF_Q4, 24, 12)
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    gpu_ctx.apply_penalty(K, p.fixed_dofs)
    F = P.random_top_loads(p, 3, 99)
    X, st = gpu_ctx.pcg(K, F, batched=0)          # one right-hand side after another
    for k in range(3):
        xk, sk = gpu_ctx.pcg(K, F[k])
        assert np.array_equal(X[k], xk)
    xw, sw = gpu_ctx.pcg(K, F[0], x0=X[0])
    assert sw.iterations == 0


@pytest.mark.parametrize("elem,variant", [(gf.GF_BRICK20, 0), (gf.GF_BRICK20, 8),
                                          (gf.GF_Q8, 0), (gf.GF_Q8, 7), (gf.GF_Q8, 9)])
@pytest.mark.parametrize("nrhs", [2, 5, 16, 32, 40])
def test_batched_pcg_matches_column_solves(gpu_ctx, nrhs, elem, variant):
    """All load cases advance together (one SpMM per iteration, on the ELL
    arrays or on the block-sliced copy; 8, 16, 32 or 64 columns wide); each
    column keeps its own scalars, so it follows its single-RHS iteration."""
    p = (P.structured(gf.GF_BRICK20, 3, 3, 2) if elem == gf.GF_BRICK20
         else P.structured(gf.GF_Q8, 9, 5))
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    gpu_ctx.apply_penalty(K, p.fixed_dofs)
    gpu_ctx.tune(K, variant)
    F = P.random_top_loads(p, nrhs, 7)
    F[nrhs // 2] = 0.0                            # an empty load case converges at once
    X, st = gpu_ctx.pcg(K, F)
    assert st.converged and st.true_relres <= 1.05e-8
    assert np.all(X[nrhs // 2] == 0)
    Xw, sw = gpu_ctx.pcg(K, F, x0=X)
    assert sw.iterations == 0
    # against the oracle both solve to 1e-10, so that the two iterations'
    # different rounding stays below the 1e-9 displacement bar
    Xt, stt = gpu_ctx.pcg(K, F, rtol=1e-10)
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    O.apply_penalty(Ko, p.dim, p.supports)
    for k in range(nrhs):
        xo, ito, _, _ = O.pcg_jacobi(Ko, F[k], rtol=1e-10)
        assert np.linalg.norm(Xt[k] - xo) <= 1e-9 * max(np.linalg.norm(xo), 1e-300)


@pytest.mark.parametrize("precision", [gf.GF_PREC_FP64, gf.GF_PREC_MIXED])
@pytest.mark.parametrize("nrhs", [32, 40])
def test_batched_solves_are_deterministic(gpu_ctx, nrhs, precision):
    """The per-column reductions (one workgroup per load case, the last to
    finish setting the iteration state) add in a fixed order: the same solve
    twice gives the same bits and iteration counts."""
    p = P.structured(gf.GF_BRICK20, 3, 3, 2)
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    gpu_ctx.apply_penalty(K, p.fixed_dofs)
    gpu_ctx.tune(K, 8)
    F = P.random_top_loads(p, nrhs, 5)
    X1, s1 = gpu_ctx.pcg(K, F, precision=precision)
    X2, s2 = gpu_ctx.pcg(K, F, precision=precision)
    assert s1.converged and s1.true_relres <= 1.05e-8
    assert np.array_equal(X1, X2) and s1.iterations == s2.iterations


@pytest.mark.parametrize("elem,variant", [(gf.GF_BRICK20, 0), (gf.GF_BRICK20, 8),
                                          (gf.GF_Q8, 0), (gf.GF_Q8, 7)])
@pytest.mark.parametrize("nrhs", [2, 5, 32, 40])
def test_batched_mixed_meets_fp64_tolerance(gpu_ctx, nrhs, elem, variant):
    """All load cases advance together with FP32 inner iterations (FP32 SpMM
    over the ELL arrays or the block-sliced copy), each column refreshed in
    FP64: every column meets the FP64 true residual, and at a tight
    tolerance the displacements agree with the FP64 oracle to 1e-9."""
    p = (P.structured(gf.GF_BRICK20, 3, 3, 2) if elem == gf.GF_BRICK20
         else P.structured(gf.GF_Q8, 9, 5))
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    gpu_ctx.apply_penalty(K, p.fixed_dofs)
    gpu_ctx.tune(K, variant)
    F = P.random_top_loads(p, nrhs, 11)
    F[nrhs // 2] = 0.0
    X, st = gpu_ctx.pcg(K, F, precision=gf.GF_PREC_MIXED)
    assert st.converged and st.true_relres <= 1.05e-8
    assert st.refinements >= 1 and st.fallbacks == 0
    assert np.all(X[nrhs // 2] == 0)
    X64, s64 = gpu_ctx.pcg(K, F)
    assert st.iterations <= 3 * s64.iterations + 10
    for k in range(nrhs):   # both at 1e-8: differences at the residual level
        assert np.linalg.norm(X[k] - X64[k]) <= 1e-5 * max(np.linalg.norm(X64[k]), 1e-300)
    Xt, stt = gpu_ctx.pcg(K, F, precision=gf.GF_PREC_MIXED, rtol=1e-10)
    assert stt.converged and stt.true_relres <= 1.05e-10
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    O.apply_penalty(Ko, p.dim, p.supports)
    for k in (0, nrhs - 1):
        xo, _, _, _ = O.pcg_jacobi(Ko, F[k], rtol=1e-12)
        assert np.linalg.norm(Xt[k] - xo) <= 1e-9 * np.linalg.norm(xo)
    # warm start from the answer: nothing left to do
    Xw, sw = gpu_ctx.pcg(K, F, x0=Xt, precision=gf.GF_PREC_MIXED, rtol=1e-8)
    assert sw.iterations == 0
    # kernel launches one by one give the graph replay's result
    Xk, sk = gpu_ctx.pcg(K, F, precision=gf.GF_PREC_MIXED, schedule=gf.GF_SCHED_KERNELS)
    assert sk.iterations == st.iterations and np.array_equal(Xk, X)


def test_batched_mixed_falls_back_to_fp64(gpu_ctx):
    """The FP32 rounding of [[1, 1-e], [1-e, 1]] (e = 1e-9) is singular: the
    batched FP32 iteration breaks down and every column is redone by FP64
    batched PCG from zero."""
    e = 1e-9
    val = np.array([[1.0, 1.0], [1 - e, 1 - e]])
    col = np.array([[0, 1], [1, 0]], dtype=np.int32)
    K = gpu_ctx.upload_ell(val, col, np.array([2, 2], dtype=np.int32))
    F = np.array([[1.0, -1.0], [2.0, -2.0], [0.5, 0.25]])
    X, st = gpu_ctx.pcg(K, F, precision=gf.GF_PREC_MIXED, rtol=1e-10)
    assert st.converged and st.fallbacks == 1
    A = np.array([[1.0, 1 - e], [1 - e, 1.0]])
    for k in range(3):
        assert np.allclose(X[k], np.linalg.solve(A, F[k]), rtol=1e-6)


def test_pcg_error_paths(gpu_ctx):
    K = upload_oracle(gpu_ctx, diag_corner(32))
    b = np.arange(1, 33, dtype=float) ** 2
    x, st = gpu_ctx.pcg(K, b, rtol=1e-14, maxit=2, raise_on_fail=False)
    assert st.code == gf.GF_ERR_NOCONV and st.iterations == 2 and not st.converged
    # indefinite matrix: p.Kp <= 0 is detected as breakdown (no fault, no hang)
    n = 4
    val = np.array([[1.0, -1.0, 1.0, 1.0]])
    col = np.arange(n, dtype=np.int32)[None, :]
    Kn = gpu_ctx.upload_ell(val, col, np.ones(n, dtype=np.int32))
    x, st = gpu_ctx.pcg(Kn, np.array([0.0, 1.0, 0.0, 0.0]), precond=gf.GF_PRECOND_NONE,
                        raise_on_fail=False)
    assert st.code == gf.GF_ERR_BREAKDOWN
    x, st = gpu_ctx.pcg(K, np.zeros(32))
    assert st.iterations == 0 and np.all(x == 0)


# ---- SpMV variants, schedules, mixed precision ------------------------------------------
def assembled(ctx, t, nx, ny, nz=0):
    p = P.structured(t, nx, ny, nz)
    mesh = ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = ctx.symbolic(mesh)
    ctx.assemble(mesh, K)
    ctx.apply_penalty(K, p.fixed_dofs)
    return p, K


def test_spmv_variants_match_oracle(gpu_ctx):
    g = gf.Grid(gf.GF_BRICK20, 5, 4, 3)
    conn, coords = g.fill()
    Ko, _ = O.assemble(3, 3, conn, coords, 210000.0, 0.3, skip_zero=False)
    K = upload_oracle(gpu_ctx, Ko)
    x = np.random.default_rng(5).normal(size=g.ndof)
    yo = Ko.axy(x)
    ys = []
    for v in range(5):
        gpu_ctx.tune(K, v)
        y, _ = gpu_ctx.spmv(K, x)
        assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max(), v
        ys.append(y)
    # one thread per row: same summation order whatever the unroll
    assert np.array_equal(ys[0], ys[1]) and np.array_equal(ys[1], ys[2])
    chosen, ms = gpu_ctx.tune(K)
    # an uploaded matrix carries no node structure: the node variants are skipped
    assert 0 <= chosen < 5 and np.all(ms[:5] > 0) and np.all(ms[5:] == -1)
    with pytest.raises(gf.GpufemError):
        gpu_ctx.tune(K, 5)


@pytest.mark.parametrize("et,dims,node_variant", [(gf.GF_Q8, (9, 5), 5), (gf.GF_Q4, (13, 7), 5),
                                                  (gf.GF_BRICK8, (6, 5, 4), 6),
                                                  (gf.GF_BRICK20, (4, 3, 3), 6)])
def test_node_block_spmv_is_bitwise_thread_per_row(gpu_ctx, et, dims, node_variant):
    p, K = assembled(gpu_ctx, et, *dims)
    x = np.random.default_rng(9).normal(size=p.ndof)
    gpu_ctx.tune(K, 0)
    y0, _ = gpu_ctx.spmv(K, x)
    gpu_ctx.tune(K, node_variant)
    y1, _ = gpu_ctx.spmv(K, x)
    assert np.array_equal(y0, y1)
    with pytest.raises(gf.GpufemError):   # the other dimension's variant does not fit
        gpu_ctx.tune(K, 11 - node_variant)
    x1, s1 = gpu_ctx.pcg(K, p.f, schedule=gf.GF_SCHED_KERNELS)
    gpu_ctx.tune(K, 0)
    x0, s0 = gpu_ctx.pcg(K, p.f, schedule=gf.GF_SCHED_KERNELS)
    assert s0.iterations == s1.iterations and np.array_equal(x0, x1)
    chosen, ms = gpu_ctx.tune(K)
    assert ms[node_variant] > 0 and ms[11 - node_variant] == -1


@pytest.mark.parametrize("et,dims,variant", [(gf.GF_Q8, (9, 5), 7), (gf.GF_Q4, (70, 33), 7),
                                             (gf.GF_BRICK8, (6, 5, 4), 8),
                                             (gf.GF_BRICK20, (4, 3, 3), 8),
                                             (gf.GF_BRICK8, (17, 16, 15), 8),
                                             (gf.GF_Q8, (9, 5), 9), (gf.GF_Q4, (70, 33), 9),
                                             (gf.GF_Q8, (130, 70), 9),
                                             (gf.GF_BRICK20, (4, 3, 3), 10),
                                             (gf.GF_BRICK8, (17, 16, 15), 10),
                                             (gf.GF_Q4, (1, 1), 9), (gf.GF_Q8, (1, 1), 7),
                                             (gf.GF_BRICK8, (1, 1, 1), 10),
                                             (gf.GF_BRICK20, (1, 1, 1), 8)])
def test_block_sliced_spmv_is_bitwise_thread_per_row(gpu_ctx, et, dims, variant):
    """Variants 7/8 sweep the slices with the PCG grid, 9/10 give each group
    of four slices its own workgroup; all add a row's entries in slot order."""
    other = {7: 8, 8: 7, 9: 10, 10: 9}[variant]    # the other dimension's variant
    p, K = assembled(gpu_ctx, et, *dims)
    x = np.random.default_rng(10).normal(size=p.ndof)
    gpu_ctx.tune(K, 0)
    y0, _ = gpu_ctx.spmv(K, x)
    gpu_ctx.tune(K, variant)
    y1, _ = gpu_ctx.spmv(K, x)
    assert np.array_equal(y0, y1)
    with pytest.raises(gf.GpufemError):
        gpu_ctx.tune(K, other)
    # the copy follows re-assembly and the penalty (value versions); the dot
    # products' partials differ, so compare two solves to 1e-10
    x1, s1 = gpu_ctx.pcg(K, p.f, rtol=1e-10, schedule=gf.GF_SCHED_KERNELS)
    gpu_ctx.tune(K, 0)
    x0, s0 = gpu_ctx.pcg(K, p.f, rtol=1e-10, schedule=gf.GF_SCHED_KERNELS)
    assert abs(s0.iterations - s1.iterations) <= 1 + s0.iterations // 100
    assert np.linalg.norm(x1 - x0) <= 1e-9 * np.linalg.norm(x0)
    gpu_ctx.tune(K, variant)
    xm, sm = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED)
    assert sm.converged and sm.true_relres <= 1e-8
    chosen, ms = gpu_ctx.tune(K)
    assert ms[variant] > 0 and ms[other] == -1


def test_block_sliced_copy_tracks_new_values(gpu_ctx):
    p = P.structured(gf.GF_Q8, 12, 6)
    mesh = gpu_ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu)
    K = gpu_ctx.symbolic(mesh)
    gpu_ctx.assemble(mesh, K)
    gpu_ctx.tune(K, 7)
    x = np.random.default_rng(3).normal(size=p.ndof)
    ya, _ = gpu_ctx.spmv(K, x)
    gpu_ctx.apply_penalty(K, p.fixed_dofs)   # changes diagonal values
    yb, _ = gpu_ctx.spmv(K, x)
    gpu_ctx.tune(K, 0)
    yc, _ = gpu_ctx.spmv(K, x)
    assert not np.array_equal(ya, yb) and np.array_equal(yb, yc)


@pytest.mark.parametrize("et,dims,variant", [(gf.GF_Q4, (64, 32), v) for v in (0, 2, 3, 4, 5, 7, 9)]
                         + [(gf.GF_BRICK8, (9, 8, 7), v) for v in (6, 8, 10)])
def test_persistent_matches_kernel_schedule(gpu_ctx, et, dims, variant):
    """The persistent kernel plays the SpMV kernel's logical workgroups in
    turn (more of them than co-resident ones for variants 9/10), so both
    schedules reduce the same partials in the same order."""
    p, K = assembled(gpu_ctx, et, *dims)
    gpu_ctx.tune(K, variant)
    x1, s1 = gpu_ctx.pcg(K, p.f, schedule=gf.GF_SCHED_KERNELS)
    x2, s2 = gpu_ctx.pcg(K, p.f, schedule=gf.GF_SCHED_PERSISTENT)
    assert s1.iterations == s2.iterations and s2.converged
    assert np.array_equal(x1, x2)


def test_profile_barrier(gpu_ctx):
    """Grid barriers of the persistent schedule on its own co-resident grid:
    a positive cost per barrier, a grid of whole XCD rounds, and repeated
    calls reset the barrier (no wait runs into its spin limit)."""
    p, K = assembled(gpu_ctx, gf.GF_BRICK8, 9, 8, 7)
    gpu_ctx.tune(K, 10)
    for nbar in (200, 50):
        us, nb = gpu_ctx.profile_barrier(K, nbar)
        assert 0.0 < us < 1000.0 and nb > 0 and nb % 8 == 0
    with pytest.raises(gf.GpufemError):
        gpu_ctx.profile_barrier(K, 0)


def test_profile_pcg_times_the_solver_spmv(gpu_ctx):
    """Events around each SpMV of real PCG iterations: the kernel's share of
    an iteration; a system solved exactly before the count is an error."""
    p, K = assembled(gpu_ctx, gf.GF_Q4, 64, 32)
    ms_spmv, ms_it = gpu_ctx.profile_pcg(K, p.f, 40)
    assert 0.0 < ms_spmv < ms_it
    n = 4
    Kd = gpu_ctx.upload_ell(np.full((1, n), 2.0), np.arange(n, dtype=np.int32)[None, :],
                            np.ones(n, dtype=np.int32))
    with pytest.raises(gf.GpufemError):
        gpu_ctx.profile_pcg(Kd, np.ones(n), 5)


def test_persistent_breakdown_and_maxit(gpu_ctx):
    n = 4
    val = np.array([[1.0, -1.0, 1.0, 1.0]])
    col = np.arange(n, dtype=np.int32)[None, :]
    Kn = gpu_ctx.upload_ell(val, col, np.ones(n, dtype=np.int32))
    x, st = gpu_ctx.pcg(Kn, np.array([0.0, 1.0, 0.0, 0.0]), precond=gf.GF_PRECOND_NONE,
                        schedule=gf.GF_SCHED_PERSISTENT, raise_on_fail=False)
    assert st.code == gf.GF_ERR_BREAKDOWN
    p, K = assembled(gpu_ctx, gf.GF_Q4, 16, 8)
    x, st = gpu_ctx.pcg(K, p.f, maxit=7, schedule=gf.GF_SCHED_PERSISTENT, raise_on_fail=False)
    assert st.code == gf.GF_ERR_NOCONV and st.iterations == 7


@pytest.mark.parametrize("t,nx,ny,nz", [(gf.GF_Q4, 64, 32, 0), (gf.GF_Q8, 32, 16, 0),
                                        (gf.GF_BRICK8, 8, 8, 8), (gf.GF_BRICK20, 4, 4, 4)])
def test_mixed_precision_meets_fp64_tolerance(gpu_ctx, t, nx, ny, nz):
    """FP32 inner iterations inside FP64 refinement reach the same true
    residual as the FP64 solve; at a tight tolerance the displacements agree
    with the FP64 oracle's to 1e-9 relative."""
    p, K = assembled(gpu_ctx, t, nx, ny, nz)
    xm, sm = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED, rtol=1e-8)
    assert sm.converged and sm.true_relres <= 1.05e-8 and sm.refinements >= 1
    x64, s64 = gpu_ctx.pcg(K, p.f, rtol=1e-8)
    # FP32 rounding of the operator and vectors delays CG on these penalized
    # systems (about 2x on the plates); the bound catches a broken refinement
    assert sm.iterations <= 3 * s64.iterations + 10
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    O.apply_penalty(Ko, p.dim, p.supports)
    xo, _, _, _ = O.pcg_jacobi(Ko, p.f, rtol=1e-12)
    xt, st = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED, rtol=1e-10)
    assert st.converged and st.true_relres <= 1.05e-10
    assert np.linalg.norm(xt - xo) <= 1e-9 * np.linalg.norm(xo)


def test_mixed_graph_replay_matches_kernel_launches(gpu_ctx):
    """The mixed solve's captured iterations (default schedule) give the
    bitwise result of launching them one by one; a second matrix on the same
    plan size recaptures (different arrays) and still agrees."""
    for t, nx, ny in [(gf.GF_Q4, 24, 12), (gf.GF_Q8, 16, 8)]:
        p, K = assembled(gpu_ctx, t, nx, ny)
        xg, sg = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED)
        xk, sk = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED, schedule=gf.GF_SCHED_KERNELS)
        xg2, sg2 = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED, schedule=gf.GF_SCHED_GRAPH)
        assert sg.iterations == sk.iterations == sg2.iterations
        assert np.array_equal(xg, xk) and np.array_equal(xg, xg2)
        assert sg.true_relres <= 1e-8


def test_mixed_warm_start_and_zero_rhs(gpu_ctx):
    p, K = assembled(gpu_ctx, gf.GF_Q4, 24, 12)
    x, s = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED)
    xw, sw = gpu_ctx.pcg(K, p.f, x0=x, precision=gf.GF_PREC_MIXED)
    assert sw.iterations == 0
    xz, sz = gpu_ctx.pcg(K, np.zeros(p.ndof), precision=gf.GF_PREC_MIXED)
    assert sz.iterations == 0 and np.all(xz == 0)


def test_mixed_falls_back_to_fp64_when_fp32_operator_is_singular(gpu_ctx):
    """[[1, 1-e], [1-e, 1]] with e = 1e-9 is SPD in FP64 but rounds to a
    singular matrix in FP32: the FP32 iteration breaks down and the solve is
    redone by zero-start FP64 PCG."""
    e = 1e-9
    val = np.array([[1.0, 1.0], [1 - e, 1 - e]])
    col = np.array([[0, 1], [1, 0]], dtype=np.int32)
    K = gpu_ctx.upload_ell(val, col, np.array([2, 2], dtype=np.int32))
    b = np.array([1.0, -1.0])
    x, st = gpu_ctx.pcg(K, b, precision=gf.GF_PREC_MIXED, rtol=1e-10)
    assert st.converged and st.fallbacks == 1
    assert np.allclose(x, np.array([1.0, -1.0]) / e, rtol=1e-6)
