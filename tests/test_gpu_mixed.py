"""GPU tests of the mixed-precision solves: the FP32 block-sliced copy in
difference form, the accuracy bar (displacements within 1e-9 of the FP64
solution converged to 1e-12) at the operating tolerance the forms quote,
and the FP64 continuation after an FP32 iteration stalls."""
import numpy as np
import pytest

import gpufem_amd as gf
import oracle as O
from gpufem_amd import forms
from gpufem_amd import problems as P

pytestmark = pytest.mark.gpu


def setup(ctx, p, variant=None):
    mesh = ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E)
    K = ctx.symbolic(mesh)
    ctx.assemble(mesh, K)
    ctx.apply_penalty(K, p.fixed_dofs)
    ctx.tune(K, -1 if variant is None else variant)
    return mesh, K


def oracle_solution(p, rtol=1e-12):
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E,
                       skip_zero=False)
    O.apply_penalty(Ko, p.dim, p.supports)
    return O.pcg_jacobi(Ko, p.f, rtol=rtol)[0]


@pytest.mark.parametrize("variant", [7, 9])
def test_difference_form_keeps_fp32_iterations_near_fp64(gpu_ctx, variant):
    """On a uniform structured plate a plain FP32 copy of the matrix repeats
    the same rounding in every row, which shifts every eigenvalue alike
    (about -3e-8 of the Jacobi-scaled spectrum, next to the plate's lowest
    4.5e-6 here), and FP32 Jacobi-PCG needed 3.1x the FP64 iterations (CPU
    emulation); the difference form (block row sums stored, off-diagonal
    blocks acting on x_J - x_R) keeps the spectrum and needs 1.8x."""
    p = P.config_q8_plate(64, 32)
    mesh, K = setup(gpu_ctx, p, variant)
    x64, s64 = gpu_ctx.pcg(K, p.f)
    xm, sm = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED)
    assert sm.converged and sm.fallbacks == 0 and sm.true_relres <= 1.05e-8
    assert sm.iterations <= 2.4 * s64.iterations, (sm.iterations, s64.iterations)


@pytest.mark.parametrize("name", ["brick8_cube_12", "q8_plate_48x24"])
def test_mixed_meets_bar_at_operating_tolerance(gpu_ctx, name):
    """The mixed form at the tolerance the forms quote it at meets the
    accuracy bar against the oracle's FP64 solution converged to 1e-12."""
    p = P.by_name(name)
    mesh, K = setup(gpu_ctx, p)
    r = forms._mixed_on(gpu_ctx, p, K, name)
    op = r["mixed"]
    assert op["meets_bar"] and op["displacement_rel_err"] <= forms.ERR_BAR
    assert op["true_relres"] <= 1.05 * op["rtol"] and op["fallbacks"] == 0
    xo = oracle_solution(p)
    xm, sm = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED, rtol=op["rtol"])
    assert np.linalg.norm(xm - xo) <= forms.ERR_BAR * np.linalg.norm(xo)


def test_batched_mixed_meets_bar_at_operating_tolerance(gpu_ctx):
    p = P.structured(gf.GF_BRICK20, 4, 4, 3)
    mesh, K = setup(gpu_ctx, p, 8)
    F = P.random_top_loads(p, 8, 21)
    X_ref, s_ref = gpu_ctx.pcg(K, F, rtol=forms.REF_RTOL)

    def solve(rtol):
        X, st = gpu_ctx.pcg(K, F, precision=gf.GF_PREC_MIXED, rtol=rtol)
        return st, X

    op, runs = forms.operating_point(solve, X_ref, "test")
    assert op["meets_bar"] and op["fallbacks"] == 0
    Ko, _ = O.assemble(p.dim, p.ngauss, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    O.apply_penalty(Ko, p.dim, p.supports)
    st, X = solve(op["rtol"])
    for k in (0, 7):
        xo = O.pcg_jacobi(Ko, F[k], rtol=1e-12)[0]
        assert np.linalg.norm(X[k] - xo) <= forms.ERR_BAR * np.linalg.norm(xo)


def test_mixed_ill_conditioned_continues_in_fp64(gpu_ctx):
    """A slender Q8 cantilever (128 x 8) whose Jacobi-scaled condition number
    is 1.5e8 (u32 kappa about 9; CPU emulation: FP32 iterations, even with
    FP64 vectors, stop converging on such plates).  The condition estimate
    from the CG coefficients crosses 0.05/u32 and the solve continues in
    FP64 (from the FP32 iterate or from zero, whichever has the lower energy
    functional): the answer still meets the tolerance and the oracle's
    displacements, and the FP32 part costs at most a few check intervals."""
    p = P.config_q8_plate(128, 8)
    mesh, K = setup(gpu_ctx, p)
    x64, s64 = gpu_ctx.pcg(K, p.f)
    assert s64.converged and s64.true_relres <= 1.05e-8
    xm, sm = gpu_ctx.pcg(K, p.f, precision=gf.GF_PREC_MIXED)
    assert sm.converged and sm.true_relres <= 1.05e-8 and sm.fallbacks == 1
    assert sm.iterations <= s64.iterations + 6 * 256, (sm.iterations, s64.iterations)
    xo = oracle_solution(p)
    assert np.linalg.norm(xm - xo) <= forms.ERR_BAR * np.linalg.norm(xo)
    assert np.linalg.norm(x64 - xo) <= forms.ERR_BAR * np.linalg.norm(xo)
    # the same plate batched: every column meets the tolerance
    F = np.stack([p.f, 0.5 * p.f, -p.f])
    X, st = gpu_ctx.pcg(K, F, precision=gf.GF_PREC_MIXED)
    assert st.converged and st.true_relres <= 1.05e-8 and st.fallbacks == 1
    for k, scale in enumerate([1.0, 0.5, -1.0]):
        assert np.linalg.norm(X[k] - scale * xo) <= forms.ERR_BAR * np.linalg.norm(xo)


def test_guarded_initial_guess(gpu_ctx):
    """use_x0 = 2 keeps a guess whose residual is below ||b|| and replaces a
    worse one by zero (the zero-start answer, bitwise)."""
    p = P.structured(gf.GF_BRICK8, 6, 5, 4)
    mesh, K = setup(gpu_ctx, p)
    x0, s0 = gpu_ctx.pcg(K, p.f, schedule=gf.GF_SCHED_KERNELS)
    good = x0 * (1 + 1e-3)
    xg, sg = gpu_ctx.pcg(K, p.f, x0=good, use_x0=2, schedule=gf.GF_SCHED_KERNELS)
    assert sg.x0_rejected == 0 and sg.iterations < s0.iterations
    bad = np.random.default_rng(1).normal(size=p.ndof) * np.abs(x0).max()
    xb, sb = gpu_ctx.pcg(K, p.f, x0=bad, use_x0=2, schedule=gf.GF_SCHED_KERNELS)
    assert sb.x0_rejected == 1 and sb.iterations == s0.iterations
    assert np.array_equal(xb, x0)
