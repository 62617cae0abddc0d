"""gf_pcg_noalloc (the allocation-free Jacobi-PCG over a static device
scratch array) against gf_pcg with the same SpMV variant and schedule:
bitwise-equal solutions and iteration counts, on two calls in a row, the
second on a smaller matrix; a matrix above the scratch array's limit is
rejected with an error."""
import ctypes as C

import numpy as np
import pytest

import gpufem_amd as gf
from gpufem_amd import DevBuf, _lib
from gpufem_amd import problems as P

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def noalloc(ctx, K, d_b, d_x, **kw):
    fn = ctx.lib.gf_pcg_noalloc
    o = gf.solve_opts(device_ptrs=1, **kw)
    st = _lib.SolveStats()
    code = fn(ctx.handle, K.handle, d_b.ptr, d_x.ptr, C.byref(o), C.byref(st))
    st.code = code
    return st


def assembled(ctx, p):
    mesh = ctx.upload_mesh(p.elem_type, p.ngauss, p.conn, p.coords, p.E, p.nu, elem_E=p.elem_E)
    K = ctx.symbolic(mesh)
    ctx.assemble(mesh, K)
    ctx.apply_penalty(K, p.fixed_dofs)
    return mesh, K


@pytest.mark.parametrize("variant", [0, 10])
def test_noalloc_matches_pcg_on_two_calls(gpu_ctx, variant):
    big = P.config_brick8_cube(24)          # 46,875 DOF
    small = P.structured(gf.GF_BRICK8, 7, 5, 4)
    for p in (big, small):
        mesh, K = assembled(gpu_ctx, p)
        gpu_ctx.tune(K, variant)
        d_b = DevBuf.from_array(gpu_ctx, p.f)
        d_x = DevBuf(gpu_ctx, p.ndof * 8)
        x_ref, s_ref = gpu_ctx.pcg(K, p.f, schedule=gf.GF_SCHED_KERNELS)
        st = noalloc(gpu_ctx, K, d_b, d_x)
        assert st.code == gf.GF_OK, gpu_ctx.lib.gf_last_error(gpu_ctx.handle)
        x = d_x.download(np.float64)
        assert st.converged and st.iterations == s_ref.iterations
        assert st.true_relres <= 1e-8
        assert np.array_equal(x, x_ref), p.name
        # warm start from the answer: no iteration left
        st2 = noalloc(gpu_ctx, K, d_b, d_x, use_x0=1)
        assert st2.code == gf.GF_OK and st2.iterations == 0


def test_noalloc_without_tuned_copies_and_limits(gpu_ctx):
    """A matrix never solved: one thread per ELL row (agrees with gf_pcg to
    rounding); a matrix above the row limit is rejected."""
    p = P.structured(gf.GF_Q8, 20, 10)
    mesh, K = assembled(gpu_ctx, p)
    d_b = DevBuf.from_array(gpu_ctx, p.f)
    d_x = DevBuf(gpu_ctx, p.ndof * 8)
    st = noalloc(gpu_ctx, K, d_b, d_x, rtol=1e-10)
    assert st.code == gf.GF_OK and st.converged
    x = d_x.download(np.float64)
    x_ref, _ = gpu_ctx.pcg(K, p.f, rtol=1e-10)
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    assert noalloc(gpu_ctx, K, d_b, d_x, precision=gf.GF_PREC_MIXED).code == gf.GF_ERR_UNSUPPORTED
    big = P.structured(gf.GF_Q4, 730, 719)    # 1,052,640 DOF > 2^20 rows
    assert big.ndof > MAX_ROWS
    mesh2, K2 = assembled(gpu_ctx, big)
    d_b2 = DevBuf.from_array(gpu_ctx, big.f)
    d_x2 = DevBuf(gpu_ctx, big.ndof * 8)
    assert noalloc(gpu_ctx, K2, d_b2, d_x2).code == gf.GF_ERR_UNSUPPORTED
