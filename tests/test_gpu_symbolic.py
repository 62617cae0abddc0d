"""The device symbolic phase (gf_ell_symbolic_device) against the host
phase: identical structure, first-fit colors (the reference's femColor
order) and slot map, so the assembled matrix is bitwise the same; on
structured meshes, shuffled connectivity and the reference's own
unstructured neutral-file mesh."""
import os

import numpy as np
import pytest

import gpufem_amd as gf
import oracle as O
from gpufem_amd import nf

pytestmark = pytest.mark.gpu

NF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nf")
CASES = [(gf.GF_Q4, 64, 32, 0), (gf.GF_Q8, 24, 12, 0), (gf.GF_BRICK8, 8, 6, 5),
         (gf.GF_BRICK20, 4, 3, 3)]


def _both(ctx, t, ng, conn, coords, nnodes):
    mesh = ctx.upload_mesh(t, ng, conn, coords, 210000.0, 0.3)
    Kh = ctx.symbolic(mesh)
    Kd = ctx.symbolic(mesh, device=True)
    assert (Kd.n, Kd.width, Kd.nnz, Kd.ncolors) == (Kh.n, Kh.width, Kh.nnz, Kh.ncolors)
    ch, cd = ctx.colors(mesh, Kh), ctx.colors(mesh, Kd)
    assert np.array_equal(ch, cd)
    col_o, nc = O.greedy_coloring(conn, nnodes)
    assert np.array_equal(cd, col_o) and Kd.ncolors == nc
    ctx.assemble(mesh, Kh)
    ctx.assemble(mesh, Kd)
    vh, colh, rnh = ctx.download(Kh)
    vd, cold, rnd = ctx.download(Kd)
    assert np.array_equal(rnh, rnd) and np.array_equal(colh, cold)
    assert np.array_equal(vh, vd)          # same colors and slots: bitwise the same sums
    return Kd


@pytest.mark.parametrize("t,nx,ny,nz", CASES)
@pytest.mark.parametrize("shuffle", [False, True])
def test_device_symbolic_equals_host(gpu_ctx, t, nx, ny, nz, shuffle):
    g = gf.Grid(t, nx, ny, nz)
    conn, coords = g.fill()
    if shuffle:
        conn = conn[np.random.default_rng(5).permutation(conn.shape[0])]
    ng = 3 if t in (gf.GF_Q8, gf.GF_BRICK20) else 2
    Kd = _both(gpu_ctx, t, ng, np.ascontiguousarray(conn), coords, g.nnodes)
    assert Kd.symbolic_timing.ms_device > 0


@pytest.mark.parametrize("name", ["rubik10_brk8.nf", "placa2_Q4.nf", "brick_brk20.nf"])
def test_device_symbolic_on_reference_meshes(gpu_ctx, name):
    p = nf.read_nf(os.path.join(NF, name))
    _both(gpu_ctx, p.elem_type, p.ngauss, p.conn, p.coords, p.coords.shape[0])


def test_device_symbolic_isolated_node(gpu_ctx):
    # a node no element uses keeps its diagonal slot, as in the host phase
    g = gf.Grid(gf.GF_Q4, 3, 2)
    conn, coords = g.fill()
    coords = np.vstack([coords, [[9.0, 9.0]]])
    _both(gpu_ctx, gf.GF_Q4, 2, conn, coords, coords.shape[0])
