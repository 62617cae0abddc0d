"""Neutral-file reader (gpufem_amd/nf.py, after src/Utils/fileIO.cc:73-404)
on the reference's own meshes (tests/golden/nf/, copied by fetch_fixtures.sh),
checked with the CPU oracle: the properties the reference's FemTest asserts
on the same files (src/Test/FemTest.cpp:127-181), positive Jacobians, and
the Brick20 node permutation against the structured generator's node order."""
import os

import numpy as np
import pytest

import gpufem_amd as gf
import oracle as O
from gpufem_amd import nf

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nf")


def path(name):
    return os.path.join(HERE, name)


def jacobian_dets(p):
    pts, _ = O.gauss_points(p.dim, p.ngauss)
    dets = []
    for e in range(p.conn.shape[0]):
        xe = p.coords[p.conn[e] - 1].T
        for g in pts:
            dN = O.dshape(p.dim, p.grid.nen, g)
            dets.append(O.det(O.jacobian(p.dim, p.grid.nen, xe, dN)))
    return np.array(dets)


def test_q4_plate_file_blocks():
    p = nf.read_nf(path("placa2_Q4.nf"))
    assert p.elem_type == gf.GF_Q4 and p.dim == 2 and p.ngauss == 2
    assert p.conn.shape[1] == 4 and p.coords.shape[1] == 2
    assert p.supports.shape[1] == 3 and p.loads.shape[1] == 3
    assert np.all(jacobian_dets(p) > 0)
    # the file's loads land in the force vector by assignment
    assert np.array_equal(p.f, O.force_vector(2, p.coords.shape[0], p.loads))


def test_duplicate_load_rows_keep_nonzero_components():
    text = """%ELEMENT.Q4
1
1 1 1 1 1 2 3 4
%NODE.COORD
4
1 0 0
2 1 0
3 1 1
4 0 1
%NODE.SUPPORT
1
1 1 1
%LOAD.CASE.NODAL.FORCES
3
3 1.5 0.0
2 0.0 -2.0
3 0.0 7.0
%END
"""
    p = nf.parse_nf(text)
    assert np.array_equal(p.f, O.force_vector(2, 4, p.loads))
    assert list(p.f[4:6]) == [1.5, 7.0]


def test_reference_femtest_properties_q4():
    # FemTest macro_test_2dstiffness_diag_nonzero / macro_test_stiffness_constraint
    p = nf.read_nf(path("placa2_Q4.nf"))
    K, _ = O.assemble(2, p.ngauss, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    k00 = K.get(0, 0)
    O.apply_penalty(K, 2, p.supports)
    assert k00 <= K.get(0, 0)


@pytest.mark.parametrize("name", ["brick_brk20v2.nf", "brick_brk20.nf"])
def test_brick20_file_diag_positive(name):
    # FemTest macro_test_3dstiffness_diag_nonzero
    p = nf.read_nf(path(name))
    assert p.elem_type == gf.GF_BRICK20 and p.conn.shape == (1, 20)
    assert np.all(jacobian_dets(p) > 0)
    K, _ = O.assemble(3, p.ngauss, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    assert all(K.get(i, i) > 0 for i in range(p.ndof))


def test_brick20_permutation_matches_element_node_order():
    # the file's unit-cube element, permuted, puts every local node where the
    # structured generator's unit-cube element has it
    p = nf.read_nf(path("brick_brk20.nf"))
    g = gf.Grid(gf.GF_BRICK20, 1, 1, 1)
    conn, coords = g.fill()
    assert np.allclose(p.coords[p.conn[0] - 1], coords[conn[0] - 1])
    # and the element stiffness matrices agree entry by entry
    Ka = O.element_stiffness(3, 20, 3, p.coords[p.conn[0] - 1].T, 210000.0, 0.3)
    Kb = O.element_stiffness(3, 20, 3, coords[conn[0] - 1].T, 210000.0, 0.3)
    assert np.allclose(Ka, Kb, rtol=1e-13, atol=1e-9)


def test_brick20_without_permutation_gives_another_element():
    # the file's raw node order (a plausible reader mistake) still has detJ > 0
    # on this cube but is a different element: its stiffness is far off
    p = nf.read_nf(path("brick_brk20.nf"))
    raw = p.conn[0][np.argsort(nf.BRICK20_FROM_FILE)]
    Ka = O.element_stiffness(3, 20, 3, p.coords[p.conn[0] - 1].T, 210000.0, 0.3)
    Kr = O.element_stiffness(3, 20, 3, p.coords[raw - 1].T, 210000.0, 0.3)
    assert np.abs(Ka - Kr).max() > 0.1 * np.abs(Ka).max()


def test_block_search_rewinds_once_at_end():
    text = """%NODE.COORD
2
1 0 0
2 1 0
%ELEMENT.Q4
1
1 1 1 1 1 2 2 1
%END
"""
    with pytest.raises(nf.NfError):
        nf.parse_nf(text)      # no support block anywhere
    ok = text.replace("%END", "%NODE.SUPPORT\n0\n%LOAD.CASE.NODAL.FORCES\n0\n%END")
    p = nf.parse_nf(ok)
    assert p.coords.shape == (2, 2) and p.supports.shape == (0, 3)


def test_rubik10_brick8_solves():
    p = nf.read_nf(path("rubik10_brk8.nf"))
    assert p.elem_type == gf.GF_BRICK8 and p.conn.shape == (1000, 8) and p.coords.shape[0] == 1331
    assert np.all(jacobian_dets(p) > 0)
    K, _ = O.assemble(3, 2, p.conn, p.coords, p.E, p.nu, skip_zero=False)
    O.apply_penalty(K, 3, p.supports)
    x, it, rel, _ = O.pcg_jacobi(K, p.f, rtol=1e-8)
    assert rel <= 1e-8 and it > 10 and np.linalg.norm(x) > 0
