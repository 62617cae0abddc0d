"""Pins the oracle's element-level functions with checks that do not restate
its formulas: quadrature exactness, compliance inversion, finite differences
of the textbook shape functions, exact integration of the rectangular
bilinear/trilinear stiffness, invariants (symmetry, rigid-body null space,
rank, linearity in E) and agreement with a higher-order rule."""
import itertools

import numpy as np
import pytest

import oracle as O

ELEMS = [(2, 4), (2, 8), (3, 8), (3, 20)]


# ---- Gauss rule (src/FEM/FemData.cc:226-316) --------------------------------------
@pytest.mark.parametrize("dim,n", [(2, 2), (2, 3), (3, 2), (3, 3)])
def test_gauss_rule_exactness(dim, n):
    pts, w = O.gauss_points(dim, n)
    assert pts.shape == (n ** dim, dim)
    assert abs(w.sum() - 2.0 ** dim) < 1e-14
    # exact for every monomial of degree <= 2n-1 per coordinate
    for powers in itertools.product(range(2 * n), repeat=dim):
        exact = np.prod([0.0 if p % 2 else 2.0 / (p + 1) for p in powers])
        approx = np.sum(w * np.prod(pts ** np.array(powers), axis=1))
        assert abs(approx - exact) < 1e-13
    # ... and not for degree 2n (so it is the n-point rule, not a finer one)
    p = [2 * n] + [0] * (dim - 1)
    exact = 2.0 / (2 * n + 1) * 2.0 ** (dim - 1)
    assert abs(np.sum(w * np.prod(pts ** np.array(p), axis=1)) - exact) > 1e-3


def test_gauss_point_order_first_coordinate_slowest():
    pts, _ = O.gauss_points(3, 2)
    a = 1 / np.sqrt(3)
    expect = np.array(list(itertools.product([-a, a], repeat=3)))
    assert np.allclose(pts, expect, atol=1e-15)


# ---- constitutive matrix (src/FEM/FemData.cc:322-365) -------------------------------
@pytest.mark.parametrize("E,nu", [(210000.0, 0.3), (1.0, 0.0), (100.0, 0.25)])
def test_constitutive_inverts_compliance(E, nu):
    # plane stress compliance: e = S s with S from uniaxial tests
    S2 = np.array([[1, -nu, 0], [-nu, 1, 0], [0, 0, 2 * (1 + nu)]]) / E
    assert np.allclose(O.constitutive(2, E, nu) @ S2, np.eye(3), atol=1e-13)
    S3 = np.zeros((6, 6))
    S3[:3, :3] = -nu
    np.fill_diagonal(S3[:3, :3], 1.0)
    S3[3:, 3:] = np.eye(3) * 2 * (1 + nu)
    S3 /= E
    assert np.allclose(O.constitutive(3, E, nu) @ S3, np.eye(6), atol=1e-13)


# ---- shape functions (src/FEM/StiffAlgo.cc:47-202) ----------------------------------
def textbook_N(dim, nen, xi_nodes, x):
    """Lagrange (Q4/Brick8) or serendipity (Q8/Brick20) shape functions."""
    N = np.zeros(nen)
    for a in range(nen):
        an = xi_nodes[a]
        lin = 1.0 + x * an
        if nen in (4,) or (dim == 3 and nen == 8):
            N[a] = np.prod(lin) / 2.0 ** dim
        elif np.all(an != 0):
            N[a] = np.prod(lin) / 2.0 ** dim * (np.dot(x, an) - (dim - 1))
        else:
            m = int(np.nonzero(an == 0)[0][0])
            others = [i for i in range(dim) if i != m]
            N[a] = (1 - x[m] ** 2) * np.prod(lin[others]) / 2.0 ** (dim - 1)
    return N


@pytest.mark.parametrize("dim,nen", ELEMS)
def test_shape_kronecker_at_nodes(dim, nen):
    xi = O.node_natural_coords(dim, nen)
    M = np.array([textbook_N(dim, nen, xi, xi[b]) for b in range(nen)])
    assert np.allclose(M, np.eye(nen), atol=1e-15)


@pytest.mark.parametrize("dim,nen", ELEMS)
def test_dshape_matches_finite_differences(dim, nen):
    rng = np.random.default_rng(7)
    xi = O.node_natural_coords(dim, nen)
    h = 1e-5
    for _ in range(20):
        x = rng.uniform(-1, 1, dim)
        dN = O.dshape(dim, nen, x)
        for k in range(dim):
            e = np.zeros(dim)
            e[k] = h
            fd = (textbook_N(dim, nen, xi, x + e) - textbook_N(dim, nen, xi, x - e)) / (2 * h)
            assert np.allclose(dN[k], fd, atol=1e-8), (k, x)


@pytest.mark.parametrize("dim,nen", ELEMS)
def test_dshape_completeness(dim, nen):
    xi = O.node_natural_coords(dim, nen)
    rng = np.random.default_rng(3)
    for _ in range(10):
        x = rng.uniform(-1, 1, dim)
        dN = O.dshape(dim, nen, x)
        assert np.allclose(dN.sum(axis=1), 0.0, atol=1e-14)          # partition of unity
        assert np.allclose(dN @ xi, np.eye(dim), atol=1e-14)          # linear completeness
        if nen in (8, 20) and not (dim == 3 and nen == 8):            # quadratic completeness
            for i in range(dim):
                for j in range(dim):
                    g = dN @ (xi[:, i] * xi[:, j])
                    expect = np.zeros(dim)
                    expect[i] += x[j]
                    expect[j] += x[i]
                    assert np.allclose(g, expect, atol=1e-13)


# ---- Jacobian / det / inverse (src/FEM/StiffAlgo.cc:207-244, LAops.cc:80-117) ---------
@pytest.mark.parametrize("dim,nen", ELEMS)
def test_jacobian_axis_aligned(dim, nen):
    xi = O.node_natural_coords(dim, nen)
    sizes = np.array([2.0, 0.5, 3.0][:dim])
    X = (xi + 1) / 2 * sizes                                          # box [0, sizes]
    dN = O.dshape(dim, nen, np.full(dim, 0.3))
    J = O.jacobian(dim, nen, X.T.copy(), dN)
    assert np.allclose(J, np.diag(sizes / 2), atol=1e-14)
    assert abs(O.det(J) - np.prod(sizes / 2)) < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_inverse_random(dim):
    rng = np.random.default_rng(11)
    for _ in range(20):
        J = rng.normal(size=(dim, dim)) + 3 * np.eye(dim)
        assert np.allclose(O.inverse(J) @ J, np.eye(dim), atol=1e-12)
        assert abs(O.det(J) - np.linalg.det(J)) < 1e-10 * abs(np.linalg.det(J))


# ---- element stiffness (src/FEM/StiffAlgoCPU.cc:210-255) -----------------------------
def exact_box_stiffness(dim, sizes, E, nu):
    """Exact integral of B^T D B for the bilinear/trilinear element on a box,
    assembled from 1-D integrals of linear functions (no quadrature)."""
    xi = O.node_natural_coords(dim, 2 ** dim)
    # 1-D linear functions phi_lo = 1 - t, phi_hi = t on [0, L]
    def m(a, b, L):   # int phi_a phi_b
        return L / 3 if a == b else L / 6
    def k(a, b, L):   # int phi_a' phi_b'
        sa, sb = (1 if a else -1), (1 if b else -1)
        return sa * sb / L
    def g(a, b, L):   # int phi_a' phi_b
        return (1 if a else -1) * 0.5
    hi = (xi > 0).astype(int)                  # which linear function per axis
    nen = 2 ** dim
    # I[p][q][a][b] = int dN_a/dx_p dN_b/dx_q
    I = np.zeros((dim, dim, nen, nen))
    for a in range(nen):
        for b in range(nen):
            for p in range(dim):
                for q in range(dim):
                    v = 1.0
                    for ax in range(dim):
                        L = sizes[ax]
                        ha, hb = hi[a, ax], hi[b, ax]
                        if ax == p and ax == q:
                            v *= k(ha, hb, L)
                        elif ax == p:
                            v *= g(ha, hb, L)
                        elif ax == q:
                            v *= g(hb, ha, L)
                        else:
                            v *= m(ha, hb, L)
                    I[p, q, a, b] = v
    D = O.constitutive(dim, E, nu)
    s = D.shape[0]
    # B rows in terms of derivative index: strain = sum_c G[s_row][c][p] du_c/dx_p
    G = np.zeros((s, dim, dim))
    for c in range(dim):
        G[c, c, c] = 1.0
    if dim == 2:
        G[2, 0, 1] = G[2, 1, 0] = 1.0
    else:
        G[3, 0, 1] = G[3, 1, 0] = 1.0
        G[4, 1, 2] = G[4, 2, 1] = 1.0
        G[5, 0, 2] = G[5, 2, 0] = 1.0
    K = np.zeros((dim * nen, dim * nen))
    for a in range(nen):
        for b in range(nen):
            for c in range(dim):
                for d in range(dim):
                    acc = 0.0
                    for p in range(dim):
                        for q in range(dim):
                            acc += (G[:, c, p] @ D @ G[:, d, q]) * I[p, q, a, b]
                    K[dim * a + c, dim * b + d] = acc
    return K


@pytest.mark.parametrize("dim,sizes", [(2, (1.0, 1.0)), (2, (2.0, 0.5)), (3, (1.0, 1.0, 1.0)),
                                       (3, (1.5, 0.7, 2.0))])
def test_ke_matches_exact_box_integral(dim, sizes):
    nen = 2 ** dim
    xi = O.node_natural_coords(dim, nen)
    X = (xi + 1) / 2 * np.array(sizes)
    Ke = O.element_stiffness(dim, nen, 2, X.T.copy(), 210000.0, 0.3)
    K = exact_box_stiffness(dim, np.array(sizes), 210000.0, 0.3)
    assert np.abs(Ke - K).max() <= 1e-12 * np.abs(K).max()


def distorted_coords(dim, nen, seed, return_map=False):
    """Nodes of an affinely distorted element: X = A xi."""
    xi = O.node_natural_coords(dim, nen)
    rng = np.random.default_rng(seed)
    A = np.eye(dim) + 0.15 * rng.normal(size=(dim, dim))
    X = xi @ A.T
    return (X, A) if return_map else X


def rigid_modes(X):
    nen, dim = X.shape
    modes = []
    for c in range(dim):
        u = np.zeros((nen, dim))
        u[:, c] = 1
        modes.append(u.ravel())
    pairs = [(0, 1)] if dim == 2 else [(0, 1), (1, 2), (0, 2)]
    for i, j in pairs:
        u = np.zeros((nen, dim))
        u[:, i] = -X[:, j]
        u[:, j] = X[:, i]
        modes.append(u.ravel())
    return np.array(modes)


@pytest.mark.parametrize("dim,nen", ELEMS)
@pytest.mark.parametrize("ng", [2, 3])
def test_ke_invariants(dim, nen, ng):
    X = distorted_coords(dim, nen, 5)
    Ke = O.element_stiffness(dim, nen, ng, X.T.copy(), 210000.0, 0.3)
    scale = np.abs(Ke).max()
    assert np.abs(Ke - Ke.T).max() <= 1e-12 * scale
    R = rigid_modes(X)
    assert np.abs(Ke @ R.T).max() <= 1e-10 * scale * np.abs(X).max()
    nrig = 3 if dim == 2 else 6
    sv = np.linalg.svd(Ke, compute_uv=False)
    rank = int(np.sum(sv > 1e-9 * sv[0]))
    full = dim * nen - nrig
    # serendipity with reduced (2-point) integration has spurious zero-energy modes
    if ng == 3 or nen in (4,) or (dim == 3 and nen == 8):
        assert rank == full
    Ke2 = O.element_stiffness(dim, nen, ng, X.T.copy(), 2 * 210000.0, 0.3)
    assert np.abs(Ke2 - 2 * Ke).max() <= 1e-14 * scale


@pytest.mark.parametrize("dim,nen,ng", [(2, 4, 2), (2, 8, 3), (3, 8, 2), (3, 20, 3)])
def test_ke_rule_is_exact_on_affine_elements(dim, nen, ng):
    """On an affine element the integrand is polynomial; the reference rule
    (2 points linear, 3 points quadratic) equals a 5-point rule."""
    X = distorted_coords(dim, nen, 9)
    Ke = O.element_stiffness(dim, nen, ng, X.T.copy(), 1.0, 0.3)
    g5, w5 = np.polynomial.legendre.leggauss(5)
    D = O.constitutive(dim, 1.0, 0.3)
    K5 = np.zeros_like(Ke)
    for idx in itertools.product(range(5), repeat=dim):
        x = g5[list(idx)]
        w = np.prod(w5[list(idx)])
        dN = O.dshape(dim, nen, x)
        J = O.jacobian(dim, nen, X.T.copy(), dN)
        dNdx = O.inverse(J) @ dN
        B = O.bmatrix(dim, nen, dNdx)
        K5 += B.T @ D @ B * O.det(J) * w
    assert np.abs(Ke - K5).max() <= 1e-12 * np.abs(K5).max()


@pytest.mark.parametrize("dim,nen,ng", [(2, 4, 2), (2, 8, 3), (3, 8, 2), (3, 20, 3)])
def test_ke_constant_strain_energy(dim, nen, ng):
    """A linear displacement field u = A x gives constant strain; the element
    strain energy u^T Ke u equals eps^T D eps times the element volume."""
    X, Amap = distorted_coords(dim, nen, 13, return_map=True)
    Ke = O.element_stiffness(dim, nen, ng, X.T.copy(), 210000.0, 0.3)
    rng = np.random.default_rng(1)
    A = rng.normal(size=(dim, dim))
    u = (X @ A.T).ravel()
    eps_t = 0.5 * (A + A.T)
    if dim == 2:
        eps = np.array([eps_t[0, 0], eps_t[1, 1], 2 * eps_t[0, 1]])
    else:
        eps = np.array([eps_t[0, 0], eps_t[1, 1], eps_t[2, 2], 2 * eps_t[0, 1],
                        2 * eps_t[1, 2], 2 * eps_t[0, 2]])
    vol = abs(np.linalg.det(Amap)) * 2 ** dim   # image of the reference cube
    D = O.constitutive(dim, 210000.0, 0.3)
    assert abs(u @ Ke @ u - eps @ D @ eps * vol) <= 1e-10 * abs(eps @ D @ eps * vol)
