"""Pins the oracle's ELL storage and solvers with the reference's own known
answers (tests/golden/reference_kats.json) and with scipy direct solves."""
import json
import os

import numpy as np
import pytest
import scipy.linalg

import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


# ---- ELLmatrix semantics (src/Test/SPRmatrixTest.cpp, SPRSearchTest.cpp) --------------
def test_stepped_search_kat():
    g = GOLD["stepped_search"]
    for val, ln, start, step, expect in g["cases"]:
        assert O.binsearch_step(g["vector"], val, ln, start, step) == expect


def test_set_get_simple():
    m = O.EllMatrix(256)
    assert m.get(0, 0) == 0 and m.get(78, 87) == 0                    # InitializesToZero
    m.set(0, 0, 4); assert m.get(0, 0) == 4
    m.set(2, 9, 8); assert m.get(2, 9) == 8
    m.set(91, 19, 4.56); assert m.get(91, 19) == 4.56
    m.set(91, 19, -4.56); assert m.get(91, 19) == -4.56


def test_out_of_bounds_dropped():
    m = O.EllMatrix(256)
    m.add(256, 256, 1)
    assert m.get(256, 256) == 0
    m.set(256, 256, 1)
    assert m.get(256, 256) == 0
    assert m.nnz() == 0


def test_set_same_row_and_zero_skip():
    m = O.EllMatrix(256)
    m.set(15, 3, 0)
    assert m.get(15, 3) == 0 and m.nnz() == 0                          # zero input dropped
    for c, v in [(6, 3), (7, 5), (4, 4)]:
        m.set(15, c, v)
        assert m.get(15, c) == v
    val, col, rn = m.export()
    assert list(col[: rn[15], 15]) == [4, 6, 7]                        # sorted insertion


def test_add_semantics():
    m = O.EllMatrix(256)
    m.set(15, 3, 1); m.add(15, 3, 2); assert m.get(15, 3) == 3
    m.add(15, 3, 8); m.add(15, 3, 32); assert m.get(15, 3) == 43
    m.set(9, 0, -5); m.add(9, 0, 10); assert m.get(9, 0) == 5
    m.set(30, 30, 6); m.add(30, 30, -10); assert m.get(30, 30) == -4


def test_full_rows_grow_the_matrix():
    n = 256
    m = O.EllMatrix(n)
    for i in range(n):
        m.add(0, i, 1)
        m.add(10, i, 2)
    for i in range(n - 1, -1, -1):
        m.add(0, i, 1)
        m.add(10, i, 2)
    assert sum(m.get(0, i) for i in range(n)) == 2 * n
    assert sum(m.get(10, i) for i in range(n)) == 4 * n
    assert m.maxrowlen > n                                          # grew by 64-slot steps
    assert m.maxrowlen % 64 == 0


def test_small_full_matrix_vs_dense():
    n = 64
    m = O.EllMatrix(n)
    A = np.zeros((n, n))
    rng = np.random.default_rng(0)
    for _ in range(3000):
        i, j = rng.integers(0, n, 2)
        v = float(rng.integers(-3, 4))
        m.add(int(i), int(j), v)
        A[i, j] += v
    assert np.array_equal(m.dense(), A)
    x = rng.normal(size=n)
    assert np.allclose(m.axy(x), A @ x, rtol=1e-14, atol=1e-12)


def test_spmv_kats():
    for n in GOLD["spmv_diagonal"]["sizes"]:
        m = O.EllMatrix(n)
        for i in range(n):
            m.add(i, i, i + 1.0)
        y = m.axy(np.arange(1, n + 1, dtype=float))
        assert np.allclose(y, np.arange(1, n + 1) ** 2, atol=1e-3)
    m = O.EllMatrix(64)
    for i in range(64):
        for j in range(8):
            m.add(i, j, 1)
    assert np.allclose(m.axy(np.full(64, 2.0)), 16.0)
    m = O.EllMatrix(64)
    for i in range(64):
        for j in range(4):
            m.add(i, i + j, 1)
    y = m.axy(np.full(64, 2.0))
    expect = np.full(64, 8.0)
    for k in range(1, 4):
        expect[64 - 4 + k] -= 2 * k
    assert np.allclose(y, expect)


# ---- CG (src/SPRmatrix/SPRmatrix.cc:405-480) -------------------------------------------
def diag_corner_matrix(n):
    m = O.EllMatrix(n)
    for i in range(n):
        m.add(i, i, float((i + 1) ** 2))
    m.set(0, 0, 2)
    m.set(0, 1, 2)
    m.set(1, 0, 2)
    return m


@pytest.mark.parametrize("n", GOLD["cg_diag_corner"]["sizes"])
def test_cg_kat_diag_corner(n):
    m = diag_corner_matrix(n)
    b = np.arange(1, n + 1, dtype=float) ** 2
    x, it, _ = O.cg(m, b, 3000, 1e-5)
    assert abs(x[0] + 1) < 1e-3 and abs(x[1] - 1.5) < 1e-3
    assert np.all(np.abs(x[2:] - 1) < 1e-3)
    xp, itp, rel, _ = O.pcg_jacobi(m, b, rtol=1e-10)
    assert abs(xp[0] + 1) < 1e-8 and abs(xp[1] - 1.5) < 1e-8 and rel <= 1e-10


def test_cg_kat_tridiag():
    m = O.EllMatrix(4)
    for i in range(4):
        m.add(i, i, 2)
        m.add(i, i + 1, -1)                       # (3, 4) is out of bounds -> dropped
        m.add(i, i - 1, -1)                       # (0, -1) is out of bounds -> dropped
    x, it, _ = O.cg(m, np.ones(4), 200, 1e-5)
    assert np.allclose(x, GOLD["cg_tridiag_4"]["x"], atol=1e-3)


def test_cg_kat_ones_plus_identity():
    n = 32
    m = O.EllMatrix(n)
    for i in range(n):
        for j in range(n):
            m.set(i, j, 1)
        m.set(i, i, 2)
    x, it, _ = O.cg(m, np.ones(n), 500, 1e-5)
    assert np.allclose(x, GOLD["cg_ones_plus_identity_32"]["x_each"], atol=1e-3)
    assert np.allclose(x, 1.0 / 33.0, rtol=1e-9)


def test_dot_kat_and_determinism():
    n = 100003
    a = np.full(n, 2.0)
    b = np.arange(1, n + 1, dtype=float)
    assert O.dot(a, b) == float(2 * n * (n + 1) // 2)
    rng = np.random.default_rng(1)
    a, b = rng.normal(size=n), rng.normal(size=n)
    assert O.dot(a, b, 1) == O.dot(a, b, 8)       # fixed-order blocked sum


@pytest.mark.parametrize("n,seed", [(50, 0), (200, 1)])
def test_cg_and_pcg_vs_direct(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.normal(size=(n, n))
    A = M @ M.T + 10 * np.eye(n)          # SPD as in src/Test/helpers.h:44-75
    m = O.EllMatrix(n)
    for i in range(n):
        for j in range(n):
            m.add(i, j, A[i, j])
    b = rng.normal(size=n)
    xd = scipy.linalg.solve(A, b, assume_a="pos")
    x, it, _ = O.cg(m, b, 10000, 1e-12)
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    xp, itp, rel, _ = O.pcg_jacobi(m, b, rtol=1e-12)
    assert rel <= 1e-12 * 1.01
    assert np.linalg.norm(xp - xd) <= 1e-9 * np.linalg.norm(xd)
    # warm start from the answer converges immediately
    xw, itw, relw, _ = O.pcg_jacobi(m, b, rtol=1e-8, x0=xd)
    assert itw == 0


def test_pcg_zero_rhs_and_maxit():
    m = diag_corner_matrix(16)
    x, it, rel, _ = O.pcg_jacobi(m, np.zeros(16))
    assert it == 0 and np.all(x == 0)
    b = np.arange(1, 17, dtype=float)
    x, it, rel, _ = O.pcg_jacobi(m, b, rtol=1e-14, maxit=1)
    assert it == 1
