"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle the HIP product path is tested against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package; the product (gpufem_amd) never does.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcEll(C.Structure):
    _fields_ = [("n", C.c_int64), ("maxrowlen", C.c_int32), ("growstep", C.c_int32),
                ("skip_zero", C.c_int32), ("pad_", C.c_int32), ("val", C.POINTER(C.c_double)),
                ("col", C.POINTER(C.c_int32)), ("rownnz", C.POINTER(C.c_int32))]


_P, _I32, _I64, _D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
_EP = C.POINTER(OrcEll)
SIGS = {
    "orc_gauss_points": (C.c_int, [C.c_int, C.c_int, _P, _P]),
    "orc_constitutive": (None, [C.c_int, _D, _D, _P]),
    "orc_dshape": (C.c_int, [C.c_int, C.c_int, _P, _P]),
    "orc_node_natural_coords": (C.c_int, [C.c_int, C.c_int, _P]),
    "orc_jacobian": (None, [C.c_int, C.c_int, _P, _P, _P]),
    "orc_det": (_D, [C.c_int, _P]),
    "orc_inverse": (None, [C.c_int, _P, _D, _P]),
    "orc_bmatrix": (None, [C.c_int, C.c_int, _P, _P]),
    "orc_element_stiffness": (C.c_int, [C.c_int, C.c_int, C.c_int, _P, _D, _D, _P]),
    "orc_ell_create": (_EP, [_I64, C.c_int]),
    "orc_ell_free": (None, [_EP]),
    "orc_ell_binsearch_step": (_I64, [_P, _I32, _I32, _I64, _I64]),
    "orc_ell_set": (None, [_EP, _I64, _I64, _D]),
    "orc_ell_add": (None, [_EP, _I64, _I64, _D]),
    "orc_ell_get": (_D, [_EP, _I64, _I64]),
    "orc_ell_reserve": (None, [_EP, _I32]),
    "orc_ell_max_rownnz": (_I32, [_EP]),
    "orc_ell_nnz": (_I64, [_EP]),
    "orc_ell_zero_values": (None, [_EP]),
    "orc_ell_axy": (None, [_EP, _P, _P, C.c_int]),
    "orc_ell_export": (C.c_int, [_EP, _I32, _P, _P, _P]),
    "orc_greedy_coloring": (C.c_int, [_I64, C.c_int, _P, _I64, _P]),
    "orc_assemble": (C.c_int, [_EP, C.c_int, C.c_int, C.c_int, _I64, _P, _P, _P, _D, _D, _P,
                               C.c_int, C.c_int, _P]),
    "orc_apply_penalty": (_D, [_EP, C.c_int, _I64, _P, _D]),
    "orc_force_vector": (None, [C.c_int, _I64, _I64, _P, _P]),
    "orc_apply_elimination": (None, [_EP, C.c_int, _I64, _P]),
    "orc_dot": (_D, [_I64, _P, _P, C.c_int]),
    "orc_cg": (C.c_int, [_EP, _P, _P, C.c_int, _D, C.c_int, _P]),
    "orc_pcg_jacobi": (C.c_int, [_EP, _P, _P, C.c_int, _D, C.c_int, C.c_int, _P, _P, C.c_int]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        for k, (r, a) in SIGS.items():
            f = getattr(lib, k)
            f.restype, f.argtypes = r, a
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_threads():
    return max(1, min(os.cpu_count() or 1, 16))


# ---- element level -----------------------------------------------------------
def gauss_points(dim, ngauss):
    pts = np.zeros((27, 3))
    w = np.zeros(27)
    n = load().orc_gauss_points(dim, ngauss, _p(pts), _p(w))
    return pts.reshape(-1)[: n * dim].reshape(n, dim).copy(), w[:n].copy()


def constitutive(dim, E, nu):
    s = 3 if dim == 2 else 6
    D = np.zeros((s, s))
    load().orc_constitutive(dim, E, nu, _p(D))
    return D


def dshape(dim, nen, rst):
    dN = np.zeros((dim, nen))
    rst = np.ascontiguousarray(rst, dtype=np.float64)
    if not load().orc_dshape(dim, nen, _p(rst), _p(dN)):
        raise ValueError("unsupported element")
    return dN


def node_natural_coords(dim, nen):
    xi = np.zeros((nen, dim))
    if not load().orc_node_natural_coords(dim, nen, _p(xi)):
        raise ValueError("unsupported element")
    return xi


def jacobian(dim, nen, xe, dN):
    J = np.zeros((dim, dim))
    load().orc_jacobian(dim, nen, _p(np.ascontiguousarray(xe)), _p(np.ascontiguousarray(dN)),
                        _p(J))
    return J


def det(J):
    J = np.ascontiguousarray(J, dtype=np.float64)
    return load().orc_det(J.shape[0], _p(J))


def inverse(J):
    J = np.ascontiguousarray(J, dtype=np.float64)
    R = np.zeros_like(J)
    load().orc_inverse(J.shape[0], _p(J), load().orc_det(J.shape[0], _p(J)), _p(R))
    return R


def bmatrix(dim, nen, dNdx):
    s = 3 if dim == 2 else 6
    B = np.zeros((s, dim * nen))
    load().orc_bmatrix(dim, nen, _p(np.ascontiguousarray(dNdx)), _p(B))
    return B


def element_stiffness(dim, nen, ngauss, xe, E, nu):
    """xe is [dim][nen] (coordinate-major)."""
    ne = dim * nen
    Ke = np.zeros((ne, ne))
    st = load().orc_element_stiffness(dim, nen, ngauss,
                                      _p(np.ascontiguousarray(xe, dtype=np.float64)), E, nu,
                                      _p(Ke))
    if st != 1:
        raise ValueError("element stiffness failed (status %d)" % st)
    return Ke


# ---- ELL -------------------------------------------------------------------------
class EllMatrix:
    """The reference's ELLmatrix semantics in FP64 (oracle)."""

    def __init__(self, n, skip_zero=True):
        self.lib = load()
        self.m = self.lib.orc_ell_create(n, 1 if skip_zero else 0)
        if not self.m:
            raise MemoryError("orc_ell_create")
        self.n = n

    def __del__(self):
        try:
            if self.m:
                self.lib.orc_ell_free(self.m)
                self.m = None
        except Exception:
            pass

    def set(self, r, c, v):
        self.lib.orc_ell_set(self.m, r, c, v)

    def add(self, r, c, v):
        self.lib.orc_ell_add(self.m, r, c, v)

    def get(self, r, c):
        return self.lib.orc_ell_get(self.m, r, c)

    @property
    def maxrowlen(self):
        return self.m.contents.maxrowlen

    def max_rownnz(self):
        return self.lib.orc_ell_max_rownnz(self.m)

    def nnz(self):
        return self.lib.orc_ell_nnz(self.m)

    def export(self, width=None):
        if width is None:
            width = self.max_rownnz()
        val = np.zeros((width, self.n))
        col = np.zeros((width, self.n), dtype=np.int32)
        rn = np.zeros(self.n, dtype=np.int32)
        if not self.lib.orc_ell_export(self.m, width, _p(val), _p(col), _p(rn)):
            raise ValueError("width smaller than the longest row")
        return val, col, rn

    def axy(self, x, nthreads=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        self.lib.orc_ell_axy(self.m, _p(x), _p(y), nthreads or default_threads())
        return y

    def dense(self):
        val, col, rn = self.export()
        A = np.zeros((self.n, self.n))
        for i in range(self.n):
            for s in range(rn[i]):
                A[i, col[s, i]] += val[s, i]
        return A


def binsearch_step(vec, val, steppedlen, startpos, step):
    v = np.ascontiguousarray(vec, dtype=np.int32)
    return load().orc_ell_binsearch_step(_p(v), val, steppedlen, startpos, step)


def greedy_coloring(conn, nnodes):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    out = np.zeros(conn.shape[0], dtype=np.int32)
    nc = load().orc_greedy_coloring(conn.shape[0], conn.shape[1], _p(conn), nnodes, _p(out))
    if nc < 0:
        raise ValueError("coloring failed (%d)" % nc)
    return out, nc


def assemble(dim, ngauss, conn, coords, E, nu, elem_E=None, skip_zero=False, nthreads=None,
             colors=None, into=None):
    """Colored FP64 assembly; returns (EllMatrix, seconds).  With `into`, the
    contributions are added to that matrix (numeric re-assembly when its
    values were zeroed), else a fresh matrix is made."""
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    nnodes = coords.shape[0]
    if colors is None:
        colors, nc = greedy_coloring(conn, nnodes)
    else:
        nc = int(colors.max())
    K = into if into is not None else EllMatrix(nnodes * dim, skip_zero=skip_zero)
    secs = C.c_double()
    eE = None if elem_E is None else np.ascontiguousarray(elem_E, dtype=np.float64)
    st = load().orc_assemble(K.m, dim, conn.shape[1], ngauss, conn.shape[0], _p(conn),
                             _p(coords), _p(eE), E, nu, _p(colors), nc,
                             nthreads or default_threads(), C.byref(secs))
    if st != 1:
        raise ValueError("assembly failed (status %d)" % st)
    return K, secs.value


def apply_penalty(K, dim, supports, coef=1e4):
    sup = np.ascontiguousarray(supports, dtype=np.int32)
    return load().orc_apply_penalty(K.m, dim, sup.shape[0], _p(sup), coef)


def apply_elimination(K, dim, supports):
    sup = np.ascontiguousarray(supports, dtype=np.int32).reshape(-1, dim + 1)
    load().orc_apply_elimination(K.m, dim, sup.shape[0], _p(sup))


def force_vector(dim, nnodes, loads):
    loads = np.ascontiguousarray(loads, dtype=np.float64)
    f = np.zeros(nnodes * dim)
    load().orc_force_vector(dim, nnodes, loads.shape[0], _p(loads), _p(f))
    return f


def dot(a, b, nthreads=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return load().orc_dot(a.size, _p(a), _p(b), nthreads or default_threads())


def cg(K, b, maxit, eps, nthreads=None):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    secs = C.c_double()
    it = load().orc_cg(K.m, _p(b), _p(x), maxit, eps, nthreads or default_threads(),
                       C.byref(secs))
    return x, it, secs.value


def pcg_jacobi(K, b, rtol=1e-8, maxit=100000, x0=None, nthreads=None, progress_every=0):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    rel = C.c_double()
    secs = C.c_double()
    it = load().orc_pcg_jacobi(K.m, _p(b), _p(x), maxit, rtol, 0 if x0 is None else 1,
                               nthreads or default_threads(), C.byref(rel), C.byref(secs),
                               int(progress_every))
    return x, it, rel.value, secs.value
