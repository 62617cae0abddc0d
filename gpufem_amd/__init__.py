"""MI355X (gfx950) engine for gpufem's assembly-and-solve path.

Thin Python layer over the C-ABI of libgpufem.so (include/gpufem.h): element
stiffness by Gauss quadrature fused with a colored, atomic-free scatter into
slot-major ELL storage, the penalty constraint, ELL SpMV and Jacobi-PCG, all
in hand-written HIP kernels.  Nothing here computes on the CPU: every numeric
entry point goes through the library and raises if it is not built.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (GF_BRICK8, GF_BRICK20, GF_ERR_ARG, GF_ERR_BREAKDOWN,  # noqa: F401
                   GF_ERR_DEGENERATE, GF_ERR_NOCONV, GF_ERR_UNSUPPORTED, GF_OK,
                   GF_PREC_FP64, GF_PREC_MIXED, GF_PRECOND_JACOBI, GF_PRECOND_NONE,
                   GF_Q4, GF_Q8, GF_SCHED_AUTO, GF_SCHED_GRAPH, GF_SCHED_KERNELS,
                   GF_SCHED_PERSISTENT, GpufemError, check, load)

ELEM_NAMES = {GF_Q4: "Q4", GF_Q8: "Q8", GF_BRICK8: "Brick8", GF_BRICK20: "Brick20"}
ELEM_DIM = {GF_Q4: 2, GF_Q8: 2, GF_BRICK8: 3, GF_BRICK20: 3}
ELEM_NEN = {GF_Q4: 4, GF_Q8: 8, GF_BRICK8: 8, GF_BRICK20: 20}


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# structured meshes (host only)
# ---------------------------------------------------------------------------
class Grid:
    """Structured grid of one element type (lexicographic node numbering)."""

    def __init__(self, elem_type, nx, ny, nz=0, h=1.0):
        self.desc = _lib.GridDesc(elem_type, nx, ny, nz, h)
        lib = load()
        nn, ne, nen, dim = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        check(lib.gf_grid_sizes(C.byref(self.desc), C.byref(nn), C.byref(ne), C.byref(nen),
                                C.byref(dim)))
        self.elem_type = elem_type
        self.nnodes, self.nelem = nn.value, ne.value
        self.nen, self.dim = nen.value, dim.value
        self.ndof = self.nnodes * self.dim

    def fill(self):
        conn = np.empty((self.nelem, self.nen), dtype=np.int32)
        coords = np.empty((self.nnodes, self.dim), dtype=np.float64)
        check(load().gf_grid_fill(C.byref(self.desc), _ptr(conn), _ptr(coords)))
        return conn, coords

    def face_nodes(self, face):
        lib = load()
        cnt = lib.gf_grid_face_nodes(C.byref(self.desc), face, None)
        if cnt < 0:
            check(int(cnt))
        out = np.empty(cnt, dtype=np.int32)
        lib.gf_grid_face_nodes(C.byref(self.desc), face, _ptr(out))
        return out

    def face_load(self, face, comp, q, f=None):
        if f is None:
            f = np.zeros(self.ndof, dtype=np.float64)
        check(load().gf_grid_face_load(C.byref(self.desc), face, comp, q, _ptr(f)))
        return f


def rng_normal(seed, n):
    out = np.empty(n, dtype=np.float64)
    check(load().gf_rng_normal(C.c_uint64(seed), n, _ptr(out)))
    return out


def rng_uniform_int(seed, n, lo, hi):
    out = np.empty(n, dtype=np.int32)
    check(load().gf_rng_uniform_int(C.c_uint64(seed), n, lo, hi, _ptr(out)))
    return out


# ---------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------
class Mesh:
    def __init__(self, ctx, handle, elem_type, ngauss, nnodes, nelem, arrays):
        self.ctx, self.handle = ctx, handle
        self.elem_type, self.ngauss = elem_type, ngauss
        self.nnodes, self.nelem = nnodes, nelem
        self.dim = ELEM_DIM[elem_type]
        self._arrays = arrays

    def free(self):
        if self.handle:
            load().gf_mesh_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Ell:
    def __init__(self, ctx, handle):
        self.ctx, self.handle = ctx, handle
        info = _lib.EllInfo()
        check(load().gf_ell_info_get(handle, C.byref(info)))
        self.n, self.width, self.ncolors = info.n, info.width, info.ncolors
        self.nnz, self.bytes = info.nnz, info.bytes

    def free(self):
        if self.handle:
            load().gf_ell_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def solve_opts(**kw):
    o = _lib.SolveOpts()
    load().gf_solve_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError("unknown solver option %r" % k)
        setattr(o, k, v)
    return o


class Context:
    """One GPU, one HIP stream (gf_ctx)."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.gf_create(C.byref(h), device))
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.gf_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, code):
        return check(code, self.handle)

    def device_info(self):
        cu = C.c_int32()
        name = C.create_string_buffer(128)
        self._check(self.lib.gf_device_info(self.handle, C.byref(cu), name, 128))
        return cu.value, name.value.decode()

    # -- mesh / matrix ------------------------------------------------------
    def upload_mesh(self, elem_type, ngauss, conn, coords, E=210000.0, nu=0.3, elem_E=None):
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        if elem_E is not None:
            elem_E = np.ascontiguousarray(elem_E, dtype=np.float64)
        d = _lib.MeshDesc(elem_type, ngauss, coords.shape[0], conn.shape[0], _ptr(conn),
                          _ptr(coords), _ptr(elem_E), E, nu)
        h = C.c_void_p()
        self._check(self.lib.gf_mesh_upload(self.handle, C.byref(d), C.byref(h)))
        return Mesh(self, h, elem_type, ngauss, coords.shape[0], conn.shape[0],
                    (conn, coords, elem_E))

    def symbolic(self, mesh):
        h = C.c_void_p()
        t = _lib.Timing()
        self._check(self.lib.gf_ell_symbolic(self.handle, mesh.handle, C.byref(h), C.byref(t)))
        m = Ell(self, h)
        m.symbolic_timing = t
        return m

    def colors(self, mesh, ell):
        out = np.empty(mesh.nelem, dtype=np.int32)
        self._check(self.lib.gf_ell_colors(self.handle, ell.handle, _ptr(out)))
        return out

    def assemble(self, mesh, ell):
        t = _lib.Timing()
        self._check(self.lib.gf_assemble(self.handle, mesh.handle, ell.handle, C.byref(t)))
        return t

    def apply_penalty(self, ell, dofs, coef=1e4):
        dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        pen = C.c_double()
        self._check(self.lib.gf_apply_penalty(self.handle, ell.handle, _ptr(dofs), dofs.size,
                                              coef, C.byref(pen)))
        return pen.value

    def upload_ell(self, val, col, rownnz):
        """Matrix from slot-major host arrays of shape (width, n)."""
        val = np.ascontiguousarray(val, dtype=np.float64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        rownnz = np.ascontiguousarray(rownnz, dtype=np.int32)
        width, n = val.shape
        h = C.c_void_p()
        self._check(self.lib.gf_ell_upload(self.handle, n, width, _ptr(val), _ptr(col),
                                           _ptr(rownnz), C.byref(h)))
        return Ell(self, h)

    def download(self, ell):
        """(val, col, rownnz) with val/col of shape (width, n): val[slot, row]."""
        val = np.empty((ell.width, ell.n), dtype=np.float64)
        col = np.empty((ell.width, ell.n), dtype=np.int32)
        rn = np.empty(ell.n, dtype=np.int32)
        self._check(self.lib.gf_ell_download(self.handle, ell.handle, _ptr(val), _ptr(col),
                                             _ptr(rn)))
        return val, col, rn

    def spmv(self, ell, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        nrhs = 1 if x.ndim == 1 else x.shape[0]
        y = np.empty_like(x)
        t = _lib.Timing()
        self._check(self.lib.gf_spmv(self.handle, ell.handle, _ptr(x), _ptr(y), nrhs,
                                     C.byref(t)))
        return y, t

    def tune(self, ell, force_variant=-1):
        """Pick (or force) the SpMV variant; returns (chosen, ms per variant)."""
        ms = np.zeros(5)
        chosen = C.c_int32()
        self._check(self.lib.gf_ell_tune(self.handle, ell.handle, force_variant, _ptr(ms),
                                         C.byref(chosen)))
        return chosen.value, ms

    def profile_spmv(self, ell, nrep=20):
        ms = C.c_double()
        self._check(self.lib.gf_profile_spmv(self.handle, ell.handle, nrep, C.byref(ms)))
        return ms.value

    def pcg(self, ell, b, x0=None, raise_on_fail=True, **opts):
        """Solve K x = b; b is (n,) or (nrhs, n).  Returns (x, SolveStats)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        nrhs = 1 if b.ndim == 1 else b.shape[0]
        if x0 is not None:
            x = np.array(x0, dtype=np.float64, copy=True).reshape(b.shape)
            opts.setdefault("use_x0", 1)
        else:
            x = np.zeros_like(b)
        o = solve_opts(**opts)
        st = _lib.SolveStats()
        code = self.lib.gf_pcg(self.handle, ell.handle, _ptr(b), _ptr(x), nrhs, C.byref(o),
                               C.byref(st))
        if raise_on_fail:
            self._check(code)
        st.code = code
        return x, st

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
