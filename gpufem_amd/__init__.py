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


def grf(seed, shape, lengthscale):
    """Unit-variance Gaussian random field on a cell grid `shape` = (nx, ny[, nz]);
    returns an array indexed [k, j, i] (x fastest in memory)."""
    nx, ny = shape[0], shape[1]
    nz = shape[2] if len(shape) > 2 else 1
    out = np.empty(nx * ny * nz, dtype=np.float64)
    check(load().gf_grf(C.c_uint64(seed), nx, ny, nz, float(lengthscale), _ptr(out)))
    return out.reshape((nz, ny, nx) if len(shape) > 2 else (ny, nx))


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

    def set_E(self, elem_E):
        """New per-element moduli (the next assembly uses them)."""
        eE = np.ascontiguousarray(elem_E, dtype=np.float64)
        self._arrays = self._arrays[:2] + (eE,)
        self.ctx._check(load().gf_mesh_set_E(self.ctx.handle, self.handle, _ptr(eE)))

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

    def symbolic(self, mesh, device=False):
        """ELL structure, coloring and slot map (host phase, or device=True)."""
        h = C.c_void_p()
        t = _lib.Timing()
        fn = self.lib.gf_ell_symbolic_device if device else self.lib.gf_ell_symbolic
        self._check(fn(self.handle, mesh.handle, C.byref(h), C.byref(t)))
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

    def assemble_tune(self, mesh, ell, force_variant=-1):
        """Pick (or force) the assembly kernel; returns (chosen, ms per variant)."""
        ms = np.full(_lib.GF_ASM_VARIANTS, -1.0)
        chosen = C.c_int32()
        self._check(self.lib.gf_assemble_tune(self.handle, mesh.handle, ell.handle, force_variant,
                                              _ptr(ms), C.byref(chosen)))
        return chosen.value, ms

    def apply_penalty(self, ell, dofs, coef=1e4):
        dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        pen = C.c_double()
        self._check(self.lib.gf_apply_penalty(self.handle, ell.handle, _ptr(dofs), dofs.size,
                                              coef, C.byref(pen)))
        return pen.value

    def apply_elimination(self, ell, dofs):
        """Zero substitution: rows and columns of the DOFs become unit vectors."""
        dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        self._check(self.lib.gf_apply_elimination(self.handle, ell.handle, _ptr(dofs), dofs.size))

    def zero_dofs(self, dev, n, dofs, nrhs=1):
        """Clears entries d of nrhs device vectors of length n (DevBuf or pointer)."""
        dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        ptr = dev.ptr if isinstance(dev, DevBuf) else dev
        self._check(self.lib.gf_zero_dofs(self.handle, ptr, n, nrhs, _ptr(dofs), dofs.size))

    def load_vector(self, dim, nnodes, loads, out=None):
        """Force vector from load rows [node, F...] on the device; returns the
        DevBuf (out, when given, is filled)."""
        loads = np.ascontiguousarray(loads, dtype=np.float64).reshape(-1, dim + 1)
        if out is None:
            out = DevBuf(self, max(nnodes * dim, 1) * 8)
        self._check(self.lib.gf_load_vector(self.handle, dim, nnodes, _ptr(loads), loads.shape[0],
                                            out.ptr))
        return out

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
        ms = np.zeros(_lib.GF_SPMV_VARIANTS)
        chosen = C.c_int32()
        self._check(self.lib.gf_ell_tune(self.handle, ell.handle, force_variant, _ptr(ms),
                                         C.byref(chosen)))
        return chosen.value, ms

    def profile_spmv(self, ell, nrep=20):
        ms = C.c_double()
        self._check(self.lib.gf_profile_spmv(self.handle, ell.handle, nrep, C.byref(ms)))
        return ms.value

    def profile_barrier(self, ell, nbar=3000):
        """Microseconds per grid barrier of the persistent schedule, and the
        persistent grid's workgroup count."""
        us = C.c_double()
        nb = C.c_int32()
        self._check(self.lib.gf_profile_barrier(self.handle, ell.handle, nbar, C.byref(us),
                                                C.byref(nb)))
        return us.value, nb.value

    def profile_pcg(self, ell, b, iters=200):
        """Average SpMV kernel time and iteration time (ms) over iters PCG
        iterations from zero for b (events around each SpMV)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        ms_spmv, ms_it = C.c_double(), C.c_double()
        self._check(self.lib.gf_profile_pcg(self.handle, ell.handle, _ptr(b), iters,
                                            C.byref(ms_spmv), C.byref(ms_it)))
        return ms_spmv.value, ms_it.value

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

    def group(self, problems):
        """Grouped batch of independent meshes (gpufem_amd.problems.Problem)."""
        return Group(self, problems)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Group:
    """gf_group: many meshes assembled and solved as one block-diagonal system."""

    def __init__(self, ctx, problems):
        self.ctx, self.lib = ctx, ctx.lib
        self.problems = list(problems)
        keep = []
        descs = (_lib.MeshDesc * len(self.problems))()
        for k, p in enumerate(self.problems):
            conn = np.ascontiguousarray(p.conn, dtype=np.int32)
            coords = np.ascontiguousarray(p.coords, dtype=np.float64)
            eE = None if p.elem_E is None else np.ascontiguousarray(p.elem_E, dtype=np.float64)
            keep += [conn, coords, eE]
            descs[k] = _lib.MeshDesc(p.elem_type, p.ngauss, coords.shape[0], conn.shape[0],
                                     _ptr(conn), _ptr(coords), _ptr(eE), p.E, p.nu)
        h = C.c_void_p()
        ctx._check(self.lib.gf_group_create(ctx.handle, len(self.problems), descs, C.byref(h)))
        self.handle = h
        n, nnz, nc = C.c_int64(), C.c_int64(), C.c_int32()
        check(self.lib.gf_group_info(h, C.byref(n), C.byref(nnz), C.byref(nc)))
        self.n, self.nnz, self.ncolors = n.value, nnz.value, nc.value
        self.offsets = np.cumsum([0] + [p.ndof for p in self.problems])

    def assemble(self):
        t = _lib.Timing()
        self.ctx._check(self.lib.gf_group_assemble(self.ctx.handle, self.handle, C.byref(t)))
        return t

    def set_E(self, elem_E_list):
        """New per-element moduli, one array per mesh (structure kept)."""
        eE = np.ascontiguousarray(np.concatenate(elem_E_list), dtype=np.float64)
        self.ctx._check(self.lib.gf_group_set_E(self.ctx.handle, self.handle, _ptr(eE)))

    def apply_penalty(self, coef=1e4):
        dofs = np.ascontiguousarray(np.concatenate([p.fixed_dofs for p in self.problems]),
                                    dtype=np.int32)
        nd = np.ascontiguousarray([p.fixed_dofs.size for p in self.problems], dtype=np.int64)
        self.ctx._check(self.lib.gf_group_apply_penalty(self.ctx.handle, self.handle, _ptr(dofs),
                                                        _ptr(nd), coef))

    def pcg(self, b=None, raise_on_fail=True, x0=None, **opts):
        """Solve every mesh; b and x0 (optional initial guess) concatenated."""
        if b is None:
            b = np.concatenate([p.f for p in self.problems])
        b = np.ascontiguousarray(b, dtype=np.float64)
        if x0 is not None:
            x = np.array(x0, dtype=np.float64, copy=True).reshape(b.shape)
            opts.setdefault("use_x0", 1)
        else:
            x = np.zeros_like(b)
        o = solve_opts(**opts)
        st = _lib.SolveStats()
        its = np.zeros(len(self.problems), dtype=np.int32)
        code = self.lib.gf_group_pcg(self.ctx.handle, self.handle, _ptr(b), _ptr(x), C.byref(o),
                                     C.byref(st), _ptr(its))
        if raise_on_fail:
            self.ctx._check(code)
        st.code = code
        return x, st, its

    def split(self, x):
        return [x[self.offsets[k]:self.offsets[k + 1]] for k in range(len(self.problems))]

    def free(self):
        if self.handle:
            self.lib.gf_group_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DevBuf:
    """Device allocation owned by a Context (gf_dev_alloc / gf_dev_free)."""

    def __init__(self, ctx, nbytes):
        self.ctx, self.lib = ctx, ctx.lib
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx._check(self.lib.gf_dev_alloc(ctx.handle, max(self.nbytes, 4), C.byref(p)))
        self.ptr = p

    @classmethod
    def from_array(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.nbytes)
        d.upload(a)
        return d

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        dst = C.c_void_p(self.ptr.value + offset)
        self.ctx._check(self.lib.gf_copy_h2d(self.ctx.handle, dst, _ptr(a), a.nbytes))

    def download(self, dtype, count=None, offset=0):
        dt = np.dtype(dtype)
        count = (self.nbytes - offset) // dt.itemsize if count is None else int(count)
        out = np.empty(count, dtype=dt)
        src = C.c_void_p(self.ptr.value + offset)
        self.ctx._check(self.lib.gf_copy_d2h(self.ctx.handle, _ptr(out), src, out.nbytes))
        return out

    def at(self, offset):
        """Device address `offset` bytes into the buffer."""
        return C.c_void_p(self.ptr.value + int(offset))

    def free(self):
        if self.ptr:
            self.lib.gf_dev_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Net:
    """gf_nn: the warm-start network on a node grid (nx, ny, nz)."""

    def __init__(self, ctx, nx, ny, nz, batch_max, seed=0x5EED0007):
        self.ctx, self.lib = ctx, ctx.lib
        self.nx, self.ny, self.nz, self.batch_max = nx, ny, nz, batch_max
        h = C.c_void_p()
        ctx._check(self.lib.gf_nn_create(ctx.handle, nx, ny, nz, batch_max, C.c_uint64(seed),
                                         C.byref(h)))
        self.handle = h
        self.nparams = self.lib.gf_nn_nparams(h)
        self.nodes = nx * ny * nz

    def layers(self):
        out = []
        info = np.zeros(8, dtype=np.int32)
        for i in range(self.lib.gf_nn_layer_count(self.handle)):
            check(self.lib.gf_nn_layer_info(self.handle, i, _ptr(info)))
            out.append(tuple(int(v) for v in info))
        return out

    def params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        self.ctx._check(self.lib.gf_nn_params_get(self.ctx.handle, self.handle, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        self.ctx._check(self.lib.gf_nn_params_set(self.ctx.handle, self.handle, _ptr(p)))

    def grads(self):
        g = np.empty(self.nparams, dtype=np.float32)
        self.ctx._check(self.lib.gf_nn_grads_get(self.ctx.handle, self.handle, _ptr(g)))
        return g

    def forward(self, in_dev, batch, out_dev=None):
        self.ctx._check(self.lib.gf_nn_forward(self.ctx.handle, self.handle, in_dev, batch,
                                               out_dev))

    def backward(self, in_dev, target_dev, batch):
        loss = C.c_float()
        self.ctx._check(self.lib.gf_nn_backward(self.ctx.handle, self.handle, in_dev, target_dev,
                                                batch, C.byref(loss)))
        return loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx._check(self.lib.gf_nn_adam(self.ctx.handle, self.handle, lr, beta1, beta2, eps))

    def train_step(self, in_dev, target_dev, batch, lr, want_loss=True):
        loss = C.c_float()
        self.ctx._check(self.lib.gf_nn_train_step(self.ctx.handle, self.handle, in_dev,
                                                  target_dev, batch, lr,
                                                  C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def to_dofs(self, pred_dev, in_dev, sample, mask_channel, scale, x_dev, base_channel=-1,
                base_weight=0.0):
        self.ctx._check(self.lib.gf_nn_to_dofs(self.ctx.handle, self.handle, pred_dev, in_dev,
                                               sample, mask_channel, base_channel,
                                               float(base_weight), float(scale), x_dev))

    def free(self):
        if self.handle:
            self.lib.gf_nn_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
