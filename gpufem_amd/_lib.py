"""ctypes binding of libgpufem.so (the C-ABI declared in include/gpufem.h).

The shared library is built in-tree by ``make -C gpufem_amd`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback for anything in this package.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpufem.so")

GF_OK = 0
GF_ERR_ARG = -1
GF_ERR_HIP = -2
GF_ERR_NOMEM = -3
GF_ERR_UNSUPPORTED = -4
GF_ERR_DEGENERATE = -5
GF_ERR_BREAKDOWN = -6
GF_ERR_NOCONV = -7

GF_Q4, GF_Q8, GF_BRICK8, GF_BRICK20 = 1, 2, 3, 4

GF_PRECOND_NONE, GF_PRECOND_JACOBI = 0, 1
GF_PREC_FP64, GF_PREC_MIXED = 0, 1
GF_SCHED_AUTO, GF_SCHED_KERNELS, GF_SCHED_GRAPH, GF_SCHED_PERSISTENT = 0, 1, 2, 3
GF_SPMV_VARIANTS = 11
GF_ASM_VARIANTS = 5  # entries gf_assemble_tune writes into ms_per_variant


class GridDesc(C.Structure):
    _fields_ = [("elem_type", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
                ("nz", C.c_int32), ("h", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("ms_device", C.c_double), ("ms_host", C.c_double), ("launches", C.c_int64)]


class MeshDesc(C.Structure):
    _fields_ = [("elem_type", C.c_int32), ("ngauss", C.c_int32), ("nnodes", C.c_int64),
                ("nelem", C.c_int64), ("conn", C.c_void_p), ("coords", C.c_void_p),
                ("elem_E", C.c_void_p), ("E", C.c_double), ("nu", C.c_double)]


class EllInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("width", C.c_int32), ("ncolors", C.c_int32),
                ("nnz", C.c_int64), ("bytes", C.c_int64)]


class SolveOpts(C.Structure):
    _fields_ = [("precond", C.c_int32), ("precision", C.c_int32), ("schedule", C.c_int32),
                ("use_x0", C.c_int32), ("maxit", C.c_int32), ("device_ptrs", C.c_int32),
                ("rtol", C.c_double), ("inner_rtol", C.c_double), ("batched", C.c_int32),
                ("verify", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32),
                ("converged", C.c_int32), ("status", C.c_int32), ("relres", C.c_double),
                ("true_relres", C.c_double), ("ms_device", C.c_double),
                ("ms_host", C.c_double), ("fallbacks", C.c_int32), ("x0_rejected", C.c_int32)]


_P = C.c_void_p
_I32, _I64, _D = C.c_int32, C.c_int64, C.c_double

# name -> (restype, argtypes); every symbol include/gpufem.h declares
SIGNATURES = {
    "gf_grid_sizes": (C.c_int, [C.POINTER(GridDesc), _P, _P, _P, _P]),
    "gf_grid_fill": (C.c_int, [C.POINTER(GridDesc), _P, _P]),
    "gf_grid_face_nodes": (C.c_int64, [C.POINTER(GridDesc), _I32, _P]),
    "gf_grid_face_load": (C.c_int, [C.POINTER(GridDesc), _I32, _I32, _D, _P]),
    "gf_rng_normal": (C.c_int, [C.c_uint64, _I64, _P]),
    "gf_rng_uniform_int": (C.c_int, [C.c_uint64, _I64, _I32, _I32, _P]),
    "gf_grf": (C.c_int, [C.c_uint64, _I64, _I64, _I64, _D, _P]),
    "gf_create": (C.c_int, [C.POINTER(_P), _I32]),
    "gf_destroy": (C.c_int, [_P]),
    "gf_last_error": (C.c_char_p, [_P]),
    "gf_device_info": (C.c_int, [_P, _P, C.c_char_p, _I32]),
    "gf_mesh_upload": (C.c_int, [_P, C.POINTER(MeshDesc), C.POINTER(_P)]),
    "gf_mesh_free": (C.c_int, [_P]),
    "gf_mesh_set_E": (C.c_int, [_P, _P, _P]),
    "gf_ell_symbolic": (C.c_int, [_P, _P, C.POINTER(_P), C.POINTER(Timing)]),
    "gf_ell_symbolic_device": (C.c_int, [_P, _P, C.POINTER(_P), C.POINTER(Timing)]),
    "gf_ell_info_get": (C.c_int, [_P, C.POINTER(EllInfo)]),
    "gf_ell_colors": (C.c_int, [_P, _P, _P]),
    "gf_ell_free": (C.c_int, [_P]),
    "gf_ell_upload": (C.c_int, [_P, _I64, _I32, _P, _P, _P, C.POINTER(_P)]),
    "gf_ell_download": (C.c_int, [_P, _P, _P, _P, _P]),
    "gf_assemble": (C.c_int, [_P, _P, _P, C.POINTER(Timing)]),
    "gf_assemble_tune": (C.c_int, [_P, _P, _P, _I32, _P, _P]),
    "gf_apply_penalty": (C.c_int, [_P, _P, _P, _I64, _D, _P]),
    "gf_apply_elimination": (C.c_int, [_P, _P, _P, _I64]),
    "gf_zero_dofs": (C.c_int, [_P, _P, _I64, _I32, _P, _I64]),
    "gf_load_vector": (C.c_int, [_P, _I32, _I64, _P, _I64, _P]),
    "gf_spmv": (C.c_int, [_P, _P, _P, _P, _I32, C.POINTER(Timing)]),
    "gf_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "gf_pcg": (C.c_int, [_P, _P, _P, _P, _I32, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "gf_ell_tune": (C.c_int, [_P, _P, _I32, _P, _P]),
    "gf_profile_spmv": (C.c_int, [_P, _P, _I32, C.POINTER(C.c_double)]),
    "gf_profile_pcg": (C.c_int, [_P, _P, _P, _I32, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    "gf_stream_copy": (C.c_int, [_P, _P, _P, _I64, _I32, C.POINTER(C.c_double)]),
    "gf_group_create": (C.c_int, [_P, _I32, C.POINTER(MeshDesc), C.POINTER(_P)]),
    "gf_group_info": (C.c_int, [_P, _P, _P, _P]),
    "gf_group_assemble": (C.c_int, [_P, _P, C.POINTER(Timing)]),
    "gf_group_set_E": (C.c_int, [_P, _P, _P]),
    "gf_group_apply_penalty": (C.c_int, [_P, _P, _P, _P, _D]),
    "gf_group_pcg": (C.c_int, [_P, _P, _P, _P, C.POINTER(SolveOpts), C.POINTER(SolveStats), _P]),
    "gf_group_free": (C.c_int, [_P]),
    "gf_nn_create": (C.c_int, [_P, _I32, _I32, _I32, _I32, C.c_uint64, C.POINTER(_P)]),
    "gf_nn_free": (C.c_int, [_P]),
    "gf_nn_nparams": (C.c_int64, [_P]),
    "gf_nn_layer_count": (C.c_int, [_P]),
    "gf_nn_layer_info": (C.c_int, [_P, _I32, _P]),
    "gf_nn_params_get": (C.c_int, [_P, _P, _P]),
    "gf_nn_params_set": (C.c_int, [_P, _P, _P]),
    "gf_nn_grads_get": (C.c_int, [_P, _P, _P]),
    "gf_nn_forward": (C.c_int, [_P, _P, _P, _I32, _P]),
    "gf_nn_backward": (C.c_int, [_P, _P, _P, _P, _I32, _P]),
    "gf_nn_adam": (C.c_int, [_P, _P, C.c_float, C.c_float, C.c_float, C.c_float]),
    "gf_nn_train_step": (C.c_int, [_P, _P, _P, _P, _I32, C.c_float, _P]),
    "gf_nn_to_dofs": (C.c_int, [_P, _P, _P, _P, _I32, _I32, _I32, _D, _D, _P]),
    "gf_nn_conv3d": (C.c_int, [_P, _P, _I32, _P, _I32, _P, _P, _I32, _I32, _I32, _I32,
                               C.c_float, _P, _P, _P, _I32]),
    "gf_nn_conv3d_wgrad": (C.c_int, [_P, _P, _I32, _P, _I32, _P, _P, _I32, _I32, _P, _P]),
    "gf_pcg_noalloc": (C.c_int, [_P, _P, _P, _P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "gf_profile_barrier": (C.c_int, [_P, _P, _I32, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "gf_dev_alloc": (C.c_int, [_P, _I64, C.POINTER(_P)]),
    "gf_dev_free": (C.c_int, [_P, _P]),
    "gf_copy_h2d": (C.c_int, [_P, _P, _P, _I64]),
    "gf_copy_d2h": (C.c_int, [_P, _P, _P, _I64]),
    "gf_copy_d2d": (C.c_int, [_P, _P, _P, _I64]),
}

_lib = None


def load():
    """Load libgpufem.so, binding every declared entry point; raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libgpufem.so is not built (%s); run `make -C gpufem_amd` or "
            "__graft_entry__.build()" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GpufemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gpufem error %d: %s" % (code, msg))
        self.code = code


def check(code, ctx=None):
    if code != GF_OK:
        msg = load().gf_last_error(ctx)
        raise GpufemError(code, msg.decode() if msg else "")
    return code
