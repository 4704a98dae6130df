"""ctypes binding of libcommit_hip.so (declared in include/commit_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C commit_amd/csrc``).  There is no fallback: if the library is missing
or fails to load, every operator call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcommit_hip.so")
SYNTH_PATH = os.path.join(_HERE, "libcommit_synth.so")

FP64 = 0
FP32 = 1
OPT_XCD_REMAP = 1
OPT_MULTI_A_WAVES = 2
OPT_A_DENSE = 3
OPT_SLAB = 4
# voxels per wavefront of the LDS-slab kernels (A x, A^T y): part of their names
SLAB_VA, SLAB_VAT = 6, 6

# every symbol include/commit_hip.h declares, with its ctypes signature
_P = ctypes.c_void_p
_I = ctypes.c_int
_U = ctypes.c_uint32
_I64P = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = {
    "commit_hip_plan_create": (_I, [ctypes.POINTER(_P), _I, _I, _I, _I, _I, _I, _I,
                                    _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _U, _U, _U, _I, _I, _I]),
    "commit_hip_plan_destroy": (_I, [_P]),
    "commit_hip_plan_set_eval": (_I, [_P, _P]),
    "commit_hip_plan_shape": (_I, [_P, _I64P, _I64P]),
    "commit_hip_plan_device_bytes": (ctypes.c_int64, [_P]),
    "commit_hip_plan_set_precision": (_I, [_P, _I]),
    "commit_hip_plan_precision": (_I, [_P]),
    "commit_hip_plan_set_option": (_I, [_P, _I, _I]),
    "commit_hip_plan_get_option": (_I, [_P, _I]),
    "commit_hip_A": (_I, [_P, _P, _P, _P]),
    "commit_hip_At": (_I, [_P, _P, _P, _P]),
    "commit_hip_A_residual": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "commit_hip_residual_combine": (_I, [_P, ctypes.c_double, _P, ctypes.c_double, _P, _P, _P, _P]),
    "commit_hip_At_combined": (_I, [_P, ctypes.c_double, _P, ctypes.c_double, _P, _P, _P, _P, _P, _P]),
    "commit_hip_At_weighted": (_I, [_P, _P, _P, _P, _P]),
    "commit_hip_At_pass": (_I, [_P, _P, _P, _P, _I, _P]),
    "commit_hip_A_multi": (_I, [_P, _P, _P, _I, _I, _P]),
    "commit_hip_At_multi": (_I, [_P, _P, _P, _P, _I, _I, _P]),
    "commit_hip_prox": (_I, [_P, ctypes.c_int64, _P, _P, ctypes.c_double, ctypes.c_double, _P, _P]),
    "commit_hip_fista_sums": (_I, [ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, ctypes.c_int64, _P, _P, _P,
                                   _P, _P, _P]),
    "commit_hip_fista_sums_noalloc": (_I, [ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, ctypes.c_int64, _P, _P,
                                           _P, _P, _P]),
    "commit_hip_axpby": (_I, [ctypes.c_int64, ctypes.c_double, _P, ctypes.c_double, _P, _P, _P]),
    "commit_hip_prox_f32": (_I, [_P, ctypes.c_int64, _P, _P, ctypes.c_double, ctypes.c_double, _P, _P]),
    "commit_hip_fista_sums_f32": (_I, [ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P, ctypes.c_int64, _P, _P, _P,
                                       _P, _P, _P]),
    "commit_hip_axpby_f32": (_I, [ctypes.c_int64, ctypes.c_double, _P, ctypes.c_double, _P, _P, _P]),
    "commit_hip_fista_multi_work": (ctypes.c_int64, [_P, ctypes.c_int64]),
    "commit_hip_prox_multi": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "commit_hip_fista_multi_sums": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "commit_hip_multi_residual": (_I, [_P, _P, _P, _I, _P, _P, _P, _P, _P]),
    "commit_hip_A_multi_residual_work": (ctypes.c_int64, [_P, _P]),
    "commit_hip_A_multi_residual": (_I, [_P, _P, _P, _P, _I, _P, _P, _P, _P, _P]),
    "commit_hip_fista_multi_control": (_I, [_P, _P, ctypes.c_double, ctypes.c_double, _I, _P, _P, _P, _P]),
    "commit_hip_fista_multi_update": (_I, [_P, _P, _P, _P, _P, _P]),
    "commit_hip_fista_multi_qf": (_I, [_P, _P, _P, _P]),
    "commit_hip_A_multi_trial": (_I, [_P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "commit_hip_At_multi_momentum": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "commit_hip_fista_multi_roles": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "commit_hip_plan_set_lut": (_I, [_P, _P, _P, _P, _P]),
    "commit_hip_lut_grad": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "commit_hip_vox_create": (_I, [ctypes.POINTER(_P), _I]),
    "commit_hip_vox_destroy": (_I, [_P]),
    "commit_hip_trk2dictionary": (_I, [_P, _P, _P, _I, ctypes.c_int64, _I, _I, _I, _I] + [ctypes.c_float] * 9 +
                                  [_P, _I, ctypes.c_float, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P, _P, _P, _P,
                                   _P]),
    "commit_hip_vox_sizes": (_I, [_P, _I64P, _I64P, _I64P, _I64P, _I64P]),
    "commit_hip_vox_fetch": (_I, [_P] * 13),
    "commit_hip_last_error": (ctypes.c_char_p, []),
    "commit_hip_version": (ctypes.c_char_p, []),
}

_lib = None

REG_NONE, REG_LASSO, REG_GROUP_LASSO, REG_SPARSE_GROUP_LASSO = 0, 1, 2, 3
SUMS_BLOCKS = 256
NOALLOC_MAX_GROUPS = 65536       # COMMIT_HIP_NOALLOC_MAX_GROUPS


class ProxParams(ctypes.Structure):
    """commit_hip_prox_params (include/commit_hip.h)."""
    _fields_ = [("start", ctypes.c_int64 * 3), ("size", ctypes.c_int64 * 3), ("reg", ctypes.c_int * 3),
                ("nn", ctypes.c_int * 3), ("weighted", ctypes.c_int * 3), ("lam", ctypes.c_double * 3),
                ("lam_group", ctypes.c_double), ("w", ctypes.c_void_p), ("ngroups", ctypes.c_int64),
                ("group_ptr", ctypes.c_void_p), ("group_idx", ctypes.c_void_p), ("group_w", ctypes.c_void_p),
                ("groups_overlap", ctypes.c_int)]


def sums_work(ngroups):
    """Doubles of workspace commit_hip_fista_sums needs (COMMIT_HIP_SUMS_WORK)."""
    return 8 * SUMS_BLOCKS + int(ngroups)


class CommitHipError(RuntimeError):
    pass


def load():
    """Load libcommit_hip.so once; raise if it is absent (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CommitHipError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() or `make -C commit_amd/csrc`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc, what):
    if rc != 0:
        msg = load().commit_hip_last_error().decode(errors="replace")
        raise CommitHipError(f"{what} failed ({rc}): {msg}")


def ptr(a):
    """Host pointer of a numpy array (None for empty/None)."""
    if a is None or a.size == 0:
        return None
    return ctypes.c_void_p(a.ctypes.data)
