"""ctypes binding of libptq2.so (the C ABI declared in include/ptq2.h).

The library is built in-tree (ptatin-gene_amd/libptq2.so).  Loading fails
loudly when it is missing: there is no CPU fallback for any operator.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libptq2.so")

OK = 0
ERR_INVALID_ARG = 1
ERR_NOT_SETUP = 2
ERR_SIZE = 3
ERR_ALLOC = 4
ERR_DEVICE = 5
ERR_NONFINITE = 6
ERR_MESH = 7
ERR_VISCOSITY = 8
ERR_UNSUPPORTED = 9

DEVICE = 0x1
FP32 = 0x2
MATMULT = 0x4
ADD = 0x8
ONTHEFLY = 0x10
STOREDGEO = 0x20
MFMA = 0x40

# every entry point of include/ptq2.h with its ctypes signature
_c_int, _c_uint, _vp, _i64 = ctypes.c_int, ctypes.c_uint, ctypes.c_void_p, ctypes.c_int64
SIGNATURES = {
    "ptq2_create": (_c_int, [ctypes.POINTER(_vp), _c_int]),
    "ptq2_destroy": (_c_int, [ctypes.POINTER(_vp)]),
    "ptq2_strerror": (ctypes.c_char_p, [_c_int]),
    "ptq2_last_error": (ctypes.c_char_p, [_vp]),
    "ptq2_set_stream": (_c_int, [_vp, _vp]),
    "ptq2_get_stream": (_vp, [_vp]),
    "ptq2_sync": (_c_int, [_vp]),
    "ptq2_setup_mesh": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "ptq2_setup_mesh_box": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                     _c_int, _c_int, _c_int, _vp, _vp]),
    "ptq2_set_viscosity": (_c_int, [_vp, _vp, ctypes.c_uint64, _c_uint]),
    "ptq2_a11_apply": (_c_int, [_vp, _vp, _vp, _c_uint]),
    "ptq2_a11_apply_noalloc": (_c_int, [_vp, _vp, _vp, _c_uint]),
    "ptq2_a11_apply_batch": (_c_int, [_vp, _c_int, _vp, _vp, _c_uint]),
    "ptq2_a11_diag": (_c_int, [_vp, _vp, _c_uint]),
    "ptq2_a12_apply": (_c_int, [_vp, _vp, _vp, _c_uint]),
    "ptq2_body_force": (_c_int, [_vp, _vp, _vp, _c_uint]),
    "ptq2_a21_apply": (_c_int, [_vp, _vp, _vp, _c_uint]),
    "ptq2_stokes_apply": (_c_int, [_vp, _vp, _vp, _vp, _vp, _c_uint]),
    "ptq2_sizes": (_c_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "ptq2_stokes_residual": (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_uint]),
    "ptq2_project_points_q1": (_c_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _c_uint]),
}

_lib = None


class PTQ2Error(RuntimeError):
    """A nonzero PTQ2_ERR_* status (PetscErrorCode analogue)."""

    def __init__(self, code, where, detail=""):
        self.code = code
        msg = "%s failed: %s" % (where, strerror(code))
        if detail:
            msg += " (%s)" % detail
        super().__init__(msg)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libptq2.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def strerror(code):
    return lib().ptq2_strerror(code).decode()


def check(code, where, ctx=None):
    if code != OK:
        detail = ""
        if ctx:
            detail = lib().ptq2_last_error(ctx).decode()
        raise PTQ2Error(code, where, detail)
