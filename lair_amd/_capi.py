"""ctypes bindings of include/lair_dev.h (the C ABI of liblair_dev.so).

The library is loaded from lair_amd/lib/liblair_dev.so.  There is no fallback:
a missing or unloadable library raises ImportError.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liblair_dev.so")

i32 = ctypes.c_int32
i64 = ctypes.c_int64
f64 = ctypes.c_double
vp = ctypes.c_void_p
P64 = ctypes.POINTER(ctypes.c_int64)
PD = ctypes.POINTER(ctypes.c_double)

LAIR_OK = 0
LAIR_ERR_SHAPE = 1
LAIR_ERR_SINGULAR = 2
LAIR_ERR_DEVICE = 3
LAIR_ERR_NOMEM = 4
LAIR_ERR_NOCONV = 5
LAIR_ERR_SYNC = 6
LAIR_ERR_UNSUPPORTED = 7

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16X6 = 2
PREC_BF16 = 3
PIVOT_TOURNAMENT = 0x100
LU_P, LU_L, LU_U, LU_PL = 0, 1, 2, 3

KCLASS = {"panel": 0, "laswp": 1, "trsm": 2, "gemm": 3, "getrs": 4, "residual": 5, "convert": 6,
          "gemm_panel": 7}


class IrInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("fell_back", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("backward_error", ctypes.c_double),
        ("history", ctypes.c_double * 32),
    ]


# (name, restype, argtypes) for every symbol declared in include/lair_dev.h.
SIGNATURES = [
    ("lair_context_create", i32, [ctypes.POINTER(vp)]),
    ("lair_context_destroy", i32, [vp]),
    ("lair_context_synchronize", i32, [vp]),
    ("lair_context_set_wait_limit", i32, [vp, i64]),
    ("lair_context_stream", vp, [vp]),
    ("lair_last_error", ctypes.c_char_p, []),
    ("lair_malloc", i32, [ctypes.POINTER(vp), ctypes.c_uint64]),
    ("lair_free", i32, [vp]),
    ("lair_memcpy_h2d", i32, [vp, vp, ctypes.c_uint64]),
    ("lair_memcpy_d2h", i32, [vp, vp, ctypes.c_uint64]),
    ("lair_memcpy_d2d", i32, [vp, vp, ctypes.c_uint64]),
    ("lair_dgetrf", i32, [i64, i64, vp, i64, i64, P64, P64]),
    ("lair_sgetrf", i32, [i64, i64, vp, i64, i64, P64, P64]),
    ("lair_dgetrs", i32, [i64, i64, vp, i64, i64, P64, vp, i64, i64, i32]),
    ("lair_sgetrs", i32, [i64, i64, vp, i64, i64, P64, vp, i64, i64, i32]),
    ("lair_dlaswp", i32, [i64, vp, i64, i64, i64, i64, P64]),
    ("lair_dsolve", i32, [i64, i64, vp, i64, i64, i64, vp, vp]),
    ("lair_zgetrf", i32, [i64, i64, vp, i64, i64, P64, P64]),
    ("lair_cgetrf", i32, [i64, i64, vp, i64, i64, P64, P64]),
    ("lair_zgetrs", i32, [i64, i64, vp, i64, i64, P64, vp, i64, i64]),
    ("lair_cgetrs", i32, [i64, i64, vp, i64, i64, P64, vp, i64, i64]),
    ("lair_zsolve", i32, [i64, i64, vp, i64, i64, i64, vp, vp]),
    ("lair_csolve", i32, [i64, i64, vp, i64, i64, i64, vp, vp]),
    ("lair_zgetrf_device", i32, [vp, i64, i64, vp, i64, vp, i64, P64]),
    ("lair_cgetrf_device", i32, [vp, i64, i64, vp, i64, vp, i64, P64]),
    ("lair_zgetrs_device", i32, [vp, i64, i64, vp, i64, vp, vp, i64]),
    ("lair_cgetrs_device", i32, [vp, i64, i64, vp, i64, vp, vp, i64]),
    ("lair_dgetrf_device", i32, [vp, i64, i64, vp, i64, vp, i64, P64]),
    ("lair_sgetrf_device", i32, [vp, i64, i64, vp, i64, vp, i64, P64]),
    ("lair_sgetrf_pivot_device", i32, [vp, i64, i64, vp, i64, vp, i64, i32, i32, P64]),
    ("lair_dgetrs_device", i32, [vp, i64, i64, vp, i64, vp, vp, i64, i32]),
    ("lair_dgetrs_device_noalloc", i32, [vp, i64, i64, vp, i64, vp, vp, i64, i32]),
    ("lair_dlu_factor", i32, [vp, i64, i64, vp, i64, i64, ctypes.POINTER(vp)]),
    ("lair_dlu_destroy", i32, [vp]),
    ("lair_dlu_info", i32, [vp, P64, P64, P64]),
    ("lair_dlu_factors", i32, [vp, ctypes.POINTER(vp), P64, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
    ("lair_dlu_solve", i32, [vp, vp, i64, vp, i64, i32]),
    ("lair_dlu_solve_backward", i32, [vp, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64]),
    ("lair_dlu_extract", i32, [vp, vp, i32, vp, i64]),
    ("lair_sgetrs_device", i32, [vp, i64, i64, vp, i64, vp, vp, i64, i32]),
    ("lair_dgemm_minus_device", i32, [vp, i64, i64, i64, vp, i64, vp, i64, vp, i64]),
    ("lair_sgemm_minus_device", i32, [vp, i32, i64, i64, i64, vp, i64, vp, i64, vp, i64]),
    ("lair_dgesv_mixed_device", i32,
     [vp, i64, i64, vp, i64, vp, i64, vp, i64, i32, i32, f64, ctypes.POINTER(IrInfo)]),
    ("lair_dgesv_device", i32, [vp, i64, i64, vp, i64, vp, i64, vp, i64]),
    ("lair_dbackward_error", i32, [vp, i64, i64, vp, i64, vp, i64, vp, i64, PD]),
    ("lair_dsolve_backward_device", i32,
     [vp, i64, i64, vp, i64, vp, vp, i64, vp, i64, vp, i64, vp, i64]),
    ("lair_dsolve_backward", i32, [i64, i64, vp, P64, vp, vp, vp, vp]),
    ("lair_gesv_batched_plan", i32,
     [vp, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), vp, vp, vp, vp, i32, ctypes.POINTER(vp)]),
    ("lair_gesv_batched_run", i32, [vp]),
    ("lair_gesv_batched_config", i32, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]),
    ("lair_gesv_batched_destroy", i32, [vp]),
    ("lair_profile_enable", i32, [vp, i32]),
    ("lair_profile_read", i32, [vp, i32, PD, P64, PD, PD]),
    ("lair_profile_reset", i32, [vp]),
]

_lib = None


def load():
    """Loads liblair_dev.so and declares every entry point; raises ImportError
    when the library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with python -m lair_amd.build")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error():
    return load().lair_last_error().decode(errors="replace")


class LairError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        super().__init__(f"{where}: status {status}: {last_error()}")


def check(status, where=""):
    if status != LAIR_OK:
        raise LairError(status, where)
    return status
