"""ctypes binding of libmimadv.so (the C ABI declared in include/mimadv.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C fdd3260_amd``).  There is no fallback: if the library is missing
or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmimadv.so")

# status codes
OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_UNSUPPORTED = -3
ERR_NOMEM = -4

BC_PER, BC_NNF, BC_OPE = 0, 1, 2
WENO, PPM, MUSCL, WENO5, QUICK, LW = 0, 1, 2, 3, 4, 5
LIM_MINMOD, LIM_MC, LIM_VL, LIM_FROMM = 0, 1, 2, 3
WENO_BOTTOM_X0_IS_X1, WENO_BOTTOM_CLOSURE = 0, 1
INT_EULER, INT_HEUN, INT_SSPRK3 = 1, 2, 3

DIR_X, DIR_Y, DIR_Z, DIR_ALL = 1, 2, 4, 7
OVERWRITE = 8
HOST_PTRS = 16
UNLIMITED = 128
POSITIVE = 2048
SORDER1 = 256
DREF_DEN0 = 512
GRAPH = 32
MOMENTUM = 64
PROFILE = 1024
F64, F32 = 0, 1
OWN_STREAM = C.c_void_p(-1 & ((1 << 64) - 1)).value

EXPORTED = (
    "mimadv_create",
    "mimadv_destroy",
    "mimadv_last_error",
    "mimadv_set_stream",
    "mimadv_synchronize",
    "mimadv_set_opts",
    "mimadv_scalar_tendency",
    "mimadv_momentum_tendency",
    "mimadv_fill_halo",
    "mimadv_fill_winds",
    "mimadv_integ_mom",
    "mimadv_integ_scal",
    "mimadv_totalav",
    "mimadv_advance",
    "mimadv_last_kernel_ms",
    "mimadv_set_tuning",
    "mimadv_batch_tendency",
    "mimadv_batch_advance",
    "mimadv_batch_budget",
    "mimadv_batch_kernel_ms",
    "mimadv_batch_budget_device",
    "mimadv_adjoint_tendency",
    "mimadv_adjoint_advance",
    "mimadv_advance_record",
    "mimadv_adjoint_advance_recorded",
    "mimadv_set_reference_density",
    "mimadv_get_momentum",
    "mimadv_get_advective",
    "mimadv_get_conserved",
    "mimadv_advection_1d",
    "mimadv_multi_create",
    "mimadv_multi_destroy",
    "mimadv_multi_last_error",
    "mimadv_multi_set_tiles",
    "mimadv_multi_get_tiles",
    "mimadv_multi_autotune",
    "mimadv_multi_advance",
    "mimadv_multi_last_ms",
)


class Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("dx", C.c_double),
        ("dy", C.c_double),
        ("dz", C.c_double),
        ("fdz0", C.POINTER(C.c_double)),
        ("dzw", C.POINTER(C.c_double)),
        ("dz1w", C.POINTER(C.c_double)),
        ("dz1", C.POINTER(C.c_double)),
        ("bcl", C.c_int * 6),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("scheme", C.c_int),
        ("muscl_limiter", C.c_int),
        ("ppm_limiter", C.c_int),
        ("lim_tol", C.c_double),
        ("weno_bottom", C.c_int),
        ("mom_ord", C.c_int),
        ("lavisc", C.c_int),
    ]


class Tuning(C.Structure):
    _fields_ = [("group32", C.c_int), ("kchunk", C.c_int), ("blocks_per_cu", C.c_int), ("tile_y", C.c_int),
                ("tile_x", C.c_int)]


_lib = None


class MimadvError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libmimadv.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MimadvError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(the HIP library has no CPU fallback)"
        )
    lib = C.CDLL(LIB_PATH)
    vp, i, d, u = C.c_void_p, C.c_int, C.c_double, C.c_uint
    pp = C.POINTER(C.c_void_p)
    ip = C.POINTER(C.c_int)
    lib.mimadv_create.argtypes = [C.POINTER(Grid), C.POINTER(Opts), C.POINTER(C.c_void_p)]
    lib.mimadv_create.restype = i
    lib.mimadv_destroy.argtypes = [vp]
    lib.mimadv_destroy.restype = None
    lib.mimadv_last_error.argtypes = [vp]
    lib.mimadv_last_error.restype = C.c_char_p
    lib.mimadv_set_stream.argtypes = [vp, vp]
    lib.mimadv_set_stream.restype = i
    lib.mimadv_synchronize.argtypes = [vp]
    lib.mimadv_synchronize.restype = i
    lib.mimadv_set_opts.argtypes = [vp, C.POINTER(Opts)]
    lib.mimadv_set_opts.restype = i
    lib.mimadv_scalar_tendency.argtypes = [vp, d, vp, vp, vp, i, pp, pp, ip, u]
    lib.mimadv_scalar_tendency.restype = i
    lib.mimadv_momentum_tendency.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, u]
    lib.mimadv_momentum_tendency.restype = i
    lib.mimadv_fill_halo.argtypes = [vp, i, pp, u]
    lib.mimadv_fill_halo.restype = i
    lib.mimadv_fill_winds.argtypes = [vp, vp, vp, vp, u]
    lib.mimadv_fill_winds.restype = i
    lib.mimadv_integ_mom.argtypes = [vp, i, i, d, C.POINTER(d), d, d, i, pp, pp, pp, pp, u]
    lib.mimadv_integ_mom.restype = i
    lib.mimadv_integ_scal.argtypes = [vp, i, i, d, i, pp, pp, pp, pp, u]
    lib.mimadv_integ_scal.restype = i
    lib.mimadv_totalav.argtypes = [vp, vp, C.POINTER(d), C.POINTER(d), u]
    lib.mimadv_totalav.restype = i
    lib.mimadv_advance.argtypes = [vp, i, i, d, vp, vp, vp, i, pp, ip, u]
    lib.mimadv_advance.restype = i
    lib.mimadv_last_kernel_ms.argtypes = [vp, C.POINTER(d)]
    lib.mimadv_last_kernel_ms.restype = i
    lib.mimadv_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.mimadv_set_tuning.restype = i
    lib.mimadv_batch_tendency.argtypes = [vp, vp, vp, vp, i, ip, pp, pp, u]
    lib.mimadv_batch_tendency.restype = i
    lib.mimadv_batch_advance.argtypes = [vp, i, i, d, vp, vp, vp, i, ip, pp, u]
    lib.mimadv_batch_advance.restype = i
    lib.mimadv_batch_budget.argtypes = [vp, i, ip, pp, C.POINTER(d)]
    lib.mimadv_batch_budget_device.argtypes = [vp, i, ip, pp, vp]
    lib.mimadv_batch_budget_device.restype = i
    lib.mimadv_batch_kernel_ms.argtypes = [vp, C.POINTER(d), ip]
    lib.mimadv_batch_kernel_ms.restype = i
    lib.mimadv_adjoint_tendency.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.mimadv_adjoint_tendency.restype = i
    lib.mimadv_adjoint_advance.argtypes = [vp, i, i, d, vp, vp, vp, vp, vp, vp, vp, vp, i]
    lib.mimadv_adjoint_advance.restype = i
    lib.mimadv_advance_record.argtypes = [vp, i, i, d, vp, vp, vp, vp, i, pp]
    lib.mimadv_advance_record.restype = i
    lib.mimadv_adjoint_advance_recorded.argtypes = [vp, i, i, d, vp, vp, vp, i, pp, vp, vp, vp, vp]
    lib.mimadv_adjoint_advance_recorded.restype = i
    lib.mimadv_batch_budget.restype = i
    dp = C.POINTER(C.c_double)
    lib.mimadv_set_reference_density.argtypes = [vp, dp, dp]
    lib.mimadv_set_reference_density.restype = i
    lib.mimadv_get_momentum.argtypes = [vp, vp, vp, vp, i, u]
    lib.mimadv_get_momentum.restype = i
    lib.mimadv_get_advective.argtypes = [vp, d, d, vp, vp, vp, vp, vp, vp, u]
    lib.mimadv_get_advective.restype = i
    lib.mimadv_get_conserved.argtypes = [vp, vp, i, pp, i, u]
    lib.mimadv_get_conserved.restype = i
    lib.mimadv_advection_1d.argtypes = [vp, C.POINTER(C.c_double), vp, i, pp, pp, u]
    lib.mimadv_advection_1d.restype = i
    lib.mimadv_multi_create.argtypes = [i, pp, pp, pp, pp, ip, C.POINTER(pp), C.POINTER(vp)]
    lib.mimadv_multi_create.restype = i
    lib.mimadv_multi_destroy.argtypes = [vp]
    lib.mimadv_multi_destroy.restype = None
    lib.mimadv_multi_last_error.argtypes = [vp]
    lib.mimadv_multi_last_error.restype = C.c_char_p
    lib.mimadv_multi_set_tiles.argtypes = [vp, ip, ip]
    lib.mimadv_multi_set_tiles.restype = i
    lib.mimadv_multi_get_tiles.argtypes = [vp, ip, ip, ip]
    lib.mimadv_multi_get_tiles.restype = i
    lib.mimadv_multi_autotune.argtypes = [vp, i]
    lib.mimadv_multi_autotune.restype = i
    lib.mimadv_multi_advance.argtypes = [vp, i, i, C.POINTER(d)]
    lib.mimadv_multi_advance.restype = i
    lib.mimadv_multi_last_ms.argtypes = [vp, C.POINTER(d)]
    lib.mimadv_multi_last_ms.restype = i
    _lib = lib
    return lib


def ptr_array(ptrs):
    arr = (C.c_void_p * max(1, len(ptrs)))()
    for n, p in enumerate(ptrs):
        arr[n] = p
    return arr
