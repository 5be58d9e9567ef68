"""Loader for the in-tree HIP engine (gradscf_amd/lib/libgradscf_jk.so).

Mirrors the reference's private native-library loader
(src/gradscf/integrals/_native/__init__.py:10-56): load only this package's own
library, with RTLD_LOCAL, and fail loudly when it is missing or stale.  There
is no CPU fallback: every compute entry point needs the library and a GPU.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libgradscf_jk.so"

EXPORTS = ("gscf_last_error", "gscf_plan_create", "gscf_plan_create_batch", "gscf_plan_get_molecules",
           "gscf_plan_destroy", "gscf_plan_get_info", "gscf_plan_get_schwarz", "gscf_plan_profile", "gscf_plan_class_times",
           "gscf_jk_build", "gscf_direct_jk", "gscf_eri_blocks", "gscf_int1e", "gscf_jk_geometry_vjp",
           "gscf_workspace_bytes", "gscf_int1e_gradient", "gscf_tuning_set", "gscf_tuning_get",
           "gscf_tuning_reset", "gscf_jk_param_vjp", "gscf_int1e_param_vjp", "gscf_options_set",
           "gscf_options_get", "gscf_fma_probe", "gscf_plan_incore", "gscf_plan_incore_bytes",
           "gscf_packed_eri", "gscf_packed_jk", "gscf_jk_build_noalloc", "gscf_density_fitting",
           "gscf_compact")

GSCF_OK, GSCF_INVALID_ARGUMENT, GSCF_INTERNAL = 0, 1, 2
FLAG_SYMMETRIC_DENSITY = 1
FLAG_MIXED_PRECISION = 2
FLAG_NO_VALIDATE = 4


class PlanInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("n_cart", ctypes.c_int32), ("nbas", ctypes.c_int32),
                ("nseg", ctypes.c_int32), ("npairs", ctypes.c_int64), ("nquartets", ctypes.c_int64),
                ("cutoff", ctypes.c_double), ("plan_ms", ctypes.c_double), ("nmol", ctypes.c_int32),
                ("nc_max", ctypes.c_int32), ("nn_total", ctypes.c_int64), ("device_bytes", ctypes.c_int64),
                ("lds_max", ctypes.c_int64), ("classes_global_atomic", ctypes.c_int32)]


class ClassTuning(ctypes.Structure):
    _fields_ = [("chunk_items", ctypes.c_int32), ("bras_max", ctypes.c_int32), ("lds_pack", ctypes.c_int32),
                ("rys_lds", ctypes.c_int32), ("target_prim", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("ket_lds", ctypes.c_int32)]


class Options(ctypes.Structure):
    _fields_ = [("general_contraction", ctypes.c_int32), ("streams", ctypes.c_int32),
                ("partial_sums", ctypes.c_int32), ("lds_limit", ctypes.c_int32),
                ("deterministic", ctypes.c_int32), ("balance", ctypes.c_int32)]


class NativeError(RuntimeError):
    """Internal engine failure (the reference's ffi::Error::Internal)."""


@lru_cache(maxsize=1)
def lib() -> ctypes.CDLL:
    if not LIB_PATH.is_file():
        raise RuntimeError(
            "The MI355X J/K engine is not built. Run `python -m gradscf_amd.build` "
            "(requires hipcc with gfx950 support).")
    handle = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_LOCAL)
    for name in EXPORTS:
        if not hasattr(handle, name):
            raise RuntimeError("The J/K engine library is stale; rebuild with python -m gradscf_amd.build")
    vp, i32, i64, dbl, u32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_uint32
    handle.gscf_last_error.restype = ctypes.c_char_p
    handle.gscf_plan_create.argtypes = [vp, i32, vp, i32, vp, i32, i32, dbl, vp, ctypes.POINTER(vp)]
    handle.gscf_plan_create_batch.argtypes = [vp, i32, vp, i32, vp, i32, i32, vp, vp, i32, dbl, vp,
                                              ctypes.POINTER(vp)]
    handle.gscf_plan_get_molecules.argtypes = [vp, i32, vp, vp]
    handle.gscf_plan_destroy.argtypes = [vp]
    handle.gscf_plan_get_info.argtypes = [vp, ctypes.POINTER(PlanInfo)]
    handle.gscf_plan_get_schwarz.argtypes = [vp, vp]
    handle.gscf_plan_profile.argtypes = [vp, i32]
    handle.gscf_plan_class_times.argtypes = [vp, i32, vp, vp, vp, vp]
    handle.gscf_jk_build.argtypes = [vp, vp, i32, vp, i32, i32, u32, dbl, vp]
    handle.gscf_jk_build_noalloc.argtypes = [vp, vp, i32, vp, i32, i32, u32, dbl, vp]
    handle.gscf_direct_jk.argtypes = [vp, i32, vp, i32, vp, i32, vp, i32, i32, vp, i32, dbl, i32, i32, i32, vp]
    handle.gscf_eri_blocks.argtypes = [vp, vp, i64, vp, vp, vp]
    handle.gscf_int1e.argtypes = [vp, vp, vp, vp, vp]
    handle.gscf_jk_geometry_vjp.argtypes = [vp, vp, vp, vp, vp, vp]
    handle.gscf_int1e_gradient.argtypes = [vp, vp, vp, vp, vp]
    handle.gscf_tuning_set.argtypes = [i32, vp]
    handle.gscf_tuning_get.argtypes = [i32, vp]
    handle.gscf_tuning_reset.argtypes = []
    handle.gscf_options_set.argtypes = [vp]
    handle.gscf_options_get.argtypes = [vp]
    handle.gscf_fma_probe.argtypes = [i32, i64, vp, vp]
    handle.gscf_plan_incore.argtypes = [vp, i32, vp]
    handle.gscf_plan_incore_bytes.argtypes = [vp, vp]
    handle.gscf_packed_eri.argtypes = [vp, i32, vp, i32, vp, i32, i32, i32, vp, i64, vp]
    handle.gscf_packed_jk.argtypes = [vp, i64, i32, vp, i32, i32, vp, i32, i32, vp]
    handle.gscf_density_fitting.argtypes = [vp, i32, vp, i32, vp, i32, i32, i32, i32, vp, i64, vp]
    handle.gscf_compact.argtypes = [vp, i32, vp, i32, vp, i32, i32, i32, i32, vp, i64, vp]
    handle.gscf_jk_param_vjp.argtypes = [vp, vp, vp, vp, vp, vp]
    handle.gscf_int1e_param_vjp.argtypes = [vp, vp, vp, vp, vp]
    handle.gscf_workspace_bytes.argtypes = [vp, i32, vp, i32, vp, i32, i32, dbl, i32, vp]
    return handle


def check(rc: int) -> None:
    if rc == GSCF_OK:
        return
    msg = lib().gscf_last_error().decode()
    if rc == GSCF_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise NativeError(msg)
