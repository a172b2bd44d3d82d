"""ctypes binding of the C ABI in include/mhd_amd.h (libmhd_amd.so, built in-tree).

The library is the only compute path: if it is missing this module raises, it never
falls back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libmhd_amd.so")

# the symbols include/mhd_amd.h declares, in header order
EXPORTS = (
    "mhd_version", "mhd_error_string", "mhd_last_error", "mhd_create", "mhd_destroy", "mhd_get_dims",
    "mhd_device_bytes", "mhd_set_grid", "mhd_generate_grid", "mhd_get_grid", "mhd_upload_prim",
    "mhd_upload_cons", "mhd_download_prim", "mhd_download_cons", "mhd_download_flux",
    "mhd_snapshot_begin", "mhd_snapshot_end", "mhd_control_timestep", "mhd_set_boundary", "mhd_evaluate_ch", "mhd_numerical_flux",
    "mhd_update_conserv", "mhd_damp_psi", "mhd_update_primitive", "mhd_run", "mhd_fused_begin_step",
    "mhd_fused_step", "mhd_run_noalloc",
    "mhd_generate_problem_kh",
    "mhd_generate_problem_shear", "mhd_diagnostics", "mhd_status", "mhd_last_run_timing",
    "mhd_profile_kernels", "mhd_kernel_names",
    "mhd_batch_create", "mhd_batch_destroy", "mhd_batch_box_bytes", "mhd_batch_free_bytes", "mhd_batch_set_tiling",
    "mhd_batch_add_box", "mhd_batch_upload_state", "mhd_batch_remove_box", "mhd_batch_step",
    "mhd_batch_last_step_ms", "mhd_batch_box_state", "mhd_batch_poll", "mhd_batch_finish_box",
    "mhd_set_tiling", "mhd_get_tiling", "mhd_set_xy_waves_per_simd", "mhd_get_xy_waves_per_simd",
)


class MhdConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("bc", C.c_int * 6),
        ("xmin", C.c_double * 3), ("xmax", C.c_double * 3),
        ("gam", C.c_double), ("cfl", C.c_double), ("alphabp", C.c_double),
        ("energy_flux", C.c_int), ("grid_shift", C.c_int), ("precision", C.c_int), ("device", C.c_int),
        ("stages", C.c_int),
    ]


class MhdIcParams(C.Structure):
    _fields_ = [
        ("rho0", C.c_double), ("p0", C.c_double), ("v0", C.c_double), ("a", C.c_double),
        ("bx0", C.c_double), ("amp", C.c_double), ("sigma", C.c_double), ("seed", C.c_uint64),
    ]


class MhdError(RuntimeError):
    pass


_lib = None


def load(path: str | None = None):
    """Load libmhd_amd.so (or the build variant named by $MHD_AMD_LIB) and declare every
    signature.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("MHD_AMD_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise MhdError(f"{path} is not built; run `python -m mhdturbulence_amd.build` (HIP path is mandatory)")
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    ip = C.POINTER(C.c_int)
    sig = {
        "mhd_version": (C.c_int, []),
        "mhd_error_string": (C.c_char_p, [C.c_int]),
        "mhd_last_error": (C.c_char_p, []),
        "mhd_create": (C.c_int, [C.POINTER(MhdConfig), C.POINTER(vp)]),
        "mhd_destroy": (C.c_int, [vp]),
        "mhd_get_dims": (C.c_int, [vp, ip, ip, ip]),
        "mhd_device_bytes": (C.c_int, [vp, C.POINTER(C.c_uint64)]),
        "mhd_set_grid": (C.c_int, [vp, dp, dp, dp, dp, dp, dp]),
        "mhd_generate_grid": (C.c_int, [vp]),
        "mhd_get_grid": (C.c_int, [vp, dp, dp, dp, dp, dp, dp]),
        "mhd_upload_prim": (C.c_int, [vp, dp]),
        "mhd_upload_cons": (C.c_int, [vp, dp]),
        "mhd_download_prim": (C.c_int, [vp, dp]),
        "mhd_download_cons": (C.c_int, [vp, dp]),
        "mhd_download_flux": (C.c_int, [vp, C.c_int, dp]),
        "mhd_snapshot_begin": (C.c_int, [vp]),
        "mhd_snapshot_end": (C.c_int, [vp, C.POINTER(dp)]),
        "mhd_control_timestep": (C.c_int, [vp, dp]),
        "mhd_set_boundary": (C.c_int, [vp]),
        "mhd_evaluate_ch": (C.c_int, [vp, dp]),
        "mhd_numerical_flux": (C.c_int, [vp, C.c_int, C.c_double]),
        "mhd_update_conserv": (C.c_int, [vp, C.c_double]),
        "mhd_damp_psi": (C.c_int, [vp, C.c_double, C.c_double]),
        "mhd_update_primitive": (C.c_int, [vp]),
        "mhd_run": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, dp, dp, ip]),
        "mhd_fused_begin_step": (C.c_int, [vp, dp, dp, ip]),
        "mhd_fused_step": (C.c_int, [vp]),
        "mhd_run_noalloc": (C.c_int, [vp, C.c_int, C.c_double, dp, dp, ip]),
        "mhd_generate_problem_kh": (C.c_int, [vp, C.c_double, C.c_double]),
        "mhd_generate_problem_shear": (C.c_int, [vp, C.POINTER(MhdIcParams)]),
        "mhd_diagnostics": (C.c_int, [vp, dp]),
        "mhd_status": (C.c_int, [vp, ip]),
        "mhd_last_run_timing": (C.c_int, [vp, dp, ip]),
        "mhd_profile_kernels": (C.c_int, [vp, C.c_int, C.c_int, dp, C.c_int, ip]),
        "mhd_kernel_names": (C.c_char_p, [C.c_int]),
        "mhd_batch_create": (C.c_int, [C.POINTER(MhdConfig), C.c_int, C.c_uint64, C.POINTER(vp)]),
        "mhd_batch_destroy": (C.c_int, [vp]),
        "mhd_batch_box_bytes": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint64)]),
        "mhd_batch_free_bytes": (C.c_int, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "mhd_batch_set_tiling": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "mhd_batch_add_box": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(MhdIcParams), C.c_double, ip]),
        "mhd_batch_upload_state": (C.c_int, [vp, C.c_int, dp, dp]),
        "mhd_batch_remove_box": (C.c_int, [vp, C.c_int]),
        "mhd_batch_step": (C.c_int, [vp, C.c_int]),
        "mhd_batch_last_step_ms": (C.c_int, [vp, dp]),
        "mhd_batch_box_state": (C.c_int, [vp, C.c_int, dp, ip, ip, ip]),
        "mhd_batch_poll": (C.c_int, [vp, ip, C.c_int, ip]),
        "mhd_batch_finish_box": (C.c_int, [vp, C.c_int, dp, dp]),
        "mhd_set_tiling": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "mhd_get_tiling": (C.c_int, [vp, ip, ip, ip]),
        "mhd_set_xy_waves_per_simd": (C.c_int, [vp, C.c_int]),
        "mhd_get_xy_waves_per_simd": (C.c_int, [vp, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        L = load()
        raise MhdError(f"{L.mhd_error_string(rc).decode()} ({rc}): {L.mhd_last_error().decode()}")
