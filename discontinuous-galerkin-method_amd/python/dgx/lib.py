"""ctypes binding of libdgx.so (include/dgx.h, include/dgx_setup.h).

The library is the product: it must be built (``make -C discontinuous-galerkin-method_amd``) before
use, and loading fails loudly when it is missing.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

PKG_DIR = pathlib.Path(__file__).resolve().parents[2]
LIB_PATH = PKG_DIR / "lib" / "libdgx.so"
OPERATOR_DIR = PKG_DIR / "input_files" / "precomputed_matrix"

DGX_FP64, DGX_FP32, DGX_FP32_COMPENSATED, DGX_FP64_REFERENCE = 0, 1, 2, 3
DGX_SRC_FORCE_Z, DGX_SRC_EXPLOSIVE = 0, 1
DGX_LAYOUT_AUTO, DGX_LAYOUT_ELEMENT, DGX_LAYOUT_SHOT_INNER = 0, 1, 2
DGX_OK, DGX_EARG, DGX_EMESH, DGX_EHIP, DGX_ENOMEM, DGX_EIO, DGX_ESTATE = 0, -1, -2, -3, -4, -5, -6
DGX_STEP_AUTO, DGX_STEP_HALF, DGX_STEP_FULL, DGX_STEP_BLOCKED = 0, 1, 2, 3
DGX_NOALLOC_SCRATCH_BYTES = 32 << 20  # staging array of dgx_get_fields_noalloc, per GPU

dptr = C.POINTER(C.c_double)
lptr = C.POINTER(C.c_long)


class Operators(C.Structure):
    _fields_ = [("V", C.c_double * 36), ("inverV", C.c_double * 36), ("Vr", C.c_double * 36),
                ("Vs", C.c_double * 36), ("S1", C.c_double * 18), ("S2", C.c_double * 18),
                ("S3", C.c_double * 18), ("FeketeNew", C.c_double * 12)]

    def array(self, name: str) -> np.ndarray:
        shape = {"V": (6, 6), "inverV": (6, 6), "Vr": (6, 6), "Vs": (6, 6), "S1": (6, 3), "S2": (6, 3),
                 "S3": (6, 3), "FeketeNew": (6, 2)}[name]
        return np.ctypeslib.as_array(getattr(self, name)).reshape(shape).copy()


class Source(C.Structure):
    _fields_ = [("elem", C.c_long), ("kind", C.c_int), ("shot", C.c_int), ("wavelet", C.c_long),
                ("fun", C.c_double * 6)]


class Receiver(C.Structure):
    _fields_ = [("elem", C.c_long), ("weight", C.c_double * 6)]


class ProblemC(C.Structure):
    _fields_ = [("tri", C.c_long), ("point_num", C.c_long), ("point", dptr), ("triangle", lptr),
                ("neighbor", lptr), ("lam", dptr), ("miu", dptr), ("den", dptr), ("absorb_fun", dptr),
                ("dt", C.c_double), ("precision", C.c_int), ("nshots", C.c_long), ("nsteps", C.c_long),
                ("nwavelets", C.c_long), ("wavelet", dptr), ("nsrc", C.c_long), ("src", C.POINTER(Source)),
                ("nrec", C.c_long), ("rec", C.POINTER(Receiver)), ("layout", C.c_int), ("shot_tile", C.c_int),
                ("step_kernel", C.c_int), ("wg_waves", C.c_int), ("block_steps", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("precision", C.c_int), ("hbm_budget", C.c_longlong), ("steps_per_block", C.c_int),
                ("group_shots", C.c_int)]


class RequestC(C.Structure):
    _fields_ = [("mesh", C.c_int), ("nshots", C.c_int), ("nsteps", C.c_long), ("src", C.POINTER(Source)),
                ("wavelet", dptr), ("nrec", C.c_long), ("rec", C.POINTER(Receiver)), ("arrival", C.c_double),
                ("lam", dptr), ("miu", dptr), ("den", dptr), ("dt", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_long), ("running", C.c_long), ("done", C.c_long), ("batches", C.c_long),
                ("bytes_in_use", C.c_longlong), ("dof_updates", C.c_longlong), ("busy_seconds", C.c_double),
                ("blocks", C.c_long), ("global_step", C.c_long), ("lane_updates", C.c_longlong),
                ("gpu_seconds", C.c_double), ("moved", C.c_longlong)]


class SpongeParams(C.Structure):
    _fields_ = [("bound_x", C.c_double), ("bound_z", C.c_double), ("model_w", C.c_double),
                ("model_h", C.c_double), ("beta", C.c_double), ("pi", C.c_double)]


_lib = None


class DgxError(RuntimeError):
    pass


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise DgxError(f"libdgx.so is not built ({LIB_PATH}); run make -C {PKG_DIR}")
    lib = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    sig = {
        "dgx_create": (C.c_int, [C.POINTER(ProblemC), C.POINTER(Operators), C.POINTER(vp)]),
        "dgx_destroy": (C.c_int, [vp]),
        "dgx_step": (C.c_int, [vp, C.c_long]),
        "dgx_absorb_boundary": (C.c_int, [vp]),
        "dgx_source_kernel": (C.c_int, [vp, C.c_double]),
        "dgx_lu_velocity": (C.c_int, [vp]),
        "dgx_lu_stress_elastic": (C.c_int, [vp]),
        "dgx_record": (C.c_int, [vp, C.c_long]),
        "dgx_get_fields": (C.c_int, [vp, C.c_long] + [dptr] * 5),
        "dgx_get_fields_noalloc": (C.c_int, [vp, C.c_long] + [dptr] * 5),
        "dgx_snapshot_begin": (C.c_int, [vp, C.c_long, C.c_int]),
        "dgx_snapshot_wait": (C.c_int, [vp] + [dptr] * 5),
        "dgx_set_fields": (C.c_int, [vp, C.c_long] + [dptr] * 5),
        "dgx_get_traces": (C.c_int, [vp, C.c_long, dptr, dptr]),
        "dgx_synchronize": (C.c_int, [vp]),
        "dgx_current_step": (C.c_long, [vp]),
        "dgx_kernel_times": (C.c_int, [vp, C.c_long, dptr, dptr]),
        "dgx_last_error": (C.c_char_p, []),
        "dgx_kernel_bench": (C.c_int, [vp, C.c_int, C.c_long, dptr]),
        "dgx_copy_probe": (C.c_int, [C.c_long, C.c_long, C.c_int, dptr]),
        "dgx_layout_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), dptr, C.c_int]),
        "dgx_tile_info": (C.c_int, [vp] + [C.POINTER(C.c_int)] * 4 + [dptr, C.c_int]),
        "dgx_compiled_operators": (C.c_int, [vp]),
        "dgx_block_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_long), C.POINTER(C.c_long),
                                     C.POINTER(C.c_long)]),
        "dgx_read_ascii": (C.c_long, [C.c_char_p, dptr, C.c_long, C.c_long]),
        "dgx_read_ascii_long": (C.c_long, [C.c_char_p, lptr, C.c_long, C.c_long]),
        "dgx_read_operators": (C.c_int, [C.c_char_p, C.POINTER(Operators)]),
        "dgx_mesh_count": (C.c_long, [C.c_char_p]),
        "dgx_read_mesh": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_long, C.c_long, dptr, lptr, lptr, lptr]),
        "dgx_save_data": (C.c_int, [C.c_char_p, dptr, C.c_long, C.c_long]),
        "dgx_save_record": (C.c_int, [C.c_char_p, dptr, C.c_long]),
        "dgx_save_binary": (C.c_int, [C.c_char_p, dptr, C.c_long]),
        "dgx_write_mesh_binary": (C.c_int, [C.c_char_p, C.c_long, C.c_long, dptr, lptr, lptr, lptr]),
        "dgx_mesh_binary_counts": (C.c_int, [C.c_char_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
        "dgx_read_mesh_binary": (C.c_int, [C.c_char_p, C.c_long, C.c_long, dptr, lptr, lptr, lptr]),
        "dgx_all_node": (None, [C.c_long, dptr, lptr, dptr, dptr]),
        "dgx_sponge": (C.c_long, [C.c_long, dptr, lptr, dptr, C.POINTER(SpongeParams), dptr]),
        "dgx_cfl_dt": (C.c_double, [C.c_long, dptr, lptr, dptr]),
        "dgx_ricker": (C.c_double, [C.c_long, C.c_double, C.c_double, C.c_double, C.c_double]),
        "dgx_ricker_delay": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "dgx_wavelet": (C.c_int, [C.c_long, C.c_double, C.c_double, C.c_double, C.c_double, dptr]),
        "dgx_locator_create": (C.c_int, [C.c_long, C.c_long, dptr, lptr, C.POINTER(vp)]),
        "dgx_locator_destroy": (None, [vp]),
        "dgx_locate_source": (C.c_int, [vp, C.POINTER(Operators), C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_long, C.POINTER(Source)]),
        "dgx_locate_receivers": (C.c_int, [vp, C.c_long, dptr, C.POINTER(Receiver)]),
        "dgx_locate_receivers_device": (C.c_int, [C.c_long, C.c_long, dptr, lptr, C.c_long, dptr,
                                                  C.POINTER(Receiver)]),
        "dgx_locate_sources_device": (C.c_int, [C.c_long, C.c_long, dptr, lptr, C.POINTER(Operators), C.c_long,
                                                dptr, C.POINTER(C.c_int), C.POINTER(C.c_int), lptr,
                                                C.POINTER(Source)]),
        "dgx_p2_basis": (None, [C.c_double, C.c_double, dptr]),
        "dgx_reference_operators": (C.c_int, [C.c_long, dptr, lptr, dptr, dptr, dptr, C.POINTER(Operators)] + [dptr] * 5),
        "dgx_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(Operators), C.POINTER(vp)]),
        "dgx_server_destroy": (C.c_int, [vp]),
        "dgx_server_add_mesh": (C.c_int, [vp, C.POINTER(ProblemC), C.POINTER(C.c_int)]),
        "dgx_server_submit": (C.c_int, [vp, C.POINTER(RequestC), C.POINTER(C.c_long)]),
        "dgx_server_advance": (C.c_int, [vp, C.c_int]),
        "dgx_server_status": (C.c_int, [vp, C.c_long, C.POINTER(C.c_int), dptr, dptr]),
        "dgx_server_result": (C.c_int, [vp, C.c_long, dptr, dptr]),
        "dgx_server_release": (C.c_int, [vp, C.c_long]),
        "dgx_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "dgx_server_now": (C.c_double, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    """Names of the C-ABI functions declared in include/*.h (the binding's table)."""
    load()
    return sorted(n for n in dir(_lib) if n.startswith("dgx_"))


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        msg = load().dgx_last_error().decode(errors="replace")
        raise DgxError(f"{what} failed ({rc}): {msg}")
    return rc


def dp(a: np.ndarray):
    return a.ctypes.data_as(dptr)


def lp(a: np.ndarray):
    return a.ctypes.data_as(lptr)


def read_operators(directory: os.PathLike | str = OPERATOR_DIR) -> Operators:
    ops = Operators()
    check(load().dgx_read_operators(str(directory).encode(), C.byref(ops)), "dgx_read_operators")
    return ops
