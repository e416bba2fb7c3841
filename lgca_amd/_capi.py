"""ctypes binding of include/lgca.h (liblgca_hip.so).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "liblgca_hip.so")

OK, ERR_INVALID_ARGUMENT, ERR_HIP, ERR_OUT_OF_MEMORY, ERR_INVALID_STATE, ERR_BUDGET = range(6)
HPP, FHP_I, FHP_II, FHP_III = range(4)
FLUID, SOLID_NO_SLIP, SOLID_SLIP = range(3)
FORCE_NONE, FORCE_X, FORCE_Y = range(3)

# Every entry point of include/lgca.h, with its ctypes signature.
_u8p = C.POINTER(C.c_uint8)
_f32p = C.POINTER(C.c_float)
_h = C.c_void_p


class Info(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("num_dir", C.c_uint32), ("dim_x", C.c_uint32),
        ("dim_y", C.c_uint32), ("coarse_radius", C.c_uint32), ("coarse_dim_x", C.c_uint32),
        ("coarse_dim_y", C.c_uint32), ("words_per_row", C.c_uint32), ("row_pitch", C.c_uint32),
        ("tile_rows", C.c_uint32), ("num_cells", C.c_uint64), ("num_coarse_cells", C.c_uint64),
        ("state_bytes", C.c_uint64), ("device_bytes", C.c_uint64), ("step", C.c_uint64),
        ("seed", C.c_uint64), ("steps_per_pass", C.c_int32), ("pass_pipeline", C.c_uint32),
    ]


GEOM_PERIODIC, GEOM_PIPE, GEOM_CYLINDER, GEOM_TYPES = range(4)
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_RELEASED, JOB_FAILED = range(5)
ALL_JOBS = 2 ** 64 - 1


class JobDesc(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("dim_x", C.c_uint32), ("dim_y", C.c_uint32),
        ("coarse_radius", C.c_uint32), ("seed", C.c_uint64), ("steps", C.c_uint64),
        ("fill", C.c_double), ("geometry", C.c_int32), ("cyl_x", C.c_int32),
        ("cyl_y", C.c_int32), ("cyl_diameter", C.c_double), ("force_dir", C.c_int32),
        ("force_p", C.c_double), ("fx0", C.c_uint32), ("fx1", C.c_uint32),
        ("fy0", C.c_uint32), ("fy1", C.c_uint32), ("cell_types", C.c_void_p),
        ("state", C.c_void_p),
    ]


class JobInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("reserved", C.c_uint32), ("step", C.c_uint64),
        ("device_bytes", C.c_uint64), ("submit_s", C.c_double), ("admit_s", C.c_double),
        ("done_s", C.c_double),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("jobs_submitted", C.c_uint64), ("jobs_done", C.c_uint64), ("jobs_active", C.c_uint64),
        ("jobs_queued", C.c_uint64), ("rounds", C.c_uint64), ("launches", C.c_uint64),
        ("site_updates", C.c_uint64), ("bytes_in_use", C.c_uint64), ("arena_bytes", C.c_uint64),
        ("steps_per_pass", C.c_int32), ("wave_target", C.c_uint32), ("tunings", C.c_uint64),
    ]


SIGNATURES = {
    "lgca_create": [C.POINTER(_h), C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int],
    "lgca_destroy": [_h],
    "lgca_get_info": [_h, C.POINTER(Info)],
    "lgca_set_stream": [_h, C.c_void_p],
    "lgca_synchronize": [_h],
    "lgca_set_seed": [_h, C.c_uint64],
    "lgca_set_step": [_h, C.c_uint64],
    "lgca_set_cell_types": [_h, C.c_void_p],
    "lgca_get_cell_types": [_h, C.c_void_p],
    "lgca_set_state_u8": [_h, C.c_void_p],
    "lgca_get_state_u8": [_h, C.c_void_p],
    "lgca_set_state_u8_device": [_h, C.c_void_p],
    "lgca_get_state_u8_device": [_h, C.c_void_p],
    "lgca_init_random": [_h, C.c_double],
    "lgca_init_random_masked": [_h, C.c_double, C.c_void_p],
    "lgca_set_forcing": [_h, C.c_int, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32],
    "lgca_step": [_h, C.c_uint64],
    "lgca_step_noalloc": [_h, C.c_uint64],
    "lgca_set_tuning": [_h, C.c_int, C.c_uint32],
    "lgca_tune": [_h],
    "lgca_set_pipeline": [_h, C.c_int],
    "lgca_apply_body_force_count": [_h, C.c_int, C.c_int64, C.POINTER(C.c_int64)],
    "lgca_snapshot": [_h],
    "lgca_post_process": [_h, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "lgca_post_process_device": [_h, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p)],
    "lgca_count_particles": [_h, C.POINTER(C.c_uint64)],
    "lgca_mean_velocity": [_h, _f32p],
    "lgca_controller_set": [_h, C.c_int, C.c_float, C.c_int64, C.c_int64],
    "lgca_controller_tick": [_h, C.c_uint64],
    "lgca_controller_get": [_h, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _f32p],
    "lgca_save_checkpoint": [_h, C.c_char_p],
    "lgca_load_checkpoint": [C.POINTER(_h), C.c_char_p, C.c_int],
    "lgca_batch_create": [C.POINTER(_h), C.c_int, C.c_uint64],
    "lgca_batch_destroy": [_h],
    "lgca_batch_job_bytes": [C.POINTER(JobDesc), C.POINTER(C.c_uint64)],
    "lgca_batch_submit": [_h, C.POINTER(JobDesc), C.POINTER(C.c_uint64)],
    "lgca_batch_wait": [_h, C.c_uint64, C.c_double],
    "lgca_batch_poll": [_h, C.POINTER(C.c_uint64), C.c_uint32, C.POINTER(C.c_uint32)],
    "lgca_batch_job_info": [_h, C.c_uint64, C.POINTER(JobInfo)],
    "lgca_batch_result": [_h, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p],
    "lgca_batch_release": [_h, C.c_uint64],
    "lgca_batch_set_tuning": [_h, C.c_int, C.c_uint32, C.c_uint64],
    "lgca_batch_get_stats": [_h, C.POINTER(BatchStats)],
}
# entry points returning text rather than a status
TEXT_ENTRY_POINTS = ("lgca_last_error", "lgca_batch_last_error")

_lib = None


class LgcaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lgca error {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """The loaded engine library; raises when it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m lgca_amd.build` "
                               "(the engine has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = C.c_int
        for name in TEXT_ENTRY_POINTS:
            fn = getattr(L, name)
            fn.argtypes = [_h]
            fn.restype = C.c_char_p
        _lib = L
    return _lib


def check(code: int, handle=None) -> None:
    if code != OK:
        msg = lib().lgca_last_error(handle)
        raise LgcaError(code, msg.decode() if msg else "")


def check_batch(code: int, batch=None) -> None:
    if code != OK:
        msg = lib().lgca_batch_last_error(batch)
        raise LgcaError(code, msg.decode() if msg else "")
