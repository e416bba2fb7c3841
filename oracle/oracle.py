"""ctypes front end of the CPU oracle (oracle/cpu_ref.c, oracle/cpu_fast.c) -- test
infrastructure only.

The oracle restates the reference's OpenMP path (src/omp_lattice.cpp,
src/lattice.cpp, src/lgca_models.h) on one byte per cell.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblgca_oracle.so")

FORCE_NONE, FORCE_X, FORCE_Y = 0, 1, 2
MODELS = {"HPP": 0, "FHP_I": 1, "FHP_II": 2, "FHP_III": 3}
NUM_DIR = {0: 4, 1: 6, 2: 7, 3: 7}

_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p_or_null = C.c_void_p

_lib = None


def build() -> None:
    """Compile the oracle library (and the reference-model driver when the
    reference checkout is present)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32, u64, i32, i64, f64 = C.c_uint32, C.c_uint64, C.c_int, C.c_int64, C.c_double
        L.oracle_philox4x32_10.argtypes = [C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
        L.oracle_threshold.argtypes = [f64]
        L.oracle_threshold.restype = u64
        L.oracle_chirality_word.argtypes = [u64, u64, u32, u32]
        L.oracle_chirality_word.restype = u32
        L.oracle_force_word.argtypes = [u64, u64, u32, u32, u64]
        L.oracle_force_word.restype = u32
        L.oracle_fill_draw.argtypes = [u64, u32, u32, u32]
        L.oracle_fill_draw.restype = u32
        L.oracle_num_dir.argtypes = [i32]
        L.oracle_num_dir.restype = i32
        for name in ("oracle_bounce_back", "oracle_bounce_forward_x", "oracle_bounce_forward_y"):
            getattr(L, name).argtypes = [i32, C.c_uint8]
            getattr(L, name).restype = C.c_uint8
        L.oracle_collide.argtypes = [i32, C.c_uint8, i32]
        L.oracle_collide.restype = C.c_uint8
        L.oracle_pull_source.argtypes = [i32, u32, u32, u32, u32, i32, C.POINTER(u32), C.POINTER(u32)]
        L.oracle_lattice_vec_bits.argtypes = [i32, i32, i32]
        L.oracle_lattice_vec_bits.restype = u32
        L.oracle_init_random.argtypes = [i32, u32, u32, _u8p_or_null, _u8p, u64, f64]
        L.oracle_step.argtypes = [i32, u32, u32, _u8p_or_null, _u8p, _u8p, u64, u64,
                                  i32, u64, u32, u32, u32, u32]
        L.oracle_run.argtypes = [i32, u32, u32, _u8p_or_null, _u8p, u64, u64, u64,
                                 i32, u64, u32, u32, u32, u32, i32]
        L.oracle_fast_run.argtypes = [i32, u32, u32, _u8p_or_null, _u8p, u64, u64, u64,
                                      i32, u64, u32, u32, u32, u32, i32]
        L.oracle_fast_run.restype = i32
        L.oracle_body_force_count.argtypes = [i32, u32, u32, _u8p_or_null, _u8p, u64, u64, i32, i32]
        L.oracle_body_force_count.restype = i64
        L.oracle_cell_post_process.argtypes = [i32, u32, u32, _u8p, _f32p, _f32p]
        L.oracle_mean_post_process.argtypes = [u32, u32, u32, _f32p, _f32p, _f32p, _f32p]
        L.oracle_coarse_velocity.argtypes = [C.c_size_t, _f32p, _f32p, _f32p]
        L.oracle_count_particles.argtypes = [C.c_size_t, _u8p]
        L.oracle_count_particles.restype = u64
        L.oracle_mean_velocity.argtypes = [u32, u32, _u8p_or_null, _f32p, _f32p, C.POINTER(f64)]
        _lib = L
    return _lib


def _types_ptr(types):
    if types is None:
        return None
    assert types.dtype == np.uint8 and types.flags.c_contiguous
    return types.ctypes.data


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def threshold(p: float) -> int:
    return int(lib().oracle_threshold(float(p)))


def collide(model: int, s: int, p: int) -> int:
    return int(lib().oracle_collide(model, s, p))


def pull_source(model, dim_x, dim_y, x, y, d):
    sx, sy = C.c_uint32(), C.c_uint32()
    lib().oracle_pull_source(model, dim_x, dim_y, x, y, d, C.byref(sx), C.byref(sy))
    return int(sx.value), int(sy.value)


def init_random(model, dim_x, dim_y, types, seed, p_fill, state=None):
    if state is None:
        state = np.zeros(dim_y * dim_x, dtype=np.uint8)
    lib().oracle_init_random(model, dim_x, dim_y, _types_ptr(types), state, seed, p_fill)
    return state


def force_args(force):
    """force = None or dict(dir='x'|'y', p=float, x0, x1, y0, y1)."""
    if not force:
        return (FORCE_NONE, 0, 0, 0, 0, 0)
    d = FORCE_X if force["dir"] == "x" else FORCE_Y
    return (d, threshold(force["p"]), force.get("x0", 0), force.get("x1", 2**32 - 1),
            force.get("y0", 0), force.get("y1", 2**32 - 1))


def step(model, dim_x, dim_y, types, state, seed, step_index, force=None):
    out = np.empty_like(state)
    lib().oracle_step(model, dim_x, dim_y, _types_ptr(types), state, out, seed, step_index,
                      *force_args(force))
    return out


def run(model, dim_x, dim_y, types, state, seed, step0, nsteps, force=None, nthreads=0):
    st = np.ascontiguousarray(state).copy()
    lib().oracle_run(model, dim_x, dim_y, _types_ptr(types), st, seed, step0, nsteps,
                     *force_args(force), nthreads)
    return st


def fast_run(model, dim_x, dim_y, types, state, seed, step0, nsteps, force=None, nthreads=0):
    """run() computed bit-sliced (cpu_fast.c), for long runs; pinned to run()."""
    st = np.ascontiguousarray(state).copy()
    if lib().oracle_fast_run(model, dim_x, dim_y, _types_ptr(types), st, seed, step0, nsteps,
                             *force_args(force), nthreads) != 0:
        raise RuntimeError("oracle_fast_run: its collision disagrees with oracle_collide")
    return st


def body_force_count(model, dim_x, dim_y, types, state, seed, call_index, direction, forcing):
    d = FORCE_X if direction == "x" else FORCE_Y
    return int(lib().oracle_body_force_count(model, dim_x, dim_y, _types_ptr(types), state,
                                             seed, call_index, d, forcing))


def cell_post_process(model, dim_x, dim_y, state):
    n = dim_x * dim_y
    cd = np.empty(n, dtype=np.float32)
    cm = np.empty(2 * n, dtype=np.float32)
    lib().oracle_cell_post_process(model, dim_x, dim_y, np.ascontiguousarray(state), cd, cm)
    return cd, cm


def mean_post_process(dim_x, dim_y, r, cd, cm):
    nc = (dim_x // (2 * r)) * (dim_y // (2 * r))
    md = np.empty(nc, dtype=np.float32)
    mm = np.empty(2 * nc, dtype=np.float32)
    lib().oracle_mean_post_process(dim_x, dim_y, r, cd, cm, md, mm)
    return md, mm


def coarse_velocity(md, mm):
    v = np.empty(2 * md.size, dtype=np.float32)
    lib().oracle_coarse_velocity(md.size, md, mm, v)
    return v


def post_process(model, dim_x, dim_y, r, state):
    cd, cm = cell_post_process(model, dim_x, dim_y, state)
    md, mm = mean_post_process(dim_x, dim_y, r, cd, cm)
    return cd, cm, md, mm


def count_particles(state) -> int:
    return int(lib().oracle_count_particles(state.size, np.ascontiguousarray(state)))


def mean_velocity(dim_x, dim_y, types, cd, cm):
    out = (C.c_double * 2)()
    lib().oracle_mean_velocity(dim_x, dim_y, _types_ptr(types), cd, cm, out)
    return float(out[0]), float(out[1])
