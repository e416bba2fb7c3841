"""Python mirror of the reference's Lattice<Model> interface over the HIP engine.

Method names, argument meaning and the order of side effects follow
src/lattice.h:32-238 and src/omp_lattice.h:27-78, so a script written against
OMP_Lattice reads the same here.  Every computation runs on the GPU through
liblgca_hip.so; host-side work is limited to what the reference does on the
host as setup (deriving sizes, rasterising boundary geometry).

Differences, all deliberate:
  * random numbers come from the counter hash (seed, step, cell) instead of
    rand(), so runs are reproducible and independent of thread count;
  * `collide_and_propagate(p)` ignores p, as OMP_Lattice does
    (src/omp_lattice.cpp:198 uses the per-cell random bit, never p);
  * explicit sizes are accepted besides the (test case, Re, Ma, r) form;
  * a per-cell body force (`set_forcing`) can be fused into every step.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import _capi as capi
from ._capi import check, lib

MODELS = {"HPP": capi.HPP, "FHP_I": capi.FHP_I, "FHP_II": capi.FHP_II, "FHP_III": capi.FHP_III}
NUM_DIR = {capi.HPP: 4, capi.FHP_I: 6, capi.FHP_II: 7, capi.FHP_III: 7}
INV_DIR0 = {capi.HPP: 2, capi.FHP_I: 3, capi.FHP_II: 3, capi.FHP_III: 3}
FLUID, SOLID_NO_SLIP, SOLID_SLIP = capi.FLUID, capi.SOLID_NO_SLIP, capi.SOLID_SLIP

f32 = np.float32


def model_id(model) -> int:
    if isinstance(model, str):
        return MODELS[model.upper().replace("-", "_")]
    return int(model)


def derive_parameters(model, test_case: str, Re: float, Ma_s: float, r: int) -> dict:
    """Sizes and physical parameters of the reference constructor
    (src/lattice.cpp:29-157), evaluated with its FP32/FP64 promotions."""
    m = model_id(model)
    q = NUM_DIR[m]
    if not Re > 1.0e-06 or not Ma_s > 1.0e-06:
        raise ValueError("Re and Ma must be positive")
    Re, Ma_s = f32(Re), f32(Ma_s)
    d = f32(f32(1.0) / f32(q))                                                   # :52
    nu = f32(1.0 / 12.0 * 1.0 / (float(d) * (1.0 - float(d)) ** 3.0) - 1.0 / 8.0)  # :55
    g = f32(2 / (2 + 2.0) * (1.0 - 2.0 * float(d)) / (1.0 - float(d)))          # :58
    nu_s = f32(nu / g)                                                           # :61
    c_s = f32(1.0 / math.sqrt(2))                                                # :64
    u = f32(Ma_s * c_s)                                                          # :67
    if test_case == "pipe":
        dim_y = int(float(f32(Re * nu_s) / u) + 0.5)                            # :73
    elif test_case == "karman":
        diameter = f32(f32(Re * nu_s) / u)
        dim_y = int(3.0 * float(diameter) + 0.5)                                 # :79
    elif test_case == "collision":
        dim_y = 8
    elif test_case in ("diffusion", "periodic", "box"):
        dim_y = int(Re)
    else:
        raise ValueError(f"invalid test case {test_case}")
    dim_y += 2 * r - dim_y % (2 * r)                                             # :100
    dim_x = 2 * dim_y if test_case in ("pipe", "karman", "collision") else dim_y
    if test_case == "collision":
        dim_x += 1
    bf_dir = "x" if test_case in ("pipe", "karman") else None
    return dict(dim_x=dim_x, dim_y=dim_y, d=d, nu=nu, g=g, nu_s=nu_s, c_s=c_s, u=u,
                Re=Re, Ma_s=Ma_s, bf_dir=bf_dir, test_case=test_case)


def disc(dim_x: int, dim_y: int, center_x: int, center_y: int, diameter: float) -> np.ndarray:
    """Cells with (float)sqrt(dx^2 + dy^2) < diameter / 2, as the reference
    tests them (src/lattice.cpp:264-279, 473-494)."""
    ys, xs = np.mgrid[0:dim_y, 0:dim_x]
    d2 = ((xs - center_x).astype(np.float64) ** 2 + (ys - center_y).astype(np.float64) ** 2)
    dist = np.sqrt(d2).astype(np.float32).astype(np.float64)
    return (dist < float(f32(diameter)) / 2.0).reshape(-1)


class Lattice:
    """One lattice on one GPU, mirroring Lattice<Model> / OMP_Lattice<Model>."""

    def __init__(self, model, test_case: str | None = None, Re: float | None = None,
                 Ma_s: float | None = None, coarse_graining_radius: int = 1, *,
                 dim_x: int | None = None, dim_y: int | None = None, seed: int = 0,
                 device: int = 0):
        self.model = model_id(model)
        self.NUM_DIR = NUM_DIR[self.model]
        r = int(coarse_graining_radius)
        if test_case is not None:
            p = derive_parameters(self.model, test_case, Re, Ma_s, r)
            dim_x, dim_y = p["dim_x"], p["dim_y"]
        else:
            p = dict(test_case=None, bf_dir=None, nu_s=f32(0), c_s=f32(1.0 / math.sqrt(2)),
                     u=f32(0), Ma_s=f32(Ma_s or 0), Re=f32(Re or 0))
        self._p = p
        self.m_bf_dir = p["bf_dir"]
        self._pending = 0
        self._hd = C.c_void_p()
        check(lib().lgca_create(C.byref(self._hd), self.model, int(dim_x), int(dim_y), r,
                                int(device)))
        self._attach(r)
        self.set_seed(seed)

    @classmethod
    def from_checkpoint(cls, path: str, device: int = 0) -> "Lattice":
        """A lattice resumed from a file written by save_checkpoint: same
        state, cell types, seed, step and forcing settings."""
        self = cls.__new__(cls)
        self._pending = 0
        self._hd = C.c_void_p()
        check(lib().lgca_load_checkpoint(C.byref(self._hd), os.fsencode(path), int(device)))
        info = self.info()
        self.model = int(info.model)
        self.NUM_DIR = NUM_DIR[self.model]
        self._p = dict(test_case=None, bf_dir=None, nu_s=f32(0), c_s=f32(1.0 / math.sqrt(2)),
                       u=f32(0), Ma_s=f32(0), Re=f32(0))
        self.m_bf_dir = None
        self._attach(int(info.coarse_radius))
        self.m_cell_type_cpu[:] = self.get_cell_types()
        return self

    def set_forcing_controller(self, forcing: int | None = None, equilibrium: int | None = None,
                               u: float | None = None) -> None:
        """The viewers' forcing controller on the GPU (apps/pipe/pipe_viewer.cpp:107-125):
        direction m_bf_dir, target velocity u (default u()), initial and equilibrium
        amounts (defaults: get_initial_forcing() / get_equilibrium_forcing())."""
        d = {"x": capi.FORCE_X, "y": capi.FORCE_Y}[self.m_bf_dir or "x"]
        f0 = self.get_initial_forcing() if forcing is None else forcing
        eq = self.get_equilibrium_forcing() if equilibrium is None else equilibrium
        uu = float(self.u()) if u is None else float(f32(u))
        self._check(lib().lgca_controller_set(self._h, d, uu, int(f0), int(eq)))

    def controlled_tick(self, steps: int) -> None:
        """One viewer tick on the GPU: decide on the snapshot's mean velocity,
        apply the count-mode force, run `steps` steps, take a new snapshot."""
        self.copy_data_to_device()
        self._check(lib().lgca_controller_tick(self._h, int(steps)))

    def controller_state(self):
        """(forcing, amount applied by the last tick or -1, (vx, vy) it saw)."""
        f, a = C.c_int64(), C.c_int64()
        v = (C.c_float * 2)()
        self._check(lib().lgca_controller_get(self._h, C.byref(f), C.byref(a), v))
        return int(f.value), int(a.value), (float(v[0]), float(v[1]))

    def save_checkpoint(self, path: str) -> None:
        self.copy_data_to_device()
        self._check(lib().lgca_save_checkpoint(self._h, os.fsencode(path)))

    def _attach(self, r: int) -> None:
        info = self.info()
        self.m_dim_x, self.m_dim_y = info.dim_x, info.dim_y
        self.m_num_cells = info.num_cells
        self.m_coarse_graining_radius = r
        self.m_coarse_dim_x, self.m_coarse_dim_y = info.coarse_dim_x, info.coarse_dim_y
        self.m_num_coarse_cells = info.num_coarse_cells
        # host-side cell types, as m_cell_type_cpu; uploaded before device work
        self.m_cell_type_cpu = np.zeros(self.m_num_cells, dtype=np.uint8)
        self._types_dirty = False
        self.m_cell_density_cpu = np.zeros(self.m_num_cells, dtype=np.float32)
        self.m_cell_momentum_cpu = np.zeros(2 * self.m_num_cells, dtype=np.float32)
        self.m_mean_density_cpu = np.zeros(self.m_num_coarse_cells, dtype=np.float32)
        self.m_mean_momentum_cpu = np.zeros(2 * self.m_num_coarse_cells, dtype=np.float32)

    # ---------------------------------------------------------------- handle
    def _flush(self) -> None:
        """Enqueues the steps collide_and_propagate has counted."""
        if self._pending:
            n, self._pending = self._pending, 0
            check(lib().lgca_step(self._hd, n), self._hd)

    @property
    def _h(self):
        """The handle for a device call, pending steps enqueued first."""
        self._flush()
        return self._hd

    def close(self) -> None:
        hd = getattr(self, "_hd", None)
        if hd:
            lib().lgca_destroy(hd)
            self._hd = C.c_void_p()
            self._pending = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def _check(self, code: int) -> None:
        check(code, self._hd)

    def info(self) -> capi.Info:
        inf = capi.Info()
        self._check(lib().lgca_get_info(self._h, C.byref(inf)))
        return inf

    # -------------------------------------------------------------- extras
    def set_seed(self, seed: int) -> None:
        self._check(lib().lgca_set_seed(self._h, int(seed) & (2**64 - 1)))

    def set_step(self, step: int) -> None:
        self._check(lib().lgca_set_step(self._h, int(step)))

    @property
    def steps(self) -> int:
        return int(self.info().step)

    def set_stream(self, stream_ptr: int | None) -> None:
        self._check(lib().lgca_set_stream(self._h, stream_ptr))

    def synchronize(self) -> None:
        self._check(lib().lgca_synchronize(self._h))

    def set_forcing(self, direction: str | None, p: float = 0.0, x0: int = 0,
                    x1: int | None = None, y0: int = 0, y1: int | None = None) -> None:
        d = {None: capi.FORCE_NONE, "x": capi.FORCE_X, "y": capi.FORCE_Y}[direction]
        x1 = self.m_dim_x if x1 is None else x1
        y1 = self.m_dim_y if y1 is None else y1
        self._check(lib().lgca_set_forcing(self._h, d, float(p), x0, x1, y0, y1))

    def set_tuning(self, steps_per_pass: int, tile_rows: int = 0) -> None:
        """Launch shape of step(): steps fused per HBM pass (1, 2, 4, 8; 0 = the
        one-step quad kernel; -1 = automatic, the default) and rows per tile
        (0 = automatic)."""
        self._check(lib().lgca_set_tuning(self._h, int(steps_per_pass), int(tile_rows)))

    def set_pipeline(self, on: bool) -> None:
        """Several passes per launch of step() (each XCD runs a band of row
        tiles pass after pass); see lgca_set_pipeline."""
        self._check(lib().lgca_set_pipeline(self._h, 1 if on else 0))

    def tune(self) -> None:
        """Times the candidate launch shapes now (without changing the state or
        the step) and keeps the fastest; see lgca_tune."""
        self.copy_data_to_device()
        self._check(lib().lgca_tune(self._h))

    def step(self, nsteps: int = 1) -> None:
        self.copy_data_to_device()
        self._check(lib().lgca_step(self._h, int(nsteps)))

    def set_cell_types(self, types: np.ndarray) -> None:
        t = np.ascontiguousarray(types, dtype=np.uint8).reshape(-1)
        assert t.size == self.m_num_cells
        self.m_cell_type_cpu[:] = t
        self._types_dirty = True
        self.copy_data_to_device()

    def get_cell_types(self) -> np.ndarray:
        out = np.empty(self.m_num_cells, dtype=np.uint8)
        self._check(lib().lgca_get_cell_types(self._h, out.ctypes.data))
        return out

    def set_state_u8(self, state: np.ndarray) -> None:
        s = np.ascontiguousarray(state, dtype=np.uint8).reshape(-1)
        assert s.size == self.m_num_cells
        self._check(lib().lgca_set_state_u8(self._h, s.ctypes.data))

    def get_state_u8(self) -> np.ndarray:
        out = np.empty(self.m_num_cells, dtype=np.uint8)
        self._check(lib().lgca_get_state_u8(self._h, out.ctypes.data))
        return out

    def apply_bc_cylinder(self, center_x: int, center_y: int, diameter: float) -> None:
        """Pipe walls plus a NO_SLIP disc (the karman set-up with an explicit
        position and size)."""
        self.apply_bc_pipe()
        self.m_cell_type_cpu[disc(self.m_dim_x, self.m_dim_y, center_x, center_y, diameter)] = \
            SOLID_NO_SLIP
        self._types_dirty = True

    def post_process_device(self):
        """Device pointers (mean density, mean momentum, velocity) of the coarse fields."""
        md, mm, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
        self._check(lib().lgca_post_process_device(self._h, C.byref(md), C.byref(mm), C.byref(v)))
        return md.value, mm.value, v.value

    def apply_body_force_count(self, direction: str, forcing: int) -> int:
        d = {"x": capi.FORCE_X, "y": capi.FORCE_Y}[direction]
        out = C.c_int64()
        self.copy_data_to_device()
        self._check(lib().lgca_apply_body_force_count(self._h, d, int(forcing), C.byref(out)))
        return int(out.value)

    # --------------------------------------------- reference: BC and set-up
    def apply_cell_type_all(self, cell_type: int) -> None:
        self.m_cell_type_cpu[:] = cell_type
        self._types_dirty = True

    def _grid(self) -> np.ndarray:
        return self.m_cell_type_cpu.reshape(self.m_dim_y, self.m_dim_x)

    def apply_boundary_cell_type_east(self, cell_type: int) -> None:
        self._grid()[:, -1] = cell_type
        self._types_dirty = True

    def apply_boundary_cell_type_north(self, cell_type: int) -> None:
        self._grid()[-1, :] = cell_type
        self._types_dirty = True

    def apply_boundary_cell_type_west(self, cell_type: int) -> None:
        self._grid()[:, 0] = cell_type
        self._types_dirty = True

    def apply_boundary_cell_type_south(self, cell_type: int) -> None:
        self._grid()[0, :] = cell_type
        self._types_dirty = True

    def apply_bc_periodic(self) -> None:
        self.apply_cell_type_all(FLUID)

    def apply_bc_pipe(self) -> None:
        self.apply_cell_type_all(FLUID)
        self.apply_boundary_cell_type_north(SOLID_NO_SLIP)
        self.apply_boundary_cell_type_south(SOLID_NO_SLIP)

    def apply_bc_reflecting(self, bounce_type: str) -> None:
        self.apply_cell_type_all(FLUID)
        if bounce_type == "back":
            t = SOLID_NO_SLIP
        elif bounce_type == "forward":
            t = SOLID_SLIP
        else:
            raise ValueError(f"Invalid bounce type {bounce_type}")
        self.apply_boundary_cell_type_east(t)
        self.apply_boundary_cell_type_north(t)
        self.apply_boundary_cell_type_west(t)
        self.apply_boundary_cell_type_south(t)

    def apply_bc_karman_vortex_street(self) -> None:
        # centre (dim_x/6, dim_y/2 + 1/10*dim_y) with 1/10 == 0 in integer
        # arithmetic; diameter dim_y/3 (src/lattice.cpp:252-280)
        self.apply_bc_cylinder(self.m_dim_x // 6, self.m_dim_y // 2, float(self.m_dim_y // 3))

    def init_zero(self) -> None:
        pass  # the state starts empty, as the reference's calloc'ed Bitset

    def init_random(self, p_fill: float | None = None) -> None:
        """src/lattice.cpp:198-217: FLUID cells, each direction occupied with
        probability 1/Q (or p_fill)."""
        self.copy_data_to_device()
        p = 1.0 / self.NUM_DIR if p_fill is None else p_fill
        self._check(lib().lgca_init_random(self._h, float(p)))

    def init_diffusion(self, p_fill: float | None = None) -> None:
        """src/lattice.cpp:465-496: init_random restricted to a central disc of
        diameter dim_y/4."""
        self.copy_data_to_device()
        region = disc(self.m_dim_x, self.m_dim_y, self.m_dim_x // 2, self.m_dim_y // 2,
                      float(self.m_dim_y // 4)).astype(np.uint8)
        p = 1.0 / self.NUM_DIR if p_fill is None else p_fill
        self._check(lib().lgca_init_random_masked(self._h, float(p), region.ctypes.data))

    def init_single(self, occupied_nodes) -> None:
        """src/lattice.cpp:297-308: set node (dir + cell * 8) for each entry."""
        st = self.get_state_u8()
        for node in occupied_nodes:
            cell, d = divmod(int(node), 8)
            st[cell] |= np.uint8(1 << d)
        self.set_state_u8(st)

    def init_single_collision(self) -> None:
        n = self.m_dim_x * self.m_dim_y
        self.init_single([(n // 2 + 1) * 8, (n // 2 + 5) * 8 + INV_DIR0[self.model]])

    # ---------------------------------------------- reference: compute path
    def setup_parallel(self) -> None:
        pass  # one GPU, one stream; nothing to configure

    def copy_data_to_device(self) -> None:
        if self._types_dirty:
            self._check(lib().lgca_set_cell_types(self._h, self.m_cell_type_cpu.ctypes.data))
            self._types_dirty = False

    def copy_data_from_device(self) -> None:
        self.synchronize()

    def collide_and_propagate(self, p: bool = False) -> None:
        """One step.  Steps are counted and enqueued together (one lgca_step
        of up to 64) before anything else reaches the device, so the tiled
        kernel advances several steps per pass; results are those of single
        steps."""
        self.copy_data_to_device()
        self._pending += 1
        if self._pending >= 64:
            self._flush()

    def apply_body_force(self, forcing: int) -> int:
        """Count-mode force in direction m_bf_dir (src/omp_lattice.cpp:254-346)."""
        if self.m_bf_dir is None:
            return 0
        return self.apply_body_force_count(self.m_bf_dir, forcing)

    def copy_data_to_output_buffer(self) -> None:
        self._check(lib().lgca_snapshot(self._h))

    def post_process(self, cells: bool = True) -> None:
        """Fills cell_density / cell_momentum (when `cells`) and the coarse
        fields from the output snapshot."""
        cd = self.m_cell_density_cpu.ctypes.data if cells else None
        cm = self.m_cell_momentum_cpu.ctypes.data if cells else None
        self._check(lib().lgca_post_process(self._h, cd, cm, self.m_mean_density_cpu.ctypes.data,
                                            self.m_mean_momentum_cpu.ctypes.data))

    def get_mean_velocity(self):
        out = (C.c_float * 2)()
        self._check(lib().lgca_mean_velocity(self._h, out))
        return [float(out[0]), float(out[1])]

    def get_n_particles(self) -> int:
        n = C.c_uint64()
        self._check(lib().lgca_count_particles(self._h, C.byref(n)))
        return int(n.value)

    def get_initial_forcing(self) -> int:
        return int(0.01 * self.m_num_cells)                                      # :450

    def get_equilibrium_forcing(self) -> int:
        p = self._p
        forcing = f32((8.0 * float(p["nu_s"]) * float(p["Ma_s"]) * float(p["c_s"]))
                      / float(f32(self.m_dim_y)) ** 2.0)                         # :458
        return int(math.ceil(0.5 * self.m_num_cells * float(forcing)))          # :460

    # ------------------------------------------------------------ getters
    def nu_s(self):
        return self._p["nu_s"]

    def c_s(self):
        return self._p["c_s"]

    def u(self):
        return self._p["u"]

    def dim_x(self) -> int:
        return self.m_dim_x

    def dim_y(self) -> int:
        return self.m_dim_y

    def num_cells(self) -> int:
        return self.m_num_cells

    def coarse_dim_x(self) -> int:
        return self.m_coarse_dim_x

    def coarse_dim_y(self) -> int:
        return self.m_coarse_dim_y

    def num_coarse_cells(self) -> int:
        return self.m_num_coarse_cells

    def cell_density(self) -> np.ndarray:
        return self.m_cell_density_cpu

    def mean_density(self) -> np.ndarray:
        return self.m_mean_density_cpu

    def cell_momentum(self) -> np.ndarray:
        return self.m_cell_momentum_cpu

    def mean_momentum(self) -> np.ndarray:
        return self.m_mean_momentum_cpu
