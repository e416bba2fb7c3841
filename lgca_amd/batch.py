"""Batched ensembles and serving over liblgca_hip.so (the lgca_batch_* entry
points of include/lgca.h).

A `Batch` owns one GPU stream and a worker thread in the library.  Jobs are
submitted at any time; the library admits them in order while their device
memory fits the HBM budget and advances every admitted job together, a few
steps per round.  A job is one run of one lattice: the set-up of the
reference's apply_bc_periodic / apply_bc_pipe / apply_bc_karman_vortex_street
(src/lattice.cpp:221-280) or given cell types, init_random
(src/lattice.cpp:198-217) from the job's own seed, `steps` steps of
collide_and_propagate (src/omp_lattice.cpp:100-249) with an optional per-cell
force, and post_process (src/omp_lattice.cpp:349-454).  Its bits equal what the
single-lattice `Lattice` gives for the same parameters.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi
from ._capi import check_batch, lib
from .lattice import model_id

GEOMETRIES = {"periodic": capi.GEOM_PERIODIC, "pipe": capi.GEOM_PIPE,
              "cylinder": capi.GEOM_CYLINDER, "types": capi.GEOM_TYPES}
FORCE_DIRS = {None: capi.FORCE_NONE, "x": capi.FORCE_X, "y": capi.FORCE_Y}
STATUS = {capi.JOB_QUEUED: "queued", capi.JOB_RUNNING: "running", capi.JOB_DONE: "done",
          capi.JOB_RELEASED: "released", capi.JOB_FAILED: "failed"}


@dataclass
class Job:
    """One lattice run.  `force_region` is (x0, x1, y0, y1), clipped to the
    lattice; None forces every cell."""
    model: object
    dim_x: int
    dim_y: int
    steps: int
    seed: int = 0
    coarse_radius: int = 1
    fill: float = 0.0
    geometry: str = "periodic"
    cyl_x: int = 0
    cyl_y: int = 0
    cyl_diameter: float = 0.0
    force_dir: str | None = None
    force_p: float = 0.0
    force_region: tuple | None = None
    cell_types: np.ndarray | None = None
    state: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    def desc(self) -> capi.JobDesc:
        d = capi.JobDesc()
        d.model = model_id(self.model)
        d.dim_x, d.dim_y = int(self.dim_x), int(self.dim_y)
        d.coarse_radius = int(self.coarse_radius)
        d.seed = int(self.seed) & (2 ** 64 - 1)
        d.steps = int(self.steps)
        d.fill = float(self.fill)
        d.geometry = GEOMETRIES[self.geometry]
        d.cyl_x, d.cyl_y = int(self.cyl_x), int(self.cyl_y)
        d.cyl_diameter = float(self.cyl_diameter)
        d.force_dir = FORCE_DIRS[self.force_dir]
        d.force_p = float(self.force_p)
        x0, x1, y0, y1 = self.force_region or (0, self.dim_x, 0, self.dim_y)
        d.fx0, d.fx1, d.fy0, d.fy1 = int(x0), int(x1), int(y0), int(y1)
        self._keep = []
        n = self.dim_x * self.dim_y
        for name in ("cell_types", "state"):
            arr = getattr(self, name)
            if arr is None:
                continue
            a = np.ascontiguousarray(arr, dtype=np.uint8).reshape(-1)
            if a.size != n:
                raise ValueError(f"{name} must hold dim_x * dim_y bytes")
            self._keep.append(a)
            setattr(d, name, a.ctypes.data)
        return d

    @property
    def num_cells(self) -> int:
        return self.dim_x * self.dim_y

    @property
    def coarse_shape(self) -> tuple:
        r2 = 2 * self.coarse_radius
        return (self.dim_y // r2, self.dim_x // r2)


def job_bytes(job: Job) -> int:
    """Device bytes the job holds while admitted (what the budget counts)."""
    out = C.c_uint64()
    d = job.desc()
    rc = lib().lgca_batch_job_bytes(C.byref(d), C.byref(out))
    if rc != capi.OK:
        raise ValueError("invalid job description")
    return out.value


class Batch:
    """Many lattices on one GPU, admitted against an HBM budget."""

    def __init__(self, device: int = 0, hbm_budget: int = 200 * 10 ** 9):
        h = C.c_void_p()
        check_batch(lib().lgca_batch_create(C.byref(h), int(device), int(hbm_budget)), None)
        self._h = h
        self._jobs: dict[int, Job] = {}

    # lifecycle -------------------------------------------------------------
    def close(self) -> None:
        if self._h:
            lib().lgca_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code: int) -> None:
        check_batch(code, self._h)

    # jobs --------------------------------------------------------------------
    def submit(self, job: Job) -> int:
        d = job.desc()
        jid = C.c_uint64()
        self._check(lib().lgca_batch_submit(self._h, C.byref(d), C.byref(jid)))
        self._jobs[jid.value] = job
        return jid.value

    def wait(self, job_id: int | None = None, timeout: float = -1.0) -> None:
        jid = capi.ALL_JOBS if job_id is None else int(job_id)
        self._check(lib().lgca_batch_wait(self._h, jid, float(timeout)))

    def poll(self, max_ids: int = 4096) -> list:
        ids = (C.c_uint64 * max_ids)()
        n = C.c_uint32()
        self._check(lib().lgca_batch_poll(self._h, ids, max_ids, C.byref(n)))
        return [ids[i] for i in range(n.value)]

    def info(self, job_id: int) -> dict:
        i = capi.JobInfo()
        self._check(lib().lgca_batch_job_info(self._h, int(job_id), C.byref(i)))
        return dict(status=STATUS[i.status], step=i.step, device_bytes=i.device_bytes,
                    submit_s=i.submit_s, admit_s=i.admit_s, done_s=i.done_s)

    def result(self, job_id: int, state: bool = True, coarse: bool = True):
        """(state bytes, coarse density, coarse momentum) of a finished job;
        parts not asked for are None."""
        job = self._jobs[job_id]
        st = np.empty(job.num_cells, dtype=np.uint8) if state else None
        ny, nx = job.coarse_shape
        md = np.empty(nx * ny, dtype=np.float32) if coarse else None
        mm = np.empty(2 * nx * ny, dtype=np.float32) if coarse else None
        ptr = lambda a: a.ctypes.data if a is not None else None  # noqa: E731
        self._check(lib().lgca_batch_result(self._h, int(job_id), ptr(st), ptr(md), ptr(mm)))
        return st, md, mm

    def release(self, job_id: int) -> None:
        self._check(lib().lgca_batch_release(self._h, int(job_id)))

    def set_tuning(self, steps_per_pass: int = 0, wave_target: int = 0,
                   max_active_cells: int = 0) -> None:
        self._check(lib().lgca_batch_set_tuning(self._h, int(steps_per_pass), int(wave_target),
                                                int(max_active_cells)))

    def stats(self) -> dict:
        s = capi.BatchStats()
        self._check(lib().lgca_batch_get_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_}

