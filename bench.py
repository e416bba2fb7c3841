"""Benchmark of the lattice-gas step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload karman|pipe|hpp]

Default workload: FHP-III Karman street, 32768 x 8192 cells (268,435,456),
NO_SLIP cylinder of diameter 512 at (4096, 4099), walls on rows 0 and 8191,
fill 0.25 per direction, per-cell x force p = 2e-3 on x < 64, hash seed 11.
A step is one collide-and-propagate time step of the whole lattice; K steps
are timed after W warm-up steps, with the state resident in HBM throughout.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public specification

WORKLOADS = {
    "karman": dict(model="FHP_III", dim_x=32768, dim_y=8192, r=32, seed=11, fill=0.25,
                   geometry=("cylinder", 4096, 4099, 512.0), force=("x", 2e-3, 0, 64),
                   label="FHP-III Karman street 32768x8192, cylinder d=512 at (4096,4099), "
                         "walls, fill 0.25, x force p=2e-3 on x<64, seed 11"),
    "pipe": dict(model="FHP_II", dim_x=2048, dim_y=512, r=16, seed=7, fill=0.3,
                 geometry=("pipe",), force=("x", 1e-3, 0, None),
                 label="FHP-II pipe 2048x512, walls, fill 0.3, x force p=1e-3, seed 7"),
    "hpp": dict(model="HPP", dim_x=256, dim_y=256, r=8, seed=1, fill=0.2, geometry=("periodic",),
                force=None, label="HPP 256x256 periodic, fill 0.2, seed 1"),
}
Q = {"HPP": 4, "FHP_I": 6, "FHP_II": 7, "FHP_III": 7}


def build_lattice(w):
    from lgca_amd import Lattice
    lat = Lattice(w["model"], dim_x=w["dim_x"], dim_y=w["dim_y"],
                  coarse_graining_radius=w["r"], seed=w["seed"])
    geo = w["geometry"]
    if geo[0] == "cylinder":
        lat.apply_bc_cylinder(geo[1], geo[2], geo[3])
    elif geo[0] == "pipe":
        lat.apply_bc_pipe()
    else:
        lat.apply_bc_periodic()
    lat.init_random(w["fill"])
    if w["force"]:
        d, p, x0, x1 = w["force"]
        lat.set_forcing(d, p, x0, x1)
    return lat


def cpu_baseline(w, lat, step0: int, budget_s: float):
    """The CPU oracle (a restatement of the reference's OpenMP step) on the same
    lattice and state, for as many steps as fit in about `budget_s` seconds."""
    from oracle import oracle as orc
    threads = min(16, os.cpu_count() or 1)
    state = lat.get_state_u8()
    types = lat.m_cell_type_cpu.copy()
    model = {"HPP": 0, "FHP_I": 1, "FHP_II": 2, "FHP_III": 3}[w["model"]]
    force = None
    if w["force"]:
        d, p, x0, x1 = w["force"]
        force = dict(dir=d, p=p, x0=x0, x1=x1 if x1 is not None else w["dim_x"])
    n = w["dim_x"] * w["dim_y"]
    t = time.perf_counter()
    state = orc.run(model, w["dim_x"], w["dim_y"], types, state, w["seed"], step0, 1, force,
                    nthreads=threads)
    t1 = time.perf_counter() - t
    more = max(0, int(budget_s / max(t1, 1e-9)) - 1)
    steps = 1
    if more:
        t = time.perf_counter()
        orc.run(model, w["dim_x"], w["dim_y"], types, state, w["seed"], step0 + 1, more, force,
                nthreads=threads)
        t1 += time.perf_counter() - t
        steps += more
    return {"value": round(n * steps / t1 / 1e9, 4), "unit": "GSUP/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} step(s) of the same {w['dim_x']}x{w['dim_y']} workload from "
                      f"step {step0}, {t1:.1f} s, oracle/cpu_ref.c with {threads} OpenMP threads"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="karman", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--steps-per-pass", type=int, default=None,
                    help="steps fused per HBM pass (1, 2, 4, 8; 0 = one-step quad kernel)")
    ap.add_argument("--tile-rows", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU y-strip decomposition is a design, not a build)", file=sys.stderr)
        return 2

    from lgca_amd import hiprt
    w = WORKLOADS[args.workload]
    lat = build_lattice(w)
    if args.steps_per_pass is not None:
        lat.set_tuning(args.steps_per_pass, args.tile_rows)
    # the step kernels run on this stream, so events recorded on it bracket them
    stream = hiprt.Stream()
    lat.set_stream(stream.ptr.value)
    n = w["dim_x"] * w["dim_y"]

    lat.step(args.warmup)
    stream.synchronize()
    n0 = lat.get_n_particles()
    start, end = hiprt.Event(), hiprt.Event()
    stream.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    lat.step(args.steps)
    end.record(stream)
    end.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_ms(end)
    n1 = lat.get_n_particles()

    sups = n * args.steps / (ms / 1e3)
    bytes_per_step = n * 2 * Q[w["model"]] / 8
    achieved = bytes_per_step * args.steps / (ms / 1e3) / 1e9
    # HBM bytes per launch of the dominant kernel, measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE for the default launch shape (see the file)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"step_kernel_{args.workload}_traffic.json")
    if os.path.exists(tfile) and args.steps_per_pass is None and args.tile_rows == 0:
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    rec = {
        "metric": "site updates/s (GSUP/s)",
        "value": round(sups / 1e9, 3),
        "unit": "GSUP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": w["label"], "cells": n, "directions": Q[w["model"]],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "particles_conserved": n0 == n1,
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(w, lat, args.warmup + args.steps, args.cpu_seconds)
    lat.close()
    start.close()
    end.close()
    stream.close()
    print(json.dumps(rec), flush=True)
    return 0 if n0 == n1 else 1


if __name__ == "__main__":
    sys.exit(main())
