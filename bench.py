"""Benchmark of the lattice-gas step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload karman|pipe|hpp]
                    [--ensemble-jobs 512] [--serve-rates 50,200,1000,2000]

Default workload: FHP-III Karman street, 32768 x 8192 cells (268,435,456),
NO_SLIP cylinder of diameter 512 at (4096, 4099), walls on rows 0 and 8191,
fill 0.25 per direction, per-cell x force p = 2e-3 on x < 64, hash seed 11.
A step is one collide-and-propagate time step of the whole lattice.  The
launch shape is tuned first (lgca_tune: timed passes that leave the state and
the step unchanged), then K steps are timed after W warm-up steps, with the
state resident in HBM throughout.

The other two single-lattice workloads are timed the same way, at their full
step counts, and ride on the line under "single_lattice":
  HPP 256x256 periodic, fill 0.2, seed 1, 1000 steps;
  FHP-II pipe 2048x512, walls, fill 0.3, x force p=1e-3, seed 7, 20000 steps.

Two more measurements ride on the same JSON line:
  "ensemble": 512 FHP-II jobs (long side 128..2048 step 64, aspect 1..4,
      periodic / pipe / cylinder of radius 8..L/8, fill 0.1-0.5, own seed,
      1000-10000 steps, 32x32 coarse fields) submitted at once to one batch;
      jobs/s against the single-lattice path run one job at a time (timed on
      an evenly spaced sample of the same jobs), with bit-equality of sampled
      jobs between the two.
  "serving": Poisson arrivals at each rate of jobs uniform over HPP, FHP-I,
      FHP-II, FHP-III, 256x256..4096x1024, periodic / pipe (x force p
      1e-4..5e-3) / cylinder (random radius and position), 500-20000 steps,
      own seed, admitted against a 200 GB budget; p50/p99 latency and jobs/s,
      next to the one-at-a-time path (each job's measured single-lattice
      service time, queued first-come first-served on the same arrivals).
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


def host_cpu():
    """(threads to use, CPUs this process may run on, CPUs of the machine, CPU
    model).  The thread count is the process's CPU share: OMP_NUM_THREADS when
    set (the GPU machines set it to their share), else the CPUs in its
    affinity mask."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    threads = affinity
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, affinity, os.cpu_count() or 1, model


def cpu_baseline(w, lat, step0: int, budget_s: float):
    """The CPU oracle (a restatement of the reference's OpenMP step) on the same
    lattice and state, for as many steps as fit in about `budget_s` seconds."""
    from oracle import oracle as orc
    threads, affinity, ncpu, cpu_model = host_cpu()
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
            "kind": "port", "nproc": affinity, "machine_cpus": ncpu, "cpu_model": cpu_model,
            "sample": f"{steps} step(s) of the same {w['dim_x']}x{w['dim_y']} workload from "
                      f"step {step0}, {t1:.1f} s, oracle/cpu_ref.c (byte per cell, pull, "
                      f"OpenMP over rows) with {threads} OpenMP threads, the process's CPU "
                      f"share ({affinity} CPUs in its affinity mask, {ncpu} in the machine)"}


def time_steps(w, steps: int, warmup: int, tuning=None, tune: bool = True):
    """One lattice of workload w: set up, tune the launch shape, warm up, then
    time `steps` steps with HIP events on the stream its kernels run on."""
    from lgca_amd import hiprt
    lat = build_lattice(w)
    if tuning is not None:
        lat.set_tuning(*tuning)
    # the step kernels run on this stream, so events recorded on it bracket them
    stream = hiprt.Stream()
    lat.set_stream(stream.ptr.value)
    if tune:
        lat.tune()
    n = w["dim_x"] * w["dim_y"]
    lat.step(warmup)
    stream.synchronize()
    n0 = lat.get_n_particles()
    start, end = hiprt.Event(), hiprt.Event()
    stream.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    lat.step(steps)
    end.record(stream)
    end.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_ms(end)
    n1 = lat.get_n_particles()
    # the coarse fields of the final state (snapshot, then the coarse kernel)
    lat.copy_data_to_output_buffer()
    lat.post_process_device()  # allocates the coarse arrays; the second call is timed
    stream.synchronize()
    start.record(stream)
    lat.post_process_device()
    end.record(stream)
    end.synchronize()
    coarse_ms = start.elapsed_ms(end)
    info = lat.info()
    lat.close()
    start.close()
    end.close()
    stream.close()
    sups = n * steps / (ms / 1e3)
    achieved = n * 2 * Q[w["model"]] / 8 * steps / (ms / 1e3) / 1e9
    return dict(ms=ms, wall=wall, sups=sups, achieved=achieved, conserved=n0 == n1,
                steps_per_pass=int(info.steps_per_pass), tile_rows=int(info.tile_rows),
                coarse_ms=coarse_ms)


def other_configs():
    """The HPP 256x256 and FHP-II 2048x512 workloads at their full step counts."""
    out = {}
    for key, name, steps in (("hpp_256x256_periodic", "hpp", 1000),
                             ("fhp2_pipe_2048x512", "pipe", 20000)):
        w = WORKLOADS[name]
        r = time_steps(w, steps, warmup=8)
        out[key] = {
            "workload": w["label"], "steps": steps, "ms_per_step": round(r["ms"] / steps, 6),
            "value": round(r["sups"] / 1e9, 3), "unit": "GSUP/s",
            "steps_per_pass": r["steps_per_pass"], "tile_rows": r["tile_rows"],
            "roofline": {"bound": "hbm", "achieved": round(r["achieved"], 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(r["achieved"] / HBM_PEAK_GBS, 4)},
            "particles_conserved": r["conserved"],
        }
        print(f"[bench] {key}: {out[key]['value']} GSUP/s", file=sys.stderr, flush=True)
    return out


# ------------------------------------------------------------ ensembles

def single_lattice_run(job, want_state: bool = False):
    """One job through the plain single-lattice path, start to finish: create,
    set-up, fill, steps, post-process, coarse results, destroy.  Returns
    (seconds, state bytes or None, coarse density, coarse momentum); the state
    bytes, fetched only to compare bits, are not timed (a batch job's state
    stays on the GPU too)."""
    from lgca_amd import Lattice
    t = time.perf_counter()
    lat = Lattice(job.model, dim_x=job.dim_x, dim_y=job.dim_y,
                  coarse_graining_radius=job.coarse_radius, seed=job.seed)
    if job.geometry == "pipe":
        lat.apply_bc_pipe()
    elif job.geometry == "cylinder":
        lat.apply_bc_cylinder(job.cyl_x, job.cyl_y, job.cyl_diameter)
    lat.init_random(job.fill)
    if job.force_dir is not None:
        lat.set_forcing(job.force_dir, job.force_p)
    lat.step(job.steps)
    lat.copy_data_to_output_buffer()
    lat.post_process(cells=False)
    md, mm = lat.mean_density().copy(), lat.mean_momentum().copy()
    lat.synchronize()
    dt = time.perf_counter() - t
    st = lat.get_state_u8() if want_state else None
    t = time.perf_counter()
    lat.close()
    return dt + time.perf_counter() - t, st, md, mm


def ensemble_bench(n_jobs: int, sample: int):
    import numpy as np
    from lgca_amd import Batch
    from lgca_amd.workloads import ensemble_jobs
    jobs = ensemble_jobs(n_jobs)
    su = sum(j.dim_x * j.dim_y * j.steps for j in jobs)
    with Batch(hbm_budget=200 * 10 ** 9) as b:
        t0 = time.perf_counter()
        ids = [b.submit(j) for j in jobs]
        b.wait(timeout=600)
        t_batch = time.perf_counter() - t0
        stats = b.stats()
        picks = list(range(0, n_jobs, max(1, n_jobs // max(1, sample))))[:sample]
        got = {i: b.result(ids[i]) for i in picks}
    t_one, identical = 0.0, True
    for k, i in enumerate(picks):
        dt, st, md, mm = single_lattice_run(jobs[i], want_state=i in got)
        t_one += dt
        if i in got:
            gst, gmd, gmm = got[i]
            identical &= bool(np.array_equal(gst, st) and np.array_equal(gmd.view(np.uint32),
                                                                         md.view(np.uint32))
                              and np.array_equal(gmm.view(np.uint32), mm.view(np.uint32)))
    batch_rate = n_jobs / t_batch
    one_rate = len(picks) / t_one
    return {
        "jobs": n_jobs, "site_updates": su, "seconds": round(t_batch, 3),
        "jobs_per_s": round(batch_rate, 2), "gsup_per_s": round(su / t_batch / 1e9, 1),
        "one_at_a_time_jobs_per_s": round(one_rate, 2),
        "one_at_a_time_sample": len(picks), "speedup": round(batch_rate / one_rate, 2),
        "bits_identical_sampled": identical, "rounds": stats["rounds"],
        "launches": stats["launches"], "tuned_steps_per_pass": stats["steps_per_pass"],
        "tuned_wave_target": stats["wave_target"],
        "workload": "512 FHP-II lattices, L 128..2048, aspect 1..4, periodic/pipe/cylinder, "
                    "fill 0.1-0.5, 1000-10000 steps" if n_jobs == 512 else f"{n_jobs} C4-style jobs",
    }


def serving_bench(rates, n_jobs: int, max_active_cells: int):
    import numpy as np
    from lgca_amd import Batch
    from lgca_amd.workloads import poisson_arrivals, serving_jobs
    jobs = serving_jobs(n_jobs)
    unit_arrivals = poisson_arrivals(n_jobs, 1.0)
    # one-at-a-time service times, measured once per job; the single-lattice
    # results of an evenly spaced sample are kept to compare bits with
    sample = set(range(0, n_jobs, max(1, n_jobs // 8)))
    service, alone = [], {}
    for i, j in enumerate(jobs):
        dt, st, md, mm = single_lattice_run(j, want_state=i in sample)
        service.append(dt)
        if i in sample:
            alone[i] = (st, md, mm)
    service = np.array(service)
    out = []
    identical = True
    for lam in rates:
        arr = unit_arrivals / lam
        with Batch(hbm_budget=200 * 10 ** 9) as b:
            b.set_tuning(0, 0, max_active_cells)
            ids = []
            t0 = time.monotonic()
            for i, j in enumerate(jobs):
                delay = t0 + arr[i] - time.monotonic()
                if delay > 0:
                    time.sleep(delay)
                ids.append(b.submit(j))
            b.wait(timeout=900)
            infos = [b.info(i) for i in ids]
            # each sampled job's bits in the mix equal its bits alone
            for i in sorted(sample):
                gst, gmd, gmm = b.result(ids[i])
                st, md, mm = alone[i]
                identical &= bool(np.array_equal(gst, st)
                                  and np.array_equal(gmd.view(np.uint32), md.view(np.uint32))
                                  and np.array_equal(gmm.view(np.uint32), mm.view(np.uint32)))
        lat = np.array([x["done_s"] - x["submit_s"] for x in infos])
        span = max(x["done_s"] for x in infos) - min(x["submit_s"] for x in infos)
        # one at a time: the same arrivals served first-come first-served
        fin = np.empty(n_jobs)
        prev = -np.inf
        for i in range(n_jobs):
            prev = max(arr[i], prev) + service[i]
            fin[i] = prev
        lat1 = fin - arr
        out.append({
            "rate": lam, "jobs": n_jobs,
            "jobs_per_s": round(n_jobs / span, 2),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
            "one_at_a_time_jobs_per_s": round(n_jobs / (fin[-1] - arr[0]), 2),
            "one_at_a_time_p50_ms": round(float(np.percentile(lat1, 50)) * 1e3, 2),
            "one_at_a_time_p99_ms": round(float(np.percentile(lat1, 99)) * 1e3, 2),
        })
        print(f"[serving] rate {lam}: {out[-1]}", file=sys.stderr, flush=True)
    return {"levels": out, "max_active_cells": max_active_cells,
            "bits_identical_sampled": identical, "bits_sample": len(sample),
            "mean_service_ms_one_at_a_time": round(float(service.mean()) * 1e3, 2),
            "workload": "Poisson arrivals; HPP/FHP-I/FHP-II/FHP-III, 256x256..4096x1024, "
                        "periodic/pipe(force)/cylinder, 500-20000 steps, 200 GB budget"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--workload", default="karman", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the HPP 256x256 and FHP-II 2048x512 single-lattice rates")
    ap.add_argument("--steps-per-pass", type=int, default=None,
                    help="steps fused per HBM pass (1, 2, 4, 8; 0 = one-step quad kernel; "
                         "-1 = automatic, the default)")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--ensemble-jobs", type=int, default=512, help="0 skips the ensemble leg")
    ap.add_argument("--ensemble-sample", type=int, default=32)
    ap.add_argument("--serve-rates", default="50,200,1000,2000", help="empty skips serving")
    ap.add_argument("--serve-jobs", type=int, default=120)
    ap.add_argument("--serve-max-active-cells", type=int, default=1 << 28)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU y-strip decomposition is a design, not a build)", file=sys.stderr)
        return 2

    w = WORKLOADS[args.workload]
    tuning = None
    if args.steps_per_pass is not None or args.tile_rows:
        tuning = (-1 if args.steps_per_pass is None else args.steps_per_pass, args.tile_rows)
    r = time_steps(w, args.steps, args.warmup, tuning)
    n = w["dim_x"] * w["dim_y"]
    ms = r["ms"]
    # HBM bytes per launch of the dominant kernel, measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE for the default launch shape (see the file)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"step_kernel_{args.workload}_traffic.json")
    if os.path.exists(tfile) and tuning is None:
        with open(tfile) as f:
            tr = json.load(f)
        if (tr.get("steps_per_pass"), tr.get("tile_rows")) == (r["steps_per_pass"], r["tile_rows"]):
            traffic = tr.get("hbm_bytes_per_launch")
    rec = {
        "metric": "site updates/s (GSUP/s)",
        "value": round(r["sups"] / 1e9, 3),
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
                   "parallelism": "single GPU",
                   "steps_per_pass": r["steps_per_pass"], "tile_rows": r["tile_rows"]},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r["achieved"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "particles_conserved": r["conserved"],
        "wall_s": round(r["wall"], 3),
        "coarse_fields_ms": round(r["coarse_ms"], 3),
    }
    if not args.no_other_configs:
        rec["single_lattice"] = other_configs()
    if args.ensemble_jobs > 0:
        print("[bench] ensemble", file=sys.stderr, flush=True)
        rec["ensemble"] = ensemble_bench(args.ensemble_jobs, args.ensemble_sample)
    rates = [float(x) for x in args.serve_rates.split(",") if x.strip()]
    if rates:
        print("[bench] serving", file=sys.stderr, flush=True)
        rec["serving"] = serving_bench(rates, args.serve_jobs, args.serve_max_active_cells)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        lat = build_lattice(w)
        rec["cpu_baseline"] = cpu_baseline(w, lat, 0, args.cpu_seconds)
        lat.close()
    print(json.dumps(rec), flush=True)
    return 0 if r["conserved"] else 1


if __name__ == "__main__":
    sys.exit(main())
