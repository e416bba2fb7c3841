"""Batched ensembles and serving (lgca_batch_* of include/lgca.h), through the C ABI.

The batch adds scheduling, not arithmetic: every job's final state bytes and
FP32 coarse fields must equal, bit for bit, the oracle's run of the same job
(set-up, hash fill, steps with the per-cell force, post-process), and a job's
bits must not depend on what runs beside it, when it was admitted, or the
launch shape the autotuner picked.  Large jobs, which the oracle would take
minutes for, are checked against the single-lattice GPU path instead, itself
pinned to the oracle by test_gpu_parity.py.
"""
import sys
import time

import numpy as np
import pytest

from lgca_amd.lattice import model_id
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

Q = {0: 4, 1: 6, 2: 7, 3: 7}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def job_types(job):
    """Cell types of a job as the reference set-up writes them."""
    from lgca_amd import disc
    from lgca_amd.batch import Job  # noqa: F401
    n = job.dim_x * job.dim_y
    if job.geometry == "types":
        return np.ascontiguousarray(job.cell_types, dtype=np.uint8).reshape(-1)
    t = np.zeros(n, dtype=np.uint8)
    if job.geometry in ("pipe", "cylinder"):
        g = t.reshape(job.dim_y, job.dim_x)
        g[0, :] = 1
        g[-1, :] = 1
    if job.geometry == "cylinder":
        t[disc(job.dim_x, job.dim_y, job.cyl_x, job.cyl_y, job.cyl_diameter)] = 1
    return t


def force_dict(job):
    if job.force_dir is None or job.force_p <= 0:
        return None
    x0, x1, y0, y1 = job.force_region or (0, job.dim_x, 0, job.dim_y)
    return dict(dir=job.force_dir, p=job.force_p, x0=x0, x1=x1, y0=y0, y1=y1)


def oracle_job(job):
    m = model_id(job.model)
    types = job_types(job)
    if job.state is not None:
        st = np.ascontiguousarray(job.state, dtype=np.uint8).reshape(-1).copy()
    else:
        st = orc.init_random(m, job.dim_x, job.dim_y, types, job.seed, job.fill)
    st = orc.run(m, job.dim_x, job.dim_y, types, st, job.seed, 0, job.steps, force=force_dict(job),
                 nthreads=8)
    _, _, md, mm = orc.post_process(m, job.dim_x, job.dim_y, job.coarse_radius, st)
    return st, md, mm


def single_lattice_job(job):
    """The same job on the single-lattice GPU path (lgca_create .. lgca_step)."""
    from lgca_amd import Lattice
    lat = Lattice(model_id(job.model), dim_x=job.dim_x, dim_y=job.dim_y,
                  coarse_graining_radius=job.coarse_radius, seed=job.seed)
    try:
        lat.set_cell_types(job_types(job))
        if job.state is not None:
            lat.set_state_u8(job.state)
        else:
            lat.init_random(job.fill)
        f = force_dict(job)
        if f:
            lat.set_forcing(f["dir"], f["p"], f["x0"], f["x1"], f["y0"], f["y1"])
        lat.step(job.steps)
        st = lat.get_state_u8()
        lat.copy_data_to_output_buffer()
        lat.post_process(cells=False)
        return st, lat.mean_density().copy(), lat.mean_momentum().copy()
    finally:
        lat.close()


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                                 b.view(np.uint32) if b.dtype == np.float32 else b)


def check_job(name, got, want):
    st, md, mm = got
    wst, wmd, wmm = want
    assert np.array_equal(st, wst), f"{name}: state differs in {(st != wst).sum()} cells"
    assert same_bits(md.reshape(-1), wmd.reshape(-1)), f"{name}: coarse density differs"
    assert same_bits(mm.reshape(-1), wmm.reshape(-1)), f"{name}: coarse momentum differs"


def mixed_jobs(rng):
    """Small jobs over every model, geometry, width class, force and step
    remainder the batch has to handle."""
    from lgca_amd import Job
    types70 = np.zeros(70 * 10, dtype=np.uint8)
    g = types70.reshape(10, 70)
    g[rng.random((10, 70)) < 0.08] = 1
    g[rng.random((10, 70)) < 0.05] = 2
    g[0, :] = 2
    g[-1, :35] = 1
    state64 = rng.integers(0, 1 << 7, 64 * 32, dtype=np.uint8)
    return [
        Job(0, 64, 48, 37, seed=1, fill=0.2, coarse_radius=8),
        Job(0, 33, 20, 10, seed=2, fill=0.3, geometry="pipe", coarse_radius=2),
        Job(0, 16, 6, 9, seed=3, fill=0.5),
        Job(1, 96, 64, 25, seed=4, fill=0.25, geometry="cylinder", cyl_x=30, cyl_y=32,
            cyl_diameter=20.0, force_dir="x", force_p=0.01, force_region=(0, 20, 0, 64),
            coarse_radius=4),
        Job(2, 200, 64, 41, seed=5, fill=0.3, geometry="pipe", force_dir="x", force_p=0.005,
            coarse_radius=16),
        Job(3, 128, 128, 64, seed=6, fill=0.2, geometry="cylinder", cyl_x=50, cyl_y=70,
            cyl_diameter=33.5, force_dir="y", force_p=0.02, force_region=(10, 90, 5, 100),
            coarse_radius=8),
        Job(2, 70, 10, 13, seed=7, fill=0.4, geometry="types", cell_types=types70),
        Job(2, 64, 32, 8, seed=8, state=state64, geometry="pipe", coarse_radius=2),
        Job(3, 40, 20, 0, seed=9, fill=0.3),
        Job(2, 2048, 64, 12, seed=10, fill=0.1, geometry="pipe", coarse_radius=16),
        Job(1, 1000, 30, 17, seed=11, fill=0.2, force_dir="x", force_p=0.003, coarse_radius=3),
        Job(2, 30, 2, 11, seed=12, fill=0.5),
        Job(3, 100, 40, 33, seed=2 ** 40 + 13, fill=0.3, geometry="cylinder", cyl_x=-3, cyl_y=20,
            cyl_diameter=17.0, force_dir="x", force_p=1.0, force_region=(60, 80, 0, 40)),
    ]


def run_batch(jobs, tuning=None, budget=8 * 10 ** 9):
    from lgca_amd import Batch
    with Batch(hbm_budget=budget) as b:
        if tuning:
            b.set_tuning(*tuning)
        ids = [b.submit(j) for j in jobs]
        b.wait(timeout=300)
        out = [b.result(i) for i in ids]
        stats = b.stats()
        for i in ids:
            assert b.info(i)["status"] == "done"
            b.release(i)
    return out, stats


def test_mixed_jobs_match_oracle():
    rng = np.random.default_rng(20)
    jobs = mixed_jobs(rng)
    say(f"[batch] {len(jobs)} mixed jobs")
    got, stats = run_batch(jobs)
    say(f"[batch] stats {stats}")
    assert stats["jobs_done"] == len(jobs)
    for i, job in enumerate(jobs):
        check_job(f"job {i} {job.model}/{job.dim_x}x{job.dim_y}/{job.geometry}", got[i],
                  oracle_job(job))


@pytest.mark.parametrize("tuning", [(2, 256, 0), (4, 64, 0), (8, 100000, 0), (8, 0, 0),
                                    (0, 0, 30000)])
def test_results_do_not_depend_on_tuning_or_mix(tuning):
    """Each job alone under a fixed launch shape (or with an active-cell cap
    that serialises the mix) gives the bits it gives in the autotuned mix."""
    rng = np.random.default_rng(20)
    jobs = mixed_jobs(rng)
    say(f"[batch] tuning {tuning}")
    want, _ = run_batch(jobs)
    got, _ = run_batch(jobs, tuning=tuning)
    for i in range(len(jobs)):
        check_job(f"job {i} tuning {tuning}", got[i], want[i])
    alone, _ = run_batch([jobs[5]], tuning=tuning)
    check_job("job 5 alone", alone[0], want[5])


def strip_jobs(rng):
    """Jobs wider than 64 words, which the batch runs in x-strips: three
    strips, two with a ragged last word, given types with solids in a few
    rows only, forcing in one column band."""
    from lgca_amd import Job
    sparse = np.zeros(3968 * 96, dtype=np.uint8)
    g = sparse.reshape(96, 3968)
    for r in (1, 40, 41, 94):
        g[r, 700:760] = 1
    g[60, -5:] = 2
    g[rng.random((96, 3968)) < 0.0005] = 1
    return [
        Job(2, 4160, 300, 29, seed=51, fill=0.3, geometry="pipe", force_dir="x", force_p=0.01,
            coarse_radius=16),
        Job(1, 2100, 140, 23, seed=52, fill=0.25, geometry="cylinder", cyl_x=2000, cyl_y=70,
            cyl_diameter=40.0, coarse_radius=5),
        Job(0, 3968, 96, 31, seed=53, fill=0.2, geometry="types", cell_types=sparse, coarse_radius=8),
        Job(3, 4096, 130, 17, seed=54, fill=0.3, geometry="cylinder", cyl_x=100, cyl_y=120,
            cyl_diameter=30.0, force_dir="y", force_p=0.05, force_region=(3000, 3100, 0, 130),
            coarse_radius=13),
    ]


def test_strip_mode_jobs_match_oracle():
    """Jobs in x-strips (rows over 64 words, the serving sizes) under the
    autotuner, under tall tiles (over 64 rows: the walk's solid-row mask is
    refreshed mid-tile) and under short 2-step tiles: every job equals the
    oracle."""
    rng = np.random.default_rng(50)
    jobs = strip_jobs(rng)
    want = [oracle_job(j) for j in jobs]
    for tuning in (None, (4, 16, 0), (2, 100000, 0), (8, 0, 0)):
        say(f"[batch] strip-mode jobs, tuning {tuning}")
        got, stats = run_batch(jobs, tuning=tuning)
        assert stats["jobs_done"] == len(jobs)
        for i, job in enumerate(jobs):
            check_job(f"strip job {i} {job.model}/{job.dim_x}x{job.dim_y}/{job.geometry} "
                      f"tuning {tuning}", got[i], want[i])


def test_budget_admission_and_release():
    """With a budget that holds two jobs, the rest wait in order and are
    admitted as finished jobs are released; a job larger than the budget is
    refused at submit."""
    from lgca_amd import Batch, Job, LgcaError, job_bytes
    jobs = [Job(2, 256, 128, 24 + 4 * i, seed=100 + i, fill=0.3, geometry="pipe",
                force_dir="x", force_p=0.002, coarse_radius=8) for i in range(6)]
    per = max(job_bytes(j) for j in jobs)
    budget = 2 * per + per // 2
    say(f"[batch] budget {budget} bytes, {per} per job")
    with Batch(hbm_budget=budget) as b:
        with pytest.raises(LgcaError) as ei:
            b.submit(Job(2, 4096, 2048, 4, fill=0.1))
        assert ei.value.code == 5
        ids = [b.submit(j) for j in jobs]
        done, results = [], {}
        t0 = time.monotonic()
        while len(done) < len(ids):
            assert time.monotonic() - t0 < 120, "jobs did not finish"
            s = b.stats()
            assert s["bytes_in_use"] <= budget
            assert s["jobs_active"] <= 2
            for i in b.poll():
                results[i] = b.result(i)
                b.release(i)
                done.append(i)
            time.sleep(0.002)
        admits = [b.info(i)["admit_s"] for i in ids]
        assert admits == sorted(admits), "admission is in submission order"
    for k, i in enumerate(ids):
        check_job(f"budget job {k}", results[i], single_lattice_job(jobs[k]))


def test_continuous_batching_mid_run_arrivals():
    """Jobs submitted while a long job runs are admitted at once, overlap with
    it on the GPU and finish before it; every result equals the job run alone."""
    from lgca_amd import Batch, Job
    # ~10^4 rounds: far longer than submitting the short jobs takes
    long_job = Job(3, 2048, 1024, 40000, seed=31, fill=0.25, geometry="cylinder", cyl_x=340,
                   cyl_y=512, cyl_diameter=170.0, force_dir="x", force_p=0.002,
                   force_region=(0, 64, 0, 1024), coarse_radius=16)
    short = [Job(m, 128 + 64 * m, 64, 40 + 8 * m, seed=40 + m, fill=0.3,
                 geometry="pipe" if m % 2 else "periodic", coarse_radius=4) for m in range(4)]
    with Batch(hbm_budget=4 * 10 ** 9) as b:
        lid = b.submit(long_job)
        t0 = time.monotonic()
        while b.info(lid)["status"] == "queued":
            assert time.monotonic() - t0 < 30
            time.sleep(0.0005)
        sids = [b.submit(j) for j in short]
        b.wait(timeout=300)
        linfo = b.info(lid)
        for i in sids:
            si = b.info(i)
            assert si["admit_s"] < linfo["done_s"]
            assert si["done_s"] < linfo["done_s"], "short jobs finish while the long one runs"
        got = {i: b.result(i) for i in [lid] + sids}
    for k, j in enumerate(short):
        check_job(f"short {k}", got[sids[k]], oracle_job(j))
    check_job("long", got[lid], single_lattice_job(long_job))


def test_ensemble_matches_single_lattice_path():
    """The whole 512-job ensemble workload (FHP-II, long side 128..2048 step
    64, aspect 1..4, periodic / pipe / cylinder, fill 0.1-0.5, own seeds), with
    steps cut to 40-120 so the comparison runs in seconds: every job's state
    and coarse fields equal its run alone on the single-lattice path, and the
    eight smallest the oracle's."""
    from lgca_amd.workloads import ensemble_jobs
    jobs = ensemble_jobs(512, steps_range=(40, 121))
    say(f"[batch] ensemble workload, {len(jobs)} jobs")
    got, stats = run_batch(jobs, budget=200 * 10 ** 9)
    say(f"[batch] stats {stats}")
    for i, j in enumerate(jobs):
        check_job(f"ensemble job {i} {j.dim_x}x{j.dim_y} {j.geometry}", got[i], single_lattice_job(j))
    for i in sorted(range(len(jobs)), key=lambda i: jobs[i].dim_x * jobs[i].dim_y)[:8]:
        check_job(f"ensemble job {i} vs oracle", got[i], oracle_job(jobs[i]))


def test_invalid_job_descriptions():
    from lgca_amd import Batch, Job, LgcaError
    bad_types = np.full(8 * 8, 3, dtype=np.uint8)
    bad_state = np.full(8 * 8, 0x80, dtype=np.uint8)
    cases = [
        (Job(7, 8, 8, 1), 1),
        (Job(2, 8, 7, 1), 1),
        (Job(2, 8, 8, 1, coarse_radius=5), 1),
        (Job(2, 8, 8, 1, fill=1.5), 1),
        (Job(2, 8, 8, 1, force_dir="x", force_p=2.0), 1),
        (Job(2, 8, 8, 1, geometry="types", cell_types=bad_types), 4),
        (Job(2, 8, 8, 1, state=bad_state), 4),
    ]
    with Batch(hbm_budget=10 ** 9) as b:
        for job, code in cases:
            with pytest.raises(LgcaError) as ei:
                b.submit(job)
            assert ei.value.code == code, (job, ei.value)
        jid = b.submit(Job(2, 8, 8, 3, fill=0.5))
        b.wait(jid, timeout=60)
        b.result(jid)
        b.release(jid)
        with pytest.raises(LgcaError):
            b.result(jid)
        with pytest.raises(LgcaError):
            b.release(jid)


def test_serving_stream_matches_single_lattice_path():
    """The serving workload as a Poisson stream into one batch with the serving
    cell cap (2^28 active cells): 48 jobs uniform over the four models,
    256x256..4096x1024, periodic / x-forced pipe (p 1e-4..5e-3) / cylinder of
    random radius and position, steps cut to 50..400.  Jobs arrive while
    others run, are admitted as cells free up and share launches with other
    models.  Every job's state bytes and coarse fields equal its run alone on
    the single-lattice path; the four smallest also equal the oracle's."""
    import time as _time

    from lgca_amd import Batch
    from lgca_amd.workloads import poisson_arrivals, serving_jobs
    jobs = serving_jobs(48, seed=2026, steps_range=(50, 401))
    assert {int(j.model) for j in jobs} == {0, 1, 2, 3}
    assert {j.geometry for j in jobs} == {"periodic", "pipe", "cylinder"}
    arr = poisson_arrivals(len(jobs), 400.0, seed=3)
    say(f"[serving] {len(jobs)} jobs over {arr[-1]:.3f} s")
    with Batch(hbm_budget=200 * 10 ** 9) as b:
        b.set_tuning(0, 0, 1 << 28)
        ids = []
        t0 = _time.monotonic()
        for i, j in enumerate(jobs):
            delay = t0 + arr[i] - _time.monotonic()
            if delay > 0:
                _time.sleep(delay)
            ids.append(b.submit(j))
        b.wait(timeout=300)
        got = [b.result(i) for i in ids]
        stats = b.stats()
    say(f"[serving] stats {stats}")
    assert stats["jobs_done"] == len(jobs)
    for i, job in enumerate(jobs):
        name = f"job {i} m{model_id(job.model)} {job.dim_x}x{job.dim_y} {job.geometry} {job.steps} steps"
        check_job(name, got[i], single_lattice_job(job))
    for i in sorted(range(len(jobs)), key=lambda i: jobs[i].dim_x * jobs[i].dim_y)[:4]:
        check_job(f"job {i} vs oracle", got[i], oracle_job(jobs[i]))


def test_new_job_in_a_freed_slot_gets_its_own_work_items():
    """A job that runs alone, finishes and frees its slot, then a job of
    another size admitted into that slot with the same steps per pass and
    tile height: its work items must be rebuilt for its own geometry (a
    round signature of slots and steps per pass alone once reused the first
    job's items)."""
    from lgca_amd import Batch, Job
    first = Job(2, 640, 128, 40, seed=31, fill=0.3, coarse_radius=4)
    later = [Job(0, 1280, 256, 36, seed=32, fill=0.2, geometry="pipe", coarse_radius=4),
             Job(1, 2496, 192, 44, seed=33, fill=0.25, coarse_radius=4),
             Job(3, 320, 64, 48, seed=34, fill=0.3, geometry="cylinder", cyl_x=90, cyl_y=30,
                 cyl_diameter=21.0, force_dir="x", force_p=0.01, coarse_radius=4)]
    with Batch(hbm_budget=8 * 10 ** 9) as b:
        b.set_tuning(4, 2048, 0)
        ids = [b.submit(first)]
        b.wait(ids[0], timeout=60)
        for j in later:
            ids.append(b.submit(j))
            b.wait(ids[-1], timeout=60)  # each runs alone, in the slot the last one freed
        got = [b.result(i) for i in ids]
    for i, job in enumerate([first] + later):
        say(f"[slot reuse] job {i} m{model_id(job.model)} {job.dim_x}x{job.dim_y}")
        check_job(f"job {i}", got[i], oracle_job(job))
