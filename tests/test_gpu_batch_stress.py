"""Batch engine edge cases: thousands of tiny jobs (the job table grows past
its first 1024 slots and slots are reused), submissions from several threads
at once, a batch destroyed with work still queued, and the tallest lattice
the engine takes (dim_y = 65534), each job checked against the oracle or the
single-lattice path."""
import threading
import time

import numpy as np
import pytest

from oracle import oracle as orc
from test_gpu_batch import check_job, oracle_job, say, single_lattice_job

pytestmark = pytest.mark.gpu


def tiny_jobs(n, seed):
    from lgca_amd import Job
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        m = int(rng.integers(0, 4))
        dx = int(rng.integers(2, 41))
        dy = int(rng.integers(1, 11)) * 2
        geom = ("periodic", "pipe", "cylinder")[i % 3]
        j = Job(m, dx, dy, int(rng.integers(0, 21)), seed=int(rng.integers(0, 2 ** 63)),
                fill=float(rng.uniform(0.0, 1.0)), geometry=geom, cyl_x=dx // 2, cyl_y=dy // 2,
                cyl_diameter=float(rng.uniform(0.0, 6.0)))
        if i % 4 == 0:
            j.force_dir, j.force_p = ("x", "y")[i % 8 // 4], float(rng.uniform(0.0, 0.3))
        jobs.append(j)
    return jobs


def test_thousands_of_tiny_jobs_with_slot_reuse():
    from lgca_amd import Batch
    jobs = tiny_jobs(2500, 1)
    say(f"[stress] {len(jobs)} tiny jobs")
    results = {}
    with Batch(hbm_budget=2 * 10 ** 9) as b:
        ids = {}
        for k, j in enumerate(jobs):
            ids[b.submit(j)] = k
            if k % 500 == 499:  # drain some, so later jobs reuse released slots
                b.wait(timeout=120)
                for i in b.poll():
                    results[ids[i]] = b.result(i)
                    b.release(i)
        b.wait(timeout=300)
        for i in b.poll():
            results[ids[i]] = b.result(i)
            b.release(i)
        s = b.stats()
    assert len(results) == len(jobs) and s["jobs_done"] == len(jobs)
    for k in range(0, len(jobs), 7):
        check_job(f"tiny {k}", results[k], oracle_job(jobs[k]))


def test_concurrent_submitters():
    from lgca_amd import Batch
    jobs = tiny_jobs(400, 2)
    out = {}
    errors = []
    with Batch(hbm_budget=2 * 10 ** 9) as b:
        def worker(part):
            try:
                for k in part:
                    out[k] = b.submit(jobs[k])
            except Exception as e:  # reported below
                errors.append(e)
        th = [threading.Thread(target=worker, args=(range(t, len(jobs), 4),)) for t in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors
        b.wait(timeout=300)
        got = {k: b.result(i) for k, i in out.items()}
    for k in range(0, len(jobs), 5):
        check_job(f"threaded {k}", got[k], oracle_job(jobs[k]))


def test_destroy_with_work_queued():
    """Closing a batch with running and queued jobs returns promptly and leaves
    the GPU usable (a fresh batch runs a job correctly afterwards)."""
    from lgca_amd import Batch, Job
    b = Batch(hbm_budget=64 * 10 ** 6)
    for i in range(40):
        b.submit(Job(2, 512, 256, 20000, seed=i, fill=0.3, geometry="pipe"))
    time.sleep(0.05)
    t0 = time.monotonic()
    b.close()
    assert time.monotonic() - t0 < 60
    job = Job(3, 64, 32, 13, seed=5, fill=0.4, geometry="cylinder", cyl_x=20, cyl_y=16,
              cyl_diameter=9.0, force_dir="x", force_p=0.05)
    with Batch(hbm_budget=10 ** 8) as b2:
        i = b2.submit(job)
        b2.wait(i, timeout=60)
        got = b2.result(i)
    check_job("after destroy", got, oracle_job(job))


def test_tallest_lattice():
    """dim_y = 65534 (the engine's limit is 65535 rows; FHP needs it even):
    batch and single-lattice paths against the oracle."""
    from lgca_amd import Job
    from test_gpu_batch import run_batch
    job = Job(2, 70, 65534, 9, seed=77, fill=0.3, geometry="cylinder", cyl_x=35, cyl_y=30000,
              cyl_diameter=40.0, force_dir="y", force_p=0.01, force_region=(0, 70, 100, 65000),
              coarse_radius=35)
    got, _ = run_batch([job])
    want = oracle_job(job)
    check_job("tall batch", got[0], want)
    check_job("tall single", single_lattice_job(job), want)
