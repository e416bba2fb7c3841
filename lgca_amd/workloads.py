"""Job streams of the batched ensemble and the serving workload, shared by
bench.py and the tests.

ensemble_jobs: FHP-II lattices with long side L in {128..2048} step 64,
aspect 1..4 (short side rounded to 64, at least 64), either orientation,
periodic / pipe / cylinder (radius 8..L/8 at (dim_x/4, dim_y/2)), fill
U(0.1, 0.5), own seed, 1000-10000 steps, r = 16.

serving_jobs: model uniform over HPP, FHP-I, FHP-II, FHP-III; dim_x in
256..4096 and dim_y in 256..1024 (multiples of 64); periodic / pipe with x
force p U(1e-4, 5e-3) / cylinder (walls plus a disc of random radius
8..dim_y/8 at a random position); fill 0.25; 500-20000 steps; own seed;
r = 16.

`steps_range` replaces the step interval [lo, hi) without changing any other
draw, so a test can run the same jobs for fewer steps.
"""
from __future__ import annotations

import numpy as np

from .batch import Job


def ensemble_jobs(n: int, seed: int = 2024, steps_range=(1000, 10001)):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        L = int(rng.integers(2, 33)) * 64
        S = max(64, int(round(L / float(rng.uniform(1.0, 4.0)) / 64.0)) * 64)
        dx, dy = (L, S) if rng.random() < 0.5 else (S, L)
        geom = ("periodic", "pipe", "cylinder")[int(rng.integers(0, 3))]
        rad = int(rng.integers(8, L // 8 + 1))
        jobs.append(Job("FHP_II", dx, dy, int(rng.integers(*steps_range)), seed=100000 + i,
                        coarse_radius=16, fill=float(rng.uniform(0.1, 0.5)), geometry=geom,
                        cyl_x=dx // 4, cyl_y=dy // 2, cyl_diameter=2.0 * rad))
    return jobs


def serving_jobs(n: int, seed: int = 2025, steps_range=(500, 20001)):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        model = int(rng.integers(0, 4))
        dx, dy = int(rng.integers(4, 65)) * 64, int(rng.integers(4, 17)) * 64
        geom = ("periodic", "pipe", "cylinder")[int(rng.integers(0, 3))]
        j = Job(model, dx, dy, int(rng.integers(*steps_range)), seed=200000 + i, coarse_radius=16,
                fill=0.25, geometry=geom)
        if geom == "pipe":
            j.force_dir, j.force_p = "x", float(rng.uniform(1e-4, 5e-3))
        elif geom == "cylinder":
            rad = int(rng.integers(8, dy // 8 + 1))
            j.cyl_diameter = 2.0 * rad
            j.cyl_x = int(rng.integers(rad, dx - rad))
            j.cyl_y = int(rng.integers(rad, dy - rad))
        jobs.append(j)
    return jobs


def poisson_arrivals(n: int, rate: float, seed: int = 7) -> np.ndarray:
    """Arrival times (seconds from the first) of a Poisson stream."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0, n))
    return (t - t[0]) / rate
