"""The timed trial of launch shapes (lgca_set_tuning with automatic rows, and
automatic steps per pass by default): it settles on one of its candidates,
never changes results, and restarts when the workload changes."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

# the fixed rule's own height (2 rows on this lattice) and the trial's others
CANDIDATES = {2, 4, 8, 12, 16, 20, 24, 32}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_lattice(dx, dy, seed):
    from lgca_amd import Lattice
    rng = np.random.default_rng(seed)
    types = np.zeros((dy, dx), dtype=np.uint8)
    types[0, :] = types[-1, :] = 1
    ys, xs = np.mgrid[0:dy, 0:dx]
    types[(xs - dx // 3) ** 2 + (ys - dy // 2) ** 2 < (dy // 5) ** 2] = 1
    state = rng.integers(0, 128, dx * dy, dtype=np.uint8)
    lat = Lattice(2, dim_x=dx, dim_y=dy, coarse_graining_radius=1, seed=seed)
    lat.set_cell_types(types.reshape(-1))
    lat.set_state_u8(state)
    lat.set_forcing("x", 0.02, 0, 200)
    return lat, types.reshape(-1), state


def test_trial_picks_a_candidate_and_keeps_results():
    dx, dy, seed = 4096, 1024, 77
    say(f"[trial] {dx}x{dy} FHP-II, cylinder, forced strip")
    lat, types, state = make_lattice(dx, dy, seed)
    ref, _, _ = make_lattice(dx, dy, seed)
    ref.set_tuning(4, 20)
    assert lat.info().tile_rows == 0
    assert lat.info().steps_per_pass == -1
    # 8 candidates x 2 groups x 4 launches x 4 steps, 2 x 2 x 4 x 2 steps at
    # K = 2 and 2 x 4 x 8 at K = 8 (this lattice's rule gives 2-row tiles):
    # 352 steps of trial, in uneven calls
    for n in (5, 3, 17, 9, 30, 1, 22, 13, 7, 11, 101, 41, 161):
        lat.step(n)
        ref.step(n)
    lat.synchronize()
    lat.step(4)  # decides now that the trial's launches have finished
    ref.step(4)
    rows, k = lat.info().tile_rows, lat.info().steps_per_pass
    say(f"[trial] chosen tile rows {rows}, steps per pass {k}")
    assert (k == 4 and rows in CANDIDATES) or (k == 2 and rows in {2, 4}) or (k, rows) == (8, 2)
    np.testing.assert_array_equal(lat.get_state_u8(), ref.get_state_u8())
    assert lat.steps == ref.steps == 425
    # and both agree with the oracle
    want = orc.run(2, dx, dy, types, state, seed, 0, 425,
                   dict(dir="x", p=0.02, x0=0, x1=200), nthreads=8)
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    # a new forcing restarts the trial; an explicit height is reported as set
    lat.set_forcing("x", 0.01, 0, 100)
    assert lat.info().tile_rows == 0
    lat.set_tuning(4, 18)
    assert lat.info().tile_rows == 18
    lat.close()
    ref.close()


def test_trial_on_a_lattice_shorter_than_its_candidates():
    """A 14-row lattice: candidates taller than the lattice are left out."""
    from lgca_amd import Lattice
    lat = Lattice(2, dim_x=1024, dim_y=14, coarse_graining_radius=1, seed=3)
    lat.apply_bc_periodic()
    lat.init_random(0.3)
    lat.step(240)  # the trial: 224 steps
    lat.synchronize()
    lat.step(4)
    assert lat.info().tile_rows in {2, 4, 8, 12}
    lat.close()



def test_automatic_steps_per_pass_on_a_small_lattice():
    """The 2048 x 512 pipe (forced, walls): the default shape trial also times
    2-step passes and 8-step passes of 2-row tiles (352 steps); whatever it
    keeps, results equal a fixed shape's and the oracle's, and an explicit
    steps_per_pass is kept."""
    from lgca_amd import Lattice
    dx, dy, seed = 2048, 512, 7

    def pipe():
        lat = Lattice(2, dim_x=dx, dim_y=dy, coarse_graining_radius=1, seed=seed)
        lat.apply_bc_pipe()
        lat.init_random(0.3)
        lat.set_forcing("x", 1e-3)
        return lat

    lat, ref = pipe(), pipe()
    ref.set_tuning(4, 8)
    types, state = lat.m_cell_type_cpu.copy(), lat.get_state_u8()
    np.testing.assert_array_equal(state, ref.get_state_u8())
    say(f"[trial] {dx}x{dy} FHP-II pipe, automatic steps per pass")
    for n in (3, 150, 1, 29, 200):
        lat.step(n)
        ref.step(n)
    lat.synchronize()
    lat.step(6)
    ref.step(6)
    info = lat.info()
    say(f"[trial] chosen steps per pass {info.steps_per_pass}, tile rows {info.tile_rows}")
    assert (info.steps_per_pass, info.tile_rows) in {(4, r) for r in CANDIDATES} | {(2, 2), (2, 4), (8, 2)}
    got = lat.get_state_u8()
    np.testing.assert_array_equal(got, ref.get_state_u8())
    want = orc.run(2, dx, dy, types, state, seed, 0, 389, dict(dir="x", p=1e-3, x0=0, x1=dx), nthreads=8)
    np.testing.assert_array_equal(got, want)
    # an explicit steps per pass keeps the trial to heights at that K
    lat.set_tuning(4, 0)
    assert lat.info().steps_per_pass == 4 and lat.info().tile_rows == 0
    lat.step(300)
    lat.synchronize()
    lat.step(4)
    assert lat.info().steps_per_pass == 4 and lat.info().tile_rows in CANDIDATES
    lat.close()
    ref.close()
