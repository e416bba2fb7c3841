"""Parity of the HIP engine with the CPU oracle, through the C ABI.

Bit-exact for every integer result (state bytes, particle counts, forcing
counts); FP32 coarse fields compared bit for bit as well, since the kernels
reproduce the reference's summation order (the stated tolerance, 1e-6
relative, is therefore not needed); the mean velocity is summed exactly on the
GPU and compared with the oracle's double sum to 1e-5 relative.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

Q = {0: 4, 1: 6, 2: 7, 3: 7}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def Lattice():
    from lgca_amd import Lattice as L
    return L


def make_types(rng, dx, dy, kind):
    t = np.zeros(dx * dy, dtype=np.uint8)
    g = t.reshape(dy, dx)
    if kind == "pipe":
        g[0, :] = 1
        g[-1, :] = 1
    elif kind == "slipbox":
        g[0, :] = g[-1, :] = g[:, 0] = g[:, -1] = 2
    elif kind == "mixed":
        g[rng.random((dy, dx)) < 0.06] = 1
        g[rng.random((dy, dx)) < 0.04] = 2
        g[0, :] = 2
        g[:, -1] = 2
        g[-1, : dx // 2] = 1
    elif kind == "cylinder":
        g[0, :] = g[-1, :] = 1
        ys, xs = np.mgrid[0:dy, 0:dx]
        g[(xs - dx // 3) ** 2 + (ys - dy // 2) ** 2 < (dy // 4) ** 2] = 1
    elif kind == "sparse":
        # solids in a few rows only, two of them next to the y wrap: most
        # waves of a tall lattice see none, the edge ones see them across it
        for r in (1, 70, 71, 130 % dy, dy - 2):
            g[r, dx // 5: dx // 5 + 9] = 1
        g[dy // 2, -3:] = 2
    return t


def random_state(rng, model, n):
    return rng.integers(0, 1 << Q[model], n, dtype=np.uint8)


def run_both(Lattice, model, dx, dy, types, state, seed, step0, nsteps, force=None, r=1,
             chunks=1, tuning=None):
    lat = Lattice(model, dim_x=dx, dim_y=dy, coarse_graining_radius=r, seed=seed)
    if tuning is not None:
        lat.set_tuning(*tuning)
    if types is not None:
        lat.set_cell_types(types)
    lat.set_state_u8(state)
    lat.set_step(step0)
    if force:
        lat.set_forcing(force["dir"], force["p"], force.get("x0", 0), force.get("x1"),
                        force.get("y0", 0), force.get("y1"))
    per = nsteps // chunks
    for i in range(chunks):
        lat.step(per if i < chunks - 1 else nsteps - per * (chunks - 1))
    got = lat.get_state_u8()
    want = orc.run(model, dx, dy, types, state, seed, step0, nsteps, force, nthreads=8)
    return lat, got, want


STEP_CASES = [
    # model, dx, dy, geometry, force
    (0, 2, 2, "periodic", None),
    (0, 3, 3, "periodic", None),
    (0, 33, 7, "mixed", dict(dir="x", p=0.3)),
    (0, 64, 64, "slipbox", dict(dir="y", p=0.5, x0=5, x1=40)),
    (1, 2, 2, "periodic", None),
    (1, 37, 10, "mixed", dict(dir="x", p=0.2, y0=2, y1=7)),
    (1, 130, 6, "cylinder", dict(dir="y", p=0.4)),
    (2, 5, 4, "periodic", dict(dir="x", p=1.0)),
    (2, 63, 12, "mixed", dict(dir="y", p=0.25)),
    (2, 128, 32, "pipe", dict(dir="x", p=0.05)),
    (2, 200, 14, "cylinder", dict(dir="x", p=0.1, x0=33, x1=97)),
    (3, 96, 16, "slipbox", None),
    (3, 8232, 6, "mixed", dict(dir="x", p=0.02, x0=8100)),  # two workgroups per row
    (2, 4096, 8, "periodic", dict(dir="x", p=0.3, x0=0, x1=40)),  # force on the wrapped words
    (3, 1000, 66, "cylinder", dict(dir="x", p=0.01, x0=0, x1=64)),
    # forced strips in some block columns only: those columns are dispatched
    # first (the first of two; the middle of three)
    (2, 8192, 24, "cylinder", dict(dir="x", p=0.3, x0=3000, x1=3100)),
    (3, 16384, 20, "mixed", dict(dir="y", p=0.2, x0=9000, x1=9100)),
    # tall narrow rows: many row tiles per wave (segment mode), idle lanes
    (2, 96, 300, "mixed", dict(dir="x", p=0.1, y0=50, y1=250)),
    (0, 700, 50, "cylinder", None),
    # solids in a few rows: waves with and without solid rows side by side
    # (segment mode with 8 and 2 rows per wave; strip mode, where 100-row
    # tiles test rows past the first 64 of the walk)
    (2, 256, 200, "sparse", None),
    (1, 1024, 130, "sparse", dict(dir="y", p=0.2, y0=100, y1=130)),
    (2, 4096, 200, "sparse", dict(dir="x", p=0.1, y0=60, y1=80)),
    # rows of exactly 64 words: one row per wave, neighbours by wave rotation
    (0, 2048, 10, "mixed", dict(dir="y", p=0.3)),
    (1, 2048, 12, "cylinder", dict(dir="x", p=0.2, x0=0, x1=100)),
    (3, 2048, 8, "slipbox", None),
]


TUNINGS = [(0, 0), (1, 0), (2, 0), (4, 0), (8, 0), (4, 5), (8, 3), (4, 100), (2, 2)]


@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: f"k{t[0]}th{t[1]}")
@pytest.mark.parametrize("case", STEP_CASES, ids=lambda c: f"m{c[0]}_{c[1]}x{c[2]}_{c[3]}")
def test_step_bit_exact(Lattice, case, tuning):
    model, dx, dy, geom, force = case
    rng = np.random.default_rng(dx * 131 + dy)
    types = make_types(rng, dx, dy, geom)
    state = random_state(rng, model, dx * dy)
    say(f"[step] model={model} {dx}x{dy} {geom} force={force} tuning={tuning}")
    lat, got, want = run_both(Lattice, model, dx, dy, types, state, seed=dx + 7 * dy,
                              step0=11, nsteps=23, force=force, chunks=3, tuning=tuning)
    np.testing.assert_array_equal(got, want)
    assert lat.steps == 11 + 23
    lat.close()


@pytest.mark.parametrize("dx", [62 * 32, 62 * 32 + 1, 63 * 32 - 3, 124 * 32, 4 * 62 * 32 + 77])
@pytest.mark.parametrize("k", [1, 4, 8])
def test_tile_strip_boundaries(Lattice, dx, k):
    """Widths around the 62-word stride of the tiled kernel's waves."""
    dy = 24
    rng = np.random.default_rng(dx + k)
    types = make_types(rng, dx, dy, "mixed")
    state = random_state(rng, 2, dx * dy)
    say(f"[tile] {dx}x{dy} k={k}")
    _, got, want = run_both(Lattice, 2, dx, dy, types, state, seed=5, step0=4, nsteps=2 * k + 1,
                            force=dict(dir="y", p=0.2, x0=dx // 3, x1=dx - 40), tuning=(k, 8))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("model,dx,dy,geom,force,tuning", [
    # forced strips without solids: the force-only walk, with decisions on
    # some rows only (zero words elsewhere), every model
    (0, 4096, 64, "periodic", dict(dir="x", p=0.2, x0=100, x1=300, y0=10, y1=50), (8, 16)),
    (1, 4096, 64, "periodic", dict(dir="x", p=0.2, x0=100, x1=300, y0=10, y1=50), (8, 16)),
    (2, 4096, 64, "periodic", dict(dir="x", p=0.2, x0=100, x1=300, y0=10, y1=50), (4, 16)),
    (3, 4096, 64, "periodic", dict(dir="x", p=0.2, x0=4000, x1=4096, y0=10, y1=50), (8, 16)),
    # a short last word keeps the general walk
    (2, 4096 + 77, 64, "periodic", dict(dir="x", p=0.2, x0=100, x1=300, y0=10, y1=50), (8, 16)),
    # walls and a cylinder: walks over solids in quarters (40 rows), halves
    # (12 rows) and whole (6 rows), beside forced and plain walks
    (3, 4096, 96, "cylinder", dict(dir="x", p=0.05, x0=0, x1=64), (8, 40)),
    (1, 4096, 96, "cylinder", dict(dir="y", p=0.1), (4, 12)),
    (2, 4096 + 5, 96, "cylinder", dict(dir="x", p=0.05, x0=0, x1=64), (8, 6)),
    # forced everywhere: every strip forced, walls in the first and last tile
    (2, 8192, 150, "pipe", dict(dir="x", p=0.01), (8, 74)),
])
def test_force_only_walks_and_solid_parts(Lattice, model, dx, dy, geom, force, tuning):
    """Strip-mode passes whose waves take the force-only walk (forced along x,
    no solids) or walk a quarter or half of a tile over solids: bits equal the
    oracle's (src/omp_lattice.cpp:100-249, 254-346) whatever walk or part a
    wave takes."""
    rng = np.random.default_rng(dx * 7 + dy + model)
    types = make_types(rng, dx, dy, geom)
    state = random_state(rng, model, dx * dy)
    say(f"[walks] model={model} {dx}x{dy} {geom} force={force} tuning={tuning}")
    lat, got, want = run_both(Lattice, model, dx, dy, types, state, seed=dx + dy, step0=6,
                              nsteps=29, force=force, chunks=2, tuning=tuning)
    np.testing.assert_array_equal(got, want)
    lat.close()


def test_single_steps_match_batched_steps(Lattice):
    rng = np.random.default_rng(5)
    dx, dy = 160, 20
    types = make_types(rng, dx, dy, "mixed")
    state = random_state(rng, 2, dx * dy)
    say("[step] chunked vs single-call forcing windows")
    force = dict(dir="x", p=0.3, x0=10, x1=150)
    _, a, want = run_both(Lattice, 2, dx, dy, types, state, 9, 0, 150, force, chunks=150)
    _, b, _ = run_both(Lattice, 2, dx, dy, types, state, 9, 0, 150, force, chunks=1)
    np.testing.assert_array_equal(a, want)
    np.testing.assert_array_equal(b, want)


def test_config_hpp_256_periodic_1000_steps(Lattice):
    """HPP 256x256 torus, each bit Bernoulli(0.2) from seed 1, 1000 steps,
    r = 8 (16x16 coarse blocks)."""
    say("[config] HPP 256x256 x 1000")
    dx = dy = 256
    state = orc.init_random(0, dx, dy, None, 1, 0.2)
    lat = Lattice(0, dim_x=dx, dim_y=dy, coarse_graining_radius=8, seed=1)
    lat.init_random(0.2)
    np.testing.assert_array_equal(lat.get_state_u8(), state)
    lat.step(1000)
    want = orc.run(0, dx, dy, None, state, 1, 0, 1000, nthreads=8)
    got = lat.get_state_u8()
    np.testing.assert_array_equal(got, want)
    assert lat.get_n_particles() == orc.count_particles(state)
    lat.copy_data_to_output_buffer()
    lat.post_process()
    cd, cm, md, mm = orc.post_process(0, dx, dy, 8, want)
    np.testing.assert_array_equal(lat.cell_density().view(np.uint32), cd.view(np.uint32))
    np.testing.assert_array_equal(lat.cell_momentum().view(np.uint32), cm.view(np.uint32))
    np.testing.assert_array_equal(lat.mean_density().view(np.uint32), md.view(np.uint32))
    np.testing.assert_array_equal(lat.mean_momentum().view(np.uint32), mm.view(np.uint32))
    lat.close()


def test_config_fhp2_pipe_20000_steps(Lattice):
    """FHP-II 2048x512 pipe, walls top and bottom, fill 0.3, per-cell x force
    p = 1e-3, seed 7, r = 16 (32x32 blocks), all 20000 steps through the
    default launch path (trial of launch shapes included)."""
    say("[config] FHP-II 2048x512 pipe x 20000")
    dx, dy, steps = 2048, 512, 20000
    lat = Lattice(2, dim_x=dx, dim_y=dy, coarse_graining_radius=16, seed=7)
    lat.apply_bc_pipe()
    lat.init_random(0.3)
    types = lat.m_cell_type_cpu.copy()
    state = orc.init_random(2, dx, dy, types, 7, 0.3)
    np.testing.assert_array_equal(lat.get_state_u8(), state)
    n0 = lat.get_n_particles()
    lat.set_forcing("x", 1e-3)
    # in chunks, as a viewer steps: the trial finishes early and later chunks
    # run the kept shape
    for n in (300, 700, 9000, 10000):
        lat.step(n)
    want = orc.run(2, dx, dy, types, state, 7, 0, steps, dict(dir="x", p=1e-3), nthreads=16)
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    assert lat.steps == steps
    assert lat.get_n_particles() == n0
    lat.copy_data_to_output_buffer()
    lat.post_process(cells=False)
    _, _, md, mm = orc.post_process(2, dx, dy, 16, want)
    np.testing.assert_array_equal(lat.mean_density().view(np.uint32), md.view(np.uint32))
    np.testing.assert_array_equal(lat.mean_momentum().view(np.uint32), mm.view(np.uint32))
    info = lat.info()
    say(f"[config] kept steps per pass {info.steps_per_pass}, tile rows {info.tile_rows}")
    lat.close()


@pytest.mark.parametrize("model", [0, 1, 2, 3])
@pytest.mark.parametrize("dims", [(64, 32, 4), (48, 24, 12), (24, 8, 4), (8, 8, 4), (70, 20, 5),
                                  (52, 22, 5)])
def test_post_process_bit_exact(Lattice, model, dims):
    dx, dy, r = dims
    rng = np.random.default_rng(dx + dy + model)
    state = random_state(rng, model, dx * dy)
    say(f"[post] model={model} {dx}x{dy} r={r}")
    lat = Lattice(model, dim_x=dx, dim_y=dy, coarse_graining_radius=r)
    lat.set_state_u8(state)
    lat.post_process()  # before any snapshot: the snapshot is empty
    assert not lat.mean_density().any()
    lat.copy_data_to_output_buffer()
    lat.post_process()
    cd, cm, md, mm = orc.post_process(model, dx, dy, r, state)
    for got, want in ((lat.cell_density(), cd), (lat.cell_momentum(), cm),
                      (lat.mean_density(), md), (lat.mean_momentum(), mm)):
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    from lgca_amd import hiprt
    _, _, vptr = lat.post_process_device()
    lat.synchronize()
    v = np.empty(2 * md.size, dtype=np.float32)
    hiprt.memcpy_d2h(v.ctypes.data, vptr, v.nbytes)
    np.testing.assert_array_equal(v.view(np.uint32),
                                  orc.coarse_velocity(md, mm).view(np.uint32))
    lat.close()


@pytest.mark.parametrize("model", [0, 1, 2])
def test_count_and_mean_velocity(Lattice, model):
    dx, dy = 96, 40
    rng = np.random.default_rng(model)
    types = make_types(rng, dx, dy, "mixed")
    state = random_state(rng, model, dx * dy)
    say(f"[reduce] model={model}")
    lat = Lattice(model, dim_x=dx, dim_y=dy, coarse_graining_radius=2)
    lat.set_cell_types(types)
    lat.set_state_u8(state)
    assert lat.get_n_particles() == orc.count_particles(state)
    lat.copy_data_to_output_buffer()
    got = lat.get_mean_velocity()
    cd, cm = orc.cell_post_process(model, dx, dy, state)
    want = orc.mean_velocity(dx, dy, types, cd, cm)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7)
    lat.close()


@pytest.mark.parametrize("model,p", [(0, 0.2), (1, 0.35), (2, 0.5), (3, 0.25)])
def test_init_random_matches_oracle(Lattice, model, p):
    dx, dy = 100, 30
    rng = np.random.default_rng(3)
    types = make_types(rng, dx, dy, "mixed")
    pre = random_state(rng, model, dx * dy)
    say(f"[init] model={model} p={p}")
    lat = Lattice(model, dim_x=dx, dim_y=dy, coarse_graining_radius=1, seed=77)
    lat.set_cell_types(types)
    lat.set_state_u8(pre)
    lat.init_random(p)
    want = orc.init_random(model, dx, dy, types, 77, p, state=pre.copy())
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    # restricted to the reference's diffusion disc
    lat.set_state_u8(pre)
    lat.init_diffusion(p)
    from lgca_amd.lattice import disc
    region = disc(dx, dy, dx // 2, dy // 2, float(dy // 4))
    eff = types.copy()
    eff[~region] = 1
    want = orc.init_random(model, dx, dy, eff, 77, p, state=pre.copy())
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    lat.close()


@pytest.mark.parametrize("model,direction", [(0, "x"), (0, "y"), (1, "y"), (2, "x"), (3, "y")])
@pytest.mark.parametrize("forcing", [0, 1, 37, 10**9, -1, -(2**31)])
def test_count_mode_force(Lattice, model, direction, forcing):
    dx, dy = 48, 16
    rng = np.random.default_rng(forcing % 1000 + model)
    types = make_types(rng, dx, dy, "mixed")
    state = random_state(rng, model, dx * dy)
    say(f"[force-count] model={model} dir={direction} forcing={forcing}")
    lat = Lattice(model, dim_x=dx, dim_y=dy, coarse_graining_radius=1, seed=123)
    lat.set_cell_types(types)
    lat.set_state_u8(state)
    want_state = state.copy()
    got_n = lat.apply_body_force_count(direction, forcing)
    want_n = orc.body_force_count(model, dx, dy, types, want_state, 123, 0, direction, forcing)
    assert got_n == want_n
    np.testing.assert_array_equal(lat.get_state_u8(), want_state)
    # the second call draws a fresh candidate sequence
    got_n = lat.apply_body_force_count(direction, forcing)
    want_n = orc.body_force_count(model, dx, dy, types, want_state, 123, 1, direction, forcing)
    assert got_n == want_n
    np.testing.assert_array_equal(lat.get_state_u8(), want_state)
    lat.close()


def test_invalid_inputs_are_rejected(Lattice):
    from lgca_amd import LgcaError
    say("[errors]")
    with pytest.raises(LgcaError):
        Lattice(2, dim_x=16, dim_y=7, coarse_graining_radius=1)  # odd dim_y for FHP
    with pytest.raises(LgcaError):
        Lattice(0, dim_x=4, dim_y=4, coarse_graining_radius=3)   # smaller than 2r
    lat = Lattice(1, dim_x=40, dim_y=6, coarse_graining_radius=1)
    bad = np.zeros(240, dtype=np.uint8)
    bad[17] = 1 << 6                                            # FHP-I has 6 directions
    with pytest.raises(LgcaError):
        lat.set_state_u8(bad)
    assert lat.get_n_particles() == 0
    t = np.zeros(240, dtype=np.uint8)
    t[3] = 5
    with pytest.raises(LgcaError):
        lat.set_cell_types(t)
    # apply_body_force takes an int: amounts beyond its range are refused
    for amount in (2 ** 31, -(2 ** 31) - 1):
        with pytest.raises(LgcaError):
            lat.apply_body_force_count("x", amount)
    lat.close()


def test_codec_round_trip_ragged(Lattice):
    for model, dx, dy in [(0, 2, 3), (1, 31, 2), (2, 33, 4), (3, 129, 2), (2, 4099, 2)]:
        say(f"[codec] {model} {dx}x{dy}")
        rng = np.random.default_rng(dx)
        st = random_state(rng, model, dx * dy)
        t = rng.integers(0, 3, dx * dy, dtype=np.uint8)
        lat = Lattice(model, dim_x=dx, dim_y=dy, coarse_graining_radius=1)
        lat.set_state_u8(st)
        np.testing.assert_array_equal(lat.get_state_u8(), st)
        lat.set_cell_types(t)
        np.testing.assert_array_equal(lat.get_cell_types(), t)
        lat.close()


def test_full_karman_size_through_the_default_path(Lattice):
    """FHP-III 32768x8192 with the cylinder of diameter 512 at (4096, 4099),
    walls, fill 0.25, x force p = 2e-3 on x < 64, seed 11, r = 32: 256 steps
    through the default launch path (the trial's candidate shapes, two
    passes each: 4-step passes of 7 heights, 8-step passes of 6 fixed heights
    and of up to 4 sized to the GPU's wave slots, 216 steps at most; then the
    kept shape) equal the oracle's bits, and so do the coarse fields of the
    final state."""
    dx, dy, steps = 32768, 8192, 256
    say("[full] karman 32768x8192")
    lat = Lattice(3, dim_x=dx, dim_y=dy, coarse_graining_radius=32, seed=11)
    lat.apply_bc_cylinder(4096, 4099, 512.0)
    lat.init_random(0.25)
    lat.set_forcing("x", 2e-3, 0, 64)
    types = lat.m_cell_type_cpu.copy()
    state = lat.get_state_u8()
    want0 = orc.init_random(3, dx, dy, types, 11, 0.25)
    np.testing.assert_array_equal(state, want0)
    del want0
    n0 = lat.get_n_particles()
    lat.step(216)
    lat.synchronize()
    lat.step(40)  # decides on the trial, then runs the kept shape
    info = lat.info()
    say(f"[full] kept steps per pass {info.steps_per_pass}, tile rows {info.tile_rows}")
    assert info.steps_per_pass in (4, 8) and info.tile_rows > 0
    got = lat.get_state_u8()
    assert lat.get_n_particles() == n0
    want = orc.fast_run(3, dx, dy, types, state, 11, 0, steps, dict(dir="x", p=2e-3, x0=0, x1=64),
                        nthreads=16)
    del state
    np.testing.assert_array_equal(got, want)
    del got
    lat.copy_data_to_output_buffer()
    lat.post_process(cells=False)
    _, _, md, mm = orc.post_process(3, dx, dy, 32, want)
    np.testing.assert_array_equal(lat.mean_density().view(np.uint32), md.view(np.uint32))
    np.testing.assert_array_equal(lat.mean_momentum().view(np.uint32), mm.view(np.uint32))
    lat.close()


@pytest.mark.parametrize("k", [2, 4, 8])
@pytest.mark.parametrize("model,dx,dy,geom", [(0, 700, 40, "cylinder"), (1, 96, 30, "mixed"),
                                              (2, 3000, 26, "mixed"), (3, 2048, 20, "cylinder")])
def test_passes_start_at_every_step_offset(Lattice, model, dx, dy, geom, k):
    """K-step passes started at steps 4m + o for o = 0..3 (the chirality group
    offset the kernels are compiled for), each run as one launch of k steps,
    with forcing on part of the lattice: every offset gives the oracle's bits."""
    rng = np.random.default_rng(dx + 13 * k + model)
    types = make_types(rng, dx, dy, geom)
    state = random_state(rng, model, dx * dy)
    force = dict(dir="x", p=0.2, x0=dx // 4, x1=dx // 2)
    for off in range(4):
        step0 = 40 + off
        say(f"[offset] model={model} {dx}x{dy} k={k} t0={step0}")
        _, got, want = run_both(Lattice, model, dx, dy, types, state, seed=99 + off, step0=step0,
                                nsteps=k, force=force, tuning=(k, 4))
        np.testing.assert_array_equal(got, want)


def test_tune_keeps_state_and_step(Lattice):
    """lgca_tune times candidate passes without changing the state or the
    step, reports a candidate shape, and later steps give the oracle's bits."""
    dx, dy, seed = 4096, 512, 21
    rng = np.random.default_rng(seed)
    types = make_types(rng, dx, dy, "cylinder")
    state = random_state(rng, 2, dx * dy)
    force = dict(dir="x", p=0.01, x0=0, x1=128)
    lat = Lattice(2, dim_x=dx, dim_y=dy, coarse_graining_radius=1, seed=seed)
    lat.set_cell_types(types)
    lat.set_state_u8(state)
    lat.set_step(5)
    lat.set_forcing("x", 0.01, 0, 128)
    say("[tune] 4096x512 FHP-II, cylinder, forced strip")
    lat.tune()
    info = lat.info()
    say(f"[tune] kept steps per pass {info.steps_per_pass}, tile rows {info.tile_rows}")
    assert (info.steps_per_pass in (2, 4) and info.tile_rows in {2, 4, 8, 12, 16, 20, 24, 32}) or \
        (info.steps_per_pass, info.tile_rows) == (8, 2)
    assert lat.steps == 5
    np.testing.assert_array_equal(lat.get_state_u8(), state)
    lat.step(37)
    want = orc.run(2, dx, dy, types, state, seed, 5, 37, force, nthreads=8)
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    assert (lat.info().steps_per_pass, lat.info().tile_rows) == (info.steps_per_pass,
                                                                  info.tile_rows)
    # a fixed height leaves nothing to tune
    lat.set_tuning(4, 10)
    lat.tune()
    assert lat.info().tile_rows == 10
    lat.close()


@pytest.mark.slow
@pytest.mark.timeout(1200)
def test_full_karman_5000_steps_bit_exact(Lattice):
    """All 5000 steps of the FHP-III Karman workload at its full size
    (32768x8192, cylinder d = 512 at (4096, 4099), walls, fill 0.25, x force
    p = 2e-3 on x < 64, seed 11, r = 32), through the tuned default launch
    shape, against the bit-sliced CPU checker (oracle/cpu_fast.c, itself pinned
    to oracle/cpu_ref.c by test_oracle.py): final state and coarse fields."""
    import os
    dx, dy, steps = 32768, 8192, 5000
    threads = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    say(f"[full 5000] karman 32768x8192, checker on {threads} threads")
    lat = Lattice(3, dim_x=dx, dim_y=dy, coarse_graining_radius=32, seed=11)
    lat.apply_bc_cylinder(4096, 4099, 512.0)
    lat.init_random(0.25)
    lat.set_forcing("x", 2e-3, 0, 64)
    types = lat.m_cell_type_cpu.copy()
    state = lat.get_state_u8()
    lat.tune()
    lat.step(steps)
    info = lat.info()
    say(f"[full 5000] GPU done: steps per pass {info.steps_per_pass}, tile rows {info.tile_rows}")
    got = lat.get_state_u8()
    lat.copy_data_to_output_buffer()
    lat.post_process(cells=False)
    md_gpu, mm_gpu = lat.mean_density().copy(), lat.mean_momentum().copy()
    lat.close()
    want = orc.fast_run(3, dx, dy, types, state, 11, 0, steps, dict(dir="x", p=2e-3, x0=0, x1=64),
                        nthreads=threads)
    del state
    say("[full 5000] checker done")
    np.testing.assert_array_equal(got, want)
    del got
    _, _, md, mm = orc.post_process(3, dx, dy, 32, want)
    np.testing.assert_array_equal(md_gpu.view(np.uint32), md.view(np.uint32))
    np.testing.assert_array_equal(mm_gpu.view(np.uint32), mm.view(np.uint32))
