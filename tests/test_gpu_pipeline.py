"""Launches of several passes (lgca_set_pipeline), switched off in this build
(see the skip reason below): each XCD runs a band of row
tiles pass after pass, every tile waiting only for its neighbours' previous
pass.  The bits must equal the oracle's (src/omp_lattice.cpp:100-249) and the
pass-by-pass path's, for every model, geometry, forcing and launch span."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SWITCHED_OFF = pytest.mark.skip(
    reason="multi-pass launches are switched off: on the GPU the first case (4096x1024 periodic, "
           "8 steps x 16 rows) ran past its 100 s limit in two runs, cause not found; "
           "lgca_set_pipeline(1) refuses")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def Lattice():
    from lgca_amd import Lattice as L
    return L


def make_types(rng, dx, dy, geom):
    t = np.zeros((dy, dx), dtype=np.uint8)
    if geom in ("cylinder", "mixed"):
        t[0, :] = t[-1, :] = 1
        ys, xs = np.mgrid[0:dy, 0:dx]
        t[(xs - dx // 3) ** 2 + (ys - dy // 2) ** 2 < (dy // 6) ** 2] = 1
    if geom == "mixed":
        t[rng.random((dy, dx)) < 0.01] = 2
    return t.reshape(-1)


@SWITCHED_OFF
@pytest.mark.parametrize("model,dx,dy,geom,k,th,steps,force", [
    (2, 4096, 1024, "periodic", 8, 16, 67, None),
    (3, 4096, 1024, "cylinder", 8, 24, 96, dict(dir="x", p=0.01, x0=0, x1=300)),
    (1, 2048, 512, "mixed", 4, 8, 41, dict(dir="y", p=0.02, x0=100, x1=900, y0=50, y1=400)),
    (0, 4096, 512, "cylinder", 8, 16, 80, dict(dir="x", p=0.02)),
    (2, 3968, 640, "mixed", 4, 12, 53, dict(dir="x", p=0.005, x0=0, x1=3968)),
])
def test_pipelined_passes_match_the_oracle(Lattice, model, dx, dy, geom, k, th, steps, force):
    rng = np.random.default_rng(dx + dy + model)
    types = make_types(rng, dx, dy, geom)
    q = orc.NUM_DIR[model]
    state = rng.integers(0, 1 << q, dx * dy, dtype=np.uint8)
    state[types != 0] = 0
    lat = Lattice(model, dim_x=dx, dim_y=dy, coarse_graining_radius=1, seed=41)
    lat.set_cell_types(types)
    lat.set_state_u8(state)
    lat.set_step(3)
    if force:
        lat.set_forcing(force["dir"], force["p"], force.get("x0", 0), force.get("x1", dx),
                        force.get("y0", 0), force.get("y1", dy))
    lat.set_tuning(k, th)
    lat.set_pipeline(True)
    assert lat.info().pass_pipeline == 1
    say(f"[pipe] model {model} {dx}x{dy} {geom} k {k} th {th} steps {steps}")
    # uneven calls: launches of several passes, single passes and short ones
    for n in (steps // 2, 1, steps - steps // 2 - 1):
        lat.step(n)
    lat.synchronize()
    got = lat.get_state_u8()
    want = orc.run(model, dx, dy, types, state, 41, 3, steps, force, nthreads=8)
    np.testing.assert_array_equal(got, want)
    assert lat.steps == 3 + steps
    lat.close()


@SWITCHED_OFF
def test_pipelined_karman_full_size(Lattice):
    """The FHP-III Karman workload at full size (32768x8192, cylinder d = 512 at
    (4096, 4099), walls, fill 0.25, x force p = 2e-3 on x < 64, seed 11): 256
    steps of 8-step passes of 64 rows, launched several passes at a time,
    against the bit-sliced checker."""
    dx, dy, steps = 32768, 8192, 256
    lat = Lattice(3, dim_x=dx, dim_y=dy, coarse_graining_radius=32, seed=11)
    lat.apply_bc_cylinder(4096, 4099, 512.0)
    lat.init_random(0.25)
    lat.set_forcing("x", 2e-3, 0, 64)
    types = lat.m_cell_type_cpu.copy()
    state = lat.get_state_u8()
    lat.set_tuning(8, 64)
    lat.set_pipeline(True)
    say("[pipe] karman 32768x8192, 256 steps")
    lat.step(steps)
    lat.synchronize()
    got = lat.get_state_u8()
    lat.close()
    want = orc.fast_run(3, dx, dy, types, state, 11, 0, steps, dict(dir="x", p=2e-3, x0=0, x1=64),
                        nthreads=16)
    np.testing.assert_array_equal(got, want)


@SWITCHED_OFF
def test_pipeline_on_a_lattice_too_small_for_bands(Lattice):
    """Too few row tiles for one band per XCD: steps run pass by pass."""
    dx, dy, steps = 1024, 32, 40
    rng = np.random.default_rng(5)
    state = rng.integers(0, 128, dx * dy, dtype=np.uint8)
    lat = Lattice(2, dim_x=dx, dim_y=dy, coarse_graining_radius=1, seed=5)
    lat.set_state_u8(state)
    lat.set_tuning(8, 8)
    lat.set_pipeline(True)
    lat.step(steps)
    lat.synchronize()
    want = orc.run(2, dx, dy, None, state, 5, 0, steps, None, nthreads=8)
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    lat.close()


def test_set_pipeline_refuses_while_switched_off(Lattice):
    """The mode is switched off: turning it on fails with a message, turning it
    off is accepted, and steps run pass by pass as before."""
    from lgca_amd._capi import LgcaError
    lat = Lattice(2, dim_x=256, dim_y=64, coarse_graining_radius=1, seed=3)
    lat.apply_bc_periodic()
    lat.init_random(0.3)
    with pytest.raises(LgcaError, match="switched off"):
        lat.set_pipeline(True)
    lat.set_pipeline(False)
    assert lat.info().pass_pipeline == 0
    state = lat.get_state_u8()
    lat.step(9)
    want = orc.run(2, 256, 64, None, state, 3, 0, 9, None, nthreads=4)
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    lat.close()
