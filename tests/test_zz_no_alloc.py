"""lgca_step_noalloc, the step entry point that allocates nothing, against
lgca_step through the C ABI: the same bits on two calls in a row, the second
on a smaller lattice, and a clear error where one pass would need more force
decisions than its fixed scratch array holds."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(model, dx, dy, seed, force, tuning=None):
    from lgca_amd import Lattice
    rng = np.random.default_rng(seed)
    types = np.zeros((dy, dx), dtype=np.uint8)
    types[0, :] = types[-1, :] = 1
    ys, xs = np.mgrid[0:dy, 0:dx]
    types[(xs - dx // 3) ** 2 + (ys - dy // 2) ** 2 < (dy // 5) ** 2] = 1
    state = rng.integers(0, 1 << (4 if model == 0 else (6 if model == 1 else 7)), dx * dy,
                         dtype=np.uint8)
    lat = Lattice(model, dim_x=dx, dim_y=dy, coarse_graining_radius=1, seed=seed)
    if tuning:
        lat.set_tuning(*tuning)
    lat.set_cell_types(types.reshape(-1))
    lat.set_state_u8(state)
    lat.set_step(3)
    lat.set_forcing(*force)
    return lat


def noalloc_step(lat, n):
    from lgca_amd._capi import check, lib
    h = lat.handle
    check(lib().lgca_step_noalloc(h, C.c_uint64(n)), h)


def test_noalloc_matches_step_on_two_calls():
    cases = [
        # a larger lattice first, then a smaller one: the second call reuses
        # (and overwrites) the decisions the first left in the scratch array
        (3, 4096, 200, 11, ("x", 0.02, 0, 700), None, 57),
        (0, 96, 40, 5, ("y", 0.3, 10, 80), None, 23),
    ]
    for model, dx, dy, seed, force, tuning, n in cases:
        say(f"[noalloc] model={model} {dx}x{dy} force={force} steps={n}")
        a = make(model, dx, dy, seed, force, tuning)
        b = make(model, dx, dy, seed, force, tuning)
        a.step(n)
        noalloc_step(b, n)
        np.testing.assert_array_equal(b.get_state_u8(), a.get_state_u8())
        assert a.steps == b.steps == 3 + n
        a.close()
        b.close()


def test_noalloc_after_tune_and_with_the_one_step_kernel():
    say("[noalloc] tuned shape and the one-step kernel")
    for tuning in ((-1, 0), (0, 0), (8, 6)):
        a = make(2, 1000, 66, 9, ("x", 0.05, 0, 64), tuning)
        b = make(2, 1000, 66, 9, ("x", 0.05, 0, 64), tuning)
        b.tune()
        a.step(41)
        noalloc_step(b, 30)
        noalloc_step(b, 11)
        np.testing.assert_array_equal(b.get_state_u8(), a.get_state_u8())
        a.close()
        b.close()


def test_noalloc_rejects_a_force_region_beyond_its_scratch():
    from lgca_amd import LgcaError
    say("[noalloc] oversized force region")
    # 8192 x 2048 forced everywhere: 256 words x 2048 rows per step; 8-step
    # passes need 2^22 words, twice the scratch
    lat = make(2, 8192, 2048, 1, ("x", 0.01, 0, 8192), (8, 16))
    with pytest.raises(LgcaError, match="fixed scratch"):
        noalloc_step(lat, 8)
    assert lat.steps == 3
    # 4-step passes fit exactly
    lat.set_tuning(4, 16)
    noalloc_step(lat, 8)
    assert lat.steps == 11
    lat.close()
