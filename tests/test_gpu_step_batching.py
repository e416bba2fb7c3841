"""Lattice.collide_and_propagate counts single steps and enqueues them
together before any other device call (as the C++ HIP_Lattice does, which
tests/native/viewer_loop.cpp covers); results must be those of single steps."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_single_steps_are_batched_without_changing_results():
    """collide_and_propagate counts steps and enqueues them together before
    any other device call; interleaving them with forcing changes, cell
    types, count-mode force and state reads gives the same states as the
    oracle stepping one at a time."""
    from lgca_amd import Lattice
    dx, dy, seed = 640, 96, 31
    rng = np.random.default_rng(seed)
    types = np.zeros(dx * dy, dtype=np.uint8)
    types[:dx] = types[-dx:] = 1
    state = rng.integers(0, 128, dx * dy, dtype=np.uint8)
    lat = Lattice(3, dim_x=dx, dim_y=dy, coarse_graining_radius=4, seed=seed)
    lat.set_cell_types(types)
    lat.set_state_u8(state)
    lat.set_forcing("x", 0.05, 0, 100)
    want = state.copy()
    force = dict(dir="x", p=0.05, x0=0, x1=100)
    step = 0
    for n in (3, 70, 1, 6):  # 70 crosses the 64-step flush
        for _ in range(n):
            lat.collide_and_propagate()
        want = orc.run(3, dx, dy, types, want, seed, step, n, force, nthreads=8)
        step += n
        np.testing.assert_array_equal(lat.get_state_u8(), want)
        assert lat.steps == step
    # cell types changed with steps pending: those steps keep the old types
    for _ in range(5):
        lat.collide_and_propagate()
    types2 = types.copy()
    types2[dy // 2 * dx + 200: dy // 2 * dx + 260] = 1
    lat.m_cell_type_cpu[:] = types2
    lat._types_dirty = True
    for _ in range(7):
        lat.collide_and_propagate()
    want = orc.run(3, dx, dy, types, want, seed, step, 5, force, nthreads=8)
    want = orc.run(3, dx, dy, types2, want, seed, step + 5, 7, force, nthreads=8)
    step += 12
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    # a forcing change with steps pending applies only to later steps
    for _ in range(4):
        lat.collide_and_propagate()
    lat.set_forcing("y", 0.1)
    for _ in range(4):
        lat.collide_and_propagate()
    want = orc.run(3, dx, dy, types2, want, seed, step, 4, force, nthreads=8)
    want = orc.run(3, dx, dy, types2, want, seed, step + 4, 4, dict(dir="y", p=0.1), nthreads=8)
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    lat.close()
