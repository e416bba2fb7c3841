"""Checkpoint and resume (lgca_save_checkpoint / lgca_load_checkpoint).

A run resumed from a checkpoint must continue bit for bit as the saved run:
state, cell types, seed, step (the hash input, including the K-step
alignment), the per-cell force settings and the count-mode force call
counter all travel in the file.  The resumed sequence is also replayed on the
oracle."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def make_lattice(seed=77):
    from lgca_amd import Lattice
    lat = Lattice("FHP_III", dim_x=300, dim_y=100, coarse_graining_radius=5, seed=seed)
    lat.apply_bc_cylinder(60, 51, 24.0)
    lat.init_random(0.3)
    lat.set_forcing("x", 0.004, 0, 40)
    return lat


def continue_run(lat):
    lat.step(29)
    r1 = lat.apply_body_force_count("x", 150)
    lat.step(11)
    return r1, lat.get_state_u8()


def test_resume_is_bit_identical(tmp_path):
    from lgca_amd import Lattice
    from lgca_amd.checkpoint import read_checkpoint
    path = str(tmp_path / "ck.bin")
    lat = make_lattice()
    lat.step(37)
    r0 = lat.apply_body_force_count("x", 200)
    lat.step(6)
    saved_state = lat.get_state_u8()
    saved_types = lat.get_cell_types()
    lat.save_checkpoint(path)
    want_r1, want = continue_run(lat)
    lat.close()

    ck = read_checkpoint(path)
    assert (ck["model"], ck["dim_x"], ck["dim_y"], ck["coarse_radius"]) == (3, 300, 100, 5)
    assert (ck["seed"], ck["step"], ck["force_calls"], ck["force_dir"]) == (77, 43, 1, 1)
    assert (ck["fx0"], ck["fx1"], ck["fy0"], ck["fy1"]) == (0, 40, 0, 100)
    assert np.array_equal(ck["state"], saved_state)
    assert np.array_equal(ck["types"], saved_types)

    res = Lattice.from_checkpoint(path)
    assert res.steps == 43
    got_r1, got = continue_run(res)
    res.close()
    assert got_r1 == want_r1
    assert np.array_equal(got, want)

    # the same continuation on the oracle, from the checkpoint's bytes
    st = ck["state"].copy()
    force = dict(dir="x", p=0.004, x0=0, x1=40, y0=0, y1=100)
    st = orc.run(3, 300, 100, ck["types"], st, 77, 43, 29, force=force, nthreads=8)
    assert orc.body_force_count(3, 300, 100, ck["types"], st, 77, 1, "x", 150) == want_r1
    st = orc.run(3, 300, 100, ck["types"], st, 77, 72, 11, force=force, nthreads=8)
    assert np.array_equal(st, want)
    assert r0 > 0


def test_damaged_checkpoints_are_refused(tmp_path):
    from lgca_amd import Lattice, LgcaError
    path = str(tmp_path / "ck.bin")
    lat = make_lattice()
    lat.step(5)
    lat.save_checkpoint(path)
    lat.close()
    raw = bytearray(open(path, "rb").read())
    bad = tmp_path / "bad.bin"
    raw[200] ^= 0x10
    bad.write_bytes(bytes(raw))
    with pytest.raises(LgcaError) as ei:
        Lattice.from_checkpoint(str(bad))
    assert ei.value.code == 4
    short = tmp_path / "short.bin"
    short.write_bytes(bytes(raw[:-3]))
    with pytest.raises(LgcaError):
        Lattice.from_checkpoint(str(short))
    with pytest.raises(LgcaError):
        Lattice.from_checkpoint(str(tmp_path / "absent.bin"))
    assert os.path.exists(path)


def test_damaged_header_is_refused(tmp_path):
    """The checksum covers the header: a changed seed or force window is
    refused, not loaded."""
    from lgca_amd import Lattice, LgcaError
    path = str(tmp_path / "ck.bin")
    lat = make_lattice()
    lat.step(3)
    lat.save_checkpoint(path)
    lat.close()
    raw = open(path, "rb").read()
    for pos in (40, 72, 80, 100):  # step, fx0, fy0, controller u
        bad = bytearray(raw)
        bad[pos] ^= 0x04
        p = tmp_path / f"bad{pos}.bin"
        p.write_bytes(bytes(bad))
        with pytest.raises(LgcaError) as ei:
            Lattice.from_checkpoint(str(p))
        assert ei.value.code == 4


@pytest.mark.parametrize("window", [(0, 0), (30, 10)])
def test_first_layout_with_an_empty_force_window(tmp_path, window):
    """A file of the first layout whose x window is empty (fx0 == fx1) or
    reversed (fx0 > fx1) loads with the forcing switched off, as
    lgca_set_forcing would have left it, and steps as the oracle does."""
    from lgca_amd import Lattice
    from lgca_amd.checkpoint import MAGIC, write_checkpoint
    dx, dy, seed = 200, 40, 5
    rng = np.random.default_rng(1)
    types = np.zeros(dx * dy, dtype=np.uint8)
    types[:dx] = types[-dx:] = 1
    state = rng.integers(0, 128, dx * dy, dtype=np.uint8)
    state[types != 0] = 0
    hd = dict(magic=MAGIC, version=1, model=2, dim_x=dx, dim_y=dy, coarse_radius=2, flags=0,
              seed=seed, step=6, force_calls=0, force_dir=1, reserved1=0,
              force_thr=1 << 31, fx0=window[0], fx1=window[1], fy0=0, fy1=dy)
    path = str(tmp_path / "v1.bin")
    write_checkpoint(path, hd, types, state)
    lat = Lattice.from_checkpoint(path)
    assert lat.steps == 6
    lat.step(21)
    want = orc.run(2, dx, dy, types, state.copy(), seed, 6, 21, None, nthreads=8)
    np.testing.assert_array_equal(lat.get_state_u8(), want)
    # its snapshot is the saved state
    lat.post_process(cells=False)
    _, _, md, mm = orc.post_process(2, dx, dy, 2, state)
    np.testing.assert_array_equal(lat.mean_density().view(np.uint32), md.view(np.uint32))
    lat.close()


def test_snapshot_and_controller_resume_bit_for_bit(tmp_path):
    """A controller run saved between ticks, with a snapshot older than the
    live state, resumes with the same snapshot (post-processing and mean
    velocity), the same controller amount and the same later ticks."""
    from lgca_amd import Lattice
    path = str(tmp_path / "ctl.bin")

    def make_pipe():
        lat = Lattice("FHP_III", "pipe", 80.0, 0.3, 10, seed=3)
        lat.apply_bc_pipe()
        lat.init_random()
        lat.copy_data_to_output_buffer()
        return lat

    def ticks(lat, n):
        log = []
        for _ in range(n):
            lat.controlled_tick(5)
            _, applied, v = lat.controller_state()
            log.append((applied, v[0], v[1]))
        return log

    a = make_pipe()
    # amounts that pass through all of the controller's branches
    a.set_forcing_controller(1000, 300, u=0.03)
    first = ticks(a, 25)
    a.step(3)  # the live state moves past the snapshot of the last tick
    a.save_checkpoint(path)
    a.post_process(cells=False)
    md_a, mm_a = a.mean_density().copy(), a.mean_momentum().copy()
    v_a = a.get_mean_velocity()
    want = ticks(a, 30)
    state_a = a.get_state_u8()
    a.close()

    b = Lattice.from_checkpoint(path)
    b.post_process(cells=False)
    np.testing.assert_array_equal(b.mean_density().view(np.uint32), md_a.view(np.uint32))
    np.testing.assert_array_equal(b.mean_momentum().view(np.uint32), mm_a.view(np.uint32))
    assert b.get_mean_velocity() == v_a
    assert b.controller_state()[1] == first[-1][0]
    got = ticks(b, 30)
    assert got == want
    assert {1000, 300, -1} <= set(x[0] for x in first + want)
    np.testing.assert_array_equal(b.get_state_u8(), state_a)
    b.close()
