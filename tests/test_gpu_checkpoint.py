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
