"""The C++ drop-in (lgca_amd/host/hip_lattice.h) driven exactly as the
reference's pipe viewer drives OMP_Lattice (apps/pipe/pipe_viewer.cpp:35-201),
by a plain C++ program linked against the C ABI; the same sequence is replayed
on the CPU oracle and must give the same bits.

The viewer's forcing decisions depend on FP32 mean velocities; the native
program logs the decisions it took, the oracle replay applies the same ones,
and the logged velocities are checked against the oracle's to 1e-5."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
NATIVE = os.path.join(ROOT, "tests", "native", "viewer_loop")


def test_pipe_viewer_sequence_through_cpp_dropin(tmp_path):
    if not os.path.exists(NATIVE):
        pytest.fail("tests/native/viewer_loop is not built (run __graft_entry__.build())")
    seed, ticks = 4242, 12
    print("[native] pipe viewer loop", flush=True)
    r = subprocess.run([NATIVE, str(tmp_path), str(seed), str(ticks)], capture_output=True,
                       text=True, timeout=120)
    sys.stdout.write(r.stdout + r.stderr)
    assert r.returncode == 0, r.stderr
    assert "Error check PASSED" in r.stderr
    lines = open(tmp_path / "ticks.txt").read().split("\n")
    dx, dy, n0, u, forcing0 = lines[0].split()
    dx, dy, n0 = int(dx), int(dy), int(n0)
    assert (dx, dy) == (1480, 740)  # the pipe viewer's lattice (Re 80, Ma 0.3, r 10)

    types = np.zeros(dx * dy, dtype=np.uint8)
    types[:dx] = 1
    types[-dx:] = 1
    state = orc.init_random(3, dx, dy, types, seed, 1.0 / 7.0)
    assert orc.count_particles(state) == n0
    calls = 0
    for t in range(ticks):
        _, v0, v1, applied = lines[1 + t].split()
        cd, cm = orc.cell_post_process(3, dx, dy, state)
        want_v = orc.mean_velocity(dx, dy, types, cd, cm)
        np.testing.assert_allclose([float(v0), float(v1)], want_v, rtol=1e-5, atol=1e-6)
        if int(applied) >= 0:
            orc.body_force_count(3, dx, dy, types, state, seed, calls, "x", int(applied))
            calls += 1
        state = orc.run(3, dx, dy, types, state, seed, 5 * t, 5, nthreads=8)
    got = np.fromfile(tmp_path / "state.bin", dtype=np.uint8)
    np.testing.assert_array_equal(got, state)
    _, _, md, mm = orc.post_process(3, dx, dy, 10, state)
    mean = np.fromfile(tmp_path / "mean.bin", dtype=np.float32)
    np.testing.assert_array_equal(mean[:md.size].view(np.uint32), md.view(np.uint32))
    np.testing.assert_array_equal(mean[md.size:].view(np.uint32), mm.view(np.uint32))
    assert calls > 0

    # the IoVti files of the last tick hold the post-processed snapshot
    from lgca_amd.checkpoint import read_checkpoint
    from lgca_amd.io_vti import read_vti
    cd, cm, md, mm = orc.post_process(3, dx, dy, 10, state)
    cell = read_vti(str(tmp_path / f"cell_res_{5 * ticks}.vti"))
    assert cell["extent"] == (0, dx, 0, dy, 0, 0) and cell["section"] == "CellData"
    np.testing.assert_array_equal(cell["arrays"]["Cell density"].reshape(-1).view(np.uint32),
                                  cd.view(np.uint32))
    np.testing.assert_array_equal(cell["arrays"]["Cell momentum"].reshape(-1).view(np.uint32),
                                  cm.view(np.uint32))
    mean_v = read_vti(str(tmp_path / f"mean_res_{5 * ticks}.vti"))
    assert mean_v["extent"] == (0, dx // 20 - 1, 0, dy // 20 - 1, 0, 0)
    assert mean_v["scalars"] == "Mean density"
    np.testing.assert_array_equal(mean_v["arrays"]["Mean density"].reshape(-1).view(np.uint32),
                                  md.view(np.uint32))
    np.testing.assert_array_equal(mean_v["arrays"]["Mean momentum"].reshape(-1).view(np.uint32),
                                  mm.view(np.uint32))
    # and the checkpoint the program saved last
    ck = read_checkpoint(str(tmp_path / "checkpoint.bin"), verify=False)
    assert (ck["seed"], ck["step"], ck["force_calls"]) == (seed, 5 * ticks, calls)
    np.testing.assert_array_equal(ck["state"], state)
    np.testing.assert_array_equal(ck["types"], types)
