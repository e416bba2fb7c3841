"""The viewers' forcing controller on the GPU (lgca_controller_*), against the
same loop driven from the host through the reference interface
(apps/pipe/pipe_viewer.cpp:107-146: get_mean_velocity, apply_body_force,
collide_and_propagate, copy_data_to_output_buffer), tick by tick: the amounts
applied, the velocities the decisions saw and the final state must agree bit
for bit, and the count-mode call counter must carry over to host calls and
checkpoints."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PP = 5


def make(seed):
    from lgca_amd import Lattice
    lat = Lattice("FHP_III", "pipe", 80.0, 0.3, 10, seed=seed)  # the pipe viewer's 1480 x 740
    lat.apply_bc_pipe()
    lat.init_random()
    lat.copy_data_to_output_buffer()
    return lat


def host_loop(lat, ticks, u=None, forcing=None, equilibrium=None):
    forcing = lat.get_initial_forcing() if forcing is None else forcing
    eq = lat.get_equilibrium_forcing() if equilibrium is None else equilibrium
    # the viewer compares v < u in float and v > 0.9 * u in double
    u = float(lat.u()) if u is None else float(np.float32(u))
    log = []
    for _ in range(ticks):
        v = lat.get_mean_velocity()
        applied = -1
        if v[0] < u:
            if v[0] > 0.9 * u:
                forcing = eq
            lat.apply_body_force(forcing)
            applied = forcing
        lat.step(PP)
        lat.copy_data_to_output_buffer()
        log.append((applied, float(v[0]), float(v[1])))
    return log


def device_loop(lat, ticks, u=None, forcing=None, equilibrium=None):
    lat.set_forcing_controller(forcing, equilibrium, u=u)
    log = []
    for _ in range(ticks):
        lat.controlled_tick(PP)
        _, applied, v = lat.controller_state()
        log.append((applied, v[0], v[1]))
    return log


@pytest.mark.parametrize("seed,ticks", [(11, 12), (5, 40)])
def test_device_controller_matches_host_loop(seed, ticks, tmp_path):
    from lgca_amd.checkpoint import read_checkpoint
    a = make(seed)
    want = host_loop(a, ticks)
    b = make(seed)
    got = device_loop(b, ticks)
    assert got == want
    assert any(x[0] >= 0 for x in want)
    assert np.array_equal(a.get_state_u8(), b.get_state_u8())
    # the call counter continues on the host path and into checkpoints
    ra = a.apply_body_force(77)
    rb = b.apply_body_force(77)
    assert ra == rb
    assert np.array_equal(a.get_state_u8(), b.get_state_u8())
    a.save_checkpoint(str(tmp_path / "a.ckpt"))
    b.save_checkpoint(str(tmp_path / "b.ckpt"))
    ca = read_checkpoint(str(tmp_path / "a.ckpt"), verify=False)
    cb = read_checkpoint(str(tmp_path / "b.ckpt"), verify=False)
    assert ca["force_calls"] == cb["force_calls"] == 1 + sum(1 for x in want if x[0] >= 0)
    a.close()
    b.close()


def test_all_three_branches_of_the_decision():
    """With gentle amounts and a low target the run passes through the initial
    forcing, the equilibrium forcing (0.9 u < v < u) and no forcing (v >= u)."""
    kw = dict(u=0.03, forcing=1000, equilibrium=300)
    a = make(3)
    want = host_loop(a, 80, **kw)
    b = make(3)
    got = device_loop(b, 80, **kw)
    assert got == want
    amounts = [x[0] for x in want]
    assert {1000, 300, -1} <= set(amounts), amounts
    assert np.array_equal(a.get_state_u8(), b.get_state_u8())
    a.close()
    b.close()


def test_controller_needs_set():
    from lgca_amd import LgcaError
    lat = make(1)
    with pytest.raises(LgcaError):
        lat.controlled_tick(5)
    lat.close()
