"""Pins the CPU oracle (oracle/cpu_ref.c) before anything is compared with it.

Sources of truth, none of which is the oracle's own formula:
  * Random123 known-answer vectors for Philox4x32-10;
  * tables dumped from the reference's own src/lgca_models.h, compiled by
    oracle/Makefile (tests/golden/ref_models.json, ref_offsets.json);
  * conservation laws and closed-form trajectories of the automaton;
  * indicator-field probes of the coarse-graining window.
"""
import json
import math
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import REF_DUMP

MODEL_NAMES = ["HPP", "FHP_I", "FHP_II", "FHP_III"]
Q = [4, 6, 7, 7]
# Integer momentum per direction: HPP unit vectors; FHP in units (1/2, sin(pi/3)).
MOM = {
    0: [(1, 0), (0, 1), (-1, 0), (0, -1)],
    1: [(2, 0), (1, 1), (-1, 1), (-2, 0), (-1, -1), (1, -1)],
}
MOM[2] = MOM[3] = MOM[1] + [(0, 0)]


def momentum(model, s):
    mx = my = 0
    for d, (vx, vy) in enumerate(MOM[model]):
        if (s >> d) & 1:
            mx += vx
            my += vy
    return mx, my


# ------------------------------------------------------------------ hash

def test_philox_known_answers(orc):
    # Random123 kat_vectors, philox4x32 10 rounds.
    assert orc.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                              0x6D5451FD]
    assert orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                      [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_threshold_closed_form(orc):
    assert orc.threshold(0.0) == 0
    assert orc.threshold(-1.0) == 0
    assert orc.threshold(1.0) == 2**32
    assert orc.threshold(0.5) == 2**31
    assert orc.threshold(0.2) == 858993459  # floor(0.2 * 2^32)
    assert orc.threshold(1e-3) == 4294967


def test_hash_streams_are_the_stated_philox_outputs(orc):
    L = orc.lib()
    seed = 0x0123456789ABCDEF
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for (y, w, step) in [(0, 0, 0), (5, 7, 3), (8191, 1023, 4999), (3, 2, 2**33 + 6)]:
        assert L.oracle_chirality_word(seed, step, y, w) == \
            orc.philox([w, y, (step & 0xFFFFFFFF) >> 2, 0], key)[step & 3]
    for (x, y, d) in [(0, 0, 0), (9, 3, 6)]:
        assert L.oracle_fill_draw(seed, x, y, d) == orc.philox([x >> 2, y, d, 2], key)[x & 3]


# ------------------------------------------------- per-cell force contract

def _leading_ones(u):
    n = 0
    while n < 32 and (u >> (31 - n)) & 1:
        n += 1
    return n


def test_force_word_half_probability_closed_form(orc):
    """At p = 1/2 the first-fire thresholds are 2^32 - 2^(31-m) exactly, so the
    first forced cell among the next n is at offset (leading one-bits of the
    uniform), none when that is >= n.  This pins the counter layout (which
    Philox outputs, in which order), the sharing of a group's extras by its
    four steps in step order, and the walk, independently of the integer
    recurrence.  At p = 1/2 nearly every step takes extras, so a later step of
    a group reading the group's extras from the start would fail here."""
    L = orc.lib()
    seed = 0xFEEDFACE12345678
    key = [seed & 0xFFFFFFFF, seed >> 32]
    thr = orc.threshold(0.5)
    assert thr == 2 ** 31
    cases = [(0, 0, 0), (1, 2, 3), (77, 5, 4999), (8191, 1023, 2 ** 33 + 7),
             (4, 9, 4), (4, 9, 5), (4, 9, 6), (4, 9, 7)]
    for (y, w, step) in cases:
        group = (step & 0xFFFFFFFF) >> 2
        prim = orc.philox([w, y, group, 1], key)
        extras = []
        for e in range(4 * 32):
            extras.append(orc.philox([w, y, group, 5 + 4 * (e >> 2)], key)[e & 3])
        used = 0
        for i in range((step & 3) + 1):  # the group's steps up to this one
            us = [prim[i]]
            want, pos, k = 0, 0, 0
            while True:
                m = _leading_ones(us[k])
                k += 1
                if m >= 32 - pos:
                    break
                want |= 1 << (pos + m)
                pos += m + 1
                if pos == 32:
                    break
                us.append(extras[used])
                used += 1
        assert L.oracle_force_word(seed, step, y, w, thr) == want, (y, w, step)


def test_force_word_edge_probabilities(orc):
    L = orc.lib()
    for w in range(50):
        assert L.oracle_force_word(7, 11, 3, w, 0) == 0
        assert L.oracle_force_word(7, 11, 3, w, 2 ** 32) == 0xFFFFFFFF


@pytest.mark.parametrize("p", [1e-3, 0.02, 0.3, 0.9])
def test_force_word_is_iid_bernoulli(orc, p):
    """Statistics over 20000 words: the fire rate is p, every bit position fires
    at rate p, adjacent cells fire independently, and the count per word
    follows Binomial(32, p).  A wrong threshold, an off-by-one in the walk or
    a reused uniform shows up here."""
    from math import comb, sqrt
    L = orc.lib()
    thr = orc.threshold(p)
    words = np.array([L.oracle_force_word(12345, 1000 + (i % 7), i // 101, i % 101, thr)
                      for i in range(20000)], dtype=np.uint64)
    bits = ((words[:, None] >> np.arange(32, dtype=np.uint64)) & np.uint64(1)).astype(np.int64)
    n = bits.size
    rate = bits.mean()
    assert abs(rate - p) < 5 * sqrt(p * (1 - p) / n)
    per_pos = bits.mean(axis=0)
    assert np.all(np.abs(per_pos - p) < 5 * sqrt(p * (1 - p) / bits.shape[0]))
    pair = (bits[:, :-1] & bits[:, 1:]).mean()
    assert abs(pair - p * p) < 5 * sqrt(p * p * (1 - p * p) / (bits.shape[0] * 31)) + 1e-4
    counts = np.bincount(bits.sum(axis=1), minlength=33)
    expect = np.array([comb(32, k) * p ** k * (1 - p) ** (32 - k) for k in range(33)]) * len(words)
    keep = expect >= 5
    chi2 = (((counts - expect) ** 2 / np.maximum(expect, 1e-12))[keep]).sum()
    dof = max(1, int(keep.sum()) - 1)
    assert chi2 < dof + 6 * sqrt(2 * dof), (chi2, dof)


# ------------------------------------------------------- rules vs reference

@pytest.mark.parametrize("model", range(4))
def test_collide_matches_reference(orc, ref_models, model):
    t = ref_models[MODEL_NAMES[model]]
    assert t["num_dir"] == Q[model]
    for p, key in ((0, "collide_p0"), (1, "collide_p1")):
        mine = [orc.collide(model, s, p) for s in range(1 << Q[model])]
        assert mine == t[key]


@pytest.mark.parametrize("model", range(4))
@pytest.mark.parametrize("rule", ["bounce_back", "bounce_forward_x", "bounce_forward_y"])
def test_bounce_matches_reference(orc, ref_models, model, rule):
    fn = getattr(orc.lib(), "oracle_" + rule)
    assert [fn(model, s) for s in range(1 << Q[model])] == ref_models[MODEL_NAMES[model]][rule]


@pytest.mark.parametrize("model", range(4))
def test_lattice_vectors_match_reference(orc, ref_models, model):
    t = ref_models[MODEL_NAMES[model]]
    L = orc.lib()
    assert [L.oracle_lattice_vec_bits(model, d, 0) for d in range(Q[model])] == t["lattice_vec_x"]
    assert [L.oracle_lattice_vec_bits(model, d, 1) for d in range(Q[model])] == t["lattice_vec_y"]
    if model:
        s = struct.unpack("<f", struct.pack("<I", t["lattice_vec_y"][1]))[0]
        assert t["lattice_vec_y"][1] == 0x3F5DB3D7
        assert s == np.float32(math.sin(math.pi / 3))


@pytest.mark.parametrize("model", range(4))
def test_collide_conserves_mass_and_momentum(orc, model):
    for p in (0, 1):
        for s in range(1 << Q[model]):
            o = orc.collide(model, s, p)
            assert bin(o).count("1") == bin(s).count("1")
            assert momentum(model, o) == momentum(model, s)


def test_collide_change_counts_and_fhp3_equals_fhp2(ref_models):
    changed = {m: sum(1 for s, o in enumerate(ref_models[m]["collide_p0"]) if s != o)
               for m in MODEL_NAMES}
    assert changed == {"HPP": 2, "FHP_I": 5, "FHP_II": 22, "FHP_III": 22}
    for k in ("collide_p0", "collide_p1"):
        assert ref_models["FHP_II"][k] == ref_models["FHP_III"][k]
    # the two chiralities differ only on the rotating head-on pairs
    p0, p1 = ref_models["FHP_II"]["collide_p0"], ref_models["FHP_II"]["collide_p1"]
    assert sorted(s for s in range(128) if p0[s] != p1[s]) == [9, 18, 36, 73, 82, 100]


def test_fixture_matches_live_reference_build():
    """The committed fixtures equal what the reference header produces now."""
    if not os.path.exists(REF_DUMP):
        pytest.skip("reference checkout not present (oracle/_ref not built)")
    from conftest import GOLDEN
    live = json.loads(subprocess.run([REF_DUMP, "tables"], check=True, capture_output=True,
                                     text=True).stdout)
    with open(os.path.join(GOLDEN, "ref_models.json")) as f:
        assert json.load(f) == live


# ------------------------------------------------------ streaming geometry

def test_pull_sources_match_reference_offset_tables(orc, ref_offsets):
    """Compose the reference's offset tables the way src/omp_lattice.cpp:141-179
    does and compare with the oracle's torus neighbour for every cell and dir."""
    for rec in ref_offsets:
        m, dx, dy = rec["model"], rec["dim_x"], rec["dim_y"]
        n = dx * dy
        for cell in range(n):
            x, y = cell % dx, cell // dx
            par = "even" if y % 2 == 0 else "odd"
            east = (cell + 1) % dx == 0
            north = cell >= n - dx
            west = cell % dx == 0
            south = cell < dx
            for d in range(rec["num_dir"]):
                inv = rec["inv_dir"][d]
                off = rec["neighbor_" + par][inv]
                if east:
                    off += rec["western_" + par][inv]
                if north:
                    off += rec["southern_" + par][inv]
                if west:
                    off += rec["eastern_" + par][inv]
                if south:
                    off += rec["northern_" + par][inv]
                src = cell + off
                sx, sy = orc.pull_source(m, dx, dy, x, y, d)
                assert 0 <= src < n, (rec["model"], dx, dy, cell, d)
                assert src == sy * dx + sx, (m, dx, dy, x, y, d)


def _single(dx, dy, x, y, d):
    st = np.zeros(dx * dy, dtype=np.uint8)
    st[y * dx + x] = 1 << d
    return st


def _where(st, dx):
    idx = np.flatnonzero(st)
    assert idx.size == 1
    c = int(idx[0])
    return c % dx, c // dx, int(st[c])


@pytest.mark.parametrize("model", range(4))
def test_lone_particle_moves_straight_and_wraps(orc, model):
    dx, dy = 10, 6
    for d in range(Q[model] if model == 0 else 6):
        st = _single(dx, dy, 3, 2, d)
        for t in range(1, 2 * dx * dy + 1):
            st = orc.step(model, dx, dy, None, st, 7, t)
            x, y, s = _where(st, dx)
            assert s == 1 << d
            if model == 0:
                vx, vy = MOM[0][d]
                assert (x, y) == ((3 + vx * t) % dx, (2 + vy * t) % dy)
            elif t % 2 == 0:
                # in the FHP units (1/2, sin(pi/3)) a particle crosses vy rows per
                # step and vx whole cells every two steps
                vx, vy = MOM[model][d]
                assert (x, y) == ((3 + vx * (t // 2)) % dx, (2 + vy * t) % dy)
        assert _where(st, dx)[:2] == (3, 2)


def test_rest_particle_stays(orc):
    st = _single(8, 4, 5, 1, 6)
    for t in range(5):
        st = orc.step(2, 8, 4, None, st, 1, t)
    assert _where(st, 8) == (5, 1, 64)


@pytest.mark.parametrize("model", range(4))
def test_periodic_run_conserves_mass_and_momentum(orc, model):
    dx, dy = 40, 24
    st = orc.init_random(model, dx, dy, None, 5, 0.35)
    n0 = orc.count_particles(st)

    def total_mom(s):
        mx = my = 0
        for d, (vx, vy) in enumerate(MOM[model]):
            b = ((s >> d) & 1).astype(np.int64)
            mx += vx * int(b.sum())
            my += vy * int(b.sum())
        return mx, my

    m0 = total_mom(st)
    out = orc.run(model, dx, dy, None, st, 5, 0, 60)
    assert orc.count_particles(out) == n0
    assert total_mom(out) == m0


def test_run_equals_repeated_step(orc):
    dx, dy = 37, 10
    types = np.zeros(dx * dy, dtype=np.uint8)
    types[:dx] = 1
    types[-dx:] = 2
    types[5 * dx + 7] = 1
    st = orc.init_random(2, dx, dy, types, 3, 0.3)
    f = dict(dir="x", p=0.05, x0=3, x1=20)
    a = st.copy()
    for t in range(25):
        a = orc.step(2, dx, dy, types, a, 3, 100 + t, f)
    b = orc.run(2, dx, dy, types, st, 3, 100, 25, f, nthreads=4)
    np.testing.assert_array_equal(a, b)


def test_no_slip_wall_reverses_particle_one_step_later(orc):
    dx, dy = 4, 4
    types = np.zeros(dx * dy, dtype=np.uint8)
    types[3 * dx:] = 1  # north row solid, bounce-back
    st = _single(dx, dy, 1, 1, 1)  # HPP, moving north
    st = orc.step(0, dx, dy, types, st, 0, 0)
    assert _where(st, dx) == (1, 2, 2)
    st = orc.step(0, dx, dy, types, st, 0, 1)
    assert _where(st, dx) == (1, 3, 8)  # inside the wall, now moving south
    st = orc.step(0, dx, dy, types, st, 0, 2)
    assert _where(st, dx) == (1, 2, 8)


def test_slip_wall_mirrors_vertical_component(orc):
    dx, dy = 8, 4
    types = np.zeros(dx * dy, dtype=np.uint8)
    types[3 * dx:] = 2  # north row solid, bounce-forward
    st = _single(dx, dy, 2, 2, 1)  # FHP, moving north-east from an even row
    st = orc.step(1, dx, dy, types, st, 0, 0)
    assert _where(st, dx) == (2, 3, 1 << 5)  # entered the slip row, turned south-east
    st = orc.step(1, dx, dy, types, st, 0, 1)
    assert _where(st, dx) == (3, 2, 1 << 5)


def test_slip_corner_uses_y_mirror(orc):
    dx, dy = 6, 4
    types = np.full(dx * dy, 2, dtype=np.uint8)
    types[dx + 1:dx + 5] = 0
    types[2 * dx + 1:2 * dx + 5] = 0
    # particle pulled into the north-east corner (x=5, y=3) moving north-east
    st = _single(dx, dy, 5, 2, 1)  # even row: dir 1 goes to (x, y+1) on the odd row
    st = orc.step(1, dx, dy, types, st, 0, 0)
    x, y, s = _where(st, dx)
    assert (x, y) == (5, 3)
    assert s == 1 << 2  # mirror in y of direction 1 is direction 2, not the x mirror 5


def test_interior_slip_cell_passes_through(orc):
    dx, dy = 8, 6
    types = np.zeros(dx * dy, dtype=np.uint8)
    types[2 * dx + 4] = 2
    st = _single(dx, dy, 3, 2, 0)
    st = orc.step(1, dx, dy, types, st, 0, 0)
    assert _where(st, dx) == (4, 2, 1)


# ---------------------------------------------------------------- forcing

def test_per_cell_force_flips_and_conserves(orc):
    dx, dy = 64, 16
    st = np.full(dx * dy, 1 << 3, dtype=np.uint8)  # everything moving west
    n0 = orc.count_particles(st)
    out = orc.step(1, dx, dy, None, st, 9, 0, dict(dir="x", p=1.0, x0=10, x1=20))
    assert orc.count_particles(out) == n0
    grid = out.reshape(dy, dx)
    assert (grid[:, 10:20] == 1).all()
    assert (grid[:, :10] == 8).all() and (grid[:, 20:] == 8).all()
    out = orc.step(1, dx, dy, None, st, 9, 0, dict(dir="x", p=0.25))
    frac = float((out == 1).mean())
    assert 0.2 < frac < 0.3


def test_force_y_rules(orc):
    dx, dy = 8, 4
    st = np.full(dx * dy, (1 << 1) | (1 << 2), dtype=np.uint8)
    # stream: uniform state is invariant under streaming; collide: {1,2} with no
    # rest turns into rest + dir 0... use FHP-I so the pair {1,2} does not collide
    out = orc.step(1, dx, dy, None, st, 0, 0, dict(dir="y", p=1.0))
    assert (out == ((1 << 5) | (1 << 4))).all()
    st = np.full(dx * dy, 1 << 1, dtype=np.uint8)
    out = orc.step(0, dx, dy, None, st, 0, 0, dict(dir="y", p=1.0))
    assert (out == 1 << 3).all()


def test_count_mode_force(orc):
    dx, dy = 16, 8
    n = dx * dy
    st = np.full(n, 1 << 3, dtype=np.uint8)
    # forcing <= 0 still makes the one try of the reference's do-while loop
    a = st.copy()
    assert orc.body_force_count(2, dx, dy, None, a, 1, 0, "x", 0) == 1
    assert int((a == 1).sum()) == 1
    # every first visit flips; repeated candidates find the cell done already
    a = st.copy()
    assert orc.body_force_count(2, dx, dy, None, a, 1, 0, "x", 40) == 40
    assert int((a == 1).sum()) == 40
    # all solid: nothing flips and the loop stops after 2N candidates
    types = np.ones(n, dtype=np.uint8)
    a = st.copy()
    assert orc.body_force_count(2, dx, dy, types, a, 1, 0, "x", 5) == 0
    np.testing.assert_array_equal(a, st)
    # y forcing reverts two particles in one try when both flips are possible
    a = np.full(n, (1 << 1) | (1 << 2), dtype=np.uint8)
    assert orc.body_force_count(1, dx, dy, None, a, 1, 0, "y", 1) == 2


def _candidates(orc, seed, call, n, stop):
    """Distinct cells in the order the count-mode candidates first reach them,
    from the stated stream: candidate i of call c is cell
    (philox({i_lo, i_hi, c, 3})[1] << 32 | [0]) mod n; `stop(distinct, i)`
    ends the walk after candidate i."""
    key = [seed & 0xFFFFFFFF, seed >> 32]
    seen, order, i = set(), [], 0
    while True:
        o = orc.philox([i & 0xFFFFFFFF, i >> 32, call, 3], key)
        cell = ((o[1] << 32) | o[0]) % n
        i += 1
        if cell not in seen:
            seen.add(cell)
            order.append(cell)
        if stop(len(order), i):
            return order


def test_count_mode_candidates_are_the_stated_philox_outputs(orc):
    """Every cell holds one particle the x force can revert, so each first
    visit reverts exactly one and a revisit none: the cells that flip are the
    first `forcing` distinct candidates of the stated stream, and a negative
    amount (taken modulo 2^32, as the reference's unsigned comparison takes
    it, src/omp_lattice.cpp:263,345) visits all 2N candidates."""
    dx, dy = 37, 11
    n = dx * dy
    seed = 0x1122334455667788
    for call in (0, 1, 9):
        for forcing in (1, 5, 60, 400, -1, -(2 ** 31)):
            st = np.full(n, 1 << 3, dtype=np.uint8)
            got = orc.body_force_count(2, dx, dy, None, st, seed, call, "x", forcing)
            if forcing > 0:
                want = _candidates(orc, seed, call, n,
                                   lambda k, i: k >= forcing or i >= 2 * n)
            else:
                want = _candidates(orc, seed, call, n, lambda k, i: i >= 2 * n)
            assert got == len(want), (call, forcing)
            assert sorted(np.flatnonzero(st == 1).tolist()) == sorted(want)
    # 400 exceeds the distinct cells 2N candidates reach: the loop ends at 2N
    st = np.full(n, 1 << 3, dtype=np.uint8)
    assert orc.body_force_count(2, dx, dy, None, st, seed, 0, "x", 400) < 400


def test_mean_velocity_closed_form(orc):
    """get_mean_velocity (src/omp_lattice.cpp:508-557) is the mean over FLUID
    cells of momentum / density.  Every FLUID cell of the lattice holds the
    same particles, or half of them hold none: the mean is the one cell's
    velocity, or half of it; solid cells with particles do not count."""
    dx, dy = 24, 10
    types = np.zeros(dx * dy, dtype=np.uint8)
    types[:dx] = 1                 # a wall row, occupied below
    types[5 * dx + 3] = 2
    for model, s, vel in [(0, 0b0001, (1.0, 0.0)), (0, 0b0011, (0.5, 0.5)),
                          (2, 0b0000011, (0.75, 0.4330127)),   # dirs 0, 1: (1.5, sin60) / 2
                          (2, 0b1000100, (-0.25, 0.4330127)),   # dir 2 + rest: (-0.5, sin60) / 2
                          (1, 0b010010, (0.0, 0.0))]:           # dirs 1, 4 cancel
        st = np.full(dx * dy, s, dtype=np.uint8)
        cd, cm = orc.cell_post_process(model, dx, dy, st)
        got = orc.mean_velocity(dx, dy, types, cd, cm)
        np.testing.assert_allclose(got, vel, rtol=2e-7, atol=1e-12)
        half = st.copy()
        half[1::2] = 0             # every other cell empty (each row has even width)
        cd, cm = orc.cell_post_process(model, dx, dy, half)
        got = orc.mean_velocity(dx, dy, types, cd, cm)
        fluid = int((types == 0).sum())
        occupied = int(((types == 0) & (half != 0)).sum())
        np.testing.assert_allclose(got, np.array(vel) * occupied / fluid, rtol=2e-7, atol=1e-12)


def test_coarse_velocity_closed_form(orc):
    """Coarse velocity = mean momentum / mean density per coarse cell, 0 where
    the block is empty: a uniform one-direction state gives the direction's
    velocity exactly in x (multiples of 1/2 add exactly in FP32) and to FP32
    rounding in y; an empty block gives 0."""
    dx, dy, r = 32, 16, 4
    for model, s, vel in [(0, 0b0010, (0.0, 1.0)), (0, 0b0101, (0.0, 0.0)),
                          (1, 0b000001, (1.0, 0.0)), (2, 0b0100000, (0.5, -0.8660254))]:
        st = np.full((dy, dx), s, dtype=np.uint8)
        st[:, : r + 1] = 0         # the window's r + 1 columns of coarse column 0
        _, _, md, mm = orc.post_process(model, dx, dy, r, st.reshape(-1))
        v = orc.coarse_velocity(md, mm).reshape(dy // (2 * r), dx // (2 * r), 2)
        np.testing.assert_array_equal(v[:, 0], 0.0)
        np.testing.assert_array_equal(v[:, 1:, 0], np.float32(vel[0]))
        np.testing.assert_allclose(v[:, 1:, 1], vel[1], rtol=1e-6, atol=0)


# ----------------------------------------------------------- post-process

def _window_weights(dx, dy, r, c):
    """Closed form of the averaging window: coarse cell c covers columns
    x0..x0+r of rows y0..y0+2r; when dim_x == 2r the x = 2r candidate of each
    row lands on column 0 of the row above."""
    cdx = dx // (2 * r)
    x0, y0 = (c % cdx) * 2 * r, (c // cdx) * 2 * r
    w = np.zeros((dy, dx), dtype=np.int64)
    for y in range(y0, min(y0 + 2 * r + 1, dy)):
        w[y, x0:x0 + r + 1] += 1
    if dx == 2 * r:
        for y in range(y0 + 1, min(y0 + 2 * r + 2, dy)):
            w[y, 0] += 1
    return w


@pytest.mark.parametrize("dims", [(8, 8, 2), (12, 4, 2), (4, 8, 2), (6, 6, 3), (2, 4, 1)])
def test_mean_window_indicator_probe(orc, dims):
    dx, dy, r = dims
    n = dx * dy
    nc = (dx // (2 * r)) * (dy // (2 * r))
    weights = [_window_weights(dx, dy, r, c) for c in range(nc)]
    counts = [int(w.sum()) for w in weights]
    if dx >= 4 * r:
        rows_last = 2 * r
        assert counts[-1] == (r + 1) * rows_last
        if nc > dx // (2 * r):
            assert counts[0] == (r + 1) * (2 * r + 1)
    cm = np.zeros(2 * n, dtype=np.float32)
    for k in range(n):
        cd = np.zeros(n, dtype=np.float32)
        cd[k] = 1.0
        md, _ = orc.mean_post_process(dx, dy, r, cd, cm)
        for c in range(nc):
            assert md[c] * counts[c] == pytest.approx(weights[c].flat[k], abs=1e-6), (k, c)


def test_cell_post_process_closed_form(orc):
    s = np.float32(math.sin(math.pi / 3))
    for model in range(4):
        states = np.arange(1 << Q[model], dtype=np.uint8)
        cd, cm = orc.cell_post_process(model, states.size, 1, states)
        for st in range(1 << Q[model]):
            ix, iy = momentum(model, st)
            assert cd[st] == bin(st).count("1")
            if model == 0:
                assert (cm[2 * st], cm[2 * st + 1]) == (ix, iy)
            else:
                assert cm[2 * st] == np.float32(ix * 0.5)
                assert cm[2 * st + 1] == np.float32(iy) * s


def test_mean_post_process_sequential_fp32_order(orc):
    """Momentum-y sums of multiples of sin(pi/3) round at every add, so the
    order is observable; check it against an explicit y-outer, x-inner loop."""
    dx, dy, r = 32, 32, 4
    rng = np.random.default_rng(1)
    st = rng.integers(0, 128, dx * dy, dtype=np.uint8)
    cd, cm = orc.cell_post_process(2, dx, dy, st)
    md, mm = orc.mean_post_process(dx, dy, r, cd, cm)
    cdx = dx // (2 * r)
    for c in range(md.size):
        x0, y0 = (c % cdx) * 2 * r, (c // cdx) * 2 * r
        acc = np.float32(0)
        cnt = 0
        for y in range(y0, min(y0 + 2 * r + 1, dy)):
            for x in range(x0, x0 + r + 1):
                acc = np.float32(acc + cm[2 * (y * dx + x) + 1])
                cnt += 1
        assert mm[2 * c + 1] == np.float32(acc / np.float32(cnt))


def test_init_random_fraction_and_solids(orc):
    dx, dy = 128, 64
    types = np.zeros(dx * dy, dtype=np.uint8)
    types[:dx] = 1
    st = orc.init_random(2, dx, dy, types, 11, 0.25)
    assert (st[:dx] == 0).all()
    bits = np.unpackbits(st[dx:, None], axis=1, bitorder="little")[:, :7]
    assert abs(bits.mean() - 0.25) < 0.01
    assert (np.unpackbits(st[:, None], axis=1, bitorder="little")[:, 7] == 0).all()


# ------------------------------------------------------ bit-sliced checker

def _types_for(rng, kind, dx, dy):
    t = np.zeros((dy, dx), dtype=np.uint8)
    if kind == "pipe":
        t[0, :] = t[-1, :] = 1
    elif kind == "mixed":
        t[rng.random((dy, dx)) < 0.07] = 1
        t[rng.random((dy, dx)) < 0.05] = 2
        t[:, 0] = 2
        t[-1, :] = 2
        t[0, : dx // 2] = 1
    elif kind == "slipbox":
        t[0, :] = t[-1, :] = t[:, 0] = t[:, -1] = 2
    return None if kind == "periodic" else t.reshape(-1)


@pytest.mark.parametrize("model", [0, 1, 2, 3])
@pytest.mark.parametrize("dims", [(1, 2), (2, 2), (3, 4), (63, 6), (64, 4), (65, 8), (130, 10),
                                  (200, 12), (257, 2)])
def test_fast_run_equals_run(orc, model, dims):
    """oracle_fast_run (64 cells per word) equals oracle_run cell for cell on
    every model, on widths around the 64-cell word (ragged last words, a
    lattice one cell wide), with walls, slip and bounce-back cells, x and y
    forcing on regions that cut words, and runs starting at every step of a
    four-step chirality group."""
    dx, dy = dims
    if model != 0 and dy % 2:
        dy += 1
    rng = np.random.default_rng(dx * 7 + dy + model)
    for kind, force, step0 in [("periodic", None, 0), ("pipe", dict(dir="x", p=0.3), 1),
                               ("mixed", dict(dir="y", p=0.4, x0=dx // 3, x1=dx - dx // 4,
                                              y0=1, y1=dy), 2),
                               ("slipbox", dict(dir="x", p=1.0, x0=1, x1=max(2, dx // 2)), 3)]:
        types = _types_for(rng, kind, dx, dy)
        st = rng.integers(0, 1 << Q[model], dx * dy, dtype=np.uint8)
        want = orc.run(model, dx, dy, types, st, 99 + step0, step0, 13, force, nthreads=4)
        got = orc.fast_run(model, dx, dy, types, st, 99 + step0, step0, 13, force, nthreads=4)
        np.testing.assert_array_equal(got, want, err_msg=f"{kind} {force}")


def test_fast_run_equals_run_on_a_karman_like_lattice(orc):
    """A scaled Karman set-up (cylinder, walls, x force on the first 64
    columns, fill 0.25) over 40 steps from step 5."""
    dx, dy = 1024, 256
    t = np.zeros((dy, dx), dtype=np.uint8)
    t[0, :] = t[-1, :] = 1
    ys, xs = np.mgrid[0:dy, 0:dx]
    t[(xs - 128) ** 2 + (ys - 131) ** 2 < 32 ** 2] = 1
    types = t.reshape(-1)
    st = orc.init_random(3, dx, dy, types, 11, 0.25)
    force = dict(dir="x", p=2e-3, x0=0, x1=64)
    want = orc.run(3, dx, dy, types, st, 11, 5, 40, force, nthreads=8)
    got = orc.fast_run(3, dx, dy, types, st, 11, 5, 40, force, nthreads=8)
    np.testing.assert_array_equal(got, want)
