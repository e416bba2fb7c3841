"""The checkpoint reader and writer (lgca_amd/checkpoint.py) on files built
here: the checksum is FNV-1a-64 (pinned by its published test vectors) over
the header and the payload, and damaged or truncated files are refused.  The
engine's writer and reader are checked against this module in
test_gpu_checkpoint.py."""
import numpy as np
import pytest

from lgca_amd.checkpoint import (HEADER, HEADER2, MAGIC, checksum_of, fnv1a64, read_checkpoint,
                                 write_checkpoint)


def test_fnv1a64_published_vectors():
    assert fnv1a64(b"") == 0xCBF29CE484222325
    assert fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a64(b"foobar") == 0x85944171F73967E8
    # continuing a hash is hashing the concatenation
    assert fnv1a64(b"bar", fnv1a64(b"foo")) == fnv1a64(b"foobar")


def header(version=2, **kw):
    hd = dict(magic=MAGIC, version=version, model=2, dim_x=6, dim_y=4, coarse_radius=1, flags=1,
              seed=9, step=10, force_calls=11, force_dir=1, reserved1=0, force_thr=123, fx0=0,
              fx1=6, fy0=0, fy1=4, ctrl_dir=1, ctrl_u=0.25, ctrl_eq=5, ctrl_forcing=7,
              ctrl_amount=-1, ctrl_reverted=3, ctrl_vx=0.125, ctrl_vy=-0.5)
    hd.update(kw)
    return hd


def test_reader_round_trip_and_damage(tmp_path):
    types = (np.arange(24, dtype=np.uint8) % 3)
    state = (np.arange(24, dtype=np.uint8) * 7) % 128
    snap = (np.arange(24, dtype=np.uint8) * 5) % 128
    p = tmp_path / "ok.bin"
    write_checkpoint(str(p), header(), types, state, snap)
    raw = p.read_bytes()
    assert len(raw) == HEADER.size + HEADER2.size + 3 * 24
    ck = read_checkpoint(str(p))
    assert (ck["seed"], ck["step"], ck["force_calls"], ck["force_thr"]) == (9, 10, 11, 123)
    assert ck["controller"] and (ck["ctrl_eq"], ck["ctrl_forcing"], ck["ctrl_amount"]) == (5, 7, -1)
    assert (ck["ctrl_u"], ck["ctrl_vx"], ck["ctrl_vy"]) == (0.25, 0.125, -0.5)
    assert np.array_equal(ck["types"], types) and np.array_equal(ck["state"], state)
    assert np.array_equal(ck["snapshot"], snap)
    # a flipped bit anywhere, header or payload, is caught
    for pos in (20, 60, 100, 130, len(raw) - 1):
        bad = bytearray(raw)
        bad[pos] ^= 1
        (tmp_path / "bad.bin").write_bytes(bytes(bad))
        with pytest.raises(ValueError, match="checksum|magic|size"):
            read_checkpoint(str(tmp_path / "bad.bin"))
    (tmp_path / "short.bin").write_bytes(raw[:-1])
    with pytest.raises(ValueError, match="size"):
        read_checkpoint(str(tmp_path / "short.bin"))
    (tmp_path / "magic.bin").write_bytes(b"XXXXXXXX" + raw[8:])
    with pytest.raises(ValueError, match="magic"):
        read_checkpoint(str(tmp_path / "magic.bin"))


def test_version_1_files_are_read(tmp_path):
    """Files of the first layout: 96-byte header, checksum of the payload
    only, no snapshot (the state stands in for it)."""
    types = np.zeros(24, dtype=np.uint8)
    state = (np.arange(24, dtype=np.uint8) * 3) % 128
    hd = header(version=1, flags=0)
    p = tmp_path / "v1.bin"
    write_checkpoint(str(p), hd, types, state)
    assert checksum_of(hd, types.tobytes() + state.tobytes()) == fnv1a64(types.tobytes() +
                                                                        state.tobytes())
    ck = read_checkpoint(str(p))
    assert ck["version"] == 1 and not ck["controller"]
    assert np.array_equal(ck["snapshot"], state)
