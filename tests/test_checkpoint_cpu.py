"""The checkpoint reader (lgca_amd/checkpoint.py) on files built here: the
checksum is FNV-1a-64 (pinned by its published test vectors), and damaged or
truncated files are refused.  The engine's writer is checked against this
reader in test_gpu_checkpoint.py."""
import numpy as np
import pytest

from lgca_amd.checkpoint import HEADER, MAGIC, fnv1a64, read_checkpoint


def test_fnv1a64_published_vectors():
    assert fnv1a64(b"") == 0xCBF29CE484222325
    assert fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a64(b"foobar") == 0x85944171F73967E8


def write(path, types, state, **kw):
    payload = types.tobytes() + state.tobytes()
    f = dict(model=2, dim_x=6, dim_y=4, r=1, seed=9, step=10, calls=11, fdir=1, thr=123,
             fx0=0, fx1=6, fy0=0, fy1=4)
    f.update(kw)
    hdr = HEADER.pack(MAGIC, 1, f["model"], f["dim_x"], f["dim_y"], f["r"], 0, f["seed"],
                      f["step"], f["calls"], f["fdir"], 0, f["thr"], f["fx0"], f["fx1"], f["fy0"],
                      f["fy1"], fnv1a64(payload))
    path.write_bytes(hdr + payload)


def test_reader_round_trip_and_damage(tmp_path):
    types = (np.arange(24, dtype=np.uint8) % 3)
    state = (np.arange(24, dtype=np.uint8) * 7) % 128
    p = tmp_path / "ok.bin"
    write(p, types, state)
    ck = read_checkpoint(str(p))
    assert (ck["seed"], ck["step"], ck["force_calls"], ck["force_thr"]) == (9, 10, 11, 123)
    assert np.array_equal(ck["types"], types) and np.array_equal(ck["state"], state)
    raw = bytearray(p.read_bytes())
    raw[-1] ^= 1
    (tmp_path / "bad.bin").write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        read_checkpoint(str(tmp_path / "bad.bin"))
    (tmp_path / "short.bin").write_bytes(bytes(raw[:-1]))
    with pytest.raises(ValueError, match="size"):
        read_checkpoint(str(tmp_path / "short.bin"))
    (tmp_path / "magic.bin").write_bytes(b"XXXXXXXX" + bytes(raw[8:]))
    with pytest.raises(ValueError, match="magic"):
        read_checkpoint(str(tmp_path / "magic.bin"))
