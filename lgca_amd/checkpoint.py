"""Reader of the engine's checkpoint files (lgca_save_checkpoint in include/lgca.h).

Layout, little-endian: the 8-byte magic b"LGCACKP1"; u32 version, i32 model,
u32 dim_x, dim_y, coarse_radius, reserved; u64 seed, step, force_calls;
i32 force_dir, u32 reserved; u64 force_thr; u32 fx0, fx1, fy0, fy1; u64
FNV-1a-64 of the payload; then the payload: num_cells cell-type bytes and
num_cells state bytes (bit d = direction d, the reference's Bitset layout,
src/lgca_bitset.h:229-231).
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"LGCACKP1"
HEADER = struct.Struct("<8sIiIIIIQQQiIQIIIIQ")
assert HEADER.size == 96

FIELDS = ("magic", "version", "model", "dim_x", "dim_y", "coarse_radius", "reserved0", "seed",
          "step", "force_calls", "force_dir", "reserved1", "force_thr", "fx0", "fx1", "fy0", "fy1",
          "checksum")


def fnv1a64(data: bytes) -> int:
    """FNV-1a, 64-bit.  The hash is sequential, so this runs at Python speed;
    read_checkpoint(verify=False) skips it for large files."""
    h = np.uint64(0xCBF29CE484222325)
    prime = np.uint64(0x100000001B3)
    with np.errstate(over="ignore"):
        for b in np.frombuffer(data, dtype=np.uint8):
            h = (h ^ np.uint64(b)) * prime
    return int(h)


def read_checkpoint(path: str, verify: bool = True) -> dict:
    """Header fields plus `types` and `state` (uint8, num_cells each)."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < HEADER.size:
        raise ValueError("not a checkpoint: too short")
    hd = dict(zip(FIELDS, HEADER.unpack_from(raw)))
    if hd["magic"] != MAGIC or hd["version"] != 1:
        raise ValueError("not a checkpoint: bad magic or version")
    n = hd["dim_x"] * hd["dim_y"]
    payload = raw[HEADER.size:]
    if len(payload) != 2 * n:
        raise ValueError("checkpoint payload has the wrong size")
    if verify and fnv1a64(payload) != hd["checksum"]:
        raise ValueError("checkpoint checksum mismatch")
    hd["types"] = np.frombuffer(payload, dtype=np.uint8, count=n).copy()
    hd["state"] = np.frombuffer(payload, dtype=np.uint8, offset=n, count=n).copy()
    return hd
