"""Reader and writer of the engine's checkpoint files (lgca_save_checkpoint /
lgca_load_checkpoint in include/lgca.h), usable without a GPU.

Layout, little-endian: the 8-byte magic b"LGCACKP1"; u32 version, i32 model,
u32 dim_x, dim_y, coarse_radius, flags (bit 0: controller state present);
u64 seed, step, force_calls; i32 force_dir, u32 reserved; u64 force_thr;
u32 fx0, fx1, fy0, fy1; u64 checksum.  Version 2 continues the header with
i32 controller direction, f32 controller u, i64 equilibrium forcing, current
forcing, last tick's amount and reverted particles, and f32 mean velocity
(x, y) the last tick saw.  The payload follows: num_cells cell-type bytes,
num_cells state bytes (bit d = direction d, the reference's Bitset layout,
src/lgca_bitset.h:229-231) and, in version 2, num_cells bytes of the output
snapshot.  The checksum is FNV-1a-64 of the header (checksum field zero) and
then the payload; version 1 files hash the payload only.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"LGCACKP1"
HEADER = struct.Struct("<8sIiIIIIQQQiIQIIIIQ")
assert HEADER.size == 96
HEADER2 = struct.Struct("<ifqqqqff")
assert HEADER2.size == 48

FIELDS = ("magic", "version", "model", "dim_x", "dim_y", "coarse_radius", "flags", "seed",
          "step", "force_calls", "force_dir", "reserved1", "force_thr", "fx0", "fx1", "fy0", "fy1",
          "checksum")
FIELDS2 = ("ctrl_dir", "ctrl_u", "ctrl_eq", "ctrl_forcing", "ctrl_amount", "ctrl_reverted",
           "ctrl_vx", "ctrl_vy")
FLAG_CONTROLLER = 1


def fnv1a64(data: bytes, h: int = 0xCBF29CE484222325) -> int:
    """FNV-1a, 64-bit, continuing from hash h.  The hash is sequential, so this
    runs at Python speed; read_checkpoint(verify=False) skips it for large
    files."""
    h = np.uint64(h)
    prime = np.uint64(0x100000001B3)
    with np.errstate(over="ignore"):
        for b in np.frombuffer(data, dtype=np.uint8):
            h = (h ^ np.uint64(b)) * prime
    return int(h)


def _header_bytes(hd: dict, checksum: int) -> bytes:
    vals = [hd[k] for k in FIELDS[:-1]] + [checksum]
    raw = HEADER.pack(*vals)
    if hd["version"] == 2:
        raw += HEADER2.pack(*[hd.get(k, 0) for k in FIELDS2])
    return raw


def checksum_of(hd: dict, payload: bytes) -> int:
    """The checksum a file with header fields hd and this payload carries."""
    if hd["version"] == 1:
        return fnv1a64(payload)
    return fnv1a64(payload, fnv1a64(_header_bytes(hd, 0)))


def read_checkpoint(path: str, verify: bool = True) -> dict:
    """Header fields plus `types`, `state` and `snapshot` (uint8, num_cells
    each; a version 1 file's snapshot is its state)."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < HEADER.size:
        raise ValueError("not a checkpoint: too short")
    hd = dict(zip(FIELDS, HEADER.unpack_from(raw)))
    if hd["magic"] != MAGIC or hd["version"] not in (1, 2):
        raise ValueError("not a checkpoint: bad magic or version")
    off = HEADER.size
    if hd["version"] == 2:
        if len(raw) < off + HEADER2.size:
            raise ValueError("checkpoint payload has the wrong size")
        hd.update(zip(FIELDS2, HEADER2.unpack_from(raw, off)))
        off += HEADER2.size
    else:
        hd["flags"] = 0
    n = hd["dim_x"] * hd["dim_y"]
    parts = 3 if hd["version"] == 2 else 2
    payload = raw[off:]
    if len(payload) != parts * n:
        raise ValueError("checkpoint payload has the wrong size")
    if verify and checksum_of(hd, payload) != hd["checksum"]:
        raise ValueError("checkpoint checksum mismatch")
    hd["types"] = np.frombuffer(payload, dtype=np.uint8, count=n).copy()
    hd["state"] = np.frombuffer(payload, dtype=np.uint8, offset=n, count=n).copy()
    hd["snapshot"] = (np.frombuffer(payload, dtype=np.uint8, offset=2 * n, count=n).copy()
                      if parts == 3 else hd["state"].copy())
    hd["controller"] = bool(hd["flags"] & FLAG_CONTROLLER)
    return hd


def write_checkpoint(path: str, hd: dict, types: np.ndarray, state: np.ndarray,
                     snapshot: np.ndarray | None = None) -> None:
    """Writes a file with header fields hd (FIELDS, and FIELDS2 for version 2;
    the checksum is computed here) and the given bytes; version 2 needs the
    snapshot."""
    payload = np.ascontiguousarray(types, np.uint8).tobytes() + \
        np.ascontiguousarray(state, np.uint8).tobytes()
    if hd["version"] == 2:
        payload += np.ascontiguousarray(snapshot, np.uint8).tobytes()
    with open(path, "wb") as f:
        f.write(_header_bytes(hd, checksum_of(hd, payload)) + payload)
