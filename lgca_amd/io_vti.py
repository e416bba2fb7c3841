"""VTK XML image files (.vti) of a lattice's output arrays, without VTK.

The Python twin of lgca_amd/host/io_vti.h (uncompressed, the two write
identical bytes).
`write_lattice_vti` produces the pair of files the reference's IoVti<Model>
writes (src/lgca_io_vti.cpp:113-144): cell_res_<step>.vti with
(dim_x + 1) x (dim_y + 1) points and CellData "Cell density" / "Cell
momentum", and mean_res_<step>.vti with coarse_dim_x x coarse_dim_y points
and PointData "Mean density" / "Mean momentum" (src/lgca_io_vti.cpp:50-87).
Arrays are raw appended binary with UInt64 byte-count headers (the C++ writer
also has the reference's LZ4 compression).  `read_vti` parses both back.
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np


def write_vti(path: str, points_x: int, points_y: int, cell_data: bool, arrays, scalars: str = "") -> None:
    """arrays: list of (name, float32 array, components)."""
    if sys.byteorder != "little":
        raise RuntimeError("the files are written little-endian")
    extent = f"0 {points_x - 1} 0 {points_y - 1} 0 0"
    section = "CellData" if cell_data else "PointData"
    head = ['<?xml version="1.0"?>\n',
            '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
            f'  <ImageData WholeExtent="{extent}" Origin="0 0 0" Spacing="1 1 1">\n',
            f'    <Piece Extent="{extent}">\n',
            f"      <{section}"]
    if any(name == scalars for name, _, _ in arrays):
        head.append(f' Scalars="{scalars}"')
    head.append(">\n")
    blobs = []
    offset = 0
    for name, data, comps in arrays:
        raw = np.ascontiguousarray(data, dtype=np.float32).tobytes()
        head.append(f'        <DataArray type="Float32" Name="{name}" NumberOfComponents="{comps}" '
                    f'format="appended" offset="{offset}"/>\n')
        blobs.append(np.uint64(len(raw)).tobytes() + raw)
        offset += 8 + len(raw)
    head += [f"      </{section}>\n", "    </Piece>\n", "  </ImageData>\n",
             '  <AppendedData encoding="raw">\n   _']
    with open(path, "wb") as f:
        f.write("".join(head).encode())
        for b in blobs:
            f.write(b)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")


def write_lattice_vti(lattice, step: int, directory: str, scalars: str = "Mean density") -> tuple:
    """IoVti::write(step, dir) for a lgca_amd.Lattice after post_process()."""
    cell = os.path.join(directory, f"cell_res_{step}.vti")
    mean = os.path.join(directory, f"mean_res_{step}.vti")
    write_vti(cell, lattice.dim_x() + 1, lattice.dim_y() + 1, True,
              [("Cell density", lattice.cell_density(), 1),
               ("Cell momentum", lattice.cell_momentum(), 2)], scalars)
    write_vti(mean, lattice.coarse_dim_x(), lattice.coarse_dim_y(), False,
              [("Mean density", lattice.mean_density(), 1),
               ("Mean momentum", lattice.mean_momentum(), 2)], scalars)
    return cell, mean


_liblz4 = None


def lz4_decompress_block(data: bytes, size: int) -> bytes:
    """One LZ4 block, decoded by the system's liblz4 (an independent decoder
    of what lgca_amd/host/io_vti.h encodes)."""
    global _liblz4
    if _liblz4 is None:
        import ctypes
        import ctypes.util
        name = ctypes.util.find_library("lz4") or "liblz4.so.1"
        _liblz4 = ctypes.CDLL(name)
        _liblz4.LZ4_decompress_safe.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                                 ctypes.c_int]
        _liblz4.LZ4_decompress_safe.restype = ctypes.c_int
    import ctypes
    out = ctypes.create_string_buffer(max(size, 1))
    got = _liblz4.LZ4_decompress_safe(data, out, len(data), size)
    if got != size:
        raise ValueError(f"LZ4 block decoded to {got} bytes, expected {size}")
    return out.raw[:size]


def read_vti(path: str) -> dict:
    """{'extent': (x0, x1, y0, y1, z0, z1), 'section': 'CellData'|'PointData',
    'scalars': name or None, 'arrays': {name: float32 array (tuples, comps)},
    'compressor': name or None}.  Reads raw or LZ4-compressed appended data."""
    raw = open(path, "rb").read()
    cut = raw.index(b'<AppendedData encoding="raw">')
    start = raw.index(b"_", cut) + 1
    xml = raw[:cut].decode()
    comp = re.search(r'compressor="([^"]+)"', xml)
    if comp and comp.group(1) != "vtkLZ4DataCompressor":
        raise ValueError(f"unsupported compressor {comp.group(1)}")
    ext = tuple(int(v) for v in re.search(r'WholeExtent="([^"]+)"', xml).group(1).split())
    sec = re.search(r"<(CellData|PointData)([^>]*)>", xml)
    scal = re.search(r'Scalars="([^"]+)"', sec.group(2))
    arrays = {}
    for m in re.finditer(r'<DataArray type="Float32" Name="([^"]+)" NumberOfComponents="(\d+)" '
                         r'format="appended" offset="(\d+)"/>', xml):
        name, comps, off = m.group(1), int(m.group(2)), int(m.group(3))
        pos = start + off
        if comp:
            nb, bs, last = (int(v) for v in np.frombuffer(raw, dtype=np.uint64, count=3, offset=pos))
            sizes = np.frombuffer(raw, dtype=np.uint64, count=nb, offset=pos + 24)
            p = pos + 24 + 8 * nb
            parts = []
            for i in range(nb):
                usize = last if (i == nb - 1 and last) else bs
                parts.append(lz4_decompress_block(raw[p:p + int(sizes[i])], usize))
                p += int(sizes[i])
            a = np.frombuffer(b"".join(parts), dtype=np.float32)
        else:
            n = int(np.frombuffer(raw, dtype=np.uint64, count=1, offset=pos)[0])
            a = np.frombuffer(raw, dtype=np.float32, count=n // 4, offset=pos + 8)
        arrays[name] = a.reshape(-1, comps)
    return {"extent": ext, "section": sec.group(1), "scalars": scal.group(1) if scal else None,
            "arrays": arrays, "compressor": comp.group(1) if comp else None}
