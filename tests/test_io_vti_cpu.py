"""The VTK-free .vti writer (lgca_amd/host/io_vti.h and lgca_amd/io_vti.py).

The files carry what the reference's IoVti writes (src/lgca_io_vti.cpp:50-144):
image sizes dim+1 points for the cell file and coarse-size points for the mean
file, the four arrays by name, component count and order, the active scalars.
The C++ and Python writers must produce identical bytes, and the XML must be
well formed with the appended block where its offsets say."""
import os
import shutil
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from conftest import ROOT
from lgca_amd.io_vti import read_vti, write_vti


class Fake:
    def dim_x(self): return 6
    def dim_y(self): return 4
    def coarse_dim_x(self): return 3
    def coarse_dim_y(self): return 2
    def cell_density(self): return 0.5 * np.arange(24, dtype=np.float32)
    def cell_momentum(self): return 0.25 * np.arange(48, dtype=np.float32) - 3.0
    def mean_density(self): return (1.0 / np.arange(1, 7, dtype=np.float32)).astype(np.float32)
    def mean_momentum(self): return -0.125 * np.arange(12, dtype=np.float32)


def test_python_writer_layout(tmp_path):
    from lgca_amd.io_vti import write_lattice_vti
    cell, mean = write_lattice_vti(Fake(), 7, str(tmp_path))
    c = read_vti(cell)
    assert c["extent"] == (0, 6, 0, 4, 0, 0) and c["section"] == "CellData"
    assert c["scalars"] is None  # "Mean density" is not a cell array
    assert list(c["arrays"]) == ["Cell density", "Cell momentum"]
    assert np.array_equal(c["arrays"]["Cell density"][:, 0], Fake().cell_density())
    assert c["arrays"]["Cell momentum"].shape == (24, 2)
    m = read_vti(mean)
    assert m["extent"] == (0, 2, 0, 1, 0, 0) and m["section"] == "PointData"
    assert m["scalars"] == "Mean density"
    assert np.array_equal(m["arrays"]["Mean momentum"].reshape(-1), Fake().mean_momentum())
    # the XML part parses, and the appended block starts with the first array's header
    raw = open(mean, "rb").read()
    cut = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    ET.fromstring(raw[:cut - 1].decode() + "</AppendedData></VTKFile>")
    assert int(np.frombuffer(raw, dtype=np.uint64, count=1, offset=cut)[0]) == 6 * 4


def _writer(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "vti_writer"
    r = subprocess.run([gxx, "-std=c++11", "-Wall", "-Wextra", "-Werror", "-O1",
                        "-I", os.path.join(ROOT, "lgca_amd", "host"),
                        os.path.join(ROOT, "tests", "native", "vti_writer.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _have_liblz4():
    try:
        from lgca_amd.io_vti import lz4_decompress_block
        lz4_decompress_block(b"\x10a", 1)
        return True
    except OSError:
        return False


def test_cpp_and_python_writers_agree(tmp_path):
    exe = _writer(tmp_path)
    out = tmp_path / "cpp"
    out.mkdir()
    subprocess.run([str(exe), str(out), "raw"], check=True)
    from lgca_amd.io_vti import write_lattice_vti
    py = tmp_path / "py"
    py.mkdir()
    write_lattice_vti(Fake(), 7, str(py))
    for name in ("cell_res_7.vti", "mean_res_7.vti"):
        assert (out / name).read_bytes() == (py / name).read_bytes(), name


def test_empty_scalars_and_single_component(tmp_path):
    p = tmp_path / "x.vti"
    write_vti(str(p), 2, 2, False, [("a", np.array([1.0, 2.0, 3.0, 4.0], np.float32), 1)], "")
    v = read_vti(str(p))
    assert v["scalars"] is None and v["arrays"]["a"].shape == (4, 1)


def test_lz4_blocks_decode_with_the_reference_decoder(tmp_path):
    """Every block the encoder writes decodes, with the system's liblz4, to the
    bytes it encoded: sizes 0..40, random, zero, periodic and float data."""
    if not _have_liblz4():
        pytest.skip("liblz4 not available")
    from lgca_amd.io_vti import lz4_decompress_block
    exe = _writer(tmp_path)
    subprocess.run([str(exe), str(tmp_path), "blocks"], check=True)
    raw = (tmp_path / "blocks.bin").read_bytes()
    pos, count, saved = 0, 0, 0
    while pos < len(raw):
        n, c = (int(v) for v in np.frombuffer(raw, dtype=np.uint64, count=2, offset=pos))
        pos += 16
        src = raw[pos:pos + n]
        pos += n
        blk = raw[pos:pos + c]
        pos += c
        assert lz4_decompress_block(blk, n) == src, (count, n, c)
        saved += n - c
        count += 1
    assert count == 46 and saved > 32768 * 2


def test_lz4_files_match_the_raw_arrays(tmp_path):
    if not _have_liblz4():
        pytest.skip("liblz4 not available")
    exe = _writer(tmp_path)
    for mode in ("raw", "lz4"):
        (tmp_path / mode).mkdir()
        subprocess.run([str(exe), str(tmp_path / mode), mode], check=True)
    for name in ("cell_res_7.vti", "mean_res_7.vti"):
        a = read_vti(str(tmp_path / "raw" / name))
        b = read_vti(str(tmp_path / "lz4" / name))
        assert a["compressor"] is None and b["compressor"] == "vtkLZ4DataCompressor"
        assert a["extent"] == b["extent"] and a["scalars"] == b["scalars"]
        assert list(a["arrays"]) == list(b["arrays"])
        for k in a["arrays"]:
            assert np.array_equal(a["arrays"][k], b["arrays"][k]), k
