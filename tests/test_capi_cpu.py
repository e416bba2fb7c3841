"""CPU-side checks of the engine's boundary: the C-ABI library loads and exports
every entry point include/lgca.h declares, the ctypes signatures cover them,
and the host-side mirror derives the reference's lattice sizes.  No compute
call is made here (there may be no GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lgca.h")


def declared_entry_points():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"LGCA_API\s+[\w\s\*]+?\b(lgca_\w+)\s*\(", text)))


def test_header_declares_the_expected_surface():
    names = declared_entry_points()
    for must in ("lgca_create", "lgca_destroy", "lgca_step", "lgca_post_process",
                 "lgca_set_cell_types", "lgca_set_state_u8", "lgca_get_state_u8",
                 "lgca_apply_body_force_count", "lgca_count_particles", "lgca_mean_velocity"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from lgca_amd import _capi
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("liblgca_hip.so not built")
    L = _capi.lib()
    for name in declared_entry_points():
        assert hasattr(L, name), name


def test_ctypes_signatures_cover_the_header():
    from lgca_amd import _capi
    names = set(declared_entry_points())
    bound = set(_capi.SIGNATURES) | set(_capi.TEXT_ENTRY_POINTS)
    assert names == bound


def test_library_refuses_to_load_when_missing(tmp_path, monkeypatch):
    from lgca_amd import _capi
    monkeypatch.setattr(_capi, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_capi, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _capi.lib()


@pytest.mark.parametrize("case,model,Re,Ma,r,dims", [
    # the reference viewers' set-ups (apps/*/ *_viewer.h) and the stale
    # periodic app's default
    ("pipe", "FHP_III", 80.0, 0.3, 10, (1480, 740)),
    ("karman", "FHP_III", 80.0, 0.3, 20, (4400, 2200)),
    ("diffusion", "FHP_III", 200.0, 0.2, 1, (202, 202)),
    ("collision", "FHP_III", 80.0, 0.2, 1, (21, 10)),
    ("periodic", "HPP", 240.0, 0.3, 8, (256, 256)),
])
def test_reference_size_derivation(case, model, Re, Ma, r, dims):
    from lgca_amd.lattice import derive_parameters
    p = derive_parameters(model, case, Re, Ma, r)
    assert (p["dim_x"], p["dim_y"]) == dims
    assert p["bf_dir"] == ("x" if case in ("pipe", "karman") else None)


def test_disc_matches_float_rounded_distance():
    from lgca_amd.lattice import disc
    m = disc(64, 32, 20, 16, 11.0).reshape(32, 64)
    ys, xs = np.nonzero(m)
    d = np.sqrt((xs - 20.0) ** 2 + (ys - 16.0) ** 2)
    assert (d < 5.5).all()
    assert m.sum() == sum(1 for y in range(32) for x in range(64)
                          if np.float32(np.sqrt((x - 20) ** 2 + (y - 16) ** 2)) < 5.5)


def test_cpp_dropin_compiles_against_the_c_abi():
    """The C++ drop-in header and the viewer-loop program compile with a plain
    host compiler against include/lgca.h alone (no HIP headers)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = os.path.join(ROOT, "tests", "native", "viewer_loop.cpp")
    r = subprocess.run([gxx, "-std=c++11", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "lgca_amd", "host"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_ctypes_structures_match_the_c_layout(tmp_path):
    """sizeof / offsetof of every struct in include/lgca.h, from a C compiler,
    equal the ctypes mirrors' (a mismatch would corrupt arguments silently)."""
    import shutil
    import subprocess
    from lgca_amd import _capi
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    structs = {"lgca_info": _capi.Info, "lgca_job_desc": _capi.JobDesc,
               "lgca_job_info": _capi.JobInfo, "lgca_batch_stats": _capi.BatchStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "lgca.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    r = subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in filter(None, out):
        cname, field, value = line.split()
        py = structs[cname]
        want = C.sizeof(py) if field == "size" else getattr(py, field).offset
        assert int(value) == want, line



def test_batch_refuses_force_regions_past_one_descriptor():
    """A pass reads its 8 steps of force decisions through one buffer
    descriptor: the batch refuses forced regions whose decisions of an
    8-step pass reach 2 GiB (rows x 32-cell words x 8 steps x 4 bytes), and
    accepts the same lattice with a smaller region."""
    from lgca_amd.batch import Job, job_bytes
    dy = 65534
    ok = Job("FHP_II", 32 * 1024, dy, 10, force_dir="x", force_p=0.01)   # 2,147,418,112 bytes
    assert job_bytes(ok) > 0
    big = Job("FHP_II", 32 * 1025, dy, 10, force_dir="x", force_p=0.01)  # 2,149,515,520 bytes
    with pytest.raises(ValueError):
        job_bytes(big)
    small = Job("FHP_II", 32 * 1025, dy, 10, force_dir="x", force_p=0.01, force_region=(0, 64, 0, dy))
    assert job_bytes(small) > 0
