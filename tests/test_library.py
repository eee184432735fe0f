"""CPU checks of the product boundary: librsfg.so loads and exports every
entry point include/rsfg.h declares; the Python mirror matches the header."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "rsfg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rsfg_[a-z_]+)\s*\(", txt)))


def test_header_declarations_are_exported():
    from interact_amd import lib
    L = lib.load()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert sorted(lib.EXPORTS) == names


def test_version_string():
    from interact_amd import lib
    assert b"gfx950" in lib.load().rsfg_version()


def test_struct_layouts_match_header():
    """ctypes mirrors of the C structs have the C sizes (checked against a
    tiny C translation unit compiled from the header)."""
    import subprocess
    import tempfile
    from interact_amd import lib
    src = r'''
#include <stdio.h>
#include "rsfg.h"
int main(void){printf("%zu %zu %zu %zu %zu\n", sizeof(rsfg_patch),
 sizeof(rsfg_assembly), sizeof(rsfg_friction), sizeof(rsfg_solver_opts),
 sizeof(rsfg_stats));return 0;}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        sizes = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    py = [ctypes.sizeof(t) for t in (lib.Patch, lib.Assembly, lib.Friction, lib.SolverOpts, lib.Stats)]
    assert sizes == py


def test_read_table_matches_the_reference_inputs(tmp_path):
    """rsf_solve's per-patch tables through the library's reader (host code,
    no GPU): the BP5 2 km rsf ("a b"), ic ("tau v") and dc files as held by
    the reference (seas_bp_tests/bp5); a short file names the failing row."""
    import numpy as np
    from interact_amd.model import read_table
    from conftest import GOLDEN
    d = os.path.join(GOLDEN, "bp5_2km")
    for name, ncols in (("rsf_bp5_2km.dat", 2), ("ic_bp5_2km.in", 2), ("dc_bp5_2km.in", 1)):
        ref = np.loadtxt(os.path.join(d, name)).reshape(1000, ncols)
        got = read_table(os.path.join(d, name), 1000, ncols)
        assert np.array_equal(got, ref), name
    short = tmp_path / "short.in"
    short.write_text("0.01 0.015\n0.01\n")
    from interact_amd import lib
    import pytest
    with pytest.raises(lib.RsfgError, match="row 1 column 1"):
        read_table(str(short), 2, 2)


def test_read_patch_file_matches_the_oracle_reader(tmp_path):
    """The reference geometry format through the library's reader (host
    code, no GPU), src/read_geometry.c:115-117 and :148-215: the BP5 2 km
    rectangles as the oracle's reader parses them, and a triangle file (the
    rough Hurst-0.8 mesh) with its vertices; a capacity below the file's
    patch count is an error that reports the full count."""
    import ctypes as C
    import numpy as np
    import oracle
    from conftest import GOLDEN
    from interact_amd import faults, lib
    from interact_amd.model import read_patch_file
    import pytest
    f = os.path.join(GOLDEN, "bp5_2km", "geom_bp5_2km.in")
    got, ref = read_patch_file(f), oracle.read_patches(f)
    assert len(got["l"]) == len(ref["l"]) == 1000
    for k in ("xyz", "strike", "dip", "l", "w", "group"):
        assert np.array_equal(got[k], ref[k]), k
    assert not got["vert"].any()
    nx, nz = 6, 4
    off = faults.rough_offset(nx, nz, 1.0)
    tg = faults.planar_tri_grid(1.0, nx, nz, offset=off)
    tf = tmp_path / "tri.in"
    faults.write_patch_file(str(tf), tg)
    got = read_patch_file(str(tf))
    assert len(got["l"]) == 2 * nx * nz
    assert (got["l"] < 0).all() and (got["w"] < 0).all()
    assert np.array_equal(got["vert"], tg["vert"])
    L = lib.load()
    arr = (lib.Patch * 10)()
    n = C.c_int(10)
    assert L.rsfg_read_patch_file(os.fsencode(f), arr, C.byref(n)) == lib.RSFG_EINVAL
    assert n.value == 1000
    assert b"more than the capacity" in L.rsfg_last_error()
    assert arr[9].y == ref["xyz"][9, 1]
    # reading ends at the first line that does not parse as a patch, as the
    # reference's while (fscanf(...) == 8) loop does
    short = tmp_path / "short.in"
    short.write_text("0 0 -1000 0 90 500 500 0\n0 0 -2000 0 90 500\n")
    assert len(read_patch_file(str(short))["l"]) == 1
    tri = tmp_path / "tri_short.in"
    tri.write_text("0 0 -1000 0 90 -1 -1 0 0 0 -1 0 1 -1\n")
    with pytest.raises(lib.RsfgError, match="truncated triangle"):
        read_patch_file(str(tri))
