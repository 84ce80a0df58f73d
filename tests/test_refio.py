"""Reference text formats (C-ABI writers/readers, host only).

The layout is the reference's (one record per line, space separated) and the
numbers use C++'s default ostream formatting, which the C++ standard fixes as
printf's %g with precision 6 -- the expected text below is produced with
Python's "%g" for that reason."""
import os

import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import refio


def test_reference_paths():
    p = refio.paths("/x", 1e7, 3)
    assert p["uvw"] == "/x/frequency_10M/uvw3frequency10M.txt"
    assert p["updated_uvw"] == "/x/frequency_10M/updated_uvw3frequency10M.txt"
    assert p["shadeM"] == "/x/frequency_10M/shadeM3frequency10M.txt"
    assert p["F"] == "/x/F_recon_10M/F3period10M.txt"
    assert refio.paths("/x", 1e6, 130)["F"] == "/x/F_recon_1M/F130period1M.txt"


def test_uvw_round_trip_and_format(tmp_path):
    uvw = np.array([[1.0, -2.5, 3.25], [123456.789, 1e-5, -0.0], [7.0, 8.0, 9.0]])
    p = str(tmp_path / "frequency_10M" / "uvw1frequency10M.txt")
    refio.write_uvw(p, uvw)
    lines = open(p).read().splitlines()
    f32 = uvw.astype(np.float32)
    assert lines == ["%g %g %g" % tuple(float(x) for x in r) for r in f32]
    back = refio.read_uvw(p)
    assert np.array_equal(back, np.array([[float("%g" % x) for x in r] for r in f32], np.float32))
    # 4-column updated_uvw with counts
    refio.write_uvw(p, uvw, counts=np.array([1, 5, 2], np.float32))
    assert open(p).read().splitlines()[1].endswith(" 5")
    u4, c4 = refio.read_uvw(p, n_cols=4)
    assert list(c4) == [1, 5, 2] and len(u4) == 3


def test_shade_round_trip_normalises_like_the_reference_reader(tmp_path):
    box = np.array([[10, 3, 7, 9], [1, 20940, 5, 2]], np.int32)
    p = str(tmp_path / "shadeM1frequency10M.txt")
    refio.write_shade(p, box)
    assert open(p).read().splitlines() == ["10 3 7 9", "1 20940 5 2"]
    back = refio.read_shade(p)
    # ordered so that S <= E, then 0-based (vissGen10M_shade.cu:385-399)
    assert back.tolist() == [[2, 9, 6, 8], [0, 20939, 1, 4]]


def test_image_round_trip(tmp_path):
    F = np.array([0.5, -1.25e-7, 3.0e12, 42.0], np.float32)
    p = str(tmp_path / "F_recon_1M" / "F1period1M.txt")
    refio.write_image(p, F)
    assert open(p).read().splitlines() == ["%g" % float(x) for x in F]
    assert np.array_equal(refio.read_image(p), np.array(["%g" % float(x) for x in F], np.float32))


def test_io_errors(tmp_path):
    with pytest.raises(vrs.VrsError) as e:
        refio.read_image(str(tmp_path / "missing.txt"), cap=4)
    assert e.value.code == "VRS_E_IO"
    p = str(tmp_path / "img.txt")
    refio.write_image(p, np.arange(10, dtype=np.float32))
    with pytest.raises(vrs.VrsError) as e:        # more rows than the capacity
        refio.read_image(p, cap=5)
    assert e.value.code == "VRS_E_ARG"
    q = str(tmp_path / "trunc.txt")
    with open(q, "w") as f:
        f.write("1 2 3 4\n5 6 7\n")
    with pytest.raises(vrs.VrsError) as e:        # truncated 4-column row
        refio.read_uvw(q, n_cols=4, cap=8)
    assert e.value.code == "VRS_E_IO"
    with pytest.raises(vrs.VrsError):
        refio.write_image(str(tmp_path / "no" / "\0bad"), np.zeros(2, np.float32))


def test_ex_expansion_to_fine_cells(tmp_path):
    # Ex file: one line per fine cell with F of its parent pixel
    # (vissGen1M_shade_accurate.cu:567-588, NX 1-based)
    F = np.array([1.5, -2.0, 3.0e-3], np.float32)
    NX = np.array([3, 1, 1, 2, 3], np.int32)
    p = refio.paths(str(tmp_path), 1e6, 2)["Ex"]
    assert p.endswith("F_recon_1M_accu/Ex2frequency1M.txt")
    refio.write_ex(p, F, NX)
    assert open(p).read().splitlines() == ["%g" % float(F[q - 1]) for q in NX]
    with pytest.raises(vrs.VrsError) as e:
        refio.write_ex(p, F, np.array([4], np.int32))      # parent out of range
    assert e.value.code == "VRS_E_ARG"


def test_binary_visibilities_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    V = (rng.standard_normal((3, 17)) + 1j * rng.standard_normal((3, 17))).astype(np.complex64)
    p = refio.paths(str(tmp_path), 1e7, 1)["V"]
    refio.write_vis(p, V)
    assert os.path.getsize(p) == 24 + 3 * 17 * 8
    with open(p, "rb") as f:
        head = f.read(24)
    assert head[:4] == b"VRSV" and np.frombuffer(head[4:8], np.uint32)[0] == 1
    assert list(np.frombuffer(head[8:24], np.int64)) == [3, 17]
    assert np.array_equal(refio.read_vis(p), V)
    with open(p, "r+b") as f:                             # corrupt the magic
        f.write(b"XXXX")
    with pytest.raises(vrs.VrsError) as e:
        refio.read_vis(p)
    assert e.value.code == "VRS_E_IO"
