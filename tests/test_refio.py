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


# ------------------------------------------------------------------ uvwMap
UVW_MAP = """5
-23 -288 -166 4
0 0 0 2
7 -1 3 8
this line is not a cell
7 -1 3 5
12 40 -9
"""


def test_uvw_map_weights_by_hand(tmp_path):
    """vissGen1M_shade_accurate.cu:248-289, 352-368 worked by hand: the header
    line is the cell count; "7 -1 3" appears twice and the later count (5) wins;
    the prose line and the three-field line are skipped; rows are truncated
    toward zero after rounding to FP32; a row without a cell weighs 1."""
    p = tmp_path / "uvwMap130.txt"
    p.write_text(UVW_MAP)
    rows = np.array([[-23, -288, -166],     # 1/4
                     [0.9, -0.7, 0.2],      # truncates to 0 0 0: 1/2
                     [7.99, -1.5, 3.0],     # 7 -1 3: 1/5 (the later line)
                     [12, 40, -9],          # only in the three-field line: 1
                     [-23, -288, -167],     # absent: 1
                     [-0.5, 0, 0]],         # -0.5 truncates to 0: 1/2
                    np.float64)
    w, info = refio.uvw_map_weights(str(p), rows)
    assert w.dtype == np.float32
    assert w.tolist() == [np.float32(1 / 4), np.float32(1 / 2), np.float32(1 / 5), 1.0, 1.0,
                          np.float32(1 / 2)]
    assert info == {"cells": 3, "skipped": 2, "found": 4}


def test_uvw_map_round_trip_with_gridded_counts(tmp_path):
    rng = np.random.default_rng(5)
    cells = np.unique(rng.integers(-500, 501, (300, 3)), axis=0).astype(np.float64)
    counts = rng.integers(1, 40, len(cells)).astype(np.int32)
    p = os.path.join(str(tmp_path), "frequency_1M", "uvwMap130.txt")
    refio.write_uvw_map(p, cells, counts)
    lines = open(p).read().splitlines()
    assert lines[0] == str(len(cells)) and len(lines) == len(cells) + 1
    assert lines[1] == " ".join(str(int(x)) for x in cells[0]) + f" {counts[0]}"
    rows = np.concatenate([cells, cells[::7] + 1000])          # the shifted rows are absent
    w, info = refio.uvw_map_weights(p, rows)
    assert np.array_equal(w[:len(cells)], (1.0 / counts.astype(np.float32)).astype(np.float32))
    assert np.all(w[len(cells):] == 1.0)
    assert info == {"cells": len(cells), "skipped": 0, "found": len(cells)}
    assert refio.paths(str(tmp_path), 1e6, 1)["uvwMap"] == p


def test_uvw_map_errors(tmp_path):
    rows = np.zeros((2, 3))
    with pytest.raises(vrs.VrsError, match="VRS_E_IO"):
        refio.uvw_map_weights(str(tmp_path / "missing.txt"), rows)
    bad = tmp_path / "nohead.txt"
    bad.write_text("cells\n1 2 3 4\n")
    with pytest.raises(vrs.VrsError, match="VRS_E_IO"):
        refio.uvw_map_weights(str(bad), rows)
    zero = tmp_path / "zero.txt"
    zero.write_text("1\n0 0 0 0\n")
    with pytest.raises(vrs.VrsError, match="VRS_E_ARG"):
        refio.uvw_map_weights(str(zero), rows)
    ok = tmp_path / "ok.txt"
    ok.write_text("1\n0 0 0 3\n")
    with pytest.raises(vrs.VrsError, match="VRS_E_ARG"):
        refio.uvw_map_weights(str(ok), np.array([[3e9, 0, 0]]))
    with pytest.raises(vrs.VrsError, match="VRS_E_ARG"):
        refio.uvw_map_weights(str(ok), np.array([[np.nan, 0, 0]]))
    with pytest.raises(vrs.VrsError, match="VRS_E_ARG"):
        refio.write_uvw_map(str(tmp_path / "w.txt"), rows, [1, 0])
    w, info = refio.uvw_map_weights(str(ok), np.zeros((0, 3)))
    assert w.shape == (0,) and info["cells"] == 1
