"""The reference pipeline's file layout: names and text formats of the
inter-stage files, so that the library's outputs can replace the reference's
file for file.

Names (directory per frequency, one file per period p, 1-based):
  frequency_{F}M/uvw{p}frequency{F}M.txt            MIncline_generate_uvw_float.cu:421
  frequency_{F}M/updated_uvw{p}frequency{F}M.txt    read at vissGen10M_shade.cu:351
  frequency_{F}M/shadeM{p}frequency{F}M.txt         shade10M.cu:269, vissGen10M_shade.cu:377
  F_recon_{F}M/F{p}period{F}M.txt                   vissGen10M_shade.cu:573
  F_recon_{F}M_accu/Ex{p}frequency{F}M.txt          vissGen1M_shade_accurate.cu:573-582
  frequency_{F}M/V{p}frequency{F}M.bin              binary complex64 (no reference file)
  frequency_1M/uvwMap{P}.txt                        cell counts over P periods, read at
                                                    vissGen1M_shade_accurate.cu:252
The readers and writers are the C-ABI's (vrs_read_* / vrs_write_*).
"""
import ctypes as C
import os

import numpy as np

from . import _lib


def _mhz(freq_hz):
    f = freq_hz / 1e6
    return str(int(round(f))) if abs(f - round(f)) < 1e-9 else ("%g" % f)


def paths(root, freq_hz, period):
    f = _mhz(freq_hz)
    d = os.path.join(root, f"frequency_{f}M")
    return {
        "uvw": os.path.join(d, f"uvw{period}frequency{f}M.txt"),
        "updated_uvw": os.path.join(d, f"updated_uvw{period}frequency{f}M.txt"),
        "shadeM": os.path.join(d, f"shadeM{period}frequency{f}M.txt"),
        "F": os.path.join(root, f"F_recon_{f}M", f"F{period}period{f}M.txt"),
        "Ex": os.path.join(root, f"F_recon_{f}M_accu", f"Ex{period}frequency{f}M.txt"),
        "V": os.path.join(d, f"V{period}frequency{f}M.bin"),
        "uvwMap": os.path.join(d, "uvwMap130.txt"),
    }


def _mk(path):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    return path.encode()


def write_uvw(path, uvw, counts=None):
    uvw = np.ascontiguousarray(uvw, np.float64).reshape(-1, 3)
    c = None if counts is None else np.ascontiguousarray(counts, np.float32)
    _lib.check(_lib.load().vrs_write_uvw_txt(_mk(path), uvw.ctypes.data, len(uvw),
                                             None if c is None else c.ctypes.data))


def read_uvw(path, n_cols=3, cap=1 << 26):
    uvw = np.empty((cap, 3), np.float64) if cap <= (1 << 20) else None
    if uvw is None:  # size from the line count first
        with open(path) as f:
            cap = sum(1 for _ in f)
        uvw = np.empty((max(cap, 1), 3), np.float64)
    col4 = np.empty(len(uvw), np.float32)
    n = C.c_int64()
    _lib.check(_lib.load().vrs_read_uvw_txt(path.encode(), n_cols, uvw.ctypes.data,
                                            col4.ctypes.data, len(uvw), C.byref(n)))
    return (uvw[:n.value], col4[:n.value]) if n_cols == 4 else uvw[:n.value]


def uvw_map_weights(map_path, uvw):
    """Weights 1/count of each uvw row's cell in a uvwMap table (1 for a cell
    the table lacks), as vissGen1M_shade_accurate.cu:248-289, 352-368.
    Returns (w float32 [n], info) with info = cells, skipped lines, rows found."""
    uvw = np.ascontiguousarray(uvw, np.float64).reshape(-1, 3)
    w = np.empty(len(uvw), np.float32)
    cells, skipped, found = C.c_int64(), C.c_int64(), C.c_int64()
    _lib.check(_lib.load().vrs_uvw_map_weights(
        map_path.encode(), uvw.ctypes.data, len(uvw), w.ctypes.data, C.byref(cells),
        C.byref(skipped), C.byref(found)))
    return w, {"cells": cells.value, "skipped": skipped.value, "found": found.value}


def write_uvw_map(path, uvw, counts):
    """A uvwMap table: the number of cells, then "u v w count" per cell."""
    uvw = np.ascontiguousarray(uvw, np.float64).reshape(-1, 3)
    counts = np.ascontiguousarray(counts, np.int32).reshape(-1)
    if len(counts) != len(uvw):
        raise ValueError(f"counts has {len(counts)} entries for {len(uvw)} rows")
    _lib.check(_lib.load().vrs_write_uvw_map_txt(_mk(path), uvw.ctypes.data, counts.ctypes.data,
                                                 len(uvw)))


def write_shade(path, box4):
    box4 = np.ascontiguousarray(box4, np.int32).reshape(-1, 4)
    _lib.check(_lib.load().vrs_write_shade_txt(_mk(path), box4.ctypes.data, len(box4)))


def read_shade(path, cap=None):
    if cap is None:
        with open(path) as f:
            cap = max(sum(1 for _ in f), 1)
    out = np.empty((cap, 4), np.int32)
    n = C.c_int64()
    _lib.check(_lib.load().vrs_read_shade_txt(path.encode(), out.ctypes.data, cap, C.byref(n)))
    return out[:n.value]


def write_image(path, F):
    F = np.ascontiguousarray(F, np.float32).reshape(-1)
    _lib.check(_lib.load().vrs_write_image_txt(_mk(path), F.ctypes.data, len(F)))


def read_image(path, cap=None):
    if cap is None:
        with open(path) as f:
            cap = max(sum(1 for _ in f), 1)
    out = np.empty(cap, np.float32)
    n = C.c_int64()
    _lib.check(_lib.load().vrs_read_image_txt(path.encode(), out.ctypes.data, cap, C.byref(n)))
    return out[:n.value]


def write_ex(path, F, NX, one_based=True):
    """One line per fine cell: F of its parent pixel (NX 1-based by default)."""
    F = np.ascontiguousarray(F, np.float32).reshape(-1)
    NX = np.ascontiguousarray(NX, np.int32).reshape(-1)
    _lib.check(_lib.load().vrs_write_ex_txt(_mk(path), F.ctypes.data, len(F), NX.ctypes.data,
                                            len(NX), 1 if one_based else 0))


def write_vis(path, V):
    """V [maps][n_vis] complex (or [..., 2] float32) -> binary complex64 file."""
    V = np.asarray(V)
    if np.iscomplexobj(V):
        V = np.stack([V.real, V.imag], -1)
    V = np.ascontiguousarray(V, np.float32)
    if V.ndim == 2:
        V = V[None]
    _lib.check(_lib.load().vrs_write_vis_bin(_mk(path), V.ctypes.data, V.shape[0], V.shape[1]))


def read_vis(path):
    """Binary complex64 visibilities -> complex64 [maps][n_vis]."""
    lib = _lib.load()
    nm, nv = C.c_int64(), C.c_int64()
    _lib.check(lib.vrs_read_vis_bin(path.encode(), None, 0, C.byref(nm), C.byref(nv)))
    out = np.empty((nm.value, nv.value, 2), np.float32)
    _lib.check(lib.vrs_read_vis_bin(path.encode(), out.ctypes.data, nm.value * nv.value,
                                    C.byref(nm), C.byref(nv)))
    return out[..., 0] + 1j * out[..., 1]
