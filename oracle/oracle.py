"""ctypes front end of the FP64 oracle (oracle.c) and the FP32 CPU baseline
(cpu_f32_baseline.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product package never does.
Each wrapper names the reference function it restates (see oracle.c).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_lib = None
_base = None

_P = np.ctypeslib.ndpointer
_f64 = _P(dtype=np.float64, flags="C_CONTIGUOUS")
_f32 = _P(dtype=np.float32, flags="C_CONTIGUOUS")
_i32 = _P(dtype=np.int32, flags="C_CONTIGUOUS")
_u32 = _P(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64 = _P(dtype=np.int64, flags="C_CONTIGUOUS")


def build():
    """Compile the checker libraries with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib, _base
    if _lib is not None:
        return
    so = os.path.join(_BUILD, "liboracle.so")
    sb = os.path.join(_BUILD, "libcpubaseline.so")
    if not (os.path.exists(so) and os.path.exists(sb)):
        build()
    L = C.CDLL(so)
    L.vo_splitmix64.restype = C.c_uint64
    L.vo_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
    L.vo_normals.argtypes = [C.c_uint64, C.c_int64, _f64]
    L.vo_healpix_ring_vec.argtypes = [C.c_int64, C.c_int64, _f64]
    L.vo_healpix_nest_vec.argtypes = [C.c_int64, C.c_int64, _f64]
    L.vo_nest2ring.restype = C.c_int64
    L.vo_nest2ring.argtypes = [C.c_int64, C.c_int64]
    L.vo_orbit_positions.argtypes = [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double,
                                     C.c_double, _f64, _f64, _f64]
    L.vo_uvw.argtypes = [C.c_int64, _f64, C.c_int, _i32, C.c_double, _f64]
    L.vo_occ_one.restype = C.c_int
    L.vo_occ_one.argtypes = [_f64, _f64, C.c_double, C.POINTER(C.c_int)]
    L.vo_pix_hash.restype = C.c_uint32
    L.vo_pix_hash.argtypes = [C.c_int64]
    L.vo_mask_stats.argtypes = [C.c_int64, _f64, _i32, C.c_int64, _f64, C.c_double, C.c_int64,
                                C.c_void_p, _i32, _u32, C.POINTER(C.c_int64)]
    L.vo_mask_bits.argtypes = [C.c_int64, _f64, _i32, C.c_int64, _f64, C.c_double, C.c_int64,
                               C.c_void_p, _u32]
    L.vo_visgen.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, _f64, C.c_int64, _f64, C.c_int, _f32,
                            C.c_double, C.c_double, C.c_int, C.c_double, C.c_int64, C.c_void_p,
                            _f64]
    L.vo_recon.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_int, _f64, C.c_int64, _f64, C.c_int,
                           _f64, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_double,
                           C.c_int64, C.c_void_p, _f64]
    L.vo_grid_keys.restype = C.c_int64
    L.vo_grid_keys.argtypes = [C.c_int64, _f64, C.c_int, _i32, C.c_double,
                               _P(dtype=np.uint64, flags="C_CONTIGUOUS")]
    L.vo_grid_unique.restype = C.c_int64
    L.vo_grid_unique.argtypes = [C.c_int64, _P(dtype=np.uint64, flags="C_CONTIGUOUS"), _f32, _i32]
    L.vo_fine_index.argtypes = [C.c_int64, _i32, C.c_int32, C.c_int64, _i32, _i64, _i32]
    L.vo_visgen_fine.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, _f64, C.c_int64, _f64, C.c_int64, _f64,
                                 _f32, _i32, _i64, _i32, C.c_double, C.c_double, C.c_int,
                                 C.c_double, C.c_int64, C.c_void_p, _f64]
    L.vo_uvw_position_index.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i64]
    L.vo_shade_box.argtypes = [C.c_int64, _f32, _f32, _f32, C.c_int, C.c_float, _i32]
    B = C.CDLL(sb)
    B.vb_visgen_f32.argtypes = [C.c_int64, _f32, _i32, _f32, C.c_int64, _f32, _f32, C.c_float,
                                C.c_float, C.c_int, C.c_float, C.c_int64, C.c_void_p, _f32]
    B.vb_recon_f32.argtypes = [C.c_int64, C.c_int, _f32, _i32, _f32, C.c_int64, _f32, _f32,
                               C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_int64, C.c_void_p,
                               _f32]
    _lib, _base = L, B


def _idx(sel):
    if sel is None:
        return None, None
    a = np.ascontiguousarray(sel, dtype=np.int64)
    return a, a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- RNG
def splitmix64(seed, n):
    _load()
    st = C.c_uint64(seed)
    return np.array([_lib.vo_splitmix64(C.byref(st)) for _ in range(n)], dtype=np.uint64)


def normals(seed, n):
    _load()
    out = np.empty(n, np.float64)
    _lib.vo_normals(seed, n, out)
    return out


# ------------------------------------------------------------ HEALPix
def healpix_ring(nside):
    _load()
    npix = 12 * nside * nside
    out = np.empty((npix, 3), np.float64)
    v = np.empty(3, np.float64)
    for p in range(npix):
        _lib.vo_healpix_ring_vec(nside, p, v)
        out[p] = v
    return out


def healpix_nest(nside):
    _load()
    npix = 12 * nside * nside
    out = np.empty((npix, 3), np.float64)
    v = np.empty(3, np.float64)
    for p in range(npix):
        _lib.vo_healpix_nest_vec(nside, p, v)
        out[p] = v
    return out


def nest2ring(nside):
    _load()
    return np.array([_lib.vo_nest2ring(nside, p) for p in range(12 * nside * nside)], np.int64)


# ------------------------------------------------------------- orbits
def orbit_positions(n_t, dt, a, incl, mu, raan, m0):
    """calculateOrbit restated (MIncline_generate_uvw_float.cu:83-110)."""
    _load()
    raan = np.ascontiguousarray(raan, np.float64)
    m0 = np.ascontiguousarray(m0, np.float64)
    pos = np.empty((len(raan), n_t, 3), np.float64)
    _lib.vo_orbit_positions(len(raan), n_t, dt, a, incl, mu, raan, m0, pos)
    return pos


def uvw(pos, pairs, lambda_m):
    """computeUVW restated without ceil (MIncline_generate_uvw_float.cu:153-160)."""
    _load()
    pairs = np.ascontiguousarray(pairs, np.int32)
    n_t = pos.shape[1]
    out = np.empty((len(pairs), n_t, 3), np.float64)
    _lib.vo_uvw(n_t, np.ascontiguousarray(pos, np.float64), len(pairs), pairs, lambda_m, out)
    return out


# --------------------------------------------------------------- mask
def occ_one(r, s, R):
    _load()
    near = C.c_int(0)
    m = _lib.vo_occ_one(np.ascontiguousarray(r, np.float64), np.ascontiguousarray(s, np.float64),
                        R, C.byref(near))
    return bool(m), bool(near.value)


def pix_hash(p):
    _load()
    return _lib.vo_pix_hash(int(p))


def mask_stats(pos, pairs, lmn, R, vis_sel=None):
    _load()
    pos = np.ascontiguousarray(pos, np.float64)
    pairs = np.ascontiguousarray(pairs, np.int32)
    lmn = np.ascontiguousarray(lmn, np.float64)
    n_t = pos.shape[1]
    sel, selp = _idx(vis_sel)
    n_sel = len(sel) if sel is not None else len(pairs) * n_t
    cnt = np.empty(n_sel, np.int32)
    h = np.empty(n_sel, np.uint32)
    near = C.c_int64(0)
    _lib.vo_mask_stats(n_t, pos, pairs, len(lmn), lmn, R, n_sel, selp, cnt, h, C.byref(near))
    return cnt, h, near.value


def mask_bits(pos, pairs, lmn, R, vis_sel=None):
    _load()
    pos = np.ascontiguousarray(pos, np.float64)
    pairs = np.ascontiguousarray(pairs, np.int32)
    lmn = np.ascontiguousarray(lmn, np.float64)
    n_t = pos.shape[1]
    sel, selp = _idx(vis_sel)
    n_sel = len(sel) if sel is not None else len(pairs) * n_t
    nw = (len(lmn) + 31) // 32
    bits = np.empty((n_sel, nw), np.uint32)
    _lib.vo_mask_bits(n_t, pos, pairs, len(lmn), lmn, R, n_sel, selp, bits)
    return bits


# ------------------------------------------------------ vis-gen / recon
def visgen(pos, pairs, uvw_, lmn, I, R, s=1.0, wfac=0, t_occ=240.0, vis_sel=None):
    """visscal restated in FP64; returns complex128 [maps][n_sel]."""
    _load()
    I = np.ascontiguousarray(np.atleast_2d(I), np.float32)
    pos = np.ascontiguousarray(pos, np.float64)
    n_t = pos.shape[1]
    sel, selp = _idx(vis_sel)
    n_sel = len(sel) if sel is not None else len(pairs) * n_t
    out = np.empty((I.shape[0], n_sel, 2), np.float64)
    pairs = np.ascontiguousarray(pairs, np.int32)
    _lib.vo_visgen(n_t, pos.ctypes.data, pairs.ctypes.data,
                   np.ascontiguousarray(uvw_, np.float64).reshape(-1), len(lmn),
                   np.ascontiguousarray(lmn, np.float64), I.shape[0], I, R, s, wfac, t_occ,
                   n_sel, selp, out)
    return out[..., 0] + 1j * out[..., 1]


def fine_index(nx, n_pix, nx_base=1):
    """Inverse index (countLoc, start, NXq) of a fine-cell sky
    (vissGen1M_shade_optim2.cu:59-119)."""
    _load()
    nx = np.ascontiguousarray(nx, np.int32)
    count = np.empty(n_pix, np.int32)
    start = np.empty(n_pix, np.int64)
    idx = np.empty(len(nx), np.int32)
    _lib.vo_fine_index(len(nx), nx, nx_base, n_pix, count, start, idx)
    return count, start, idx


def visgen_fine(pos, pairs, uvw_, lmn, cell_lmn, ff, nx, R, s=1.0, wfac=0, t_occ=240.0,
                nx_base=1, vis_sel=None):
    """visscal over a fine-cell sky with sub-pixel shading, FP64
    (vissGen10M_shade.cu:108-145); returns complex128 [n_sel]."""
    _load()
    pos = np.ascontiguousarray(pos, np.float64)
    n_t = pos.shape[1]
    count, start, idx = fine_index(nx, len(lmn), nx_base)
    sel, selp = _idx(vis_sel)
    n_sel = len(sel) if sel is not None else len(pairs) * n_t
    out = np.empty((n_sel, 2), np.float64)
    pairs = np.ascontiguousarray(pairs, np.int32)
    _lib.vo_visgen_fine(n_t, pos.ctypes.data, pairs.ctypes.data,
                        np.ascontiguousarray(uvw_, np.float64).reshape(-1), len(lmn),
                        np.ascontiguousarray(lmn, np.float64), len(cell_lmn),
                        np.ascontiguousarray(cell_lmn, np.float64),
                        np.ascontiguousarray(ff, np.float32), count, start, idx, R, s, wfac,
                        t_occ, n_sel, selp, out)
    return out[:, 0] + 1j * out[:, 1]


def recon(pos, pairs, uvw_, lmn, V, R, s=1.0, wfac=0, kappa=None, weights=None, pix_sel=None):
    """imagerecon restated in FP64; V complex [maps][n_vis]; returns [maps][n_sel]."""
    _load()
    V = np.atleast_2d(np.asarray(V))
    Vr = np.ascontiguousarray(np.stack([V.real, V.imag], -1), np.float64)
    pos = np.ascontiguousarray(pos, np.float64)
    n_t = pos.shape[1]
    n_vis = len(pairs) * n_t
    if kappa is None:
        kappa = 1.0 / n_vis
    sel, selp = _idx(pix_sel)
    n_sel = len(sel) if sel is not None else len(lmn)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    wp = None if w is None else w.ctypes.data_as(C.c_void_p)
    out = np.empty((V.shape[0], n_sel), np.float64)
    pairs = np.ascontiguousarray(pairs, np.int32)
    _lib.vo_recon(n_t, pos.ctypes.data, pairs.ctypes.data, len(pairs),
                  np.ascontiguousarray(uvw_, np.float64).reshape(-1), len(lmn),
                  np.ascontiguousarray(lmn, np.float64), V.shape[0], Vr, wp, R, s, wfac,
                  kappa, n_sel, selp, out)
    return out


# ------------------------------------- the reference's stage-3/4 inputs
def visgen_rows(uvw_rows, lmn, I, s=1.0, wfac=0, vis_sel=None):
    """visscal over explicit uvw rows (vissGen10M_shade.cu:351-373 read them from
    updated_uvw) with no occultation, FP64; complex128 [maps][n_sel]."""
    _load()
    I = np.ascontiguousarray(np.atleast_2d(I), np.float32)
    uvw_rows = np.ascontiguousarray(uvw_rows, np.float64).reshape(-1, 3)
    n = len(uvw_rows)
    sel, selp = _idx(vis_sel)
    n_sel = len(sel) if sel is not None else n
    out = np.empty((I.shape[0], n_sel, 2), np.float64)
    _lib.vo_visgen(n, None, None, uvw_rows.reshape(-1), len(lmn),
                   np.ascontiguousarray(lmn, np.float64), I.shape[0], I, 0.0, s, wfac, 0.0,
                   n_sel, selp, out)
    return out[..., 0] + 1j * out[..., 1]


def recon_rows(uvw_rows, lmn, V, s=1.0, wfac=0, kappa=None, weights=None, pix_sel=None):
    """imagerecon over explicit uvw rows with no occultation, FP64; kappa defaults
    to 1 / n_rows (the reference divides by its row count, vissGen10M_shade.cu:173)."""
    _load()
    V = np.atleast_2d(np.asarray(V))
    Vr = np.ascontiguousarray(np.stack([V.real, V.imag], -1), np.float64)
    uvw_rows = np.ascontiguousarray(uvw_rows, np.float64).reshape(-1, 3)
    n = len(uvw_rows)
    kappa = 1.0 / n if kappa is None else kappa
    sel, selp = _idx(pix_sel)
    n_sel = len(sel) if sel is not None else len(lmn)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    wp = None if w is None else w.ctypes.data_as(C.c_void_p)
    out = np.empty((V.shape[0], n_sel), np.float64)
    _lib.vo_recon(n, None, None, 1, uvw_rows.reshape(-1), len(lmn),
                  np.ascontiguousarray(lmn, np.float64), V.shape[0], Vr, wp, 0.0, s, wfac,
                  kappa, n_sel, selp, out)
    return out


def visgen_fine_rows(uvw_rows, lmn, cell_lmn, ff, nx, s=1.0, wfac=0, nx_base=1, vis_sel=None):
    """visscal's fine-cell sum (vissGen10M_shade.cu:121-138) over explicit uvw rows,
    no occultation, FP64; complex128 [n_sel]."""
    _load()
    uvw_rows = np.ascontiguousarray(uvw_rows, np.float64).reshape(-1, 3)
    n = len(uvw_rows)
    count, start, idx = fine_index(nx, len(lmn), nx_base)
    sel, selp = _idx(vis_sel)
    n_sel = len(sel) if sel is not None else n
    out = np.empty((n_sel, 2), np.float64)
    _lib.vo_visgen_fine(n, None, None, uvw_rows.reshape(-1), len(lmn),
                        np.ascontiguousarray(lmn, np.float64), len(cell_lmn),
                        np.ascontiguousarray(cell_lmn, np.float64),
                        np.ascontiguousarray(ff, np.float32), count, start, idx, 0.0, s, wfac,
                        0.0, n_sel, selp, out)
    return out[:, 0] + 1j * out[:, 1]


def uvw_position_index(orbit_res, orbit_counts, interval, span):
    """uvwPosition's windows (MIncline_generate_uvw_float.cu:113-151): source sample
    of each kept step, -1 where zero-filled."""
    _load()
    out = np.empty(orbit_counts * (orbit_res // 3), np.int64)
    _lib.vo_uvw_position_index(orbit_res, orbit_counts, interval, span, out)
    return out


def grid_uvw(pos, pairs, lambda_m, vis_pairs=None):
    """Gridded, sorted, unique uvw with per-cell counts (MIncline_generate_uvw_float.cu:
    153-160, 374-397).  Returns (uvw_unique [n][3] float32, counts [n] int32,
    vis_count [len(vis_pairs) * n_t] or None)."""
    _load()
    pos = np.ascontiguousarray(pos, np.float64)
    pairs = np.ascontiguousarray(pairs, np.int32)
    n_t = pos.shape[1]
    keys = np.empty(len(pairs) * n_t, np.uint64)
    sc = 1e3 / lambda_m
    if _lib.vo_grid_keys(n_t, pos, len(pairs), pairs, sc, keys) < 0:
        raise ValueError("gridded uvw out of range")
    uu = np.empty((len(keys), 3), np.float32)
    cc = np.empty(len(keys), np.int32)
    sorted_keys = keys.copy()
    nu = _lib.vo_grid_unique(len(keys), sorted_keys, uu, cc)
    vc = None
    if vis_pairs is not None:
        vp = np.ascontiguousarray(vis_pairs, np.int32)
        vk = np.empty(len(vp) * n_t, np.uint64)
        _lib.vo_grid_keys(n_t, pos, len(vp), vp, sc, vk)
        uk = np.unique(sorted_keys)
        idx = np.searchsorted(uk, vk)
        vc = cc[:nu][idx]
    return uu[:nu], cc[:nu], vc


def shade_box(u, v, w, res, r_moon_ref=1737.74):
    """ShadeMatrix box restated in FP32 (shade10M.cu:62-105, 218-224)."""
    _load()
    u = np.ascontiguousarray(u, np.float32)
    out = np.empty((len(u), 4), np.int32)
    _lib.vo_shade_box(len(u), u, np.ascontiguousarray(v, np.float32),
                      np.ascontiguousarray(w, np.float32), res, r_moon_ref, out)
    return out


# ------------------------------------------------------ FP32 CPU baseline
def cpu_visgen_f32(pos, pairs, uvw_, lmn, I, R, s=1.0, wfac=0, t_occ=240.0, vis_sel=None):
    _load()
    pos_f = np.ascontiguousarray(pos, np.float32)
    n_t = pos.shape[1]
    sel, selp = _idx(vis_sel)
    n_sel = len(sel) if sel is not None else len(pairs) * n_t
    out = np.empty((n_sel, 2), np.float32)
    _base.vb_visgen_f32(n_t, pos_f, np.ascontiguousarray(pairs, np.int32),
                        np.ascontiguousarray(uvw_, np.float32).reshape(-1), len(lmn),
                        np.ascontiguousarray(lmn, np.float32), np.ascontiguousarray(I, np.float32),
                        R, s, wfac, t_occ, n_sel, selp, out)
    return out[:, 0] + 1j * out[:, 1]


def cpu_recon_f32(pos, pairs, uvw_, lmn, V, R, s=1.0, wfac=0, weights=None, pix_sel=None):
    _load()
    pos_f = np.ascontiguousarray(pos, np.float32)
    n_t = pos.shape[1]
    sel, selp = _idx(pix_sel)
    n_sel = len(sel) if sel is not None else len(lmn)
    Vf = np.ascontiguousarray(np.stack([V.real, V.imag], -1), np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    wp = None if w is None else w.ctypes.data_as(C.c_void_p)
    out = np.empty(n_sel, np.float32)
    _base.vb_recon_f32(n_t, len(pairs), pos_f, np.ascontiguousarray(pairs, np.int32),
                       np.ascontiguousarray(uvw_, np.float32).reshape(-1), len(lmn),
                       np.ascontiguousarray(lmn, np.float32), Vf, wp, R, s, wfac, n_sel, selp,
                       out)
    return out
