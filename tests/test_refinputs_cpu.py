"""The reference's stage-1 sampling and stage-3/4 inputs, on the CPU.

- uvwPosition's one-third-orbit windows (MIncline_generate_uvw_float.cu:113-151):
  the library's host helper against the oracle's restatement of the kernel's
  index arithmetic, and both against windows worked out by hand.
- The oracle's vis-gen / recon over explicit uvw rows (the updated_uvw input of
  vissGen10M_shade.cu:351-373, no occultation) pinned by brute force in numpy,
  a single-pixel closed form and the adjoint identity.
"""
import numpy as np
import pytest

from viss_recon_shade_amd import _lib
from viss_recon_shade_amd import workload as W


def lib_windows(OR, OC, interval, span):
    out = np.empty(OC * span, np.int64)
    _lib.check(_lib.load().vrs_uvw_position_index(OR, OC, interval, span, out.ctypes.data))
    return out


# ----------------------------------------------------------- orbit windows
def test_windows_by_hand(oracle):
    # OR = 12, OC = 3, interval 5, span 4: orbit 2 starts at 12 + 5 = 17, orbit 3 at
    # 24 + 10 = 34 and runs off the 36-sample series: 34, 35, then zero-filled (the
    # reference reads sample 36, one past the series, quirk Q11; here it is -1)
    want = [0, 1, 2, 3, 17, 18, 19, 20, 34, 35, -1, -1]
    assert lib_windows(12, 3, 5, 4).tolist() == want
    assert oracle.uvw_position_index(12, 3, 5, 4).tolist() == want
    # interval 8: orbit 3 would start at 2 * 20 = 40 > 36, so it is wrapped back by
    # reset = 36 // 20 + 1 = 2 orbits: (2 - 2) * 12 + 16 = 16
    want = [0, 1, 2, 3, 20, 21, 22, 23, 16, 17, 18, 19]
    assert lib_windows(12, 3, 8, 4).tolist() == want
    assert oracle.uvw_position_index(12, 3, 8, 4).tolist() == want


@pytest.mark.parametrize("OR,OC", [(12, 7), (720, 40), (8640, 140), (360, 150)])
def test_windows_match_oracle(oracle, OR, OC):
    interval, span = (OR // 360) * 10 if OR >= 360 else 5, OR // 3
    a = lib_windows(OR, OC, interval, span)
    b = oracle.uvw_position_index(OR, OC, interval, span)
    assert np.array_equal(a, b)
    total = OR * OC
    blocks = a.reshape(OC, span)
    for blk in blocks:
        live = blk[blk >= 0]
        assert np.all(np.diff(live) == 1)          # each window is a run of samples
        assert live.size == 0 or live.max() < total
        assert np.all(blk[live.size:] == -1)       # zero fill only at the end of a window


def test_reference_constants_wrap_and_zero_fill():
    """The reference's own constants at stride 1 (OrbitRes = 2.4 * 3600 = 8640,
    interval 240, span 2880, 140 orbits): the last orbits wrap back and one window
    runs off the end of the series."""
    idx = W.uvw_position_index(8640, 140)
    blocks = idx.reshape(140, 2880)
    starts = blocks[:, 0]
    assert starts[1] - starts[0] == 8640 + 240
    assert np.any(idx < 0)                        # zero-filled tail of one window
    assert np.any(np.diff(starts) < 0)            # wrapped windows


def test_window_helper_rejects_bad_arguments():
    lib = _lib.load()
    out = np.empty(16, np.int64)
    assert lib.vrs_uvw_position_index(12, 3, 5, 3, out.ctypes.data) == -1   # span != OR / 3
    assert lib.vrs_uvw_position_index(2, 3, 5, 1, out.ctypes.data) == -1
    assert lib.vrs_uvw_position_index(12, 3, 5, 4, None) == -1


# ------------------------------------------------- uvw rows, no occultation
def brute_visgen(uvw, lmn, I, s, wfac):
    ph = (uvw[:, 0:1] * lmn[None, :, 0] + uvw[:, 1:2] * lmn[None, :, 1]
          + uvw[:, 2:3] * (lmn[None, :, 2] - 1.0)) * s
    V = (np.exp(-2j * np.pi * ph) * I[None, :].astype(np.float64)).sum(1)
    if wfac:
        V = V * np.exp(-2j * np.pi * s * uvw[:, 2])
    return V


def brute_recon(uvw, lmn, V, w, kappa, s, wfac):
    n = lmn[:, 2] if wfac else lmn[:, 2] - 1.0
    ph = (uvw[:, 0:1] * lmn[None, :, 0] + uvw[:, 1:2] * lmn[None, :, 1]
          + uvw[:, 2:3] * n[None, :]) * s
    return kappa * np.real((w[:, None] * V[:, None]) * np.exp(2j * np.pi * ph)).sum(0)


@pytest.fixture(scope="module")
def rows():
    rng = np.random.default_rng(12)
    uvw = rng.integers(-40, 41, (57, 3)).astype(np.float64)    # integer rows, as updated_uvw
    lmn = np.empty((12 * 16, 3))
    _lib.load().vrs_healpix_lmn(4, lmn.ctypes.data)
    I = rng.lognormal(0.0, 1.0, len(lmn)).astype(np.float32)
    return uvw, lmn, I


@pytest.mark.parametrize("s,wfac", [(1.0, 0), (0.5, 1)])
def test_rows_visgen_and_recon_brute_force(oracle, rows, s, wfac):
    uvw, lmn, I = rows
    V = oracle.visgen_rows(uvw, lmn, I, s=s, wfac=wfac)[0]
    Vb = brute_visgen(uvw, lmn, I, s, wfac)
    assert np.abs(V - Vb).max() <= 1e-11 * np.abs(Vb).max()
    w = (1.0 / (1 + np.arange(len(uvw)) % 4)).astype(np.float32)
    F = oracle.recon_rows(uvw, lmn, V, s=s, wfac=wfac, weights=w)[0]
    Fb = brute_recon(uvw, lmn, V, w.astype(np.float64), 1.0 / len(uvw), s, wfac)
    assert np.abs(F - Fb).max() <= 1e-11 * np.abs(Fb).max()


def test_rows_single_pixel_and_adjoint(oracle, rows):
    uvw, lmn, _ = rows
    p = 37
    I = np.zeros(len(lmn), np.float32)
    I[p] = 1.0
    V = oracle.visgen_rows(uvw, lmn, I)[0]
    ph = uvw @ np.array([lmn[p, 0], lmn[p, 1], lmn[p, 2] - 1.0])
    assert np.allclose(V, np.exp(-2j * np.pi * ph), rtol=0, atol=1e-12)
    # <A I, y> = <I, A^H y>: recon with kappa = 1 is Re(A^H y)
    rng = np.random.default_rng(3)
    I = rng.random(len(lmn)).astype(np.float32)
    y = rng.standard_normal(len(uvw)) + 1j * rng.standard_normal(len(uvw))
    AI = oracle.visgen_rows(uvw, lmn, I)[0]
    AHy = oracle.recon_rows(uvw, lmn, y, kappa=1.0)[0]
    lhs = np.real(np.vdot(y, AI))
    rhs = float(I.astype(np.float64) @ AHy)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_rows_fine_sky_is_pixel_mean(oracle, rows):
    """With nothing occulted, the fine-cell sum of a pixel is its cells' mean FF
    at the pixel's direction (vissGen10M_shade.cu:121-138)."""
    uvw, lmn, _ = rows
    cell_lmn, ff, nx = W.fine_sky(4, factor=2, seed=5)
    Vf = oracle.visgen_fine_rows(uvw, lmn, cell_lmn, ff, nx)
    mean = np.zeros(len(lmn))
    np.add.at(mean, nx - 1, ff.astype(np.float64))
    mean /= np.bincount(nx - 1, minlength=len(lmn))
    Vb = brute_visgen(uvw, lmn, mean, 1.0, 0)
    assert np.abs(Vf - Vb).max() <= 1e-11 * np.abs(Vb).max()
