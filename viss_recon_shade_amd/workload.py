"""Synthetic workloads: the array, the orbits and the skies of the five
benchmark configurations, plus tiny ones for tests.

The reference hard-wires its constants per frequency fork
(MIncline_generate_uvw_float.cu:22-30, 185-190; shade10M.cu:26-32;
vissGen10M_shade.cu:56, 183-185).  Here one orbit model serves every
configuration: circular orbits of radius R_moon + 300 km, inclination 30 deg,
RAAN spread uniformly over the satellites, along-track offsets U(0, 100 km)
drawn with seed 0, time step 5 s.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

R_MOON_KM = 1737.4
ALTITUDE_KM = 300.0
A_KM = R_MOON_KM + ALTITUDE_KM
INCL_RAD = math.radians(30.0)
MU_MOON = 4902.793455          # km^3 / s^2 (MIncline_generate_uvw_float.cu:23)
DT_S = 5.0
MAX_OFFSET_KM = 100.0
ORBIT_SEED = 0
C_LIGHT = 299792458.0


@dataclass
class Workload:
    name: str
    n_sat: int
    n_t: int
    freq_hz: float
    nside: int
    n_maps: int = 1
    sky: list = field(default_factory=list)   # per map: (seed, sigma, n_src, s_min, s_max)
    note: str = ""

    @property
    def n_base(self):
        return self.n_sat * (self.n_sat - 1) // 2

    @property
    def n_vis(self):
        return self.n_base * self.n_t

    @property
    def n_pix(self):
        return 12 * self.nside * self.nside

    @property
    def terms_per_direction(self):
        return self.n_maps * self.n_vis * self.n_pix


def _wl(name, n_sat, n_t, freq, nside, sky, note=""):
    return Workload(name, n_sat, n_t, freq, nside, len(sky), sky, note)


CONFIGS = {
    # 1 MHz plain FP32: 8 satellites, 28 baselines x 1440 steps, Nside 32, log-normal(0, 1) seed 1
    "c1": _wl("c1", 8, 1440, 1e6, 32, [(1, 1.0, 0, 1.0, 1e3)],
              "plain FP32, 1 MHz, 28 x 1440, Nside 32"),
    # 10 MHz at scale: 16 satellites, 120 x 8640, Nside 128, log-normal(0, 1.5) + 200 S^-2 sources
    "c2": _wl("c2", 16, 8640, 1e7, 128, [(1, 1.5, 200, 1.0, 1e3)],
              "plain FP32 at scale, 10 MHz, 120 x 8640, Nside 128, 200 point sources S in [1, 1e3]"),
    # precision form: c2 geometry, Nside 64, 64 maps with seeds 100..163
    "c3": _wl("c3", 16, 8640, 1e7, 64, [(100 + k, 1.0, 0, 1.0, 1e3) for k in range(64)],
              "64 maps, 10 MHz, 120 x 8640, Nside 64"),
    # differentiable form: 120 x 4320, Nside 64 ground truth
    "c5": _wl("c5", 16, 4320, 1e7, 64, [(1, 1.0, 0, 1.0, 1e3)],
              "least squares, 10 MHz, 120 x 4320, Nside 64"),
    # tiny cases for tests
    "tiny": _wl("tiny", 3, 16, 1e6, 4, [(7, 1.0, 0, 1.0, 1e3)], "3 sats, 16 steps, Nside 4"),
    "small": _wl("small", 4, 96, 3e6, 16, [(11, 1.0, 3, 1.0, 1e3)],
                 "4 sats, 96 steps, Nside 16, 3 point sources"),
}


def c4_channels(n_ch=32, seed=2):
    """Batched form: channel frequencies log-spaced over 1-30 MHz, per-channel
    steps T_c ~ U{2000..8640} and spectral index alpha_c ~ N(-2.5, 0.1),
    both from one SplitMix64 stream seeded with `seed`."""
    lib = _lib.load()
    freqs = 1e6 * 30.0 ** (np.arange(n_ch) / (n_ch - 1))
    u = np.empty(n_ch, np.float64)
    _lib.check(lib.vrs_rng_uniforms(seed, n_ch, u.ctypes.data))
    n_t = (2000 + np.floor(u * 6641)).astype(np.int64)
    z = np.empty(n_ch, np.float64)
    _lib.check(lib.vrs_rng_normals(seed + 1000003, n_ch, z.ctypes.data))
    alpha = -2.5 + 0.1 * z
    return freqs, n_t, alpha


def orbit_params(n_sat, seed=ORBIT_SEED, a_km=A_KM, max_offset_km=MAX_OFFSET_KM):
    lib = _lib.load()
    raan = np.empty(n_sat, np.float64)
    m0 = np.empty(n_sat, np.float64)
    _lib.check(lib.vrs_orbit_offsets(n_sat, seed, a_km, max_offset_km, raan.ctypes.data,
                                     m0.ctypes.data))
    return raan, m0


def make_sky(wl):
    """[n_maps][n_pix] float32 intensities in RING order."""
    lib = _lib.load()
    out = np.empty((wl.n_maps, wl.n_pix), np.float32)
    for k, (seed, sigma, n_src, s_min, s_max) in enumerate(wl.sky):
        row = np.empty(wl.n_pix, np.float32)
        _lib.check(lib.vrs_sky_lognormal(seed, sigma, wl.n_pix, n_src, s_min, s_max,
                                         row.ctypes.data))
        out[k] = row
    return out


def all_pairs(n_sat):
    return np.array([(a, b) for a in range(n_sat) for b in range(a + 1, n_sat)], np.int32)


def fine_sky(nside, factor=4, seed=7, sigma=1.0):
    """A fine-cell sky under an nside sky: the HEALPix pixels of nside * factor
    as cells (RING order), lognormal intensities, and each cell's parent RING
    pixel, 1-based as the reference's NX file (NESTED: parent = child >> 2 log2 f).
    Returns (cell_lmn [n][3] float64, FF [n] float32, NX [n] int32)."""
    if factor < 1 or factor & (factor - 1):
        raise ValueError("factor must be a power of two")
    lib = _lib.load()
    nf = nside * factor
    n_cells, n_pix = 12 * nf * nf, 12 * nside * nside
    lmn = np.empty((n_cells, 3), np.float64)
    _lib.check(lib.vrs_healpix_lmn(nf, lmn.ctypes.data))
    ring_of_nest_f = np.empty(n_cells, np.int64)
    _lib.check(lib.vrs_healpix_nest2ring(nf, ring_of_nest_f.ctypes.data))
    ring_of_nest_c = np.empty(n_pix, np.int64)
    _lib.check(lib.vrs_healpix_nest2ring(nside, ring_of_nest_c.ctypes.data))
    shift = 2 * (factor.bit_length() - 1)
    nx = np.empty(n_cells, np.int32)
    nx[ring_of_nest_f] = (ring_of_nest_c[np.arange(n_cells) >> shift] + 1).astype(np.int32)
    ff = np.empty(n_cells, np.float32)
    _lib.check(lib.vrs_sky_lognormal(seed, sigma, n_cells, 0, 1.0, 1.0, ff.ctypes.data))
    return lmn, ff, nx


C4_NSIDE = 64
C4_N_SAT = 16


def c4_skies(freqs, alpha, nside=C4_NSIDE, seed=1):
    """Per-channel skies of the batched form: I_c = I0 (nu_c / 1 MHz)^alpha_c,
    I0 log-normal(0, 1) from `seed` (RING order), [n_ch][n_pix] float32."""
    n_pix = 12 * nside * nside
    I0 = make_sky(Workload("c4", C4_N_SAT, 1, 1e6, nside, 1, [(seed, 1.0, 0, 1.0, 1e3)]))[0]
    return np.stack([I0.astype(np.float64) * (f / 1e6) ** a
                     for f, a in zip(freqs, alpha)]).astype(np.float32).reshape(len(freqs), n_pix)


def uvw_position_index(orbit_res, orbit_counts, interval=None, span=None):
    """The reference's orbit windows (uvwPosition, MIncline_generate_uvw_float.cu:
    113-151, constants :185-190): source sample index of every kept step, -1 for
    zero-filled ones.  interval defaults to (orbit_res / 360) * 10, span to
    orbit_res / 3, both in the reference's integer arithmetic."""
    lib = _lib.load()
    interval = (orbit_res // 360) * 10 if interval is None else interval
    span = orbit_res // 3 if span is None else span
    out = np.empty(orbit_counts * span, np.int64)
    _lib.check(lib.vrs_uvw_position_index(orbit_res, orbit_counts, interval, span,
                                          out.ctypes.data))
    return out
