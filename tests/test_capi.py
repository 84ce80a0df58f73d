"""The C-ABI library on the CPU: it loads, exports every entry point that
include/vrs.h declares, validates arguments without a GPU, and its host-side
helpers (RNG, sky, orbit offsets, HEALPix) agree with the oracle."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import _lib, workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "vrs.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vrs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = vrs.load()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding table covers the same set
    assert sorted(_lib.SIGNATURES) == names


def test_version_and_knob_presets():
    lib = vrs.load()
    assert b"gfx950" in lib.vrs_version()
    k = vrs.knobs()
    assert (k.phase_scale, k.wfac, k.t_occ, k.kappa, k.precision) == (1.0, 0, 240.0, 0.0, 0)
    r = vrs.knobs("reference")
    assert (r.phase_scale, r.wfac, r.t_occ) == (0.5, 1, 240.0)
    assert vrs.knobs(precision="mixed").precision == vrs.VRS_MIXED_F64PHASE


def test_no_device_is_a_clean_error():
    lib = vrs.load()
    if lib.vrs_device_count() > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    st = lib.vrs_create(0, C.byref(h))
    assert st == -5 and not h.value
    assert b"no HIP device" in lib.vrs_last_error(None)
    with pytest.raises(vrs.VrsError) as e:
        vrs.Context(0)
    assert e.value.code == "VRS_E_NODEV"


def test_null_arguments_rejected_before_any_gpu_work():
    lib = vrs.load()
    k = vrs.knobs()
    assert lib.vrs_create(0, None) == -1
    assert lib.vrs_visgen(None, None, C.byref(k), 1, None, None) == -1
    assert b"ctx is NULL" in lib.vrs_last_error(None)
    assert lib.vrs_recon(None, None, C.byref(k), 1, None, None, None) == -1
    assert lib.vrs_adjoint(None, None, C.byref(k), None, None) == -1
    assert lib.vrs_shade(None, None, None, None, None, None) == -1
    assert lib.vrs_geom_create(None, None, None) == -1
    assert lib.vrs_orbit_uvw(None, None, None, None) == -1
    assert lib.vrs_set_stream(None, None) == -1
    assert lib.vrs_timing_last_ms(None, b"visgen") == -1.0
    assert lib.vrs_geom_info_get(None, None) == -1
    lib.vrs_destroy(None)
    lib.vrs_geom_destroy(None)


def test_host_helper_argument_checks():
    lib = vrs.load()
    out = np.empty(4, np.float32)
    assert lib.vrs_sky_lognormal(1, 1.0, 0, 0, 1.0, 2.0, out.ctypes.data) == -1
    assert lib.vrs_sky_lognormal(1, -1.0, 4, 0, 1.0, 2.0, out.ctypes.data) == -1
    assert lib.vrs_sky_lognormal(1, 1.0, 4, 5, 1.0, 2.0, out.ctypes.data) == -1
    assert lib.vrs_sky_lognormal(1, 1.0, 4, 1, 2.0, 1.0, out.ctypes.data) == -1
    assert lib.vrs_healpix_lmn(0, out.ctypes.data) == -1
    assert lib.vrs_healpix_nest2ring(3, out.ctypes.data) == -1
    d = np.empty(2, np.float64)
    assert lib.vrs_orbit_offsets(0, 0, 2000.0, 100.0, d.ctypes.data, d.ctypes.data) == -1
    assert lib.vrs_orbit_offsets(2, 0, 2000.0, 5000.0, d.ctypes.data, d.ctypes.data) == -1


def test_rng_matches_oracle(oracle):
    lib = vrs.load()
    n = 1001
    z = np.empty(n, np.float64)
    _lib.check(lib.vrs_rng_normals(42, n, z.ctypes.data))
    assert np.array_equal(z, oracle.normals(42, n))
    u = np.empty(5, np.float64)
    _lib.check(lib.vrs_rng_uniforms(1234567, 5, u.ctypes.data))
    raw = oracle.splitmix64(1234567, 5)
    assert np.array_equal(u, (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)


def test_sky_recipe(oracle):
    n = 3072
    I = np.empty(n, np.float32)
    vrs.load().vrs_sky_lognormal(1, 1.5, n, 0, 1.0, 1e3, I.ctypes.data)
    assert np.array_equal(I, np.exp(1.5 * oracle.normals(1, n)).astype(np.float32))
    # point sources: 10 distinct pixels gain S in [1, 1e3]
    J = np.empty(n, np.float32)
    vrs.load().vrs_sky_lognormal(1, 1.5, n, 10, 1.0, 1e3, J.ctypes.data)
    d = J.astype(np.float64) - I
    hit = np.nonzero(d != 0)[0]
    assert len(hit) == 10
    assert d[hit].min() >= 1.0 - 1e-3 and d[hit].max() <= 1e3 * (1 + 1e-6)


def test_orbit_offsets_recipe(oracle):
    raan, m0 = W.orbit_params(16)
    assert np.allclose(raan, 2 * np.pi * np.arange(16) / 16)
    u = np.empty(16)
    vrs.load().vrs_rng_uniforms(0, 16, u.ctypes.data)
    assert np.allclose(m0, 2 * np.arcsin(100.0 * u / (2 * W.A_KM)), rtol=1e-15)
    assert (m0 >= 0).all() and (m0 < 100.0 / W.A_KM + 1e-9).all()


@pytest.mark.parametrize("nside", [1, 2, 8, 32])
def test_healpix_matches_oracle(oracle, nside):
    lib = vrs.load()
    n = 12 * nside * nside
    lmn = np.empty((n, 3), np.float64)
    _lib.check(lib.vrs_healpix_lmn(nside, lmn.ctypes.data))
    assert np.abs(lmn - oracle.healpix_ring(nside)).max() <= 2.3e-16
    n2r = np.empty(n, np.int64)
    _lib.check(lib.vrs_healpix_nest2ring(nside, n2r.ctypes.data))
    assert np.array_equal(n2r, oracle.nest2ring(nside))


@pytest.mark.parametrize("nside", [1, 4, 16, 128])
def test_healpix_antipode(oracle, nside):
    """Pixel centres are antipodally symmetric: the map is an involution onto
    the pixel set and each pair of directions sums to zero (to FP64 rounding of
    pix2vec).  Checked against the oracle's HEALPix vectors."""
    lib = vrs.load()
    n = 12 * nside * nside
    a = np.empty(n, np.int64)
    _lib.check(lib.vrs_healpix_antipode(nside, a.ctypes.data))
    assert np.array_equal(np.sort(a), np.arange(n)) and np.array_equal(a[a], np.arange(n))
    assert not np.any(a == np.arange(n))
    v = oracle.healpix_ring(nside)
    assert np.abs(v + v[a]).max() <= 2e-15
    # the antipode of a NESTED face 0-5 pixel lies in faces 6-11 (the chunk layout
    # pairs the two halves chunk by chunk)
    n2r = np.empty(n, np.int64)
    _lib.check(lib.vrs_healpix_nest2ring(nside, n2r.ctypes.data))
    r2n = np.empty(n, np.int64)
    r2n[n2r] = np.arange(n)
    assert np.all(r2n[a[n2r[: n // 2]]] >= n // 2)
    assert lib.vrs_healpix_antipode(3, a.ctypes.data) == -1


def test_pix_hash_matches_oracle(oracle):
    lib = vrs.load()
    for p in [0, 1, 7, 196607, 2 ** 31 + 5]:
        assert lib.vrs_pix_hash(p) == oracle.pix_hash(p)


def test_workload_sizes():
    c1, c2, c3 = W.CONFIGS["c1"], W.CONFIGS["c2"], W.CONFIGS["c3"]
    assert (c1.n_base, c1.n_vis, c1.n_pix) == (28, 40320, 12288)
    assert (c2.n_base, c2.n_vis, c2.n_pix) == (120, 1036800, 196608)
    assert c2.terms_per_direction == 1036800 * 196608
    assert c3.n_maps == 64 and c3.n_pix == 49152
    f, t, a = W.c4_channels()
    assert len(f) == 32 and abs(f[0] - 1e6) < 1e-3 and abs(f[-1] - 3e7) < 1e-2
    assert t.min() >= 2000 and t.max() <= 8640
    assert abs(a.mean() + 2.5) < 0.1
    assert math.isclose(W.A_KM, 2037.4)
