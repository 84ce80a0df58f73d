"""The reference's own inputs and sampling on the GPU, through the C-ABI.

- A uvw-row geometry (vrs_geom_create_uvw): the u v w count rows of an
  updated_uvw file, as vissGen10M_shade.cu:351-373 reads them, imaged with the
  reference's knobs (s = 0.5, trailing exp(-2 pi i s w), 1/count weights,
  kappa = 1/rows).  No satellites, so no occultation: the reference's own
  masks never shade on these inputs (vissGen10M_shade.cu:130, 165).
- Orbit geometries with the reference's gridded uvw (ceil, lambda = int(3e8/f),
  MIncline_generate_uvw_float.cu:156-158, 190) and its one-third-orbit
  windows (uvwPosition, :113-151), then the whole reference chain
  orbits -> gridded unique rows + counts -> rows -> V -> F file.
Bars: uvw and counts bit-exact; V, F within 1e-5 (FP64 phase) of the FP64 oracle.
"""
import math
import os

import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import refio
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu

TOL = 1e-5


def cplx(V):
    return V[..., 0].astype(np.float64) + 1j * V[..., 1].astype(np.float64)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


# 24 hand-written rows: integer u v w in wavelengths and a cell count
UPDATED_UVW = """\
-37 12 5 1
-21 -40 3 2
-15 7 -9 1
-8 33 0 4
-3 -3 -3 1
0 0 0 6
0 19 -27 1
2 -11 14 3
4 4 4 1
5 -36 8 2
7 21 -6 1
9 0 31 1
11 -29 -2 5
13 13 -13 1
17 -5 22 2
19 38 1 1
22 -17 -30 1
25 6 9 3
28 -24 17 1
30 35 -11 2
31 -8 0 1
34 15 26 1
36 -33 -7 2
39 2 12 1
"""


@pytest.fixture(scope="module")
def rows_file(tmp_path_factory):
    p = tmp_path_factory.mktemp("ref") / "updated_uvw1frequency10M.txt"
    p.write_text(UPDATED_UVW)
    return str(p)


def test_updated_uvw_rows_visgen_recon(ctx, oracle, rows_file, tmp_path, report):
    uvw, count = refio.read_uvw(rows_file, n_cols=4)
    want = np.array([[float(x) for x in ln.split()[:3]] for ln in UPDATED_UVW.splitlines()])
    assert np.array_equal(uvw, want)
    g = ctx.uvw_geometry(uvw, nside=8)
    assert g.n_vis == len(want) and g.n_pix == 768
    assert np.array_equal(g.uvw(), want)                         # bit-exact rows on the device
    assert int(g.ShadeMatrix()["count"].sum()) == 0               # no satellites: nothing shaded
    lmn = g.lmn()
    rng = np.random.default_rng(4)
    I = rng.lognormal(0.0, 1.0, (2, g.n_pix)).astype(np.float32)
    w = (1.0 / count).astype(np.float32)
    for prec, tol in (("mixed", TOL), ("plain", 1e-4)):
        k = vrs.knobs("reference", precision=prec)                 # s = 0.5, wfac = 1, kappa = 1/rows
        V = g.visscal(I, k)
        Vo = oracle.visgen_rows(uvw, lmn, I, s=k.phase_scale, wfac=k.wfac)
        F = g.imagerecon(V, w, k)
        Fo = oracle.recon_rows(uvw, lmn, cplx(V), s=k.phase_scale, wfac=k.wfac, weights=w)
        ev, ef = rel(cplx(V), Vo), rel(F, Fo)
        print(f"[updated_uvw rows {prec}] V rel {ev:.2e} F rel {ef:.2e}", flush=True)
        assert ev <= tol and ef <= tol, (prec, ev, ef)
        if prec == "mixed":
            report("updated_uvw rows (24 hand-written rows, Nside 8) vs FP64", V=ev, F=ef)
            path = os.path.join(str(tmp_path), "F_recon_10M", "F1period10M.txt")
            refio.write_image(path, F[0])
            back = refio.read_image(path)
            assert np.allclose(back, F[0], rtol=1e-5, atol=0)       # 6-digit text, as ostream
    g.close()


def test_updated_uvw_rows_fine_sky(ctx, oracle, rows_file):
    uvw, count = refio.read_uvw(rows_file, n_cols=4)
    g = ctx.uvw_geometry(uvw, nside=8)
    cell_lmn, ff, nx = W.fine_sky(8, factor=2, seed=9)
    g.fine_sky(cell_lmn, ff, nx)
    k = vrs.knobs("reference", precision="mixed")
    V = cplx(g.visscal_fine(k))
    Vo = oracle.visgen_fine_rows(uvw, g.lmn(), cell_lmn, ff, nx, s=k.phase_scale, wfac=k.wfac)
    e = rel(V, Vo)
    print(f"[updated_uvw rows fine sky] V rel {e:.2e}", flush=True)
    assert e <= TOL
    g.close()


def test_explicit_direction_sky(ctx, oracle):
    """A (theta, phi) grid of directions in the caller's order, as the reference's
    l, m, n files give them (vissGen10M_shade.cu:199-248)."""
    th = (np.arange(30) + 0.5) * math.pi / 30
    ph = np.arange(50) * 2 * math.pi / 50
    T, P = np.meshgrid(th, ph, indexing="ij")
    lmn = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    rng = np.random.default_rng(8)
    uvw = rng.integers(-300, 301, (200, 3)).astype(np.float64)
    g = ctx.uvw_geometry(uvw, lmn=lmn)
    assert g.n_pix == 1500 and np.array_equal(g.lmn(), lmn)
    I = rng.lognormal(0.0, 1.0, g.n_pix).astype(np.float32)
    k = vrs.knobs(precision="mixed")
    V = g.visscal(I, k)
    Vo = oracle.visgen_rows(uvw, lmn, I)
    F = g.imagerecon(V, None, k)
    Fo = oracle.recon_rows(uvw, lmn, cplx(V))
    assert rel(cplx(V), Vo) <= TOL and rel(F, Fo) <= TOL, (rel(cplx(V), Vo), rel(F, Fo))
    g.close()


def test_uvw_geometry_state_errors(ctx):
    g = ctx.uvw_geometry(np.zeros((4, 3)), nside=2)
    with pytest.raises(vrs.VrsError, match="VRS_E_STATE"):
        g.MIncline()                                    # no satellite positions
    with pytest.raises(vrs.VrsError, match="VRS_E_STATE"):
        g.grid_uvw()
    with pytest.raises(vrs.VrsError, match="VRS_E_STATE"):
        ctx.batch(g, [1e6], [4])
    with pytest.raises(vrs.VrsError, match="VRS_E_ARG"):
        ctx.uvw_geometry(np.zeros((4, 3)), lmn=np.ones((3, 3)))   # not unit vectors
    with pytest.raises(vrs.VrsError, match="VRS_E_ARG"):
        ctx.uvw_geometry(np.full((2, 3), np.nan), nside=2)
    g.close()


# ------------------------------------------------------- gridded orbit uvw
def test_ceil_lambda_int_uvw_bit_exact(ctx):
    wl = W.CONFIGS["c1"]
    g = ctx.geometry(wl.n_sat, wl.n_t, wl.freq_hz, wl.nside, uvw_ceil=True, lambda_int=True)
    assert g.lambda_m == 300.0                          # int(3e8 / 1e6)
    pos, uvw = g.MIncline()
    sc = 1e3 / 300.0
    want = np.ceil((pos[g.pairs[:, 1]] - pos[g.pairs[:, 0]]) * sc)
    assert np.array_equal(uvw, want)
    g.close()


def test_reference_chain_orbits_to_image(ctx, oracle, tmp_path, report):
    """orbits -> ceil-gridded uvw of every ordered pair, sorted and made unique
    with cell counts (the updated_uvw content) -> rows -> V -> F with 1/count
    weights and kappa = 1/rows -> F file."""
    wl = W.CONFIGS["small"]
    g = ctx.geometry(wl.n_sat, wl.n_t, wl.freq_hz, wl.nside)
    pos, _ = g.MIncline()
    gr = g.grid_uvw(ordered=True, lambda_int=True, weights=False)
    ordered = np.array([(m, n) for m in range(wl.n_sat) for n in range(wl.n_sat) if m != n], np.int32)
    ou, oc, _ = oracle.grid_uvw(pos, ordered, float(int(3e8 / wl.freq_hz)))
    assert np.array_equal(gr["uvw"], ou) and np.array_equal(gr["counts"], oc)
    p = refio.paths(str(tmp_path), wl.freq_hz, 1)
    refio.write_uvw(p["updated_uvw"], gr["uvw"], gr["counts"])
    rows, count = refio.read_uvw(p["updated_uvw"], n_cols=4)
    assert np.array_equal(rows, gr["uvw"].astype(np.float64))
    assert np.array_equal(count, gr["counts"].astype(np.float32))
    gu = ctx.uvw_geometry(rows, nside=wl.nside)
    I = W.make_sky(wl)
    k = vrs.knobs("reference", precision="mixed")
    V = gu.visscal(I, k)
    w = (1.0 / count).astype(np.float32)
    F = gu.imagerecon(V, w, k)
    lmn = gu.lmn()
    Vo = oracle.visgen_rows(rows, lmn, I, s=k.phase_scale, wfac=k.wfac)
    Fo = oracle.recon_rows(rows, lmn, cplx(V), s=k.phase_scale, wfac=k.wfac, weights=w)
    ev, ef = rel(cplx(V), Vo), rel(F, Fo)
    print(f"[reference chain] {len(rows)} unique rows of {len(ordered) * wl.n_t} samples: "
          f"V rel {ev:.2e} F rel {ef:.2e}", flush=True)
    report("reference chain orbits -> updated_uvw -> V -> F vs FP64", V=ev, F=ef)
    assert ev <= TOL and ef <= TOL
    refio.write_image(p["F"], F[0])
    assert os.path.getsize(p["F"]) > 0
    gu.close()
    g.close()


def test_orbit_windows_sample_index(ctx, oracle):
    """uvwPosition windows at OrbitRes 720 over 26 orbits: the last window runs off
    the propagated span and is zero-filled; positions elsewhere are the orbit at
    the window's sample times."""
    OR, OC, n_sat = 720, 26, 3
    idx = W.uvw_position_index(OR, OC)
    assert (idx < 0).any()
    period = 2 * math.pi * math.sqrt(W.A_KM ** 3 / W.MU_MOON)
    dt = period / OR
    raan, m0 = W.orbit_params(n_sat)
    g = ctx.geometry(n_sat, len(idx), 1e6, 2, raan=raan, m0=m0, dt_s=dt, sample_index=idx)
    pos, uvw = g.MIncline()
    full = oracle.orbit_positions(int(idx.max()) + 1, dt, W.A_KM, W.INCL_RAD, W.MU_MOON, raan, m0)
    live = idx >= 0
    assert np.abs(pos[:, live] - full[:, idx[live]]).max() <= 1e-12 * W.A_KM
    assert np.all(pos[:, ~live] == 0) and np.all(uvw[:, ~live] == 0)
    st = g.ShadeMatrix()
    dead = np.concatenate([b * len(idx) + np.nonzero(~live)[0] for b in range(g.n_base)])
    assert np.all(st["count"][dead] == 0)
    g.close()


def test_uvw_map_weighted_rows_1mhz(ctx, oracle, tmp_path, report):
    """The 1 MHz pipeline's stage 3-4 inputs (vissGen1M_shade_accurate.cu:248-289,
    345-381): one period's gridded samples as the rows of its uvw file, weighted
    1/count through a uvwMap table of cell counts over three periods (the
    reference's table counts over 130).  The orbits repeat each period, so most
    cells count 3.  The looked-up weights equal the gridding's own per-sample
    weights bit for bit; V and F against the FP64 oracle."""
    wl, n_per = W.CONFIGS["small"], 3
    period = 2 * math.pi * math.sqrt(W.A_KM ** 3 / W.MU_MOON)
    raan, m0 = W.orbit_params(wl.n_sat)
    g = ctx.geometry(wl.n_sat, n_per * wl.n_t, 1e6, wl.nside, raan=raan, m0=m0,
                     dt_s=period / wl.n_t, uvw_ceil=True, lambda_int=True)
    _, uvw = g.MIncline()
    gr = g.grid_uvw(ordered=False, lambda_int=True, weights=True)
    assert np.median(gr["counts"]) == n_per
    rows = uvw[:, :wl.n_t].reshape(-1, 3).astype(np.float64)       # period 1
    want_w = gr["weight"].reshape(len(g.pairs), -1)[:, :wl.n_t].reshape(-1)
    p = refio.paths(str(tmp_path), 1e6, 1)
    refio.write_uvw_map(p["uvwMap"], gr["uvw"], gr["counts"])
    refio.write_uvw(p["uvw"], rows)
    rows_in = refio.read_uvw(p["uvw"], n_cols=3)
    assert np.array_equal(rows_in, rows)
    w, info = refio.uvw_map_weights(p["uvwMap"], rows_in)
    assert info["found"] == len(rows) and info["cells"] == len(gr["counts"])
    assert np.array_equal(w, want_w)
    gu = ctx.uvw_geometry(rows_in, nside=wl.nside)
    I = W.make_sky(wl)
    k = vrs.knobs("reference", precision="mixed")
    V = gu.visscal(I, k)
    F = gu.imagerecon(V, w, k)
    lmn = gu.lmn()
    Vo = oracle.visgen_rows(rows_in, lmn, I, s=k.phase_scale, wfac=k.wfac)
    Fo = oracle.recon_rows(rows_in, lmn, cplx(V), s=k.phase_scale, wfac=k.wfac, weights=w)
    ev, ef = rel(cplx(V), Vo), rel(F, Fo)
    print(f"[1 MHz uvwMap rows] {len(rows)} rows, {info['cells']} cells: V rel {ev:.2e} "
          f"F rel {ef:.2e}", flush=True)
    report("1 MHz uvw rows weighted through uvwMap vs FP64", V=ev, F=ef)
    assert ev <= TOL and ef <= TOL
    gu.close()
    g.close()
