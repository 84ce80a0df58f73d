"""Split-FP16 matrix-core path (VRS_MFMA_SPLIT16) against the FP64 oracle.

Per-map relative L2 error bound: 1e-5 at 1-3 MHz, 1e-4 at 10 MHz (the
three-product hi/lo split carries ~22 significand bits; the phase is FP64).
"""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu


def cplx(V):
    return V[..., 0].astype(np.float64) + 1j * V[..., 1].astype(np.float64)


def per_map_rel(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


@pytest.fixture(scope="module")
def small(ctx):
    wl = W.CONFIGS["small"]
    g = ctx.workload_geometry(wl)
    pos, uvw = g.MIncline()
    rng = np.random.default_rng(5)
    # 40 maps: not a multiple of the 32-row tile; intensities spanning 1e-3 .. 1e3
    I = np.exp(rng.standard_normal((40, wl.n_pix)) * 2.0).astype(np.float32)
    I[3] *= 1e-3
    I[7] *= 1e3
    return wl, g, pos, uvw, g.lmn(), I


@pytest.mark.parametrize("preset", ["canonical", "reference"])
def test_mfma_visgen_matches_oracle(ctx, oracle, small, preset):
    wl, g, pos, uvw, lmn, I = small
    k = vrs.knobs(preset, precision="mfma")
    V = cplx(g.visscal(I, k))
    Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I, g.r_moon_km, s=k.phase_scale, wfac=k.wfac,
                       t_occ=k.t_occ)
    err = per_map_rel(V, Vo)
    print(f"[mfma visgen {preset}] per-map rel err max {err.max():.2e} median {np.median(err):.2e}")
    assert err.max() <= 1e-5, err


def test_mfma_recon_matches_oracle(ctx, oracle, small):
    wl, g, pos, uvw, lmn, I = small
    k = vrs.knobs(precision="mfma")
    V = g.visscal(I, k)
    w = (1.0 / (1.0 + np.arange(wl.n_vis) % 7)).astype(np.float32)
    F = g.imagerecon(V, w, k)
    Fo = oracle.recon(pos, g.pairs, uvw, lmn, cplx(V), g.r_moon_km, weights=w)
    err = per_map_rel(F.astype(np.float64), Fo)
    print(f"[mfma recon] per-map rel err max {err.max():.2e} median {np.median(err):.2e}")
    assert err.max() <= 1e-5, err


def test_mfma_single_map_and_adjoint_consistency(ctx, small):
    wl, g, pos, uvw, lmn, I = small
    km = vrs.knobs(precision="mfma")
    kx = vrs.knobs(precision="mixed")
    one = np.ascontiguousarray(I[:1])
    Vm, Vx = cplx(g.visscal(one, km)), cplx(g.visscal(one, kx))
    assert per_map_rel(Vm, Vx).max() <= 2e-5
    Fm = g.imagerecon(g.visscal(one, km), None, km)
    Fx = g.imagerecon(g.visscal(one, kx), None, kx)
    assert per_map_rel(Fm.astype(np.float64), Fx.astype(np.float64)).max() <= 2e-5


def test_mfma_c3_shape_sampled(ctx, oracle):
    wl = W.CONFIGS["c3"]
    g = ctx.workload_geometry(wl)
    pos, uvw = g.MIncline()
    lmn = g.lmn()
    I = W.make_sky(wl)
    k = vrs.knobs(precision="mfma")
    V = g.visscal(I, k)
    sel = np.linspace(0, wl.n_vis - 1, 24).astype(np.int64)
    maps = [0, 17, 63]
    Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I[maps], g.r_moon_km, vis_sel=sel)
    err = per_map_rel(cplx(V)[maps][:, sel], Vo)
    print(f"[c3 mfma visgen] sampled per-map rel err {err}")
    assert err.max() <= 1e-4
    F = g.imagerecon(V, None, k)
    psel = np.linspace(0, wl.n_pix - 1, 6).astype(np.int64)
    Fo = oracle.recon(pos, g.pairs, uvw, lmn, cplx(V)[maps], g.r_moon_km, pix_sel=psel)
    ferr = per_map_rel(F[maps][:, psel].astype(np.float64), Fo)
    print(f"[c3 mfma recon] sampled per-map rel err {ferr}")
    assert ferr.max() <= 1e-4
