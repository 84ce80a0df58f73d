"""Fine-cell sky with sub-pixel shading on the GPU against the FP64 oracle.

The reference's visscal averages FF over each pixel's fine cells, a shaded
cell counting as 240 (vissGen10M_shade.cu:121-138), with the inverse index
countLoc / NXq built by scans (vissGen1M_shade_optim2.cu:59-119).  The index
must match the oracle's counting pass exactly; V within the vis-gen
tolerances (mixed <= 1e-5, plain <= 2e-2 relative L2).
"""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-5
TOL_PLAIN = 2e-2


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cplx(V):
    return V[..., 0].astype(np.float64) + 1j * V[..., 1].astype(np.float64)


_G = {}


def fine_geom(ctx, name, factor=4):
    key = (name, factor)
    if key not in _G:
        wl = W.CONFIGS[name]
        print(f"[fine geom] {name} nside={wl.nside} factor={factor}", flush=True)
        g = ctx.workload_geometry(wl)
        pos, uvw = g.MIncline()
        cl, ff, nx = W.fine_sky(wl.nside, factor=factor, seed=21)
        g.fine_sky(cl, ff, nx)
        _G[key] = (wl, g, pos, uvw, g.lmn(), cl, ff, nx)
    return _G[key]


@pytest.mark.parametrize("name", ["tiny", "small"])
def test_fine_index_matches_oracle(ctx, oracle, name):
    wl, g, pos, uvw, lmn, cl, ff, nx = fine_geom(ctx, name)
    cnt, nxq, mean = g.fine_index()
    ocnt, ostart, oidx = oracle.fine_index(nx, wl.n_pix, nx_base=1)
    assert np.array_equal(cnt, ocnt)
    assert np.array_equal(nxq, oidx)
    # per-pixel mean FF in the index's order, FP32 sums
    ref = np.array([ff[oidx[s:s + c]].astype(np.float64).mean() for s, c in zip(ostart, ocnt)])
    assert np.abs(mean - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("name,prec", [("tiny", "mixed"), ("small", "mixed"), ("small", "plain")])
def test_visgen_fine_parity(ctx, oracle, name, prec):
    wl, g, pos, uvw, lmn, cl, ff, nx = fine_geom(ctx, name)
    k = vrs.knobs("reference", precision=prec)
    V = cplx(g.visscal_fine(k))
    ref = oracle.visgen_fine(pos, g.pairs, uvw, lmn, cl, ff, nx, g.r_moon_km, s=k.phase_scale,
                             wfac=k.wfac, t_occ=k.t_occ)
    err = rel(V, ref)
    print(f"[fine {name} {prec}] rel L2 {err:.3e}", flush=True)
    assert err <= (TOL_MIXED if prec == "mixed" else TOL_PLAIN)
    # the sub-pixel mask matters: pixel-level shading of the mean sky differs
    _, _, mean = g.fine_index()
    Vp = cplx(g.visscal(mean, k))[0]
    assert rel(Vp, ref) > 10 * TOL_MIXED


def test_visgen_fine_c1_sampled(ctx, oracle):
    wl, g, pos, uvw, lmn, cl, ff, nx = fine_geom(ctx, "c1")
    k = vrs.knobs("reference", precision="mixed")
    V = cplx(g.visscal_fine(k))
    sel = np.random.default_rng(9).choice(wl.n_vis, 300, replace=False)
    ref = oracle.visgen_fine(pos, g.pairs, uvw, lmn, cl, ff, nx, g.r_moon_km, s=k.phase_scale,
                             wfac=k.wfac, t_occ=k.t_occ, vis_sel=sel)
    err = rel(V[sel], ref)
    print(f"[fine c1 sampled] rel L2 {err:.3e}", flush=True)
    assert err <= TOL_MIXED


def test_one_cell_per_pixel_is_pixel_visgen(ctx):
    # cells at the pixel centres, one per pixel, in shuffled order: the
    # fine path must give the pixel-level vis-gen
    wl = W.CONFIGS["small"]
    g = ctx.workload_geometry(wl)
    lmn = g.lmn()
    I = W.make_sky(wl)[0]
    perm = np.random.default_rng(4).permutation(wl.n_pix)
    g.fine_sky(lmn[perm], I[perm], (perm + 1).astype(np.int32))
    for prec in ("mixed", "plain"):
        k = vrs.knobs("reference", precision=prec)
        Vf = cplx(g.visscal_fine(k))
        Vp = cplx(g.visscal(I, k))[0]
        assert rel(Vf, Vp) <= 1e-6
    g.close()


def test_fine_errors(ctx):
    wl = W.CONFIGS["tiny"]
    g = ctx.workload_geometry(wl)
    with pytest.raises(RuntimeError):
        g.fine_index()
    with pytest.raises(vrs.VrsError):   # nothing attached
        g.visscal_fine()
    cl, ff, nx = W.fine_sky(wl.nside, factor=2)
    bad = nx.copy()
    bad[0] = wl.n_pix + 1                 # parent out of range
    with pytest.raises(vrs.VrsError):
        g.fine_sky(cl, ff, bad)
    lonely = nx.copy()
    lonely[lonely == 1] = 2               # pixel 1 without cells
    with pytest.raises(vrs.VrsError):
        g.fine_sky(cl, ff, lonely)
    with pytest.raises(vrs.VrsError):     # not unit vectors
        g.fine_sky(cl * 1.01, ff, nx)
    g.fine_sky(cl, ff, nx)
    with pytest.raises(vrs.VrsError):
        g.visscal_fine(vrs.knobs("reference", precision="mfma"))
    g.close()
