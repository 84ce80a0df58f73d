"""Edge cases on the GPU against the FP64 oracle: the smallest geometry (two
satellites, one step, Nside 1: a single chunk that is mostly padding), an
explicit baseline list given in reverse order, a large sky with few
visibilities (Nside 256), and a map count that crosses the matrix-core
path's 64-map groups.  Same tolerances as the parity tests: mask bit-exact,
mixed <= 1e-5, plain <= 2e-2, split-FP16 <= 1e-4 relative L2."""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cplx(V):
    return V[..., 0].astype(np.float64) + 1j * V[..., 1].astype(np.float64)


def check_geometry(g, oracle, I, precisions=("mixed", "plain")):
    pos, uvw = g.MIncline()
    lmn = g.lmn()
    st = g.ShadeMatrix(bits=True)
    assert np.array_equal(st["bits"], oracle.mask_bits(pos, g.pairs, lmn, g.r_moon_km))
    for prec in precisions:
        k = vrs.knobs("reference", precision=prec)
        V = g.visscal(I, k)
        Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I, g.r_moon_km, s=k.phase_scale,
                           wfac=k.wfac, t_occ=k.t_occ)
        tol = 1e-5 if prec == "mixed" else 2e-2
        assert rel(cplx(V), Vo) <= tol, (prec, rel(cplx(V), Vo))
        F = g.imagerecon(V, None, k)
        Fo = oracle.recon(pos, g.pairs, uvw, lmn, cplx(V), g.r_moon_km, s=k.phase_scale,
                          wfac=k.wfac)
        assert rel(F, Fo) <= tol, (prec, rel(F, Fo))
    return pos, uvw, lmn


def test_smallest_geometry(ctx, oracle):
    print("[edges] 2 satellites, 1 step, Nside 1", flush=True)
    g = ctx.geometry(2, 1, 1e6, 1)
    assert g.n_vis == 1 and g.n_pix == 12
    I = np.linspace(0.5, 3.0, 12, dtype=np.float32)[None]
    pos, uvw, lmn = check_geometry(g, oracle, I)
    # the matrix-core path on 3 maps of the same geometry
    I3 = np.stack([I[0], 2 * I[0], I[0][::-1]]).astype(np.float32)
    k = vrs.knobs(precision="mfma")
    V3 = g.visscal(I3, k)
    Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I3, g.r_moon_km)
    assert rel(cplx(V3), Vo) <= 1e-4
    # fine-cell sky: 4 cells per pixel
    cl, ff, nx = W.fine_sky(1, factor=2, seed=3)
    g.fine_sky(cl, ff, nx)
    kf = vrs.knobs("reference", precision="mixed")
    Vf = oracle.visgen_fine(pos, g.pairs, uvw, lmn, cl, ff, nx, g.r_moon_km, s=kf.phase_scale,
                            wfac=kf.wfac, t_occ=kf.t_occ)
    assert rel(cplx(g.visscal_fine(kf)), Vf) <= 1e-5
    # gridding of the two ordered pairs
    gr = g.grid_uvw()
    assert int(gr["counts"].sum()) == 2
    g.close()


def test_explicit_reversed_baselines(ctx, oracle):
    print("[edges] explicit baselines (2,0), (3,1), (1,2)", flush=True)
    g = ctx.geometry(4, 24, 3e6, 8, pairs=[(2, 0), (3, 1), (1, 2)])
    assert g.n_vis == 3 * 24
    I = W.make_sky(W.Workload("e", 4, 24, 3e6, 8, 1, [(5, 1.0, 2, 1.0, 1e3)]))
    check_geometry(g, oracle, I)
    g.close()


def test_large_sky_few_visibilities(ctx, oracle):
    print("[edges] Nside 256, 2 satellites x 8 steps", flush=True)
    g = ctx.geometry(2, 8, 1e7, 256)
    assert g.n_pix == 786432
    I = W.make_sky(W.Workload("e", 2, 8, 1e7, 256, 1, [(9, 1.0, 5, 1.0, 1e3)]))
    pos, uvw = g.MIncline()
    lmn = g.lmn()
    st = g.ShadeMatrix()
    cnt, h, near = oracle.mask_stats(pos, g.pairs, lmn, g.r_moon_km)
    assert np.array_equal(st["count"], cnt) and np.array_equal(st["hash"], h)
    k = vrs.knobs(precision="mixed")
    V = g.visscal(I, k)
    assert rel(cplx(V), oracle.visgen(pos, g.pairs, uvw, lmn, I, g.r_moon_km)) <= 1e-4
    psel = np.linspace(0, g.n_pix - 1, 64).astype(np.int64)
    F = g.imagerecon(V, None, k)
    Fo = oracle.recon(pos, g.pairs, uvw, lmn, cplx(V), g.r_moon_km, pix_sel=psel)
    assert rel(F[0, psel], Fo[0]) <= 1e-4
    g.close()


def test_map_count_across_matrix_core_groups(ctx, oracle):
    print("[edges] 65 maps on the matrix cores", flush=True)
    wl = W.CONFIGS["tiny"]
    g = ctx.workload_geometry(wl)
    pos, uvw = g.MIncline()
    lmn = g.lmn()
    rng = np.random.default_rng(12)
    I = rng.lognormal(0.0, 1.0, (65, wl.n_pix)).astype(np.float32)
    k = vrs.knobs(precision="mfma")
    V = g.visscal(I, k)
    sel = [0, 63, 64]
    Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I[sel], g.r_moon_km)
    for i, m in enumerate(sel):
        assert rel(cplx(V[m]), Vo[i]) <= 1e-4, m
    F = g.imagerecon(V, None, k)
    Fo = oracle.recon(pos, g.pairs, uvw, lmn, cplx(V[sel]), g.r_moon_km)
    for i, m in enumerate(sel):
        assert rel(F[m], Fo[i]) <= 1e-4, m
    g.close()


def test_term_ceiling_bounds(ctx):
    """The measured bare-term rate: positive, below the VALU issue bound (one
    wave64 VALU instruction per 4 clocks per SIMD at 2.4 GHz: 1024 SIMDs x
    2.4e9 x 16 lanes / 6 non-transcendental instructions), and above the
    plain vis-gen kernel's own rate, which runs the same term plus shading."""
    print("[edges] term ceiling", flush=True)
    c = ctx.term_ceiling()
    assert 0 < c < 1024 * 2.4e9 * 16 / 6 / 1e9, c
    wl = W.CONFIGS["c1"]
    g = ctx.workload_geometry(wl)
    I = W.make_sky(wl)
    k = vrs.knobs(precision="plain")
    ctx.timing(True)
    g.visscal(I, k)
    ms = ctx.last_ms("visgen")
    ctx.timing(False)
    rate = wl.n_vis * wl.n_pix / (ms * 1e-3) / 1e9
    assert rate < 1.05 * c, (rate, c)
    g.close()


def test_matrix_core_recon_every_pixel(ctx):
    """The matrix-core recon against the FP64-phase recon of the same V on every
    pixel of c1 with 64 maps.  Each half-wave of the recon kernel takes its own
    8 visibilities of a K-step; a K-step where only one half holds shaded or
    straddling pairs must still mask them (a wave-uniform test on lane 0 alone
    left ~7e-4 relative L2 here, 1e-2 in the worst map)."""
    print("[edges] matrix-core recon vs FP64 recon, c1, 64 maps, all pixels", flush=True)
    wl = W.CONFIGS["c1"]
    g = ctx.workload_geometry(wl)
    rng = np.random.default_rng(5)
    I = rng.lognormal(0.0, 1.0, (64, wl.n_pix)).astype(np.float32)
    kx = vrs.knobs(precision="mixed")
    V = g.visscal(I, kx)
    Fm = g.imagerecon(V, None, vrs.knobs(precision="mfma")).astype(np.float64)
    Fx = g.imagerecon(V, None, kx).astype(np.float64)
    assert rel(Fm, Fx) <= 5e-5, rel(Fm, Fx)
    worst = float(np.max(np.abs(Fm - Fx).max(1) / np.abs(Fx).max(1)))
    assert worst <= 1e-4, worst
    g.close()


def test_matrix_core_visgen_every_visibility(ctx):
    """The matrix-core vis-gen against the FP64-phase vis-gen on every visibility
    of c1 with 64 maps (wave tiles that shade a whole chunk skip the products and
    keep only the T_occ term; mixed tiles mask per pixel)."""
    print("[edges] matrix-core vis-gen vs FP64 vis-gen, c1, 64 maps, all visibilities", flush=True)
    wl = W.CONFIGS["c1"]
    g = ctx.workload_geometry(wl)
    rng = np.random.default_rng(6)
    I = rng.lognormal(0.0, 1.0, (64, wl.n_pix)).astype(np.float32)
    Vm = cplx(g.visscal(I, vrs.knobs(precision="mfma")))
    Vx = cplx(g.visscal(I, vrs.knobs(precision="mixed")))
    assert rel(Vm, Vx) <= 2e-5, rel(Vm, Vx)
    worst = float(np.max(np.abs(Vm - Vx).max(1) / np.abs(Vx).max(1)))
    assert worst <= 5e-5, worst
    g.close()
