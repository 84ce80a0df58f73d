"""Antipodal pixel pairs (vrs_knobs.antipodal, on by default).

On a HEALPix sky of Nside >= 16 the kernels pair each pixel p with its
antipode -p, whose phase is 2 wo - phi(p), so one fract and one v_sin / v_cos
serve both terms: E(-p) = conj(E(p)) exp(-4 pi i wo).  Shading is still decided
per pixel.  Checked here:
  - FP64 phase: the paired and the direct kernels agree to FP32 rounding
    (<= 2e-6) and both meet the 1e-5 bar against the FP64 oracle;
  - plain FP32: the paired kernel's error against FP64 stays within the plain
    bar and within 1.5x the direct kernel's (the antipode's phase carries p's
    FP32 rounding instead of its own);
  - skies without the pairing (Nside < 16, explicit directions) give bit-identical
    results with the knob on and off.
"""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cplx(V):
    return V[..., 0].astype(np.float64) + 1j * V[..., 1].astype(np.float64)


@pytest.fixture(scope="module")
def c1(ctx):
    wl = W.CONFIGS["c1"]
    g = ctx.workload_geometry(wl)
    pos, uvw = g.MIncline()
    yield wl, g, pos, uvw, g.lmn(), W.make_sky(wl)
    g.close()


@pytest.mark.parametrize("prec", ["mixed", "plain"])
def test_visgen_paired_vs_direct_c1(ctx, oracle, c1, prec, report):
    wl, g, pos, uvw, lmn, I = c1
    Vp = g.visscal(I, vrs.knobs(precision=prec, antipodal=1))
    Vd = g.visscal(I, vrs.knobs(precision=prec, antipodal=0))
    sel = np.arange(0, wl.n_vis, 61, dtype=np.int64)
    Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I, g.r_moon_km, vis_sel=sel)
    ep, ed = rel(cplx(Vp)[..., sel], Vo), rel(cplx(Vd)[..., sel], Vo)
    epd = rel(cplx(Vp), cplx(Vd))
    print(f"[c1 {prec}] paired {ep:.2e} direct {ed:.2e} vs FP64; paired vs direct {epd:.2e}",
          flush=True)
    report(f"c1 vis-gen {prec}: antipodal pairs vs FP64 (direct kernel beside it)",
           paired=ep, direct=ed)
    if prec == "mixed":
        assert ep <= 1e-5 and ed <= 1e-5
        assert epd <= 2e-6
    else:
        assert ep <= 2e-2 and ep <= 1.5 * ed


@pytest.mark.parametrize("prec", ["mixed", "plain"])
def test_recon_ignores_the_knob_c1(ctx, oracle, c1, prec):
    """Recon runs per pixel with the knob on or off: identical bits, and the
    FP64 bar on both."""
    wl, g, pos, uvw, lmn, I = c1
    V = g.visscal(I, vrs.knobs(precision="mixed"))
    w = (1.0 / (1.0 + np.arange(wl.n_vis) % 3)).astype(np.float32)
    Fp = g.imagerecon(V, w, vrs.knobs(precision=prec, antipodal=1))
    Fd = g.imagerecon(V, w, vrs.knobs(precision=prec, antipodal=0))
    psel = np.arange(0, wl.n_pix, 37, dtype=np.int64)
    Fo = oracle.recon(pos, g.pairs, uvw, lmn, cplx(V), g.r_moon_km, weights=w, pix_sel=psel)
    ep, ed = rel(Fp[..., psel], Fo), rel(Fd[..., psel], Fo)
    print(f"[c1 recon {prec}] paired {ep:.2e} direct {ed:.2e} vs FP64", flush=True)
    assert np.array_equal(Fp, Fd)
    assert ep <= (1e-5 if prec == "mixed" else 2e-2)


def test_unpaired_skies_ignore_the_knob(ctx):
    wl = W.CONFIGS["tiny"]                      # Nside 4: no antipodal chunk layout
    g = ctx.workload_geometry(wl)
    I = W.make_sky(wl)
    for prec in ("plain", "mixed"):
        a = g.visscal(I, vrs.knobs(precision=prec, antipodal=1))
        b = g.visscal(I, vrs.knobs(precision=prec, antipodal=0))
        assert np.array_equal(a, b)
        Fa = g.imagerecon(a, None, vrs.knobs(precision=prec, antipodal=1))
        Fb = g.imagerecon(a, None, vrs.knobs(precision=prec, antipodal=0))
        assert np.array_equal(Fa, Fb)
    g.close()
    rng = np.random.default_rng(2)
    lmn = rng.standard_normal((600, 3))
    lmn /= np.linalg.norm(lmn, axis=1, keepdims=True)
    gu = ctx.uvw_geometry(rng.integers(-200, 201, (300, 3)).astype(np.float64), lmn=lmn)
    I = rng.random(600).astype(np.float32)
    a = gu.visscal(I, vrs.knobs(precision="mixed", antipodal=1))
    b = gu.visscal(I, vrs.knobs(precision="mixed", antipodal=0))
    assert np.array_equal(a, b)
    gu.close()


def test_knob_is_validated(ctx, c1):
    wl, g, pos, uvw, lmn, I = c1
    with pytest.raises(vrs.VrsError, match="VRS_E_ARG"):
        g.visscal(I, vrs.knobs(antipodal=2))
