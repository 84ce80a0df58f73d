"""The batched configuration at its own size: 32 channels log-spaced over
1-30 MHz, 16 satellites (120 baselines), per-channel steps T_c ~ U{2000..8640}
(seed 2), line-of-sight compaction, Nside 64, per-channel spectral index.

One vrs_batch run (one captured hipGraph: chunk-order sky gather, one
persistent vis-gen launch over all channels' kept samples, one persistent
recon launch) replaces the reference's one-frequency-per-program loop
(vissGen10M_shade.cu:291-298).  Checked against the FP64 oracle
(vissGen10M_shade_optim1.cu:130-272 restated):
  V: every channel, 64 kept samples x all pixels, per-channel rel. L2 <= 1e-5
  F: channels 0, 10, 21, 31, 16 pixels x all kept samples, kappa = 1/kept,
     per-channel rel. L2 <= 1e-5
  mask: shade counts and hashes of kept samples of every channel, bit-exact.
"""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu

TOL = 1e-5
C = 299792458.0


def cplx(V):
    return V[..., 0].astype(np.float64) + 1j * V[..., 1].astype(np.float64)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def c4(ctx):
    freqs, n_t, alpha = W.c4_channels()
    print(f"[c4] building: {len(freqs)} channels, T_c in [{n_t.min()}, {n_t.max()}]", flush=True)
    g = ctx.geometry(W.C4_N_SAT, int(n_t.max()), 1e7, W.C4_NSIDE)
    b = ctx.batch(g, freqs, n_t, compact=True, precision="mixed", autotune=True)
    I = W.c4_skies(freqs, alpha)
    V, F = b.run(I, vrs.knobs(precision="mixed"))
    pos, _ = g.MIncline()
    print(f"[c4] kept {int(b.kept.sum())} of {120 * int(n_t.sum())} samples, info {b.info}",
          flush=True)
    yield freqs, n_t, g, b, I, V, F, pos, g.lmn()
    b.close()
    g.close()


def _kept(b, g, c):
    sl = slice(b.offsets[c], b.offsets[c + 1])
    bt = b.sample_bt[sl]
    real = bt >= 0
    return sl, real, bt[real]


def test_c4_layout(c4):
    freqs, n_t, g, b, I, V, F, pos, lmn = c4
    assert len(freqs) == 32 and freqs[0] == 1e6 and abs(freqs[-1] - 3e7) < 1e-3
    assert n_t.min() >= 2000 and n_t.max() <= 8640
    assert b.info["n_vis_total"] == int(b.kept.sum())
    assert (b.kept > 0).all()
    for c in range(len(freqs)):
        sl, real, bt = _kept(b, g, c)
        assert real.sum() == b.kept[c]
        assert np.all(V[sl][~real] == 0)          # padding stays zero
        t = bt % g.n_t
        assert t.max() < n_t[c]                   # channel c uses its first n_t[c] steps
        assert np.all(np.diff(bt) > 0)            # (baseline, t) order, no repeats
    # compaction: each kept sample's satellites see each other past the Moon
    assert b.kept.sum() < 120 * int(n_t.sum())


def test_c4_visibilities_every_channel(c4, oracle, report):
    freqs, n_t, g, b, I, V, F, pos, lmn = c4
    worst = 0.0
    for c in range(len(freqs)):
        sl, real, bt = _kept(b, g, c)
        pick = np.linspace(0, len(bt) - 1, 64).astype(np.int64)
        uvw = oracle.uvw(pos, g.pairs, C / freqs[c])
        Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I[c], g.r_moon_km, vis_sel=bt[pick])[0]
        e = rel(cplx(V[sl][real][pick]), Vo)
        worst = max(worst, e)
        assert e <= TOL, (c, freqs[c], e)
    print(f"[c4] V worst channel rel L2 {worst:.2e}", flush=True)
    report("c4 V (32 channels x 64 kept samples x all pixels vs FP64)", max_rel_l2=worst)


@pytest.mark.parametrize("c", [0, 10, 21, 31])
def test_c4_image_channels(c4, oracle, report, c):
    freqs, n_t, g, b, I, V, F, pos, lmn = c4
    sl, real, bt = _kept(b, g, c)
    uvw = oracle.uvw(pos, g.pairs, C / freqs[c])
    Vfull = np.zeros(g.n_vis, np.complex128)
    Vfull[bt] = cplx(V[sl][real])
    psel = np.linspace(0, g.n_pix - 1, 16).astype(np.int64)
    Fo = oracle.recon(pos, g.pairs, uvw, lmn, Vfull, g.r_moon_km, kappa=1.0 / b.kept[c],
                      pix_sel=psel)[0]
    e = rel(F[c, psel], Fo)
    print(f"[c4] channel {c} ({freqs[c] / 1e6:.2f} MHz, {b.kept[c]} samples): F rel {e:.2e}",
          flush=True)
    report(f"c4 F channel {c} (16 pixels x all kept samples vs FP64)", rel_l2=e)
    assert e <= TOL, e


def test_c4_mask_of_kept_samples(c4, oracle):
    freqs, n_t, g, b, I, V, F, pos, lmn = c4
    st = g.ShadeMatrix()
    for c in (0, 15, 31):
        sl, real, bt = _kept(b, g, c)
        pick = bt[::97]
        cnt, h, _ = oracle.mask_stats(pos, g.pairs, lmn, g.r_moon_km, vis_sel=pick)
        assert np.array_equal(st["count"][pick], cnt), c
        assert np.array_equal(st["hash"][pick], h), c
