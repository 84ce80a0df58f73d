"""Ragged multi-channel batch (one captured hipGraph per run) against
per-channel single-geometry runs and the FP64 oracle.

The batch computes exactly the per-visibility sums that a single-channel
geometry computes (same positions, same uvw formula, same chunk order), so
vis-gen must agree with per-channel vrs_visgen to FP32 rounding; recon is
checked against the oracle with kappa = 1 / kept samples.  The graph is
replayed several times with eager work and synchronisations in between,
each replay against the eager result.
"""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu

FREQS = [1e6, 4e6, 2.5e6]
NT = [96, 40, 70]
# orbit planes 0.25 rad apart: satellites see each other part of the time, so the
# line-of-sight compaction keeps an uneven share of samples per channel
RAAN = np.array([0.0, 0.25, 0.5, 0.75])
M0 = np.array([0.0, 0.3, 0.6, 0.9])


def cplx(V):
    return V[..., 0].astype(np.float64) + 1j * V[..., 1].astype(np.float64)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def setup(ctx):
    wl = W.CONFIGS["small"]
    g = ctx.geometry(wl.n_sat, wl.n_t, wl.freq_hz, wl.nside, raan=RAAN, m0=M0)
    rng = np.random.default_rng(9)
    I = np.exp(rng.standard_normal((len(FREQS), wl.n_pix))).astype(np.float32)
    return wl, g, I


@pytest.mark.parametrize("prec", ["mixed", "plain"])
def test_batch_matches_per_channel_runs(ctx, oracle, setup, prec):
    wl, g, I = setup
    b = ctx.batch(g, FREQS, NT, compact=True, precision=prec, autotune=True)
    k = vrs.knobs(precision=prec)
    V, F = b.run(I, k)
    print(f"[batch {prec}] info {b.info} kept {list(b.kept)}", flush=True)
    assert b.info["n_vis_total"] == int(b.kept.sum())
    assert (b.kept > 0).all()
    assert b.kept.sum() < g.n_base * sum(NT)   # compaction dropped samples
    lmn = g.lmn()
    for c, (f, nt) in enumerate(zip(FREQS, NT)):
        gc = ctx.geometry(wl.n_sat, nt, f, wl.nside, raan=RAAN, m0=M0)
        Vc = cplx(gc.visscal(np.ascontiguousarray(I[c:c + 1]), k))[0]
        sl = slice(b.offsets[c], b.offsets[c + 1])
        bt = b.sample_bt[sl]
        real = bt >= 0
        bl, t = bt[real] // g.n_t, bt[real] % g.n_t
        idx = bl * nt + t
        assert len(idx) == b.kept[c]
        assert np.all(V[sl][~real] == 0)
        e = rel(cplx(V[sl][real]), Vc[idx])
        assert e <= 1e-6, (c, e)
        # recon: kappa = 1 / kept over the kept samples only
        pos, uvw = gc.MIncline()
        Vfull = np.zeros(gc.n_vis, np.complex128)
        Vfull[idx] = cplx(V[sl][real])
        Fo = oracle.recon(pos, gc.pairs, uvw, lmn, Vfull, gc.r_moon_km, kappa=1.0 / b.kept[c])[0]
        tol = 1e-5 if prec == "mixed" else 2e-2
        ef = rel(F[c], Fo)
        print(f"[batch {prec}] channel {c}: {b.kept[c]} samples, V rel {e:.1e}, F rel {ef:.1e}")
        assert ef <= tol, (c, ef)
        gc.close()
    b.close()


def test_batch_graph_replays(ctx, setup):
    torch = pytest.importorskip("torch")
    wl, g, I = setup
    b = ctx.batch(g, FREQS, NT, compact=True, precision="mixed", autotune=False)
    k = vrs.knobs(precision="mixed")
    ref = [b.run(I * s, k) for s in (1.0, 2.0, 0.5)]   # eager-free reference set (host arrays)
    It = torch.from_numpy(I).cuda()
    for rep in range(3):
        for j, s in enumerate((1.0, 2.0, 0.5)):
            Vd, Fd = b.run(It * s, k)
            # eager work between replays
            x = torch.randn(1 << 20, device="cuda").sum()
            torch.cuda.synchronize()
            assert np.array_equal(Vd.cpu().numpy(), ref[j][0]), (rep, j)
            assert np.array_equal(Fd.cpu().numpy(), ref[j][1]), (rep, j)
            del x
    # knob change re-captures the graph
    k2 = vrs.knobs(precision="mixed", t_occ=0.0)
    V2, F2 = b.run(I, k2)
    assert not np.array_equal(V2, ref[0][0])
    b.close()


def test_batch_argument_errors(ctx, setup):
    wl, g, I = setup
    with pytest.raises(vrs.VrsError):
        ctx.batch(g, [1e6], [g.n_t + 1])
    with pytest.raises(vrs.VrsError):
        ctx.batch(g, [-1.0], [10])
    with pytest.raises(vrs.VrsError):
        ctx.batch(g, [1e6], [10], precision="mfma")
    b = ctx.batch(g, [1e6], [10], autotune=False)
    with pytest.raises(vrs.VrsError):
        b.run(I[:1], vrs.knobs(precision="plain"))   # precision differs from the plan
    with pytest.raises(ValueError):
        b.run(I)
    b.close()
