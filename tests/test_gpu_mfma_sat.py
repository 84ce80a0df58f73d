"""Matrix-core kernels on per-satellite phase factors (knob sat_factors = 1,
the default: E_xy = E_y conj(E_x), each satellite's own occultation zeroing its
factor) on every output, against the FP64-phase VALU kernels and against the
per-entry matrix-core kernels (sat_factors = 0).

Geometries: c1 (8 satellites, 28 baselines: four steps per vis-gen block), the
4-satellite `small` geometry, the c2 orbits at Nside 32 (16 satellites: two
steps per block), an explicit baseline list in reverse order, and orbit windows
with zero-filled steps.  T_occ = 240 exercises the full-sky Gram sums of the
shaded-pixel term, T_occ = 0 the kernel without it.  Bounds: relative L2 over
every output 2e-5 (vis-gen) and 5e-5 (recon), worst per-map max-abs / max 5e-5
and 1e-4, as the per-entry matrix-core kernels' every-output tests."""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def worst(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b).max(-1) / np.abs(b).max(-1)))


def cplx(V):
    return V[..., 0].astype(np.float64) + 1j * V[..., 1].astype(np.float64)


def check(g, I, report, name, preset="canonical", t_occ=240.0):
    km = vrs.knobs(preset, precision="mfma", t_occ=t_occ)
    k0 = vrs.knobs(preset, precision="mfma", t_occ=t_occ, sat_factors=0)
    kx = vrs.knobs(preset, precision="mixed", t_occ=t_occ)
    Vs = cplx(g.visscal(I, km))
    Vx = cplx(g.visscal(I, kx))
    V0 = cplx(g.visscal(I, k0))
    ev, wv, e0 = rel(Vs, Vx), worst(Vs, Vx), rel(Vs, V0)
    # recon of one V (the FP64-phase one) by both kernels
    Vin = g.visscal(I, kx)
    Fs = g.imagerecon(Vin, None, km).astype(np.float64)
    Fx = g.imagerecon(Vin, None, kx).astype(np.float64)
    ef, wf = rel(Fs, Fx), worst(Fs, Fx)
    report(f"sat factors {name}", V_rel=ev, V_worst=wv, V_vs_entry=e0, F_rel=ef, F_worst=wf)
    print(f"[sat {name}] V {ev:.2e} (worst {wv:.2e}, vs per-entry {e0:.2e}) F {ef:.2e} "
          f"(worst {wf:.2e})", flush=True)
    assert ev <= 2e-5 and wv <= 5e-5, (ev, wv)
    assert e0 <= 2e-5, e0
    assert ef <= 5e-5 and wf <= 1e-4, (ef, wf)


@pytest.mark.parametrize("t_occ", [240.0, 0.0])
def test_sat_factors_c1_every_output(ctx, report, t_occ):
    print(f"[sat] c1, 64 maps, t_occ {t_occ}", flush=True)
    wl = W.CONFIGS["c1"]
    g = ctx.workload_geometry(wl)
    rng = np.random.default_rng(11)
    I = rng.lognormal(0.0, 1.0, (64, wl.n_pix)).astype(np.float32)
    check(g, I, report, f"c1 t_occ={t_occ:g}", t_occ=t_occ)
    g.close()


def test_sat_factors_small_reference_preset(ctx, report):
    print("[sat] small geometry, 40 maps, reference conventions", flush=True)
    wl = W.CONFIGS["small"]
    g = ctx.workload_geometry(wl)
    rng = np.random.default_rng(12)
    I = np.exp(rng.standard_normal((40, wl.n_pix)) * 2.0).astype(np.float32)
    check(g, I, report, "small reference", preset="reference")
    g.close()


def test_sat_factors_sixteen_satellites(ctx, report):
    print("[sat] c2 orbits (16 satellites), 1 MHz, 96 steps, Nside 32, 70 maps", flush=True)
    g = ctx.geometry(16, 96, 1e6, 32)
    rng = np.random.default_rng(13)
    I = rng.lognormal(0.0, 1.0, (70, g.n_pix)).astype(np.float32)
    check(g, I, report, "16 satellites, 70 maps")
    g.close()


def test_sat_factors_reversed_pairs_and_windows(ctx, report):
    print("[sat] explicit reversed baselines; orbit windows with zero-filled steps", flush=True)
    rng = np.random.default_rng(14)
    g = ctx.geometry(4, 40, 3e6, 16, pairs=[(2, 0), (3, 1), (1, 2), (3, 0)])
    I = rng.lognormal(0.0, 1.0, (8, g.n_pix)).astype(np.float32)
    check(g, I, report, "reversed pairs")
    g.close()
    idx = np.arange(60, dtype=np.int64) * 7
    idx[50:] = -1  # zero-filled samples
    g = ctx.geometry(6, 60, 3e6, 16, sample_index=idx)
    I = rng.lognormal(0.0, 1.0, (8, g.n_pix)).astype(np.float32)
    check(g, I, report, "windows")
    g.close()


@pytest.mark.parametrize("nside,n_t", [(4, 40), (2, 40), (8, 41)])
def test_sat_factors_star_layout_with_padding(ctx, report, nside, n_t):
    """16 satellites (star-ordered waves and their K-step skips) on skies smaller
    than one 256-pixel chunk, so the stages hold padding pixels, and with an odd
    step count, so the last block holds one step (its skip masks from one row)."""
    print(f"[sat] 16 satellites, Nside {nside}, {n_t} steps", flush=True)
    rng = np.random.default_rng(15 + nside)
    g = ctx.geometry(16, n_t, 1e7, nside)
    I = rng.lognormal(0.0, 1.0, (24, g.n_pix)).astype(np.float32)
    for t_occ in (240.0, 0.0):
        check(g, I, report, f"16 satellites, Nside {nside}, {n_t} steps, t_occ={t_occ:g}", t_occ=t_occ)
    g.close()
