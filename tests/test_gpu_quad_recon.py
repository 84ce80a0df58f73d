"""One-map matrix-core vis-gen and recon on per-satellite factors (precision
"mfma", one map, an orbit geometry of at most 16 satellites): vis-gen as Gram
matrices of the step's satellite factors over the pixels (masked and weighted by
I - t_occ, plus the full-sky sums), recon as a Hermitian form in them, F =
sum_x c_x Zr_x + s_x Zi_x with Z = W e.

Vis-gen: every visibility against the FP64-phase VALU kernel and against the
multi-map matrix-core kernels (map 0 of two), relative L2 2e-5, worst 5e-5.

Recon: every output against the FP64-phase VALU recon (which the oracle pins) and
against the multi-map matrix-core kernels run on the same V as map 0 of two:
relative L2 over every pixel 5e-5, worst max-abs / max 1e-4 (the bounds of the
other matrix-core recon tests).  Geometries: c1 (8 satellites), 16 satellites
at Nside 32 (step segments of which the last are empty), an explicit pair list
in reverse order with a duplicated pair, orbit windows with zero-filled steps,
skies with padding pixels and an odd step count, both knob presets, with and
without weights."""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def worst(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def check(g, report, name, seed, preset="canonical", weights=True):
    rng = np.random.default_rng(seed)
    km = vrs.knobs(preset, precision="mfma")
    kx = vrs.knobs(preset, precision="mixed")
    V = rng.standard_normal((g.n_vis, 2)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, g.n_vis).astype(np.float32) if weights else None
    Fq = g.imagerecon(V, w, km).astype(np.float64)[0]
    Fx = g.imagerecon(V, w, kx).astype(np.float64)[0]
    V2 = np.stack([V, rng.standard_normal((g.n_vis, 2)).astype(np.float32)])
    Fm = g.imagerecon(V2, w, km).astype(np.float64)[0]
    e, wo, em = rel(Fq, Fx), worst(Fq, Fx), rel(Fq, Fm)
    report(f"one-map quadratic-form recon {name}", F_rel=e, F_worst=wo, F_vs_multimap=em)
    print(f"[quad {name}] F {e:.2e} (worst {wo:.2e}, vs multi-map {em:.2e})", flush=True)
    assert e <= 5e-5 and wo <= 1e-4, (e, wo)
    assert em <= 5e-5, em
    for t_occ in (240.0, 0.0):
        km = vrs.knobs(preset, precision="mfma", t_occ=t_occ)
        kx = vrs.knobs(preset, precision="mixed", t_occ=t_occ)
        I = rng.lognormal(0.0, 1.0, (1, g.n_pix)).astype(np.float32)
        Vg, Vx = g.visscal(I, km)[0].astype(np.float64), g.visscal(I, kx)[0].astype(np.float64)
        I2 = np.concatenate([I, rng.lognormal(0.0, 1.0, (1, g.n_pix)).astype(np.float32)])
        Vm = g.visscal(I2, km)[0].astype(np.float64)
        ev, wv, evm = rel(Vg, Vx), worst(Vg, Vx), rel(Vg, Vm)
        report(f"one-map Gram vis-gen {name} t_occ={t_occ:g}", V_rel=ev, V_worst=wv, V_vs_multimap=evm)
        print(f"[gram {name} t_occ={t_occ:g}] V {ev:.2e} (worst {wv:.2e}, vs multi-map {evm:.2e})",
              flush=True)
        assert ev <= 2e-5 and wv <= 5e-5, (ev, wv)
        assert evm <= 2e-5, evm


@pytest.mark.parametrize("preset", ["canonical", "reference"])
def test_quad_recon_c1(ctx, report, preset):
    print(f"[quad] c1, {preset} preset", flush=True)
    g = ctx.workload_geometry(W.CONFIGS["c1"])
    check(g, report, f"c1 {preset}", 21, preset=preset)
    check(g, report, f"c1 {preset} unweighted", 22, preset=preset, weights=False)
    g.close()


def test_quad_recon_sixteen_satellites(ctx, report):
    print("[quad] 16 satellites, 1 MHz, 96 steps, Nside 32", flush=True)
    g = ctx.geometry(16, 96, 1e6, 32)
    check(g, report, "16 satellites Nside 32", 23)
    g.close()


def test_quad_recon_pair_lists_and_windows(ctx, report):
    print("[quad] reversed pairs with a duplicate; zero-filled steps", flush=True)
    g = ctx.geometry(4, 40, 3e6, 16, pairs=[(2, 0), (3, 1), (1, 2), (3, 0), (2, 0)])
    check(g, report, "reversed pairs", 24)
    g.close()
    idx = np.arange(60, dtype=np.int64) * 7
    idx[50:] = -1
    g = ctx.geometry(6, 60, 3e6, 16, sample_index=idx)
    check(g, report, "windows", 25)
    g.close()


@pytest.mark.parametrize("nside,n_t", [(2, 41), (8, 40)])
def test_quad_recon_padding(ctx, report, nside, n_t):
    print(f"[quad] 16 satellites, Nside {nside}, {n_t} steps", flush=True)
    g = ctx.geometry(16, n_t, 1e7, nside)
    check(g, report, f"Nside {nside} {n_t} steps", 26 + nside)
    g.close()
