"""Reference-compatibility rows on the GPU: uvw gridding / dedup / counts
(bit-exact integer work against the oracle), the reference's ShadeMatrix box
(bit-exact against the oracle), and the reference file layout written from
GPU results."""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import refio
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["tiny", "c1", "c2", "coarse"])
def test_grid_uvw_bit_exact(ctx, oracle, name):
    if name == "coarse":
        # 20 kHz: lambda = 15 km, so cells hold many samples and the counts matter
        wl = W.Workload("coarse", 8, 4000, 2e4, 8, [(1, 1.0, 0, 1.0, 1e3)])
    else:
        wl = W.CONFIGS[name]
    g = ctx.workload_geometry(wl)
    pos, _ = g.MIncline()
    out = g.grid_uvw(ordered=True, lambda_int=True)
    ordered = [(m, n) for m in range(wl.n_sat) for n in range(wl.n_sat) if m != n]
    lam = float(int(3e8 / wl.freq_hz))
    u, c, vc = oracle.grid_uvw(pos, ordered, lam, vis_pairs=g.pairs)
    print(f"[{name}] {len(c)} unique cells from {c.sum()} samples, max count {c.max()}", flush=True)
    if name == "coarse":
        assert c.max() > 10 and len(c) < c.sum() // 4
    assert np.array_equal(out["uvw"], u)
    assert np.array_equal(out["counts"], c)
    assert np.array_equal(out["vis_count"], vc)
    assert np.array_equal(out["weight"], (1.0 / vc.astype(np.float32)).astype(np.float32))
    # unordered baselines with c / f
    out2 = g.grid_uvw(ordered=False, lambda_int=False)
    u2, c2, _ = oracle.grid_uvw(pos, g.pairs, 299792458.0 / wl.freq_hz)
    assert np.array_equal(out2["uvw"], u2) and np.array_equal(out2["counts"], c2)
    g.close()


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_shade_box_matches_reference_formula(ctx, oracle, report, name):
    """Every row of the reference box (ShadeMatrix, shade10M.cu:62-105, 218-224)
    identical on the GPU and in the oracle: both take the angles in FP64, sum
    |uvw| in the same FP64 order and round each index once with the same tie
    rule."""
    wl = W.CONFIGS[name]
    g = ctx.workload_geometry(wl)
    uvw = g.grid_uvw(ordered=True, lambda_int=True, weights=False)["uvw"]
    for res in (2094, 20940):
        box = ctx.shade_box(uvw, res)
        ref = oracle.shade_box(uvw[:, 0], uvw[:, 1], uvw[:, 2], res)
        same = np.all(box == ref, axis=1)
        print(f"[shade box {name} res {res}] {len(uvw)} rows, {same.mean() * 100:.4f} % identical",
              flush=True)
        report(f"shade box {name} res {res} ({len(uvw)} gridded rows)", identical_rows=int(same.sum()),
               rows=len(uvw))
        assert same.all(), np.nonzero(~same)[0][:10]
    with pytest.raises(vrs.VrsError):
        ctx.shade_box(uvw, 1)
    g.close()


def test_reference_file_layout_round_trip(ctx, tmp_path):
    wl = W.CONFIGS["small"]
    g = ctx.workload_geometry(wl)
    gr = g.grid_uvw(ordered=True, lambda_int=True)
    box = ctx.shade_box(gr["uvw"], 2094)
    I = W.make_sky(wl)
    k = vrs.knobs("reference")
    F = g.imagerecon(g.visscal(I, k), gr["weight"], k)[0]
    p = refio.paths(str(tmp_path), wl.freq_hz, 1)
    refio.write_uvw(p["uvw"], gr["uvw"])
    refio.write_uvw(p["updated_uvw"], gr["uvw"], counts=gr["counts"].astype(np.float32))
    refio.write_shade(p["shadeM"], box)
    refio.write_image(p["F"], F)
    assert np.array_equal(refio.read_uvw(p["uvw"]), gr["uvw"])       # integer-valued
    u4, c4 = refio.read_uvw(p["updated_uvw"], n_cols=4)
    assert np.array_equal(c4, gr["counts"].astype(np.float32))
    sb = refio.read_shade(p["shadeM"])
    assert np.array_equal(sb[:, [0, 2]], np.minimum(box[:, [0, 2]], box[:, [1, 3]]) - 1)
    Fr = refio.read_image(p["F"])
    assert np.allclose(Fr, F, rtol=1e-5, atol=1e-6 * np.abs(F).max())
    g.close()
