"""Committed oracle fixtures (tests/golden/oracle_*.npz, written by
tests/golden/make_fixtures.py): the oracle must reproduce them (CPU), and the
GPU path must match them through the C-ABI -- shade counts and hashes bit for
bit, V, F, the fine-sky V and the adjoint gradient within the mixed-precision
tolerance (1e-5 relative L2)."""
import os

import numpy as np
import pytest

from viss_recon_shade_amd import workload as W

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["tiny", "small"]
TOL_MIXED = 1e-5


def load(name):
    with np.load(os.path.join(HERE, f"oracle_{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(oracle, name):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixtures",
                                                  os.path.join(HERE, "make_fixtures.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    fx, now = load(name), mk.fixture(name)
    for key in ("pairs", "mask_count", "mask_hash", "near_limb", "I", "raan", "m0"):
        assert np.array_equal(np.asarray(now[key]), fx[key]), key
    assert np.array_equal(now["resid"], fx["resid"])
    for key in ("pos", "uvw", "V", "F", "V_fine", "grad"):
        scale = np.abs(fx[key]).max()
        assert np.abs(np.asarray(now[key]) - fx[key]).max() <= 1e-12 * scale, key


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(ctx, name):
    import viss_recon_shade_amd as vrs
    fx = load(name)
    print(f"[golden] {name}", flush=True)
    g = ctx.geometry(int(fx["n_sat"]), int(fx["n_t"]), float(fx["freq_hz"]), int(fx["nside"]),
                     raan=fx["raan"], m0=fx["m0"])
    pos, uvw = g.MIncline()
    assert np.abs(pos - fx["pos"]).max() <= 1e-12 * W.A_KM
    sh = g.ShadeMatrix()
    assert np.array_equal(sh["count"], fx["mask_count"])
    assert np.array_equal(sh["hash"], fx["mask_hash"])
    assert sh["near_limb"] == int(fx["near_limb"])
    k = vrs.knobs("reference", precision="mixed")
    V = g.visscal(fx["I"], k)[0]
    assert rel(V[:, 0] + 1j * V[:, 1], fx["V"]) <= TOL_MIXED
    Vin = np.stack([fx["V"].real, fx["V"].imag], -1).astype(np.float32)[None]
    F = g.imagerecon(Vin, None, k)[0]
    assert rel(F, fx["F"]) <= TOL_MIXED
    # adjoint: FP64 gradient of the committed residual (canonical knobs, kappa = 2)
    r32 = np.stack([fx["resid"].real, fx["resid"].imag], -1).astype(np.float32)
    G = g.adjoint(r32, vrs.knobs(precision="mixed"))
    assert rel(G, fx["grad"]) <= TOL_MIXED
    cl, ff, nx = W.fine_sky(int(fx["nside"]), factor=int(fx["fine_factor"]), seed=21)
    g.fine_sky(cl, ff, nx)
    Vf = g.visscal_fine(k)
    assert rel(Vf[:, 0] + 1j * Vf[:, 1], fx["V_fine"]) <= TOL_MIXED
    g.close()
