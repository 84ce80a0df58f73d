"""Writes the committed oracle fixtures (tests/golden/*.npz).

The reference ships no test vectors (its data files are git-ignored and its
output folders hold only placeholders), so these fixtures are the FP64
oracle's outputs on small configurations, each cross-checked by the closed-form
and brute-force tests in tests/test_oracle.py.  They lock the oracle against
regressions and let GPU tests compare against stored vectors.

    python tests/golden/make_fixtures.py        # rewrites tests/golden/*.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from viss_recon_shade_amd import workload as W  # noqa: E402

C_LIGHT = 299792458.0
# reference conventions: s = 1/dl = 0.5, trailing exp(-2 pi i s w), T_occ = 240
S, WFAC, T_OCC = 0.5, 1, 240.0


def fixture(name, fine_factor=2):
    wl = W.CONFIGS[name]
    raan, m0 = W.orbit_params(wl.n_sat)
    pos = oracle.orbit_positions(wl.n_t, W.DT_S, W.A_KM, W.INCL_RAD, W.MU_MOON, raan, m0)
    pairs = W.all_pairs(wl.n_sat)
    uvw = oracle.uvw(pos, pairs, C_LIGHT / wl.freq_hz)
    lmn = oracle.healpix_ring(wl.nside)
    I = W.make_sky(wl)[0]
    count, hsh, near = oracle.mask_stats(pos, pairs, lmn, W.R_MOON_KM)
    V = oracle.visgen(pos, pairs, uvw, lmn, I, W.R_MOON_KM, s=S, wfac=WFAC, t_occ=T_OCC)[0]
    F = oracle.recon(pos, pairs, uvw, lmn, V[None], W.R_MOON_KM, s=S, wfac=WFAC)[0]
    # adjoint (gradient of sum |V(I) - V_obs|^2) for a fixed residual, canonical knobs
    rng = np.random.default_rng(17)
    resid = (rng.standard_normal(len(V)) + 1j * rng.standard_normal(len(V))).astype(np.complex64)
    grad = oracle.recon(pos, pairs, uvw, lmn, resid.astype(np.complex128)[None], W.R_MOON_KM,
                        kappa=2.0)[0]
    cl, ff, nx = W.fine_sky(wl.nside, factor=fine_factor, seed=21)
    Vf = oracle.visgen_fine(pos, pairs, uvw, lmn, cl, ff, nx, W.R_MOON_KM, s=S, wfac=WFAC,
                            t_occ=T_OCC)
    return dict(n_sat=wl.n_sat, n_t=wl.n_t, freq_hz=wl.freq_hz, nside=wl.nside, raan=raan,
                m0=m0, fine_factor=fine_factor, I=I, pos=pos, pairs=pairs, uvw=uvw,
                mask_count=count, mask_hash=hsh, near_limb=near, V=V, F=F, V_fine=Vf,
                resid=resid, grad=grad)


def main():
    for name in ("tiny", "small"):
        path = os.path.join(HERE, f"oracle_{name}.npz")
        np.savez_compressed(path, **fixture(name))
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
