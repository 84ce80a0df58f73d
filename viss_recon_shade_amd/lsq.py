"""Least-squares sky recovery on the GPU: minimise sum_v |V(I)_v - V_obs_v|^2
over I >= 0 with the vis-gen forward model and its hand-written adjoint.

Methods (all on torch CUDA tensors, every operator application is a HIP
kernel through the C-ABI):
  "pg": projected gradient with an exact line search.  The loss is quadratic
        in I, so along a direction d the minimiser is alpha = <g, d>/(2 |A d|^2);
        one extra forward per iteration with t_occ = 0 gives A d.
  "cg": projected conjugate gradient (Polak-Ribiere+ direction restricted to
        the free set, exact line search, projection after each step).
Each iteration records the loss; the step never increases it before projection.
"""
import math

import torch

from .api import knobs as make_knobs


def _fwd(geom, I, k):
    return geom.visscal(I.contiguous().unsqueeze(0), k)[0]


def recover(geom, V_obs, n_iter=100, method="pg", I0=None, k=None, callback=None):
    """Returns (I, history) with history = list of loss values (float)."""
    k = k or make_knobs(precision="mixed")
    k_lin = make_knobs(precision=k.precision, phase_scale=k.phase_scale, wfac=k.wfac, t_occ=0.0)
    dev = V_obs.device
    I = (torch.zeros(geom.n_pix, device=dev, dtype=torch.float32) if I0 is None
         else I0.clone().float().clamp_min(0))
    hist = []
    d_prev = None
    g_prev = None
    for it in range(n_iter):
        r = _fwd(geom, I, k) - V_obs                       # residual
        loss = float((r.double() ** 2).sum())
        hist.append(loss)
        g = geom.adjoint(r.contiguous(), k)                 # 2 A^H r
        # projected gradient: free set = pixels not pinned at the bound
        free = (I > 0) | (g < 0)
        gp = torch.where(free, g, torch.zeros_like(g))
        if method == "cg" and d_prev is not None:
            y = gp - g_prev
            beta = max(0.0, float((gp.double() * y.double()).sum() /
                                  max((g_prev.double() ** 2).sum().item(), 1e-300)))
            d = -gp + beta * torch.where(free, d_prev, torch.zeros_like(d_prev))
            if float((d.double() * gp.double()).sum()) >= 0:   # not a descent direction
                d = -gp
        else:
            d = -gp
        Ad = _fwd(geom, d, k_lin)
        den = 2.0 * float((Ad.double() ** 2).sum())
        num = -float((gp.double() * d.double()).sum())
        if den <= 0 or not math.isfinite(num / den) or num <= 0:
            break
        alpha = num / den
        I = (I + alpha * d).clamp_min(0)
        d_prev, g_prev = d, gp
        if callback:
            callback(it, I, loss)
    r = _fwd(geom, I, k) - V_obs
    hist.append(float((r.double() ** 2).sum()))
    return I, hist


def add_noise(V, snr, seed):
    """Complex Gaussian noise with per-visibility SNR: n = (|V|/snr)/sqrt2 (z1 + i z2),
    z from SplitMix64/Box-Muller with `seed`."""
    import numpy as np

    from . import _lib
    lib = _lib.load()
    n = V.shape[0]
    z = np.empty(2 * n, np.float64)
    _lib.check(lib.vrs_rng_normals(seed, 2 * n, z.ctypes.data))
    z = torch.from_numpy(z.reshape(n, 2)).to(V.device)
    amp = torch.sqrt((V.double() ** 2).sum(-1, keepdim=True)) / snr / math.sqrt(2.0)
    return (V.double() + amp * z).float()
