"""torch.autograd.Function over the C-ABI: forward = vis-gen (vrs_visgen),
backward = the shaded back-projection (vrs_adjoint).

The forward model is affine in the sky: V(I) = A I + t_occ G with
(A I)_v = sum_p (1 - M_vp) I_p exp(-2 pi i s phi_vp).  For a real loss L the
gradient is dL/dI = A^H dL/dV, i.e. grad_p = sum_v (1 - M_vp) Re(g_v exp(+2 pi i s phi_vp))
with g = dL/dVr + i dL/dVi; vrs_adjoint(resid) returns 2 A^H resid, so the
backward passes resid = g / 2.  The reference has no gradient; this is the
differentiation layer the least-squares sky recovery runs on.
"""
import torch

from .api import knobs as make_knobs


class VisGen(torch.autograd.Function):
    """V[n_vis, 2] (float32, re/im) = visgen(I[n_pix]) on the geometry's device."""

    @staticmethod
    def forward(fctx, I, geom, k):
        if not (I.is_cuda and I.dtype == torch.float32 and I.dim() == 1):
            raise ValueError("I must be a 1-D float32 CUDA tensor")
        V = geom.visscal(I.contiguous().unsqueeze(0), k)[0]
        fctx.geom = geom
        fctx.k = k
        return V

    @staticmethod
    def backward(fctx, gV):
        if gV is None:
            return None, None, None
        resid = (0.5 * gV).contiguous()
        grad = fctx.geom.adjoint(resid, fctx.k)
        return grad, None, None


def visgen(I, geom, k=None):
    """Differentiable vis-gen: I [n_pix] -> V [n_vis, 2]."""
    return VisGen.apply(I, geom, k or make_knobs(precision="mixed"))


def loss_sq(I, geom, V_obs, k=None):
    """L(I) = sum_v |V(I)_v - V_obs_v|^2 with its autograd graph."""
    V = visgen(I, geom, k)
    return ((V - V_obs) ** 2).sum()
