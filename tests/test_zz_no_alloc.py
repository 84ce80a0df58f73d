"""vrs_recon_noalloc: the allocation-free recon entry point, through the C-ABI.

Its partial images live in a fixed per-device scratch array instead of a
context workspace, on the same kernels, so its output must equal vrs_recon's
bit for bit.  Two calls in a row on the same device share that scratch; the
second, on a smaller input, must not see the first's contents.
"""
import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import _lib
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu


def _device_inputs(torch, g, n_maps, seed):
    rng = np.random.default_rng(seed)
    I = rng.lognormal(0.0, 1.0, (n_maps, g.n_pix)).astype(np.float32)
    It = torch.from_numpy(I).cuda()
    V = g.visscal(It, vrs.knobs(precision="mixed"))
    w = torch.from_numpy((1.0 / (1.0 + np.arange(g.n_vis) % 3)).astype(np.float32)).cuda()
    return V, w


@pytest.mark.parametrize("prec", ["mixed", "plain"])
def test_noalloc_equals_recon_twice_second_smaller(ctx, prec):
    torch = pytest.importorskip("torch")
    k = vrs.knobs(precision=prec)
    big = ctx.workload_geometry(W.CONFIGS["c1"])      # 40 320 visibilities, Nside 32
    small = ctx.workload_geometry(W.CONFIGS["small"])  # 576 visibilities, Nside 16
    print(f"[noalloc {prec}] c1 x 3 maps, then small x 1 map", flush=True)
    V1, w1 = _device_inputs(torch, big, 3, 1)
    V2, w2 = _device_inputs(torch, small, 1, 2)
    F1_ref = big.imagerecon(V1, w1, k)
    F2_ref = small.imagerecon(V2, None, k)
    torch.cuda.synchronize()
    F1 = big.imagerecon_noalloc(V1, w1, k)
    F2 = small.imagerecon_noalloc(V2, None, k)
    torch.cuda.synchronize()
    assert torch.equal(F1, F1_ref)
    assert torch.equal(F2, F2_ref)
    # and back to the larger input, into a caller-provided output
    out = torch.full_like(F1_ref, float("nan"))
    big.imagerecon_noalloc(V1, w1, k, out=out)
    torch.cuda.synchronize()
    assert torch.equal(out, F1_ref)
    big.close()
    small.close()


def test_noalloc_rejects_host_arrays_and_oversize(ctx):
    torch = pytest.importorskip("torch")
    lib = _lib.load()
    g = ctx.workload_geometry(W.CONFIGS["c1"])
    k = vrs.knobs(precision="mixed")
    Vh = np.zeros((g.n_vis, 2), np.float32)
    Fh = np.zeros(g.n_pix, np.float32)
    st = lib.vrs_recon_noalloc(ctx.handle, g.handle, vrs.api.C.byref(k), 1,
                               Vh.ctypes.data, None, Fh.ctypes.data)
    assert st == -1  # VRS_E_ARG: host arrays would need staging
    # more maps than the scratch holds: refused on the host before any launch
    n_maps = 64
    Vd = torch.zeros((n_maps, g.n_vis, 2), device="cuda")
    with pytest.raises(vrs.VrsError, match="vrs_recon_noalloc holds"):
        g.imagerecon_noalloc(Vd, None, k)
    with pytest.raises(vrs.VrsError):
        g.imagerecon_noalloc(Vd[:1], None, vrs.knobs(precision="mfma"))
    g.close()
