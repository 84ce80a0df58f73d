"""GPU parity of the four stages against the FP64 oracle, through the C-ABI.

The oracle is fed the geometry the GPU used (positions and uvw exported by
vrs_orbit_uvw, lmn by vrs_geom_lmn), so the shade mask must match bit for bit
and vis-gen / recon within the stated relative L2 tolerances:
  VRS_MIXED_F64PHASE  <= 1e-5 (tiny/small/c1), <= 1e-4 at 10 MHz
  VRS_PLAIN_F32       <= 2e-2 (the reference's FP32 phase; the CPU FP32
                       restatement has the same error class); at 10 MHz
                       <= 3e-2 and <= 1.25x the CPU FP32 restatement's own
                       error on the same sample
"""
import math

import numpy as np
import pytest

import viss_recon_shade_amd as vrs
from viss_recon_shade_amd import workload as W

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-5
TOL_MIXED_10MHZ = 1e-4
TOL_PLAIN = 2e-2
# 10 MHz plain FP32: |u| reaches 1.4e5 wavelengths, so an FP32 phase carries
# ~2^-7 revolutions of rounding; the CPU restatement of the reference's own FP32
# arithmetic measures 2.6e-2 on the c2 sample below, so the bound there is 3e-2
# together with "no worse than 1.25x the reference arithmetic on the same sample"
TOL_PLAIN_10MHZ = 3e-2


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cplx(V):
    return V[..., 0].astype(np.float64) + 1j * V[..., 1].astype(np.float64)


_GEOMS = {}


def geom(ctx, name):
    if name not in _GEOMS:
        wl = W.CONFIGS[name]
        print(f"[geom] {name}: n_sat={wl.n_sat} n_t={wl.n_t} nside={wl.nside}", flush=True)
        g = ctx.workload_geometry(wl)
        pos, uvw = g.MIncline()
        _GEOMS[name] = (wl, g, pos, uvw, g.lmn(), W.make_sky(wl))
    return _GEOMS[name]


# ------------------------------------------------------------------ stage 1
@pytest.mark.parametrize("name", ["tiny", "c1"])
def test_orbit_uvw_matches_oracle(ctx, oracle, name):
    wl, g, pos, uvw, lmn, _ = geom(ctx, name)
    opos = oracle.orbit_positions(wl.n_t, W.DT_S, W.A_KM, W.INCL_RAD, W.MU_MOON, g.raan, g.m0)
    assert np.abs(pos - opos).max() <= 1e-12 * W.A_KM
    ouvw = oracle.uvw(pos, g.pairs, 299792458.0 / wl.freq_hz)
    assert np.abs(uvw - ouvw).max() <= 1e-12 * np.abs(ouvw).max()
    assert np.abs(lmn - oracle.healpix_ring(wl.nside)).max() <= 2.3e-16


# ------------------------------------------------------------------ stage 2
@pytest.mark.parametrize("name", ["tiny", "small"])
def test_shade_bitmap_bit_exact(ctx, oracle, name):
    wl, g, pos, uvw, lmn, _ = geom(ctx, name)
    out = g.ShadeMatrix(bits=True)
    ref = oracle.mask_bits(pos, g.pairs, lmn, g.r_moon_km)
    assert np.array_equal(out["bits"], ref)
    cnt, h, near = oracle.mask_stats(pos, g.pairs, lmn, g.r_moon_km)
    assert np.array_equal(out["count"], cnt)
    assert np.array_equal(out["hash"], h)
    assert out["near_limb"] == near
    assert out["shaded_pairs"] == int(cnt.sum())


def test_shade_counts_and_hashes_c1(ctx, oracle):
    wl, g, pos, uvw, lmn, _ = geom(ctx, "c1")
    out = g.ShadeMatrix()
    cnt, h, near = oracle.mask_stats(pos, g.pairs, lmn, g.r_moon_km)
    assert np.array_equal(out["count"], cnt)
    assert np.array_equal(out["hash"], h)
    assert out["near_limb"] == near
    print(f"[c1 shade] shaded fraction {cnt.sum() / (wl.n_vis * wl.n_pix):.4f} "
          f"near-limb {near} fp64 tests {out['fp64_tests']}", flush=True)


# ------------------------------------------------------------------ stage 3 + 4
CASES = [("tiny", "plain", "canonical"), ("tiny", "mixed", "canonical"),
         ("small", "plain", "reference"), ("small", "mixed", "reference"),
         ("small", "mixed", "canonical")]


@pytest.mark.parametrize("name,prec,preset", CASES)
def test_visgen_recon_parity_small(ctx, oracle, name, prec, preset):
    wl, g, pos, uvw, lmn, I = geom(ctx, name)
    k = vrs.knobs(preset, precision=prec)
    V = g.visscal(I, k)
    Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I, g.r_moon_km, s=k.phase_scale, wfac=k.wfac,
                       t_occ=k.t_occ)
    tol = TOL_MIXED if prec == "mixed" else TOL_PLAIN
    e_v = rel(cplx(V), Vo)
    assert e_v <= tol, e_v
    w = (1.0 / (1.0 + np.arange(wl.n_vis) % 5)).astype(np.float32)
    F = g.imagerecon(V, w, k)
    Fo = oracle.recon(pos, g.pairs, uvw, lmn, cplx(V), g.r_moon_km, s=k.phase_scale,
                      wfac=k.wfac, weights=w)
    e_f = rel(F, Fo)
    assert e_f <= tol, e_f
    print(f"[{name} {prec} {preset}] V rel {e_v:.2e}  F rel {e_f:.2e}", flush=True)


@pytest.mark.parametrize("prec", ["plain", "mixed"])
def test_c1_parity_subsets(ctx, oracle, prec):
    wl, g, pos, uvw, lmn, I = geom(ctx, "c1")
    k = vrs.knobs(precision=prec)
    V = g.visscal(I, k)
    sel = np.arange(0, wl.n_vis, 97, dtype=np.int64)
    Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I, g.r_moon_km, vis_sel=sel)
    e_v = rel(cplx(V)[0, sel], Vo[0])
    F = g.imagerecon(V, None, k)
    psel = np.arange(0, wl.n_pix, 61, dtype=np.int64)
    Fo = oracle.recon(pos, g.pairs, uvw, lmn, cplx(V), g.r_moon_km, pix_sel=psel)
    e_f = rel(F[0, psel], Fo[0])
    print(f"[c1 {prec}] V rel {e_v:.2e}  F rel {e_f:.2e}", flush=True)
    tol = TOL_MIXED if prec == "mixed" else TOL_PLAIN
    assert e_v <= tol and e_f <= tol, (e_v, e_f)


def test_adjoint_dot_product_and_oracle(ctx, oracle):
    wl, g, pos, uvw, lmn, I = geom(ctx, "small")
    k = vrs.knobs(precision="mixed", t_occ=0.0)
    rng = np.random.default_rng(7)
    y = rng.standard_normal((wl.n_vis, 2)).astype(np.float32)
    AI = cplx(g.visscal(I, k))[0]
    grad = g.adjoint(y, k)
    # grad = 2 A^H y  and  Re<A I, y> = <I, A^H y>
    lhs = float(np.real(np.vdot(AI, cplx(y))))
    rhs = 0.5 * float(I[0].astype(np.float64) @ grad.astype(np.float64))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    go = oracle.recon(pos, g.pairs, uvw, lmn, cplx(y), g.r_moon_km, kappa=2.0)[0]
    assert rel(grad, go) <= TOL_MIXED


def test_wfac_convention_leaves_image_invariant(ctx):
    wl, g, pos, uvw, lmn, I = geom(ctx, "small")
    F0 = g.imagerecon(g.visscal(I, vrs.knobs(precision="mixed")), None,
                      vrs.knobs(precision="mixed"))
    k1 = vrs.knobs(precision="mixed", wfac=1)
    F1 = g.imagerecon(g.visscal(I, k1), None, k1)
    assert rel(F1, F0) <= 1e-5


def test_multi_map_equals_single_maps(ctx):
    wl, g, pos, uvw, lmn, I = geom(ctx, "small")
    I3 = np.stack([I[0], 2 * I[0], np.roll(I[0], 5)]).astype(np.float32)
    k = vrs.knobs(precision="mixed")
    V3 = g.visscal(I3, k)
    for m in range(3):
        Vm = g.visscal(np.ascontiguousarray(I3[m:m + 1]), k)
        assert np.array_equal(V3[m], Vm[0])
    F3 = g.imagerecon(V3, None, k)
    assert np.array_equal(F3[1], g.imagerecon(np.ascontiguousarray(V3[1:2]), None, k)[0])


def test_device_pointers_match_host_pointers(ctx):
    torch = pytest.importorskip("torch")
    wl, g, pos, uvw, lmn, I = geom(ctx, "small")
    k = vrs.knobs(precision="mixed")
    Vh = g.visscal(I, k)
    It = torch.from_numpy(I).cuda()
    Vd = g.visscal(It, k)
    torch.cuda.synchronize()
    assert np.array_equal(Vd.cpu().numpy(), Vh)
    Fd = g.imagerecon(Vd, None, k)
    torch.cuda.synchronize()
    assert np.array_equal(Fd.cpu().numpy(), g.imagerecon(Vh, None, k))


def test_argument_errors_on_gpu(ctx):
    wl, g, pos, uvw, lmn, I = geom(ctx, "tiny")
    with pytest.raises(vrs.VrsError) as e:
        g.visscal(I, vrs.knobs(phase_scale=float("nan")))
    assert e.value.code == "VRS_E_ARG"
    with pytest.raises(vrs.VrsError):
        g.visscal(I, vrs.knobs(wfac=3))
    with pytest.raises(vrs.VrsError):
        ctx.geometry(3, 8, 1e6, 3)            # nside not a power of two
    with pytest.raises(vrs.VrsError):
        ctx.geometry(3, 8, 1e6, 4, a_km=1000.0)  # orbit inside the Moon
    with pytest.raises(vrs.VrsError):
        ctx.geometry(3, 8, 1e6, 4, pairs=[(0, 0)])
    with pytest.raises(ValueError):
        g.visscal(np.zeros((1, 7), np.float32))


# ------------------------------------------------------------------ 10 MHz at scale
def test_c2_full_size_sampled_parity(ctx, oracle, report):
    wl, g, pos, uvw, lmn, I = geom(ctx, "c2")
    k = vrs.knobs(precision="mixed")
    V = g.visscal(I, k)
    assert np.isfinite(V).all()
    sel = np.linspace(0, wl.n_vis - 1, 48).astype(np.int64)
    Vo = oracle.visgen(pos, g.pairs, uvw, lmn, I, g.r_moon_km, vis_sel=sel)
    e_v = rel(cplx(V)[0, sel], Vo[0])
    F = g.imagerecon(V, None, k)
    psel = np.linspace(0, wl.n_pix - 1, 12).astype(np.int64)
    Fo = oracle.recon(pos, g.pairs, uvw, lmn, cplx(V), g.r_moon_km, pix_sel=psel)
    e_f = rel(F[0, psel], Fo[0])
    print(f"[c2 mixed] V rel {e_v:.2e}  F rel {e_f:.2e}", flush=True)
    assert e_v <= TOL_MIXED_10MHZ and e_f <= TOL_MIXED_10MHZ, (e_v, e_f)
    # plain FP32 at 10 MHz: the reference's FP32 arithmetic (vissGen10M_shade_optim1.cu:
    # 183-195, complexExp of vissGen10M_shade.cu:35-39) restated on the host cores has its
    # own error against FP64 on the same sample; the GPU plain path must sit in that
    # error class (<= 1.25x of it) and under the stated 10 MHz bound
    kp = vrs.knobs(precision="plain")
    Vp = g.visscal(I, kp)
    e_p = rel(cplx(Vp)[0, sel], Vo[0])
    Vcpu = oracle.cpu_visgen_f32(pos, g.pairs, uvw, lmn, I[0], g.r_moon_km, vis_sel=sel)
    e_cpu = rel(Vcpu, Vo[0])
    print(f"[c2 plain] V rel {e_p:.2e} (CPU FP32 restatement {e_cpu:.2e})", flush=True)
    report("c2 plain FP32 V (48 vis x all pix vs FP64)", gpu_rel_l2=e_p, cpu_f32_rel_l2=e_cpu)
    assert e_p <= TOL_PLAIN_10MHZ and e_p <= 1.25 * e_cpu, (e_p, e_cpu)
    # mask totals: every visibility shades between one and two caps' worth of pixels
    st = g.ShadeMatrix()
    c = math.sqrt(1 - (W.R_MOON_KM / W.A_KM) ** 2)
    frac = st["count"] / wl.n_pix
    assert frac.min() >= (1 - c) / 2 - 0.01 and frac.max() <= (1 - c) + 0.01
    ssel = np.arange(0, wl.n_vis, 4099, dtype=np.int64)
    cnt, h, _ = oracle.mask_stats(pos, g.pairs, lmn, g.r_moon_km, vis_sel=ssel)
    assert np.array_equal(st["count"][ssel], cnt) and np.array_equal(st["hash"][ssel], h)


def test_c2_full_size_linearity_and_adjoint(ctx):
    """Size-independent properties on every c2 output: with T_occ = 0 vis-gen is
    linear in I (dyadic intensities keep 2 I1 - 3 I2 exact in FP32), and the
    adjoint satisfies <A I, y> = <I, A^H y> with A^H = vrs_adjoint / 2."""
    wl, g, pos, uvw, lmn, I = geom(ctx, "c2")
    rng = np.random.default_rng(31)
    I1 = (rng.integers(0, 1024, wl.n_pix) / 256).astype(np.float32)
    I2 = (rng.integers(0, 1024, wl.n_pix) / 256).astype(np.float32)
    k = vrs.knobs(precision="mixed", t_occ=0.0)
    V = cplx(g.visscal(np.stack([I1, I2, 2 * I1 - 3 * I2]), k))
    e_lin = rel(V[2], 2 * V[0] - 3 * V[1])
    print(f"[c2 linearity] rel {e_lin:.2e}", flush=True)
    assert e_lin <= 1e-5
    y = (rng.standard_normal(wl.n_vis) + 1j * rng.standard_normal(wl.n_vis)).astype(np.complex64)
    y2 = np.stack([y.real, y.imag], -1).astype(np.float32)
    AhY = g.adjoint(y2, k).astype(np.float64) / 2.0   # gradient = 2 Re(A^H r)
    lhs = np.real(np.vdot(y.astype(np.complex128), V[0]))
    rhs = float(np.dot(I1.astype(np.float64), AhY))
    e_adj = abs(lhs - rhs) / abs(lhs)
    print(f"[c2 adjoint identity] rel {e_adj:.2e}", flush=True)
    assert e_adj <= 1e-5
