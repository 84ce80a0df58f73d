"""Benchmark of the hot path on one MI355X.

Workload (the 10 MHz plain-FP32 configuration at scale): 16 satellites on
circular 300 km lunar orbits -> 120 baselines x 8640 steps (1 036 800
visibilities), HEALPix Nside 128 sky (196 608 pixels, log-normal(0, 1.5) plus
200 S^-2 point sources), synthetic.  One step = stage 1 (orbits, caps, uvw on
the device) + vis-gen + recon, inputs resident in HBM.
value = (vis-gen terms + recon terms) / step time, terms = N_vis x N_pix.

The plain step pairs each pixel with its antipode (one phase and sin/cos serve
both terms, vrs_knobs.antipodal); the same step with every term on its own is
reported beside it (plain_direct).
Also reported: the FP64-phase path on the same workload, the relative error
of both against the FP64 oracle on sampled outputs, per-kernel times from HIP
events on the context stream, the dominant kernel's roofline fraction, and the
FP32 OpenMP CPU restatement timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3          # MI355X published FP32 vector peak: packed FP32 FMA (2 per lane)
PEAK_FP32_VALU_TFLOPS = 78.6      # the same VALU issuing one FP32 FMA per lane per clock, which
                                  # is what the plain kernels issue (built without packed FP32)
TRANS_PEAK = 256 * 2.4e9 * 16     # v_sin/v_cos issue: 16 lanes/clk/CU (quarter rate)
FLOPS_PER_TERM = 10               # 3 FMA phase + 2 FMA accumulate
N_CU, N_XCD = 256, 8              # SQ counters sum over CUs, GRBM_GUI_ACTIVE over XCDs


def visgen_name(f64, vt, pair=False):
    """Kernel name of a vis-gen tile configuration (pair: antipodal pixel pairs)."""
    if pair:
        return f"visgen_pair_kernel<{'true' if f64 else 'false'}, {vt}>"
    return f"visgen_kernel<{'true' if f64 else 'false'}, {vt}, false>"


def algorithmic_bytes(wl, kernel):
    """Minimum bytes one plain-FP32 launch moves: uvw (FP64, 24 B) and the
    satellites' caps (32 B per satellite and step) once, V (8 B) once, the sky
    (l, m, n FP32 + I for vis-gen; l, m, n + F + 8 partial images for recon)."""
    base = wl.n_vis * 24 + wl.n_sat * wl.n_t * 32 + wl.n_vis * 8
    if kernel == "visgen":
        return base + wl.n_pix * 16
    return base + wl.n_pix * (12 + 4 + 8 * 4)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=32)
    ap.add_argument("--mask-stride", type=int, default=32,
                    help="compare the shade counts and hashes of every n-th visibility "
                         "with the oracle (1: all of them)")
    ap.add_argument("--no-mfma", action="store_true", help="skip the multi-map matrix-core legs")
    ap.add_argument("--no-batch", action="store_true", help="skip the multi-channel batch leg")
    ap.add_argument("--no-c5", action="store_true", help="skip the least-squares recovery leg")
    ap.add_argument("--no-grid", action="store_true", help="skip the uvw gridding leg")
    ap.add_argument("--grid-ref", action="store_true",
                    help="also grid at the reference's own stage-1 scale (8 satellites x 40.32 M "
                         "windowed samples, 56 ordered pairs: 2.26e9 keys, ~110 GB of HBM)")
    return ap.parse_args()


def timed_steps(torch, step, steps, warmup):
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def kernel_avg_ms(ctx, step, names, reps):
    ctx.timing(True)
    acc = {n: 0.0 for n in names}
    for _ in range(reps):
        step()
        for n in names:
            acc[n] += ctx.last_ms(n)
    ctx.timing(False)
    return {n: acc[n] / reps for n in names}


def cpu_model():
    """CPU model name of the host (first 'model name' of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_error(g, wl, pos, uvw, lmn, I, V, refs):
    """Error of the FP32 CPU restatement against the FP64 oracle on the same
    deterministic samples as the GPU accuracy figures (refs from accuracy())."""
    import oracle as o
    Vc = o.cpu_visgen_f32(pos, g.pairs, uvw, lmn, I, g.r_moon_km, vis_sel=refs["vis_sel"])
    Fc = o.cpu_recon_f32(pos, g.pairs, uvw, lmn, V, g.r_moon_km, pix_sel=refs["pix_sel"])
    return {"V_rel_l2": float(np.linalg.norm(Vc - refs["Vo"]) / np.linalg.norm(refs["Vo"])),
            "F_rel_l2": float(np.linalg.norm(Fc - refs["Fo"]) / np.linalg.norm(refs["Fo"])),
            "sample": "the plain-FP32 accuracy sample, F from the GPU's V"}


def cpu_baseline(g, wl, pos, uvw, lmn, I, V, stride):
    import oracle as o
    sel = np.arange(0, wl.n_vis, stride, dtype=np.int64)
    psel = np.arange(0, wl.n_pix, stride, dtype=np.int64)
    o.cpu_visgen_f32(pos, g.pairs, uvw, lmn, I, g.r_moon_km, vis_sel=sel[:8])  # warm
    t0 = time.perf_counter()
    o.cpu_visgen_f32(pos, g.pairs, uvw, lmn, I, g.r_moon_km, vis_sel=sel)
    t1 = time.perf_counter()
    o.cpu_recon_f32(pos, g.pairs, uvw, lmn, V, g.r_moon_km, pix_sel=psel)
    t2 = time.perf_counter()
    terms = len(sel) * wl.n_pix + len(psel) * wl.n_vis
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(terms / (t2 - t0) / 1e9, 4), "unit": "Gterms/s", "cores": cores,
            "kind": "port", "cpu_model": cpu_model(),
            "omp": {"OMP_NUM_THREADS": cores, "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND")},
            "sample": f"every {stride}th visibility x all {wl.n_pix} pixels (vis-gen) + every "
                      f"{stride}th pixel x all {wl.n_vis} visibilities (recon): {terms:.3e} terms "
                      f"in {t2 - t0:.1f} s; FP32 OpenMP restatement of visscal/imagerecon"}


def mfma_legs(torch, ctx, vrs, W, wl, g, pos, uvw, lmn, args):
    """Multi-map contraction on the matrix cores: 64 maps on the same 10 MHz
    geometry (Nside 128), and the 64-map Nside-64 precision configuration.
    Terms are map x vis x pix per direction."""
    import oracle as o
    res = {}
    km = vrs.knobs(precision="mfma")
    for name, geom, nside in (("c2_geometry_64maps", g, wl.nside), ("c3", None, 64)):
        gg = geom
        wl3 = W.CONFIGS["c3"]
        if gg is None:
            gg = ctx.workload_geometry(wl3)
        n_pix = 12 * nside * nside
        I = W.make_sky(W.Workload("m", wl.n_sat, wl.n_t, wl.freq_hz, nside, 64, wl3.sky))
        It = torch.from_numpy(I).cuda()
        st = {}

        def step():
            gg.MIncline(export=False)
            V = gg.visscal(It, km)
            st["V"] = V
            st["F"] = gg.imagerecon(V, None, km)

        ms = timed_steps(torch, step, max(2, args.steps // 2), 1)
        kt = kernel_avg_ms(ctx, step, ["mfma_visgen", "mfma_recon"], 2)
        terms = 2 * 64 * gg.n_vis * n_pix
        entry = {"value": round(terms / (ms * 1e-3) / 1e9, 1), "unit": "Gterms/s (map x vis x pix)",
                 "ms_per_step": round(ms, 3), "kernels_ms": {k: round(v, 3) for k, v in kt.items()},
                 "n_maps": 64, "nside": nside}
        # MFMA roofline: 4 algorithmic flops per map-term (two real MACs), dense fp16 peak
        mf = max(kt.values())
        entry["roofline"] = {"bound": "mfma", "achieved": round(4 * 64 * gg.n_vis * n_pix / (mf * 1e-3) / 1e12, 2),
                             "peak": 2500.0, "unit": "TFLOP/s"}
        entry["roofline"]["frac"] = round(entry["roofline"]["achieved"] / 2500.0, 4)
        if not args.no_accuracy:
            V = st["V"].cpu().numpy()
            F = st["F"].cpu().numpy()
            p2, u2 = gg.MIncline()
            l2 = gg.lmn()
            sel = np.linspace(0, gg.n_vis - 1, 256).astype(np.int64)
            Vo = o.visgen(p2, gg.pairs, u2, l2, I, gg.r_moon_km, vis_sel=sel)
            Vc = V[:, sel, 0].astype(np.float64) + 1j * V[:, sel, 1]
            psel = np.linspace(0, n_pix - 1, 64).astype(np.int64)
            Vfull = V[..., 0].astype(np.float64) + 1j * V[..., 1]
            Fo = o.recon(p2, gg.pairs, u2, l2, Vfull, gg.r_moon_km, pix_sel=psel)
            ev = np.linalg.norm(Vc - Vo, axis=1) / np.linalg.norm(Vo, axis=1)
            ef = np.linalg.norm(F[:, psel] - Fo, axis=1) / np.linalg.norm(Fo, axis=1)
            entry["accuracy"] = {
                "V_rel_l2_max": float(ev.max()), "V_rel_l2_median": float(np.median(ev)),
                "F_rel_l2_max": float(ef.max()), "F_rel_l2_median": float(np.median(ef)),
                "sample": "all 64 maps; 256 visibilities x all pixels (V); 64 pixels x all "
                          "visibilities (F); per-map relative L2 against the FP64 oracle"}
            del Vfull, V
        res[name] = entry
        del It, st
        torch.cuda.empty_cache()
    return res


def batch_leg(torch, ctx, vrs, W, args):
    """Batched form: 32 channels log-spaced over 1-30 MHz, per-channel steps
    T_c ~ U{2000..8640} (seed 2), line-of-sight compaction, Nside 64,
    I_c = I0 (nu_c / 1 MHz)^alpha_c with alpha_c ~ N(-2.5, 0.1); one captured
    hipGraph per step, autotuned tiles.  Terms = sum_c kept_c x N_pix per direction."""
    freqs, n_t, alpha = W.c4_channels()
    nside = W.C4_NSIDE
    g = ctx.geometry(W.C4_N_SAT, int(n_t.max()), 1e7, nside)
    b = ctx.batch(g, freqs, n_t, compact=True, precision="mixed", autotune=True)
    I = W.c4_skies(freqs, alpha, nside)
    It = torch.from_numpy(I).cuda()
    k = vrs.knobs(precision="mixed")

    def step():
        b.run(It, k, want_V=False)

    ms = timed_steps(torch, step, max(2, args.steps // 2), 1)
    kt = kernel_avg_ms(ctx, step, ["batch"], 2)
    terms = 2 * int(b.kept.sum()) * 12 * nside * nside
    out = {"value": round(terms / (ms * 1e-3) / 1e9, 1), "unit": "Gterms/s (vis x pix)",
           "ms_per_step": round(ms, 3), "graph_ms": round(kt["batch"], 3),
           "channels": len(freqs), "kept_samples": int(b.kept.sum()),
           "samples_before_compaction": int(120 * n_t.sum()),
           "tiles": {"vis_tile": b.info["vis_tile"], "recon_seg": b.info["recon_seg"]},
           "precision": "VRS_MIXED_F64PHASE"}
    if not args.no_accuracy:
        V, F = b.run(I, k)
        out["accuracy"] = batch_accuracy(g, b, freqs, I, V, F)
    b.close()
    g.close()
    return out


def c5_leg(torch, ctx, vrs, W):
    """Differentiable form: 16 satellites, 120 x 4320 visibilities at 10 MHz,
    Nside 64 ground truth, V_obs = V(I_true) + complex Gaussian noise at
    per-visibility SNR 20 (seed 3); 100 projected-CG iterations of
    min sum |V(I) - V_obs|^2 over I >= 0, each one forward vis-gen, one shaded
    back-projection (the hand-written adjoint) and one forward of the search
    direction (FP64 phase).  Reported with T_occ = 240 and with T_occ = 0."""
    from viss_recon_shade_amd import lsq
    wl = W.CONFIGS["c5"]
    g = ctx.workload_geometry(wl)
    I = torch.from_numpy(W.make_sky(wl)[0]).cuda()
    out = {"workload": wl.note, "n_vis": wl.n_vis, "n_pix": wl.n_pix, "method": "projected CG"}
    for t_occ in (240.0, 0.0):
        k = vrs.knobs(precision="mixed", t_occ=t_occ)
        V_true = g.visscal(I[None], k)[0]
        V_obs = lsq.add_noise(V_true, 20.0, 3)
        floor = float(((V_true - V_obs).double() ** 2).sum())
        lsq.recover(g, V_obs, n_iter=2, method="cg", k=k)           # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Irec, hist = lsq.recover(g, V_obs, n_iter=100, method="cg", k=k)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        n_it = len(hist) - 1
        passes = 3 * n_it + 1          # forward, adjoint, forward of the direction; final forward
        terms = passes * wl.n_vis * wl.n_pix
        out[f"t_occ_{int(t_occ)}"] = {
            "iterations": n_it, "seconds": round(dt, 3), "ms_per_iteration": round(dt / n_it * 1e3, 3),
            "value": round(terms / dt / 1e9, 1), "unit": "Gterms/s (vis x pix, all passes)",
            "loss_initial_over_floor": hist[0] / floor, "loss_final_over_floor": hist[-1] / floor,
            "sky_rel_err": float(torch.linalg.norm(Irec - I) / torch.linalg.norm(I))}
    g.close()
    return out


def _radix_passes(uvw):
    """Digit passes the radix sort runs: 8-bit digits of the 63-bit cell keys that
    are not the same for every cell."""
    k = ((((uvw[:, 0].astype(np.int64) + (1 << 20)) << 42) | ((uvw[:, 1].astype(np.int64) + (1 << 20)) << 21)
          | (uvw[:, 2].astype(np.int64) + (1 << 20)))).astype(np.uint64)
    return int(sum(len(np.unique((k >> np.uint64(8 * d)) & np.uint64(255))) > 1 for d in range(8)))


def grid_leg(ctx, vrs, W, g_c2, wl, ref_scale):
    """uvw gridding (ceil, lambda = int(3e8/f)), sort, unique and counts of every
    ordered satellite pair: c2's 16 satellites x 8640 steps, and optionally the
    reference's own stage-1 scale (uvwPosition windows, OrbitRes 864000 x 140
    orbits -> 40.32 M samples per satellite, 8 satellites, 56 ordered pairs).
    Algorithmic bytes per key: 8 (key written) + 8 (AND/OR digit check) +
    24 per radix pass (histogram read, scatter read + write) + 8 (run count)."""
    import ctypes as C
    import math
    lib = ctx._lib
    out = {}

    def run(g, n_keys, label):
        ctx.timing(True)
        n_u = C.c_int64()
        for _ in range(2):
            ctx.check(lib.vrs_grid_uvw(ctx.handle, g.handle, 3, 0, None, None, C.byref(n_u), None, None))
            ms = ctx.last_ms("grid")
        ctx.timing(False)
        uu = np.empty((n_u.value, 3), np.float32)
        ctx.check(lib.vrs_grid_uvw(ctx.handle, g.handle, 3, n_u.value, uu.ctypes.data, None,
                                   C.byref(n_u), None, None))
        passes = _radix_passes(uu)
        alg = n_keys * (8 + 8 + 24 * passes + 8)
        out[label] = {"keys": n_keys, "unique_cells": int(n_u.value), "ms": round(ms, 3),
                      "gkeys_per_s": round(n_keys / (ms * 1e-3) / 1e9, 2), "radix_passes": passes,
                      "algorithmic_bytes": alg,
                      "hbm_frac": round(alg / (ms * 1e-3) / 8e12, 3)}

    run(g_c2, wl.n_sat * (wl.n_sat - 1) * wl.n_t, "c2_ordered_pairs")
    if ref_scale:
        OR, OC = 864000, 140
        idx = W.uvw_position_index(OR, OC)
        period = 2 * math.pi * math.sqrt(W.A_KM ** 3 / W.MU_MOON)
        g = ctx.geometry(8, len(idx), 1e7, 1, dt_s=period / OR, sample_index=idx)
        run(g, 8 * 7 * len(idx), "reference_stage1_scale")
        g.close()
    out["note"] = ("ms = keys + digit check + radix sort + unique count (HIP events, 'grid'); "
                   "hbm_frac = algorithmic bytes / ms / 8 TB/s")
    return out


def batch_accuracy(g, b, freqs, I, V, F, n_vis_sample=64, n_pix_sample=16,
                   image_channels=(0, 10, 21, 31)):
    """Per-channel errors of one batch run against the FP64 oracle: V on
    n_vis_sample kept samples x all pixels of every channel, F (kappa = 1/kept)
    on n_pix_sample pixels x all kept samples of `image_channels`."""
    import oracle as o
    pos, _ = g.MIncline()
    lmn = g.lmn()
    ev, ef = [], {}
    for c in range(len(freqs)):
        sl = slice(b.offsets[c], b.offsets[c + 1])
        bt = b.sample_bt[sl]
        real = bt >= 0
        bt = bt[real]
        Vc = V[sl][real][:, 0].astype(np.float64) + 1j * V[sl][real][:, 1]
        uvw = o.uvw(pos, g.pairs, 299792458.0 / freqs[c])
        pick = np.linspace(0, len(bt) - 1, n_vis_sample).astype(np.int64)
        Vo = o.visgen(pos, g.pairs, uvw, lmn, I[c], g.r_moon_km, vis_sel=bt[pick])[0]
        ev.append(float(np.linalg.norm(Vc[pick] - Vo) / np.linalg.norm(Vo)))
        if c in image_channels:
            Vfull = np.zeros(g.n_vis, np.complex128)
            Vfull[bt] = Vc
            psel = np.linspace(0, g.n_pix - 1, n_pix_sample).astype(np.int64)
            Fo = o.recon(pos, g.pairs, uvw, lmn, Vfull, g.r_moon_km, kappa=1.0 / b.kept[c],
                         pix_sel=psel)[0]
            ef[str(c)] = float(np.linalg.norm(F[c, psel] - Fo) / np.linalg.norm(Fo))
    return {"V_rel_l2_max": max(ev), "V_rel_l2_per_channel": [round(x, 9) for x in ev],
            "F_rel_l2": ef, "F_rel_l2_max": max(ef.values()),
            "sample": f"V: {n_vis_sample} kept samples x all pixels of every channel; F: "
                      f"{n_pix_sample} pixels x all kept samples of channels "
                      f"{list(image_channels)}, kappa = 1/kept; FP64 oracle"}


def _errs(x, ref):
    d = np.abs(np.asarray(x) - ref)
    scale = np.abs(ref).max()
    return {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(ref)),
            "max_abs_over_inf": float(d.max() / scale),
            "p50_abs_over_inf": float(np.quantile(d, 0.5) / scale),
            "p99_abs_over_inf": float(np.quantile(d, 0.99) / scale)}


def accuracy(g, wl, pos, uvw, lmn, I, V, F, n_vis_sample=4096, n_pix_sample=256, refs=None):
    """Errors against the FP64 oracle on deterministic subsamples: evenly spaced
    visibilities x all pixels (V), evenly spaced pixels x all visibilities (F).
    `refs` (a dict) receives the samples and the oracle's values."""
    import oracle as o
    sel = np.linspace(0, wl.n_vis - 1, n_vis_sample).astype(np.int64)
    Vo = o.visgen(pos, g.pairs, uvw, lmn, I, g.r_moon_km, vis_sel=sel)[0]
    Vc = V[0, sel, 0].astype(np.float64) + 1j * V[0, sel, 1]
    psel = np.linspace(0, wl.n_pix - 1, n_pix_sample).astype(np.int64)
    Vfull = V[0, :, 0].astype(np.float64) + 1j * V[0, :, 1]
    Fo = o.recon(pos, g.pairs, uvw, lmn, Vfull, g.r_moon_km, pix_sel=psel)[0]
    ev, ef = _errs(Vc, Vo), _errs(F[0, psel], Fo)
    if refs is not None:
        refs.update(vis_sel=sel, pix_sel=psel, Vo=Vo, Fo=Fo)
    return {"V_rel_l2": ev["rel_l2"], "F_rel_l2": ef["rel_l2"], "V": ev, "F": ef,
            "sample": f"{n_vis_sample} visibilities x all pixels; {n_pix_sample} pixels x all "
                      f"visibilities (F from the GPU's own V)"}


def mask_check(g, wl, pos, lmn, stride=32, block=65536):
    """Shade counts and XOR hashes of every stride-th visibility against the oracle,
    in blocks of visibilities with a progress line on stderr after each."""
    import oracle as o
    st = g.ShadeMatrix()
    sel = np.arange(0, wl.n_vis, stride, dtype=np.int64)
    n_cnt = n_hash = near = 0
    t0 = time.perf_counter()
    for b0 in range(0, len(sel), block):
        sb = sel[b0:b0 + block]
        cnt, h, nb = o.mask_stats(pos, g.pairs, lmn, g.r_moon_km, vis_sel=sb)
        n_cnt += int((st["count"][sb] != cnt).sum())
        n_hash += int((st["hash"][sb] != h).sum())
        near += int(nb)
        if len(sel) > block:
            print(f"mask check: {b0 + len(sb)} / {len(sel)} visibilities, {n_cnt} count and "
                  f"{n_hash} hash mismatches, {time.perf_counter() - t0:.0f} s",
                  file=sys.stderr, flush=True)
    return {"visibilities_compared": int(len(sel)),
            "pixel_pairs_compared": int(len(sel)) * wl.n_pix,
            "count_mismatches": n_cnt,
            "hash_mismatches": n_hash,
            "near_limb_pairs_sample": near,
            "near_limb_pairs_all": int(st["near_limb"]),
            "fp64_tests_all": int(st["fp64_tests"])}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this benchmark measures one GPU only; the work of several GPUs "
                                   "runs on one MI355X at full size (use --gpus 1)"}))
        sys.exit(2)
    # the CPU legs run the OpenMP restatements with threads bound close
    os.environ.setdefault("OMP_PROC_BIND", "close")
    import torch
    import viss_recon_shade_amd as vrs
    from viss_recon_shade_amd import workload as W

    wl = W.CONFIGS[args.workload]
    ctx = vrs.Context(0)
    g = ctx.workload_geometry(wl)
    pos, uvw = g.MIncline()
    lmn = g.lmn()
    I = W.make_sky(wl)
    It = torch.from_numpy(I).cuda()
    kp = vrs.knobs(precision="plain")
    km = vrs.knobs(precision="mixed")
    # recon of the one map on the matrix cores (per-satellite factors, the pairs of a
    # step as a Hermitian form): more accurate than FP32 and faster than the plain kernel
    kq = vrs.knobs(precision="mfma")
    state = {}

    def step_with(k, kr=None):
        def step():
            g.MIncline(export=False)
            V = g.visscal(It, k)
            state["V"] = V
            state["F"] = g.imagerecon(V, None, kr or k)
        return step

    terms = 2 * wl.terms_per_direction
    tile_plain = g.autotune(It, kp)
    tile_mixed = g.autotune(It, km)
    ms_plain = timed_steps(torch, step_with(kp, kq), args.steps, args.warmup)
    kt_plain = kernel_avg_ms(ctx, step_with(kp, kq), ["orbit", "visgen", "mfma_recon"],
                             max(2, args.steps // 2))
    kt_plain["recon"] = kt_plain.pop("mfma_recon")
    Vp, Fp = state["V"].cpu().numpy(), state["F"].cpu().numpy()
    # the same step with the plain FP32 recon kernel
    ms_allplain = timed_steps(torch, step_with(kp), max(2, args.steps // 2), 1)
    kt_allplain = kernel_avg_ms(ctx, step_with(kp), ["visgen", "recon"], 2)
    # the whole step on the matrix cores: one-map Gram vis-gen + quadratic-form recon
    ms_onemap = timed_steps(torch, step_with(kq), max(2, args.steps // 2), 1)
    kt_onemap = kernel_avg_ms(ctx, step_with(kq), ["mfma_visgen", "mfma_recon"], 2)
    Vq, Fq = state["V"].cpu().numpy(), state["F"].cpu().numpy()
    ms_mixed = timed_steps(torch, step_with(km), max(2, args.steps // 2), 1)
    kt_mixed = kernel_avg_ms(ctx, step_with(km), ["visgen", "recon"], 2)
    Vm, Fm = state["V"].cpu().numpy(), state["F"].cpu().numpy()
    # the same plain step with every term forming its own phase (no antipodal pairs)
    kd = vrs.knobs(precision="plain", antipodal=0)
    tile_direct = g.autotune(It, kd)
    ms_direct = timed_steps(torch, step_with(kd), max(2, args.steps // 2), 1)
    kt_direct = kernel_avg_ms(ctx, step_with(kd), ["visgen", "recon"], 2)

    value = terms / (ms_plain * 1e-3) / 1e9
    # the same step with host (numpy) buffers at the boundary: PCIe staging included
    def host_step():
        g.MIncline(export=False)
        Vh = g.visscal(I, kp)
        g.imagerecon(Vh, None, kq)
    host_step()
    t0 = time.perf_counter()
    for _ in range(2):
        host_step()
    ms_host = (time.perf_counter() - t0) / 2 * 1e3
    traffic = None
    issue = None
    dom = max(("visgen", "recon"), key=lambda n: kt_plain[n])
    dom_ms = kt_plain[dom]
    paired = wl.nside >= 16   # HEALPix skies from Nside 16 on take the antipodal layout
    kname = {"visgen": visgen_name(False, tile_plain[0], paired),
             "recon": "quad_recon_kernel"}[dom]
    tfile = os.path.join(ROOT, "profiles", "traffic_c2.json")
    if wl.name == "c2" and os.path.exists(tfile):
        kern = json.load(open(tfile))["kernels"]
        if kname in kern:
            traffic = kern[kname]["traffic_bytes"]
    mfile = os.path.join(ROOT, "profiles", "instmix_c2.json")
    if wl.name == "c2" and os.path.exists(mfile):
        mix = json.load(open(mfile)).get(kname)
        if mix and "SQ_ACTIVE_INST_VALU" in mix and "GRBM_GUI_ACTIVE" in mix:
            per_term = 64.0 / wl.terms_per_direction
            issue = {"valu_busy": round(mix["SQ_ACTIVE_INST_VALU"] / N_CU
                                        / (mix["GRBM_GUI_ACTIVE"] / N_XCD), 3),
                     "valu_insts_per_term": round(mix["SQ_INSTS_VALU"] * per_term, 2),
                     "trans_insts_per_term": round(mix["SQ_INSTS_VALU_TRANS_F32"] * per_term, 2),
                     "source": "profiles/instmix_c2.json (rocprofv3 SQ_ACTIVE_INST_VALU, "
                               "GRBM_GUI_ACTIVE, SQ_INSTS_VALU[_TRANS_F32])"}
    achieved = FLOPS_PER_TERM * wl.terms_per_direction / (dom_ms * 1e-3) / 1e12
    # measured VALU/transcendental ceiling: the bare term, registers only, full occupancy
    ceil = ctx.term_ceiling()
    ceiling = {"value": round(ceil, 1), "unit": "Gterms/s",
               "frac": {k: round(wl.terms_per_direction / (v * 1e-3) / 1e9 / ceil, 4)
                        for k, v in kt_plain.items() if k in ("visgen", "recon")},
               "note": "vrs_term_ceiling: 3 FMA + fract + v_cos + v_sin + 2 FMA per term, no "
                       "shading, no memory; recon's figure counts all (vis, pix) pairs; the "
                       "one-map matrix-core recon forms one phase per (satellite, step, pixel), "
                       "not per term, so it can exceed the term ceiling"}
    out = {
        "metric": "Gterms/s (vis x pix) vis-gen + recon",
        "value": round(value, 2),
        "unit": "Gterms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_plain, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{wl.name}: {wl.note}", "n_vis": wl.n_vis, "n_pix": wl.n_pix,
                   "n_maps": wl.n_maps, "freq_hz": wl.freq_hz,
                   "precision": {"visgen": "VRS_PLAIN_F32", "recon": "VRS_MFMA_SPLIT16 (one map)"},
                   "step": "orbit+uvw+caps, vis-gen, recon"},
        # compute roofline of a VALU kernel: a single-map DFT has no matrix-core work; each
        # term is 10 FP32 flops plus two quarter-rate transcendentals
        "roofline": {"bound": "valu", "engine": "VALU (FP32 FMA + v_sin/v_cos)",
                     "kernel": dom, "achieved": round(achieved, 2),
                     "peak": PEAK_FP32_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_VALU_TFLOPS, 4),
                     "peak_note": "MI355X FP32 VALU at one FMA per lane per clock (256 CUs x 64 "
                                  "lanes x 2 flops x 2.4 GHz); the published 157.3 TFLOP/s counts "
                                  "packed FP32 FMAs, which this build does not issue",
                     "frac_of_published_157": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic,
                     "traffic_source": "profiles/traffic_c2.json (rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                                       "copy-calibrated)" if traffic else None,
                     "algorithmic_bytes": algorithmic_bytes(wl, dom),
                     "flops_per_term": FLOPS_PER_TERM,
                     "trans_frac": round(2 * wl.terms_per_direction / (dom_ms * 1e-3) / TRANS_PEAK, 4),
                     "kernel_ms": round(dom_ms, 3), "kernel_name": kname, "issue": issue,
                     "valu_trans_ceiling": ceiling},
        "kernels_ms": {k: round(v, 3) for k, v in kt_plain.items()},
        "visgen_tile": {"plain": tile_plain[0], "mixed": tile_mixed[0]},
        "host_buffers": {"value": round(terms / (ms_host * 1e-3) / 1e9, 2), "ms_per_step": round(ms_host, 3),
                         "note": "same step with numpy I, V, F at the C-ABI (staged over PCIe)"},
        "plain_recon": {"value": round(terms / (ms_allplain * 1e-3) / 1e9, 2),
                        "ms_per_step": round(ms_allplain, 3),
                        "kernels_ms": {k: round(v, 3) for k, v in kt_allplain.items()},
                        "note": "the headline step with the plain FP32 recon kernel (every "
                                "(visibility, pixel) term forms its own phase, as the reference)"},
        "one_map_split16": {"value": round(terms / (ms_onemap * 1e-3) / 1e9, 2),
                            "ms_per_step": round(ms_onemap, 3),
                            "kernels_ms": {k: round(v, 3) for k, v in kt_onemap.items()},
                            "precision": "VRS_MFMA_SPLIT16",
                            "note": "one map on the matrix cores: vis-gen as Gram matrices of the "
                                    "per-satellite factors (gram_visgen_kernel), recon as a "
                                    "Hermitian form in them (quad_recon_kernel)"},
        "mixed_f64phase": {"value": round(terms / (ms_mixed * 1e-3) / 1e9, 2),
                           "ms_per_step": round(ms_mixed, 3),
                           "kernels_ms": {k: round(v, 3) for k, v in kt_mixed.items()}},
        "plain_direct": {"value": round(terms / (ms_direct * 1e-3) / 1e9, 2),
                         "ms_per_step": round(ms_direct, 3),
                         "kernels_ms": {k: round(v, 3) for k, v in kt_direct.items()},
                         "visgen_tile": tile_direct[0],
                         "note": "plain FP32 with antipodal=0: every (visibility, pixel) term "
                                 "forms its own phase and sin/cos (the headline pairs a pixel "
                                 "with its antipode, one sin/cos for both)"},
    }
    if not args.no_mfma:
        out["mfma_split16"] = mfma_legs(torch, ctx, vrs, W, wl, g, pos, uvw, lmn, args)
    if not args.no_batch:
        out["batch_c4"] = batch_leg(torch, ctx, vrs, W, args)
    if not args.no_c5:
        out["c5_lsq"] = c5_leg(torch, ctx, vrs, W)
    if not args.no_grid:
        out["gridding"] = grid_leg(ctx, vrs, W, g, wl, args.grid_ref)
    refs = {}
    if not args.no_accuracy:
        out["accuracy"] = {"plain": accuracy(g, wl, pos, uvw, lmn, I, Vp, Fp, refs=refs),
                           "mixed": accuracy(g, wl, pos, uvw, lmn, I, Vm, Fm),
                           "one_map_split16": accuracy(g, wl, pos, uvw, lmn, I, Vq, Fq),
                           "mask": mask_check(g, wl, pos, lmn, args.mask_stride)}
    if not args.no_cpu:
        Vc = Vp[0, :, 0].astype(np.float64) + 1j * Vp[0, :, 1]
        out["cpu_baseline"] = cpu_baseline(g, wl, pos, uvw, lmn, I[0], Vc, args.cpu_stride)
        if refs:
            out["cpu_baseline"]["accuracy"] = cpu_error(g, wl, pos, uvw, lmn, I[0], Vc, refs)
    print(json.dumps(out))


if __name__ == "__main__":
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    main()
