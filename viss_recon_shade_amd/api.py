"""Python mirror of the hot-path interface over the C-ABI.

Names follow the reference's stages: MIncline (orbits and uvw,
MIncline_generate_uvw_float.cu:184), ShadeMatrix (the occultation mask,
shade10M.cu:73), visscal (vis-gen, vissGen10M_shade.cu:108) and imagerecon
(recon, vissGen10M_shade.cu:149), as methods of Geometry.  Arrays may be
numpy arrays (host) or torch CUDA tensors (device, used in place on torch's
current stream); outputs follow the input's kind.
"""
import ctypes as C
import weakref

import numpy as np

from . import _lib
from . import workload as W
from ._lib import Knobs, VrsError  # noqa: F401  (re-exported)

try:
    import torch
except ImportError:  # torch is optional for host-only use
    torch = None


def knobs(preset="canonical", **over):
    """Scalar knobs: 'canonical' (s = 1, n - 1, T_occ = 240, kappa = 1/N_vis)
    or 'reference' (s = 0.5, trailing exp(-2 pi i s w), T_occ = 240)."""
    lib = _lib.load()
    k = Knobs()
    (lib.vrs_knobs_reference if preset == "reference" else lib.vrs_knobs_default)(C.byref(k))
    for name, val in over.items():
        if name == "precision" and isinstance(val, str):
            val = {"plain": 0, "f32": 0, "mixed": 1, "f64phase": 1, "mfma": 2}[val]
        setattr(k, name, val)
    return k


def _is_torch(x):
    return torch is not None and isinstance(x, torch.Tensor)


def _ptr(x):
    if _is_torch(x):
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return C.c_void_p(x.data_ptr())
    if not x.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(x.ctypes.data)


def _f32_input(x, shape, name, device=None):
    """Validate an input array before its pointer crosses the C-ABI, which reads
    float32 of exactly `shape` (None entries are free).  numpy inputs are
    converted to C-contiguous float32 (a float64 array would otherwise be read
    as float32 bit patterns); torch tensors must already be contiguous float32,
    on the context's device when they live on a GPU."""
    if _is_torch(x):
        if x.dtype != torch.float32:
            raise TypeError(f"{name} must be float32, got {x.dtype}")
        if x.is_cuda and device is not None and x.device.index != device:
            raise ValueError(f"{name} is on {x.device}, the context on cuda:{device}")
        if not x.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    else:
        x = np.asarray(x)
        if x.dtype.kind not in "fiu":
            raise TypeError(f"{name} must be a real array, got {x.dtype}")
        x = np.ascontiguousarray(x, np.float32)
    got = tuple(x.shape)
    if len(got) != len(shape) or any(s is not None and s != g for s, g in zip(shape, got)):
        want = "[" + "][".join("*" if s is None else str(s) for s in shape) + "]"
        raise ValueError(f"{name} must have shape {want}, got {list(got)}")
    return x


def _empty_like_kind(ref, shape, dtype):
    if _is_torch(ref):
        tdt = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32,
               np.uint32: torch.int32}[dtype]
        return torch.empty(shape, dtype=tdt, device=ref.device)
    return np.empty(shape, dtype)


class Context:
    """One device + one stream (vrs_ctx)."""

    def __init__(self, device=0):
        self._lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self._lib.vrs_create(device, C.byref(h)))
        self.handle = h
        self.device = device
        self._stream = None
        self._geoms = weakref.WeakSet()

    def _sync_stream(self, *arrays):
        """Run on torch's current stream when any argument is a torch tensor."""
        if any(_is_torch(a) for a in arrays if a is not None):
            s = torch.cuda.current_stream(self.device).cuda_stream
            if s != self._stream:
                self.check(self._lib.vrs_set_stream(self.handle, C.c_void_p(s)))
                self._stream = s

    def check(self, status):
        _lib.check(status, self.handle)

    def synchronize(self):
        self.check(self._lib.vrs_synchronize(self.handle))

    def timing(self, on=True):
        self.check(self._lib.vrs_timing_enable(self.handle, 1 if on else 0))

    def last_ms(self, kernel):
        return self._lib.vrs_timing_last_ms(self.handle, kernel.encode())

    def term_ceiling(self):
        """Measured Gterms/s of the bare plain-FP32 term (registers only, no
        shading): the VALU/transcendental ceiling of the vis-gen and recon terms."""
        out = C.c_double(0.0)
        self.check(self._lib.vrs_term_ceiling(self.handle, C.byref(out)))
        return out.value

    def geometry(self, n_sat, n_t, freq_hz, nside, raan=None, m0=None, dt_s=W.DT_S,
                 a_km=W.A_KM, incl_rad=W.INCL_RAD, mu=W.MU_MOON, r_moon_km=W.R_MOON_KM,
                 pairs=None, sample_index=None, uvw_ceil=False, lambda_int=False):
        """Orbit geometry.  sample_index [n_t] (optional): step j at time
        sample_index[j] * dt_s, negative = zero-filled (see uvw_position_index);
        uvw_ceil / lambda_int: the reference's gridded uvw, ceil(dx * 1e3 / lambda)
        with lambda = int(3e8 / f) (MIncline_generate_uvw_float.cu:156-158, 190)."""
        flags = (_lib.VRS_UVW_CEIL if uvw_ceil else 0) | (_lib.VRS_UVW_LAMBDA_INT if lambda_int else 0)
        return Geometry(self, n_sat, n_t, freq_hz, nside, raan, m0, dt_s, a_km, incl_rad, mu,
                        r_moon_km, pairs, sample_index, flags)

    def uvw_geometry(self, uvw, nside=None, lmn=None):
        """Geometry from explicit uvw rows [n][3] (wavelengths) -- the reference's
        updated_uvw input of stages 3-4 (vissGen10M_shade.cu:351-373) -- over a
        HEALPix RING sky of `nside` or explicit unit directions lmn [n_pix][3].
        No satellites, hence no occultation (the reference's own masks are no-ops
        on these inputs, vissGen10M_shade.cu:130, 165)."""
        return Geometry.from_uvw(self, uvw, nside, lmn)

    def workload_geometry(self, wl):
        return self.geometry(wl.n_sat, wl.n_t, wl.freq_hz, wl.nside)

    def shade_box(self, uvw, res, r_moon_ref_km=1737.74):
        """The reference's ShadeMatrix index boxes (1-based [thetaS thetaE phiS phiE])
        for FP32 uvw [n][3] on a res x res grid (shade10M.cu:73-105)."""
        uvw = _f32_input(uvw, (None, 3), "uvw", self.device)
        n = uvw.shape[0]
        out = _empty_like_kind(uvw, (n, 4), np.int32)
        self._sync_stream(uvw)
        self.check(self._lib.vrs_shade_box(self.handle, _ptr(uvw), n, res, r_moon_ref_km,
                                           _ptr(out)))
        return out

    ShadeMatrix_box = shade_box

    def batch(self, geom, freqs_hz, n_t, compact=True, precision="mixed", autotune=True):
        return Batch(self, geom, freqs_hz, n_t, compact, precision, autotune)

    def close(self):
        if self.handle:
            for g in list(self._geoms):   # geometries first: they live on this context
                g.close()
            self._lib.vrs_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Geometry:
    """Orbits, uvw, shade caps and the chunked HEALPix sky on one device (vrs_geom)."""

    def __init__(self, ctx, n_sat, n_t, freq_hz, nside, raan, m0, dt_s, a_km, incl_rad, mu,
                 r_moon_km, pairs, sample_index=None, uvw_flags=0):
        self.ctx = ctx
        lib = ctx._lib
        if raan is None or m0 is None:
            r0, mm = W.orbit_params(n_sat, a_km=a_km)
            raan = r0 if raan is None else raan
            m0 = mm if m0 is None else m0
        self.raan = np.ascontiguousarray(raan, np.float64)
        self.m0 = np.ascontiguousarray(m0, np.float64)
        d = _lib.GeomDesc()
        d.orbit.n_sat = n_sat
        d.orbit.n_t = n_t
        d.orbit.dt_s = dt_s
        d.orbit.a_km = a_km
        d.orbit.incl_rad = incl_rad
        d.orbit.mu_km3_s2 = mu
        d.orbit.raan_rad = self.raan.ctypes.data_as(C.POINTER(C.c_double))
        d.orbit.m0_rad = self.m0.ctypes.data_as(C.POINTER(C.c_double))
        d.freq_hz = freq_hz
        d.r_moon_km = r_moon_km
        if pairs is not None:
            self._pairs = np.ascontiguousarray(pairs, np.int32)
            d.n_base = len(self._pairs)
            d.pairs = self._pairs.ctypes.data_as(C.POINTER(C.c_int32))
        else:
            self._pairs = W.all_pairs(n_sat)
            d.n_base = 0
        d.nside = nside
        d.uvw_flags = uvw_flags
        if sample_index is not None:
            self.sample_index = np.ascontiguousarray(sample_index, np.int64)
            if self.sample_index.shape != (n_t,):
                raise ValueError(f"sample_index must have {n_t} entries")
            d.orbit.sample_index = self.sample_index.ctypes.data_as(C.POINTER(C.c_int64))
        h = C.c_void_p()
        ctx.check(lib.vrs_geom_create(ctx.handle, C.byref(d), C.byref(h)))
        self._adopt(h, r_moon_km)
        self.pairs = self._pairs

    def _adopt(self, h, r_moon_km):
        ctx = self.ctx
        self.handle = h
        ctx._geoms.add(self)
        info = _lib.GeomInfo()
        ctx.check(ctx._lib.vrs_geom_info_get(h, C.byref(info)))
        self.n_vis, self.n_pix = info.n_vis, info.n_pix
        self.n_sat, self.n_t, self.n_base = info.n_sat, info.n_t, info.n_base
        self.nside, self.n_chunk, self.lambda_m = info.nside, info.n_chunk, info.lambda_m
        self.r_moon_km = r_moon_km

    @classmethod
    def from_uvw(cls, ctx, uvw, nside=None, lmn=None):
        self = cls.__new__(cls)
        self.ctx = ctx
        self.uvw_rows = np.ascontiguousarray(uvw, np.float64).reshape(-1, 3)
        d = _lib.UvwGeomDesc()
        d.n_rows = len(self.uvw_rows)
        d.uvw = self.uvw_rows.ctypes.data_as(C.POINTER(C.c_double))
        if lmn is not None:
            self._lmn = np.ascontiguousarray(lmn, np.float64).reshape(-1, 3)
            d.n_pix = len(self._lmn)
            d.lmn = self._lmn.ctypes.data_as(C.POINTER(C.c_double))
        elif nside is None:
            raise ValueError("give nside or lmn")
        else:
            d.nside = nside
        h = C.c_void_p()
        ctx.check(ctx._lib.vrs_geom_create_uvw(ctx.handle, C.byref(d), C.byref(h)))
        self._adopt(h, 0.0)
        self.pairs = None
        self.raan = self.m0 = None
        return self

    def uvw(self):
        """uvw of every visibility [n_vis][3] (FP64 wavelengths), as the kernels use them."""
        out = np.empty((self.n_vis, 3), np.float64)
        self.ctx.check(self.ctx._lib.vrs_orbit_uvw(self.ctx.handle, self.handle, None, _ptr(out)))
        return out

    # ------------------------------------------------------------- stage 1
    def MIncline(self, export=True):
        """Re-run orbit propagation + uvw on the device; returns FP64
        (positions [n_sat][n_t][3] km, uvw [n_base][n_t][3] wavelengths)."""
        lib = self.ctx._lib
        if not export:
            self.ctx.check(lib.vrs_orbit_uvw(self.ctx.handle, self.handle, None, None))
            return None
        pos = np.empty((self.n_sat, self.n_t, 3), np.float64)
        uvw = np.empty((self.n_base, self.n_t, 3), np.float64)
        self.ctx.check(lib.vrs_orbit_uvw(self.ctx.handle, self.handle, _ptr(pos), _ptr(uvw)))
        return pos, uvw

    orbit_uvw = MIncline

    def grid_uvw(self, ordered=True, lambda_int=True, weights=True):
        """Gridded (ceil), de-duplicated, lexicographically sorted uvw with cell
        counts -- the reference's uvw / updated_uvw content -- and per-visibility
        1/count weights and counts.  Returns a dict."""
        lib = self.ctx._lib
        flags = (_lib.VRS_GRID_ORDERED if ordered else 0) | (_lib.VRS_GRID_LAMBDA_INT if lambda_int else 0)
        n = C.c_int64()
        self.ctx.check(lib.vrs_grid_uvw(self.ctx.handle, self.handle, flags, 0, None, None,
                                        C.byref(n), None, None))
        uvw = np.empty((max(n.value, 1), 3), np.float32)
        cnt = np.empty(max(n.value, 1), np.int32)
        w = np.empty(self.n_vis, np.float32) if weights else None
        vc = np.empty(self.n_vis, np.int32) if weights else None
        self.ctx.check(lib.vrs_grid_uvw(self.ctx.handle, self.handle, flags, len(cnt), _ptr(uvw),
                                        _ptr(cnt), C.byref(n), _ptr(w) if weights else None,
                                        _ptr(vc) if weights else None))
        return {"uvw": uvw[:n.value], "counts": cnt[:n.value], "weight": w, "vis_count": vc}

    def lmn(self):
        out = np.empty((self.n_pix, 3), np.float64)
        self.ctx.check(self.ctx._lib.vrs_geom_lmn(self.ctx.handle, self.handle, _ptr(out)))
        return out

    # ------------------------------------------------------------- stage 2
    def ShadeMatrix(self, bits=False):
        """Per-visibility shaded-pixel counts and hashes (and optionally the bitmap)."""
        lib = self.ctx._lib
        cnt = np.empty(self.n_vis, np.int32)
        h = np.empty(self.n_vis, np.uint32)
        b = np.empty((self.n_vis, (self.n_pix + 31) // 32), np.uint32) if bits else None
        st = _lib.ShadeStats()
        self.ctx.check(lib.vrs_shade(self.ctx.handle, self.handle, _ptr(cnt), _ptr(h),
                                     _ptr(b) if b is not None else None, C.byref(st)))
        out = {"count": cnt, "hash": h, "shaded_pairs": st.shaded_pairs,
               "near_limb": st.near_limb, "fp64_tests": st.fp64_tests}
        if bits:
            out["bits"] = b
        return out

    shade = ShadeMatrix

    # ------------------------------------------------------------- stage 3
    def visscal(self, I, k=None):
        """V[map][v] (complex64 as [..., 2] float32) from I[map][pix] float32."""
        k = k or knobs()
        I = _f32_input(I, (None,) * max(I.ndim, 1), "I", self.ctx.device)
        if I.ndim not in (1, 2) or I.shape[-1] != self.n_pix:
            raise ValueError(f"I must be [{self.n_pix}] or [maps][{self.n_pix}], got {list(I.shape)}")
        n_maps = 1 if I.ndim == 1 else I.shape[0]
        V = _empty_like_kind(I, (n_maps, self.n_vis, 2), np.float32)
        self.ctx._sync_stream(I)
        self.ctx.check(self.ctx._lib.vrs_visgen(self.ctx.handle, self.handle, C.byref(k), n_maps,
                                                _ptr(I), _ptr(V)))
        return V

    visgen = visscal

    def autotune(self, I, k=None):
        """Time the vis-gen tile candidates on input I; returns (tile, ms)."""
        k = k or knobs()
        I = _f32_input(I, (None,) * max(I.ndim, 1), "I", self.ctx.device)
        if I.ndim not in (1, 2) or I.shape[-1] != self.n_pix:
            raise ValueError(f"I must be [{self.n_pix}] or [maps][{self.n_pix}], got {list(I.shape)}")
        n_maps = 1 if I.ndim == 1 else I.shape[0]
        t = C.c_int32()
        ms = C.c_double()
        self.ctx._sync_stream(I)
        self.ctx.check(self.ctx._lib.vrs_autotune_visgen(self.ctx.handle, self.handle, C.byref(k),
                                                         n_maps, _ptr(I), C.byref(t),
                                                         C.byref(ms)))
        return t.value, ms.value

    # ------------------------------------------------------ fine-cell sky
    def fine_sky(self, lmn, FF, NX, one_based=True):
        """Attach a fine-cell sky: cell directions lmn [n][3], intensities FF [n],
        parent RING pixel NX [n] (1-based as the reference's NX file by default).
        The inverse index (countLoc / NXq) is built on the GPU."""
        lmn = np.ascontiguousarray(lmn, np.float64).reshape(-1, 3) if not _is_torch(lmn) else lmn
        FF = np.ascontiguousarray(FF, np.float32).reshape(-1) if not _is_torch(FF) else FF
        NX = np.ascontiguousarray(NX, np.int32).reshape(-1) if not _is_torch(NX) else NX
        n = lmn.shape[0]
        if FF.shape[0] != n or NX.shape[0] != n:
            raise ValueError("lmn, FF and NX must have one row per cell")
        self.ctx._sync_stream(lmn, FF, NX)
        self.ctx.check(self.ctx._lib.vrs_geom_fine_sky(self.ctx.handle, self.handle, n, _ptr(lmn),
                                                       _ptr(FF), _ptr(NX), 1 if one_based else 0))
        self.n_cells = n

    def fine_index(self):
        """(countLoc [n_pix], NXq [n_cells], per-pixel mean FF [n_pix]) of the attached sky."""
        n_cells = getattr(self, "n_cells", 0)
        if not n_cells:
            raise RuntimeError("no fine-cell sky attached")
        cnt = np.empty(self.n_pix, np.int32)
        nxq = np.empty(n_cells, np.int32)
        mean = np.empty(self.n_pix, np.float32)
        self.ctx.check(self.ctx._lib.vrs_fine_index(self.ctx.handle, self.handle, _ptr(cnt),
                                                    _ptr(nxq), _ptr(mean)))
        return cnt, nxq, mean

    def visscal_fine(self, k=None, device=False):
        """V[v] (complex64 as [..., 2] float32) of the attached fine-cell sky with
        sub-pixel shading."""
        k = k or knobs()
        if device:
            import torch
            V = torch.empty((self.n_vis, 2), dtype=torch.float32, device=f"cuda:{self.ctx.device}")
        else:
            V = np.empty((self.n_vis, 2), np.float32)
        self.ctx._sync_stream(V)
        self.ctx.check(self.ctx._lib.vrs_visgen_fine(self.ctx.handle, self.handle, C.byref(k),
                                                     _ptr(V)))
        return V

    # ------------------------------------------------------------- stage 4
    def imagerecon(self, V, w=None, k=None):
        """F[map][pix] float32 from V[map][v] (complex64 as [..., 2] float32)."""
        k = k or knobs()
        if V.ndim not in (2, 3):
            raise ValueError(f"V must be [{self.n_vis}][2] or [maps][{self.n_vis}][2]")
        V = _f32_input(V, (None,) * (V.ndim - 2) + (self.n_vis, 2), "V", self.ctx.device)
        if w is not None:
            w = _f32_input(w, (self.n_vis,), "w", self.ctx.device)
        n_maps = 1 if V.ndim == 2 else V.shape[0]
        F = _empty_like_kind(V, (n_maps, self.n_pix), np.float32)
        self.ctx._sync_stream(V, w)
        self.ctx.check(self.ctx._lib.vrs_recon(self.ctx.handle, self.handle, C.byref(k), n_maps,
                                               _ptr(V), _ptr(w) if w is not None else None,
                                               _ptr(F)))
        return F

    recon = imagerecon

    def imagerecon_noalloc(self, V, w=None, k=None, out=None):
        """imagerecon through vrs_recon_noalloc: device tensors only, no device
        allocation inside the library (its partial images live in a fixed
        per-device scratch array).  `out` (float32 [maps][n_pix] on the device)
        receives F; without it torch allocates F."""
        k = k or knobs()
        if not _is_torch(V) or not V.is_cuda:
            raise ValueError("imagerecon_noalloc takes device tensors")
        if V.ndim not in (2, 3):
            raise ValueError(f"V must be [{self.n_vis}][2] or [maps][{self.n_vis}][2]")
        V = _f32_input(V, (None,) * (V.ndim - 2) + (self.n_vis, 2), "V", self.ctx.device)
        if w is not None:
            w = _f32_input(w, (self.n_vis,), "w", self.ctx.device)
        n_maps = 1 if V.ndim == 2 else V.shape[0]
        F = out if out is not None else _empty_like_kind(V, (n_maps, self.n_pix), np.float32)
        _f32_input(F, (n_maps, self.n_pix), "out", self.ctx.device)
        self.ctx._sync_stream(V, w, F)
        self.ctx.check(self.ctx._lib.vrs_recon_noalloc(
            self.ctx.handle, self.handle, C.byref(k), n_maps, _ptr(V),
            _ptr(w) if w is not None else None, _ptr(F)))
        return F

    def adjoint(self, resid, k=None):
        """Gradient of sum |V(I) - V_obs|^2 given resid = V(I) - V_obs ([n_vis][2])."""
        k = k or knobs()
        resid = _f32_input(resid, (self.n_vis, 2), "resid", self.ctx.device)
        g = _empty_like_kind(resid, (self.n_pix,), np.float32)
        self.ctx._sync_stream(resid)
        self.ctx.check(self.ctx._lib.vrs_adjoint(self.ctx.handle, self.handle, C.byref(k),
                                                 _ptr(resid), _ptr(g)))
        return g

    def close(self):
        if self.handle:
            if self.ctx.handle:
                self.ctx._lib.vrs_geom_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """All channels of one array in one captured hipGraph (vrs_batch): channel c
    uses the first n_t[c] steps of `geom`'s orbits at freqs_hz[c], compacted to
    the mutually visible (baseline, t) samples when `compact`."""

    def __init__(self, ctx, geom, freqs_hz, n_t, compact, precision, autotune):
        self.ctx, self.geom = ctx, geom
        lib = ctx._lib
        self.freqs = np.ascontiguousarray(freqs_hz, np.float64)
        self.n_t = np.ascontiguousarray(n_t, np.int32)
        self.n_ch = len(self.freqs)
        self.precision = knobs(precision=precision).precision
        d = _lib.BatchDesc(self.n_ch, self.freqs.ctypes.data_as(C.POINTER(C.c_double)),
                           self.n_t.ctypes.data_as(C.POINTER(C.c_int32)), 1 if compact else 0,
                           self.precision, 1 if autotune else 0)
        h = C.c_void_p()
        ctx.check(lib.vrs_batch_create(ctx.handle, geom.handle, C.byref(d), C.byref(h)))
        self.handle = h
        info = _lib.BatchInfo()
        ctx.check(lib.vrs_batch_info_get(h, C.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in _lib.BatchInfo._fields_}
        self.offsets = np.empty(self.n_ch + 1, np.int64)
        self.kept = np.empty(self.n_ch, np.int64)
        self.sample_bt = np.empty(self.info["n_vis_padded"], np.int32)
        ctx.check(lib.vrs_batch_layout(h, _ptr(self.offsets), _ptr(self.kept), _ptr(self.sample_bt)))

    def run(self, I, k=None, want_V=True):
        """I [n_ch][n_pix] -> (V [n_vis_padded][2] or None, F [n_ch][n_pix])."""
        k = k or knobs(precision=self.precision)
        I = _f32_input(I, (self.n_ch, self.geom.n_pix), "I", self.ctx.device)
        V = _empty_like_kind(I, (self.info["n_vis_padded"], 2), np.float32) if want_V else None
        F = _empty_like_kind(I, (self.n_ch, self.geom.n_pix), np.float32)
        self.ctx._sync_stream(I)
        self.ctx.check(self.ctx._lib.vrs_batch_run(self.ctx.handle, self.handle, C.byref(k),
                                                   _ptr(I), _ptr(V) if V is not None else None,
                                                   _ptr(F)))
        return V, F

    def close(self):
        if self.handle:
            self.ctx._lib.vrs_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
