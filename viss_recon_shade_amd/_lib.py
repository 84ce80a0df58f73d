"""ctypes bindings of libvrs.so (include/vrs.h).

The library is built in-tree by viss_recon_shade_amd.build; loading fails
loudly when it is missing -- there is no fallback implementation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvrs.so")

VRS_OK = 0
STATUS_NAMES = {0: "VRS_OK", -1: "VRS_E_ARG", -2: "VRS_E_OOM", -3: "VRS_E_HIP", -4: "VRS_E_IO",
                -5: "VRS_E_NODEV", -6: "VRS_E_STATE"}

VRS_PLAIN_F32 = 0
VRS_MIXED_F64PHASE = 1
VRS_MFMA_SPLIT16 = 2
VRS_GRID_ORDERED = 1
VRS_GRID_LAMBDA_INT = 2
VRS_UVW_CEIL = 1
VRS_UVW_LAMBDA_INT = 2


class VrsError(RuntimeError):
    """A vrs_* call returned a non-zero vrs_status."""

    def __init__(self, status, message):
        self.status = status
        self.code = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.code}: {message}")


class OrbitDesc(C.Structure):
    _fields_ = [("n_sat", C.c_int32), ("n_t", C.c_int32), ("dt_s", C.c_double),
                ("a_km", C.c_double), ("incl_rad", C.c_double), ("mu_km3_s2", C.c_double),
                ("raan_rad", C.POINTER(C.c_double)), ("m0_rad", C.POINTER(C.c_double)),
                ("sample_index", C.POINTER(C.c_int64))]


class GeomDesc(C.Structure):
    _fields_ = [("orbit", OrbitDesc), ("freq_hz", C.c_double), ("r_moon_km", C.c_double),
                ("n_base", C.c_int32), ("pairs", C.POINTER(C.c_int32)), ("nside", C.c_int32),
                ("uvw_flags", C.c_int32)]


class UvwGeomDesc(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("uvw", C.POINTER(C.c_double)), ("nside", C.c_int32),
                ("n_pix", C.c_int64), ("lmn", C.POINTER(C.c_double))]


class GeomInfo(C.Structure):
    _fields_ = [("n_vis", C.c_int64), ("n_pix", C.c_int64), ("n_sat", C.c_int32),
                ("n_t", C.c_int32), ("n_base", C.c_int32), ("nside", C.c_int32),
                ("n_chunk", C.c_int32), ("chunk_pixels", C.c_int32), ("lambda_m", C.c_double)]


class Knobs(C.Structure):
    _fields_ = [("phase_scale", C.c_double), ("wfac", C.c_int32), ("t_occ", C.c_float),
                ("kappa", C.c_double), ("precision", C.c_int32), ("antipodal", C.c_int32),
                ("sat_factors", C.c_int32)]


class BatchDesc(C.Structure):
    _fields_ = [("n_ch", C.c_int32), ("freq_hz", C.POINTER(C.c_double)),
                ("n_t", C.POINTER(C.c_int32)), ("compact", C.c_int32), ("precision", C.c_int32),
                ("autotune", C.c_int32)]


class BatchInfo(C.Structure):
    _fields_ = [("n_vis_total", C.c_int64), ("n_vis_padded", C.c_int64), ("n_slots", C.c_int64),
                ("vis_tile", C.c_int32), ("recon_seg", C.c_int32), ("tune_visgen_ms", C.c_double),
                ("tune_recon_ms", C.c_double)]


class ShadeStats(C.Structure):
    _fields_ = [("shaded_pairs", C.c_int64), ("near_limb", C.c_int64), ("fp64_tests", C.c_int64)]


# Every function declared in include/vrs.h: name -> (restype, argtypes)
_vp = C.c_void_p
_ctx = C.c_void_p
SIGNATURES = {
    "vrs_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "vrs_destroy": (None, [_ctx]),
    "vrs_last_error": (C.c_char_p, [_ctx]),
    "vrs_set_stream": (C.c_int, [_ctx, _vp]),
    "vrs_synchronize": (C.c_int, [_ctx]),
    "vrs_device_count": (C.c_int, []),
    "vrs_version": (C.c_char_p, []),
    "vrs_knobs_default": (None, [C.POINTER(Knobs)]),
    "vrs_knobs_reference": (None, [C.POINTER(Knobs)]),
    "vrs_timing_enable": (C.c_int, [_ctx, C.c_int]),
    "vrs_timing_last_ms": (C.c_double, [_ctx, C.c_char_p]),
    "vrs_term_ceiling": (C.c_int, [_ctx, C.POINTER(C.c_double)]),
    "vrs_geom_create": (C.c_int, [_ctx, C.POINTER(GeomDesc), C.POINTER(C.c_void_p)]),
    "vrs_geom_create_uvw": (C.c_int, [_ctx, C.POINTER(UvwGeomDesc), C.POINTER(C.c_void_p)]),
    "vrs_geom_destroy": (None, [_vp]),
    "vrs_geom_info_get": (C.c_int, [_vp, C.POINTER(GeomInfo)]),
    "vrs_orbit_uvw": (C.c_int, [_ctx, _vp, _vp, _vp]),
    "vrs_geom_lmn": (C.c_int, [_ctx, _vp, _vp]),
    "vrs_shade": (C.c_int, [_ctx, _vp, _vp, _vp, _vp, C.POINTER(ShadeStats)]),
    "vrs_pix_hash": (C.c_uint32, [C.c_int64]),
    "vrs_visgen": (C.c_int, [_ctx, _vp, C.POINTER(Knobs), C.c_int32, _vp, _vp]),
    "vrs_recon": (C.c_int, [_ctx, _vp, C.POINTER(Knobs), C.c_int32, _vp, _vp, _vp]),
    "vrs_adjoint": (C.c_int, [_ctx, _vp, C.POINTER(Knobs), _vp, _vp]),
    "vrs_recon_noalloc": (C.c_int, [_ctx, _vp, C.POINTER(Knobs), C.c_int32, _vp, _vp, _vp]),
    "vrs_autotune_visgen": (C.c_int, [_ctx, _vp, C.POINTER(Knobs), C.c_int32, _vp,
                                      C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "vrs_batch_create": (C.c_int, [_ctx, _vp, C.POINTER(BatchDesc), C.POINTER(C.c_void_p)]),
    "vrs_batch_destroy": (None, [_vp]),
    "vrs_batch_info_get": (C.c_int, [_vp, C.POINTER(BatchInfo)]),
    "vrs_batch_layout": (C.c_int, [_vp, _vp, _vp, _vp]),
    "vrs_batch_run": (C.c_int, [_ctx, _vp, C.POINTER(Knobs), _vp, _vp, _vp]),
    "vrs_shade_box": (C.c_int, [_ctx, _vp, C.c_int64, C.c_int32, C.c_float, _vp]),
    "vrs_geom_fine_sky": (C.c_int, [_ctx, _vp, C.c_int64, _vp, _vp, _vp, C.c_int32]),
    "vrs_fine_index": (C.c_int, [_ctx, _vp, _vp, _vp, _vp]),
    "vrs_visgen_fine": (C.c_int, [_ctx, _vp, _vp, _vp]),
    "vrs_grid_uvw": (C.c_int, [_ctx, _vp, C.c_int32, C.c_int64, _vp, _vp, C.POINTER(C.c_int64),
                               _vp, _vp]),
    "vrs_write_uvw_txt": (C.c_int, [C.c_char_p, _vp, C.c_int64, _vp]),
    "vrs_uvw_map_weights": (C.c_int, [C.c_char_p, _vp, C.c_int64, _vp, _vp, _vp, _vp]),
    "vrs_write_uvw_map_txt": (C.c_int, [C.c_char_p, _vp, _vp, C.c_int64]),
    "vrs_read_uvw_txt": (C.c_int, [C.c_char_p, C.c_int32, _vp, _vp, C.c_int64,
                                   C.POINTER(C.c_int64)]),
    "vrs_write_shade_txt": (C.c_int, [C.c_char_p, _vp, C.c_int64]),
    "vrs_read_shade_txt": (C.c_int, [C.c_char_p, _vp, C.c_int64, C.POINTER(C.c_int64)]),
    "vrs_write_image_txt": (C.c_int, [C.c_char_p, _vp, C.c_int64]),
    "vrs_read_image_txt": (C.c_int, [C.c_char_p, _vp, C.c_int64, C.POINTER(C.c_int64)]),
    "vrs_write_ex_txt": (C.c_int, [C.c_char_p, _vp, C.c_int64, _vp, C.c_int64, C.c_int32]),
    "vrs_write_vis_bin": (C.c_int, [C.c_char_p, _vp, C.c_int64, C.c_int64]),
    "vrs_read_vis_bin": (C.c_int, [C.c_char_p, _vp, C.c_int64, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64)]),
    "vrs_rng_normals": (C.c_int, [C.c_uint64, C.c_int64, _vp]),
    "vrs_rng_uniforms": (C.c_int, [C.c_uint64, C.c_int64, _vp]),
    "vrs_sky_lognormal": (C.c_int, [C.c_uint64, C.c_double, C.c_int64, C.c_int32, C.c_double,
                                    C.c_double, _vp]),
    "vrs_orbit_offsets": (C.c_int, [C.c_int32, C.c_uint64, C.c_double, C.c_double, _vp, _vp]),
    "vrs_uvw_position_index": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _vp]),
    "vrs_healpix_lmn": (C.c_int, [C.c_int32, _vp]),
    "vrs_healpix_nest2ring": (C.c_int, [C.c_int32, _vp]),
    "vrs_healpix_antipode": (C.c_int, [C.c_int32, _vp]),
}

_lib = None


def load():
    """Load libvrs.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing; build it with "
                           "`python -m viss_recon_shade_amd.build` (no fallback exists)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status, ctx=None):
    if status != VRS_OK:
        msg = load().vrs_last_error(ctx)
        raise VrsError(status, msg.decode() if msg else "")
