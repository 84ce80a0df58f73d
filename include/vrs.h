/*
 * vrs.h -- C-ABI of the MI355X-native viss_recon_shade hot path.
 *
 * The reference (SSshuishui/viss_recon_shade) has no library interface: its
 * boundary is a set of CUDA kernels taking raw device pointers plus scalars,
 * chained through text files.  Each entry point below names the reference
 * code it replaces.  All types are C99; no C++ or torch types cross this line.
 *
 * Conventions
 *  - Every call returns vrs_status; nothing calls exit().  vrs_last_error()
 *    gives a message for the last failing call on that context (or on the
 *    calling thread when ctx is NULL).
 *  - Array arguments may be device pointers (hipMalloc / torch tensors) or
 *    plain host pointers; the kind is detected.  Device arrays are used in
 *    place and the call is stream-ordered on the context stream; host arrays
 *    are staged and the call returns after the results are copied back.
 *  - Complex arrays are complex64, interleaved (re, im) float pairs.
 *  - A context is bound to one device and one stream and is not thread-safe;
 *    use one context per thread.
 */
#ifndef VRS_H
#define VRS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum vrs_status {
    VRS_OK = 0,
    VRS_E_ARG = -1,   /* invalid argument (checked on the host before any GPU work) */
    VRS_E_OOM = -2,   /* device allocation failed */
    VRS_E_HIP = -3,   /* a HIP runtime call or kernel launch failed */
    VRS_E_IO = -4,    /* file could not be read or written */
    VRS_E_NODEV = -5, /* no usable GPU */
    VRS_E_STATE = -6  /* object used in the wrong state */
} vrs_status;

typedef enum vrs_precision {
    VRS_PLAIN_F32 = 0,      /* reference arithmetic: phase, sin/cos, sums in FP32 */
    VRS_MIXED_F64PHASE = 1, /* phase formed and range-reduced in FP64, FP32 sincos + sums */
    VRS_MFMA_SPLIT16 = 2    /* FP64 phase; phase factors split hi+lo FP16, contracted on MFMA */
} vrs_precision;

typedef struct vrs_ctx vrs_ctx;
typedef struct vrs_geom vrs_geom;

/* Circular inclined orbits (e = 0, argument of perigee 0).
 * Replaces the constants and sampling of MIncline_generate_uvw_float.cu:22-30,185-190. */
typedef struct vrs_orbit_desc {
    int32_t n_sat;
    int32_t n_t;           /* time steps t_j = j * dt_s, j = 0 .. n_t-1 */
    double dt_s;
    double a_km;           /* orbit radius (R_moon + altitude) */
    double incl_rad;
    double mu_km3_s2;      /* lunar GM; mean motion sqrt(mu / a^3) */
    const double *raan_rad; /* [n_sat], host memory */
    const double *m0_rad;   /* [n_sat], host memory: mean anomaly at t = 0 */
    /* [n_t] host memory or NULL (then step j is at j * dt_s): step j is taken at
     * time sample_index[j] * dt_s; a negative entry is a zero-filled sample
     * (position 0, shades nothing), as the reference writes for orbit windows
     * past the end of its propagated span.  vrs_uvw_position_index() gives the
     * reference's one-third-orbit windows. */
    const int64_t *sample_index;
} vrs_orbit_desc;

/* Everything that fixes the (visibility, pixel) term set. */
typedef struct vrs_geom_desc {
    vrs_orbit_desc orbit;
    double freq_hz;        /* lambda = c / freq, c = 299792458 m/s */
    double r_moon_km;      /* occulting sphere radius */
    int32_t n_base;        /* 0: all pairs a < b in lexicographic order */
    const int32_t *pairs;  /* [n_base][2] satellite indices (host), or NULL with n_base = 0 */
    int32_t nside;         /* HEALPix RING sky, N_pix = 12 nside^2 */
    int32_t uvw_flags;     /* VRS_UVW_CEIL | VRS_UVW_LAMBDA_INT, or 0 */
} vrs_geom_desc;

/* uvw_flags: the reference's gridded uvw (computeUVW,
 * MIncline_generate_uvw_float.cu:156-158, 190): VRS_UVW_CEIL rounds each
 * component up to an integer number of wavelengths, VRS_UVW_LAMBDA_INT uses
 * lambda = (int)(3e8 / f) metres instead of c / f. */
#define VRS_UVW_CEIL 1
#define VRS_UVW_LAMBDA_INT 2

/* A geometry from an explicit list of uvw rows instead of orbits: the
 * reference's stage-3/4 input (the u v w [count] rows of its updated_uvw files,
 * read by vissGen10M_shade.cu:351-373), one visibility per row (v = row).
 * There are no satellites, so no occultation: the reference's own masking in
 * visscal and imagerecon never shades a pixel on these inputs (its fine-grid
 * shade loop runs only when M3 <= 0, vissGen10M_shade.cu:130, and its recon
 * skip only at grid pixel (1,1), :165).  Sky: HEALPix RING of `nside` when lmn
 * is NULL, else n_pix explicit unit directions (pixel p = row p of lmn). */
typedef struct vrs_uvw_geom_desc {
    int64_t n_rows;
    const double *uvw;     /* [n_rows][3] wavelengths, host memory */
    int32_t nside;         /* used when lmn is NULL */
    int64_t n_pix;         /* used when lmn is non-NULL */
    const double *lmn;     /* [n_pix][3] host memory, or NULL */
} vrs_uvw_geom_desc;

typedef struct vrs_geom_info {
    int64_t n_vis;         /* n_base * n_t; visibility index v = b * n_t + t */
    int64_t n_pix;
    int32_t n_sat, n_t, n_base, nside;
    int32_t n_chunk;       /* pixel chunks the kernels classify against the shade caps */
    int32_t chunk_pixels;
    double lambda_m;
} vrs_geom_info;

/* Scalar knobs.  The reference's conventions are knob values, not code paths:
 * reference preset = { 0.5, 1, 240, 0 (=1/N_vis), 1/count weights passed to recon };
 * canonical preset = { 1, 0, 240, 0 }.                                         */
typedef struct vrs_knobs {
    double phase_scale;    /* s: 1/dl of vissGen10M_shade.cu:183-185 (reference 0.5) */
    int32_t wfac;          /* 1: trailing exp(-2 pi i s w) on V (vissGen10M_shade.cu:143) and n in recon */
    float t_occ;           /* intensity of shaded pixels in vis-gen (vissGen10M_shade.cu:132: 240) */
    double kappa;          /* recon normalisation; 0 selects 1/N_vis (vissGen10M_shade.cu:156,173) */
    int32_t precision;     /* vrs_precision */
    int32_t antipodal;     /* 1 (default): on a HEALPix sky of Nside >= 16 the vis-gen
                            * kernels let one phase and one sin/cos serve a pixel and its
                            * antipode, whose phase is 2 wo - phi (E(-p) = conj(E(p))
                            * exp(-4 pi i wo)); 0: every (visibility, pixel) term forms its
                            * own phase.  Shading is decided per pixel either way; recon
                            * (most pixel pairs have one pixel wholly shaded) always runs
                            * per pixel. */
    int32_t sat_factors;   /* VRS_MFMA_SPLIT16 only.  1 (default): on a geometry built from
                            * orbits whose u v w are satellite position differences (no
                            * VRS_UVW_CEIL, at most 16 satellites) the matrix-core kernels form
                            * each pair's phase factor as E_b conj(E_a) from per-satellite
                            * factors (one FP64 phase and sin/cos per satellite and step,
                            * each satellite's own occultation zeroing its factor; with
                            * t_occ != 0 the shaded term is t_occ times each pair's full-sky
                            * sum, the maps entering as I - t_occ); 0: every (visibility,
                            * pixel) entry forms its own FP64 phase and sin/cos.  With one
                            * map, vis-gen on such a geometry forms each step's visibilities
                            * as Gram matrices of the satellite factors and recon treats
                            * them as a Hermitian form in the factors. */
} vrs_knobs;

typedef struct vrs_shade_stats {
    int64_t shaded_pairs;  /* total (vis, pixel) pairs shaded */
    int64_t near_limb;     /* pairs within 1e-6 R_moon of either satellite's limb */
    int64_t fp64_tests;    /* pixel tests that the FP32 guard band sent to FP64 */
} vrs_shade_stats;

/* ---------------------------------------------------------------- context */
/* Replaces the per-thread cudaSetDevice / CHECK pattern (vissGen10M_shade.cu:291-298, error.cuh:4-18). */
vrs_status vrs_create(int device, vrs_ctx **out);
void vrs_destroy(vrs_ctx *ctx);
const char *vrs_last_error(const vrs_ctx *ctx);
/* Run on `hip_stream` from now on; NULL selects the null (legacy default) stream. */
vrs_status vrs_set_stream(vrs_ctx *ctx, void *hip_stream);
vrs_status vrs_synchronize(vrs_ctx *ctx);
int vrs_device_count(void);
const char *vrs_version(void);
void vrs_knobs_default(vrs_knobs *k);
void vrs_knobs_reference(vrs_knobs *k);
/* Per-kernel timing with HIP events on the context stream (off by default). */
vrs_status vrs_timing_enable(vrs_ctx *ctx, int on);
/* Milliseconds of the last launch of `kernel` ("orbit", "shade", "visgen", "recon", "mfma_visgen",
 * "mfma_recon", "batch"); -1 when unknown.  Synchronizes the stream. */
double vrs_timing_last_ms(vrs_ctx *ctx, const char *kernel);
/* Measured VALU/transcendental ceiling of the plain FP32 term on this device:
 * a register-only kernel runs the bare term (3 FMA phase, fract, v_cos, v_sin,
 * 2 FMA accumulate; no shading, no memory) at full occupancy.  *gterms_per_s =
 * terms per second / 1e9, best of two timed launches (~2 ms each).  Measurement
 * only: it has no counterpart in the reference. */
vrs_status vrs_term_ceiling(vrs_ctx *ctx, double *gterms_per_s);

/* --------------------------------------------------------------- geometry */
/* Stage 1 and the static sky: propagates the orbits on the GPU, forms uvw and the
 * per-(satellite, t) shade caps, lays the HEALPix pixels out in spatially compact
 * chunks.  Replaces MIncline() (MIncline_generate_uvw_float.cu:184-440) up to the
 * uvw it writes, and the l,m,n inputs of vissGen (vissGen10M_shade.cu:199-248). */
vrs_status vrs_geom_create(vrs_ctx *ctx, const vrs_geom_desc *desc, vrs_geom **out);
/* The uvw-list geometry (vrs_uvw_geom_desc).  vrs_visgen, vrs_recon (with the
 * reference's 1/count weights and kappa = 1/n_rows), vrs_adjoint,
 * vrs_recon_noalloc, the fine-cell sky and vrs_shade (all counts 0) accept it;
 * the orbit-based calls (satellite positions, vrs_grid_uvw, batches) return
 * VRS_E_STATE. */
vrs_status vrs_geom_create_uvw(vrs_ctx *ctx, const vrs_uvw_geom_desc *desc, vrs_geom **out);
/* Destroy a geometry; call before destroying the context that created it. */
void vrs_geom_destroy(vrs_geom *g);
vrs_status vrs_geom_info_get(const vrs_geom *g, vrs_geom_info *out);
/* Re-runs stage 1 on the device (calculateOrbit + computeUVW,
 * MIncline_generate_uvw_float.cu:83-110,153-160) and optionally exports
 * positions [n_sat][n_t][3] km and uvw [n_base][n_t][3] wavelengths (FP64). */
vrs_status vrs_orbit_uvw(vrs_ctx *ctx, vrs_geom *g, double *sat_pos, double *uvw);
/* Pixel direction cosines in RING order [n_pix][3] (FP64), as the kernels use them. */
vrs_status vrs_geom_lmn(vrs_ctx *ctx, const vrs_geom *g, double *lmn);

/* ------------------------------------------------------------------ shade */
/* Stage 2: the occultation mask M[v][p] = occ(sat_a(t), p) | occ(sat_b(t), p),
 * never materialised in full.  Replaces ShadeMatrix + save_Shade (shade10M.cu:73-127).
 * count/hash: [n_vis] per-visibility shaded-pixel count and XOR of pix_hash(p)
 * over shaded RING pixels p (either may be NULL).  bits: optional full bitmap
 * [n_vis][ceil(n_pix/32)] (bit p%32 of word p/32), only when it fits max_bits_bytes. */
vrs_status vrs_shade(vrs_ctx *ctx, const vrs_geom *g, int32_t *count, uint32_t *hash,
                     uint32_t *bits, vrs_shade_stats *stats);
uint32_t vrs_pix_hash(int64_t p);

/* ----------------------------------------------------------- vis-gen / recon */
/* Stage 3, visscal (vissGen10M_shade_optim1.cu:130-206, vissGen10M_shade.cu:108-145):
 * V[map][v] = sum_p I'[map][p] exp(-2 pi i s (u l + v m + w (n-1)))  [* exp(-2 pi i s w) if wfac]
 * with I' = t_occ where M[v][p], else I.  I: [n_maps][n_pix] float (RING order),
 * V: [n_maps][n_vis] complex64. */
vrs_status vrs_visgen(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, int32_t n_maps,
                      const float *I, float *V);
/* Stage 4, imagerecon (vissGen10M_shade_optim1.cu:210-272, vissGen10M_shade.cu:149-175):
 * F[map][p] = kappa sum_v w[v] (1 - M[v][p]) Re(V[map][v] exp(+2 pi i s psi)),
 * psi = u l + v m + w (n - 1 + wfac).  w: [n_vis] float or NULL (all 1; the
 * reference's 1/count weights, vissGen10M_shade.cu:364). F: [n_maps][n_pix] float. */
vrs_status vrs_recon(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, int32_t n_maps,
                     const float *V, const float *w, float *F);
/* Times the vis-gen tile candidates (8, 16, 32 visibilities per block) on this
 * geometry with input I and caches the fastest for (n_vis, n_pix, n_maps,
 * precision) in the context; later vrs_visgen calls of that shape use it.
 * tile_out / ms_out (nullable) report the choice and its time.  Results do not
 * depend on the tile (each visibility's sums run in the same order). */
vrs_status vrs_autotune_visgen(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k,
                               int32_t n_maps, const float *I, int32_t *tile_out, double *ms_out);
/* Gradient of L(I) = sum_v |V(I)[v] - V_obs[v]|^2 given resid = V(I) - V_obs
 * (complex64 [n_vis]): grad[p] = 2 sum_v (1 - M) Re(resid[v] exp(+2 pi i s psi)).
 * The shaded back-projection; no reference counterpart (the reference has no
 * gradient), built on the imagerecon kernel with kappa = 2 and unit weights. */
vrs_status vrs_adjoint(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, const float *resid,
                       float *grad);
/* vrs_recon without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results (identical bits) as vrs_recon, but V, F and w (when
 * non-NULL) must be device memory and the precision VRS_PLAIN_F32 or
 * VRS_MIXED_F64PHASE.  Its partial images and its kernel's work counter live
 * in one fixed scratch array per device of VRS_NOALLOC_SCRATCH_BYTES, held
 * until the process ends; partial images need segments x n_maps x padded
 * pixels x 4 bytes (~6.3 MB per map at Nside 128) and may use all but the last
 * 256 bytes, a call needing more fails with VRS_E_ARG.  Every call on a device shares that array,
 * so calls that overlap in time on one device (other streams or contexts) race;
 * each call writes all of the scratch it reads, nothing carries over. */
#define VRS_NOALLOC_SCRATCH_BYTES (64u << 20)
vrs_status vrs_recon_noalloc(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, int32_t n_maps,
                             const float *V, const float *w, float *F);

/* ----------------------------------------------------------- ragged batches */
/* All frequency channels of one array in one go: channel c uses the first
 * n_t[c] steps of the geometry's orbits at wavelength c / freq_hz[c] and keeps
 * the (baseline, t) samples whose two satellites see each other (the segment
 * between them misses the Moon) when `compact` is set, so per-channel sizes
 * are uneven.  vrs_batch_run() replays one captured hipGraph: a chunk-order
 * gather of the per-channel skies, ONE persistent vis-gen launch over the
 * concatenated sample lists of all channels and ONE persistent recon launch
 * over (channel, visibility range, pixel chunk) work items.  Tile shapes are
 * chosen by timing candidates on the batch itself (autotune) and cached per
 * context.  No reference counterpart: the reference runs one period and one
 * frequency per program (vissGen10M_shade.cu:291-298, shade10M.cu:152). */
typedef struct vrs_batch vrs_batch;
typedef struct vrs_batch_desc {
    int32_t n_ch;
    const double *freq_hz;   /* [n_ch] */
    const int32_t *n_t;      /* [n_ch] steps used per channel, each <= the geometry's n_t */
    int32_t compact;         /* 1: keep only mutually visible (baseline, t) samples */
    int32_t precision;       /* VRS_PLAIN_F32 or VRS_MIXED_F64PHASE */
    int32_t autotune;        /* 1: time candidate tiles and keep the fastest */
} vrs_batch_desc;
typedef struct vrs_batch_info {
    int64_t n_vis_total;     /* kept samples over all channels */
    int64_t n_vis_padded;    /* length of the concatenated V array (channels padded to 32) */
    int64_t n_slots;         /* recon visibility ranges */
    int32_t vis_tile;        /* chosen vis-gen tile (visibilities per block) */
    int32_t recon_seg;       /* chosen recon range length */
    double tune_visgen_ms;   /* time of the chosen vis-gen launch during tuning (0: not tuned) */
    double tune_recon_ms;
} vrs_batch_info;
vrs_status vrs_batch_create(vrs_ctx *ctx, const vrs_geom *g, const vrs_batch_desc *d,
                            vrs_batch **out);
void vrs_batch_destroy(vrs_batch *b);
vrs_status vrs_batch_info_get(const vrs_batch *b, vrs_batch_info *out);
/* offsets: [n_ch + 1] start of each channel's samples in V; kept: [n_ch] real samples;
 * sample_vt: [n_vis_padded] (baseline * n_t + t) of each entry, -1 for padding (nullable). */
vrs_status vrs_batch_layout(const vrs_batch *b, int64_t *offsets, int64_t *kept, int32_t *sample_bt);
/* I: [n_ch][n_pix] float (RING); V: [n_vis_padded] complex64 (padding entries are zero);
 * F: [n_ch][n_pix] float.  kappa = 0 normalises each channel by 1 / kept[c]. */
vrs_status vrs_batch_run(vrs_ctx *ctx, vrs_batch *b, const vrs_knobs *k, const float *I, float *V,
                         float *F);

/* ------------------------------------------------- reference compatibility */
/* The reference's box predicate (ShadeMatrix, shade10M.cu:73-105, with
 * Cart2SphKernel :62-70 and the mean-|uvw| set-up :218-224) in FP32 on the
 * GPU: per uvw the 1-based index box [thetaS thetaE phiS phiE] on a res x res
 * (theta, phi) grid around the antipode of the baseline direction.  uvw: [n][3]
 * float, box4: [n][4] int32 (host or device).  r_moon_ref_km: 1737.74 in the
 * reference.  Kept for shadeM file interchange; the physical mask is vrs_shade. */
vrs_status vrs_shade_box(vrs_ctx *ctx, const float *uvw, int64_t n, int32_t res,
                         float r_moon_ref_km, int32_t *box4);

/* uvw gridding, de-duplication and cell counts: uvw = ceil(dx * 1e3 / lambda)
 * (computeUVW, MIncline_generate_uvw_float.cu:153-160), all samples sorted
 * lexicographically by (u, v, w) and made unique (:374-397); counts[i] = number
 * of gridded samples in unique cell i (the 4th column of the updated_uvw files,
 * whose 1/count is the recon weight, vissGen10M_shade.cu:357-366).
 * flags: VRS_GRID_ORDERED = grid every ordered satellite pair m != n as the
 * reference does (else the geometry's baselines); VRS_GRID_LAMBDA_INT =
 * lambda = (int)(3e8 / f) as the reference (MIncline_generate_uvw_float.cu:190)
 * (else c / f).  Positions are the geometry's FP64 ones.
 * uvw_unique [cap][3] and counts [cap] may be NULL (then only *n_unique is
 * returned); weight / vis_count [n_vis] (nullable) give each visibility
 * (baseline a < b, t) 1/count and count of its own cell.  Host or device pointers. */
#define VRS_GRID_ORDERED 1
#define VRS_GRID_LAMBDA_INT 2
vrs_status vrs_grid_uvw(vrs_ctx *ctx, const vrs_geom *g, int32_t flags, int64_t cap,
                        float *uvw_unique, int32_t *counts, int64_t *n_unique, float *weight,
                        int32_t *vis_count);

/* Fine-cell sky with sub-pixel shading (the reference's visscal over FF / NX /
 * countLoc / NXq, vissGen10M_shade.cu:121-138; index built on the GPU where the
 * reference scans all cells per pixel, vissGen1M_shade_optim2.cu:59-119).
 * Attaches n_cells cells to the geometry: unit direction lmn [n][3] (FP64),
 * intensity ff [n], parent RING pixel nx [n] (0- or 1-based per nx_base; the
 * reference's NX is 1-based).  Every sky pixel needs at least one cell (the
 * reference divides by its cell count).  Replaces an attached sky; freed with
 * the geometry.  Host or device pointers. */
vrs_status vrs_geom_fine_sky(vrs_ctx *ctx, vrs_geom *g, int64_t n_cells, const double *lmn,
                             const float *ff, const int32_t *nx, int32_t nx_base);
/* The inverse index: count_loc [n_pix] cells per pixel, nxq [n_cells] cell
 * indices grouped by pixel (ascending pixel, ascending cell), pixel_mean
 * [n_pix] = mean FF of each pixel (all nullable). */
vrs_status vrs_fine_index(vrs_ctx *ctx, const vrs_geom *g, int32_t *count_loc, int32_t *nxq,
                          float *pixel_mean);
/* V [n_vis] complex: pixel q contributes (1/count_q) sum over its cells c of
 * (c shaded for this visibility ? t_occ : FF_c) at the pixel's direction;
 * shading is the exact geometric test of each cell's own direction.
 * Precision VRS_PLAIN_F32 or VRS_MIXED_F64PHASE. */
vrs_status vrs_visgen_fine(vrs_ctx *ctx, const vrs_geom *g, const vrs_knobs *k, float *V);

/* Text files of the reference pipeline, same layout and default ostream
 * formatting (host memory only; no GPU needed):
 *   uvw[/updated_uvw]  "u v w[ count]"  MIncline_generate_uvw_float.cu:420-427, read at
 *                                       vissGen10M_shade.cu:351-366 (col4 nullable)
 *   shadeM             "tS tE pS pE"    shade10M.cu:108-127; the reader orders each pair
 *                                       and makes it 0-based (vissGen10M_shade.cu:385-399)
 *   F image            one value/line   vissGen10M_shade.cu:563-583
 * Readers stop with VRS_E_ARG instead of writing past `cap` rows. */
vrs_status vrs_write_uvw_txt(const char *path, const double *uvw, int64_t n, const float *col4);
vrs_status vrs_read_uvw_txt(const char *path, int32_t n_cols, double *uvw, float *col4,
                            int64_t cap, int64_t *n_read);
vrs_status vrs_write_shade_txt(const char *path, const int32_t *box4, int64_t n);
vrs_status vrs_read_shade_txt(const char *path, int32_t *box4, int64_t cap, int64_t *n_read);
vrs_status vrs_write_image_txt(const char *path, const float *F, int64_t n);
vrs_status vrs_read_image_txt(const char *path, float *F, int64_t cap, int64_t *n_read);
/* The 1 MHz pipeline's weights (vissGen1M_shade_accurate.cu:248-289, 352-368):
 * map_path is a uvwMap file (first line: the number of cells; then one
 * "u v w count" line of integers per cell; lines not starting with four
 * integers are skipped, a later line for a cell replaces an earlier one).
 * w[i] = 1 / count of row i's cell, 1 where the table has no such cell; the
 * cell of a row is its u v w rounded to FP32 and truncated toward zero.
 * n_cells, n_skipped, n_found (nullable): cells in the table, lines skipped,
 * rows whose cell was found.  Unlike the reference, a file that does not open
 * (VRS_E_IO; the reference silently weights every row 1), a first line
 * without a count (VRS_E_IO), a count <= 0 (VRS_E_ARG; the reference divides
 * by it) and a row outside the int32 range (VRS_E_ARG) are errors.
 * vrs_write_uvw_map_txt writes such a table from integer-valued rows and
 * their counts (e.g. vrs_grid_uvw's). */
vrs_status vrs_uvw_map_weights(const char *map_path, const double *uvw, int64_t n, float *w,
                               int64_t *n_cells, int64_t *n_skipped, int64_t *n_found);
vrs_status vrs_write_uvw_map_txt(const char *path, const double *uvw, const int32_t *counts,
                                 int64_t n);
/* Image expanded to fine cells, one line per cell c: F[nx[c] - nx_base]
 * (the Ex files, vissGen1M_shade_accurate.cu:567-588). */
vrs_status vrs_write_ex_txt(const char *path, const float *F, int64_t n_pix, const int32_t *nx,
                            int64_t n_cells, int32_t nx_base);
/* Binary visibilities (the reference never writes V): "VRSV", uint32 version 1,
 * int64 n_maps, int64 n_vis, then [n_maps][n_vis] complex64 little-endian.
 * The reader with V == NULL returns the header only; cap counts complex values. */
vrs_status vrs_write_vis_bin(const char *path, const float *V, int64_t n_maps, int64_t n_vis);
vrs_status vrs_read_vis_bin(const char *path, float *V, int64_t cap, int64_t *n_maps,
                            int64_t *n_vis);

/* ------------------------------------------------------- host-side helpers */
/* Synthetic inputs (no reference counterpart; the reference reads MATLAB files).
 * These run on the host and need no GPU. */
/* SplitMix64 stream; normals by Box-Muller (cos branch first). */
vrs_status vrs_rng_normals(uint64_t seed, int64_t n, double *out);
vrs_status vrs_rng_uniforms(uint64_t seed, int64_t n, double *out);
/* I_p = exp(sigma z_p), z ~ N(0,1) from `seed`; then n_src point sources at
 * distinct pixels drawn from the same stream, flux from p(S) ~ S^-2 on
 * [s_min, s_max] by inverse CDF, added to I. */
vrs_status vrs_sky_lognormal(uint64_t seed, double sigma, int64_t n_pix, int32_t n_src,
                             double s_min, double s_max, float *out);
/* RAAN spread uniformly (2 pi k / n_sat); along-track offsets d_k ~ U(0, max_offset_km)
 * from `seed`, as mean-anomaly offsets 2 asin(d_k / (2 a)) (MIncline_generate_uvw_float.cu:51-58). */
vrs_status vrs_orbit_offsets(int32_t n_sat, uint64_t seed, double a_km, double max_offset_km,
                             double *raan_rad, double *m0_rad);
/* The reference's orbit windows (uvwPosition, MIncline_generate_uvw_float.cu:113-151,
 * constants :185-190: span = orbit_res / 3 samples of each orbit, starting
 * interval = orbit_res / 36 samples later on every successive orbit, wrapped
 * back by whole orbits once past the propagated span, zero-filled where a window
 * runs off its end).  Writes orbit_counts * (orbit_res / 3) entries: the source
 * sample index (orbit k, step po -> k * orbit_res + po) of each kept step, or -1
 * for a zero-filled one; feed it to vrs_orbit_desc.sample_index.  The reference
 * tests the end with `>` (:134), reading one sample past the span; here that
 * sample is zero-filled too. */
vrs_status vrs_uvw_position_index(int32_t orbit_res, int32_t orbit_counts, int32_t interval,
                                  int32_t span, int64_t *out);
/* HEALPix RING pix2vec [12 nside^2][3] (Gorski et al. 2005). */
vrs_status vrs_healpix_lmn(int32_t nside, double *lmn);
/* NESTED -> RING index map [12 nside^2]. */
vrs_status vrs_healpix_nest2ring(int32_t nside, int64_t *ring_of_nest);
/* RING index of each RING pixel's antipode [12 nside^2] (pixel centres are
 * antipodally symmetric: ring R and ring 4 nside - R, azimuth phi + pi); the
 * vis-gen and recon kernels pair a pixel with its antipode under
 * vrs_knobs.antipodal. */
vrs_status vrs_healpix_antipode(int32_t nside, int64_t *antipode_of_ring);

#ifdef __cplusplus
}
#endif

#endif /* VRS_H */
