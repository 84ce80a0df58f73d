/*
 * oracle.c -- FP64 CPU restatement of the viss_recon_shade hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it, and only as the checker.
 *
 * What each function restates (reference = SSshuishui/viss_recon_shade):
 *   vo_splitmix64 / vo_normals ....... synthetic-input RNG (no reference counterpart;
 *                                      SplitMix64 + Box-Muller, cos branch first)
 *   vo_healpix_ring_vec / _nest_vec .. HEALPix pix2vec (Gorski et al. 2005, ApJ 622:759, sec. 4);
 *                                      stands in for the absent MATLAB l,m,n inputs
 *                                      (vissGen10M_shade.cu:199-248)
 *   vo_orbit_positions ............... calculateOrbit (MIncline_generate_uvw_float.cu:83-110),
 *                                      e = 0 so E = M = nu; FP64; time via mean motion sqrt(mu/a^3)
 *   vo_uvw ........................... computeUVW (MIncline_generate_uvw_float.cu:153-160),
 *                                      without ceil, lambda = c/f
 *   vo_uvw_position_index ............ uvwPosition's one-third-orbit windows
 *                                      (MIncline_generate_uvw_float.cu:113-151)
 *   (pos == NULL in vo_visgen, vo_visgen_fine, vo_recon: the reference's
 *    stage-3/4 inputs, uvw rows without satellites, on which its own shading
 *    never masks a pixel, vissGen10M_shade.cu:130, 165)
 *   vo_occ / vo_mask_* ............... the occultation ("shade") predicate: ray from a
 *                                      satellite towards a pixel hits the closed lunar sphere.
 *                                      Replaces ShadeMatrix's box (shade10M.cu:73-105) with the
 *                                      geometric cap that the box approximates.
 *   vo_visgen ........................ visscal (vissGen10M_shade_optim1.cu:130-206;
 *                                      vissGen10M_shade.cu:108-145): V = sum_p I'_p exp(-2 pi i s phi),
 *                                      phi = u l + v m + w (n - 1), I' = T_occ where shaded,
 *                                      optional trailing exp(-2 pi i s w) (vissGen10M_shade.cu:143)
 *   vo_recon ......................... imagerecon (vissGen10M_shade_optim1.cu:210-272;
 *                                      vissGen10M_shade.cu:149-175): F = kappa sum_v w_v (1 - M)
 *                                      Re(V exp(+2 pi i s psi)), psi uses n (wfac=1) or n-1 (wfac=0)
 *   vo_shade_box ..................... ShadeMatrix + Cart2SphKernel (shade10M.cu:62-105,218-224)
 *   vo_fine_index .................... computeLocCount / computeLocViss
 *                                      (vissGen1M_shade_optim2.cu:59-119) as one counting pass
 *   vo_visgen_fine ................... visscal's fine-cell sum with 240 replacement
 *                                      (vissGen10M_shade.cu:121-138), cell-level geometric mask
 *
 * Every floating-point decision that must be bit-exact against the GPU (the
 * occultation predicate and its near-limb band) is written with explicit fma()
 * so that the compiler cannot re-associate or contract it; build with
 * -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define VO_PI 3.14159265358979323846

/* ------------------------------------------------------------------ RNG */
uint64_t vo_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* uniform in [0,1) with 53 random bits */
double vo_uniform(uint64_t *state)
{
    return (double)(vo_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* n standard normals from a generator seeded with `seed`.  Pairs (u1,u2):
 * rad = sqrt(-2 ln(1-u1)), z0 = rad cos(2 pi u2), z1 = rad sin(2 pi u2). */
void vo_normals(uint64_t seed, int64_t n, double *out)
{
    uint64_t st = seed;
    int64_t i = 0;
    while (i < n) {
        double u1 = vo_uniform(&st), u2 = vo_uniform(&st);
        double rad = sqrt(-2.0 * log(1.0 - u1));
        out[i++] = rad * cos(2.0 * VO_PI * u2);
        if (i < n) out[i++] = rad * sin(2.0 * VO_PI * u2);
    }
}

/* ------------------------------------------------------------- HEALPix */
void vo_healpix_ring_vec(int64_t nside, int64_t p, double v[3])
{
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    double z, phi;
    if (p < ncap) {                       /* north polar cap */
        int64_t i = (int64_t)((1.0 + sqrt(1.0 + 2.0 * (double)p)) / 2.0);
        /* guard the floor of the square root against rounding */
        while (2 * i * (i - 1) > p) --i;
        while (2 * (i + 1) * i <= p) ++i;
        int64_t j = p + 1 - 2 * i * (i - 1);
        z = 1.0 - (double)(i * i) / (3.0 * (double)(nside * nside));
        phi = ((double)j - 0.5) * VO_PI / (2.0 * (double)i);
    } else if (p < npix - ncap) {         /* equatorial belt */
        int64_t ip = p - ncap;
        int64_t i = ip / (4 * nside) + nside;
        int64_t j = ip % (4 * nside) + 1;
        double s = ((i + nside) & 1) ? 1.0 : 0.5;
        z = 4.0 / 3.0 - 2.0 * (double)i / (3.0 * (double)nside);
        phi = ((double)j - s) * VO_PI / (2.0 * (double)nside);
    } else {                              /* south polar cap (mirror) */
        int64_t ps = npix - p;
        int64_t i = (int64_t)((1.0 + sqrt(2.0 * (double)ps - 1.0)) / 2.0);
        while (2 * i * (i - 1) >= ps) --i;
        while (2 * (i + 1) * i < ps) ++i;
        int64_t j = 4 * i + 1 - (ps - 2 * i * (i - 1));
        z = -1.0 + (double)(i * i) / (3.0 * (double)(nside * nside));
        phi = ((double)j - 0.5) * VO_PI / (2.0 * (double)i);
    }
    double st = sqrt((1.0 - z) * (1.0 + z));
    v[0] = st * cos(phi);
    v[1] = st * sin(phi);
    v[2] = z;
}

static int64_t vo_compact_bits(int64_t x)
{
    x &= 0x5555555555555555LL;
    x = (x | (x >> 1)) & 0x3333333333333333LL;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0FLL;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFLL;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFLL;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFLL;
    return x;
}

static const int vo_jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
static const int vo_jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

/* NESTED pixel -> (ring number ir in 1..4nside-1, position jp in 1..nr, nr, z).
 * Independent of the RING formulas above: it goes face -> (ix,iy) -> ring. */
static void vo_nest_ring_coords(int64_t nside, int64_t p, int64_t *ir, int64_t *jp,
                                int64_t *nr, int *shifted)
{
    int64_t npface = nside * nside;
    int face = (int)(p / npface);
    int64_t ipf = p % npface;
    int64_t ix = vo_compact_bits(ipf), iy = vo_compact_bits(ipf >> 1);
    int64_t jr = (int64_t)vo_jrll[face] * nside - ix - iy - 1;
    int64_t n, kshift;
    if (jr < nside) {
        n = jr; kshift = 0;
    } else if (jr > 3 * nside) {
        n = 4 * nside - jr; kshift = 0;
    } else {
        n = nside; kshift = (jr - nside) & 1;
    }
    int64_t j = ((int64_t)vo_jpll[face] * n + ix - iy + 1 + kshift) / 2;
    if (j > 4 * n) j -= 4 * n;
    if (j < 1) j += 4 * n;
    *ir = jr; *jp = j; *nr = n; *shifted = (int)kshift;
}

int64_t vo_nest2ring(int64_t nside, int64_t p)
{
    int64_t ir, jp, nr; int sh;
    vo_nest_ring_coords(nside, p, &ir, &jp, &nr, &sh);
    int64_t ncap = 2 * nside * (nside - 1), npix = 12 * nside * nside;
    if (ir < nside) return 2 * ir * (ir - 1) + jp - 1;
    if (ir > 3 * nside) {
        int64_t irs = 4 * nside - ir;
        return npix - 2 * irs * (irs + 1) + jp - 1;
    }
    return ncap + (ir - nside) * 4 * nside + jp - 1;
}

void vo_healpix_nest_vec(int64_t nside, int64_t p, double v[3])
{
    int64_t ir, jp, nr; int sh;
    vo_nest_ring_coords(nside, p, &ir, &jp, &nr, &sh);
    double z, phi;
    double fn = (double)nside;
    if (ir < nside) {
        z = 1.0 - (double)(ir * ir) / (3.0 * fn * fn);
        phi = ((double)jp - 0.5) * VO_PI / (2.0 * (double)nr);
    } else if (ir > 3 * nside) {
        int64_t irs = 4 * nside - ir;
        z = -1.0 + (double)(irs * irs) / (3.0 * fn * fn);
        phi = ((double)jp - 0.5) * VO_PI / (2.0 * (double)nr);
    } else {
        z = 4.0 / 3.0 - 2.0 * (double)ir / (3.0 * fn);
        phi = ((double)jp - (sh ? 1.0 : 0.5)) * VO_PI / (2.0 * fn);
    }
    double st = sqrt((1.0 - z) * (1.0 + z));
    v[0] = st * cos(phi);
    v[1] = st * sin(phi);
    v[2] = z;
}

/* --------------------------------------------------------------- orbits */
/* Positions [n_sat][n_t][3] in km.  Circular orbit (e = 0), argument of
 * perigee 0: w = nu = M = m0[k] + sqrt(mu/a^3) t_j, t_j = j dt.
 * x = a (cos w cos O - sin w cos i sin O), y = a (cos w sin O + sin w cos i cos O),
 * z = a sin w sin i  (MIncline_generate_uvw_float.cu:106-108). */
void vo_orbit_positions(int n_sat, int64_t n_t, double dt, double a, double incl,
                        double mu, const double *raan, const double *m0, double *pos)
{
    double nmot = sqrt(mu / (a * a * a));
    double ci = cos(incl), si = sin(incl);
    for (int k = 0; k < n_sat; ++k) {
        double cO = cos(raan[k]), sO = sin(raan[k]);
        for (int64_t j = 0; j < n_t; ++j) {
            double w = m0[k] + nmot * ((double)j * dt);
            double cw = cos(w), sw = sin(w);
            double *r = pos + ((int64_t)k * n_t + j) * 3;
            r[0] = a * (cw * cO - sw * ci * sO);
            r[1] = a * (cw * sO + sw * ci * cO);
            r[2] = a * (sw * si);
        }
    }
}

/* uvw[b][t][3] = (pos[pb][t] - pos[pa][t]) * 1e3 / lambda, lambda in metres. */
void vo_uvw(int64_t n_t, const double *pos, int n_base, const int *pairs, double lambda_m,
            double *uvw)
{
    double sc = 1e3 / lambda_m;
    for (int b = 0; b < n_base; ++b) {
        const double *ra = pos + (int64_t)pairs[2 * b] * n_t * 3;
        const double *rb = pos + (int64_t)pairs[2 * b + 1] * n_t * 3;
        for (int64_t t = 0; t < n_t; ++t)
            for (int c = 0; c < 3; ++c)
                uvw[((int64_t)b * n_t + t) * 3 + c] = (rb[t * 3 + c] - ra[t * 3 + c]) * sc;
    }
}

/* uvwPosition (MIncline_generate_uvw_float.cu:113-151) as index arithmetic:
 * for each output step g of orbit k (1-based), the source sample in the
 * propagated series of OR * OC samples, or -1 where the reference writes 0.
 * The reference zero-fills only circleIdx > OR*OC (:134) and so reads one
 * sample past the series at circleIdx == OR*OC; here that sample is -1 too. */
void vo_uvw_position_index(int OR, int OC, int interval, int span, int64_t *out)
{
    const int64_t n_out = (int64_t)OC * (OR / 3);
    for (int64_t g = 0; g < n_out; ++g) {
        int64_t k = g / (OR / 3) + 1;
        int64_t offset = (k - 1) * interval;
        int64_t endIdx = (k - 1) * OR + offset + span;
        int64_t startIdx = (k - 1) * OR + offset;
        if ((k - 1) * ((int64_t)OR + interval) > (int64_t)OR * OC) {
            int64_t reset_index = ((int64_t)OR * OC) / ((int64_t)OR + interval) + 1;
            startIdx = (k - 1 - reset_index) * OR + offset;
            endIdx = (k - 1 - reset_index) * OR + offset + span;
        }
        int64_t circleIdx = startIdx + (g - (k - 1) * span);
        if (endIdx > (int64_t)OR * OC && circleIdx >= (int64_t)OR * OC)
            out[g] = -1;
        else
            out[g] = circleIdx;
    }
}

/* ------------------------------------------------------------------ mask */
/* Ray from r (km) along unit s hits the closed sphere of radius R:
 *   d = r.s < 0  and  q = |r|^2 - d^2 <= R^2.
 * near-limb: q in ((R-e)^2, (R+e)^2) with e = 1e-6 R, front side only. */
typedef struct { double R2, lo2, hi2; } vo_limb;

static vo_limb vo_make_limb(double R)
{
    vo_limb L;
    double e = 1e-6 * R;
    L.R2 = R * R;
    L.lo2 = (R - e) * (R - e);
    L.hi2 = (R + e) * (R + e);
    return L;
}

static inline double vo_r2(const double *r)
{
    return fma(r[2], r[2], fma(r[1], r[1], r[0] * r[0]));
}

static inline int vo_occ(const double *r, double r2, const double *s, const vo_limb *L,
                         int *near)
{
    double d = fma(r[2], s[2], fma(r[1], s[1], r[0] * s[0]));
    if (!(d < 0.0)) return 0;
    double q = fma(-d, d, r2);
    if (q > L->lo2 && q < L->hi2) *near = 1;
    return q <= L->R2;
}

int vo_occ_one(const double *r, const double *s, double R, int *near)
{
    vo_limb L = vo_make_limb(R);
    *near = 0;
    return vo_occ(r, vo_r2(r), s, &L, near);
}

uint32_t vo_pix_hash(int64_t p)
{
    uint64_t z = (uint64_t)p + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (uint32_t)(z ^ (z >> 32));
}

/* Per selected visibility v = b*n_t + t: number of shaded pixels, XOR of
 * vo_pix_hash over shaded pixel indices, and number of near-limb pairs.
 * vis_idx == NULL selects all visibilities. */
void vo_mask_stats(int64_t n_t, const double *pos, const int *pairs, int64_t n_pix,
                   const double *lmn, double R, int64_t n_sel, const int64_t *vis_idx,
                   int32_t *count, uint32_t *hash, int64_t *near_total)
{
    vo_limb L = vo_make_limb(R);
    int64_t near_sum = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : near_sum)
    for (int64_t k = 0; k < n_sel; ++k) {
        int64_t v = vis_idx ? vis_idx[k] : k;
        int64_t b = v / n_t, t = v % n_t;
        const double *ra = pos + ((int64_t)pairs[2 * b] * n_t + t) * 3;
        const double *rb = pos + ((int64_t)pairs[2 * b + 1] * n_t + t) * 3;
        double ra2 = vo_r2(ra), rb2 = vo_r2(rb);
        int32_t c = 0; uint32_t h = 0; int64_t nl = 0;
        for (int64_t p = 0; p < n_pix; ++p) {
            int na = 0, nb = 0;
            int m = vo_occ(ra, ra2, lmn + 3 * p, &L, &na) | vo_occ(rb, rb2, lmn + 3 * p, &L, &nb);
            if (m) { ++c; h ^= vo_pix_hash(p); }
            nl += (na | nb);
        }
        count[k] = c; hash[k] = h; near_sum += nl;
    }
    *near_total = near_sum;
}

/* Full shade bitmap [n_sel][ceil(n_pix/32)] (bit p%32 of word p/32). */
void vo_mask_bits(int64_t n_t, const double *pos, const int *pairs, int64_t n_pix,
                  const double *lmn, double R, int64_t n_sel, const int64_t *vis_idx,
                  uint32_t *bits)
{
    vo_limb L = vo_make_limb(R);
    int64_t nw = (n_pix + 31) / 32;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t k = 0; k < n_sel; ++k) {
        int64_t v = vis_idx ? vis_idx[k] : k;
        int64_t b = v / n_t, t = v % n_t;
        const double *ra = pos + ((int64_t)pairs[2 * b] * n_t + t) * 3;
        const double *rb = pos + ((int64_t)pairs[2 * b + 1] * n_t + t) * 3;
        double ra2 = vo_r2(ra), rb2 = vo_r2(rb);
        uint32_t *row = bits + k * nw;
        memset(row, 0, (size_t)nw * 4);
        for (int64_t p = 0; p < n_pix; ++p) {
            int na = 0, nb = 0;
            if (vo_occ(ra, ra2, lmn + 3 * p, &L, &na) | vo_occ(rb, rb2, lmn + 3 * p, &L, &nb))
                row[p >> 5] |= 1u << (p & 31);
        }
    }
}

/* ------------------------------------------------------------ vis-gen */
/* V[map][sel] (complex, interleaved re/im) for the selected visibilities.
 *   phi = u l + v m + w (n - 1)       (vissGen10M_shade_optim1.cu:193)
 *   V   = sum_p I'_p exp(-2 pi i s phi) [* exp(-2 pi i s w) if wfac]
 *   I'  = t_occ where shaded, else I  (vissGen10M_shade.cu:130-134, intended)
 * Phase is reduced in FP64 to r = phi - rint(phi) before sin/cos(2 pi r). */
void vo_visgen(int64_t n_t, const double *pos, const int *pairs, const double *uvw,
               int64_t n_pix, const double *lmn, int n_maps, const float *I, double R,
               double s, int wfac, double t_occ, int64_t n_sel, const int64_t *vis_idx,
               double *V)
{
    vo_limb L = vo_make_limb(R);
    double *lmn_nm1 = (double *)malloc(sizeof(double) * (size_t)n_pix);
    for (int64_t p = 0; p < n_pix; ++p) lmn_nm1[p] = lmn[3 * p + 2] - 1.0;
#pragma omp parallel for schedule(dynamic, 2)
    for (int64_t k = 0; k < n_sel; ++k) {
        int64_t v = vis_idx ? vis_idx[k] : k;
        int64_t b = v / n_t, t = v % n_t;
        /* pos == NULL: uvw rows without satellites, nothing is occulted */
        const double *ra = pos ? pos + ((int64_t)pairs[2 * b] * n_t + t) * 3 : NULL;
        const double *rb = pos ? pos + ((int64_t)pairs[2 * b + 1] * n_t + t) * 3 : NULL;
        double ra2 = pos ? vo_r2(ra) : 0.0, rb2 = pos ? vo_r2(rb) : 0.0;
        double u = uvw[3 * v] * s, vv = uvw[3 * v + 1] * s, w = uvw[3 * v + 2] * s;
        double wre = 1.0, wim = 0.0;
        if (wfac) {
            double r = w - rint(w);
            wre = cos(2.0 * VO_PI * r);
            wim = -sin(2.0 * VO_PI * r);
        }
        /* every map sums its pixels in ascending order; the mask and the phase
         * factor of a (visibility, pixel) pair are formed once for all maps */
        double *acc = (double *)calloc(2 * (size_t)n_maps, sizeof(double));
        for (int64_t p = 0; p < n_pix; ++p) {
            int na = 0, nb = 0;
            int m = pos && (vo_occ(ra, ra2, lmn + 3 * p, &L, &na) |
                            vo_occ(rb, rb2, lmn + 3 * p, &L, &nb));
            double phi = u * lmn[3 * p] + vv * lmn[3 * p + 1] + w * lmn_nm1[p];
            double r = phi - rint(phi);
            double c = cos(2.0 * VO_PI * r), sn = sin(2.0 * VO_PI * r);
            for (int mp = 0; mp < n_maps; ++mp) {
                double ip = m ? t_occ : (double)I[(int64_t)mp * n_pix + p];
                acc[2 * mp] += ip * c;
                acc[2 * mp + 1] -= ip * sn;
            }
        }
        for (int mp = 0; mp < n_maps; ++mp) {
            double are = acc[2 * mp], aim = acc[2 * mp + 1];
            double *o = V + ((int64_t)mp * n_sel + k) * 2;
            o[0] = are * wre - aim * wim;
            o[1] = are * wim + aim * wre;
        }
        free(acc);
    }
    free(lmn_nm1);
}

/* -------------------------------------------------------------- recon */
/* Inverse index of a fine-cell sky: count[q] = number of cells whose parent
 * is q (computeLocCount, vissGen1M_shade_optim2.cu:59-82), start[q] = its
 * exclusive prefix sum and idx = cell indices grouped by parent in ascending
 * cell order (computeLocViss, :86-119), here by one counting pass. */
void vo_fine_index(int64_t n_cells, const int32_t *nx, int32_t nx_base, int64_t n_pix,
                   int32_t *count, int64_t *start, int32_t *idx)
{
    for (int64_t q = 0; q < n_pix; ++q) count[q] = 0;
    for (int64_t c = 0; c < n_cells; ++c) ++count[nx[c] - nx_base];
    int64_t acc = 0;
    for (int64_t q = 0; q < n_pix; ++q) { start[q] = acc; acc += count[q]; }
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_pix > 0 ? n_pix : 1));
    for (int64_t q = 0; q < n_pix; ++q) fill[q] = start[q];
    for (int64_t c = 0; c < n_cells; ++c) idx[fill[nx[c] - nx_base]++] = (int32_t)c;
    free(fill);
}

/* Vis-gen over a fine-cell sky (visscal, vissGen10M_shade.cu:108-145): pixel q
 * contributes C_q(v) = (1/count_q) sum over its cells c of (c shaded ? t_occ :
 * FF_c) at the pixel's direction lmn[q]; a cell is shaded when its own
 * direction is occulted for either satellite (the geometric mask replaces the
 * reference's index box, :129-134).  Same phase conventions as vo_visgen. */
void vo_visgen_fine(int64_t n_t, const double *pos, const int *pairs, const double *uvw,
                    int64_t n_pix, const double *lmn, int64_t n_cells, const double *cell_lmn,
                    const float *ff, const int32_t *count, const int64_t *start,
                    const int32_t *idx, double R, double s, int wfac, double t_occ,
                    int64_t n_sel, const int64_t *vis_idx, double *V)
{
    (void)n_cells;
    vo_limb L = vo_make_limb(R);
#pragma omp parallel for schedule(dynamic, 2)
    for (int64_t k = 0; k < n_sel; ++k) {
        int64_t v = vis_idx ? vis_idx[k] : k;
        int64_t b = v / n_t, t = v % n_t;
        /* pos == NULL: uvw rows without satellites, nothing is occulted */
        const double *ra = pos ? pos + ((int64_t)pairs[2 * b] * n_t + t) * 3 : NULL;
        const double *rb = pos ? pos + ((int64_t)pairs[2 * b + 1] * n_t + t) * 3 : NULL;
        double ra2 = pos ? vo_r2(ra) : 0.0, rb2 = pos ? vo_r2(rb) : 0.0;
        double u = uvw[3 * v] * s, vv = uvw[3 * v + 1] * s, w = uvw[3 * v + 2] * s;
        double wre = 1.0, wim = 0.0;
        if (wfac) {
            double r = w - rint(w);
            wre = cos(2.0 * VO_PI * r);
            wim = -sin(2.0 * VO_PI * r);
        }
        double are = 0.0, aim = 0.0;
        for (int64_t p = 0; p < n_pix; ++p) {
            double sum = 0.0;
            for (int32_t j = 0; j < count[p]; ++j) {
                int32_t c = idx[start[p] + j];
                int na = 0, nb = 0;
                int m = pos && (vo_occ(ra, ra2, cell_lmn + 3 * (int64_t)c, &L, &na) |
                                vo_occ(rb, rb2, cell_lmn + 3 * (int64_t)c, &L, &nb));
                sum += m ? t_occ : (double)ff[c];
            }
            double ip = count[p] > 0 ? sum / count[p] : 0.0;
            double phi = u * lmn[3 * p] + vv * lmn[3 * p + 1] + w * (lmn[3 * p + 2] - 1.0);
            double r = phi - rint(phi);
            are += ip * cos(2.0 * VO_PI * r);
            aim -= ip * sin(2.0 * VO_PI * r);
        }
        V[2 * k] = are * wre - aim * wim;
        V[2 * k + 1] = are * wim + aim * wre;
    }
}

/* F[map][sel] for the selected pixels (pix_idx == NULL: all):
 *   F_p = kappa sum_v w_v (1 - M_vp) Re(V_v exp(+2 pi i s psi_vp))
 *   psi = u l + v m + w n   (wfac = 1, vissGen10M_shade_optim1.cu:260)
 *   psi = u l + v m + w (n-1) (wfac = 0)
 * V is [n_maps][n_vis] complex (interleaved), w may be NULL (all ones). */
void vo_recon(int64_t n_t, const double *pos, const int *pairs, int n_base, const double *uvw,
              int64_t n_pix, const double *lmn, int n_maps, const double *V, const float *wgt,
              double R, double s, int wfac, double kappa, int64_t n_sel, const int64_t *pix_idx,
              double *F)
{
    vo_limb L = vo_make_limb(R);
    int64_t n_vis = (int64_t)n_base * n_t;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t k = 0; k < n_sel; ++k) {
        int64_t p = pix_idx ? pix_idx[k] : k;
        const double *sp = lmn + 3 * p;
        double nterm = wfac ? sp[2] : sp[2] - 1.0;
        /* every map sums its visibilities in ascending order; the mask and the
         * phase factor of a (visibility, pixel) pair are formed once for all maps */
        double *acc = (double *)calloc((size_t)n_maps, sizeof(double));
        for (int64_t v = 0; v < n_vis; ++v) {
            int64_t b = v / n_t, t = v % n_t;
            if (pos) { /* pos == NULL: uvw rows without satellites, nothing is occulted */
                const double *ra = pos + ((int64_t)pairs[2 * b] * n_t + t) * 3;
                const double *rb = pos + ((int64_t)pairs[2 * b + 1] * n_t + t) * 3;
                int na = 0, nb = 0;
                if (vo_occ(ra, vo_r2(ra), sp, &L, &na) | vo_occ(rb, vo_r2(rb), sp, &L, &nb))
                    continue;
            }
            double u = uvw[3 * v] * s, vv = uvw[3 * v + 1] * s, w = uvw[3 * v + 2] * s;
            double phi = u * sp[0] + vv * sp[1] + w * nterm;
            double r = phi - rint(phi);
            double c = cos(2.0 * VO_PI * r), sn = sin(2.0 * VO_PI * r);
            double wv = wgt ? (double)wgt[v] : 1.0;
            for (int mp = 0; mp < n_maps; ++mp) {
                const double *Vm = V + (int64_t)mp * n_vis * 2;
                acc[mp] += wv * (Vm[2 * v] * c - Vm[2 * v + 1] * sn);
            }
        }
        for (int mp = 0; mp < n_maps; ++mp) F[(int64_t)mp * n_sel + k] = kappa * acc[mp];
        free(acc);
    }
}

/* ------------------------------------------------ reference box predicate */
/* ShadeMatrix (shade10M.cu:73-105) with Cart2SphKernel (shade10M.cu:62-70)
 * and the r0/dtheta/dphi set-up (shade10M.cu:218-224).  The reference works in
 * FP32 with CUDA's atan2f / asinf; this restatement takes the angles in FP64
 * and rounds each grid index once (vo_round_shared), the evaluation the GPU
 * box kernel shares.  Output per uvw: thetaIndS, thetaIndE, phiIndS, phiIndE. */
/* |uvw| in FP64 (products of FP32 inputs are exact in FP64). */
static double vo_row_norm(float u, float v, float w)
{
    double uu = (double)u * u, vv = (double)v * v, ww = (double)w * w;
    return sqrt((uu + vv) + ww);
}

/* Nearest integer of x >= 0; within 1e-9 of a half-integer it goes up (roundf's
 * direction at an exact tie), so that two FP64 libms that differ in the last
 * ulp of atan2 / asin / atan still pick the same index. */
static int vo_round_shared(double x)
{
    double f = floor(x), d = x - f;
    if (fabs(d - 0.5) <= 1e-9) return (int)f + 1;
    return (int)f + (d > 0.5 ? 1 : 0);
}

void vo_shade_box(int64_t n, const float *u, const float *v, const float *w, int res,
                  float r_moon_ref, int32_t *out4)
{
    /* mean |uvw|: FP64 partial sums in a fixed order -- 256 x 256 strided
     * accumulators, a halving tree over each 256, then the 256 partials in turn */
    double sum = 0.0;
    for (int b = 0; b < 256; ++b) {
        double s[256];
        for (int t = 0; t < 256; ++t) {
            double acc = 0.0;
            for (int64_t i = (int64_t)b * 256 + t; i < n; i += 65536)
                acc += vo_row_norm(u[i], v[i], w[i]);
            s[t] = acc;
        }
        for (int o = 128; o > 0; o >>= 1)
            for (int t = 0; t < o; ++t) s[t] += s[t + o];
        sum += s[0];
    }
    double mean_r = sum / (double)n;
    double R = (double)r_moon_ref;
    double dtheta = 2.0 * VO_PI / res, dphi = VO_PI / res;
    int tR = vo_round_shared(atan(R / (mean_r + R)) / dtheta);
    int pR = tR / 2; /* phiR = roundf(thetaR / 2) with integer division */
    for (int64_t i = 0; i < n; ++i) {
        double r = vo_row_norm(u[i], v[i], w[i]);
        double th = atan2((double)v[i], (double)u[i]);
        double ph = r > 0.0 ? asin((double)w[i] / r) : 0.0;
        /* antipode of the baseline direction (shade10M.cu:78-83) */
        int ti = vo_round_shared((2.0 * VO_PI - (th + VO_PI)) / dtheta);
        int pi_ = vo_round_shared((VO_PI - (ph + VO_PI / 2)) / dphi);
        int tS = ti - tR, tE = ti + tR, pS = pi_ - pR, pE = pi_ + pR;
        if (tS <= 0) tS += res;
        if (tE > res) tE -= res;
        if (pS <= 0) pS += res;
        if (pE > res) pE -= res;
        out4[4 * i] = tS;
        out4[4 * i + 1] = tE;
        out4[4 * i + 2] = pS;
        out4[4 * i + 3] = pE;
    }
}

/* ------------------------------------------------ uvw gridding and counts */
/* computeUVW with ceil (MIncline_generate_uvw_float.cu:156-158) for the given
 * pairs, sorted lexicographically by (u, v, w) (SortCompare, :60-70) and made
 * unique (:374-397); counts = samples per unique cell.  Keys pack
 * (u + 2^20, v + 2^20, w + 2^20) into 21-bit fields.  Returns the number of
 * unique cells, or -1 when a component leaves (-2^20, 2^20). */
static int vo_cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y);
}

int64_t vo_grid_keys(int64_t n_t, const double *pos, int n_pairs, const int *pairs, double sc,
                     uint64_t *keys)
{
    for (int p = 0; p < n_pairs; ++p)
        for (int64_t t = 0; t < n_t; ++t) {
            const double *ra = pos + ((int64_t)pairs[2 * p] * n_t + t) * 3;
            const double *rb = pos + ((int64_t)pairs[2 * p + 1] * n_t + t) * 3;
            double u = ceil((rb[0] - ra[0]) * sc), v = ceil((rb[1] - ra[1]) * sc),
                   w = ceil((rb[2] - ra[2]) * sc);
            const double B = 1048576.0;
            if (!(fabs(u) < B && fabs(v) < B && fabs(w) < B)) return -1;
            keys[(int64_t)p * n_t + t] = ((uint64_t)((int64_t)u + 1048576) << 42) |
                                         ((uint64_t)((int64_t)v + 1048576) << 21) |
                                         (uint64_t)((int64_t)w + 1048576);
        }
    return (int64_t)n_pairs * n_t;
}

int64_t vo_grid_unique(int64_t n, uint64_t *keys, float *uvw_unique, int32_t *counts)
{
    qsort(keys, (size_t)n, sizeof(uint64_t), vo_cmp_u64);
    int64_t nu = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (i == 0 || keys[i] != keys[i - 1]) {
            const uint64_t k = keys[i];
            const uint64_t m = (1ull << 21) - 1;
            uvw_unique[3 * nu] = (float)((int64_t)((k >> 42) & m) - 1048576);
            uvw_unique[3 * nu + 1] = (float)((int64_t)((k >> 21) & m) - 1048576);
            uvw_unique[3 * nu + 2] = (float)((int64_t)(k & m) - 1048576);
            counts[nu] = 0;
            ++nu;
        }
        counts[nu - 1] += 1;
    }
    return nu;
}
