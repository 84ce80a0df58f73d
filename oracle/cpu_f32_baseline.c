/*
 * cpu_f32_baseline.c -- the reference's plain FP32 vis-gen / recon arithmetic
 * restated for host cores with OpenMP.  It is the CPU baseline that bench.py
 * times beside the GPU; it is test/benchmark infrastructure, never part of
 * the product library.
 *
 * Follows vissGen10M_shade_optim1.cu:130-206 (visscal) and :210-272
 * (imagerecon): every quantity is float, the phase is formed in float, the
 * complex exponential is complexExp (vissGen10M_shade.cu:35-39), i.e. cosf and
 * sinf of the radian argument -2*pi*phase formed as the float product
 * (-2*pi)*phase, accumulation is a float complex sum, and the image is
 * acc / N_vis (vissGen10M_shade.cu:173).  The shade mask is the geometric
 * occultation test evaluated in float (the reference's fine-grid box loop is
 * a no-op, vissGen10M_shade.cu:130).
 *
 * Knobs: s = phase scale (reference 1/dl = 0.5, vissGen10M_shade.cu:183-185),
 * wfac = trailing exp(-2 pi i s w) on V and n (not n-1) in recon.
 */
#include <math.h>
#include <stdint.h>

#define VB_PI 3.14159265358979323846

static inline int vb_occ(const float *r, float r2, const float *s, float R2)
{
    float d = r[0] * s[0] + r[1] * s[1] + r[2] * s[2];
    return d < 0.0f && (r2 - d * d) <= R2;
}

/* V[sel] complex64 interleaved; pos_f [n_sat][n_t][3] km; uvw_f [n_vis][3];
 * lmn_f [n_pix][3]; I [n_pix]. */
void vb_visgen_f32(int64_t n_t, const float *pos_f, const int *pairs, const float *uvw_f,
                   int64_t n_pix, const float *lmn_f, const float *I, float R, float s,
                   int wfac, float t_occ, int64_t n_sel, const int64_t *vis_idx, float *V)
{
    const float R2 = R * R;
    const float m2pi = -(2.0f * (float)VB_PI);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n_sel; ++k) {
        int64_t v = vis_idx ? vis_idx[k] : k;
        int64_t b = v / n_t, t = v % n_t;
        const float *ra = pos_f + ((int64_t)pairs[2 * b] * n_t + t) * 3;
        const float *rb = pos_f + ((int64_t)pairs[2 * b + 1] * n_t + t) * 3;
        float ra2 = ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2];
        float rb2 = rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2];
        float u = uvw_f[3 * v] * s, vv = uvw_f[3 * v + 1] * s, w = uvw_f[3 * v + 2] * s;
        float are = 0.0f, aim = 0.0f;
        for (int64_t p = 0; p < n_pix; ++p) {
            const float *sp = lmn_f + 3 * p;
            float c = (vb_occ(ra, ra2, sp, R2) | vb_occ(rb, rb2, sp, R2)) ? t_occ : I[p];
            float phase = u * sp[0] + vv * sp[1] + w * (sp[2] - 1.0f);
            float arg = m2pi * phase;
            are += c * cosf(arg);
            aim += c * sinf(arg);
        }
        if (wfac) {
            float arg = m2pi * w;
            float cr = cosf(arg), ci = sinf(arg);
            float tr = are * cr - aim * ci;
            aim = are * ci + aim * cr;
            are = tr;
        }
        V[2 * k] = are;
        V[2 * k + 1] = aim;
    }
}

/* F[sel] = Re( sum_v w_v (1-M) V_v exp(+2 pi i s psi) ) / N_vis */
void vb_recon_f32(int64_t n_t, int n_base, const float *pos_f, const int *pairs,
                  const float *uvw_f, int64_t n_pix, const float *lmn_f, const float *V,
                  const float *wgt, float R, float s, int wfac, int64_t n_sel,
                  const int64_t *pix_idx, float *F)
{
    const float R2 = R * R;
    const float p2pi = 2.0f * (float)VB_PI;
    const int64_t n_vis = (int64_t)n_base * n_t;
    (void)n_pix;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n_sel; ++k) {
        int64_t p = pix_idx ? pix_idx[k] : k;
        const float *sp = lmn_f + 3 * p;
        float l = sp[0] * s, m = sp[1] * s, n = (wfac ? sp[2] : sp[2] - 1.0f) * s;
        float are = 0.0f;
        for (int64_t v = 0; v < n_vis; ++v) {
            int64_t b = v / n_t, t = v % n_t;
            const float *ra = pos_f + ((int64_t)pairs[2 * b] * n_t + t) * 3;
            const float *rb = pos_f + ((int64_t)pairs[2 * b + 1] * n_t + t) * 3;
            float ra2 = ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2];
            float rb2 = rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2];
            if (vb_occ(ra, ra2, sp, R2) | vb_occ(rb, rb2, sp, R2)) continue;
            float phase = uvw_f[3 * v] * l + uvw_f[3 * v + 1] * m + uvw_f[3 * v + 2] * n;
            float arg = p2pi * phase;
            float wv = wgt ? wgt[v] : 1.0f;
            are += wv * (V[2 * v] * cosf(arg) - V[2 * v + 1] * sinf(arg));
        }
        F[k] = are / (float)n_vis;
    }
}
