// Host-side synthetic inputs and HEALPix geometry.  None of this exists in the
// reference, which reads its sky (l, m, n, FF, NX ...) and uvw counts from
// MATLAB-produced text files (vissGen10M_shade.cu:199-286).  These functions
// need no GPU.
#include <cmath>
#include <cstdint>
#include <vector>

#include "host_sky.hpp"

namespace vrs {

namespace {
constexpr double kPi = 3.14159265358979323846;

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }
};

// Box-Muller over (u1, u2); emits the cos branch, then the sin branch.
struct NormalStream {
    SplitMix64 g;
    bool have = false;
    double spare = 0.0;
    double next() {
        if (have) {
            have = false;
            return spare;
        }
        double u1 = g.uniform(), u2 = g.uniform();
        double rad = std::sqrt(-2.0 * std::log(1.0 - u1));
        spare = rad * std::sin(2.0 * kPi * u2);
        have = true;
        return rad * std::cos(2.0 * kPi * u2);
    }
};

int64_t spread_bits_inverse(int64_t v) {  // keep even bits and pack them
    v &= 0x5555555555555555LL;
    v = (v | (v >> 1)) & 0x3333333333333333LL;
    v = (v | (v >> 2)) & 0x0F0F0F0F0F0F0F0FLL;
    v = (v | (v >> 4)) & 0x00FF00FF00FF00FFLL;
    v = (v | (v >> 8)) & 0x0000FFFF0000FFFFLL;
    v = (v | (v >> 16)) & 0x00000000FFFFFFFFLL;
    return v;
}
}  // namespace

void rng_normals(uint64_t seed, int64_t n, double *out) {
    NormalStream ns{SplitMix64{seed}};
    for (int64_t i = 0; i < n; ++i) out[i] = ns.next();
}

void rng_uniforms(uint64_t seed, int64_t n, double *out) {
    SplitMix64 g{seed};
    for (int64_t i = 0; i < n; ++i) out[i] = g.uniform();
}

void sky_lognormal(uint64_t seed, double sigma, int64_t n_pix, int32_t n_src, double s_min,
                   double s_max, float *out) {
    NormalStream ns{SplitMix64{seed}};
    std::vector<double> I(n_pix);
    for (int64_t p = 0; p < n_pix; ++p) I[p] = std::exp(sigma * ns.next());
    // point sources continue the same underlying SplitMix64 stream
    SplitMix64 &g = ns.g;
    std::vector<char> used(n_pix, 0);
    for (int32_t k = 0; k < n_src; ++k) {
        int64_t p;
        do {
            p = (int64_t)(g.uniform() * (double)n_pix);
            if (p >= n_pix) p = n_pix - 1;
        } while (used[p]);
        used[p] = 1;
        double u = g.uniform();
        // inverse CDF of p(S) ~ S^-2 on [s_min, s_max]
        double S = 1.0 / (1.0 / s_min - u * (1.0 / s_min - 1.0 / s_max));
        I[p] += S;
    }
    for (int64_t p = 0; p < n_pix; ++p) out[p] = (float)I[p];
}

void orbit_offsets(int32_t n_sat, uint64_t seed, double a_km, double max_offset_km,
                   double *raan, double *m0) {
    SplitMix64 g{seed};
    for (int32_t k = 0; k < n_sat; ++k) {
        raan[k] = 2.0 * kPi * (double)k / (double)n_sat;
        double d = max_offset_km * g.uniform();
        m0[k] = 2.0 * std::asin(d / (2.0 * a_km));
    }
}

// RING pix2vec, Gorski et al. 2005 sec. 4.
void healpix_ring_vec(int64_t nside, int64_t p, double v[3]) {
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    const double ns2 = (double)nside * (double)nside;
    double z, phi;
    if (p < ncap) {
        int64_t i = (int64_t)(0.5 * (1.0 + std::sqrt(1.0 + 2.0 * (double)p)));
        while (2 * i * (i - 1) > p) --i;
        while (2 * i * (i + 1) <= p) ++i;
        int64_t j = p + 1 - 2 * i * (i - 1);
        z = 1.0 - (double)(i * i) / (3.0 * ns2);
        phi = ((double)j - 0.5) * kPi / (2.0 * (double)i);
    } else if (p < npix - ncap) {
        int64_t q = p - ncap;
        int64_t i = q / (4 * nside) + nside;
        int64_t j = q % (4 * nside) + 1;
        double sft = ((i + nside) % 2 == 1) ? 1.0 : 0.5;
        z = 4.0 / 3.0 - 2.0 * (double)i / (3.0 * (double)nside);
        phi = ((double)j - sft) * kPi / (2.0 * (double)nside);
    } else {
        int64_t q = npix - p;
        int64_t i = (int64_t)(0.5 * (1.0 + std::sqrt(2.0 * (double)q - 1.0)));
        while (2 * i * (i - 1) >= q) --i;
        while (2 * i * (i + 1) < q) ++i;
        int64_t j = 4 * i + 1 - (q - 2 * i * (i - 1));
        z = -1.0 + (double)(i * i) / (3.0 * ns2);
        phi = ((double)j - 0.5) * kPi / (2.0 * (double)i);
    }
    double st = std::sqrt((1.0 - z) * (1.0 + z));
    v[0] = st * std::cos(phi);
    v[1] = st * std::sin(phi);
    v[2] = z;
}

// RING index of the antipode of RING pixel p.  HEALPix centres are antipodally
// symmetric: ring R (1 .. 4 nside - 1, from the north pole) and ring
// 4 nside - R hold the same number of pixels at the same azimuths, and pixel j
// of ring R faces pixel j + (ring size) / 2 of ring 4 nside - R (phi + pi).
int64_t healpix_ring_antipode(int64_t nside, int64_t p) {
    const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    auto start = [&](int64_t R) {
        if (R < nside) return 2 * R * (R - 1);
        if (R <= 3 * nside) return ncap + (R - nside) * 4 * nside;
        const int64_t i = 4 * nside - R;
        return npix - 2 * i * (i + 1);
    };
    auto size = [&](int64_t R) {
        if (R < nside) return 4 * R;
        if (R <= 3 * nside) return 4 * nside;
        return 4 * (4 * nside - R);
    };
    int64_t R;
    if (p < ncap) {
        R = (int64_t)(0.5 * (1.0 + std::sqrt(1.0 + 2.0 * (double)p)));
        while (2 * R * (R - 1) > p) --R;
        while (2 * R * (R + 1) <= p) ++R;
    } else if (p < npix - ncap) {
        R = (p - ncap) / (4 * nside) + nside;
    } else {
        const int64_t q = npix - p;
        int64_t i = (int64_t)(0.5 * (1.0 + std::sqrt(2.0 * (double)q - 1.0)));
        while (2 * i * (i - 1) >= q) --i;
        while (2 * i * (i + 1) < q) ++i;
        R = 4 * nside - i;
    }
    const int64_t j = p - start(R), n = size(R);
    const int64_t R2 = 4 * nside - R;
    return start(R2) + (j + n / 2) % n;
}

int64_t healpix_nest2ring(int64_t nside, int64_t p) {
    static const int jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
    static const int jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
    const int64_t nf = nside * nside;
    const int face = (int)(p / nf);
    const int64_t q = p % nf;
    const int64_t ix = spread_bits_inverse(q), iy = spread_bits_inverse(q >> 1);
    const int64_t jr = (int64_t)jrll[face] * nside - ix - iy - 1;
    int64_t nr, before, shift;
    if (jr < nside) {
        nr = jr;
        before = 2 * nr * (nr - 1);
        shift = 0;
    } else if (jr > 3 * nside) {
        nr = 4 * nside - jr;
        before = 12 * nf - 2 * (nr + 1) * nr;
        shift = 0;
    } else {
        nr = nside;
        before = 2 * nside * (nside - 1) + (jr - nside) * 4 * nside;
        shift = (jr - nside) & 1;
    }
    int64_t jp = ((int64_t)jpll[face] * nr + ix - iy + 1 + shift) / 2;
    if (jp > 4 * nr) jp -= 4 * nr;
    else if (jp < 1) jp += 4 * nr;
    return before + jp - 1;
}

}  // namespace vrs

namespace vrs {

// The reference's one-third-orbit windows (uvwPosition,
// MIncline_generate_uvw_float.cu:113-151).  Output block k (0-based orbit) holds
// span = orbit_res / 3 steps starting at k * (orbit_res + interval) in the
// propagated series of orbit_counts orbits; once such a start lies past the
// series, whole orbits are taken off again (the first "reset" orbits), keeping
// the k * interval offset.  A step past the end of the series is zero-filled (-1).
void uvw_position_index(int32_t orbit_res, int32_t orbit_counts, int32_t interval, int32_t span,
                        int64_t *out) {
    const int64_t total = (int64_t)orbit_res * orbit_counts;
    const int64_t reset = total / ((int64_t)orbit_res + interval) + 1;
    for (int64_t k = 0; k < orbit_counts; ++k) {
        int64_t first = k * orbit_res + k * interval;
        if (k * ((int64_t)orbit_res + interval) > total) first -= reset * orbit_res;
        const bool runs_off = first + span > total;
        for (int64_t j = 0; j < span; ++j) {
            const int64_t src = first + j;
            out[k * span + j] = (runs_off && src >= total) ? -1 : src;
        }
    }
}

}  // namespace vrs
