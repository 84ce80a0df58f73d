// Host-side synthetic inputs and HEALPix geometry (no GPU needed).
#pragma once
#include <cstdint>

namespace vrs {
void rng_normals(uint64_t seed, int64_t n, double *out);
void rng_uniforms(uint64_t seed, int64_t n, double *out);
void sky_lognormal(uint64_t seed, double sigma, int64_t n_pix, int32_t n_src, double s_min,
                   double s_max, float *out);
void orbit_offsets(int32_t n_sat, uint64_t seed, double a_km, double max_offset_km,
                   double *raan, double *m0);
void healpix_ring_vec(int64_t nside, int64_t p, double v[3]);
int64_t healpix_nest2ring(int64_t nside, int64_t p);
int64_t healpix_ring_antipode(int64_t nside, int64_t p);
void uvw_position_index(int32_t orbit_res, int32_t orbit_counts, int32_t interval, int32_t span,
                        int64_t *out);
}  // namespace vrs
