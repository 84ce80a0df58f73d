// Reference text formats (io.cpp).  Return 0, -1 (open/IO), -2 (capacity), -3 (truncated).
#pragma once
#include <cstdint>

namespace vrs {
int write_uvw_txt(const char *path, const double *uvw, int64_t n, const float *col4);
int read_uvw_txt(const char *path, int n_cols, double *uvw, float *col4, int64_t cap,
                 int64_t *n_read);
int write_shade_txt(const char *path, const int32_t *box4, int64_t n);
int read_shade_txt(const char *path, int32_t *box4, int64_t cap, int64_t *n_read);
int write_image_txt(const char *path, const float *F, int64_t n);
int read_image_txt(const char *path, float *F, int64_t cap, int64_t *n_read);
int write_ex_txt(const char *path, const float *F, int64_t n_pix, const int32_t *nx,
                 int64_t n_cells, int32_t nx_base);
int write_vis_bin(const char *path, const float *V, int64_t n_maps, int64_t n_vis);
// uvwMap table: 0, -1 (open/IO), -3 (no count on the first line), -4 (a row
// outside the int32 range or not finite), -5 (a count <= 0 in the table).
int uvw_map_weights(const char *path, const double *uvw, int64_t n, float *w, int64_t *n_cells,
                    int64_t *n_skipped, int64_t *n_found);
int write_uvw_map_txt(const char *path, const double *uvw, const int32_t *counts, int64_t n);
int read_vis_bin(const char *path, float *V, int64_t cap, int64_t *n_maps, int64_t *n_vis);
}  // namespace vrs
