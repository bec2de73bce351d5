// Shared device-side declarations of libswp_core.
#pragma once

#include "core.h"

namespace swp {

// Arguments of the per-shot point kernel: an optional scatter-add of
// w[b, s, it] (times sign) into u_add at add_loc, then an optional gather of
// u_get at get_loc into out[b, r, it].  Coordinates are (x[, y], z).
struct PointArgs {
    float *u_add;
    const float *add_w;
    const int32_t *add_loc;
    int nadd;
    float sign;
    const float *u_get;
    float *get_out;
    const int32_t *get_loc;
    int nget;
    int get_accumulate;
    int it, nt, ndim;
    int nz, ny, nx;
    int px;      // row pitch of u_add / u_get (0: nx)
    int64_t n;   // cells per shot
};

void launch_points(const PointArgs &a, int B);

// Row-pitched copies between the caller's arrays and the core's: rows rows
// of width floats, dst row r at dst + r * dpitch, src row r at src + r * spitch.
// pitch_copy writes dst (and zeroes dst's columns [width, dpitch)),
// pitch_add adds src into dst.
void pitch_copy(float *dst, int64_t dpitch, const float *src, int64_t spitch, int64_t rows, int width);
void pitch_add(float *dst, int64_t dpitch, const float *src, int64_t spitch, int64_t rows, int width);

// Whether equation 0 (acoustic2d), 1 (acoustic3d), 2 (elastic2d) is built.
bool eq_available(int eq);

}  // namespace swp
