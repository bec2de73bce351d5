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
    int64_t n;   // cells per shot
};

void launch_points(const PointArgs &a, int B);

// Whether equation 0 (acoustic2d), 1 (acoustic3d), 2 (elastic2d) is built.
bool eq_available(int eq);

}  // namespace swp
