// 2-D isotropic elastic path (velocity-stress, staggered grid, cpmls CPML):
// geometry and driver entry points.
#pragma once

#include "common.h"

namespace swp {
namespace e2d {

constexpr int kMaxM = 16;

// Field indices of one state (forward or adjoint), per shot on the runtime grid.
enum Field { VX = 0, VZ, SXX, SZZ, SXZ, M_VXX, M_VZZ, M_VXZ, M_VZX, M_SXXX, M_SZZZ, M_SXZX, M_SXZZ, NSTATE };

struct Geo {
    int nz, nx, B, M, nt;
    int pzl, pzh, pxl, pxh;
    int X0, X1, Z0, Z1;        // physical box
    int64_t n;                 // cells per shot
    int64_t mstride;           // model stride per shot (0: shared models)
    float dt, dx, dz, inv_dx, inv_dz;
};

// Staggered first-derivative weights c_1..c_M (runtime orders > 8).
struct SCoef {
    float c[kMaxM];
};

struct Prof {
    const float *az, *bz, *azh, *bzh, *ax, *bx, *axh, *bxh;
};

// Device pointers of one state's 13 fields (already offset to shot 0).
struct State {
    float *f[NSTATE];
};

bool available();
int64_t workspace_bytes(const swp_problem &p, int kind);
int64_t tape_bytes(const swp_problem &p);
void forward(const swp_problem &p, float *record);
void backward(const swp_problem &p, const swp_backward_io &io);

}  // namespace e2d
}  // namespace swp
