// 3-D constant-density acoustic path: geometry, fields and driver entry points.
#pragma once

#include "acoustic2d.h"

namespace swp {
namespace a3d {

using a2d::Coef;

// Runtime-grid geometry of one call (x fastest, then y, then z).
struct Geo {
    int nz, ny, nx, B, M, nt;
    int pzl, pzh, pyl, pyh, pxl, pxh;   // PML pad per face
    int X0, X1, Y0, Y1, Z0, Z1;         // physical box [pad_lo + M, n - pad_hi - M) per axis
    int abcn;
    int W;                              // boundary strip width M + 1
    int nxp, nyp, nzp;                  // physical box extents
    int64_t sxy;                        // nx * ny (z stride)
    int64_t n;                          // cells per shot
    int64_t mstride;                    // model stride per shot (0: shared model)
    int64_t ns;                         // strip cells per shot per step (six faces)
    float dt, dx, dy, dz;
    float inv_dx2, inv_dy2, inv_dz2, inv_dx, inv_dy, inv_dz, dt2, inv_dt2;
};

// CPML profiles per axis (a, b, db on the runtime length of the axis).
struct Prof {
    const float *a[3], *b[3], *db[3];   // axis 0 = x, 1 = y, 2 = z
};

struct StepFields {
    const float *vp;
    const float *u_prev, *u_now;
    float *u_next;
    const float *psi[3];
    float *psin[3];
    float *zeta[3];                     // updated in place
};

struct AdjFields {
    const float *vp;
    const float *l_prev, *l_now;
    float *l_next;
    const float *psi[3], *zeta[3];
    float *psio[3], *zetao[3];
};

// Widths of the six strip faces actually imaged/restored (a box thinner than
// two strips images no cell twice).
struct StripW {
    int zl, zh, yl, yh, xl, xh;
};

struct Ctx {
    Geo g;
    Coef c;
    Prof pr;
    bool zm;        // z-marching kernels (fixed orders 2..8)
    int tz;         // z-segment height of the z-marching kernels
};

bool available();
Geo make_geo(const swp_problem &p);
Ctx make_ctx(const swp_problem &p);
int64_t workspace_bytes(const swp_problem &p, int kind);
int64_t tape_bytes(const swp_problem &p);
void forward(const swp_problem &p, float *record);
void backward(const swp_problem &p, const swp_backward_io &io);

}  // namespace a3d
}  // namespace swp
