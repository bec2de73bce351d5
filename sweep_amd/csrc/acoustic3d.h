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
    int px;                             // row pitch of the core's arrays (nx rounded up to 32)
    int64_t sxy;                        // px * ny (z stride)
    int64_t n;                          // cells per shot in the core's arrays (nz * sxy)
    int64_t nm;                         // cells per shot in the caller's arrays (nz * ny * nx)
    int64_t mstride;                    // model stride per shot (0: shared model)
    int64_t ns;                         // strip cells per shot per step (six faces)
    int sbf16;                          // strips stored as bfloat16 (SWP_FLAG_BF16_STRIPS)
    float dt, dx, dy, dz;
    // 1/h^2, 1/h and 1/dt^2 with h itself (fd::Scale): the forward and
    // reconstruction stencils, the profile derivative and the u_tt of imaging
    // scale through fd::scaled; the adjoint's stencil sums multiply by .inv
    fd::Scale inv_dx2, inv_dy2, inv_dz2, inv_dx, inv_dy, inv_dz;
    float dt2;
    fd::Scale inv_dt2;
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
    bool zm;             // z-marching kernels (fixed orders 2..8)
    bool shell_percell;  // with zm: the CPML shell on the per-cell kernels
    int tz;              // z-segment height of the z-marching kernels
};

// A box of cells [x0, x1) x [y0, y1) x [z0, z1).
struct Box {
    int x0, x1, y0, y1, z0, z1;
};

// Strip layout of one shot and step (ns cells): faces z-lo, z-hi as
// (W, nyp, nxp), y-lo, y-hi as (nzp, W, nxp), x-lo, x-hi as (nzp, nyp, W);
// a cell on an edge or corner is stored in each face that holds it.
template <bool bf>
__device__ __forceinline__ void save_strips_as(const Geo &g, float *__restrict__ s, int ix, int iy, int iz, float v)
{
    const int64_t W = g.W, nxp = g.nxp, nyp = g.nyp, nzp = g.nzp;
    const int64_t x = ix - g.X0, y = iy - g.Y0, z = iz - g.Z0;
    const int64_t fz = W * nyp * nxp, fy = nzp * W * nxp, fx = nzp * nyp * W;
    if (z < W) strip_put(s, (z * nyp + y) * nxp + x, v, bf);
    if (iz >= g.Z1 - g.W) strip_put(s, fz + ((iz - (g.Z1 - g.W)) * nyp + y) * nxp + x, v, bf);
    if (y < W) strip_put(s, 2 * fz + (z * W + y) * nxp + x, v, bf);
    if (iy >= g.Y1 - g.W) strip_put(s, 2 * fz + fy + (z * W + (iy - (g.Y1 - g.W))) * nxp + x, v, bf);
    if (x < W) strip_put(s, 2 * fz + 2 * fy + (z * nyp + y) * W + x, v, bf);
    if (ix >= g.X1 - g.W) strip_put(s, 2 * fz + 2 * fy + fx + (z * nyp + y) * W + (ix - (g.X1 - g.W)), v, bf);
}

// The stencil kernels save float32 (with bf16 strips the driver packs each
// step's strips afterwards, pack_bf16).
__device__ __forceinline__ void save_strips(const Geo &g, float *__restrict__ s, int ix, int iy, int iz, float v)
{
    if (ix < g.X0 || ix >= g.X1 || iy < g.Y0 || iy >= g.Y1 || iz < g.Z0 || iz >= g.Z1) return;
    save_strips_as<false>(g, s, ix, iy, iz, v);
}

__device__ __forceinline__ float load_strip(const Geo &g, const float *__restrict__ s, int ix, int iy, int iz)
{
    const int64_t W = g.W, nxp = g.nxp, nyp = g.nyp, nzp = g.nzp;
    const int64_t x = ix - g.X0, y = iy - g.Y0, z = iz - g.Z0;
    const int64_t fz = W * nyp * nxp, fy = nzp * W * nxp, fx = nzp * nyp * W;
    const bool bf = g.sbf16;
    if (z < W) return strip_get(s, (z * nyp + y) * nxp + x, bf);
    if (iz >= g.Z1 - g.W) return strip_get(s, fz + ((iz - (g.Z1 - g.W)) * nyp + y) * nxp + x, bf);
    if (y < W) return strip_get(s, 2 * fz + (z * W + y) * nxp + x, bf);
    if (iy >= g.Y1 - g.W) return strip_get(s, 2 * fz + fy + (z * W + (iy - (g.Y1 - g.W))) * nxp + x, bf);
    if (x < W) return strip_get(s, 2 * fz + 2 * fy + (z * nyp + y) * W + x, bf);
    return strip_get(s, 2 * fz + 2 * fy + fx + (z * nyp + y) * W + (ix - (g.X1 - g.W)), bf);
}

bool available();
Geo make_geo(const swp_problem &p);
Ctx make_ctx(const swp_problem &p);
int64_t workspace_bytes(const swp_problem &p, int kind);
int64_t tape_bytes(const swp_problem &p);
void forward(const swp_problem &p, float *record);
void backward(const swp_problem &p, const swp_backward_io &io);

}  // namespace a3d
}  // namespace swp
