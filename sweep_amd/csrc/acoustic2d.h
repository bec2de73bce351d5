// 2-D acoustic path: geometry, coefficient block and driver entry points.
#pragma once

#include "common.h"
#include "fd.h"

namespace swp {
namespace a2d {

constexpr int kMaxM = 16;

// Runtime-grid geometry of one call, passed to every kernel by value.
struct Geo {
    int nz, nx, B, M, nt;
    int pzl, pzh, pxl, pxh;    // PML pad per face
    int X0, X1, Z0, Z1;        // physical box [pad_lo + M, n - pad_hi - M)
    int abcn;
    int W;                     // boundary strip width M + 1
    int nxp, nzp;              // physical box extents
    int px;                    // row pitch of the core's arrays (nx rounded up to 32)
    int64_t n;                 // cells per shot in the core's arrays (nz * px)
    int64_t nm;                // cells per shot in the caller's arrays (nz * nx)
    int64_t mstride;           // model stride per shot in the core's copy (0: shared model)
    int64_t ns;                // strip cells per shot per step
    int sbf16;                 // strips stored as bfloat16 (SWP_FLAG_BF16_STRIPS)
    float dt, dx, dz;
    // 1/h^2, 1/h and 1/dt^2 with h itself (fd::Scale): the forward and
    // reconstruction stencils and the u_tt of imaging scale through
    // fd::scaled; the adjoint multiplies by .inv
    fd::Scale inv_dx2, inv_dz2, inv_dx, inv_dz;
    float dt2;
    fd::Scale inv_dt2;
};

// FD weights: the fixed-order forward constants and the caller's tables.
struct Coef {
    float slap[4];
    float slap_c0;
    float sgrad[4];
    float lap[kMaxM + 1];
    float grad[kMaxM + 1];
};

struct Prof {
    const float *az, *bz, *dbz, *ax, *bx, *dbx;
};

struct StepFields {
    const float *vp;
    const float *u_prev, *u_now;
    float *u_next;
    const float *psix, *psiz;
    float *psixn, *psizn;
    float *zetax, *zetaz;
};

struct AdjFields {
    const float *vp;
    const float *l_prev, *l_now;
    float *l_next;
    const float *psix, *psiz, *zetax, *zetaz;
    float *psixo, *psizo, *zetaxo, *zetazo;
};

// A rectangle of blocks of the z-marching kernels: 64-column strips
// [s0, s1) times the rows [z0, z1) cut into segments of tz rows.
struct Rect {
    int s0, s1, z0, z1, tz;
};

// Up to four rectangles enumerated back to back by blockIdx.x.
struct BlockSet {
    Rect r[4];
    int x0;   // first column of strip 0 (0: strips start on 128-byte lines)
};

// Tiling of the computed box [M, n-M) into 64-column strips x row
// segments, split into the interior blocks (no CPML term, no strip cell, no
// db leak cell reaches them; tall segments) and the band blocks around them
// (short segments: more waves for the heavier, lower-occupancy band kernels).
struct ZmPlan {
    int tz, tz_band, nstrip;
    BlockSet inner, band;
    int ninner, nband;
};

struct Ctx {
    Geo g;
    Coef c;
    Prof pr;
    bool zm;        // z-marching kernels (fixed orders 2..8)
    ZmPlan plan;
    int tz_recon;   // segment height of the reconstruction kernel
};

Geo make_geo(const swp_problem &p);
// Coefficient block from the caller's (M+1) tables plus the fixed-order
// constants; shared with the 3-D path.
Coef make_coef(const swp_problem &p, const float *lap_host, const float *grad_host);
Ctx make_ctx(const swp_problem &p);
int64_t workspace_bytes(const swp_problem &p, int kind);
int64_t tape_bytes(const swp_problem &p);
void forward(const Ctx &x, const swp_problem &p, float *record);
void backward(const Ctx &x, const swp_problem &p, const swp_backward_io &io);

}  // namespace a2d
}  // namespace swp
