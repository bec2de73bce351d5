// Wave64 z-marching kernels of the 2-D acoustic path (see acoustic2d_zm.hip).
#pragma once

#include "acoustic2d.h"

namespace swp {
namespace a2d {

// Smallest launch (shots x runtime-grid cells) that runs on the z-marching
// kernels by default; smaller ones run one thread per cell.
constexpr int64_t kZmMinCells = int64_t(4) << 20;

ZmPlan make_plan(const Geo &g, int tz, int tz_band);
bool zm_supported(const Geo &g);
void launch_fwd_zm(const Ctx &x, const ZmPlan &p, const StepFields &f, float *u_this, float *strip);
void launch_adj_zm(const Ctx &x, const ZmPlan &p, const AdjFields &f, const float *img, float *grad);
void launch_recon_zm(const Ctx &x, const float *vp, const float *u_prev, const float *u_now, float *u_next,
                     const float *adj, float *grad, int x_lo, int x_hi, int z_lo, int z_hi, int tz);

}  // namespace a2d
}  // namespace swp
