// z-marching kernels of the 3-D acoustic path (see acoustic3d_zm.hip).
#pragma once

#include "acoustic3d.h"

namespace swp {
namespace a3d {

bool zm_supported(const Geo &g);
void launch_fwd_zm(const Ctx &x, const StepFields &f, float *u_this, float *strip);
void launch_adj_zm(const Ctx &x, const AdjFields &f, const float *img, float *grad);
void launch_recon_zm(const Ctx &x, const float *vp, const float *u_prev, const float *u_now, float *u_next,
                     const float *adj, float *grad, int x0, int x1, int y0, int y1, int z0, int z1);

}  // namespace a3d
}  // namespace swp
