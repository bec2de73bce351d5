// z-marching kernels of the 3-D acoustic path (see acoustic3d_zm.hip).
#pragma once

#include "acoustic3d.h"

namespace swp {
namespace a3d {

// Orders 2..8, and a z segment of tz planes plus its 2M + 1 halo and prefetch
// planes addressable with 32-bit byte offsets (tz is clamped to fit).
// Smallest launch (shots x runtime-grid cells) that runs on the z-marching
// kernels by default; smaller ones run one thread per cell.
constexpr int64_t kZmMinCells3 = int64_t(32) << 20;

bool zm_supported(const Geo &g, int &tz);
// The slabs of the computed box around the lean interior [lo + M + 1,
// hi - M - 1) per axis (the whole box when that interior is empty); returns
// how many of out[0..5] it filled.
int shell_boxes(const Geo &g, Box out[6]);
// Interior z-march launch, plus the z-march shell when `shell` is set.
void launch_fwd_zm(const Ctx &x, const StepFields &f, float *u_this, float *strip, bool shell);
void launch_adj_zm(const Ctx &x, const AdjFields &f, const float *img, float *grad, bool shell);
void launch_recon_zm(const Ctx &x, const float *vp, const float *u_prev, const float *u_now, float *u_next,
                     const float *adj, float *grad, int x0, int x1, int y0, int y1, int z0, int z1);

}  // namespace a3d
}  // namespace swp
