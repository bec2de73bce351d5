// z-marching kernels of the 3-D acoustic path (not enabled yet: the per-cell
// kernels of acoustic3d.hip run every order).
#include "acoustic3d_zm.h"

namespace swp {
namespace a3d {

bool zm_supported(const Geo &) { return false; }
void launch_fwd_zm(const Ctx &, const StepFields &, float *, float *) { fail("acoustic3d: z-march disabled"); }
void launch_adj_zm(const Ctx &, const AdjFields &, const float *, float *) { fail("acoustic3d: z-march disabled"); }
void launch_recon_zm(const Ctx &, const float *, const float *, const float *, float *, const float *, float *, int,
                     int, int, int, int, int)
{
    fail("acoustic3d: z-march disabled");
}

}  // namespace a3d
}  // namespace swp
