// acoustic3d path on gfx950 (not built yet: every entry fails loudly).
#include "acoustic3d.h"

namespace swp {
namespace a3d {

bool available() { return false; }
int64_t workspace_bytes(const swp_problem &, int) { fail("acoustic3d is not available in this core"); }
int64_t tape_bytes(const swp_problem &) { fail("acoustic3d is not available in this core"); }
void forward(const swp_problem &, float *) { fail("acoustic3d is not available in this core"); }
void backward(const swp_problem &, const swp_backward_io &) { fail("acoustic3d is not available in this core"); }

}  // namespace a3d
}  // namespace swp
