// elastic2d path on gfx950 (not built yet: every entry fails loudly).
#include "elastic2d.h"

namespace swp {
namespace e2d {

bool available() { return false; }
int64_t workspace_bytes(const swp_problem &, int) { fail("elastic2d is not available in this core"); }
int64_t tape_bytes(const swp_problem &) { fail("elastic2d is not available in this core"); }
void forward(const swp_problem &, float *) { fail("elastic2d is not available in this core"); }
void backward(const swp_problem &, const swp_backward_io &) { fail("elastic2d is not available in this core"); }

}  // namespace e2d
}  // namespace swp
