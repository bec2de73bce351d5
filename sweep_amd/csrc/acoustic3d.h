// acoustic3d path: driver entry points.
#pragma once

#include "common.h"

namespace swp {
namespace a3d {

bool available();
int64_t workspace_bytes(const swp_problem &p, int kind);
int64_t tape_bytes(const swp_problem &p);
void forward(const swp_problem &p, float *record);
void backward(const swp_problem &p, const swp_backward_io &io);

}  // namespace a3d
}  // namespace swp
