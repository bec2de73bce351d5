// Record wire format and the L2 misfit (records.hip).
#pragma once

#include "core.h"

namespace swp {
namespace rec {

constexpr int kMaxFields = 5;             // elastic: vx, vz, sxx, szz, sxz
constexpr int kMisfitPartials = SWP_MISFIT_PARTIALS;

// raw (nf, B, nrec, nt) -> canonical (B, nt, nrec, nf), and back.
void to_canonical(const float *raw, float *canon, int nf, int B, int nrec, int nt);
void to_raw(const float *canon, float *raw, int nf, int B, int nrec, int nt);

// Over canonical records: residual = (pred - obs) * mask[b, r] (mask may be
// null; residual may be null), loss[0] = 0.5 * sum residual^2 accumulated in
// float64, block partial sums in partials[0 .. misfit_blocks(n)) summed in a
// fixed order (the result does not depend on scheduling).
int misfit_blocks(int64_t n);
void l2_misfit(const float *pred, const float *obs, const float *mask, int nf, int B, int nrec, int nt,
               float *residual, double *loss, double *partials);
// The same, with the partial sums in a static device array of the library
// (kMisfitPartials doubles per GPU): no workspace argument, no allocation.
void l2_misfit_noalloc(const float *pred, const float *obs, const float *mask, int nf, int B, int nrec, int nt,
                       float *residual, double *loss);

}  // namespace rec
}  // namespace swp
