// Binomial checkpoint reversal of a time loop (host code).
//
// A backward over nt steps needs the forward state at every step in reverse
// order.  With s snapshot slots (slot 0 holding the state before step 0),
// Griewank's bound says l steps can be reversed with r forward re-runs of
// each step when C(s + r, s) >= l (A. Griewank, "Achieving logarithmic
// growth of temporal and spatial complexity in reverse automatic
// differentiation", Optim. Methods Softw. 1, 1992): the first snapshot after
// t0 goes m <= C(s + r - 1, s) steps in, the right part (at most
// C(s - 1 + r, s - 1) steps) is reversed with s - 1 slots, then the left
// part with s.  The forward pass stores the snapshots of the rightmost path
// of that recursion as it goes (spine()), so the backward starts there.
#pragma once

#include <cstdint>
#include <vector>

namespace swp {

enum RevCode : int32_t {
    REV_STEP = 0,      // advance the working state one step (it -> it + 1), nothing kept
    REV_SAVE = 1,      // working state -> slot
    REV_LOAD = 2,      // slot -> working state
    REV_REVERSE = 3,   // working state at it: re-run step it keeping what the adjoint needs, then the adjoint of it
};

struct RevOp {
    int32_t code, arg;   // arg: the step (STEP, REVERSE) or the slot (SAVE, LOAD)
};

struct SpinePt {
    int32_t slot, step;   // the forward pass saves its state before `step` into `slot`
};

// The snapshots the forward pass takes, slot 0 at step 0 first.
std::vector<SpinePt> revolve_spine(int nt, int slots);
// The backward's operations; REVERSE runs for it = nt - 1 down to 0.
std::vector<RevOp> revolve_schedule(int nt, int slots);

}  // namespace swp
