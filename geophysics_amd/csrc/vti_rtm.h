// vti_rtm.h — reverse-sweep driver and imaging kernels shared by RTM and
// the adjoint / gradient path.
#pragma once

#include <functional>

#include "vti_internal.h"

namespace vti {

// Callbacks of a reverse sweep over the forward states X_0 .. X_nt
// (X_k = state before step k) held in one working state.
struct SweepOps {
    std::function<int()> zero;            // working state <- X_0
    std::function<int(int)> store;        // checkpoint slot i <- working state
    std::function<int(int)> restore;      // working state <- checkpoint slot i
    std::function<int(int)> step;         // working X_it -> X_{it+1} (recomputation)
    std::function<int(int)> first_step;   // the same in the first sweep (record, illumination)
    std::function<int()> after_first;     // between the first sweep and the first consume
    std::function<int(int)> consume;      // working state holds X_k, k = nt .. 1
    // optional: X_k read from checkpoint slot i without restoring it (a
    // checkpointed state is consumed once, then its slot is freed)
    std::function<int(int, int)> consume_slot;
};

// Runs the first sweep, then calls consume(k) for k = nt, nt-1, ..., 1 in
// that order, recomputing states from at most nslot full-state checkpoints
// with the binomial schedule.  *fwd_steps: forward steps taken in total.
int reverse_sweep(int nt, int nslot, const SweepOps &op, long long *fwd_steps);

hipError_t launch_inject_rec(vti_ctx *c, const VtiState &s, const float *row, int iz);
hipError_t launch_image(vti_ctx *c, const float *sP, const float *gP, float *img);
hipError_t launch_illum(vti_ctx *c, const float *P, const float *Q, float *illum, int clamp);
hipError_t launch_divide(vti_ctx *c, float *img, const float *illum, size_t n, int guard);
void laplace_adj(int nz, int nx, int ny, const float *in, float *out);

}  // namespace vti
