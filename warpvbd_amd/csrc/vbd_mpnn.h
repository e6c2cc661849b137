// vbd_mpnn.h -- hooks between the step driver and the warm-start network.
#pragma once
#include "vbd_internal.h"

namespace vbd {
// Before the step's init overwrites tgt: keep the previous step's converged
// correction (uprev - tgt) when the solver still holds it.  defer: leave the
// pass that reads it to mpnn_init (which runs it, or folds it into the init).
int mpnn_before_init(vbd_mpnn *m, double dt, bool defer = false);
// In place of the step's plain init (single-body launches): decides this
// step's guess; a step that takes the history guess without a comparison gets
// one fused pass (correction, inertia target, guess), any other the deferred
// correction pass and the plain init.  k_step_init's arguments.
// *done: the fused pass ran (else the caller runs the plain init).
int mpnn_init(vbd_mpnn *m, int n, float4 *u, const double4 *uprev, const float4 *vel, double4 *tgt,
              const float *mass, const uint8_t *active, float dt, float4 g, Ctrl *ctrl, bool *done);
// After the plain initial guess: evaluate the network, keep its guess when its
// first-sweep residual is smaller.
int mpnn_apply_guess(vbd_mpnn *m, double dt);
int mpnn_used_guess(vbd_mpnn *m, int *used);
}  // namespace vbd
