// vbd_mpnn.h -- hooks between the step driver and the warm-start network.
#pragma once
#include "vbd_internal.h"

namespace vbd {
// Before the step's init overwrites tgt: keep the previous step's converged
// correction (uprev - tgt) when the solver still holds it.
int mpnn_before_init(vbd_mpnn *m, double dt);
// After the plain initial guess: evaluate the network, keep its guess when its
// first-sweep residual is smaller.
int mpnn_apply_guess(vbd_mpnn *m, double dt);
int mpnn_used_guess(vbd_mpnn *m, int *used);
}  // namespace vbd
