// Learned correction of the split IC(0)-pattern factor (filled in by the
// message-passing network inference).  Until weights are loaded the learned
// preconditioner is unavailable and the solve reports an error (the caller
// falls back to FE_PC_IC0 / FE_PC_JACOBI explicitly).
#include "fe_hip_internal.h"

int feLearnedCorrect(fe_hip_ctx *ctx)
{
  FE_REQUIRE(ctx, false, "no learned preconditioner weights are loaded");
}
