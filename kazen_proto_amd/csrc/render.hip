// Wavefront path integrator (placeholder until the integrator lands).
#include "kp_ctx.h"

namespace kp {
void render_state_free(kp_ctx*) {}
}  // namespace kp
