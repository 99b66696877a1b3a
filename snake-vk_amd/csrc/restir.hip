// ReSTIR DI (placeholder until the resampling kernels land).
#include "render_modes.h"

namespace snkpt {

struct RestirState {};
void restir_state_free(RestirState* r) { delete r; }

snkpt_status restir_pass(snkpt_ctx* ctx, const snkpt_camera*, const CamDev&, const snkpt_settings*, uint64_t,
                         uint32_t, hipStream_t) {
    return fail(ctx, SNKPT_ERR_NOT_BUILT, "ReSTIR mode: not available in this build");
}

}  // namespace snkpt
