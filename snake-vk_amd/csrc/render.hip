// Path-trace pass (placeholder until the wavefront kernels land).
#include "snkpt_host.h"

namespace snkpt {

struct RenderState {};
struct CacheState {};

snkpt_status render(snkpt_ctx* ctx, const snkpt_camera*, const snkpt_settings*, void*, uint32_t, uint32_t, uint32_t,
                    hipStream_t) {
    return fail(ctx, SNKPT_ERR_NOT_BUILT, "render: not available in this build");
}
snkpt_status oracle_accumulate(snkpt_ctx* ctx, const snkpt_camera*, const snkpt_settings*, double*, uint32_t,
                               uint32_t, uint32_t, hipStream_t) {
    return fail(ctx, SNKPT_ERR_NOT_BUILT, "oracle_accumulate: not available in this build");
}
void render_state_free(RenderState* r) { delete r; }
void cache_state_free(CacheState* c) { delete c; }

}  // namespace snkpt
