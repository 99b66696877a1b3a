// MLP radiance cache (placeholder until the cache kernels land).
#include "render_modes.h"

namespace snkpt {

struct CacheState {};
void cache_state_free(CacheState* c) { delete c; }

snkpt_status cache_pass(snkpt_ctx* ctx, const CamDev&, const snkpt_settings*, uint64_t, uint32_t, hipStream_t) {
    return fail(ctx, SNKPT_ERR_NOT_BUILT, "cache mode: not available in this build");
}

}  // namespace snkpt
