// Wavefront path tracer (queues of ray, shadow and material work).
#include <hip/hip_runtime.h>

#include "internal.h"

namespace clr {

int renderWavefront(clr_ctx* ctx, const clr_scene_dev&, const clr_render_params&, uint32_t*, clr_float3*,
                    hipStream_t) {
    return fail(ctx, CLR_ERR_UNSUPPORTED, "wavefront renderer not available in this build");
}

}  // namespace clr
