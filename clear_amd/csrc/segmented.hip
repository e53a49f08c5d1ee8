// Segmented LBVH build: many independent scenes in one set of launches.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace clr {

size_t lbvhSegScratchBytes(uint32_t, uint32_t, uint32_t) { return 0; }

int lbvhBuildSegmented(clr_ctx* ctx, const float*, const uint32_t*, uint32_t, uint32_t,
                       clr_lbvh_inode*, clr_lbvh_leaf*, hipStream_t) {
    return fail(ctx, CLR_ERR_UNSUPPORTED, "segmented build is not available in this build");
}

}  // namespace clr
