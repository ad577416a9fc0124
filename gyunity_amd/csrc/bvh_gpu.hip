// GPU BVH build (placeholder until the binned-SAH kernels land): host build.
#include "runtime.h"

namespace gy {
int BuildBVHDevice(Ctx* c, const float* pos, int32_t n_tri, int max_leaf, BVHBuild* out) {
    (void)c;
    BuildBVHHost(pos, n_tri, max_leaf, out);
    return GY_OK;
}
}  // namespace gy
