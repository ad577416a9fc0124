// Wavefront integrator (placeholder until the staged kernels land).
#include "runtime.h"

namespace gy {
int WavefrontPass(Ctx* c, Scene* sc, const PathParams& P, const gy_render_params& p, int s0, int n_s) {
    (void)c; (void)sc; (void)P; (void)p; (void)s0; (void)n_s;
    return Fail(GY_ERR_UNSUPPORTED, "wavefront integrator not built yet");
}
void WavefrontFree(Scene* sc) { (void)sc; }
}  // namespace gy
