// Per-render launch autotuner: times short pilot renders of the same scene,
// view and depth over the launch choices and keeps the fastest.  The estimator
// (and so the image) does not depend on any of them: megakernel and wavefront
// films are equal bit for bit, and the choices only move work between kernels.
//
// Choices: integrator form (megakernel / wavefront), compared first; then,
// within the faster form, the megakernel workgroup size (64, 128, 256 lanes;
// its LDS traversal stack follows the BVH depth) or the wavefront's refill
// threshold for its persistent trace waves (automatic: 48 busy lanes; or 16,
// or 32).  No counterpart in the reference, whose TiledIntegrator fixes 32x32
// tiles (integrator/TiledIntegrator.cpp:13).
#include <algorithm>
#include <chrono>
#include <cstring>

#include "gy_internal.h"

extern "C" int gy_render_stats_get(gy_context* ctx, gy_render_stats* out);
extern "C" int gy_scene_context(gy_scene* scene, gy_context** ctx);

extern "C" int gy_autotune(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* base, int32_t pilot_spp,
                           gy_render_params* best, double* best_ms_per_mpath) {
    using namespace gy;
    if (!scene || !cam || !base || !best) return Fail(GY_ERR_INVALID, "gy_autotune: null argument");
    gy_context* ctx = nullptr;
    int rc = gy_scene_context(scene, &ctx);
    if (rc != GY_OK) return rc;
    gy_render_params pilot = *base;
    // The pilot is the render itself up to 2^23 paths, else the fewest samples
    // per pixel that reach 2^23 paths: fixed per-launch costs weigh on small
    // pilots and would favour the megakernel.
    const int64_t pixels = (int64_t)cam->width * cam->height;
    int32_t auto_spp = base->spp;
    if (pixels * base->spp > (1ll << 23)) auto_spp = (int32_t)std::min<int64_t>(base->spp, ((1ll << 23) + pixels - 1) / pixels);
    pilot.spp = pilot_spp > 0 ? std::min(pilot_spp, base->spp) : std::max(1, auto_spp);
    pilot.samples_per_pass = pilot.spp;
    struct Cand {
        int mode, block, refill;
    };
    const double mpaths = (double)pixels * pilot.spp / 1e6;
    double best_t = 1e300;
    auto trial = [&](const Cand& k) -> int {
        gy_render_params p = pilot;
        p.mode = k.mode;
        p.block_size = k.block;
        p.refill_threshold = k.refill;
        double t = 1e300;
        for (int rep = 0; rep < 2; ++rep) {  // the first run allocates
            int r = gy_render(scene, cam, &p, nullptr, nullptr);
            if (r != GY_OK) return r;
            gy_render_stats st;
            r = gy_render_stats_get(ctx, &st);
            if (r != GY_OK) return r;
            t = st.total_ms;
        }
        if (t < best_t) {
            best_t = t;
            *best = *base;
            best->mode = k.mode;
            best->block_size = k.block;
            best->refill_threshold = k.refill;
        }
        return GY_OK;
    };
    // the two forms first, then the other launch shapes of the faster one
    const Cand first[] = {{GY_RENDER_WAVEFRONT, 0, 0}, {GY_RENDER_MEGAKERNEL, 128, 0}};
    for (const Cand& k : first)
        if ((rc = trial(k)) != GY_OK) return rc;
    const Cand more_wf[] = {{GY_RENDER_WAVEFRONT, 0, 16}, {GY_RENDER_WAVEFRONT, 0, 32}};
    const Cand more_mk[] = {{GY_RENDER_MEGAKERNEL, 64, 0}, {GY_RENDER_MEGAKERNEL, 256, 0}};
    const Cand(&more)[2] = best->mode == GY_RENDER_WAVEFRONT ? more_wf : more_mk;
    for (const Cand& k : more)
        if ((rc = trial(k)) != GY_OK) return rc;
    if (best_ms_per_mpath) *best_ms_per_mpath = best_t / mpaths;
    return GY_OK;
}
