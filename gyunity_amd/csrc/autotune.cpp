// Per-render launch autotuner: times short pilot renders of the same scene,
// view and depth over the launch choices and keeps the fastest.  The estimator
// (and so the image) does not depend on any of them: megakernel and wavefront
// films are equal bit for bit, and the choices only move work between kernels.
//
// Choices: integrator form (megakernel / wavefront); megakernel workgroup size
// (64, 128, 256 lanes; its LDS traversal stack follows the BVH depth); the
// wavefront's refill threshold for its persistent trace waves (automatic: 1
// for closest-hit rays, 16 for shadow and MIS rays; or 1, or 16 for all).  No counterpart in the reference, whose TiledIntegrator fixes 32x32
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
    pilot.spp = pilot_spp > 0 ? std::min(pilot_spp, base->spp) : std::min(4, base->spp);
    pilot.samples_per_pass = pilot.spp;
    struct Cand {
        int mode, block, refill;
    };
    const Cand cands[] = {{GY_RENDER_MEGAKERNEL, 64, 0}, {GY_RENDER_MEGAKERNEL, 128, 0}, {GY_RENDER_MEGAKERNEL, 256, 0},
                          {GY_RENDER_WAVEFRONT, 0, 0},   {GY_RENDER_WAVEFRONT, 0, 1},   {GY_RENDER_WAVEFRONT, 0, 16}};
    double best_t = 1e300;
    const double mpaths = (double)cam->width * cam->height * pilot.spp / 1e6;
    for (const Cand& k : cands) {
        gy_render_params p = pilot;
        p.mode = k.mode;
        p.block_size = k.block;
        p.refill_threshold = k.refill;
        double t = 1e300;
        for (int rep = 0; rep < 2; ++rep) {  // the first run allocates
            rc = gy_render(scene, cam, &p, nullptr, nullptr);
            if (rc != GY_OK) return rc;
            gy_render_stats st;
            rc = gy_render_stats_get(ctx, &st);
            if (rc != GY_OK) return rc;
            t = st.total_ms;
        }
        if (t < best_t) {
            best_t = t;
            *best = *base;
            best->mode = k.mode;
            best->block_size = k.block;
            best->refill_threshold = k.refill;
        }
    }
    if (best_ms_per_mpath) *best_ms_per_mpath = best_t / mpaths;
    return GY_OK;
}
