// Per-render launch autotuner: times short pilot renders of the same scene,
// view and depth over the launch choices and keeps the fastest.  The estimator
// (and so the image) does not depend on any of them: megakernel and wavefront
// films are equal bit for bit, and the choices only move work between kernels.
//
// Choices: integrator form (megakernel / wavefront), compared first; then,
// within the faster form, the megakernel workgroup size (64, 128, 256 lanes;
// its LDS traversal stack follows the BVH depth) or, for the wavefront, the
// refill threshold of its persistent trace waves (automatic: 48 busy lanes; or
// 16, or 32) and then the order of its closest-hit queue (as generated, or
// partitioned by ray-direction octant over the whole queue or within runs of
// 1024 entries).  The choice is kept per scene signature (scene id and
// revision of its device buffers, camera, depth, roulette, sampler, samples
// per pixel, pilot size): the same question asked again is answered without
// timing.  No counterpart in the reference, whose TiledIntegrator fixes 32x32
// tiles (integrator/TiledIntegrator.cpp:13).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <mutex>

#include "gy_internal.h"

extern "C" int gy_render_stats_get(gy_context* ctx, gy_render_stats* out);
extern "C" int gy_scene_context(gy_scene* scene, gy_context** ctx);

namespace {

struct TuneKey {
    uint64_t uid, revision;
    gy_camera_desc cam;
    int32_t spp, max_depth, rr_depth, sampler, pilot_spp;
    bool operator==(const TuneKey& o) const {
        return uid == o.uid && revision == o.revision && std::memcmp(&cam, &o.cam, sizeof cam) == 0 && spp == o.spp &&
               max_depth == o.max_depth && rr_depth == o.rr_depth && sampler == o.sampler && pilot_spp == o.pilot_spp;
    }
};
struct TuneChoice {
    int32_t mode, block_size, refill_threshold, ray_sort;
    double ms_per_mpath;
};

// Most recent choices first; a bounded list (a process tunes few views).
constexpr size_t kCacheEntries = 1024;
std::mutex g_mu;
std::deque<std::pair<TuneKey, TuneChoice>> g_cache;

bool Lookup(const TuneKey& k, TuneChoice* out) {
    std::lock_guard<std::mutex> lock(g_mu);
    for (const auto& e : g_cache)
        if (e.first == k) {
            *out = e.second;
            return true;
        }
    return false;
}
void Store(const TuneKey& k, const TuneChoice& c) {
    std::lock_guard<std::mutex> lock(g_mu);
    for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
        if (it->first == k) {
            g_cache.erase(it);
            break;
        }
    g_cache.emplace_front(k, c);
    if (g_cache.size() > kCacheEntries) g_cache.pop_back();
}

}  // namespace

extern "C" int gy_autotune_clear_cache(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_cache.clear();
    return GY_OK;
}

extern "C" int gy_autotune(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* base, int32_t pilot_spp,
                           gy_render_params* best, double* best_ms_per_mpath) {
    using namespace gy;
    if (!scene || !cam || !base || !best) return Fail(GY_ERR_INVALID, "gy_autotune: null argument");
    gy_context* ctx = nullptr;
    int rc = gy_scene_context(scene, &ctx);
    if (rc != GY_OK) return rc;
    TuneKey key;
    std::memset(&key, 0, sizeof key);  // padding takes part in the comparison
    if ((rc = SceneSignature(scene, &key.uid, &key.revision)) != GY_OK) return rc;
    key.cam = *cam;
    key.spp = base->spp;
    key.max_depth = base->max_depth;
    key.rr_depth = base->rr_depth;
    key.sampler = base->sampler;
    key.pilot_spp = pilot_spp;
    TuneChoice hit;
    if (Lookup(key, &hit)) {
        *best = *base;
        best->mode = hit.mode;
        best->block_size = hit.block_size;
        best->refill_threshold = hit.refill_threshold;
        best->ray_sort = hit.ray_sort;
        if (best_ms_per_mpath) *best_ms_per_mpath = hit.ms_per_mpath;
        return GY_OK;
    }
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
        int mode, block, refill, sort;
    };
    const double mpaths = (double)pixels * pilot.spp / 1e6;
    double best_t = 1e300;
    Cand won = {GY_RENDER_WAVEFRONT, 0, 0, GY_RAY_SORT_NONE};
    auto trial = [&](const Cand& k) -> int {
        gy_render_params p = pilot;
        p.mode = k.mode;
        p.block_size = k.block;
        p.refill_threshold = k.refill;
        p.ray_sort = k.sort;
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
            won = k;
        }
        return GY_OK;
    };
    // the two forms first, then the other launch shapes of the faster one
    const Cand first[] = {{GY_RENDER_WAVEFRONT, 0, 0, GY_RAY_SORT_NONE}, {GY_RENDER_MEGAKERNEL, 128, 0, GY_RAY_SORT_NONE}};
    for (const Cand& k : first)
        if ((rc = trial(k)) != GY_OK) return rc;
    if (won.mode == GY_RENDER_WAVEFRONT) {
        const Cand refills[] = {{GY_RENDER_WAVEFRONT, 0, 16, GY_RAY_SORT_NONE}, {GY_RENDER_WAVEFRONT, 0, 32, GY_RAY_SORT_NONE}};
        for (const Cand& k : refills)
            if ((rc = trial(k)) != GY_OK) return rc;
        // the queue orders at the best refill threshold
        const int refill = won.refill;
        if ((rc = trial({GY_RENDER_WAVEFRONT, 0, refill, GY_RAY_SORT_OCTANT})) != GY_OK) return rc;
        if ((rc = trial({GY_RENDER_WAVEFRONT, 0, refill, GY_RAY_SORT_OCTANT_LOCAL})) != GY_OK) return rc;
    } else {
        const Cand blocks[] = {{GY_RENDER_MEGAKERNEL, 64, 0, GY_RAY_SORT_NONE}, {GY_RENDER_MEGAKERNEL, 256, 0, GY_RAY_SORT_NONE}};
        for (const Cand& k : blocks)
            if ((rc = trial(k)) != GY_OK) return rc;
    }
    *best = *base;
    best->mode = won.mode;
    best->block_size = won.block;
    best->refill_threshold = won.refill;
    best->ray_sort = won.sort;
    Store(key, {won.mode, won.block, won.refill, won.sort, best_t / mpaths});
    if (best_ms_per_mpath) *best_ms_per_mpath = best_t / mpaths;
    return GY_OK;
}
