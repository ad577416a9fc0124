// C ABI of batched rendering: many (scene, view, sample budget) renders traced
// together by the wavefront kernels in one lane space (wavefront.hip,
// BatchRender), recorded once into a hipGraph and replayed.
//
// No counterpart in the reference: its TiledIntegrator renders one image per
// Render() call (integrator/TiledIntegrator.cpp:7-60).  Each entry's film is
// exactly what gy_render gives for that entry alone (same estimator, same
// per-pixel sample order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "runtime.h"

namespace gy {
int MakePathParams(Ctx* c, Scene* sc, const gy_camera_desc& cam, const gy_render_params& p, PathParams* P);
int ValidateRender(const gy_camera_desc* cam, const gy_render_params* p);
void TimersCollect(Ctx* c);

namespace {

void BatchDropGraph(Batch* b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
}

// Table of entries, lane space, films: rebuilt after entries change.
int BatchBuild(Ctx* c, Batch* b, int64_t lanes_req) {
    if (b->entries.empty()) return Fail(GY_ERR_INVALID, "gy_batch_render: empty batch");
    BatchDropGraph(b);
    std::vector<BatchEntryDev> tab(b->entries.size());
    int64_t first = 0, image_off = 0, adj_blocks = 0;
    int depth_max = 0;
    b->rendered = false;
    for (size_t i = 0; i < b->entries.size(); ++i) {
        BatchEntry& en = b->entries[i];
        en.scene_revision = en.scene->revision;
        int rc = MakePathParams(c, en.scene, en.cam, en.params, &en.P);
        if (rc != GY_OK) return rc;
        const int W = en.cam.width, H = en.cam.height;
        rc = FilmPrepareState(c, en.film, W, H, (int64_t)W * H * en.params.spp);
        if (rc != GY_OK) return rc;
        BatchEntryDev& d = tab[i];
        std::memset(&d, 0, sizeof(d));
        d.S = en.scene->dev;
        d.P = en.P;
        d.F = FilmViewState(en.film);
        d.tilesX = (W + 7) >> 3;
        d.tilesY = (H + 7) >> 3;
        d.paths = (int64_t)d.tilesX * d.tilesY * 64 * en.params.spp;
        d.first = first;
        d.image_off = image_off;
        d.adj_block0 = adj_blocks;
        en.first = first;
        en.paths = d.paths;
        en.image_off = image_off;
        first += d.paths;
        image_off += (int64_t)W * H * 3;
        adj_blocks += (d.paths + kBatchAdjointBlock - 1) / kBatchAdjointBlock;
        depth_max = std::max(depth_max, en.params.max_depth);
    }
    DevFree(b->table);
    b->table = nullptr;
    int rc = Upload(c, &b->table, tab.data(), tab.size() * sizeof(BatchEntryDev));
    if (rc != GY_OK) return rc;
    if (!b->counters) {
        rc = DevAlloc(c, (void**)&b->counters, 16 * sizeof(unsigned long long));
        if (rc != GY_OK) return rc;
    }
    b->total_paths = first;
    b->image_floats = image_off;
    b->adj_blocks = adj_blocks;
    b->depth_max = depth_max;
    const int64_t want = lanes_req > 0 ? lanes_req : (int64_t)64 << 20;
    b->lanes = std::min(want, (first + 255) / 256 * 256);
    if (b->lanes >= (1ll << 31)) return Fail(GY_ERR_UNSUPPORTED, "batch pass larger than 2^31 paths");
    b->dirty = false;
    return GY_OK;
}

struct BatchHandle {
    Batch b;
    int64_t lanes_req = 0;
};

}  // namespace
}  // namespace gy

using namespace gy;

extern "C" int gy_batch_create(gy_context* ctx, const gy_batch_params* p, gy_batch** out) {
    if (!ctx || !out) return Fail(GY_ERR_INVALID, "gy_batch_create: null argument");
    if (p && (p->lanes_per_pass < 0 || p->refill_threshold < 0 || p->refill_threshold > 64))
        return Fail(GY_ERR_INVALID, "gy_batch_create: lanes_per_pass >= 0 and 0 <= refill_threshold <= 64");
    if (p && (p->ray_sort < GY_RAY_SORT_NONE || p->ray_sort > GY_RAY_SORT_OCTANT_LOCAL))
        return Fail(GY_ERR_INVALID, "gy_batch_create: unknown ray_sort");
    BatchHandle* h = new BatchHandle;
    h->b.ctx = reinterpret_cast<Ctx*>(ctx);
    if (p) {
        h->lanes_req = p->lanes_per_pass;
        h->b.refill = p->refill_threshold > 0 ? p->refill_threshold : 0;  // 0: per-kind automatic
        h->b.use_graph = p->use_graph != 0;
        h->b.ray_sort = p->ray_sort;
    }
    *out = reinterpret_cast<gy_batch*>(h);
    return GY_OK;
}

extern "C" int gy_batch_add(gy_batch* bh, gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, int32_t* entry) {
    if (!bh || !scene) return Fail(GY_ERR_INVALID, "gy_batch_add: null batch or scene");
    int rc = ValidateRender(cam, p);
    if (rc != GY_OK) return rc;
    if (p->max_depth < 1) return Fail(GY_ERR_INVALID, "gy_batch_add: max_depth >= 1");
    if ((int64_t)cam->width * cam->height * p->spp >= (1ll << 32)) return Fail(GY_ERR_UNSUPPORTED, "gy_batch_add: more than 2^32 paths in one entry");
    Batch& b = reinterpret_cast<BatchHandle*>(bh)->b;
    Scene* sc = reinterpret_cast<Scene*>(scene);
    if (sc->ctx != b.ctx) return Fail(GY_ERR_INVALID, "gy_batch_add: scene belongs to another context");
    if (sc->stats.max_depth + 1 > 64) return Fail(GY_ERR_UNSUPPORTED, "gy_batch_add: BVH deeper than the traversal stack");
    if (sc->dev.n_sph > 0) return Fail(GY_ERR_UNSUPPORTED, "gy_batch_add: scenes with spheres (the batch runs the wavefront kernels)");
    if (sc->has_no_mat) return Fail(GY_ERR_UNSUPPORTED, "gy_batch_add: scenes with material-less triangles (PathTracing only)");
    BatchEntry en;
    en.scene = sc;
    en.cam = *cam;
    en.params = *p;
    b.entries.push_back(en);
    b.dirty = true;
    ++sc->batch_refs;
    if (entry) *entry = (int32_t)b.entries.size() - 1;
    return GY_OK;
}

extern "C" int gy_batch_render(gy_batch* bh) {
    if (!bh) return Fail(GY_ERR_INVALID, "gy_batch_render: null batch");
    BatchHandle* h = reinterpret_cast<BatchHandle*>(bh);
    Batch& b = h->b;
    Ctx* c = b.ctx;
    (void)hipSetDevice(c->device);
    int rc = GY_OK;
    // a scene whose buffers were replaced (gy_scene_set_environment) invalidates
    // the entry table and the recorded graph, which hold its device pointers
    for (const BatchEntry& en : b.entries)
        if (en.scene->revision != en.scene_revision) b.dirty = true;
    if (b.dirty) rc = BatchBuild(c, &b, h->lanes_req);
    if (rc != GY_OK) return rc;
    std::memset(&c->stats, 0, sizeof(c->stats));
    c->ev_used = 0;
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    (void)hipEventRecord(t0, c->stream);
    rc = BatchRender(c, &b);
    (void)hipEventRecord(t1, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "batch sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    TimersCollect(c);
    c->stats.total_ms = ms;
    for (const BatchEntry& en : b.entries) c->stats.paths += (int64_t)en.cam.width * en.cam.height * en.params.spp;
    if (rc == GY_OK) {
        unsigned long long cnt[8];
        e = hipMemcpy(cnt, b.counters, sizeof(cnt), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return HipFail(e, "batch counters");
        c->stats.rays = (int64_t)(cnt[0] + cnt[1] + cnt[2]);
        c->stats.shadow_rays = (int64_t)cnt[1];
        c->stats.mis_rays = (int64_t)cnt[2];
        for (const BatchEntry& en : b.entries) {
            rc = FilmCheck(en.film);
            if (rc != GY_OK) return rc;
        }
        for (BatchEntry& en : b.entries) en.tex_revision = en.scene->tex_revision;
        b.rendered = true;
    }
    return rc;
}

extern "C" int64_t gy_batch_image_floats(gy_batch* bh) {
    if (!bh) return -1;
    const Batch& b = reinterpret_cast<BatchHandle*>(bh)->b;
    return b.dirty ? -1 : b.image_floats;
}

extern "C" int gy_batch_images_device(gy_batch* bh, float* images_device) {
    if (!bh || !images_device) return Fail(GY_ERR_INVALID, "gy_batch_images_device: null argument");
    Batch& b = reinterpret_cast<BatchHandle*>(bh)->b;
    (void)hipSetDevice(b.ctx->device);
    return BatchImages(b.ctx, &b, images_device);
}

extern "C" int gy_batch_adjoint_device(gy_batch* bh, const float* dimages_device, float* grad_tex_device, float* grad_le_device) {
    if (!bh || !dimages_device || !grad_tex_device) return Fail(GY_ERR_INVALID, "gy_batch_adjoint_device: null argument");
    Batch& b = reinterpret_cast<BatchHandle*>(bh)->b;
    Ctx* c = b.ctx;
    (void)hipSetDevice(c->device);
    c->ev_used = 0;
    int rc = BatchAdjoint(c, &b, dimages_device, grad_tex_device, grad_le_device);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "batch adjoint sync");
    TimersCollect(c);
    return rc;
}

extern "C" int gy_batch_adjoint(gy_batch* bh, const float* dimages, float* grad_tex, float* grad_le) {
    if (!bh || !dimages || !grad_tex) return Fail(GY_ERR_INVALID, "gy_batch_adjoint: null argument");
    Batch& b = reinterpret_cast<BatchHandle*>(bh)->b;
    Ctx* c = b.ctx;
    if (b.entries.empty() || b.dirty || !b.rendered) return Fail(GY_ERR_INVALID, "gy_batch_adjoint: render the batch first");
    (void)hipSetDevice(c->device);
    Scene* sc = b.entries[0].scene;
    const int64_t n_le = (int64_t)sc->area_tri.size() * 3;
    void *dI = nullptr, *dg = nullptr, *dl = nullptr;
    int rc = DevAlloc(c, &dI, (size_t)b.image_floats * 4);
    if (rc == GY_OK) rc = DevAlloc(c, &dg, (size_t)std::max<int64_t>(sc->tex_floats, 1) * 4);
    if (rc == GY_OK && grad_le) rc = DevAlloc(c, &dl, (size_t)std::max<int64_t>(n_le, 1) * 4);
    if (rc == GY_OK) {
        hipError_t e = hipMemcpyAsync(dI, dimages, (size_t)b.image_floats * 4, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = HipFail(e, "batch adjoint H2D");
    }
    if (rc == GY_OK) rc = Fill32(c, dg, 0u, std::max<int64_t>(sc->tex_floats, 1));
    if (rc == GY_OK && dl) rc = Fill32(c, dl, 0u, std::max<int64_t>(n_le, 1));
    c->ev_used = 0;
    if (rc == GY_OK) rc = BatchAdjoint(c, &b, (const float*)dI, (float*)dg, (float*)dl);
    if (rc == GY_OK) {
        hipError_t e = hipMemcpyAsync(grad_tex, dg, (size_t)sc->tex_floats * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && dl) e = hipMemcpyAsync(grad_le, dl, (size_t)n_le * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = HipFail(e, "batch adjoint D2H");
    } else {
        (void)hipStreamSynchronize(c->stream);
    }
    TimersCollect(c);
    DevFree(dI);
    DevFree(dg);
    DevFree(dl);
    return rc;
}

extern "C" int gy_batch_film_download(gy_batch* bh, int32_t entry, float* color, float* weight) {
    if (!bh) return Fail(GY_ERR_INVALID, "gy_batch_film_download: null batch");
    Batch& b = reinterpret_cast<BatchHandle*>(bh)->b;
    if (entry < 0 || entry >= (int32_t)b.entries.size()) return Fail(GY_ERR_INVALID, "gy_batch_film_download: no such entry");
    const FilmState& f = b.entries[entry].film;
    if (!f.film || b.dirty) return Fail(GY_ERR_INVALID, "gy_batch_film_download: batch not rendered");
    const int64_t WH = (int64_t)f.width * f.height;
    std::vector<float4> hbuf(WH);
    hipError_t e = hipMemcpy(hbuf.data(), f.film, WH * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return HipFail(e, "batch film D2H");
    for (int64_t i = 0; i < WH; ++i) {
        if (color) {
            color[3 * i] = hbuf[i].x;
            color[3 * i + 1] = hbuf[i].y;
            color[3 * i + 2] = hbuf[i].z;
        }
        if (weight) weight[i] = hbuf[i].w;
    }
    return GY_OK;
}

// Launch choices of the batch as a whole: the batch's entries share one lane
// space and every launch, so a choice holds for all of them.  Times the
// batch's own renders (the second of two per choice: the first re-records the
// graph) over the refill threshold (automatic, 16, 32) and then, at the best
// threshold, the closest-hit ray-sort keys; keeps the fastest.
extern "C" int gy_batch_autotune(gy_batch* bh, gy_batch_params* best, double* best_ms) {
    if (!bh || !best) return Fail(GY_ERR_INVALID, "gy_batch_autotune: null argument");
    BatchHandle* h = reinterpret_cast<BatchHandle*>(bh);
    Batch& b = h->b;
    if (b.entries.empty()) return Fail(GY_ERR_INVALID, "gy_batch_autotune: empty batch");
    struct Cand {
        int refill, sort;
    };
    Cand won = {b.refill, b.ray_sort};
    double best_t = 1e300;
    auto trial = [&](const Cand& k) -> int {
        b.refill = k.refill;
        b.ray_sort = k.sort;
        b.dirty = true;
        double t = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
            const int r = gy_batch_render(bh);
            if (r != GY_OK) return r;
            t = b.ctx->stats.total_ms;
        }
        if (t < best_t) {
            best_t = t;
            won = k;
        }
        return GY_OK;
    };
    int rc = GY_OK;
    const Cand refills[] = {{0, GY_RAY_SORT_NONE}, {16, GY_RAY_SORT_NONE}, {32, GY_RAY_SORT_NONE}};
    for (const Cand& k : refills)
        if (rc == GY_OK) rc = trial(k);
    const int refill = won.refill;
    if (rc == GY_OK) rc = trial({refill, GY_RAY_SORT_OCTANT});
    if (rc == GY_OK) rc = trial({refill, GY_RAY_SORT_OCTANT_LOCAL});
    b.refill = won.refill;
    b.ray_sort = won.sort;
    b.dirty = true;
    if (rc != GY_OK) return rc;
    std::memset(best, 0, sizeof(*best));
    best->lanes_per_pass = h->lanes_req;
    best->refill_threshold = won.refill;
    best->use_graph = b.use_graph ? 1 : 0;
    best->ray_sort = won.sort;
    if (best_ms) *best_ms = best_t;
    return GY_OK;
}

extern "C" int gy_batch_destroy(gy_batch* bh) {
    if (!bh) return GY_OK;
    BatchHandle* h = reinterpret_cast<BatchHandle*>(bh);
    Batch& b = h->b;
    (void)hipSetDevice(b.ctx->device);
    (void)hipStreamSynchronize(b.ctx->stream);
    BatchDropGraph(&b);
    for (BatchEntry& en : b.entries) {
        FilmFree(&en.film);
        --en.scene->batch_refs;
    }
    DevFree(b.table);
    DevFree(b.counters);
    WavefrontStateFree(&b.wf);
    delete h;
    return GY_OK;
}
