// Host-side runtime structures shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device.h"
#include "gy_internal.h"

namespace gy {

enum KernelCategory {
    kKernMega = 0,
    kKernFilm = 1,
    kKernRaygen = 2,
    kKernTraceClosest = 3,
    kKernShade = 4,
    kKernTraceShadow = 5,
    kKernTraceMis = 6,
    kKernOther = 7
};

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct Ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    const uint32_t* d_sobol = nullptr;  // Sobol byte tables (sobol::ByteTables)
    void* d_halton = nullptr;  // digit tables, then per-dimension uint4 info and float scales
    size_t halton_dim_off = 0, halton_scale_off = 0;
    int fault = 0;
    DevBuf rays, hits, ctr;
    double last_kernel_ms = 0.0;
    gy_render_stats stats = {};
    std::vector<hipEvent_t> events;
    std::vector<int> ev_cat;
    int ev_used = 0;
    bool capturing = false;  // a hipGraph capture is recording the stream: no timer events
    void* scratch = nullptr;  // this device's copy of the no-alloc scratch array (gy_render_noalloc)
};

struct FilmState {
    float4* film = nullptr;  // rgb sums + weight, one per pixel
    float4* pass = nullptr;  // one per (pass sample, pixel): rgb + footprint mask
    int64_t pass_slots = 0;
    uint32_t* border = nullptr;  // per pixel: a sample of this pass reaches other pixels; [W*H] error count
    unsigned long long* counters = nullptr;
    int width = 0, height = 0;
};

struct FilmDev {
    float4* film;
    float4* pass;
    uint32_t* border;
    int width, height;
};

// Wavefront queues and path state (wavefront.hip).
struct WavefrontState {
    int64_t capacity = 0;
    void* mem = nullptr;
    size_t bytes = 0;
    void* spill = nullptr;
    size_t spill_bytes = 0;
    bool batch = false;  // carries the per-lane entry and pass-slot arrays
};

struct Scene {
    Ctx* ctx = nullptr;
    int32_t n_tri = 0;
    int32_t batch_refs = 0;  // batch entries that render this scene
    gyd::SceneDev dev = {};
    void *d_tri = nullptr, *d_shade = nullptr, *d_nodes = nullptr, *d_mats = nullptr, *d_tex = nullptr, *d_texinfo = nullptr;
    void *d_ltri = nullptr, *d_lle = nullptr, *d_lcdf = nullptr, *d_lfunc = nullptr;
    void* d_env = nullptr;  // environment light: texels and sampling tables in one block
    void* d_sph = nullptr;  // analytic spheres (gy_scene_set_spheres)
    // participating media (gy_scene_set_media): 2 float4 per medium, one word
    // per triangle (inside + 1, outside + 1 << 16, no material << 31)
    void *d_media = nullptr, *d_tri_media = nullptr;
    int32_t n_media = 0;
    bool has_no_mat = false;  // some triangle has no material (a medium boundary)
    uint64_t revision = 0;  // bumped when device buffers the kernels point at are replaced
    uint64_t uid = 0;       // unique over the process (never reused, unlike the address)
    float world_lo[3] = {0, 0, 0}, world_hi[3] = {0, 0, 0};
    std::vector<int32_t> area_tri;  // area lights as created (the environment light follows them)
    std::vector<float> area_le, area_func;
    int stack = 64;
    gy_scene_stats stats = {};
    double upload_ms = 0.0;
    std::vector<uint32_t> host_nodes;
    std::vector<uint32_t> tri_order;
    FilmState film;
    WavefrontState wf;
    WavefrontState vwf;  // VolPathTracing's wavefront state
    int64_t tex_floats = 0;     // floats of texture data on the device
    uint64_t tex_revision = 0;  // bumped whenever the texels are replaced
    bool film_valid = false;    // the film holds the last gy_render of film_cam / film_params
    bool film_single_pass = false;  // ... rendered in one pass: the pass buffer holds every sample
    gy_camera_desc film_cam = {};
    gy_render_params film_params = {};
};

int HipFail(hipError_t e, const char* what);
int DevAlloc(Ctx* c, void** p, size_t bytes);
int Upload(Ctx* c, void** p, const void* src, size_t bytes);
void DevFree(void* p);
int Grow(Ctx* c, DevBuf* b, size_t bytes);
int Fill32(Ctx* c, void* p, uint32_t v, int64_t n);
int FillF4(Ctx* c, float4* p, int64_t n);  // zero n float4
int CheckLaunch(Ctx* c, const char* what);
void TimerBegin(Ctx* c, int cat);
void TimerEnd(Ctx* c);
void TimersCollect(Ctx* c);
void SceneFree(Scene* sc);
void FilmFree(FilmState* f);
FilmDev FilmView(Scene* sc);
FilmDev FilmViewState(const FilmState& f);
int FilmPrepareState(Ctx* c, FilmState& f, int W, int H, int64_t pass_slots);
int FilmResolveState(Ctx* c, const FilmState& f, int n_s, bool timed = true);
int FilmClear(Ctx* c, const FilmState& f);  // zero film, border flags, counters
int FilmCheck(const FilmState& f);           // footprint error of the last render

// Per-render constants of the integrator kernels.
struct PathParams {
    gyd::CamDev cam;
    gyd::EnumDev en;
    gyd::SampleSrc smp;
    int32_t max_depth, rr_depth;
};
int WavefrontPass(Ctx* c, Scene* sc, const PathParams& P, const gy_render_params& p, int s0, int n_s);
void WavefrontFree(Scene* sc);
void WavefrontStateFree(WavefrontState* wf);

// A batch of renders (scene, view, sample budget) traced in one lane space.
struct BatchEntryDev {
    gyd::SceneDev S;
    PathParams P;
    FilmDev F;
    int64_t first;  // first global path index of the entry
    int64_t paths;  // tile-padded paths: tilesX * tilesY * 64 * spp
    int64_t image_off;  // the entry's first float in the batch's concatenated images (W * H * 3 per entry)
    int64_t adj_block0; // the entry's first workgroup in the batch adjoint launch (kBatchAdjointBlock lanes each)
    int32_t tilesX, tilesY;
};
constexpr int kBatchAdjointBlock = 128;
struct BatchEntry {
    Scene* scene = nullptr;
    gy_camera_desc cam = {};
    gy_render_params params = {};
    PathParams P = {};
    FilmState film;
    int64_t first = 0, paths = 0;
    int64_t image_off = 0;
    uint64_t tex_revision = 0;  // the scene's texture revision the last batch render saw
    uint64_t scene_revision = 0;  // the scene's revision when the table was built
};
struct Batch {
    Ctx* ctx = nullptr;
    std::vector<BatchEntry> entries;
    void* table = nullptr;  // BatchEntryDev per entry (device)
    unsigned long long* counters = nullptr;
    WavefrontState wf;
    int64_t lanes = 0;        // paths per pass
    int64_t total_paths = 0;
    int depth_max = 0;
    int refill = 0;  // 0: per-kind automatic threshold
    int ray_sort = GY_RAY_SORT_NONE;
    bool use_graph = true;
    bool dirty = true;        // entries changed since the table / graph were built
    bool rendered = false;    // the films hold a render of the current table
    int64_t image_floats = 0; // floats of all entries' images, concatenated
    int64_t adj_blocks = 0;   // workgroups of the batch adjoint launch
    hipGraphExec_t graph_exec = nullptr;
};
int BatchRender(Ctx* c, Batch* b);
// Path-replay adjoint of every entry of the last batch render, in one launch
// (adjoint.hip); all entries share one scene.
int BatchAdjoint(Ctx* c, Batch* b, const float* dimages, float* grad_tex, float* grad_le);
int BatchImages(Ctx* c, Batch* b, float* images);
// GPU binned-SAH BVH build (bvh_gpu.hip).
int BuildBVHDevice(Ctx* c, const float* pos, int32_t n_tri, int max_leaf, BVHBuild* out);

}  // namespace gy
