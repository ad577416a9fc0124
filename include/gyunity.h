/* gyunity.h - C ABI of the MI355X path-tracing integrator (libgyunity.so).
 *
 * This is the only seam between host code and the HIP kernels.  Plain C types,
 * host pointers unless a name ends in _device, int status codes (GY_OK = 0),
 * and a thread-local message from gy_last_error().  No call throws or exits.
 *
 * The reference has no C ABI; its integrator is the C++ virtual
 *     class Integrator { virtual void Render(const Scene&, const Camera&) = 0; }
 * (Gyunity/src/integrator/Integrator.h:17-27), implemented on the CPU by
 *     PathTracing(spp, maxDepth, sampler, samplerEnum)  (integrator/PT.h:12)
 *     TiledIntegrator::Render                           (integrator/TiledIntegrator.cpp:7-60)
 * Each entry point below names the reference interface it stands in for.
 * The C++ class gy::HipPathTracing (gyunity_amd/include/gyunity/HipPathTracing.h)
 * implements that Integrator interface on top of these calls.
 */
#ifndef GYUNITY_H
#define GYUNITY_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GY_OK = 0,
    GY_ERR_INVALID = 1,   /* bad argument or descriptor */
    GY_ERR_DEVICE = 2,    /* HIP runtime error (no device, launch failure) */
    GY_ERR_ALLOC = 3,     /* device or host allocation failed */
    GY_ERR_UNSUPPORTED = 4,
    GY_ERR_INJECTED = 5   /* failure forced by gy_debug_inject_fault */
};

typedef struct gy_context gy_context;
typedef struct gy_scene gy_scene;

/* Material record: 32 four-byte words.
 *   word 0  type: GY_MAT_LAMBERT (DiffuseMaterial, material/Diffuse.h:27-31),
 *                 GY_MAT_CONDUCTOR (RoughnessMaterial, material/Roughness.h:36-46),
 *                 GY_MAT_GLASS (GlassMaterial, material/Glass.h:30-34),
 *                 GY_MAT_ROUGH_DIELECTRIC (RoughDielectricMaterial, material/RoughDielectric.h:21-41)
 *   word 1  flags: GY_MATF_ALPHA_DIRECT (alpha given; else RoughnessToAlpha(roughness)),
 *                  GY_MATF_SCHLICK_F0 (conductor Fresnel from F0; else FrConductor(eta, k))
 *   2-4 R, 5-7 T, 8-9 alpha (or roughness) x/y, 10-12 eta, 13-15 k, 16-18 F0,
 *   19 dielectric eta outside, 20 eta inside, 21 albedo texture id, 22 alpha texture id. */
#define GY_MAT_WORDS 32
enum { GY_MAT_LAMBERT = 0, GY_MAT_CONDUCTOR = 1, GY_MAT_GLASS = 2, GY_MAT_ROUGH_DIELECTRIC = 3 };
enum { GY_MATF_ALPHA_DIRECT = 1, GY_MATF_SCHLICK_F0 = 2 };

/* Flattened scene: what Scene::AddPrimitive / AddLight (visual/Scene.h:39-65)
 * accumulate.  Primitive ids are triangle indices (Scene.h:43). */
typedef struct {
    int32_t n_tri;
    const float* pos;        /* n_tri*9  p0,p1,p2 */
    const float* nrm;        /* n_tri*9  vertex normals n0,n1,n2 */
    const float* face_n;     /* n_tri*3  face normal (Triangle ctor argument) */
    const float* uv;         /* n_tri*6 or NULL (all zero, as shape/Triangle.h:93-105 gives) */
    const uint32_t* mat;     /* n_tri material ids */
    const int32_t* light;    /* n_tri light ids, -1 if not emissive (NULL = none) */
    int32_t n_mat;
    const uint32_t* mats;    /* n_mat*GY_MAT_WORDS */
    int32_t n_light;
    const int32_t* light_tri;  /* triangle of each AreaLight (light/AreaLight.h:12-85) */
    const float* light_le;     /* n_light*3 emitted radiance */
    int32_t n_tex;
    const int32_t* tex_info;   /* n_tex*4: width, height, channels, float offset into tex_data */
    const float* tex_data;
} gy_scene_desc;

/* PinHoleCamera(film, position, lookAt, up, fovy, filmDistance)
 * (camera/PinholeCamera.h:16-20, camera/Camera.cpp:68-72) on a width x height film. */
typedef struct {
    float pos[3], look_at[3], up[3];
    float fovy;
    float film_dist;
    int32_t width, height;
} gy_camera_desc;

enum { GY_RENDER_MEGAKERNEL = 0, GY_RENDER_WAVEFRONT = 1 };

/* PathTracing(spp, maxDepth, SobolSampler, SobolEnum) (integrator/PT.h:12) plus
 * the Russian-roulette threshold (reference: bounce i > 5, PT.cpp:53). */
typedef struct {
    int32_t spp;
    int32_t max_depth;
    int32_t rr_depth;
    int32_t mode;            /* GY_RENDER_MEGAKERNEL or GY_RENDER_WAVEFRONT */
    int32_t samples_per_pass;/* 0 = automatic */
    int32_t block_size;      /* megakernel and light-tracing block size: 0 = automatic; else 64, 128 or 256 */
    int32_t refill_threshold;/* wavefront traversal: refill rays when fewer lanes of a
                                wave than this are busy (0 = automatic, 48; 1..64) */
    int32_t sampler;         /* GY_SAMPLER_SOBOL: SobolSampler + SobolEnum;
                                GY_SAMPLER_HALTON: HaltonSampler (Faure) + HaltonEnum */
    int32_t ray_sort;        /* wavefront: order of the closest-hit queue from the
                                second bounce on: GY_RAY_SORT_NONE (as generated),
                                GY_RAY_SORT_OCTANT (stable partition of the whole queue
                                by direction octant), GY_RAY_SORT_OCTANT_LOCAL (the same
                                within each run of 1024 entries).  The image does not
                                depend on it. */
} gy_render_params;
enum { GY_RAY_SORT_NONE = 0, GY_RAY_SORT_OCTANT = 1, GY_RAY_SORT_OCTANT_LOCAL = 2 };
enum { GY_SAMPLER_SOBOL = 0, GY_SAMPLER_HALTON = 1 };

/* BVH build knobs (BVHAccel(prims, maxPrims, "sah"), accelerator/BVHAccel.h:50-51). */
typedef struct {
    int32_t max_leaf_size;   /* 0 = 4 */
    int32_t builder;         /* 0 = GPU binned SAH, 1 = host binned SAH */
    int32_t reserved[2];
} gy_bvh_params;

typedef struct {
    int64_t n_nodes;
    int64_t n_leaves;
    int32_t max_depth;
    int32_t max_leaf_size;
    double build_ms;
    double sah_cost;
} gy_scene_stats;

/* One ray: origin, tmin, direction, tmax (math/Ray.h:9-29 without the
 * constructor's RayEps shift: callers pass the final origin). 32 bytes. */
typedef struct { float org[3]; float tmin; float dir[3]; float tmax; } gy_ray;
/* Closest hit: t, primitive id (-1 = miss), barycentrics b1, b2
 * (shape/Triangle.cpp:71-103, visual/Primitive.cpp:5-18). 16 bytes. */
typedef struct { float t; int32_t prim; float b1; float b2; } gy_hit;

typedef struct {
    int64_t paths;
    int64_t rays;            /* every traced ray: camera/continuation, shadow, MIS */
    int64_t shadow_rays;
    int64_t mis_rays;
    double total_ms;         /* wall time of the render call on the device stream */
    double kernel_ms[8];     /* per-kernel time, see gy_kernel_name */
    int64_t kernel_launches[8];
    int64_t node_visits;     /* filled by gy_count_traffic: all rays */
    int64_t tri_visits;
    int64_t shade_points;
    int64_t node_visits_kind[3];  /* by ray kind: 0 camera/continuation, 1 shadow, 2 MIS */
    int64_t tri_visits_kind[3];
} gy_render_stats;

/* ---- context ------------------------------------------------------------ */
int gy_open(int device, gy_context** out);
int gy_close(gy_context* ctx);
const char* gy_last_error(void);
const char* gy_version(void);
int gy_device_name(gy_context* ctx, char* buf, int len);

/* ---- scene (Scene::Initialize + BVHAccel::BVHAccel, visual/Scene.h:26-33,
 *      accelerator/BVHAccel.cpp:154-200) ------------------------------------ */
int gy_scene_create(gy_context* ctx, const gy_scene_desc* desc, const gy_bvh_params* bvh, gy_scene** out);
/* Fails (GY_ERR_INVALID, the scene kept) while a batch holds the scene:
 * destroy the batch first. */
int gy_scene_destroy(gy_scene* scene);
/* Environment light: EnvironmentLight(file, rotate, scale) + Scene::AddLight
 * (light/EnvLight.h:14-17, visual/Scene.h:59-64) after the area lights, with the
 * image already loaded: rgb is the latitude-longitude image, width*height*3
 * floats, top row first (as ImageIO::LoadTexture reads it through stbi_loadf,
 * image/ImageIO.cpp:100-114).  The map is importance-sampled as EnvironmentMap
 * does (light/EnvMap.h:7-134) and seen by camera rays and rays after a delta
 * bounce that leave the scene (integrator/PT.cpp:18-21).  A scene may be
 * created with no area light and take its light from here.  Calling it again
 * replaces the map.  Not supported by gy_adjoint. */
int gy_scene_set_environment(gy_scene* scene, const float* rgb, int32_t width, int32_t height, float scale);
/* Analytic spheres: Sphere(Transform(), Transform(), radius, position, inside)
 * (shape/Sphere.h:21-27) with a material, and for emissive ones an AreaLight
 * (light/AreaLight.h:12-85) appended after the scene's area lights.  Primitive
 * ids follow the triangles' (n_tri + i).  Once per scene, before any
 * environment light.  The reference's sphere is restated with its defects:
 * a closest hit leaves the default (zero) intersection record but its uv, so
 * a path ends there; any-hit queries are occluded by every sphere whose box
 * the segment passes; a sphere light's Pdf_Li is 0.  Scenes with spheres
 * render with the per-path kernels (gy_render's wavefront form runs the
 * megakernel; the batch, light tracing and the adjoint refuse them). */
typedef struct {
    float center[3];
    float radius;
    int32_t inside;     /* normals point inwards (GetNorm, shape/Sphere.h:338-342) */
    uint32_t material;
    int32_t emissive;   /* 1: an AreaLight with emitted radiance le */
    float le[3];
} gy_sphere;
int gy_scene_set_spheres(gy_scene* scene, int32_t n, const gy_sphere* spheres);
/* Participating media for VolPathTracing: HomogeneousMedium(sigma_a, sigma_s, g)
 * (volume/Homogeneous.h:11-12, Henyey-Greenstein phase, visual/Medium.h:39-48)
 * and, per triangle (scene order), Scene::AddPrimitive's MediumInterface(inside,
 * outside) (visual/Medium.h:51-59, -1: none) and whether the primitive has no
 * material (a medium boundary, crossed by rays).  Calling it again replaces
 * the media.  Scenes with material-less triangles render only through
 * gy_render_vpt / gy_vpt_li_samples. */
typedef struct {
    float sigma_a[3];
    float sigma_s[3];
    float g;              /* (-1, 1) */
} gy_medium;
typedef struct {
    int32_t inside, outside;
    int32_t no_material;
} gy_tri_medium;
int gy_scene_set_media(gy_scene* scene, int32_t n_media, const gy_medium* media, const gy_tri_medium* tri);
/* VolPathTracing(spp, maxDepth, sampler, samplerEnum, equiAngular)
 * (integrator/VPT.h:12-19, VPT.cpp:7-171) through TiledIntegrator::Render:
 * params as gy_render's (spp, max_depth, sampler, samples_per_pass, mode:
 * GY_RENDER_WAVEFRONT runs the wavefront form, GY_RENDER_MEGAKERNEL one path
 * per thread, with equal films; rr_depth, block_size, refill_threshold and
 * ray_sort are not used: the reference's VPT has no working roulette).  equiangular 1:
 * equi-angular distance sampling about one light point per bounce
 * (HomogeneousMedium::EquiAngularSampling, the reference's default); 0:
 * free-flight sampling with DirectIllumination at medium points.  Not
 * supported: spheres; equi-angular sampling with an environment light. */
int gy_render_vpt(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* params, int32_t equiangular,
                  float* color, float* weight);
/* SPPM(iterations, nPhotonsPerStage, maxDepth, initialRadius, alpha, batchShrink,
 * sampler, samplerEnum, traceGlossyRay) (integrator/SPPM.h:48-49, SPPM.cpp:198-277):
 * rgb = width x height x 3, the image SPPM::Render gives Film::SetImage;
 * radius2 and photon_count (NULL: not wanted) = each pixel's final radius^2
 * and photon count N.  Triangle scenes with area lights; traceGlossyRay is not
 * supported.  The reference updates hit points from concurrent photons without
 * locks; here each hit point applies its photons in photon-index order. */
typedef struct {
    int32_t iterations;
    int32_t max_depth;
    int64_t photons_per_iteration;
    float initial_radius;
    float alpha;
    int32_t batch_shrink;
    int32_t sampler;          /* GY_SAMPLER_SOBOL or GY_SAMPLER_HALTON (sampler and enumeration) */
    int32_t trace_glossy_ray; /* must be 0 */
    int64_t photons_per_launch; /* photon paths traced per launch, over whole iterations
                                   (0 = automatic, about 4M); the image does not depend on it */
} gy_sppm_params;
int gy_render_sppm(gy_scene* scene, const gy_camera_desc* cam, const gy_sppm_params* params, float* rgb, float* radius2,
                   float* photon_count);
/* Per-sample VolPathTracing::Li on n (x, y, sample) indices: radiance (3n),
 * film position (2n) and dimensions drawn (n). */
int gy_vpt_li_samples(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* params, int32_t equiangular,
                      int32_t n, const int32_t* x, const int32_t* y, const int32_t* sample, float* L, float* raster,
                      int32_t* dims);
int gy_scene_stats_get(gy_scene* scene, gy_scene_stats* out);
/* The context a scene was created on. */
int gy_scene_context(gy_scene* scene, gy_context** ctx);
/* Copy the device BVH back: nodes (n_nodes*16 words) and leaf triangle order. */
int gy_scene_bvh_download(gy_scene* scene, uint32_t* nodes, int64_t max_nodes, int32_t* tri_order, int64_t max_tris);

/* ---- ray queries (Scene::Intersect, visual/Scene.h:83-89 ->
 *      BVHAccel::Intersect closest / any, accelerator/BVHAccel.cpp:401-487) -- */
int gy_trace_closest(gy_scene* scene, const gy_ray* rays, int64_t n, gy_hit* hits);
int gy_trace_any(gy_scene* scene, const gy_ray* rays, int64_t n, int32_t* occluded);
int gy_trace_closest_device(gy_scene* scene, const gy_ray* rays_device, int64_t n, gy_hit* hits_device);
int gy_trace_any_device(gy_scene* scene, const gy_ray* rays_device, int64_t n, int32_t* occluded_device);
/* Device time (ms) of the last trace kernel, measured with HIP events on the
 * stream it ran on. */
double gy_last_kernel_ms(gy_context* ctx);
/* Count BVH nodes and triangles the closest-hit traversal visits. */
int gy_count_visits(gy_scene* scene, const gy_ray* rays, int64_t n, int64_t* node_visits, int64_t* tri_visits);

/* ---- rendering (Integrator::Render + Film::AddSample,
 *      integrator/TiledIntegrator.cpp:7-60, visual/Film.h:61-87) ---------- */
/* color: width*height*3 sums, weight: width*height sample counts (either may
 * be NULL to leave the film on the device; gy_film_download fetches it). */
int gy_render(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, float* color, float* weight);
int gy_film_download(gy_scene* scene, float* color, float* weight);
/* gy_render that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (tight serving loops,
 * harnesses that forbid hipMalloc).  Megakernel form only (p->mode
 * GY_RENDER_MEGAKERNEL; the wavefront's path state is allocated).  film_device:
 * width*height float4 on the scene's device (rgb sums, weight), written whole;
 * its film equals gy_render's bit for bit.  Counters, border flags and pass
 * buffer live in a 64 MiB array each device holds for the life of the process:
 * at most 3,355,392 pixels (GY_ERR_UNSUPPORTED beyond, or when an explicit
 * samples_per_pass does not fit; automatic passes take what fits).  Calls that
 * overlap on one device share that array and race.  Synchronous; the scene's
 * own film and gy_render_stats_get are left as they were. */
int gy_render_noalloc(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, float* film_device);
int gy_render_stats_get(gy_context* ctx, gy_render_stats* out);
const char* gy_kernel_name(int index);
/* Radiance of individual camera samples (x, y, s): PathTracing::Li for the path
 * a tile worker traces (TiledIntegrator.cpp:40-47).  L: 3n floats, raster: 2n
 * floats (x+u, y+v), dims: n sample dimensions consumed. */
int gy_li_samples(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, int32_t n,
                  const int32_t* x, const int32_t* y, const int32_t* s, float* L, float* raster, int32_t* dims);
/* Algorithmic HBM traffic of a render: node/triangle visits of every traced ray
 * and shading points, counted by an instrumented copy of the kernels. */
int gy_count_traffic(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, gy_render_stats* out);

/* ---- launch autotuner --------------------------------------------------------
 * No counterpart in the reference (TiledIntegrator fixes 32x32 tiles,
 * integrator/TiledIntegrator.cpp:13).  Times pilot renders of this scene and
 * view (pilot_spp samples per pixel; 0 = the render itself up to 2^23 paths,
 * else the fewest samples that reach 2^23 paths): the integrator form first,
 * then the megakernel workgroup size, or the wavefront refill threshold and
 * closest-hit queue order (ray_sort), of the faster form; writes base with the
 * fastest choice into best.  The image does not depend on the choice.  Choices
 * are kept per scene signature (the scene, the revision of its environment
 * light / spheres / media, camera, spp, depth, roulette, sampler, pilot_spp):
 * asking again returns the kept choice without timing. */
int gy_autotune(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* base, int32_t pilot_spp,
                gy_render_params* best, double* best_ms_per_mpath);
/* Forget every kept choice: the next gy_autotune times again. */
int gy_autotune_clear_cache(void);

/* ---- batched rendering ------------------------------------------------------
 * No counterpart in the reference (TiledIntegrator::Render draws one image per
 * call, integrator/TiledIntegrator.cpp:7-60).  A batch holds many renders
 * (scene, view, sample budget, depth) of uneven size; gy_batch_render traces all
 * of their paths in one lane space with the wavefront kernels (persistent trace
 * waves steal queue entries of any render), records the whole sequence into a
 * hipGraph on the first call and replays it afterwards.  Each entry's film
 * equals what gy_render gives for that entry alone.  Scenes must outlive the
 * batch; adding an entry re-records the graph. */
typedef struct gy_batch gy_batch;
typedef struct {
    int64_t lanes_per_pass;   /* paths in flight per pass: 0 = automatic (64M) */
    int32_t refill_threshold; /* persistent trace refill (0 = automatic, 1..64) */
    int32_t use_graph;        /* 1: record once into a hipGraph and replay */
    int32_t ray_sort;         /* GY_RAY_SORT_*, as in gy_render_params */
    int32_t reserved;
} gy_batch_params;
int gy_batch_create(gy_context* ctx, const gy_batch_params* p, gy_batch** out);
/* The batch holds the scene until gy_batch_destroy (gy_scene_destroy refuses
 * it meanwhile); changes to the scene's textures after a render are seen by the
 * next gy_batch_render. */
int gy_batch_add(gy_batch* b, gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, int32_t* entry);
int gy_batch_render(gy_batch* b);
int gy_batch_film_download(gy_batch* b, int32_t entry, float* color, float* weight);
int gy_batch_destroy(gy_batch* b);
/* Launch choices for the batch as a whole (its entries share one lane space and
 * every launch): times the batch's own renders over the refill threshold
 * (automatic, 16, 32), then the ray-sort keys at the best threshold; keeps the
 * fastest in the batch (the graph is re-recorded at the next render) and writes
 * it into best, with best_ms the time of one render.  The films then hold a
 * render of the batch; they do not depend on the choice.  No counterpart in the
 * reference. */
int gy_batch_autotune(gy_batch* b, gy_batch_params* best, double* best_ms);
/* Differentiable batch: every entry of the last gy_batch_render renders the
 * same textured scene (views of one scene).  Images are concatenated in entry
 * order, width*height*3 floats each (gy_batch_image_floats in total; -1 before
 * the entries are built).  gy_batch_adjoint replays every sample of every entry
 * in one launch (see gy_adjoint): dimages holds dl/dI of all images; grad_tex
 * one float per texture float, grad_le 3 per area light (NULL: none);
 * gy_batch_adjoint writes them, the _device form adds to them. */
int64_t gy_batch_image_floats(gy_batch* b);
int gy_batch_images_device(gy_batch* b, float* images_device);
int gy_batch_adjoint(gy_batch* b, const float* dimages, float* grad_tex, float* grad_le);
int gy_batch_adjoint_device(gy_batch* b, const float* dimages_device, float* grad_tex_device, float* grad_le_device);

/* ---- samplers (SobolSampler::Sample, sampler/Sobol.h:66-77;
 *      SobolEnum::GetIndex, sampler/SobolEnum.h:24-28; HaltonSampler,
 *      HaltonEnum) ------------------------------------------------------------ */
float gy_sobol_sample(uint64_t index, uint32_t dim);
uint64_t gy_sobol_enum_index(uint32_t log2_res, uint32_t frame, uint32_t px, uint32_t py);
int gy_camera_matrix(const gy_camera_desc* cam, float* raster_to_world16);
/* HaltonSampler::sample with Faure permutations (sampler/Halton.h:595-625,
 * 648-910; 0 for dim >= 256) and HaltonEnum::get_index for a width x height
 * film (sampler/HaltonEnum.h:67-103). */
float gy_halton_sample(uint64_t index, uint32_t dim);
uint64_t gy_halton_enum_index(uint32_t width, uint32_t height, uint64_t i, uint32_t x, uint32_t y);

/* ---- film output (Film::SaveImage, visual/Film.cpp:6-9; ImageIO::WriteImage,
 *      image/ImageIO.cpp:120-149) ------------------------------------------
 * gy_image_write picks the format by suffix: .hdr (Radiance RGBE, run-length
 * scanlines), .bmp (24-bit) or .png (8-bit RGB); 8-bit values are
 * ImageIO::toInt, int(pow(clamp(x, 0, 1), 1 / 2.2) * 255 + 0.5).  rgb holds
 * width*height*3 floats, rows top to bottom.  gy_film_save writes the scene's
 * last render as color / weight (0 where the weight is 0). */
int gy_image_write(const char* path, const float* rgb, int32_t width, int32_t height);
int gy_image_to_bytes(const float* rgb, int64_t n, uint8_t* out);
int gy_film_save(gy_scene* scene, const char* path);
/* RMSE of the last render's pixels (colour / weight, the pixel values
 * Film::SaveImage writes, visual/Film.cpp:6-9) against a reference image, over
 * every channel of the given image rows: reference holds n_rows * width * 3
 * floats, row rows[i] at i; rows NULL: all rows, reference the whole image.
 * The harness of "Mpaths/s at fixed RMSE". */
int gy_film_rmse(gy_scene* scene, const float* reference, const int32_t* rows, int32_t n_rows, double* rmse);
/* LightTracing (integrator/LT.h:11-241): LightTracing(sampler, maxDepth, spp)
 * ::Render with p->spp, p->max_depth and p->sampler.  Light paths from the
 * area lights, every vertex connected to the pinhole camera and splatted
 * (Film::AddSplat, visual/Film.h:89-103); rgb gets the W x H x 3 splat sums.
 * An environment light's paths carry zero throughput (GenerateLightPath takes
 * Le = Emission(), zero for it: light/EnvLight.h:20-22) but are traced, so
 * their NaNs show; a scene with no light is GY_ERR_INVALID.  Dielectrics
 * scatter light paths in the importance transport mode.  p->mode picks the form:
 * GY_RENDER_WAVEFRONT traces every path's rays of one bounce together (queues
 * and persistent trace kernels), GY_RENDER_MEGAKERNEL one path per thread
 * (p->block_size 64, 128 or 256; 0 = 64).  Float atomics add the splats: the
 * per-path values are fixed, their summation order is not. */
int gy_render_light_tracing(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, float* rgb);
/* Reads a Radiance .hdr file as stbi_loadf does for ImageIO::LoadTexture
 * (image/ImageIO.cpp:100-114), the loader of EnvironmentMap(filename): rgb gets
 * width*height*3 floats, top row first.  With rgb NULL only the size is
 * returned.  Other formats: GY_ERR_UNSUPPORTED. */
int gy_image_read(const char* path, int32_t* width, int32_t* height, float* rgb, int64_t max_floats);

/* ---- OBJ meshes (Mesh::LoadFromFile, mesh/Mesh.cpp:7-84, through
 *      tinyobjloader 2.0.0 with triangulation; Scene::AddMesh,
 *      visual/Scene.h:68-77, and Triangle::SetTransform, shape/Triangle.h:78-88)
 * Loads the triangles a Gyunity scene holds for the file: one vertex triple
 * per face (in (0, 2, 1) order when reverse_vertices is set), the file's
 * normals or a generated face normal, uv (0, 0) where absent; points through
 * object_to_world (16 floats, row-major; NULL = identity), normals through the
 * transpose of world_to_object (NULL = the inverse of object_to_world) and
 * normalised; face normal Cross(p1 - p0, p2 - p1), normalised.  The arrays have
 * the gy_scene_desc layout: pos and nrm 9 floats, face_n 3, uv 6 per triangle.
 * gy_mesh_arrays copies into the buffers that are not NULL. */
typedef struct gy_mesh gy_mesh;
int gy_mesh_load_obj(const char* path, int32_t reverse_vertices, const float* object_to_world,
                     const float* world_to_object, gy_mesh** out);
int64_t gy_mesh_triangles(gy_mesh* mesh);
int gy_mesh_arrays(gy_mesh* mesh, float* pos, float* nrm, float* face_n, float* uv);
int gy_mesh_free(gy_mesh* mesh);

/* ---- procedural scenes (the benchmark configurations) ---------------------- */
typedef struct gy_generated gy_generated;
enum { GY_SCENE_CORNELL = 0, GY_SCENE_ICOSPHERE_GGX = 1, GY_SCENE_BATCH_MEMBER = 2, GY_SCENE_ADJOINT = 3,
       GY_SCENE_ICOSPHERE_TEXTURED = 4 };
int gy_generate_scene(int kind, uint32_t seed, int32_t param, gy_generated** out);
int gy_generated_desc(gy_generated* g, gy_scene_desc* desc, gy_camera_desc* cam, gy_render_params* params);
int gy_generated_free(gy_generated* g);

/* ---- measurement ---------------------------------------------------------
 * No counterpart in the reference.  Copies `bytes` (a multiple of 16) between
 * two fresh device buffers with one plain copy kernel (k_copy_probe), twice,
 * and returns the second launch's device time.  Known-size reference for HBM
 * counter runs and the copy bandwidth the device reaches. */
int gy_copy_probe(gy_context* ctx, int64_t bytes, double* ms);

/* ---- path-replay adjoint ----------------------------------------------------
 * No counterpart in the reference (its renderer has no derivatives).  Gradients
 * of an image loss with respect to texture texels: albedo textures of Lambert
 * and GGX conductor materials and GGX roughness textures.  dimage holds
 * dl/dI for every pixel value I = colour / weight (width*height*3 floats); the
 * gradient has one float per texture float (the layout of
 * gy_scene_desc.tex_data).  The film on the device must hold gy_render of the
 * same camera and sample parameters: every camera path is traced again with its
 * own sample index and replayed (sampling pdfs, MIS weights and roulette held
 * fixed).  gy_adjoint writes grad; gy_adjoint_device adds to it. */
int64_t gy_scene_texture_floats(gy_scene* scene);
/* Pixel values colour / weight of the film on the device into image_device
 * (width*height*3 floats, device memory). */
int gy_film_image_device(gy_scene* scene, float* image_device);
int gy_scene_set_textures(gy_scene* scene, const float* tex_data, int64_t n);
int gy_scene_set_textures_device(gy_scene* scene, const float* tex_data_device, int64_t n);
int gy_adjoint(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, const float* dimage, float* grad);
int gy_adjoint_device(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, const float* dimage_device,
                      float* grad_device);
/* The same, plus the gradient with respect to each area light's emitted
 * radiance Le (AreaLight::Emission, light/AreaLight.h:16, 33-43): grad_le holds
 * 3 floats per area light in light order (NULL: none).  The light-selection
 * probabilities (Scene::ComputeLightPowerDistribution) are sampling
 * probabilities and held fixed, as every pdf of the replay. */
int gy_adjoint_emission(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p, const float* dimage,
                        float* grad, float* grad_le);
int gy_adjoint_emission_device(gy_scene* scene, const gy_camera_desc* cam, const gy_render_params* p,
                               const float* dimage_device, float* grad_device, float* grad_le_device);
/* The roughness terms of the adjoint, evaluated on the host (tests): d log f / d
 * alpha of the isotropic GGX reflection lobe for local directions wo, wi, and
 * d alpha / d roughness of RoughnessToAlpha (visual/Microfacet.h:49-54). */
float gy_ggx_dlogf_dalpha(float alpha, const float* wo, const float* wi);
float gy_dalpha_droughness(float roughness);

/* ---- failure injection (host-forced, for tests) --------------------------- */
enum { GY_FAULT_NONE = 0, GY_FAULT_ALLOC = 1, GY_FAULT_LAUNCH = 2, GY_FAULT_SCENE = 3 };
int gy_debug_inject_fault(gy_context* ctx, int kind);

#ifdef __cplusplus
}
#endif
#endif
