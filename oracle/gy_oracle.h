/* CPU oracle for the Gyunity path-tracing hot path: TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference CPU integrator (PathTracing over
 * TiledIntegrator, Sobol sampling with per-pixel enumeration, the pinhole
 * camera, the 12-bucket SAH BVH and brute-force closest/any hit, the
 * Lambert / GGX conductor / smooth and rough dielectric BSDFs, NEE + MIS to
 * triangle area lights and an environment map, and box-filter film
 * accumulation).  Each function in
 * gy_oracle.cpp cites the reference file:line it follows.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline; the
 * product path never calls it.
 *
 * What pins each part (the checks in tests/, none of which retypes the
 * formula under test):
 *   Sobol + enumeration  dim-0 = bit reversal, (0,2)-net, brute-force
 *                        enumeration, the reference's own tables (test_oracle_pins.py)
 *   Halton               base-2 bit reversal, Faure permutations' known values,
 *                        radical inverse, CRT strata (test_oracle_pins.py)
 *   camera, matrices     closed-form rays; the reference's MatrixTest known answers
 *   BVH + traversal      brute force over the same triangle test (level a)
 *   BSDFs, Fresnel, GGX  furnace, normalisation, chi-square of Sample vs Pdf,
 *                        Fresnel closed forms (test_oracle_pins.py)
 *   textures             affine and clamp closed forms
 *   integrator           closed-form furnace box; FP64 16k-spp image (level c)
 *   environment light    lat-long orientation closed forms, Sample pdf = PdfDir,
 *                        chi-square of the 2-D inversion, uniform-map sphere
 *                        sampling, constant-sky furnace a Le cos^2(pi / 2h)
 *                        (test_env.py)
 *   light tracing        light-only scene's splat sum Le W H; with a map, that
 *                        times the area lights' share of pdfLight; LT and PT
 *                        estimates agree; importance-mode transmission closed
 *                        forms (test_lt.py)
 *   OBJ, image files     the reference's own tinyobjloader and stb built into
 *                        oracle/_ref (test_obj.py, test_image.py)
 * Parity unpinned: the MSVC CRT's last-ulp behaviour of sin/cos/acos/atan2/
 * log/pow (here: double evaluation rounded once) and the TBB-ordered sums at
 * film border pixels (here: a fixed order).
 *
 * The oracle runs in two precisions:
 *   OR_FP32  bit-faithful float mode (SSE-order dot products, no FMA,
 *            IEEE divide/sqrt; compile with -ffp-contract=off)
 *   OR_FP64  the same algorithm in double, used for the converged reference
 */
#ifndef GY_ORACLE_H
#define GY_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material record: 32 four-byte words (the same layout the library's
 * gy_material uses).  Words:
 *   0 type (0 Lambert, 1 GGX conductor, 2 smooth dielectric, 3 rough dielectric)
 *   1 flags (bit0: alpha given directly, else RoughnessToAlpha(roughness);
 *            bit1: Schlick F0 Fresnel for the conductor, else FrConductor(eta,k))
 *   2-4 R (Lambert kd / conductor reflectance / dielectric Kr)
 *   5-7 T (dielectric Kt)
 *   8-9 alpha_x, alpha_y (or roughness_x, roughness_y when flag bit0 is clear)
 *   10-12 conductor eta, 13-15 conductor k, 16-18 Schlick F0
 *   19 dielectric eta1 (outside), 20 eta2 (inside)
 *   21 albedo texture id (int32, -1 none), 22 roughness texture id (-1 none)
 */
#define OR_MAT_WORDS 32

typedef struct {
    int32_t n_tri;
    const float* pos;      /* n_tri*9: p0,p1,p2 */
    const float* nrm;      /* n_tri*9: n0,n1,n2 (vertex normals) */
    const float* face_n;   /* n_tri*3: face normal as given to Triangle() */
    const float* uv;       /* n_tri*6 or NULL: per-vertex uv (NULL = all zero) */
    const uint32_t* mat;   /* n_tri material ids */
    const int32_t* light;  /* n_tri light ids (-1 = not emissive) */
    int32_t n_mat;
    const uint32_t* mats;  /* n_mat*32 words */
    int32_t n_light;
    const int32_t* light_tri; /* n_light */
    const float* light_le;    /* n_light*3 */
    int32_t n_tex;
    const int32_t* tex_info;  /* n_tex*4: width, height, channels, float offset */
    const float* tex_data;
} or_scene_desc;

typedef struct {
    float pos[3], look_at[3], up[3];
    float fovy;        /* degrees */
    float film_dist;   /* 1 in every reference scene */
    int32_t width, height;
} or_camera_desc;

typedef struct {
    int32_t spp;
    int32_t max_depth;
    int32_t rr_depth;    /* Russian roulette when bounce i > rr_depth (reference: 5) */
    int32_t precision;   /* 0 = FP32 bit-faithful, 1 = FP64 */
    int32_t nthreads;    /* tile worker threads (0 = hardware concurrency) */
    int32_t use_bvh;     /* 1 = reference BVH traversal, 0 = brute force */
    int32_t sampler;     /* 0 = SobolSampler + SobolEnum, 1 = HaltonSampler (Faure) + HaltonEnum */
    int32_t integrator;  /* 0 = PathTracing, 1 = VolPathTracing (free-flight sampling), 2 = VolPathTracing (equi-angular) */
} or_render_params;

/* Sobol tables (fixtures extracted from the reference's data tables). */
int or_load_sobol(const char* golden_dir);
float or_sobol_sample(uint64_t index, uint32_t dim);
uint64_t or_sobol_enum_index(uint32_t m, uint32_t frame, uint32_t px, uint32_t py);

/* Camera: RasterToWorld (column-major 4x4, 16 floats, element (i,j) at [j*4+i])
 * and GenerateRay for raster sample positions. */
int or_camera_matrix(const or_camera_desc* cam, float* raster_to_world16);
int or_camera_rays(const or_camera_desc* cam, int32_t n, const int32_t* px, const int32_t* py,
                   const float* u, const float* v, float* org, float* dir);

/* An analytic sphere (shape/Sphere.h:10-351, identity transform), as gy_sphere. */
typedef struct {
    float center[3];
    float radius;
    int32_t inside;    /* normals point inwards */
    uint32_t material;
    int32_t emissive;  /* an AreaLight on the sphere, after the existing lights */
    float le[3];
} or_sphere;

void* or_scene_create(const or_scene_desc* desc);
/* Appends spheres (primitive ids after the triangles); before any environment light. */
int or_scene_set_spheres(void* scene, int32_t n, const or_sphere* spheres);
/* Participating media for VolPathTracing: HomogeneousMedium(sigma_a, sigma_s, g)
 * (volume/Homogeneous.h:9-24) and, per triangle, its MediumInterface(inside,
 * outside) (-1: none) and whether it has no material (a medium boundary). */
typedef struct {
    float sigma_a[3];
    float sigma_s[3];
    float g;
} or_medium;
typedef struct {
    int32_t inside, outside, no_material;
} or_tri_medium;
int or_scene_set_media(void* scene, int32_t n_media, const or_medium* media, int32_t n_tri, const or_tri_medium* tri);
/* Test hooks (precision 0: FP32, 1: FP64).  HenyeyGreenstein::Sample_p(wo, u):
 * wi[3], p[0] = the returned value, p[1] = p(wo, wi).  Medium probe for one ray
 * (origin as given, no RayEps shift), kind 0: Sample (u = channel, distance),
 * 1: EquiAngularSampling towards light_pos (u[0]), 2: Tr; out = t, sampled,
 * weight[3], pdf (kind 1). */
/* Test-only: 1 = Scene::IntersectTr queries each crossed material-less surface
 * and continues in the medium beyond its normal (0 = the reference's rule). */
int or_set_vpt_query_crossings(int on);
/* SPPM(iterations, photons, maxDepth, initialRadius, alpha, batchShrink, sampler)
 * (integrator/SPPM.h:48-49), traceGlossyRay off; triangle scenes with area
 * lights.  color = W x H x 3 (Film::SetImage's image); radius2 and photons
 * (optional) = each pixel's final radius^2 and photon count N. */
typedef struct {
    int32_t iterations;
    int32_t max_depth;
    int64_t photons;        /* per iteration */
    float initial_radius;
    float alpha;
    int32_t batch_shrink;
    int32_t sampler;        /* 0 Sobol, 1 Halton */
    int32_t precision;      /* 0 FP32, 1 FP64 */
    int32_t nthreads;       /* eye pass threads (0: all) */
} or_sppm_params;
int or_render_sppm(void* scene, const or_camera_desc* cam, const or_sppm_params* p, double* color, double* radius2, double* photons);
/* Test hook: HashGrid over n points (all used, radii^2 r2); per query point the
 * bucket it looks up: counts[nq], then the point indices in idx (duplicates
 * kept; idx NULL: counts only). */
int or_sppm_grid_query(int32_t n, const float* pos, const float* r2, int32_t nq, const float* q, int32_t* counts, int64_t* idx,
                       int64_t cap);
/* Test hook: Scene::IntersectTr from o along d (origin moved RayEps by the Ray
 * constructor) starting in medium (-1: none); out = hit primitive (-1), Tr[3]. */
int or_vpt_intersect_tr(void* scene, int precision, const double* o, const double* d, int32_t medium, double* out);
int or_hg_probe(int precision, double g, const double* wo, const double* u, double* wi, double* p);
int or_medium_probe(int precision, int kind, const double* sigma_a, const double* sigma_s, const double* o, const double* d,
                    double tmax, const double* u, const double* light_pos, double* out);
/* Test hook: from a shading point at ref (zero error bound), sphere i's light
 * sample at u = (u0, u1) and Pdf_Li towards wi; out = pos[3], n[3], pdf, pdf_li.
 * precision 0: FP32, 1: FP64. */
int or_sphere_light_probe(void* scene, int32_t i, int precision, const double* ref, const double* u, const double* wi,
                          double* out);
void or_scene_destroy(void* scene);
/* EnvironmentLight appended after the area lights: rgb is the lat-long image,
 * top row first, 3 floats a texel (as stbi_loadf returns it). */
int or_scene_set_environment(void* scene, const float* rgb, int32_t w, int32_t h, float scale);
/* kind 0: LookupRadiance(a, b); 1: Sample(a, b) -> rgb, dir, pdf; 2: PdfDir((a, b, c));
 * 3: Emission((a, b, c)). */
int or_env_query(void* scene, int32_t n, const float* a, const float* b, const float* c, float* out_rgb,
                 float* out_dir, float* out_pdf, int kind);
int or_env_info(void* scene, float* radius, float* light_func, int32_t max_lights);
/* LightTracing (integrator/LT.h) in float mode: splat = W x H x 3 sums of
 * Film::AddSplat.  -1: the scene has no light. */
int or_render_lt(void* scene, const or_camera_desc* cam, const or_render_params* p, float* splat);
/* Reference BVH statistics: node count, max depth. */
int or_bvh_stats(void* scene, int64_t* n_nodes, int32_t* max_depth);

/* Closest / any hit for n rays given as org (3n), dir (3n), tmin (n), tmax (n).
 * Closest: out_t (n), out_prim (n, -1 miss), out_b1, out_b2.  use_bvh selects
 * the reference BVH traversal (1) or brute force in scene order (0). */
int or_trace_closest(void* scene, int32_t n, const float* org, const float* dir, const float* tmin,
                     const float* tmax, int use_bvh, float* out_t, int32_t* out_prim, float* out_b1,
                     float* out_b2);
int or_trace_any(void* scene, int32_t n, const float* org, const float* dir, const float* tmin,
                 const float* tmax, int use_bvh, int32_t* out_hit);
/* Count BVH nodes and triangles visited by the reference closest-hit traversal. */
int or_trace_counts(void* scene, int32_t n, const float* org, const float* dir, const float* tmin,
                    const float* tmax, int64_t* n_node, int64_t* n_tri);

/* Per-sample radiance Li for n (x, y, s) triples (the path a tile worker
 * would trace for sample s of pixel (x, y)).  out_L: 3n doubles.
 * out_raster: 2n floats, the film position (x+u, y+v) of each sample.
 * out_dims: n ints, the number of sample dimensions the path consumed. */
int or_li_samples(void* scene, const or_camera_desc* cam, const or_render_params* p, int32_t n,
                  const int32_t* x, const int32_t* y, const int32_t* s, double* out_L,
                  float* out_raster, int32_t* out_dims);

/* TiledIntegrator::Render into a film: color (W*H*3) and weight (W*H) sums.
 * rows [row0, row1) only (row1 <= 0 means all rows) so a bounded sample of the
 * image can be timed.  Returns paths traced; *seconds gets wall time. */
int64_t or_render(void* scene, const or_camera_desc* cam, const or_render_params* p, int32_t row0,
                  int32_t row1, double* color, double* weight, double* seconds);

/* Test pins: 4x4 GLM inverse and SSE-order mat*vec (row-major in/out),
 * BSDF evaluation / sampling in the local frame (double), Fresnel terms and
 * the GGX distribution. */
int or_mat4_inverse(const float* rows16, float* out_rows16);
int or_mat4_mul_vec(const float* rows16, const float* v4, float* out4);
int or_bsdf_eval(const uint32_t* mat_words, int32_t n, const double* wo, const double* wi, double* f, double* pdf);
int or_bsdf_sample(const uint32_t* mat_words, int32_t n, const double* wo, const double* u, double* wi, double* f, double* pdf);
/* The same in TransportMode::Importance when importance != 0 (light paths). */
int or_bsdf_sample_mode(const uint32_t* mat_words, int32_t n, const double* wo, const double* u, double* wi, double* f,
                        double* pdf, int importance);
double or_fr_dielectric(double cosI, double etaI, double etaT);
int or_fr_conductor(double cosI, const double* eta, const double* k, double* out);
/* GGXDistribution::RoughnessToAlpha (visual/Microfacet.h:49-54) in double. */
double or_roughness_to_alpha(double r);
double or_ggx_D(double ax, double ay, const double* m);
int or_tex_lookup(int w, int h, int c, const float* data, int n, const float* u, const float* v, int ch, float* out);
/* Test-only switch: end NEE shadow segments ShadowRayEps before the light point
 * (the reference's segment reaches it, see spawn_to in gy_oracle.cpp). */
int or_set_shadow_short(int on);
/* Test-only detached evaluation: paths keep being sampled with the scene's own
 * texels and emitted radiance, while BSDF values and emission are taken from
 * tex_eval (the layout of tex_data) and le_eval (3 floats per light); NULL
 * keeps the scene's own; both NULL switches it off.  Differences of renders
 * in this mode are what a path-replay adjoint estimates. */
int or_set_detached(void* scene, const float* tex_eval, const float* le_eval);
/* HaltonSampler::sample (Faure) and HaltonEnum::get_index; Faure permutation of a base. */
float or_halton_sample(uint64_t index, uint32_t dim);
uint64_t or_halton_enum_index(uint32_t w, uint32_t h, uint64_t i, uint32_t x, uint32_t y);
int or_faure_perm(uint32_t base, uint16_t* out);
int or_li_path(void* scene, const or_camera_desc* cam, const or_render_params* p, int x, int y, int smp, int max_rows,
               double* rows);

#ifdef __cplusplus
}
#endif
#endif
