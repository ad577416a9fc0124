// HipPathTracing.h - the MI355X path-tracing integrator behind Gyunity's
// Integrator interface (header-only, on top of the C ABI in gyunity.h).
//
// Mirrors, name for name, what the reference's render loop touches:
//   class Integrator { virtual void Render(const Scene&, const Camera&) = 0; }
//                                             (integrator/Integrator.h:17-27)
//   PathTracing(spp, maxDepth, pSampler, pSamplerEnum)   (integrator/PT.h:12)
//   Film{mResX, mResY}, pixel = color / weight           (visual/Film.h:20-136)
//   PinHoleCamera(film, position, lookAt, up, fovy, filmDistance)
//                                             (camera/PinholeCamera.h:16-20)
//   Scene::AddPrimitive / AddLight / Initialize          (visual/Scene.h:26-65)
//   SobolSampler, SobolEnum(film)          (sampler/Sobol.h, sampler/SobolEnum.h)
//
// The Scene here is the flattened form the device reads: triangles with
// vertex/face normals and uvs, a material table, area lights by triangle.  An
// adapter from Gyunity's own Scene fills it from GetPrimitives()/GetLights().
//
// Errors: every failing C call throws gy::Error with gy_last_error()'s text
// (the reference logs and exits instead, common/Core.h:160-166).
#pragma once

#include <gyunity.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace gy {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void Check(int rc, const char* where) {
    if (rc != GY_OK) throw Error(rc, std::string(where) + ": " + gy_last_error());
}

using Vec3 = std::array<float, 3>;

// ---------------------------------------------------------------- device
// One HIP device and stream.  Not re-entrant: one render at a time.
class Device {
public:
    explicit Device(int ordinal = 0) { Check(gy_open(ordinal, &ctx_), "gy_open"); }
    ~Device() { gy_close(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    gy_context* Handle() const { return ctx_; }
    std::string Name() const {
        char buf[256] = {0};
        Check(gy_device_name(ctx_, buf, sizeof(buf)), "gy_device_name");
        return buf;
    }
    gy_render_stats Stats() const {
        gy_render_stats s;
        Check(gy_render_stats_get(ctx_, &s), "gy_render_stats_get");
        return s;
    }

private:
    gy_context* ctx_ = nullptr;
};

// ---------------------------------------------------------------- film
class Film {
public:
    Film(int w, int h)
        : mResX(w), mResY(h), color((size_t)w * h * 3, 0.f), weight((size_t)w * h, 0.f), splat((size_t)w * h * 3, 0.f) {}
    int mResX, mResY;
    std::vector<float> color;   // rgb sums, row-major
    std::vector<float> weight;  // box-filter weights
    std::vector<float> splat;   // Film::AddSplat sums (light tracing), rgb
    // Pixel value as Film::WriteToPixelBuffer forms it: color / weight + splat
    // (visual/Film.h:128-135).
    Vec3 Pixel(int x, int y) const {
        const size_t i = (size_t)y * mResX + x;
        const float w = weight[i];
        return {color[3 * i] / w + splat[3 * i], color[3 * i + 1] / w + splat[3 * i + 1], color[3 * i + 2] / w + splat[3 * i + 2]};
    }
};

// ---------------------------------------------------------------- camera
class Camera {
public:
    virtual ~Camera() = default;
    std::shared_ptr<Film> GetFilm() const { return mpFilm; }
    virtual gy_camera_desc Desc() const = 0;

protected:
    explicit Camera(std::shared_ptr<Film> f) : mpFilm(std::move(f)) {}
    std::shared_ptr<Film> mpFilm;
};

class PinHoleCamera : public Camera {
public:
    PinHoleCamera(std::shared_ptr<Film> film, const Vec3& position, const Vec3& lookAt, const Vec3& up, float fovy,
                  float filmDistance = 1.f)
        : Camera(std::move(film)), mPos(position), mLookAt(lookAt), mUp(up), mFovy(fovy), mFilmDistance(filmDistance) {}
    gy_camera_desc Desc() const override {
        gy_camera_desc d;
        for (int k = 0; k < 3; ++k) {
            d.pos[k] = mPos[k];
            d.look_at[k] = mLookAt[k];
            d.up[k] = mUp[k];
        }
        d.fovy = mFovy;
        d.film_dist = mFilmDistance;
        d.width = mpFilm->mResX;
        d.height = mpFilm->mResY;
        return d;
    }
    Vec3 mPos, mLookAt, mUp;
    float mFovy, mFilmDistance;
};

// ---------------------------------------------------------------- samplers
// The device implements SobolSampler + SobolEnum (unscrambled,
// sampler/SobolEnum.h:24-25) and HaltonSampler with Faure permutations +
// HaltonEnum (sampler/Halton.h, sampler/HaltonEnum.h).  The objects carry the
// choice; the sampler decides it.
class Sampler {
public:
    virtual ~Sampler() = default;
    virtual int Kind() const = 0;  // GY_SAMPLER_*
};
class SobolSampler : public Sampler {
public:
    int Kind() const override { return GY_SAMPLER_SOBOL; }
};
class HaltonSampler : public Sampler {
public:
    HaltonSampler(bool faure = true) {
        if (!faure)  // init_random shuffles with the STL engine (Halton.h:627-646): not reproducible
            throw Error(GY_ERR_UNSUPPORTED, "HaltonSampler: only Faure permutations are implemented");
    }
    int Kind() const override { return GY_SAMPLER_HALTON; }
};
class SamplerEnum {
public:
    virtual ~SamplerEnum() = default;
};
class SobolEnum : public SamplerEnum {
public:
    explicit SobolEnum(const Film&) {}
};
class HaltonEnum : public SamplerEnum {
public:
    HaltonEnum(uint32_t, uint32_t) {}
};

// ---------------------------------------------------------------- materials
struct Material {
    std::array<uint32_t, GY_MAT_WORDS> words{};
    static uint32_t Bits(float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    }
    static void Put3(uint32_t* w, const Vec3& v) {
        for (int k = 0; k < 3; ++k) w[k] = Bits(v[k]);
    }
    // DiffuseMaterial(R) (material/Diffuse.h:27-31)
    static Material Diffuse(const Vec3& R, int albedoTex = -1) {
        Material m;
        m.words[0] = GY_MAT_LAMBERT;
        Put3(&m.words[2], R);
        m.words[21] = (uint32_t)albedoTex;
        m.words[22] = (uint32_t)-1;
        return m;
    }
    // RoughnessMaterial(R, roughness, eta, k) (material/Roughness.h:36-46)
    static Material Conductor(const Vec3& R, float roughness, const Vec3& eta, const Vec3& k, int roughnessTex = -1) {
        Material m;
        m.words[0] = GY_MAT_CONDUCTOR;
        Put3(&m.words[2], R);
        m.words[8] = m.words[9] = Bits(roughness);
        Put3(&m.words[10], eta);
        Put3(&m.words[13], k);
        m.words[21] = (uint32_t)-1;
        m.words[22] = (uint32_t)roughnessTex;
        return m;
    }
    // GGX conductor with alpha given directly and Schlick F0 (benchmark material).
    static Material ConductorF0(const Vec3& R, float alpha, const Vec3& F0) {
        Material m;
        m.words[0] = GY_MAT_CONDUCTOR;
        m.words[1] = GY_MATF_ALPHA_DIRECT | GY_MATF_SCHLICK_F0;
        Put3(&m.words[2], R);
        m.words[8] = m.words[9] = Bits(alpha);
        Put3(&m.words[16], F0);
        m.words[21] = m.words[22] = (uint32_t)-1;
        return m;
    }
    // GlassMaterial(R, T, eta) (material/Glass.h:30-34)
    static Material Glass(const Vec3& R, const Vec3& T, float etaOut = 1.f, float etaIn = 1.5f) {
        Material m;
        m.words[0] = GY_MAT_GLASS;
        Put3(&m.words[2], R);
        Put3(&m.words[5], T);
        m.words[19] = Bits(etaOut);
        m.words[20] = Bits(etaIn);
        m.words[21] = m.words[22] = (uint32_t)-1;
        return m;
    }
    // RoughDielectricMaterial(R, T, roughness, eta) (material/RoughDielectric.h:21-41)
    static Material RoughDielectric(const Vec3& R, const Vec3& T, float roughness, float etaOut = 1.f, float etaIn = 1.5f) {
        Material m = Glass(R, T, etaOut, etaIn);
        m.words[0] = GY_MAT_ROUGH_DIELECTRIC;
        m.words[8] = m.words[9] = Bits(roughness);
        return m;
    }
};

// ---------------------------------------------------------------- scene
// A process-unique stamp: every Scene object, copy and change gets a new one,
// so a device copy cached under it is never taken for another scene, even one
// built at the same address.
struct SceneStamp {
    uint64_t v = Next();
    SceneStamp() = default;
    SceneStamp(const SceneStamp&) : v(Next()) {}
    SceneStamp& operator=(const SceneStamp&) {
        v = Next();
        return *this;
    }
    static uint64_t Next() {
        static std::atomic<uint64_t> n{0};
        return ++n;
    }
};

class Scene {
public:
    // Scene::AddPrimitive for one triangle (primitive id = insertion order,
    // visual/Scene.h:43).  Returns the primitive id.
    int AddTriangle(const std::array<Vec3, 3>& p, const std::array<Vec3, 3>& n, const Vec3& faceNormal, int material,
                    const std::array<float, 6>& uv = {0, 0, 0, 0, 0, 0}) {
        Touch();
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) {
                pos.push_back(p[v][k]);
                nrm.push_back(n[v][k]);
            }
        for (int k = 0; k < 3; ++k) face_n.push_back(faceNormal[k]);
        for (float u : uv) uvs.push_back(u);
        mat.push_back((uint32_t)material);
        light.push_back(-1);
        return (int)mat.size() - 1;
    }
    int AddMaterial(const Material& m) {
        Touch();
        mats.insert(mats.end(), m.words.begin(), m.words.end());
        return (int)(mats.size() / GY_MAT_WORDS) - 1;
    }
    // Scene::AddLight(AreaLight(shape = triangle, Le)) (light/AreaLight.h:12-85)
    int AddAreaLight(int triangle, const Vec3& Le) {
        Touch();
        light[(size_t)triangle] = (int32_t)light_tri.size();
        light_tri.push_back(triangle);
        for (float v : Le) light_le.push_back(v);
        return (int)light_tri.size() - 1;
    }
    // Scene::AddLight(EnvironmentLight(file, rotate, scale)) after the area
    // lights (light/EnvLight.h:14-17): the latitude-longitude image, top row
    // first, w x h x 3 floats.
    void SetEnvironmentLight(int w, int h, const std::vector<float>& rgb, float scale = 1.f) {
        Touch();
        env_w = w;
        env_h = h;
        env_rgb = rgb;
        env_scale = scale;
    }
    // ImageTexture (texture/ImageTexture.h:50-69): width x height x channels floats.
    int AddTexture(int w, int h, int channels, const std::vector<float>& data) {
        Touch();
        tex_info.insert(tex_info.end(), {w, h, channels, (int32_t)tex_data.size()});
        tex_data.insert(tex_data.end(), data.begin(), data.end());
        return (int)(tex_info.size() / 4) - 1;
    }
    // Sphere(Transform(), Transform(), radius, position, inside) with a
    // material (shape/Sphere.h:21-27), after every triangle; emissive: an
    // AreaLight on it with emitted radiance Le, after the triangle lights.
    int AddSphere(const Vec3& center, float radius, int material, bool inside = false, bool emissive = false,
                  const Vec3& Le = Vec3{0, 0, 0}) {
        Touch();
        gy_sphere q = {};
        for (int k = 0; k < 3; ++k) {
            q.center[k] = center[k];
            q.le[k] = Le[k];
        }
        q.radius = radius;
        q.inside = inside ? 1 : 0;
        q.material = (uint32_t)material;
        q.emissive = emissive ? 1 : 0;
        spheres.push_back(q);
        return (int)(mat.size() + spheres.size()) - 1;
    }
    // HomogeneousMedium(sigma_a, sigma_s, scale 1, g) (volume/Homogeneous.h:11-12).
    // Returns the medium id.
    int AddMedium(const Vec3& sigma_a, const Vec3& sigma_s, float g) {
        Touch();
        gy_medium m = {};
        for (int k = 0; k < 3; ++k) {
            m.sigma_a[k] = sigma_a[k];
            m.sigma_s[k] = sigma_s[k];
        }
        m.g = g;
        media.push_back(m);
        return (int)media.size() - 1;
    }
    // The primitive's MediumInterface(inside, outside) (visual/Medium.h:51-59;
    // -1: no medium) and, with noMaterial, a null material: the triangle only
    // bounds a medium (visual/Primitive.cpp:11-14).
    void SetTriangleMedium(int triangle, int inside, int outside, bool noMaterial) {
        Touch();
        if (tri_media.size() < mat.size()) tri_media.resize(mat.size(), gy_tri_medium{-1, -1, 0});
        tri_media[(size_t)triangle] = gy_tri_medium{inside, outside, noMaterial ? 1 : 0};
    }
    // Call after changing the arrays below directly: the next Render uploads again.
    void Touch() { stamp = SceneStamp(); }
    uint64_t Stamp() const { return stamp.v; }

    gy_scene_desc Desc() const {
        gy_scene_desc d = {};
        d.n_tri = (int32_t)mat.size();
        d.pos = pos.data();
        d.nrm = nrm.data();
        d.face_n = face_n.data();
        d.uv = uvs.data();
        d.mat = mat.data();
        d.light = light.data();
        d.n_mat = (int32_t)(mats.size() / GY_MAT_WORDS);
        d.mats = mats.data();
        d.n_light = (int32_t)light_tri.size();
        d.light_tri = light_tri.data();
        d.light_le = light_le.data();
        d.n_tex = (int32_t)(tex_info.size() / 4);
        d.tex_info = tex_info.data();
        d.tex_data = tex_data.data();
        return d;
    }
    // Copy a flattened description (for example a generated benchmark scene).
    static Scene FromDesc(const gy_scene_desc& d) {
        Scene s;
        const size_t n = (size_t)d.n_tri;
        s.pos.assign(d.pos, d.pos + 9 * n);
        s.nrm.assign(d.nrm, d.nrm + 9 * n);
        s.face_n.assign(d.face_n, d.face_n + 3 * n);
        if (d.uv) s.uvs.assign(d.uv, d.uv + 6 * n);
        else s.uvs.assign(6 * n, 0.f);
        s.mat.assign(d.mat, d.mat + n);
        if (d.light) s.light.assign(d.light, d.light + n);
        else s.light.assign(n, -1);
        s.mats.assign(d.mats, d.mats + (size_t)d.n_mat * GY_MAT_WORDS);
        s.light_tri.assign(d.light_tri, d.light_tri + d.n_light);
        s.light_le.assign(d.light_le, d.light_le + 3 * (size_t)d.n_light);
        if (d.n_tex > 0) {
            s.tex_info.assign(d.tex_info, d.tex_info + 4 * (size_t)d.n_tex);
            size_t total = 0;
            for (int i = 0; i < d.n_tex; ++i) {
                const int32_t* t = d.tex_info + 4 * i;
                total = std::max(total, (size_t)t[3] + (size_t)t[0] * t[1] * t[2]);
            }
            s.tex_data.assign(d.tex_data, d.tex_data + total);
        }
        return s;
    }
    size_t TriangleCount() const { return mat.size(); }

    std::vector<float> pos, nrm, face_n, uvs;
    std::vector<uint32_t> mat;
    std::vector<int32_t> light;
    std::vector<uint32_t> mats;
    std::vector<int32_t> light_tri;
    std::vector<float> light_le;
    std::vector<int32_t> tex_info;
    std::vector<float> tex_data;
    int env_w = 0, env_h = 0;  // env_w == 0: no environment light
    std::vector<float> env_rgb;
    float env_scale = 1.f;
    std::vector<gy_sphere> spheres;
    std::vector<gy_medium> media;
    std::vector<gy_tri_medium> tri_media;  // per triangle, or empty: no media

private:
    SceneStamp stamp;
};

// Scene resident on a device: triangles, BVH (built on the GPU), shading data.
class DeviceScene {
public:
    DeviceScene(const Device& dev, const Scene& s, const gy_bvh_params* bvh = nullptr) {
        const gy_scene_desc d = s.Desc();
        Check(gy_scene_create(dev.Handle(), &d, bvh, &h_), "gy_scene_create");
        if (!s.spheres.empty()) {
            const int rc = gy_scene_set_spheres(h_, (int32_t)s.spheres.size(), s.spheres.data());
            if (rc != GY_OK) {
                gy_scene_destroy(h_);
                Check(rc, "gy_scene_set_spheres");
            }
        }
        if (s.env_w > 0) {
            const int rc = gy_scene_set_environment(h_, s.env_rgb.data(), s.env_w, s.env_h, s.env_scale);
            if (rc != GY_OK) {
                gy_scene_destroy(h_);
                Check(rc, "gy_scene_set_environment");
            }
        }
        if (!s.tri_media.empty()) {
            std::vector<gy_tri_medium> tri(s.tri_media);
            tri.resize(s.TriangleCount(), gy_tri_medium{-1, -1, 0});
            const int rc = gy_scene_set_media(h_, (int32_t)s.media.size(), s.media.data(), tri.data());
            if (rc != GY_OK) {
                gy_scene_destroy(h_);
                Check(rc, "gy_scene_set_media");
            }
        }
    }
    ~DeviceScene() { gy_scene_destroy(h_); }
    DeviceScene(const DeviceScene&) = delete;
    DeviceScene& operator=(const DeviceScene&) = delete;
    gy_scene* Handle() const { return h_; }
    gy_scene_stats Stats() const {
        gy_scene_stats st;
        Check(gy_scene_stats_get(h_, &st), "gy_scene_stats_get");
        return st;
    }

private:
    gy_scene* h_ = nullptr;
};

// ---------------------------------------------------------------- integrators
class Integrator {
public:
    virtual void Render(const Scene& scene, const Camera& camera) = 0;
    virtual ~Integrator() {}
};

// Settings the reference fixes in code: Russian roulette starts after bounce
// rrStartDepth (5 in PathTracing::Li, integrator/PT.cpp:53).
struct RenderParams {
    int rrStartDepth = 5;
    int mode = GY_RENDER_WAVEFRONT;  // or GY_RENDER_MEGAKERNEL (same film, bit for bit)
    int samplesPerPass = 0;          // 0 = automatic
    int blockSize = 0;               // megakernel block size, 0 = automatic
    int refillThreshold = 0;         // wavefront traversal refill, 0 = automatic
    int raySort = GY_RAY_SORT_NONE;  // wavefront closest-hit queue order (same film either way)
};

// PathTracing on the GPU.  Render() uploads the scene on first use (or when a
// different Scene object is passed), renders, and fills camera.GetFilm().
class HipPathTracing : public Integrator {
public:
    HipPathTracing(int samplePerPixel, int maxDepth, const std::shared_ptr<Sampler>& pSampler,
                   const std::shared_ptr<SamplerEnum>& pSamplerEnum, RenderParams params = RenderParams(),
                   std::shared_ptr<Device> device = nullptr)
        : spp(samplePerPixel), mMaxDepth(maxDepth), mpSampler(pSampler), mpSamplerEnum(pSamplerEnum), mParams(params),
          mpDevice(device ? std::move(device) : std::make_shared<Device>(0)) {
        if (!mpSampler) throw Error(GY_ERR_INVALID, "HipPathTracing: no sampler");
    }

    void Render(const Scene& scene, const Camera& camera) override {
        if (mSceneStamp != scene.Stamp() || !mpScene) {
            mpScene.reset();
            mpScene.reset(new DeviceScene(*mpDevice, scene));
            mSceneStamp = scene.Stamp();
        }
        const gy_camera_desc cam = camera.Desc();
        gy_render_params p = Params();
        auto film = camera.GetFilm();
        Check(gy_render(mpScene->Handle(), &cam, &p, film->color.data(), film->weight.data()), "gy_render");
    }

    gy_render_params Params() const {
        gy_render_params p = {};
        p.spp = spp;
        p.max_depth = mMaxDepth;
        p.rr_depth = mParams.rrStartDepth;
        p.mode = mParams.mode;
        p.samples_per_pass = mParams.samplesPerPass;
        p.block_size = mParams.blockSize;
        p.refill_threshold = mParams.refillThreshold;
        p.sampler = mpSampler->Kind();
        p.ray_sort = mParams.raySort;
        return p;
    }
    const Device& GetDevice() const { return *mpDevice; }
    const DeviceScene* GetDeviceScene() const { return mpScene.get(); }

private:
    int spp;
    int mMaxDepth;
    std::shared_ptr<Sampler> mpSampler;
    std::shared_ptr<SamplerEnum> mpSamplerEnum;
    RenderParams mParams;
    std::shared_ptr<Device> mpDevice;
    std::unique_ptr<DeviceScene> mpScene;
    uint64_t mSceneStamp = 0;  // the Scene::Stamp() of mpScene
};

// LightTracing on the GPU: LightTracing(pSampler, maxDepth, spp)
// (integrator/LT.h:14-16).  Render() uploads the scene as HipPathTracing does
// and adds the light paths' splats into camera.GetFilm()->splat, as
// Film::AddSplat does (visual/Film.h:89-103).  params.mode picks the form
// (wavefront or one path per thread: the same splats, summed in another
// order); params.blockSize is the per-path form's block size.
class HipLightTracing : public Integrator {
public:
    HipLightTracing(const std::shared_ptr<Sampler>& pSampler, int maxDepth, int spp, RenderParams params = RenderParams(),
                    std::shared_ptr<Device> device = nullptr)
        : mpSampler(pSampler), mMaxDepth(maxDepth), mSpp(spp), mParams(params),
          mpDevice(device ? std::move(device) : std::make_shared<Device>(0)) {
        if (!mpSampler) throw Error(GY_ERR_INVALID, "HipLightTracing: no sampler");
    }

    void Render(const Scene& scene, const Camera& camera) override {
        if (mSceneStamp != scene.Stamp() || !mpScene) {
            mpScene.reset();
            mpScene.reset(new DeviceScene(*mpDevice, scene));
            mSceneStamp = scene.Stamp();
        }
        const gy_camera_desc cam = camera.Desc();
        const gy_render_params p = Params();
        auto film = camera.GetFilm();
        std::vector<float> rgb(film->splat.size());
        Check(gy_render_light_tracing(mpScene->Handle(), &cam, &p, rgb.data()), "gy_render_light_tracing");
        for (size_t i = 0; i < rgb.size(); ++i) film->splat[i] = film->splat[i] + rgb[i];
    }

    gy_render_params Params() const {
        gy_render_params p = {};
        p.spp = mSpp;
        p.max_depth = mMaxDepth;
        p.rr_depth = mParams.rrStartDepth;
        p.mode = mParams.mode;
        p.block_size = mParams.blockSize;
        p.sampler = mpSampler->Kind();
        return p;
    }

private:
    std::shared_ptr<Sampler> mpSampler;
    int mMaxDepth;
    int mSpp;
    RenderParams mParams;
    std::shared_ptr<Device> mpDevice;
    std::unique_ptr<DeviceScene> mpScene;
    uint64_t mSceneStamp = 0;  // the Scene::Stamp() of mpScene
};

// VolPathTracing(spp, maxDepth, sampler, samplerEnum, equiAngular)
// (integrator/VPT.h:12-19) on the GPU: fills camera.GetFilm() as
// TiledIntegrator::Render does.  params.mode picks the wavefront form or one
// path per thread (equal films).
class HipVolPathTracing : public Integrator {
public:
    HipVolPathTracing(int samplePerPixel, int maxDepth, const std::shared_ptr<Sampler>& pSampler,
                      const std::shared_ptr<SamplerEnum>& pSamplerEnum, bool equiAngular = true,
                      RenderParams params = RenderParams(), std::shared_ptr<Device> device = nullptr)
        : spp(samplePerPixel), mMaxDepth(maxDepth), mpSampler(pSampler), mpSamplerEnum(pSamplerEnum), mEquiAngular(equiAngular),
          mParams(params), mpDevice(device ? std::move(device) : std::make_shared<Device>(0)) {
        if (!mpSampler) throw Error(GY_ERR_INVALID, "HipVolPathTracing: no sampler");
    }

    void Render(const Scene& scene, const Camera& camera) override {
        if (mSceneStamp != scene.Stamp() || !mpScene) {
            mpScene.reset();
            mpScene.reset(new DeviceScene(*mpDevice, scene));
            mSceneStamp = scene.Stamp();
        }
        const gy_camera_desc cam = camera.Desc();
        const gy_render_params p = Params();
        auto film = camera.GetFilm();
        Check(gy_render_vpt(mpScene->Handle(), &cam, &p, mEquiAngular ? 1 : 0, film->color.data(), film->weight.data()),
              "gy_render_vpt");
    }

    gy_render_params Params() const {
        gy_render_params p = {};
        p.spp = spp;
        p.max_depth = mMaxDepth;
        p.rr_depth = mParams.rrStartDepth;
        p.mode = mParams.mode;
        p.samples_per_pass = mParams.samplesPerPass;
        p.sampler = mpSampler->Kind();
        return p;
    }

private:
    int spp, mMaxDepth;
    std::shared_ptr<Sampler> mpSampler;
    std::shared_ptr<SamplerEnum> mpSamplerEnum;
    bool mEquiAngular;
    RenderParams mParams;
    std::shared_ptr<Device> mpDevice;
    std::unique_ptr<DeviceScene> mpScene;
    uint64_t mSceneStamp = 0;
};

// SPPM(iterations, nPhotonsPerStage, maxDepth, initialRadius, alpha,
// batchShrink, sampler, samplerEnum) (integrator/SPPM.h:48-49) on the GPU.
// Render sets the film's image as Film::SetImage does: color = the image,
// weight = 1, no splats, so Film::Pixel returns it.
class HipSPPM : public Integrator {
public:
    HipSPPM(int iterations, int64_t photonsPerStage, int maxDepth, float initialRadius, float alpha, bool batchShrink,
            const std::shared_ptr<Sampler>& pSampler, const std::shared_ptr<SamplerEnum>& pSamplerEnum,
            std::shared_ptr<Device> device = nullptr)
        : mpSampler(pSampler), mpSamplerEnum(pSamplerEnum), mpDevice(device ? std::move(device) : std::make_shared<Device>(0)) {
        if (!mpSampler) throw Error(GY_ERR_INVALID, "HipSPPM: no sampler");
        mParams = gy_sppm_params{};
        mParams.iterations = iterations;
        mParams.max_depth = maxDepth;
        mParams.photons_per_iteration = photonsPerStage;
        mParams.initial_radius = initialRadius;
        mParams.alpha = alpha;
        mParams.batch_shrink = batchShrink ? 1 : 0;
        mParams.sampler = mpSampler->Kind();
    }

    void Render(const Scene& scene, const Camera& camera) override {
        if (mSceneStamp != scene.Stamp() || !mpScene) {
            mpScene.reset();
            mpScene.reset(new DeviceScene(*mpDevice, scene));
            mSceneStamp = scene.Stamp();
        }
        const gy_camera_desc cam = camera.Desc();
        auto film = camera.GetFilm();
        Check(gy_render_sppm(mpScene->Handle(), &cam, &mParams, film->color.data(), nullptr, nullptr), "gy_render_sppm");
        std::fill(film->weight.begin(), film->weight.end(), 1.f);
        std::fill(film->splat.begin(), film->splat.end(), 0.f);
    }

    const gy_sppm_params& Params() const { return mParams; }

private:
    std::shared_ptr<Sampler> mpSampler;
    std::shared_ptr<SamplerEnum> mpSamplerEnum;
    gy_sppm_params mParams;
    std::shared_ptr<Device> mpDevice;
    std::unique_ptr<DeviceScene> mpScene;
    uint64_t mSceneStamp = 0;
};

}  // namespace gy
