// Procedural benchmark scenes.
//
//  GY_SCENE_CORNELL: the Cornell box in [0,1]^3 built the way the reference's
//    scene library builds its boxes (Gyunity_exe/CornellBox.h:766-926): two
//    triangles per wall with inward face normals given to the Triangle
//    constructor, a quad light made of two one-sided triangle AreaLights with a
//    black Lambert material, Lambert walls.  Two axis-aligned white boxes of 12
//    triangles each stand on the floor.  Light 0.25 x 0.25 at y = 0.99 facing
//    down, Le = (17, 12, 4).  Pinhole camera, fovy 40 deg, 256^2, 16 spp,
//    depth 5, Russian roulette for bounces i > 3.
//  GY_SCENE_ICOSPHERE_GGX: the same room without the boxes, plus an icosphere
//    subdivided `param` times (8 -> 1,310,720 triangles), radius 0.3, displaced
//    radially by 4-octave value noise of amplitude 0.05 seeded by `seed`, with
//    flat (face) normals.  GGX conductor, alpha 0.2 given directly, Schlick F0
//    (0.95, 0.64, 0.54).  1024^2, 64 spp, depth 8, roulette for i > 5.
//  GY_SCENE_BATCH_MEMBER: member `seed` of the 64-scene batch workload.  The
//    room and light of the Cornell box; a displaced icosphere whose triangle
//    count is drawn log-uniform in [1e3, 2e6] (the first N triangles of the
//    smallest subdivision level that has N); 0-20 random boxes; every object is
//    a GGX conductor (alpha U[0.05, 0.8], Schlick F0) or, with probability 1/4,
//    a rough dielectric of IOR 1.5 (alpha U[0.05, 0.8]); resolution from
//    {128, 256, 512, 1024}^2, spp from {4, 8, 16, 32, 64}, depth U{4..12},
//    roulette for i > 5.  param > 0 shrinks a member for tests: at most `param`
//    triangles, resolution / 8 (at least 16), spp <= 4.
//  GY_SCENE_ADJOINT: the room with a seeded RGB albedo texture on every wall
//    (param texels per side, default 64) and a central 20,480-triangle sphere
//    of GGX conductor with a seeded roughness texture in [0.05, 0.8]; 256^2,
//    64 spp, depth 5, roulette for i > 3.
//  GY_SCENE_ICOSPHERE_TEXTURED: the GGX icosphere scene (8 subdivisions) with
//    an RGB albedo and a roughness texture on the sphere (param texels per
//    side, default 1024: 4.2M texture floats); 512^2, 32 spp, depth 6.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "gy_internal.h"

namespace gy {

struct Generated {
    std::vector<float> pos, nrm, fn, uv;
    std::vector<int32_t> texi;
    std::vector<float> texd;
    std::vector<uint32_t> mat, mats;
    std::vector<int32_t> light, ltri;
    std::vector<float> le;
    gy_camera_desc cam;
    gy_render_params params = {};
};

namespace {

struct V {
    float x, y, z;
};
inline V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V norm(V a) {
    float l = std::sqrt((a.x * a.x + a.y * a.y) + (a.z * a.z + 0.f));
    return {a.x / l, a.y / l, a.z / l};
}

uint32_t AddMaterial(Generated& g, uint32_t type, uint32_t flags, const float R[3]) {
    uint32_t w[GY_MAT_WORDS];
    std::memset(w, 0, sizeof(w));
    w[0] = type;
    w[1] = flags;
    std::memcpy(&w[2], R, 12);
    const float one = 1.f, eta_out = 1.f, eta_in = 1.5f;
    for (int c = 0; c < 3; ++c) std::memcpy(&w[5 + c], &one, 4);
    std::memcpy(&w[19], &eta_out, 4);
    std::memcpy(&w[20], &eta_in, 4);
    int32_t none = -1;
    std::memcpy(&w[21], &none, 4);
    std::memcpy(&w[22], &none, 4);
    g.mats.insert(g.mats.end(), w, w + GY_MAT_WORDS);
    return (uint32_t)(g.mats.size() / GY_MAT_WORDS - 1);
}

void SetWord(Generated& g, uint32_t m, int word, float v) { std::memcpy(&g.mats[(size_t)m * GY_MAT_WORDS + word], &v, 4); }

void AddTri(Generated& g, V p0, V p1, V p2, V n, uint32_t mat, int32_t light) {
    const V p[3] = {p0, p1, p2};
    for (int k = 0; k < 3; ++k) {
        g.pos.push_back(p[k].x);
        g.pos.push_back(p[k].y);
        g.pos.push_back(p[k].z);
        g.nrm.push_back(n.x);
        g.nrm.push_back(n.y);
        g.nrm.push_back(n.z);
    }
    g.fn.push_back(n.x);
    g.fn.push_back(n.y);
    g.fn.push_back(n.z);
    g.mat.push_back(mat);
    g.light.push_back(light);
    for (int k = 0; k < 6; ++k) g.uv.push_back(0.f);
}
void SetLastUV(Generated& g, float u0, float v0, float u1, float v1, float u2, float v2) {
    float* t = &g.uv[g.uv.size() - 6];
    t[0] = u0; t[1] = v0; t[2] = u1; t[3] = v1; t[4] = u2; t[5] = v2;
}
int32_t AddTexture(Generated& g, int w, int h, int c, const std::vector<float>& data) {
    g.texi.insert(g.texi.end(), {w, h, c, (int32_t)g.texd.size()});
    g.texd.insert(g.texd.end(), data.begin(), data.end());
    return (int32_t)(g.texi.size() / 4 - 1);
}

// Quad a-b-c-d as triangles (a, b, c) and (a, c, d), the split the reference's
// boxes use for their light (Gyunity_exe/CornellBox.h:908-913).
void AddQuad(Generated& g, V a, V b, V c, V d, V n, uint32_t mat) {
    AddTri(g, a, b, c, n, mat, -1);
    AddTri(g, a, c, d, n, mat, -1);
}

void AddBox(Generated& g, V lo, V hi, uint32_t mat) {
    const float x0 = lo.x, y0 = lo.y, z0 = lo.z, x1 = hi.x, y1 = hi.y, z1 = hi.z;
    AddQuad(g, {x0, y1, z0}, {x0, y1, z1}, {x1, y1, z1}, {x1, y1, z0}, {0, 1, 0}, mat);    // top
    AddQuad(g, {x0, y0, z0}, {x1, y0, z0}, {x1, y0, z1}, {x0, y0, z1}, {0, -1, 0}, mat);   // bottom
    AddQuad(g, {x0, y0, z0}, {x0, y0, z1}, {x0, y1, z1}, {x0, y1, z0}, {-1, 0, 0}, mat);   // -x
    AddQuad(g, {x1, y0, z0}, {x1, y1, z0}, {x1, y1, z1}, {x1, y0, z1}, {1, 0, 0}, mat);    // +x
    AddQuad(g, {x0, y0, z0}, {x0, y1, z0}, {x1, y1, z0}, {x1, y0, z0}, {0, 0, -1}, mat);   // -z
    AddQuad(g, {x0, y0, z1}, {x1, y0, z1}, {x1, y1, z1}, {x0, y1, z1}, {0, 0, 1}, mat);    // +z
}

void AddRoomAndLight(Generated& g, bool boxes) {
    const float white[3] = {0.73f, 0.73f, 0.73f}, red[3] = {0.63f, 0.065f, 0.05f}, green[3] = {0.14f, 0.45f, 0.091f},
                black[3] = {0.f, 0.f, 0.f};
    uint32_t mw = AddMaterial(g, GY_MAT_LAMBERT, 0, white);
    uint32_t mr = AddMaterial(g, GY_MAT_LAMBERT, 0, red);
    uint32_t mg = AddMaterial(g, GY_MAT_LAMBERT, 0, green);
    uint32_t mlight = AddMaterial(g, GY_MAT_LAMBERT, 0, black);
    // walls: inward normals
    AddQuad(g, {0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, mw);    // floor
    AddQuad(g, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}, {0, -1, 0}, mw);   // ceiling
    AddQuad(g, {0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 1}, mw);    // back
    AddQuad(g, {0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 0, 0}, mr);    // left (x = 0)
    AddQuad(g, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, {-1, 0, 0}, mg);   // right (x = 1)
    if (boxes) {
        AddBox(g, {0.53f, 0.f, 0.45f}, {0.83f, 0.3f, 0.75f}, mw);
        AddBox(g, {0.17f, 0.f, 0.15f}, {0.47f, 0.6f, 0.45f}, mw);
    }
    // light: two one-sided triangles facing down
    const float y = 0.99f, a = 0.375f, b = 0.625f;
    const V l0{b, y, b}, l1{b, y, a}, l2{a, y, a}, l3{a, y, b};
    const V ln{0, -1, 0};
    int32_t first = (int32_t)g.ltri.size();
    AddTri(g, l0, l1, l2, ln, mlight, first);
    g.ltri.push_back((int32_t)(g.fn.size() / 3 - 1));
    AddTri(g, l0, l2, l3, ln, mlight, first + 1);
    g.ltri.push_back((int32_t)(g.fn.size() / 3 - 1));
    for (int i = 0; i < 2; ++i) {
        g.le.push_back(17.f);
        g.le.push_back(12.f);
        g.le.push_back(4.f);
    }
}

void SetCamera(Generated& g, int res) {
    gy_camera_desc& c = g.cam;
    c.pos[0] = 0.5f; c.pos[1] = 0.5f; c.pos[2] = 2.35f;
    c.look_at[0] = 0.5f; c.look_at[1] = 0.5f; c.look_at[2] = 0.f;
    c.up[0] = 0.f; c.up[1] = 1.f; c.up[2] = 0.f;
    c.fovy = 40.f;
    c.film_dist = 1.f;
    c.width = res;
    c.height = res;
}

// 3-D value noise: hashed lattice values in [-1, 1], smoothstep-weighted
// trilinear interpolation; fbm of 4 octaves (frequency x2, amplitude x0.5).
uint32_t Hash(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    uint32_t h = seed * 0x9E3779B1u;
    h ^= (uint32_t)x * 0x85EBCA77u;
    h = (h << 13) | (h >> 19);
    h ^= (uint32_t)y * 0xC2B2AE3Du;
    h = (h << 17) | (h >> 15);
    h ^= (uint32_t)z * 0x27D4EB2Fu;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}
double Lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    return (double)Hash(x, y, z, seed) * (2.0 / 4294967295.0) - 1.0;
}
double ValueNoise(double x, double y, double z, uint32_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    double tx = x - fx, ty = y - fy, tz = z - fz;
    auto s = [](double t) { return t * t * (3.0 - 2.0 * t); };
    double sx = s(tx), sy = s(ty), sz = s(tz);
    double acc = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                double w = (dx ? sx : 1.0 - sx) * (dy ? sy : 1.0 - sy) * (dz ? sz : 1.0 - sz);
                acc += w * Lattice(ix + dx, iy + dy, iz + dz, seed);
            }
    return acc;
}
double Fbm(double x, double y, double z, uint32_t seed) {
    double sum = 0.0, amp = 1.0, freq = 4.0, norm_ = 0.0;
    for (int o = 0; o < 4; ++o) {
        sum += amp * ValueNoise(x * freq, y * freq, z * freq, seed + 7919u * (uint32_t)o);
        norm_ += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm_;
}

void AddIcosphere(Generated& g, int levels, uint32_t seed, double cx, double cy, double cz, double radius, double amp,
                  uint32_t mat, size_t max_tris = (size_t)-1, bool uvs = false) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<double> vx = {-1, 1, -1, 1, 0, 0, 0, 0, t, t, -t, -t};
    std::vector<double> vy = {t, t, -t, -t, -1, 1, -1, 1, 0, 0, 0, 0};
    std::vector<double> vz = {0, 0, 0, 0, t, t, -t, -t, -1, 1, -1, 1};
    for (size_t i = 0; i < vx.size(); ++i) {
        double l = std::sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
        vx[i] /= l; vy[i] /= l; vz[i] /= l;
    }
    std::vector<uint32_t> faces = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                                   3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int lv = 0; lv < levels; ++lv) {
        std::unordered_map<uint64_t, uint32_t> mid;
        mid.reserve(faces.size());
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double x = vx[a] + vx[b], y = vy[a] + vy[b], z = vz[a] + vz[b];
            double l = std::sqrt(x * x + y * y + z * z);
            vx.push_back(x / l); vy.push_back(y / l); vz.push_back(z / l);
            uint32_t idx = (uint32_t)vx.size() - 1;
            mid.emplace(key, idx);
            return idx;
        };
        std::vector<uint32_t> nf;
        nf.reserve(faces.size() * 4);
        for (size_t f = 0; f < faces.size(); f += 3) {
            uint32_t a = faces[f], b = faces[f + 1], c = faces[f + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), tri, tri + 12);
        }
        faces.swap(nf);
    }
    // displaced FP32 vertex positions
    std::vector<V> P(vx.size());
    for (size_t i = 0; i < vx.size(); ++i) {
        double r = radius + amp * Fbm(vx[i], vy[i], vz[i], seed);
        P[i] = {(float)(cx + r * vx[i]), (float)(cy + r * vy[i]), (float)(cz + r * vz[i])};
    }
    const size_t nf = std::min(faces.size(), max_tris * 3);
    for (size_t f = 0; f < nf; f += 3) {
        V p0 = P[faces[f]], p1 = P[faces[f + 1]], p2 = P[faces[f + 2]];
        V n = norm(cross(sub(p1, p0), sub(p2, p0)));
        // orient outward (away from the centre)
        double ox = (p0.x + p1.x + p2.x) / 3.0 - cx, oy = (p0.y + p1.y + p2.y) / 3.0 - cy, oz = (p0.z + p1.z + p2.z) / 3.0 - cz;
        uint32_t i0 = faces[f], i1 = faces[f + 1], i2 = faces[f + 2];
        if (n.x * ox + n.y * oy + n.z * oz < 0) {
            std::swap(p1, p2);
            std::swap(i1, i2);
            n = norm(cross(sub(p1, p0), sub(p2, p0)));
        }
        AddTri(g, p0, p1, p2, n, mat, -1);
        if (uvs) {
            // spherical coordinates of the undisplaced direction
            auto su = [&](uint32_t i) { return (float)(0.5 + std::atan2(vz[i], vx[i]) / (2.0 * 3.14159265358979)); };
            auto sv = [&](uint32_t i) { return (float)(std::acos(std::max(-1.0, std::min(1.0, vy[i]))) / 3.14159265358979); };
            SetLastUV(g, su(i0), sv(i0), su(i1), sv(i1), su(i2), sv(i2));
        }
    }
}

// splitmix64: the member's random stream
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uniform(); }
    int pick(int n) { return (int)(next() % (uint64_t)n); }
};

uint32_t AddObjectMaterial(Generated& g, Rng& r) {
    const float one[3] = {1.f, 1.f, 1.f};
    const float alpha = (float)r.range(0.05, 0.8);
    if (r.uniform() < 0.25) {
        uint32_t m = AddMaterial(g, GY_MAT_ROUGH_DIELECTRIC, GY_MATF_ALPHA_DIRECT, one);
        SetWord(g, m, 8, alpha);
        SetWord(g, m, 9, alpha);
        return m;
    }
    uint32_t m = AddMaterial(g, GY_MAT_CONDUCTOR, GY_MATF_ALPHA_DIRECT | GY_MATF_SCHLICK_F0, one);
    SetWord(g, m, 8, alpha);
    SetWord(g, m, 9, alpha);
    for (int c = 0; c < 3; ++c) SetWord(g, m, 16 + c, (float)r.range(0.5, 0.98));
    return m;
}

void BatchMember(Generated& g, uint32_t seed, int32_t small) {
    Rng r{0x6A09E667F3BCC909ull ^ ((uint64_t)seed * 0x2545F4914F6CDD1Dull)};
    AddRoomAndLight(g, false);
    double n_tri = std::exp(r.range(std::log(1e3), std::log(2e6)));
    if (small > 0) n_tri = std::min(n_tri, (double)small);
    int level = 0;
    while (20.0 * std::pow(4.0, level) < n_tri) ++level;
    const double cx = r.range(0.3, 0.7), cy = r.range(0.25, 0.45), cz = r.range(0.3, 0.7);
    const double rad = r.range(0.12, 0.25);
    AddIcosphere(g, level, seed, cx, cy, cz, rad, 0.03, AddObjectMaterial(g, r), (size_t)n_tri);
    const int boxes = r.pick(21);
    for (int b = 0; b < boxes; ++b) {
        const double sx = r.range(0.03, 0.15), sy = r.range(0.03, 0.15), sz = r.range(0.03, 0.15);
        const double x = r.range(0.01, 0.99 - sx), y = r.range(0.005, 0.6), z = r.range(0.01, 0.99 - sz);
        AddBox(g, {(float)x, (float)y, (float)z}, {(float)(x + sx), (float)(y + sy), (float)(z + sz)}, AddObjectMaterial(g, r));
    }
    const int res[4] = {128, 256, 512, 1024}, spp[5] = {4, 8, 16, 32, 64};
    int rr = res[r.pick(4)], sp = spp[r.pick(5)];
    const int depth = 4 + r.pick(9);
    if (small > 0) {
        rr = std::max(16, rr / 8);
        sp = std::min(sp, 4);
    }
    SetCamera(g, rr);
    g.params.spp = sp;
    g.params.max_depth = depth;
    g.params.rr_depth = 5;
}

// The textured Cornell room of the adjoint workload: every wall Lambert with its
// own RGB albedo texture (uv over the wall), the quad light, and a central
// sphere (icosphere, 5 subdivisions: 20,480 triangles) of GGX conductor whose
// roughness (alpha, given directly) comes from a one-channel texture in
// [0.05, 0.8].  Texels seeded by `seed`; `res` texels per side.
void AdjointScene(Generated& g, uint32_t seed, int res) {
    Rng r{0xB5AD4ECEDA1CE2A9ull ^ ((uint64_t)seed * 0x9E3779B97F4A7C15ull)};
    std::vector<float> tex;
    auto rgb_texture = [&]() {
        tex.resize((size_t)res * res * 3);
        for (float& t : tex) t = (float)r.range(0.2, 0.9);
        return AddTexture(g, res, res, 3, tex);
    };
    auto wall = [&](V a, V b, V c, V d, V n) {
        const float white[3] = {1.f, 1.f, 1.f};
        uint32_t m = AddMaterial(g, GY_MAT_LAMBERT, 0, white);
        const int32_t t = rgb_texture();
        std::memcpy(&g.mats[(size_t)m * GY_MAT_WORDS + 21], &t, 4);
        AddTri(g, a, b, c, n, m, -1);
        SetLastUV(g, 0, 0, 1, 0, 1, 1);
        AddTri(g, a, c, d, n, m, -1);
        SetLastUV(g, 0, 0, 1, 1, 0, 1);
    };
    wall({0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0});    // floor
    wall({0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}, {0, -1, 0});   // ceiling
    wall({0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 1});    // back
    wall({0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 0, 0});    // left
    wall({1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, {-1, 0, 0});   // right
    const float black[3] = {0.f, 0.f, 0.f};
    uint32_t mlight = AddMaterial(g, GY_MAT_LAMBERT, 0, black);
    const float y = 0.99f, a = 0.375f, b = 0.625f;
    const V l0{b, y, b}, l1{b, y, a}, l2{a, y, a}, l3{a, y, b}, ln{0, -1, 0};
    int32_t first = (int32_t)g.ltri.size();
    AddTri(g, l0, l1, l2, ln, mlight, first);
    g.ltri.push_back((int32_t)(g.fn.size() / 3 - 1));
    AddTri(g, l0, l2, l3, ln, mlight, first + 1);
    g.ltri.push_back((int32_t)(g.fn.size() / 3 - 1));
    for (int i = 0; i < 2; ++i) g.le.insert(g.le.end(), {17.f, 12.f, 4.f});
    const float one[3] = {1.f, 1.f, 1.f}, f0[3] = {0.95f, 0.64f, 0.54f};
    uint32_t ms = AddMaterial(g, GY_MAT_CONDUCTOR, GY_MATF_ALPHA_DIRECT | GY_MATF_SCHLICK_F0, one);
    for (int c = 0; c < 3; ++c) SetWord(g, ms, 16 + c, f0[c]);
    tex.resize((size_t)res * res);
    for (float& t : tex) t = (float)r.range(0.05, 0.8);
    const int32_t tr = AddTexture(g, res, res, 1, tex);
    std::memcpy(&g.mats[(size_t)ms * GY_MAT_WORDS + 22], &tr, 4);
    AddIcosphere(g, 5, seed, 0.5, 0.3, 0.5, 0.22, 0.0, ms, (size_t)-1, true);
}

// The 1.31M-triangle scene with textures on the icosphere (uv from spherical
// coordinates): RGB albedo and GGX roughness (alpha given directly), `res`
// texels per side, smooth seeded patterns: albedo 0.5 + 0.35 sin(.) sin(.),
// roughness in [0.05, 0.8].
void TexturedIcosphere(Generated& g, uint32_t seed, int res) {
    Rng r{0x3C6EF372FE94F82Bull ^ ((uint64_t)seed * 0x9E3779B97F4A7C15ull)};
    AddRoomAndLight(g, false);
    double fq[6];
    for (double& f : fq) f = r.range(2.0, 9.0);
    std::vector<float> alb((size_t)res * res * 3), rough((size_t)res * res);
    for (int y = 0; y < res; ++y)
        for (int x = 0; x < res; ++x) {
            const double u = (x + 0.5) / res, v = (y + 0.5) / res;
            for (int c = 0; c < 3; ++c)
                alb[((size_t)y * res + x) * 3 + c] =
                    (float)(0.5 + 0.35 * std::sin(2 * 3.14159265358979 * fq[c] * u + c) * std::sin(2 * 3.14159265358979 * fq[3] * v));
            rough[(size_t)y * res + x] = (float)(0.425 + 0.375 * std::sin(2 * 3.14159265358979 * (fq[4] * u + fq[5] * v)));
        }
    const int32_t ta = AddTexture(g, res, res, 3, alb), tr = AddTexture(g, res, res, 1, rough);
    const float one[3] = {1.f, 1.f, 1.f}, f0[3] = {0.95f, 0.64f, 0.54f};
    uint32_t m = AddMaterial(g, GY_MAT_CONDUCTOR, GY_MATF_ALPHA_DIRECT | GY_MATF_SCHLICK_F0, one);
    for (int c = 0; c < 3; ++c) SetWord(g, m, 16 + c, f0[c]);
    std::memcpy(&g.mats[(size_t)m * GY_MAT_WORDS + 21], &ta, 4);
    std::memcpy(&g.mats[(size_t)m * GY_MAT_WORDS + 22], &tr, 4);
    AddIcosphere(g, 8, seed, 0.5, 0.4, 0.5, 0.3, 0.05, m, (size_t)-1, true);
}

}  // namespace
}  // namespace gy

extern "C" int gy_generate_scene(int kind, uint32_t seed, int32_t param, gy_generated** out) {
    using namespace gy;
    if (!out) return Fail(GY_ERR_INVALID, "gy_generate_scene: null output");
    Generated* g = new Generated();
    std::memset(&g->params, 0, sizeof(g->params));
    if (kind == GY_SCENE_CORNELL) {
        AddRoomAndLight(*g, true);
        SetCamera(*g, param > 0 ? param : 256);
        g->params.spp = 16;
        g->params.max_depth = 5;
        g->params.rr_depth = 3;
    } else if (kind == GY_SCENE_ICOSPHERE_GGX) {
        int levels = param > 0 ? param : 8;
        if (levels > 10) {
            delete g;
            return Fail(GY_ERR_INVALID, "gy_generate_scene: icosphere levels must be <= 10");
        }
        AddRoomAndLight(*g, false);
        const float f0[3] = {0.95f, 0.64f, 0.54f}, one[3] = {1.f, 1.f, 1.f};
        uint32_t m = AddMaterial(*g, GY_MAT_CONDUCTOR, GY_MATF_ALPHA_DIRECT | GY_MATF_SCHLICK_F0, one);
        SetWord(*g, m, 8, 0.2f);
        SetWord(*g, m, 9, 0.2f);
        for (int c = 0; c < 3; ++c) SetWord(*g, m, 16 + c, f0[c]);
        AddIcosphere(*g, levels, seed, 0.5, 0.4, 0.5, 0.3, 0.05, m);
        SetCamera(*g, 1024);
        g->params.spp = 64;
        g->params.max_depth = 8;
        g->params.rr_depth = 5;
    } else if (kind == GY_SCENE_BATCH_MEMBER) {
        BatchMember(*g, seed, param);
    } else if (kind == GY_SCENE_ICOSPHERE_TEXTURED) {
        TexturedIcosphere(*g, seed, param > 0 ? param : 1024);
        SetCamera(*g, 512);
        g->params.spp = 32;
        g->params.max_depth = 6;
        g->params.rr_depth = 5;
    } else if (kind == GY_SCENE_ADJOINT) {
        AdjointScene(*g, seed, param > 0 ? param : 64);
        SetCamera(*g, 256);
        g->params.spp = 64;
        g->params.max_depth = 5;
        g->params.rr_depth = 3;
    } else {
        delete g;
        return Fail(GY_ERR_INVALID, "gy_generate_scene: unknown scene kind");
    }
    g->params.mode = GY_RENDER_WAVEFRONT;
    *out = reinterpret_cast<gy_generated*>(g);
    return GY_OK;
}

extern "C" int gy_generated_desc(gy_generated* gg, gy_scene_desc* d, gy_camera_desc* cam, gy_render_params* params) {
    using namespace gy;
    if (!gg) return Fail(GY_ERR_INVALID, "gy_generated_desc: null handle");
    Generated* g = reinterpret_cast<Generated*>(gg);
    if (d) {
        std::memset(d, 0, sizeof(*d));
        d->n_tri = (int32_t)g->mat.size();
        d->pos = g->pos.data();
        d->nrm = g->nrm.data();
        d->face_n = g->fn.data();
        d->uv = g->uv.data();
        d->n_tex = (int32_t)(g->texi.size() / 4);
        d->tex_info = g->texi.empty() ? nullptr : g->texi.data();
        d->tex_data = g->texd.empty() ? nullptr : g->texd.data();
        d->mat = g->mat.data();
        d->light = g->light.data();
        d->n_mat = (int32_t)(g->mats.size() / GY_MAT_WORDS);
        d->mats = g->mats.data();
        d->n_light = (int32_t)g->ltri.size();
        d->light_tri = g->ltri.data();
        d->light_le = g->le.data();
    }
    if (cam) *cam = g->cam;
    if (params) *params = g->params;
    return GY_OK;
}

extern "C" int gy_generated_free(gy_generated* g) {
    delete reinterpret_cast<gy::Generated*>(g);
    return GY_OK;
}
