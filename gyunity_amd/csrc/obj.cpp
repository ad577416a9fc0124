// The number parser (ParseNumber) and the ear clipping (Triangulate) follow
// tinyobjloader's steps closely, so its notice is kept here:
//
//   The MIT License (MIT)
//
//   Copyright (c) 2012-2018 Syoyo Fujita and many contributors.
//
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to permit
//   persons to whom the Software is furnished to do so, subject to the
//   following conditions:
//
//   The above copyright notice and this permission notice shall be included
//   in all copies or substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS
//   OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
//   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN
//   NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
//   DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR
//   OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE
//   USE OR OTHER DEALINGS IN THE SOFTWARE.
//
// OBJ mesh ingest: the triangles Gyunity's scene holds after
// Mesh::LoadFromFile (mesh/Mesh.cpp:7-84) and Scene::AddMesh
// (visual/Scene.h:68-77), as flat per-triangle arrays for gy_scene_desc.
//
// Parsing restates tinyobjloader 2.0.0 (external/tinyobjloader, the loader
// the reference calls with triangulation on) wherever the result depends on it:
//   - numbers: the integer digits accumulate in a double (m = 10 m + d), each
//     fraction digit k adds d * 10^-k (a table for k < 8, pow beyond), an
//     exponent e gives ldexp(m * 5^e, e), and the double is cast to float;
//   - face indices: 1-based, negative counts back from the current end, 0 is
//     an error; "v", "v/t", "v//n", "v/t/n";
//   - polygons with more than three corners: ear clipping in the plane of the
//     two axes other than the dominant normal axis of the first corner whose
//     cross product exceeds FLT_EPSILON; a corner is an ear when it turns with
//     the polygon's signed area and no other corner lies inside it (crossing
//     test); an ear emits (k, k+1, k+2) and drops k+1; three corners are
//     emitted as they are.
// Mesh::LoadFromFile then gives every face its own three vertices (reversed
// to (0, 2, 1) on request), with the file's normals or, when the face has
// none, Cross(v1 - v0, v2 - v0) normalised when its length exceeds 1e-6, and
// uv (0, 0) where a corner has none.  A face that mixes corners with and
// without normals is rejected (the reference asserts on it).  Triangle::SetTransform
// (shape/Triangle.h:78-88) maps points through the object-to-world matrix
// (Transform::operator(), math/Transform.h:106-115), normals through the
// transposed inverse, normalised (Transform.h:247-251), and sets the face
// normal to Cross(p1 - p0, p2 - p1), normalised.  FP32 throughout, in the
// reference's operation order (host_math.h).
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gy_internal.h"
#include "host_math.h"

namespace gy {
namespace {

using hm::M4;

bool IsDigit(char c) { return c >= '0' && c <= '9'; }

// The number in [s, e); false when it does not start as a number.
bool ParseNumber(const char* s, const char* e, double* out) {
    if (s >= e) return false;
    double m = 0.0;
    int ex = 0;
    bool neg = false, eneg = false, lead_dot = false;
    const char* c = s;
    if (*c == '+' || *c == '-') {
        neg = *c == '-';
        ++c;
        if (c != e && *c == '.') lead_dot = true;
    } else if (*c == '.') {
        lead_dot = true;
    } else if (!IsDigit(*c)) {
        return false;
    }
    int read = 0;
    if (!lead_dot) {
        while (c != e && IsDigit(*c)) {
            m = m * 10;
            m += (int)(*c - '0');
            ++c;
            ++read;
        }
        if (read == 0) return false;
    }
    if (c == e) goto done;
    if (*c == '.') {
        static const double kTenth[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
        ++c;
        read = 1;
        while (c != e && IsDigit(*c)) {
            m += (int)(*c - '0') * (read < 8 ? kTenth[read] : std::pow(10.0, -read));
            ++read;
            ++c;
        }
    } else if (*c != 'e' && *c != 'E') {
        goto done;
    }
    if (c == e) goto done;
    if (*c == 'e' || *c == 'E') {
        ++c;
        if (c != e && (*c == '+' || *c == '-')) {
            eneg = *c == '-';
            ++c;
        } else if (c == e || !IsDigit(*c)) {
            return false;
        }
        read = 0;
        while (c != e && IsDigit(*c)) {
            ex = ex * 10 + (int)(*c - '0');
            ++c;
            ++read;
        }
        if (eneg) ex = -ex;
        if (read == 0) return false;
    }
done:
    *out = (neg ? -1 : 1) * (ex ? std::ldexp(m * std::pow(5.0, ex), ex) : m);
    return true;
}

// One whitespace-separated number of a line (absent or malformed: dflt).
float NextFloat(const char** p, double dflt) {
    const char* s = *p + std::strspn(*p, " \t");
    const char* e = s + std::strcspn(s, " \t\r");
    double v = dflt;
    ParseNumber(s, e, &v);
    *p = e;
    return (float)v;
}

struct Corner {
    int v, t, n;  // 0-based, -1 = none
};

// tinyobjloader's index rule: positive is 1-based, negative counts back from
// n (the number read so far), zero is invalid.
bool FixIndex(int idx, int n, int* out) {
    if (idx > 0) {
        *out = idx - 1;
        return true;
    }
    if (idx == 0) return false;
    *out = n + idx;
    return true;
}

bool ParseCorner(const char** p, int nv, int nn, int nt, Corner* c) {
    c->v = c->t = c->n = -1;
    const char* s = *p;
    if (!FixIndex(std::atoi(s), nv, &c->v)) return false;
    s += std::strcspn(s, "/ \t\r");
    if (*s == '/') {
        ++s;
        if (*s == '/') {  // v//n
            ++s;
            if (!FixIndex(std::atoi(s), nn, &c->n)) return false;
            s += std::strcspn(s, "/ \t\r");
        } else {  // v/t or v/t/n
            if (!FixIndex(std::atoi(s), nt, &c->t)) return false;
            s += std::strcspn(s, "/ \t\r");
            if (*s == '/') {
                ++s;
                if (!FixIndex(std::atoi(s), nn, &c->n)) return false;
                s += std::strcspn(s, "/ \t\r");
            }
        }
    }
    *p = s;
    return true;
}

// Even-odd crossing test of (x, y) against the triangle (vx, vy).
bool InTriangle(const float* vx, const float* vy, float x, float y) {
    bool in = false;
    for (int i = 0, j = 2; i < 3; j = i++)
        if (((vy[i] > y) != (vy[j] > y)) && (x < (vx[j] - vx[i]) * (y - vy[i]) / (vy[j] - vy[i]) + vx[i])) in = !in;
    return in;
}

// Triangles of one face, by ear clipping (three corners: as they are).
void Triangulate(const std::vector<Corner>& face, const std::vector<float>& v, std::vector<Corner>* out) {
    const size_t n0 = face.size();
    if (n0 < 3) return;
    auto ok = [&](int vi) { return vi >= 0 && (size_t)vi * 3 + 2 < v.size(); };
    size_t ax[2] = {1, 2};
    for (size_t k = 0; k < n0; ++k) {
        const Corner a = face[k], b = face[(k + 1) % n0], c = face[(k + 2) % n0];
        if (!ok(a.v) || !ok(b.v) || !ok(c.v)) continue;
        const float e0x = v[3 * b.v] - v[3 * a.v], e0y = v[3 * b.v + 1] - v[3 * a.v + 1], e0z = v[3 * b.v + 2] - v[3 * a.v + 2];
        const float e1x = v[3 * c.v] - v[3 * b.v], e1y = v[3 * c.v + 1] - v[3 * b.v + 1], e1z = v[3 * c.v + 2] - v[3 * b.v + 2];
        const float cx = std::fabs(e0y * e1z - e0z * e1y), cy = std::fabs(e0z * e1x - e0x * e1z), cz = std::fabs(e0x * e1y - e0y * e1x);
        if (cx > FLT_EPSILON || cy > FLT_EPSILON || cz > FLT_EPSILON) {
            if (!(cx > cy && cx > cz)) {
                ax[0] = 0;
                if (cz > cx && cz > cy) ax[1] = 1;
            }
            break;
        }
    }
    float area = 0.f;
    for (size_t k = 0; k < n0; ++k) {
        const int a = face[k].v, b = face[(k + 1) % n0].v;
        if (!ok(a) || !ok(b)) continue;
        area += (v[3 * a + ax[0]] * v[3 * b + ax[1]] - v[3 * a + ax[1]] * v[3 * b + ax[0]]) * 0.5f;
    }
    std::vector<Corner> rest = face;
    size_t guess = 0, tries = rest.size(), last = rest.size();
    while (rest.size() > 3 && tries > 0) {
        const size_t n = rest.size();
        if (guess >= n) guess -= n;
        if (last != n) {
            last = n;
            tries = n;
        } else {
            --tries;
        }
        Corner ind[3];
        float vx[3], vy[3];
        for (int k = 0; k < 3; ++k) {
            ind[k] = rest[(guess + k) % n];
            const bool good = ind[k].v >= 0 && (size_t)ind[k].v * 3 + ax[0] < v.size() && (size_t)ind[k].v * 3 + ax[1] < v.size();
            vx[k] = good ? v[3 * ind[k].v + ax[0]] : 0.f;
            vy[k] = good ? v[3 * ind[k].v + ax[1]] : 0.f;
        }
        const float cross = (vx[1] - vx[0]) * (vy[2] - vy[1]) - (vy[1] - vy[0]) * (vx[2] - vx[1]);
        if (cross * area < 0.f) {
            ++guess;
            continue;
        }
        bool inside = false;
        for (size_t o = 3; o < n && !inside; ++o) {
            const int ov = rest[(guess + o) % n].v;
            if (ov < 0 || (size_t)ov * 3 + ax[0] >= v.size() || (size_t)ov * 3 + ax[1] >= v.size()) continue;
            inside = InTriangle(vx, vy, v[3 * ov + ax[0]], v[3 * ov + ax[1]]);
        }
        if (inside) {
            ++guess;
            continue;
        }
        out->push_back(ind[0]);
        out->push_back(ind[1]);
        out->push_back(ind[2]);
        rest.erase(rest.begin() + (long)((guess + 1) % n));
    }
    if (rest.size() == 3) out->insert(out->end(), rest.begin(), rest.end());
}

}  // namespace

struct Mesh {
    std::vector<float> pos, nrm, face_n, uv;  // 9, 9, 3, 6 floats per triangle
};

int LoadObj(const char* path, bool reverse, const float* o2w, const float* w2o, Mesh* out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return Fail(GY_ERR_INVALID, std::string("gy_mesh_load_obj: cannot open ") + path);
    std::vector<float> v, vn, vt;
    std::vector<Corner> tris;  // three corners per triangle, file order
    std::vector<Corner> face;
    std::string line;
    char buf[4096];
    int lineno = 0;
    bool bad = false;
    while (std::fgets(buf, sizeof buf, f)) {
        line = buf;
        while (!line.empty() && line.back() != '\n' && std::fgets(buf, sizeof buf, f)) line += buf;  // long lines
        ++lineno;
        const char* p = line.c_str();
        p += std::strspn(p, " \t");
        if (p[0] == 'v' && (p[1] == ' ' || p[1] == '\t')) {
            p += 2;
            const float x = NextFloat(&p, 0.0), y = NextFloat(&p, 0.0), z = NextFloat(&p, 0.0);
            v.insert(v.end(), {x, y, z});
        } else if (p[0] == 'v' && p[1] == 'n' && (p[2] == ' ' || p[2] == '\t')) {
            p += 3;
            const float x = NextFloat(&p, 0.0), y = NextFloat(&p, 0.0), z = NextFloat(&p, 0.0);
            vn.insert(vn.end(), {x, y, z});
        } else if (p[0] == 'v' && p[1] == 't' && (p[2] == ' ' || p[2] == '\t')) {
            p += 3;
            const float x = NextFloat(&p, 0.0), y = NextFloat(&p, 0.0);
            vt.insert(vt.end(), {x, y});
        } else if (p[0] == 'f' && (p[1] == ' ' || p[1] == '\t')) {
            p += 2;
            p += std::strspn(p, " \t");
            face.clear();
            while (*p && *p != '\r' && *p != '\n' && *p != '#') {
                Corner c;
                if (!ParseCorner(&p, (int)(v.size() / 3), (int)(vn.size() / 3), (int)(vt.size() / 2), &c)) {
                    bad = true;
                    break;
                }
                face.push_back(c);
                p += std::strspn(p, " \t");
            }
            if (bad) break;
            Triangulate(face, v, &tris);
        }
    }
    std::fclose(f);
    if (bad) return Fail(GY_ERR_INVALID, std::string("gy_mesh_load_obj: bad face index at line ") + std::to_string(lineno));
    const size_t nt = tris.size() / 3;
    for (size_t k = 0; k < tris.size(); ++k) {
        const Corner& c = tris[k];
        if (c.v < 0 || (size_t)c.v * 3 + 2 >= v.size() || (c.n >= 0 && (size_t)c.n * 3 + 2 >= vn.size()) ||
            (c.t >= 0 && (size_t)c.t * 2 + 1 >= vt.size()))
            return Fail(GY_ERR_INVALID, "gy_mesh_load_obj: face index out of range");
    }
    // Mesh::LoadFromFile: per-face vertices, normals and uvs.
    using hm::V3;
    M4 m, inv_t;
    {
        static const float kId[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        m = hm::Rows(o2w ? o2w : kId);
        const M4 inv = w2o ? hm::Rows(w2o) : hm::Inverse(m);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) inv_t.c[i][j] = inv.c[j][i];
    }
    auto point = [&](V3 p) {  // Transform::operator()(p)
        const float a[4] = {p.x, p.y, p.z, 1.f};
        float r[4];
        hm::MulVec(m, a, r);
        if (r[3] == 1.f || r[3] == 0.f) return V3{r[0], r[1], r[2]};
        return V3{r[0] / r[3], r[1] / r[3], r[2] / r[3]};
    };
    auto normal = [&](V3 n) {  // Transform::TransformNormal(n).Norm()
        const float a[4] = {n.x, n.y, n.z, 0.f};
        float r[4];
        hm::MulVec(inv_t, a, r);
        return hm::Norm(V3{r[0], r[1], r[2]});
    };
    out->pos.resize(nt * 9);
    out->nrm.resize(nt * 9);
    out->face_n.resize(nt * 3);
    out->uv.resize(nt * 6);
    for (size_t t = 0; t < nt; ++t) {
        V3 p[3], n[3];
        float uv[3][2];
        int with_n = 0;
        for (int k = 0; k < 3; ++k) {
            const Corner& c = tris[3 * t + k];
            p[k] = V3{v[3 * c.v], v[3 * c.v + 1], v[3 * c.v + 2]};
            if (c.n >= 0) {
                n[k] = V3{vn[3 * c.n], vn[3 * c.n + 1], vn[3 * c.n + 2]};
                ++with_n;
            }
            uv[k][0] = c.t >= 0 ? vt[2 * c.t] : 0.f;
            uv[k][1] = c.t >= 0 ? vt[2 * c.t + 1] : 0.f;
        }
        // The reference asserts that a face's corners all have normals or none
        // has (debug builds); a release build then misaligns its normal list
        // for every later face.  Such files are rejected here.
        if (with_n != 0 && with_n != 3)
            return Fail(GY_ERR_INVALID, "gy_mesh_load_obj: a face mixes corners with and without normals");
        if (with_n == 0) {
            V3 g = hm::Cross(hm::Sub(p[1], p[0]), hm::Sub(p[2], p[0]));
            if (std::sqrt(hm::Dot(g, g)) > 1e-6f) g = hm::Norm(g);
            n[0] = n[1] = n[2] = g;
        }
        const int order[3] = {0, reverse ? 2 : 1, reverse ? 1 : 2};
        V3 q[3];
        for (int k = 0; k < 3; ++k) {
            const int s = order[k];
            q[k] = point(p[s]);
            const V3 nn = normal(n[s]);
            out->pos[9 * t + 3 * k] = q[k].x;
            out->pos[9 * t + 3 * k + 1] = q[k].y;
            out->pos[9 * t + 3 * k + 2] = q[k].z;
            out->nrm[9 * t + 3 * k] = nn.x;
            out->nrm[9 * t + 3 * k + 1] = nn.y;
            out->nrm[9 * t + 3 * k + 2] = nn.z;
            out->uv[6 * t + 2 * k] = uv[s][0];
            out->uv[6 * t + 2 * k + 1] = uv[s][1];
        }
        const V3 fn = hm::Norm(hm::Cross(hm::Sub(q[1], q[0]), hm::Sub(q[2], q[1])));
        out->face_n[3 * t] = fn.x;
        out->face_n[3 * t + 1] = fn.y;
        out->face_n[3 * t + 2] = fn.z;
    }
    return GY_OK;
}

}  // namespace gy

extern "C" int gy_mesh_load_obj(const char* path, int32_t reverse_vertices, const float* object_to_world,
                                const float* world_to_object, gy_mesh** out) {
    if (!path || !out) return gy::Fail(GY_ERR_INVALID, "gy_mesh_load_obj: null argument");
    auto* m = new gy::Mesh();
    const int rc = gy::LoadObj(path, reverse_vertices != 0, object_to_world, world_to_object, m);
    if (rc != GY_OK) {
        delete m;
        return rc;
    }
    *out = reinterpret_cast<gy_mesh*>(m);
    return GY_OK;
}

extern "C" int64_t gy_mesh_triangles(gy_mesh* mesh) {
    return mesh ? (int64_t)(reinterpret_cast<gy::Mesh*>(mesh)->face_n.size() / 3) : -1;
}

extern "C" int gy_mesh_arrays(gy_mesh* mesh, float* pos, float* nrm, float* face_n, float* uv) {
    if (!mesh) return gy::Fail(GY_ERR_INVALID, "gy_mesh_arrays: null mesh");
    const gy::Mesh& m = *reinterpret_cast<gy::Mesh*>(mesh);
    if (pos) std::memcpy(pos, m.pos.data(), m.pos.size() * 4);
    if (nrm) std::memcpy(nrm, m.nrm.data(), m.nrm.size() * 4);
    if (face_n) std::memcpy(face_n, m.face_n.data(), m.face_n.size() * 4);
    if (uv) std::memcpy(uv, m.uv.data(), m.uv.size() * 4);
    return GY_OK;
}

extern "C" int gy_mesh_free(gy_mesh* mesh) {
    delete reinterpret_cast<gy::Mesh*>(mesh);
    return GY_OK;
}
