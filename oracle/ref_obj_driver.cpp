// Test infrastructure, not part of the product: dumps what the reference's own
// OBJ loader (tinyobjloader 2.0.0, external/tinyobjloader/tiny_obj_loader.h,
// called as in mesh/Mesh.cpp:13 with triangulation on) reads from a file, so
// that tests can compare gy_mesh_load_obj's parse with it value for value.
// Built by oracle/Makefile from the header where it lies under the reference
// tree; the output is binary little-endian:
//   int64 n_v, n_vn, n_vt, n_tri
//   float v[3 n_v], vn[3 n_vn], vt[2 n_vt]
//   int32 (v, vt, vn) per corner, three corners per triangle, shapes in order (-1 = none)
#define TINYOBJLOADER_IMPLEMENTATION
#include "tiny_obj_loader.h"

#include <cstdint>
#include <cstdio>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.obj out.bin\n", argv[0]);
        return 2;
    }
    tinyobj::attrib_t attrib;
    std::vector<tinyobj::shape_t> shapes;
    std::vector<tinyobj::material_t> materials;
    std::string warn, err;
    if (!tinyobj::LoadObj(&attrib, &shapes, &materials, &warn, &err, argv[1])) {
        std::fprintf(stderr, "load failed: %s\n", err.c_str());
        return 1;
    }
    std::vector<int32_t> corners;
    for (const auto& s : shapes) {
        size_t off = 0;
        for (size_t f = 0; f < s.mesh.num_face_vertices.size(); ++f) {
            const int fv = s.mesh.num_face_vertices[f];
            for (int k = 0; k < fv; ++k) {
                const auto& idx = s.mesh.indices[off + k];
                corners.insert(corners.end(), {idx.vertex_index, idx.texcoord_index, idx.normal_index});
            }
            off += fv;
        }
    }
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 1;
    const int64_t h[4] = {(int64_t)attrib.vertices.size() / 3, (int64_t)attrib.normals.size() / 3,
                          (int64_t)attrib.texcoords.size() / 2, (int64_t)corners.size() / 9};
    std::fwrite(h, 8, 4, o);
    std::fwrite(attrib.vertices.data(), 4, attrib.vertices.size(), o);
    std::fwrite(attrib.normals.data(), 4, attrib.normals.size(), o);
    std::fwrite(attrib.texcoords.data(), 4, attrib.texcoords.size(), o);
    std::fwrite(corners.data(), 4, corners.size(), o);
    std::fclose(o);
    return 0;
}
