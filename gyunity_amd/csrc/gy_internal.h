// Internal declarations shared by the host C++ sources and the HIP kernels.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/gyunity.h"

namespace gy {

// Thread-local error message behind gy_last_error().
void SetError(const std::string& msg);
int Fail(int code, const std::string& msg);

// A scene's process-unique id and the revision of its device buffers
// (environment light, spheres, media); texel updates keep the revision.
int SceneSignature(gy_scene* s, uint64_t* uid, uint64_t* revision);

namespace sobol {
constexpr int kDims = 1024;       // dimensions of the Joe-Kuo table
constexpr int kCols = 52;         // index bits per dimension
constexpr int kMaxLog2Res = 16;   // enumeration up to 65536^2 pixels
struct EnumTables {
    int m;
    uint64_t inv[64];    // A^-1 by columns (target bit -> low-index bits)
    uint64_t frame[64];  // pixel bits reached by frame bit j
};
const uint32_t* Matrices();        // kDims * kCols
// Byte tables of the matrices: entry [dim][k][v] is the XOR of the columns
// 8k + j (j < 8, 8k + j < kCols) for the set bits j of v, so a sample is the
// XOR of one entry per index byte.  kDims * kBytes * 256 words.
constexpr int kBytes = 7;
const uint32_t* ByteTables();
const EnumTables& Enum(int m);
float Sample(uint64_t index, uint32_t dim);
uint64_t EnumIndex(uint32_t m, uint32_t frame, uint32_t px, uint32_t py);
}  // namespace sobol

namespace halton {
constexpr int kDims = 256;  // HaltonSampler::sample covers dimensions 0..255
struct DimInfo {
    uint32_t base, tsize, groups, offset;  // prime, digit-table size, digit groups, table offset
    float scale;                           // float((1 - 2^-23) / tsize^groups)
};
const std::vector<DimInfo>& Dims();
const std::vector<uint16_t>& Tables();
float Sample(uint64_t index, uint32_t dim);
struct EnumParams {
    uint32_t p2, p3;        // 2^p2 >= width, 3^p3 >= height
    uint64_t mx, my, inc;   // CRT multipliers, 2^p2 * 3^p3
    float sx, sy;           // 2^p2, 3^p3
};
EnumParams Enum(uint32_t width, uint32_t height);
uint64_t EnumIndex(const EnumParams& e, uint64_t i, uint32_t x, uint32_t y);
}  // namespace halton

// Camera matrix in the reference's FP32 operation order (camera/Camera.cpp:6-38).
struct CameraRT {
    float raster_to_world[16];  // column-major, (i, j) at [j*4 + i]
    float pos[3];
    int32_t width, height;
    int32_t res;       // RoundUpPow2(max(width, height))  (sampler/SobolEnum.h:18)
    int32_t log2res;
};
int BuildCamera(const gy_camera_desc& d, CameraRT* out);
// The pinhole camera as LightTracing sees it: WorldToRaster = CameraToRaster *
// WorldToCamera (integrator/LT.h:118), the film distance and film area
// (camera/Camera.cpp:42-44) for PinHoleCamera::We.
struct LightCameraRT {
    float world_to_raster[16];  // column-major, (i, j) at [j*4 + i]
    float pos[3], cz[3];
    float dis, area;
};
int BuildLightCamera(const gy_camera_desc& d, LightCameraRT* out);

// BVH in the device layout: 16 words per node.
//   words 0-5   child 0 AABB (min xyz, max xyz)
//   words 6-11  child 1 AABB
//   word 12/13  child 0/1 reference: node index (interior), first triangle (leaf),
//               0xffffffff (empty slot)
//   word 14/15  child 0/1 triangle count (0 = interior)
struct BVHBuild {
    std::vector<uint32_t> nodes;     // n_nodes * 16
    std::vector<uint32_t> tri_order; // leaf order -> primitive id
    int32_t max_depth = 0;
    int64_t n_leaves = 0;
    int32_t max_leaf = 0;
    double sah_cost = 0;
};
constexpr uint32_t kEmptyRef = 0xffffffffu;
// Host binned-SAH build over triangle bounds.
void BuildBVHHost(const float* pos, int32_t n_tri, int max_leaf, BVHBuild* out);
// ImageIO::WriteImage: .hdr, .bmp or .png by suffix (image.cpp).
int WriteImage(const char* path, const float* rgb, int32_t w, int32_t h);
double BVHSahCost(const BVHBuild& b);

struct Generated;

}  // namespace gy
