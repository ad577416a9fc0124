// Test infrastructure, not part of the product: writes an image through the
// reference's own writer (stb_image_write, external/stb/stb_image_write.h, as
// image/ImageIO.cpp:120-149 calls it) so that tests can compare gy_image_write's
// files with it.  Built by oracle/Makefile from the header where it lies under
// the reference tree.  Input: raw little-endian int32 width, height, then
// width*height*3 values: bytes for .bmp/.png, floats for .hdr.
#define STB_IMAGE_WRITE_IMPLEMENTATION
#include "stb_image_write.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.raw out.{bmp,png,hdr}\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 1;
    int32_t wh[2];
    if (std::fread(wh, 4, 2, f) != 2) return 1;
    const size_t n = (size_t)wh[0] * wh[1] * 3;
    const size_t len = std::strlen(argv[2]);
    const char* suf = argv[2] + (len >= 4 ? len - 4 : 0);
    int ok = 0;
    if (std::strcmp(suf, ".hdr") == 0) {
        std::vector<float> px(n);
        if (std::fread(px.data(), 4, n, f) != n) return 1;
        ok = stbi_write_hdr(argv[2], wh[0], wh[1], 3, px.data());
    } else {
        std::vector<unsigned char> px(n);
        if (std::fread(px.data(), 1, n, f) != n) return 1;
        if (std::strcmp(suf, ".bmp") == 0) ok = stbi_write_bmp(argv[2], wh[0], wh[1], 3, px.data());
        else if (std::strcmp(suf, ".png") == 0) ok = stbi_write_png(argv[2], wh[0], wh[1], 3, px.data(), wh[0] * 3);
    }
    std::fclose(f);
    return ok ? 0 : 1;
}
