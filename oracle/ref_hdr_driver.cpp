// Test infrastructure, not part of the product: reads an image through the
// reference's own loader (stb_image, external/stb/stb_image.h, called as
// stbi_loadf(file, &w, &h, &channels, 0) by ImageIO::LoadTexture,
// image/ImageIO.cpp:100-114) so that tests can compare gy_image_read with it.
// Built by oracle/Makefile from the header where it lies under the reference
// tree.  Output: little-endian int32 width, height, channels, then the floats.
#define STB_IMAGE_IMPLEMENTATION
#define STBI_ONLY_HDR
#include "stb_image.h"

#include <cstdio>

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.hdr out.raw\n", argv[0]);
        return 2;
    }
    int w = 0, h = 0, c = 0;
    float* d = stbi_loadf(argv[1], &w, &h, &c, 0);
    if (!d) {
        std::fprintf(stderr, "%s\n", stbi_failure_reason());
        return 1;
    }
    FILE* f = std::fopen(argv[2], "wb");
    if (!f) return 3;
    const int head[3] = {w, h, c};
    std::fwrite(head, 4, 3, f);
    std::fwrite(d, 4, (size_t)w * h * c, f);
    std::fclose(f);
    stbi_image_free(d);
    return 0;
}
