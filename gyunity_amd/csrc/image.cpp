// Film output: Film::SaveImage (visual/Film.cpp:6-9) writes the pixel buffer
// color / weight + splat (visual/Film.h:128-136; splat stays zero for path
// tracing) through ImageIO::WriteImage (image/ImageIO.cpp:120-149), which picks
// the format by suffix:
//   .hdr  Radiance RGBE with run-length scanlines;
//   .bmp  24-bit, bottom-up rows, BGR, rows padded to four bytes;
//   .png  8-bit RGB.
// 8-bit values are ImageIO::toInt(x) = int(pow(clamp(x, 0, 1), 1 / 2.2f) * 255
// + 0.5f) (ImageIO.cpp:66-68; WriteImage calls it with its default gamma).
// The reference writes through stb_image_write (external/stb); the BMP layout
// and the RGBE conversion and run-length coding below follow that writer's
// published format so that the files carry the same pixel bytes.  The power is
// evaluated in double and rounded to float, as every transcendental here.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gy_internal.h"

namespace gy {
namespace {

uint8_t ToByte(float x) {
    float c = x;
    if (c < 0.f) c = 0.f;
    if (c > 1.f) c = 1.f;  // NaN passes through, as in Clamp (math/MathUtils.h:48-54)
    const float p = (float)std::pow((double)c, (double)(1.f / 2.2f));
    const float v = p * 255 + .5f;
    if (!(v >= 0.f)) return 0;  // NaN: the reference's int conversion is undefined there
    return (uint8_t)(int)v;
}

struct Out {
    FILE* f = nullptr;
    bool ok = true;
    void put(const void* p, size_t n) {
        if (ok && n && std::fwrite(p, 1, n, f) != n) ok = false;
    }
    void u8(uint8_t v) { put(&v, 1); }
    void le16(uint16_t v) {
        const uint8_t b[2] = {(uint8_t)v, (uint8_t)(v >> 8)};
        put(b, 2);
    }
    void le32(uint32_t v) {
        const uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
        put(b, 4);
    }
    void be32(uint32_t v) {
        const uint8_t b[4] = {(uint8_t)(v >> 24), (uint8_t)(v >> 16), (uint8_t)(v >> 8), (uint8_t)v};
        put(b, 4);
    }
};

void WriteBmp(Out& o, const std::vector<uint8_t>& px, int w, int h) {
    const int pad = (-w * 3) & 3;
    o.put("BM", 2);
    o.le32((uint32_t)(14 + 40 + (w * 3 + pad) * h));
    o.le16(0);
    o.le16(0);
    o.le32(14 + 40);
    o.le32(40);
    o.le32((uint32_t)w);
    o.le32((uint32_t)h);
    o.le16(1);
    o.le16(24);
    for (int k = 0; k < 6; ++k) o.le32(0);
    const uint8_t zero[3] = {0, 0, 0};
    for (int y = h - 1; y >= 0; --y) {
        for (int x = 0; x < w; ++x) {
            const uint8_t* d = &px[((size_t)y * w + x) * 3];
            const uint8_t bgr[3] = {d[2], d[1], d[0]};
            o.put(bgr, 3);
        }
        o.put(zero, (size_t)pad);
    }
}

void Rgbe(const float* rgb, uint8_t* e) {
    const float m = rgb[0] > (rgb[1] > rgb[2] ? rgb[1] : rgb[2]) ? rgb[0] : (rgb[1] > rgb[2] ? rgb[1] : rgb[2]);
    if (m < 1e-32f) {
        e[0] = e[1] = e[2] = e[3] = 0;
        return;
    }
    int ex;
    const float scale = (float)std::frexp(m, &ex) * 256.0f / m;
    e[0] = (uint8_t)(rgb[0] * scale);
    e[1] = (uint8_t)(rgb[1] * scale);
    e[2] = (uint8_t)(rgb[2] * scale);
    e[3] = (uint8_t)(ex + 128);
}

void WriteHdr(Out& o, const float* rgb, int w, int h) {
    char head[160];
    const int n = std::snprintf(head, sizeof head, "#?RADIANCE\n# gyunity-amd film\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
    o.put(head, (size_t)n);
    std::vector<uint8_t> plane((size_t)w * 4);
    for (int y = 0; y < h; ++y) {
        const float* row = rgb + (size_t)y * w * 3;
        if (w < 8 || w >= 32768) {  // flat RGBE
            for (int x = 0; x < w; ++x) {
                uint8_t e[4];
                Rgbe(row + 3 * x, e);
                o.put(e, 4);
            }
            continue;
        }
        for (int x = 0; x < w; ++x) {
            uint8_t e[4];
            Rgbe(row + 3 * x, e);
            for (int c = 0; c < 4; ++c) plane[(size_t)c * w + x] = e[c];
        }
        const uint8_t lead[4] = {2, 2, (uint8_t)((w & 0xff00) >> 8), (uint8_t)(w & 0xff)};
        o.put(lead, 4);
        // per component: literal spans of at most 128 bytes, runs (three or
        // more equal bytes) of at most 127
        for (int c = 0; c < 4; ++c) {
            const uint8_t* p = &plane[(size_t)c * w];
            int x = 0;
            while (x < w) {
                int r = x;
                while (r + 2 < w && !(p[r] == p[r + 1] && p[r] == p[r + 2])) ++r;
                const bool run = r + 2 < w;
                if (!run) r = w;
                while (x < r) {
                    const int len = r - x > 128 ? 128 : r - x;
                    o.u8((uint8_t)len);
                    o.put(p + x, (size_t)len);
                    x += len;
                }
                if (run) {
                    while (r < w && p[r] == p[x]) ++r;
                    while (x < r) {
                        const int len = r - x > 127 ? 127 : r - x;
                        o.u8((uint8_t)(len + 128));
                        o.u8(p[x]);
                        x += len;
                    }
                }
            }
        }
    }
}

bool WritePng(Out& o, const std::vector<uint8_t>& px, int w, int h) {
    std::vector<uint8_t> raw;
    raw.reserve(((size_t)w * 3 + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), px.begin() + (size_t)y * w * 3, px.begin() + (size_t)(y + 1) * w * 3);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return false;
    auto chunk = [&](const char* type, const uint8_t* data, size_t len) {
        o.be32((uint32_t)len);
        uLong crc = crc32(0L, (const Bytef*)type, 4);
        if (len) crc = crc32(crc, data, (uInt)len);
        o.put(type, 4);
        o.put(data, len);
        o.be32((uint32_t)crc);
    };
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    o.put(sig, 8);
    const uint8_t ihdr[13] = {(uint8_t)(w >> 24), (uint8_t)(w >> 16), (uint8_t)(w >> 8), (uint8_t)w,
                              (uint8_t)(h >> 24), (uint8_t)(h >> 16), (uint8_t)(h >> 8), (uint8_t)h,
                              8, 2, 0, 0, 0};
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), zlen);
    chunk("IEND", nullptr, 0);
    return true;
}

// ---- reading: Radiance .hdr as stbi_loadf returns it (ImageIO::LoadTexture,
// image/ImageIO.cpp:100-114, reads environment maps through stbi_loadf with
// the file's 3 channels).  The decoding follows stb_image's HDR reader
// (external/stb/stb_image.h): header lines up to an empty one, which must
// include FORMAT=32-bit_rle_rgbe; a "-Y h +X w" line; flat RGBE pixels when
// w < 8 or w >= 32768, else per-scanline run-length data per component, with
// a scanline that does not open with 2 2 switching the rest of the file to
// flat pixels from the first one on; texel = c * 2^(e - 136), 0 when e = 0.
struct In {
    std::vector<uint8_t> d;
    size_t p = 0;
    bool eof() const { return p >= d.size(); }
    int get8() { return eof() ? 0 : d[p++]; }
};

std::string HdrToken(In& in) {
    std::string t;
    int c = in.get8();
    while (!in.eof() && c != '\n') {
        t.push_back((char)c);
        if (t.size() == 1023) {
            while (!in.eof() && in.get8() != '\n') {
            }
            break;
        }
        c = in.get8();
    }
    return t;
}

void HdrConvert(float* out, const uint8_t* e) {
    if (e[3] != 0) {
        const float f1 = (float)std::ldexp(1.0f, e[3] - (int)(128 + 8));
        out[0] = e[0] * f1;
        out[1] = e[1] * f1;
        out[2] = e[2] * f1;
    } else {
        out[0] = out[1] = out[2] = 0.f;
    }
}

int ReadHdr(const char* path, int32_t* width, int32_t* height, float* rgb, int64_t max_floats) {
    In in;
    {
        FILE* f = std::fopen(path, "rb");
        if (!f) return Fail(GY_ERR_INVALID, std::string("gy_image_read: cannot open ") + path);
        uint8_t buf[1 << 16];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) in.d.insert(in.d.end(), buf, buf + n);
        std::fclose(f);
    }
    const std::string id = HdrToken(in);
    if (id != "#?RADIANCE" && id != "#?RGBE") return Fail(GY_ERR_INVALID, "gy_image_read: not a Radiance HDR file");
    bool valid = false;
    for (;;) {
        const std::string t = HdrToken(in);
        if (t.empty()) break;
        if (t == "FORMAT=32-bit_rle_rgbe") valid = true;
    }
    if (!valid) return Fail(GY_ERR_UNSUPPORTED, "gy_image_read: unsupported HDR format");
    const std::string res = HdrToken(in);
    const char* tok = res.c_str();
    if (std::strncmp(tok, "-Y ", 3) != 0) return Fail(GY_ERR_UNSUPPORTED, "gy_image_read: unsupported HDR data layout");
    char* end = nullptr;
    const long h = std::strtol(tok + 3, &end, 10);
    tok = end;
    while (*tok == ' ') ++tok;
    if (std::strncmp(tok, "+X ", 3) != 0) return Fail(GY_ERR_UNSUPPORTED, "gy_image_read: unsupported HDR data layout");
    const long w = std::strtol(tok + 3, nullptr, 10);
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 28)) return Fail(GY_ERR_UNSUPPORTED, "gy_image_read: HDR size out of range");
    *width = (int32_t)w;
    *height = (int32_t)h;
    const int64_t nf = (int64_t)w * h * 3;
    if (!rgb) return GY_OK;  // size query
    if (max_floats < nf) return Fail(GY_ERR_INVALID, "gy_image_read: output buffer too small");
    auto need = [&](size_t k) { return in.p + k <= in.d.size(); };
    auto flat_from = [&](int64_t first) {  // flat RGBE pixels from pixel `first` to the end
        for (int64_t i = first; i < (int64_t)w * h; ++i) {
            if (!need(4)) return false;
            HdrConvert(rgb + 3 * i, &in.d[in.p]);
            in.p += 4;
        }
        return true;
    };
    if (w < 8 || w >= 32768) {
        if (!flat_from(0)) return Fail(GY_ERR_INVALID, "gy_image_read: truncated HDR data");
        return GY_OK;
    }
    std::vector<uint8_t> line((size_t)w * 4);
    for (long j = 0; j < h; ++j) {
        if (!need(4)) return Fail(GY_ERR_INVALID, "gy_image_read: truncated HDR data");
        const uint8_t c1 = in.d[in.p], c2 = in.d[in.p + 1], l0 = in.d[in.p + 2];
        if (c1 != 2 || c2 != 2 || (l0 & 0x80)) {
            // not run-length coded: these four bytes are the first pixel of
            // the image, and the rest of the file is flat pixels
            HdrConvert(rgb, &in.d[in.p]);
            in.p += 4;
            if (!flat_from(1)) return Fail(GY_ERR_INVALID, "gy_image_read: truncated HDR data");
            return GY_OK;
        }
        const int len = (l0 << 8) | in.d[in.p + 3];
        in.p += 4;
        if (len != w) return Fail(GY_ERR_INVALID, "gy_image_read: invalid decoded scanline length");
        for (int k = 0; k < 4; ++k) {
            long i = 0;
            while (i < w) {
                if (!need(1)) return Fail(GY_ERR_INVALID, "gy_image_read: truncated HDR data");
                int count = in.d[in.p++];
                const long left = w - i;
                if (count > 128) {
                    if (!need(1)) return Fail(GY_ERR_INVALID, "gy_image_read: truncated HDR data");
                    const uint8_t v = in.d[in.p++];
                    count -= 128;
                    if (count > left) return Fail(GY_ERR_INVALID, "gy_image_read: bad RLE data in HDR");
                    for (int z = 0; z < count; ++z) line[(size_t)(i++) * 4 + k] = v;
                } else {
                    if (count > left) return Fail(GY_ERR_INVALID, "gy_image_read: bad RLE data in HDR");
                    if (!need((size_t)count)) return Fail(GY_ERR_INVALID, "gy_image_read: truncated HDR data");
                    for (int z = 0; z < count; ++z) line[(size_t)(i++) * 4 + k] = in.d[in.p++];
                }
            }
        }
        for (long i = 0; i < w; ++i) HdrConvert(rgb + ((int64_t)j * w + i) * 3, &line[(size_t)i * 4]);
    }
    return GY_OK;
}

bool EndsWith(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

}  // namespace

int WriteImage(const char* path, const float* rgb, int32_t w, int32_t h) {
    if (!path || !rgb || w <= 0 || h <= 0) return Fail(GY_ERR_INVALID, "gy_image_write: need a path, pixels and a positive size");
    const std::string p(path);
    const int kind = EndsWith(p, ".hdr") ? 0 : EndsWith(p, ".bmp") ? 1 : EndsWith(p, ".png") ? 2 : -1;
    if (kind < 0) return Fail(GY_ERR_INVALID, "gy_image_write: the file name must end in .hdr, .bmp or .png");
    std::vector<uint8_t> px;
    if (kind > 0) {
        px.resize((size_t)w * h * 3);
        for (size_t i = 0; i < px.size(); ++i) px[i] = ToByte(rgb[i]);
    }
    Out o;
    o.f = std::fopen(path, "wb");
    if (!o.f) return Fail(GY_ERR_INVALID, "gy_image_write: cannot open " + p);
    bool ok = true;
    if (kind == 0) WriteHdr(o, rgb, w, h);
    else if (kind == 1) WriteBmp(o, px, w, h);
    else ok = WritePng(o, px, w, h);
    ok = ok && o.ok;
    if (std::fclose(o.f) != 0) ok = false;
    return ok ? GY_OK : Fail(GY_ERR_INVALID, "gy_image_write: writing " + p + " failed");
}

}  // namespace gy

extern "C" int gy_image_write(const char* path, const float* rgb, int32_t width, int32_t height) {
    return gy::WriteImage(path, rgb, width, height);
}

extern "C" int gy_image_read(const char* path, int32_t* width, int32_t* height, float* rgb, int64_t max_floats) {
    if (!path || !width || !height) return gy::Fail(GY_ERR_INVALID, "gy_image_read: null argument");
    const std::string p(path);
    if (!gy::EndsWith(p, ".hdr")) return gy::Fail(GY_ERR_UNSUPPORTED, "gy_image_read: only Radiance .hdr files are read");
    return gy::ReadHdr(path, width, height, rgb, max_floats);
}

extern "C" int gy_image_to_bytes(const float* rgb, int64_t n, uint8_t* out) {
    if ((!rgb || !out) && n > 0) return gy::Fail(GY_ERR_INVALID, "gy_image_to_bytes: null argument");
    for (int64_t i = 0; i < n; ++i) out[i] = gy::ToByte(rgb[i]);
    return GY_OK;
}
