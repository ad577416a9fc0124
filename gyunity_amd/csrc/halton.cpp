// Halton sampler tables and the Halton pixel enumeration (host side).
//
// HaltonSampler with Faure permutations (sampler/Halton.h:39-46, 595-625,
// 648-923): dimension d uses the (d+1)-th prime b.  Base 2 is the bit-reversed
// index written into a float mantissa (Halton.h:1441-1452); every other base
// reads a digit table of b^k entries (b^k the largest power <= 512), each entry
// the Faure-permuted, digit-reversed value of its k digits, over G groups of k
// digits (b^(kG) the largest such power <= 2^32), and scales the integer sum by
// float((1 - 2^-23) / b^(kG)).  The device evaluates the same sum from these
// tables (device.h, halton_sample).
//
// HaltonEnum (sampler/HaltonEnum.h:61-143): the i-th sample of pixel (x, y) is
// the index whose base-2 and base-3 radical inverses fall in that pixel of a
// 2^p2 x 3^p3 grid, found with the Chinese remainder theorem.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "gy_internal.h"

namespace gy {
namespace halton {
namespace {

std::vector<uint32_t> FirstPrimes(int n) {
    std::vector<uint32_t> p;
    for (uint32_t c = 2; (int)p.size() < n; ++c) {
        bool prime = true;
        for (uint32_t q : p) {
            if (q * q > c) break;
            if (c % q == 0) {
                prime = false;
                break;
            }
        }
        if (prime) p.push_back(c);
    }
    return p;
}

// Faure's permutations: identity for bases 1-3; an even base 2c interleaves
// 2*sigma_c and 2*sigma_c + 1; an odd base 2c+1 inserts the centre c into
// sigma_2c, shifting the values >= c up by one.
std::vector<std::vector<uint16_t>> FaurePermutations(uint32_t max_base) {
    std::vector<std::vector<uint16_t>> s(max_base + 1);
    for (uint32_t b = 1; b <= 3 && b <= max_base; ++b) {
        s[b].resize(b);
        for (uint32_t i = 0; i < b; ++i) s[b][i] = (uint16_t)i;
    }
    for (uint32_t b = 4; b <= max_base; ++b) {
        s[b].assign(b, 0);
        const uint32_t c = b / 2;
        if (b % 2 == 0) {
            for (uint32_t i = 0; i < c; ++i) {
                s[b][i] = (uint16_t)(2 * s[c][i]);
                s[b][c + i] = (uint16_t)(2 * s[c][i] + 1);
            }
        } else {
            const std::vector<uint16_t>& e = s[b - 1];
            for (uint32_t i = 0; i < b - 1; ++i) {
                const uint32_t to = i < c ? i : i + 1;
                s[b][to] = (uint16_t)(e[i] < c ? e[i] : e[i] + 1);
            }
            s[b][c] = (uint16_t)c;
        }
    }
    return s;
}

struct State {
    std::vector<DimInfo> dims;
    std::vector<uint16_t> tables;
};

const State& Get() {
    static State st;
    static std::once_flag once;
    std::call_once(once, [] {
        const std::vector<uint32_t> primes = FirstPrimes(kDims);
        const auto perms = FaurePermutations(primes.back());
        st.dims.resize(kDims);
        st.dims[0] = DimInfo{2, 0, 0, 0, 0.f};  // base 2: bit reversal, no table
        for (int d = 1; d < kDims; ++d) {
            const uint32_t b = primes[d];
            uint32_t k = 1, tsize = b;
            while ((uint64_t)tsize * b <= 512) {
                tsize *= b;
                ++k;
            }
            uint32_t groups = 1;
            uint64_t total = tsize;
            while (total * tsize <= (1ull << 32)) {
                total *= tsize;
                ++groups;
            }
            DimInfo& di = st.dims[d];
            di.base = b;
            di.tsize = tsize;
            di.groups = groups;
            di.offset = (uint32_t)st.tables.size();
            di.scale = (float)(0x1.fffffcp-1 / (double)total);
            // entry i: the k digits of i, least significant first, permuted and
            // read back most significant first
            for (uint32_t i = 0; i < tsize; ++i) {
                uint32_t v = 0, x = i;
                for (uint32_t j = 0; j < k; ++j) {
                    v = v * b + perms[b][x % b];
                    x /= b;
                }
                st.tables.push_back((uint16_t)v);
            }
        }
    });
    return st;
}

uint64_t ReverseBits64(uint64_t v) {
    uint64_t r = 0;
    for (int i = 0; i < 64; ++i) {
        r = (r << 1) | (v & 1);
        v >>= 1;
    }
    return r;
}

}  // namespace

const std::vector<DimInfo>& Dims() { return Get().dims; }
const std::vector<uint16_t>& Tables() { return Get().tables; }

float Sample(uint64_t index, uint32_t dim) {
    if (dim == 0) {
        const uint32_t bits = 0x3f800000u | (uint32_t)(ReverseBits64(index) >> 41);
        float f;
        std::memcpy(&f, &bits, 4);
        return f - 1.f;
    }
    if (dim >= (uint32_t)kDims) return 0.f;
    const State& st = Get();
    const DimInfo& di = st.dims[dim];
    const uint16_t* t = st.tables.data() + di.offset;
    uint64_t sum = 0, x = index;
    for (uint32_t g = 0; g < di.groups; ++g) {
        sum = sum * di.tsize + t[x % di.tsize];
        x /= di.tsize;
    }
    return (float)sum * di.scale;
}

EnumParams Enum(uint32_t width, uint32_t height) {
    EnumParams e = {};
    uint32_t w = 1, h = 1;
    while (w < width) {
        w *= 2;
        ++e.p2;
    }
    while (h < height) {
        h *= 3;
        ++e.p3;
    }
    e.sx = (float)w;
    e.sy = (float)h;
    e.inc = (uint64_t)w * h;
    // inverses: inv2 = h^-1 mod w, inv3 = w^-1 mod h (extended Euclid)
    auto inverse = [](int64_t a, int64_t m) -> int64_t {
        if (m == 1) return 0;
        int64_t r0 = m, r1 = a % m, s0 = 0, s1 = 1;
        while (r1) {
            const int64_t q = r0 / r1, r2 = r0 - q * r1, s2 = s0 - q * s1;
            r0 = r1;
            r1 = r2;
            s0 = s1;
            s1 = s2;
        }
        return ((s0 % m) + m) % m;
    };
    e.mx = (uint64_t)h * (uint64_t)inverse(h, w);
    e.my = (uint64_t)w * (uint64_t)inverse(w, h);
    return e;
}

uint64_t EnumIndex(const EnumParams& e, uint64_t i, uint32_t x, uint32_t y) {
    const uint64_t hx = e.p2 ? ReverseBits64(x) >> (64 - e.p2) : 0;
    uint64_t hy = 0, yy = y;
    for (uint32_t d = 0; d < e.p3; ++d) {
        hy = hy * 3 + yy % 3;
        yy /= 3;
    }
    return (hx * e.mx + hy * e.my) % e.inc + i * e.inc;
}

}  // namespace halton
}  // namespace gy

extern "C" float gy_halton_sample(uint64_t index, uint32_t dim) { return gy::halton::Sample(index, dim); }
extern "C" uint64_t gy_halton_enum_index(uint32_t width, uint32_t height, uint64_t i, uint32_t x, uint32_t y) {
    if (width == 0 || height == 0) return 0;
    return gy::halton::EnumIndex(gy::halton::Enum(width, height), i, x, y);
}
