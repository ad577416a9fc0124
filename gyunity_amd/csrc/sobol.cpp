// Sobol generator matrices and per-pixel sample enumeration (host side).
//
// Matrices: the Bratley-Fox recurrence over the Joe-Kuo (2008) parameters in
// sobol_joe_kuo.inc, expanded to 52 columns of 32 bits per dimension: the same
// numbers the reference tabulates in Gyunity/src/sampler/Sobol.cpp:38-10689 and
// samples with sobol::sample (sampler/Sobol.h:46-61).
//
// Enumeration: SobolEnum::look_up (sampler/SobolEnum.h:42-72) returns, for a
// 2^m x 2^m pixel grid, the index of the frame-th Sobol point whose first two
// coordinates fall in pixel (px, py).  Here that index is derived directly from
// the generator matrices: for index i = (frame << 2m) | lo, the top m bits of
// dimensions 0 and 1 are a GF(2)-linear function of lo and frame,
//     (px << m | py) = A lo  xor  F frame,
// so lo = A^-1 ((px << m | py) xor F frame).  A is invertible because the first
// two Sobol dimensions form a (0,2)-sequence; the index is unique, so any
// correct method returns the reference's index.
#include "gy_internal.h"

#include <cstdio>
#include <cstring>
#include <mutex>

namespace gy {
namespace sobol {

static const uint16_t kParams[] = {
#include "sobol_joe_kuo.inc"
};

static std::once_flag g_once;
static std::vector<uint32_t> g_matrices;  // kDims * 52
static EnumTables g_enum[kMaxLog2Res + 1];

static void build() {
    g_matrices.assign((size_t)kDims * kCols, 0u);
    for (int c = 0; c < 32; ++c) g_matrices[c] = 1u << (31 - c);  // dimension 0: van der Corput
    size_t p = 0;
    for (int d = 1; d < kDims; ++d) {
        int s = kParams[p++];
        uint32_t a = kParams[p++];
        uint64_t m[kCols];
        for (int k = 0; k < s; ++k) m[k] = kParams[p++];
        // m_k = 2 a_1 m_{k-1} ^ 4 a_2 m_{k-2} ^ ... ^ 2^s m_{k-s} ^ m_{k-s}
        for (int k = s; k < kCols; ++k) {
            uint64_t v = m[k - s] ^ (m[k - s] << s);
            for (int j = 1; j < s; ++j)
                if ((a >> (s - 1 - j)) & 1u) v ^= m[k - j] << j;
            m[k] = v;
        }
        // column c holds the 32-bit binary fraction of m_{c+1} / 2^{c+1}
        for (int c = 0; c < kCols; ++c) {
            unsigned __int128 w = (unsigned __int128)m[c] << 32;
            g_matrices[(size_t)d * kCols + c] = (uint32_t)(w >> (c + 1));
        }
    }
    // Enumeration tables for every resolution 2^m.
    for (int m = 1; m <= kMaxLog2Res; ++m) {
        EnumTables& t = g_enum[m];
        const int n = 2 * m;
        auto top = [&](int dim, int col) -> uint64_t { return g_matrices[(size_t)dim * kCols + col] >> (32 - m); };
        // Columns of A: bit layout (x << m) | y of the pixel reached by index bit k.
        uint64_t A[64];
        for (int k = 0; k < n; ++k) A[k] = (top(0, k) << m) | top(1, k);
        // Invert A over GF(2): solve A x = e_b for every unit vector e_b.
        // Gaussian elimination on the augmented rows.
        uint64_t rowsA[64], rowsI[64];
        for (int r = 0; r < n; ++r) {
            uint64_t row = 0;
            for (int k = 0; k < n; ++k)
                if ((A[k] >> r) & 1u) row |= 1ull << k;
            rowsA[r] = row;
            rowsI[r] = 1ull << r;
        }
        for (int col = 0; col < n; ++col) {
            int piv = -1;
            for (int r = col; r < n; ++r)
                if ((rowsA[r] >> col) & 1u) { piv = r; break; }
            if (piv < 0) {
                std::fprintf(stderr, "gyunity: Sobol enumeration matrix singular at m=%d\n", m);
                std::abort();
            }
            std::swap(rowsA[piv], rowsA[col]);
            std::swap(rowsI[piv], rowsI[col]);
            for (int r = 0; r < n; ++r)
                if (r != col && ((rowsA[r] >> col) & 1u)) {
                    rowsA[r] ^= rowsA[col];
                    rowsI[r] ^= rowsI[col];
                }
        }
        // Now row r of A^-1 is rowsI[r]: lo bit r = parity(rowsI[r] & target).
        // Store A^-1 by columns: inv[b] = lo contribution of target bit b.
        for (int b = 0; b < n; ++b) {
            uint64_t col = 0;
            for (int r = 0; r < n; ++r)
                if ((rowsI[r] >> b) & 1u) col |= 1ull << r;
            t.inv[b] = col;
        }
        for (int b = n; b < 64; ++b) t.inv[b] = 0;
        for (int j = 0; j < kCols - n && j < 64; ++j) t.frame[j] = (top(0, n + j) << m) | top(1, n + j);
        for (int j = kCols - n; j < 64; ++j) t.frame[j] = 0;
        t.m = m;
    }
}

const uint32_t* Matrices() {
    std::call_once(g_once, build);
    return g_matrices.data();
}

const uint32_t* ByteTables() {
    static std::vector<uint32_t> t;
    static std::once_flag once;
    std::call_once(once, [] {
        const uint32_t* M = Matrices();
        t.assign((size_t)kDims * kBytes * 256, 0u);
        for (int d = 0; d < kDims; ++d)
            for (int k = 0; k < kBytes; ++k)
                for (int v = 1; v < 256; ++v) {
                    uint32_t r = 0;
                    for (int j = 0; j < 8; ++j)
                        if (((v >> j) & 1) && 8 * k + j < kCols) r ^= M[(size_t)d * kCols + 8 * k + j];
                    t[((size_t)d * kBytes + k) * 256 + v] = r;
                }
    });
    return t.data();
}

const EnumTables& Enum(int m) {
    std::call_once(g_once, build);
    return g_enum[m];
}

float Sample(uint64_t index, uint32_t dim) {
    const uint32_t* M = Matrices() + (size_t)dim * kCols;
    uint32_t r = 0;
    for (int c = 0; index; index >>= 1, ++c)
        if (index & 1u) r ^= M[c];
    return (float)r * (1.f / 4294967296.0f);
}

uint64_t EnumIndex(uint32_t m, uint32_t frame, uint32_t px, uint32_t py) {
    if (m == 0) return 0;
    const EnumTables& t = Enum((int)m);
    uint64_t delta = 0;
    for (uint32_t j = 0, f = frame; f; f >>= 1, ++j)
        if (f & 1u) delta ^= t.frame[j];
    uint64_t target = (((uint64_t)px << m) | py) ^ delta;
    uint64_t lo = 0;
    for (uint32_t b = 0; target; target >>= 1, ++b)
        if (target & 1u) lo ^= t.inv[b];
    return ((uint64_t)frame << (2 * m)) | lo;
}

}  // namespace sobol
}  // namespace gy

extern "C" float gy_sobol_sample(uint64_t index, uint32_t dim) {
    if (dim >= (uint32_t)gy::sobol::kDims) return -1.f;
    return gy::sobol::Sample(index, dim);
}

extern "C" uint64_t gy_sobol_enum_index(uint32_t m, uint32_t frame, uint32_t px, uint32_t py) {
    if (m > (uint32_t)gy::sobol::kMaxLog2Res) return ~0ull;
    return gy::sobol::EnumIndex(m, frame, px, py);
}
