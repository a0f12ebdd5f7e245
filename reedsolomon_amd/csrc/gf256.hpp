// gf256.hpp — host-side GF(2^8) arithmetic and matrix planning for the
// classic engine: the coding matrices of New (reedsolomon.go:232-413,
// 461-488), Gauss-Jordan inversion (matrix.go:192-265) and the decode rows of
// reconstruct (reedsolomon.go:620-673, 1714-1774).  Byte work stays on the
// GPU; this file only plans the small k x k matrices.
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

namespace rsx {

// Field of 256 elements, generator 2, polynomial x^8 + x^4 + x^3 + x^2 + 1
// (the reference's generatingPolynomial 29, galois.go:26).
struct Gf256 {
    uint8_t exp[510];
    uint8_t log[256];
    uint8_t mul[256][256];

    Gf256() {
        unsigned x = 1;
        for (int i = 0; i < 255; ++i) {
            exp[i] = exp[i + 255] = static_cast<uint8_t>(x);
            log[x] = static_cast<uint8_t>(i);
            x <<= 1;
            if (x & 0x100) x ^= 0x11d;
        }
        log[0] = 0;
        for (int a = 0; a < 256; ++a)
            for (int b = 0; b < 256; ++b)
                mul[a][b] = (a && b) ? exp[log[a] + log[b]] : 0;
    }
    uint8_t inv(uint8_t a) const { return a ? exp[255 - log[a]] : 0; }
    // a**n with a**0 = 1 and 0**n = 0 for n > 0 (galois.go:898-912).
    uint8_t pow(uint8_t a, int n) const {
        if (n == 0) return 1;
        if (a == 0) return 0;
        return exp[(log[a] * n) % 255];
    }
};

const Gf256 &gf();

// Row-major byte matrix.
struct Matrix {
    int rows = 0, cols = 0;
    std::vector<uint8_t> v;
    Matrix() = default;
    Matrix(int r, int c) : rows(r), cols(c), v(static_cast<size_t>(r) * c, 0) {}
    uint8_t *row(int r) { return v.data() + static_cast<size_t>(r) * cols; }
    const uint8_t *row(int r) const { return v.data() + static_cast<size_t>(r) * cols; }
    uint8_t &at(int r, int c) { return v[static_cast<size_t>(r) * cols + c]; }
    uint8_t at(int r, int c) const { return v[static_cast<size_t>(r) * cols + c]; }
};

Matrix multiply(const Matrix &a, const Matrix &b);
// Returns false when singular.
bool invert(const Matrix &m, Matrix &out);
// Gauss-Jordan on an n x 2n row-major array [m | I] (matrix.go:192-265):
// on success the right half holds inv(m).  Allocates nothing.
bool invert_in_place(uint8_t *aug, int n);

// Coding matrix (k+m) x k of New; kind is one of RSX_MATRIX_* (fast_one_parity
// applies only when m == 1).  custom, when non-null, holds the parity rows.
int build_coding_matrix(int k, int m, int kind, bool fast_one_parity, const uint8_t *custom,
                        int custom_rows, int custom_cols, Matrix &out);

// Bit-group product tables for the device kernels.  A byte x is split into
// bit groups [0:3), [3:6), [6:8); c*x is the XOR of three table lookups
// c*(x & 7), c*(x & 0x38), c*(x & 0xC0), each a byte select of v_perm_b32.
// Per coefficient: 5 dwords {lo0..3, lo4..7, mid0..3, mid4..7, hi0..3}.
constexpr int kTableDwords = 5;
void coefficient_table(uint8_t c, uint32_t out[kTableDwords]);

}  // namespace rsx
