// gf256.cpp — see gf256.hpp.
#include "gf256.hpp"

#include "../../include/rsx.h"

namespace rsx {

const Gf256 &gf() {
    static const Gf256 field;
    return field;
}

Matrix multiply(const Matrix &a, const Matrix &b) {
    const Gf256 &f = gf();
    Matrix out(a.rows, b.cols);
    for (int r = 0; r < a.rows; ++r)
        for (int c = 0; c < b.cols; ++c) {
            uint8_t acc = 0;
            for (int i = 0; i < a.cols; ++i) acc ^= f.mul[a.at(r, i)][b.at(i, c)];
            out.at(r, c) = acc;
        }
    return out;
}

// Gauss-Jordan on [m | I] (matrix.go:192-265): pivot by swapping with the
// first lower row that has a nonzero entry, scale the pivot row to 1,
// eliminate below, then above.
bool invert_in_place(uint8_t *a, int n) {
    const Gf256 &f = gf();
    const int w = 2 * n;
    auto A = [&](int r, int c) -> uint8_t & { return a[static_cast<size_t>(r) * w + c]; };
    for (int r = 0; r < n; ++r) {
        if (A(r, r) == 0) {
            for (int below = r + 1; below < n; ++below)
                if (A(below, r) != 0) {
                    for (int c = 0; c < w; ++c) std::swap(A(r, c), A(below, c));
                    break;
                }
        }
        if (A(r, r) == 0) return false;
        if (A(r, r) != 1) {
            const uint8_t s = f.inv(A(r, r));
            for (int c = 0; c < w; ++c) A(r, c) = f.mul[A(r, c)][s];
        }
        for (int below = r + 1; below < n; ++below) {
            const uint8_t s = A(below, r);
            if (s)
                for (int c = 0; c < w; ++c) A(below, c) ^= f.mul[s][A(r, c)];
        }
    }
    for (int d = 0; d < n; ++d)
        for (int above = 0; above < d; ++above) {
            const uint8_t s = A(above, d);
            if (s)
                for (int c = 0; c < w; ++c) A(above, c) ^= f.mul[s][A(d, c)];
        }
    return true;
}

bool invert(const Matrix &m, Matrix &out) {
    const int n = m.rows;
    const int w = 2 * n;
    std::vector<uint8_t> a(static_cast<size_t>(n) * w, 0);
    for (int r = 0; r < n; ++r) {
        std::memcpy(&a[static_cast<size_t>(r) * w], m.row(r), n);
        a[static_cast<size_t>(r) * w + n + r] = 1;
    }
    if (!invert_in_place(a.data(), n)) return false;
    out = Matrix(n, n);
    for (int r = 0; r < n; ++r) std::memcpy(out.row(r), &a[static_cast<size_t>(r) * w + n], n);
    return true;
}

static Matrix vandermonde(int rows, int cols) {
    const Gf256 &f = gf();
    Matrix v(rows, cols);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) v.at(r, c) = f.pow(static_cast<uint8_t>(r), c);
    return v;
}

// reedsolomon.go:263-342: the matrix Jerasure builds (first parity row and
// column all ones).
static Matrix jerasure(int k, int total) {
    const Gf256 &f = gf();
    Matrix vm = vandermonde(total, k);
    vm.at(0, 0) = 1;
    for (int i = 1; i < k; ++i) vm.at(0, i) = 0;
    for (int i = 0; i < k - 1; ++i) vm.at(total - 1, i) = 0;
    vm.at(total - 1, k - 1) = 1;
    for (int i = 0; i < k; ++i) {
        int r = i;
        while (r < total && vm.at(r, i) == 0) ++r;
        if (r != i)
            for (int c = 0; c < k; ++c) std::swap(vm.at(r, c), vm.at(i, c));
        if (vm.at(i, i) != 1) {
            const uint8_t t = f.inv(vm.at(i, i));
            for (int j = 0; j < total; ++j) vm.at(j, i) = f.mul[vm.at(j, i)][t];
        }
        for (int j = 0; j < k; ++j) {
            const uint8_t t = vm.at(i, j);
            if (j != i && t)
                for (int rr = 0; rr < total; ++rr) vm.at(rr, j) ^= f.mul[t][vm.at(rr, i)];
        }
    }
    for (int j = 0; j < k; ++j) {
        uint8_t t = vm.at(k, j);
        if (t != 1) {
            t = f.inv(t);
            for (int i = k; i < total; ++i) vm.at(i, j) = f.mul[vm.at(i, j)][t];
        }
    }
    for (int i = k + 1; i < total; ++i) {
        uint8_t t = vm.at(i, 0);
        if (t != 1) {
            t = f.inv(t);
            for (int j = 0; j < k; ++j) vm.at(i, j) = f.mul[vm.at(i, j)][t];
        }
    }
    return vm;
}

int build_coding_matrix(int k, int m, int kind, bool fast_one_parity, const uint8_t *custom,
                        int custom_rows, int custom_cols, Matrix &out) {
    const Gf256 &f = gf();
    const int total = k + m;
    out = Matrix(total, k);
    if (custom) {  // reedsolomon.go:462-477
        if (custom_rows < m || custom_cols < k) return RSX_ERR_MATRIX;
        for (int i = 0; i < k; ++i) out.at(i, i) = 1;
        for (int r = 0; r < m; ++r)
            std::memcpy(out.row(k + r), custom + static_cast<size_t>(r) * custom_cols, k);
        return RSX_OK;
    }
    if (fast_one_parity && m == 1) {  // reedsolomon.go:391-413
        for (int i = 0; i < k; ++i) out.at(i, i) = 1;
        for (int c = 0; c < k; ++c) out.at(k, c) = 1;
        return RSX_OK;
    }
    switch (kind) {
    case RSX_MATRIX_CAUCHY:  // reedsolomon.go:369-389
        for (int r = 0; r < total; ++r)
            for (int c = 0; c < k; ++c)
                out.at(r, c) = r < k ? (r == c) : f.inv(static_cast<uint8_t>(r ^ c));
        return RSX_OK;
    case RSX_MATRIX_PAR1:  // reedsolomon.go:348-367
        for (int r = 0; r < total; ++r)
            for (int c = 0; c < k; ++c)
                out.at(r, c) = r < k ? (r == c) : f.pow(static_cast<uint8_t>(c + 1), r - k);
        return RSX_OK;
    case RSX_MATRIX_JERASURE:
        out = jerasure(k, total);
        return RSX_OK;
    default: {  // reedsolomon.go:232-256: V * inv(V[:k])
        Matrix vm = vandermonde(total, k);
        Matrix top(k, k), top_inv;
        for (int r = 0; r < k; ++r) std::memcpy(top.row(r), vm.row(r), k);
        if (!invert(top, top_inv)) return RSX_ERR_SINGULAR;
        out = multiply(vm, top_inv);
        return RSX_OK;
    }
    }
}

void coefficient_table(uint8_t c, uint32_t out[kTableDwords]) {
    const Gf256 &f = gf();
    uint8_t t[20];
    for (int v = 0; v < 8; ++v) {
        t[v] = f.mul[c][v];
        t[8 + v] = f.mul[c][v << 3];
    }
    for (int v = 0; v < 4; ++v) t[16 + v] = f.mul[c][v << 6];
    std::memcpy(out, t, sizeof t);
}

}  // namespace rsx
