// leopard_field.cpp — see leopard_field.hpp.
#include "leopard_field.hpp"

#include <cstring>

namespace rsx {

int ceil_pow2(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

LeoField::LeoField(int b) : bits(b), order(1u << b), modulus((1u << b) - 1) {
    static const uint16_t cantor16[16] = {0x0001, 0xACCA, 0x3C0E, 0x163E, 0xC582, 0xED2E,
                                          0x914C, 0x4012, 0x6C98, 0x10D8, 0x6A72, 0xB900,
                                          0xFDB8, 0xFB34, 0xFF38, 0x991E};
    static const uint16_t cantor8[8] = {1, 214, 152, 146, 86, 200, 88, 230};
    const uint16_t *cantor = bits == 16 ? cantor16 : cantor8;
    const unsigned poly = bits == 16 ? 0x1002D : 0x11D;
    exp.assign(order, 0);
    log.assign(order, 0);
    // LFSR, then conversion of the log table to the Cantor basis
    // (leopard.go:1104-1146).
    unsigned state = 1;
    for (unsigned i = 0; i < modulus; ++i) {
        exp[state] = static_cast<uint16_t>(i);
        state <<= 1;
        if (state >= order) state ^= poly;
    }
    exp[0] = static_cast<uint16_t>(modulus);
    log[0] = 0;
    for (int i = 0; i < bits; ++i) {
        const int width = 1 << i;
        for (int j = 0; j < width; ++j) log[j + width] = log[j] ^ cantor[i];
    }
    for (unsigned i = 0; i < order; ++i) log[i] = exp[log[i]];
    for (unsigned i = 0; i < order; ++i) exp[log[i]] = static_cast<uint16_t>(i);
    exp[modulus] = exp[0];

    // FFT skew (leopard.go:1150-1195).
    skew.assign(modulus, 0);
    walsh.assign(order, 0);
    std::vector<uint16_t> temp(bits - 1);
    for (int i = 1; i < bits; ++i) temp[i - 1] = static_cast<uint16_t>(1u << i);
    for (int m = 0; m < bits - 1; ++m) {
        const int step = 1 << (m + 1);
        skew[(1 << m) - 1] = 0;
        for (int i = m; i < bits - 1; ++i) {
            const int s = 1 << (i + 1);
            for (int j = (1 << m) - 1; j < s; j += step) skew[j + s] = skew[j] ^ temp[i];
        }
        temp[m] = static_cast<uint16_t>(modulus - log[mul_log(temp[m], log[temp[m] ^ 1])]);
        for (int i = m + 1; i < bits - 1; ++i) {
            const uint16_t sum = add_mod(log[temp[i] ^ 1], temp[m]);
            temp[i] = mul_log(temp[i], sum);
        }
    }
    for (unsigned i = 0; i < modulus; ++i) skew[i] = log[skew[i]];
    for (unsigned i = 0; i < order; ++i) walsh[i] = log[i];
    walsh[0] = 0;
    fwht(walsh.data(), static_cast<int>(order));
}

uint16_t LeoField::add_mod(uint32_t a, uint32_t b) const {
    // leopard.go:1001-1006: ffe(sum + sum>>bitwidth), truncated to the field width.
    const uint64_t sum = static_cast<uint64_t>(a) + b;
    return static_cast<uint16_t>((sum + (sum >> bits)) & (order - 1));
}

uint16_t LeoField::mul_log(uint16_t a, uint16_t log_b) const {
    if (a == 0) return 0;
    return exp[add_mod(log[a], log_b)];
}

void LeoField::fwht(uint16_t *data, int mtrunc) const {
    // 32-bit lanes (sums stay below 2^17), so the loop over `off` vectorizes;
    // a - b wraps modulo 2^32 and folds to the same residue the reference's
    // 64-bit uint arithmetic gives (leopard.go:1082-1091).
    const uint32_t mask = order - 1;
    const int sh = bits;
    int dist = 1, dist4 = 4;
    while (dist4 <= static_cast<int>(order)) {
        for (int r = 0; r < mtrunc; r += dist4) {
            uint16_t *p0 = data + r, *p1 = p0 + dist, *p2 = p1 + dist, *p3 = p2 + dist;
            for (int off = 0; off < dist; ++off) {
                const uint32_t t0 = p0[off], t1 = p1[off], t2 = p2[off], t3 = p3[off];
                uint32_t s01 = t0 + t1, d01 = t0 - t1, s23 = t2 + t3, d23 = t2 - t3;
                s01 = (s01 + (s01 >> sh)) & mask;
                d01 = (d01 + (d01 >> sh)) & mask;
                s23 = (s23 + (s23 >> sh)) & mask;
                d23 = (d23 + (d23 >> sh)) & mask;
                uint32_t a = s01 + s23, c = s01 - s23, b = d01 + d23, d = d01 - d23;
                p0[off] = static_cast<uint16_t>((a + (a >> sh)) & mask);
                p2[off] = static_cast<uint16_t>((c + (c >> sh)) & mask);
                p1[off] = static_cast<uint16_t>((b + (b >> sh)) & mask);
                p3[off] = static_cast<uint16_t>((d + (d >> sh)) & mask);
            }
        }
        dist = dist4;
        dist4 <<= 2;
    }
}

std::vector<uint32_t> LeoField::mul_tables() const {
    const int tw = table_dwords();
    std::vector<uint32_t> out(static_cast<size_t>(order) * tw);
    for (unsigned lm = 0; lm < order; ++lm) {
        uint32_t *rec = &out[static_cast<size_t>(lm) * tw];
        auto prod = [&](unsigned x) { return mul_log(static_cast<uint16_t>(x), static_cast<uint16_t>(lm)); };
        if (bits == 8) {
            uint8_t t[20];
            for (int v = 0; v < 8; ++v) {
                t[v] = static_cast<uint8_t>(prod(v));
                t[8 + v] = static_cast<uint8_t>(prod(v << 3));
            }
            for (int v = 0; v < 4; ++v) t[16 + v] = static_cast<uint8_t>(prod(v << 6));
            std::memcpy(rec, t, sizeof t);
            continue;
        }
        // bits 16: for each source byte (lo, hi) and group (0-2, 3-5, 6-7):
        // tables of the product's lo byte and hi byte.
        uint8_t t[80];
        int pos = 0;
        for (int src = 0; src < 2; ++src) {
            const int base = src * 8;
            for (int grp = 0; grp < 3; ++grp) {
                const int shift = base + grp * 3;
                const int entries = grp == 2 ? 4 : 8;
                for (int ob = 0; ob < 2; ++ob)
                    for (int v = 0; v < entries; ++v)
                        t[pos++] = static_cast<uint8_t>(prod(static_cast<unsigned>(v) << shift) >> (8 * ob));
            }
        }
        std::memcpy(rec, t, sizeof t);
    }
    return out;
}

const LeoField &leo_field(int bits) {
    static const LeoField f16(16), f8(8);
    return bits == 16 ? f16 : f8;
}

int leo_erasure_indicator(int k, int p, const uint8_t *erased, std::vector<uint8_t> &ind) {
    const int m = ceil_pow2(p);
    ind.assign(static_cast<size_t>(m + k), 0);
    int ic = p;
    for (int i = 0; i < p; ++i)
        if (erased[k + i]) ind[i] = 1;
    for (int i = p; i < m; ++i) ind[i] = 1;
    for (int i = 0; i < k; ++i) {
        if (erased[i]) ind[i + m] = 1;
        else ic = m + i + 1;
    }
    return ic;
}

std::vector<uint16_t> leo_error_locators(const LeoField &f, int k, int p, const uint8_t *erased,
                                         int *input_count) {
    const int m = ceil_pow2(p);
    const int n = ceil_pow2(m + k);
    std::vector<uint8_t> ind;
    const int ic = leo_erasure_indicator(k, p, erased, ind);
    std::vector<uint16_t> e(f.order, 0);
    for (size_t i = 0; i < ind.size(); ++i) e[i] = ind[i];
    f.fwht(e.data(), m + k);
    // (e * walsh) % modulus (leopard.go:506-508), reduced by folding: 2^bits
    // is 1 modulo 2^bits - 1, so no division per entry.
    const uint32_t fmask = f.modulus;
    for (unsigned i = 0; i < f.order; ++i) {
        uint32_t x = static_cast<uint32_t>(e[i]) * f.walsh[i];
        x = (x & fmask) + (x >> f.bits);
        x = (x & fmask) + (x >> f.bits);
        if (x >= f.modulus) x -= f.modulus;
        e[i] = static_cast<uint16_t>(x);
    }
    f.fwht(e.data(), static_cast<int>(f.order));
    if (input_count) *input_count = ic;
    e.resize(n);
    return e;
}

}  // namespace rsx
