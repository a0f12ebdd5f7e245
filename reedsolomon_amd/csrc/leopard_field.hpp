// leopard_field.hpp — host-side Leopard GF(2^16) / GF(2^8) field tables and
// error-locator evaluation, restated from leopard.go:984-1274 and
// leopard8.go:919-1158.  Byte work runs on the GPU; these tables are built
// once per process and uploaded.
#pragma once

#include <cstdint>
#include <vector>

namespace rsx {

// One Leopard field: bits = 16 (polynomial 0x1002D, Cantor basis of
// leopard.go:1105-1110) or 8 (polynomial 0x11D, leopard8.go:1030-1032).
struct LeoField {
    int bits = 0;
    unsigned order = 0, modulus = 0;
    std::vector<uint16_t> exp, log, skew, walsh;  // skew has `modulus` entries

    explicit LeoField(int bits);
    uint16_t add_mod(uint32_t a, uint32_t b) const;  // partial reduction, may return modulus
    uint16_t mul_log(uint16_t a, uint16_t log_b) const;
    // Decimation-in-time FWHT over the first `mtrunc` nonzero entries
    // (leopard.go:1025-1058; mod-modulus partial reduction).
    void fwht(uint16_t *data, int mtrunc) const;

    // Device multiply tables, one record per log value (order records):
    //  bits 8:  5 dwords  {c*(0..3), c*(4..7), c*((0..3)<<3), c*((4..7)<<3), c*((0..3)<<6)}
    //  bits 16: 20 dwords, for the split lo/hi byte layout of a 64-byte chunk:
    //    [0,1] lo-out of lo bits 0-2, [2,3] hi-out of lo bits 0-2,
    //    [4,5] lo-out of lo bits 3-5, [6,7] hi-out of lo bits 3-5,
    //    [8]   lo-out of lo bits 6-7, [9]   hi-out of lo bits 6-7,
    //    [10..19] the same for the hi byte.
    int table_dwords() const { return bits == 16 ? 20 : 5; }
    std::vector<uint32_t> mul_tables() const;
};

const LeoField &leo_field(int bits);

// Error locator logs of leopard.go:459-510 (GF16) / leopard8.go:437-477
// (GF8) for one erasure pattern: erased[i] for shard i (data first, then
// parity).  Returns the first `n` entries of errLocs.
std::vector<uint16_t> leo_error_locators(const LeoField &f, int k, int p, const uint8_t *erased,
                                         int *input_count);

// The first transform's input of leo_error_locators: ind[i] (m + k entries,
// m = ceilPow2(p)) is 1 for an erased parity row, a padding row [p, m) or an
// erased data row (at m + i).  Returns the IFFT's input count (one past the
// last present data row, or p).
int leo_erasure_indicator(int k, int p, const uint8_t *erased, std::vector<uint8_t> &ind);

int ceil_pow2(int n);

}  // namespace rsx
