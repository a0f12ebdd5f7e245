// kernels.hpp — device kernels of the engine and their host launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rsx {

// Inputs of one GF(2^8) matrix pass (the device form of reedsolomon.go:882
// codeSomeShards(matrixRows, inputs, outputs, byteCount, clear)).
constexpr int kMaxIn = 256;   // classic engine: k <= 256
constexpr int kMaxOut = 16;   // outputs per pass; more outputs => more passes

// Outputs of the next per-call pass when `remaining` output rows of `nin`
// inputs are left: 9-16 run together in the 16-output kernel while its
// coefficient tables (nin x 16 x 20 bytes) fit 64 KiB of LDS, which reads
// every input once per 16 outputs instead of once per 8 (50+20 x 1 MiB
// encode 0.91 -> 0.99 TB/s); otherwise at most 8.
inline int pass_outputs(int remaining, int nin) {
    const bool wide = remaining > 8 && nin * 16 * 20 <= (64 << 10);
    return remaining < (wide ? 16 : 8) ? remaining : (wide ? 16 : 8);
}

// What a GF(2^8) pass does with its outputs.
enum Gf8Mode : int {
    kGf8Clear = 0,       // out = M*in
    kGf8Accumulate = 1,  // out ^= M*in
    kGf8Compare = 2,     // *flag = flag_value where out != M*in; out is not written
};

struct Gf8Pass {
    const uint8_t *in[kMaxIn];
    uint8_t *out[kMaxOut];
    const uint32_t *tables;  // [k][mtot][5] coefficient tables (gf256.hpp)
    size_t len;              // bytes per shard
    int k;                   // inputs
    int mtot;                // outputs described by tables
    int o0;                  // first output of this pass in tables
    int nout;                // outputs of this pass (<= kMaxOut)
    int mode;                // Gf8Mode
    uint32_t *flag;          // compare mode: set to flag_value where out != M*in
    uint32_t flag_value;
};

// Launches the pass on `stream`.  All pointers must be device pointers.
hipError_t launch_gf8_pass(const Gf8Pass &p, hipStream_t stream);

// Sets n dwords to v (used instead of hipMemsetAsync so the work stays a
// plain kernel node if a caller captures it in a HIP graph).
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t stream);

}  // namespace rsx
