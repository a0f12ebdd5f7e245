// leo_shapes.cpp — host-side shape choices of the Leopard kernels: LDS
// slice widths and thread counts of the LDS-resident kernels (leo_fused.hip),
// pass shapes, grids and column chunks of the three-pass kernels
// (leo_multipass.hip).  Plain host code, shared by the planners.
#include <algorithm>
#include <vector>

#include "leo_kernels.hpp"

namespace rsx {

int leo_fused_slice(int bits, int rows, int regions, int step_rows, size_t size, size_t lds_budget) {
    const int unit = bits == 16 ? 32 : 16;
    auto bytes = [&](int slice) { return static_cast<size_t>(regions) * slice * (rows + 1); };
    // Enough column units that one radix-4 step gives every lane an item
    // while the shard still spans 512 slices (two per CU), at least 128 bytes
    // per row read (64 where nothing wider fits) ...
    int slice = 128;
    while ((step_rows / 4) * (slice / unit) < kLeoTileThreads && slice < 4096 && size / (2 * slice) >= 512) slice *= 2;
    // ... within the LDS budget.
    while (slice > 64 && bytes(slice) > lds_budget) slice /= 2;
    // GF(2^16) tiles of whole chunks that leave room for fewer than four
    // workgroups per CU take one 32-byte unit per workgroup instead.
    if (bits == 16 && slice >= 64 && bytes(64) > (40u << 10) && bytes(32) <= lds_budget) slice = 32;
    return bytes(slice) <= lds_budget ? slice : 0;
}

// 512 threads when LDS leaves room for one workgroup per CU and one radix-4
// step over `rows` rows has that many items; otherwise 256 and several
// workgroups per CU.
unsigned leo_tile_threads(int bits, int rows, int slice, size_t lds) {
    const bool alone = lds > (80u << 10);
    const int units = std::max(1, slice / (bits == 16 ? 32 : 16));
    return alone && (rows / 4) * units >= kLeoTileThreadsMax ? kLeoTileThreadsMax : kLeoTileThreads;
}

// Only while two workgroups fit per CU: at n = 8192 (one 512-thread
// workgroup per CU) the three-pass schedule is faster (1.79 vs 1.87 ms for
// k=4096 m=1024 64 KiB with the passes in 8 column chunks, 1.55 ms in one);
// at n = 4096 the quarter tile is 3-6% faster.
int leo_quarter_slice(int n) {
    const size_t lds = static_cast<size_t>(16) * (n + 1);
    const unsigned threads = leo_tile_threads(16, n, 16, lds);
    const size_t deriv = static_cast<size_t>(n) * 16;
    return lds <= (80u << 10) && deriv <= static_cast<size_t>(kLeoDerivMax) * 16 * threads ? 16 : 0;
}

int leo_quarter_encode_slice(int m, int groups) {
    const size_t lds = static_cast<size_t>(16) * (m + 1);
    const unsigned threads = leo_tile_threads(16, m, 16, lds);
    const size_t acc = groups > 1 ? static_cast<size_t>(m) * 16 : 0;
    return lds <= kLeoFusedEncodeLds && acc <= static_cast<size_t>(kLeoDerivMax) * 16 * threads ? 16 : 0;
}

void leo_pass_blocks(const LeoMultipassArgs &a, uint32_t blocks[3]) {
    const uint32_t sl = static_cast<uint32_t>(a.slices);
    blocks[0] = sl * static_cast<uint32_t>(a.n / a.C);
    blocks[1] = sl * static_cast<uint32_t>(a.P / a.A);
    blocks[2] = sl * static_cast<uint32_t>(a.n / a.P);
}

uint32_t leo_enc_pass_blocks(const LeoEncPassArgs &a, int pass) {
    const uint32_t sl = static_cast<uint32_t>(a.slices);
    return sl * static_cast<uint32_t>(pass == 1 ? a.m / a.C : pass == 2 ? a.P / a.A : a.m / a.P);
}

bool leo_multipass_shape(int n, int *C, int *P, int *A) {
    if (n < 2048) return false;
    int p = n >> 2;
    while (p > 1024) p >>= 2;
    const int J = n / p;
    if (J > 64) return false;  // pass 3's row blocks fit one 64-bit mask (never for n <= 65536)
    *C = 1024;
    *P = p;
    *A = 512 / J;
    return true;
}

uint64_t leo_pass3_blocks(int P, const std::vector<int> &out_rows) {
    uint64_t b = 0;
    for (int r : out_rows) b |= uint64_t{1} << (r / P);  // n / P <= 64 (leo_multipass_shape)
    return b;
}

int leo_pass3_grid(uint64_t blocks) {
    int g = 0;
    while (g < 64 && (blocks >> g) != 0) ++g;
    return g;
}

// Column chunk of one pass sequence: each work array is at most 512 MiB, so
// k=4096 m=1024 shards of 64 KiB run as one chunk (three launches per call).
// Fewer, larger launches beat Infinity-Cache-sized chunks: 8 chunks of 8 KiB
// (64 MiB arrays) took 1.69 ms for that reconstruct, one chunk 1.52 ms.
size_t leo_multipass_chunk(int n, size_t size) {
    size_t cw = (size_t{512} << 20) / static_cast<size_t>(n);
    cw -= cw % kLeoPassSlice;
    cw = std::max<size_t>(cw, kLeoPassSlice);
    return std::min(cw, size);
}

}  // namespace rsx
