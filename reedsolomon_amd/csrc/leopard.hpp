// leopard.hpp — the Leopard GF(2^16) / GF(2^8) engines (leopard.go, leopard8.go).
#pragma once

#include <cstddef>
#include <cstdint>

#include <vector>

#include "codec.hpp"
#include "leo_kernels.hpp"

namespace rsx {

// k + ceilPow2(m) <= 256: the shapes the reference's GF(2^8) tables cover.
bool leopard8_shape_ok(int k, int m);
// Shapes inside the tables of the reference (see leopard.cpp).
bool leopard_shape_ok(int bits, int k, int m);
int leopard_init(rsx_codec *c);
// Field tables of `bits` (8 or 16) resident on `device` (uploaded on first use).
int leo_device_tables(int device, int bits, hipStream_t stream, const LeoDev **out);
int leopard_encode(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards);
int leopard_verify(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards, int *ok);
int leopard_reconstruct(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                        bool recover_all, uint8_t *filled);

// One Leopard item of a batched plan.  Pointer fields of `task` are left
// null; the offsets say where their contents were appended: `ptr_off` into
// the plan's pointer pool (encode: groups*m data rows then p parity rows;
// reconstruct: n work rows then nout outputs), `errl_off` / `rows_off` /
// `logs_off` (bytes) into the plan's auxiliary buffer.
struct LeoItemPlan {
    int bits = 0;
    // batch launch: 0 GF(2^8), 1 GF(2^16), 2 GF(2^16) quarter-chunk tiles,
    // 3 GF(2^16) three-pass reconstruct (n >= 8192 work rows)
    int kind = 0;
    // 4 GF(2^16) three-pass encode (m above the tiles)
    LeoMultipassArgs mp{};    // kind 3: pointers are pool / aux offsets until the plan patches them
    LeoEncPassArgs emp{};     // kind 4: per-group fields and pointers are set by the plan
    int k = 0, groups = 0;    // kind 4: data shards, m-row groups
    size_t outidx_off = 0;    // kind 3: aux offset of the n-entry output index
    size_t work_bytes = 0;    // kinds 3 and 4: bytes of its two work arrays
    const LeoDev *dev = nullptr;
    LeoBatchTask task{};
    size_t lds = 0;
    size_t regs = 0;  // bytes a workgroup holds in registers (derivative / group accumulator)
    uint32_t slices = 0;
    size_t ptr_off = 0, errl_off = 0, rows_off = 0, logs_off = 0, need_off = 0;
    size_t parity_pool = 0;  // encode: pool index of the first parity pointer
    int verify_item = -1;    // Verify: the plan item whose flag a parity mismatch sets
    bool empty = false;      // nothing to do (no erasure to fill)
};

// Validates one batched Leopard item as the per-stripe method would (op:
// RSX_OP_ENCODE / _VERIFY (planned as an encode) / _RECONSTRUCT /
// _RECONSTRUCT_DATA) and plans it for the LDS-resident batch kernel.
// Returns the item's status; RSX_ERR_NOT_SUPPORTED where the batched form
// does not apply (shards not 16-byte aligned).  GF(2^16) reconstructs with
// 8192 or more work rows and encodes above the tiles are planned as three-pass
// items (kinds 3 and 4).  `ranges` caches the device allocations the plan
// has already classified.
int leopard_plan_item(rsx_codec *c, int op, uint8_t *const *shards, const size_t *lens, int nshards,
                      std::vector<const uint8_t *> &pool, std::vector<uint8_t> &aux, LeoItemPlan &out,
                      uint8_t *filled, DeviceRanges &ranges);

}  // namespace rsx
