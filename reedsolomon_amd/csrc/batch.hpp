// batch.hpp — device work list of the batched GF(2^8) path.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rsx.h"

#include <cstddef>
#include <cstdint>

namespace rsx {

constexpr int kBatchMaxIn = 256;  // inputs per task (classic engine: k <= 256)

// One matrix pass of one stripe: nout (<= 8) outputs from k inputs.
struct BatchTask {
    const uint32_t *tables;  // [k][mtot][5] coefficient tables
    uint64_t len;            // bytes per shard
    uint32_t tile0;          // first global tile index of this task
    uint32_t in_off;         // k input pointers at ptrs[in_off]
    uint32_t out_off;        // nout output pointers at ptrs[out_off]
    uint16_t k, nout, mtot, o0;
    uint8_t mode;            // Gf8Mode: clear, accumulate, or compare (Verify)
    uint8_t aligned;         // every pointer 16-byte aligned
    uint8_t pad[2];
    uint32_t item;           // compare mode: flags[item] = 1 where out != M*in
};

struct BatchLaunch {
    const BatchTask *tasks;
    const uint8_t *const *ptrs;
    uint32_t *flags;            // per item: Verify mismatch
    const uint32_t *tile_task;  // task of every tile (null: binary search over tile0)
    int ntasks;
    int maxk;       // largest k among the tasks (sizes the LDS tables)
    uint32_t ntiles;
    uint32_t tile;  // bytes per tile (multiple of 16)
};

// Leopard Verify items: recomputed parity `a` vs the item's parity `b`.
struct CompareEntry {
    const uint8_t *a;
    const uint8_t *b;
    uint64_t len;
    uint32_t item;
    uint32_t pad;
};

// flags[e.item] = 1 for every entry whose bytes differ.
hipError_t launch_batch_compare(const CompareEntry *entries, int n, uint32_t *flags, hipStream_t s);

// Launches the persistent pass for tasks of exactly m_class outputs.
hipError_t launch_batch_pass(const BatchLaunch &L, int m_class, unsigned grid, hipStream_t s);

// Request-loop settings of a batch (no C-ABI counterpart).  batch_reserve
// sizes the plan's staging and device buffers to `bytes` each and its Verify
// flags to max_items, so later plans up to that size never regrow them (a
// regrowth frees memory, which waits for the whole device).
// batch_set_class_streams(false) runs every class in order on the batch's
// stream: no fork / join events per launch, for callers that already run
// several batches side by side.
int batch_reserve(rsx_batch *b, size_t bytes, int max_items);
void batch_set_class_streams(rsx_batch *b, bool on);

}  // namespace rsx
