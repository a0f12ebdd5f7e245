// codec.hpp — internal state of an rsx_codec (one Encoder of the reference).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rsx.h"
#include "gf256.hpp"

namespace rsx {

// Error helpers: set the thread's last-error message and return the code.
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define RSX_HIP(expr)                                  \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

// Device allocation owned by the codec.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer();
    int reserve(size_t n);  // grows (never shrinks); contents are not kept
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// One cached reconstruct plan of the classic engine: which shards are read,
// which are written, and the device-resident coefficient tables of the
// decode rows (reedsolomon.go:634-673 getDecodeMatrix + 1732-1774).
struct DecodePlan {
    std::vector<int> inputs;   // k surviving shard indices (first k present)
    std::vector<int> outputs;  // shard indices written, in row order
    DeviceBuffer tables;       // [k][outputs.size()][5]
    std::vector<uint32_t> host_tables;  // the same tables in host memory
};

// Memory class of one call's shards.
enum class MemKind { Device, Host };
int classify_all(const std::vector<const void *> &ptrs, MemKind &kind);

// Classifies many pointers with one runtime query per device allocation:
// the address range of each device allocation seen is remembered, so a
// pointer inside a known range needs no query.  Valid while the caller
// frees nothing, i.e. for the span of one call.
class DeviceRanges {
  public:
    bool is_device(const void *p);

  private:
    std::vector<std::pair<uintptr_t, uintptr_t>> ranges_;  // [base, end)
    size_t last_ = 0;
};

// Shared helpers of the classic engine (codec.cpp).
size_t shard_size(const size_t *lens, int n);
int check_shards(const size_t *lens, int n, bool nilok);
// out = rows * in (mode 0), out ^= rows * in (mode 1), or *flag = flag_value
// where out != rows * in (mode 2, out is only read); `tables` is [nin][nrows][5].
int run_matmul(rsx_codec *c, const uint32_t *tables, const std::vector<const uint8_t *> &in,
               const std::vector<uint8_t *> &out, size_t len, int mode, uint32_t *flag = nullptr,
               uint32_t flag_value = 0);
// Coefficient tables of `rows` (nrows rows of nin bytes) into dst ([nin][nrows][5]).
int upload_tables(const std::vector<const uint8_t *> &rows, int nin, DeviceBuffer &dst,
                  std::vector<uint32_t> &host, hipStream_t stream);
// Verify's mismatch flag: a pinned host word mapped into the device's address
// space.  verify_flag_begin returns its device address and the value a
// mismatch of this call stores (a per-codec epoch, so the word needs no reset
// between calls); after the stream is synchronised, verify_flag_mismatch
// tells whether that value was stored.  Caller holds c->mu.
int verify_flag_begin(rsx_codec *c, uint32_t **dev_flag, uint32_t *epoch);
bool verify_flag_mismatch(const rsx_codec *c, uint32_t epoch);
// Staging slot pointers of `len` bytes in the codec's HBM buffer (nslots slots).
int stage_slots(rsx_codec *c, size_t len, size_t nslots, std::vector<uint8_t *> &dev);

struct LeopardState;  // leopard.cpp
struct LeopardDeleter {
    void operator()(LeopardState *s) const;
};

}  // namespace rsx

struct rsx_codec {
    int k = 0, m = 0, total = 0;
    int engine = 0;  // 0 classic, RSX_LEOPARD_GF16, RSX_LEOPARD_GF8
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool inversion_cache = true;

    // classic engine
    rsx::Matrix matrix;              // (k+m) x k
    rsx::DeviceBuffer parity_tables; // [k][m][5]
    std::vector<uint32_t> parity_tables_host;

    std::mutex mu;  // guards the caches and the staging buffers below
    // Plans are shared so a batch planner that copied a pointer under `mu`
    // keeps its plan alive after a flush.
    std::unordered_map<std::string, std::shared_ptr<const rsx::DecodePlan>> decode_cache;
    uint64_t decode_epoch = 0;  // incremented whenever decode_cache is flushed
    // WithInversionCache(false): the plan of the last call, held until the
    // next call replaces it (its device tables are read after the call returns).
    std::shared_ptr<const rsx::DecodePlan> uncached_plan;
    rsx::DeviceBuffer staging;  // host-memory calls: shards staged in HBM
    uint32_t *host_flag = nullptr;     // Verify: pinned word, the epoch of the last mismatch
    uint32_t *dev_flag = nullptr;      // Verify: host_flag's device address
    uint32_t verify_epoch = 0;         // Verify: value a mismatch of the current call stores
    rsx::DeviceBuffer aux_tables;  // Update / DecodeIdx: per-call coefficient tables
    // rsx_reconstruct_noalloc: host workspace sized at New (k x 2k inversion
    // array + m decode rows of k bytes), so that call allocates nothing.
    std::vector<uint8_t> noalloc_work;
    rsx::DeviceBuffer one_table;   // the table of coefficient 1 (XOR passes)

    std::unique_ptr<rsx::LeopardState, rsx::LeopardDeleter> leo;
};
