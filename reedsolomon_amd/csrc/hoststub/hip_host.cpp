// hip_host.cpp — host implementation of the HIP runtime subset declared in
// hoststub/hip/hip_runtime.h, for the sanitizer build of the library's host
// code.  Work is synchronous: a copy is done when its call returns, so every
// stream and event is always complete.  Allocations are tracked so pointer
// classification (device or host memory) behaves as on a GPU, and so ASan
// sees every buffer with its exact size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

struct ihipStream_t {
    bool capturing = false;
};
struct ihipEvent_t {
    unsigned flags = 0;
};
struct ihipGraph {};
struct hipGraphExec {};

namespace {

enum class Kind { Device, Host };

struct Block {
    size_t size;
    Kind kind;
};

// Never destroyed: the library frees device memory from static destructors
// at exit, which may run after this file's statics would have been.
std::mutex &mu = *new std::mutex;
std::map<uintptr_t, Block> &blocks = *new std::map<uintptr_t, Block>;  // base -> block
thread_local hipError_t last = hipSuccess;

hipError_t fail(hipError_t e) {
    last = e;
    return e;
}

const std::pair<const uintptr_t, Block> *find_block(const void *p) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    auto it = blocks.upper_bound(a);
    if (it == blocks.begin()) return nullptr;
    --it;
    return a < it->first + it->second.size ? &*it : nullptr;
}

hipError_t alloc(void **ptr, size_t size, Kind kind) {
    if (!ptr) return fail(hipErrorInvalidValue);
    // 256-byte alignment, as the device allocator gives
    void *p = std::aligned_alloc(256, (std::max<size_t>(size, 1) + 255) / 256 * 256);
    if (!p) return fail(hipErrorOutOfMemory);
    std::lock_guard<std::mutex> lock(mu);
    blocks[reinterpret_cast<uintptr_t>(p)] = Block{std::max<size_t>(size, 1), kind};
    *ptr = p;
    return hipSuccess;
}

hipError_t release(void *p, Kind kind) {
    if (!p) return hipSuccess;
    {
        std::lock_guard<std::mutex> lock(mu);
        const auto it = blocks.find(reinterpret_cast<uintptr_t>(p));
        if (it == blocks.end() || it->second.kind != kind) return fail(hipErrorInvalidValue);
        blocks.erase(it);
    }
    std::free(p);
    return hipSuccess;
}

}  // namespace

hipError_t hipGetDeviceCount(int *count) {
    *count = 1;
    return hipSuccess;
}
hipError_t hipSetDevice(int device) { return device == 0 ? hipSuccess : fail(hipErrorInvalidValue); }
hipError_t hipDeviceGetAttribute(int *value, hipDeviceAttribute_t attr, int) {
    if (attr != hipDeviceAttributeMultiprocessorCount) return fail(hipErrorInvalidValue);
    *value = 256;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() {
    const hipError_t e = last;
    last = hipSuccess;
    return e;
}
const char *hipGetErrorString(hipError_t e) {
    switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorInvalidValue: return "hipErrorInvalidValue";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorNotReady: return "hipErrorNotReady";
    default: return "hipError (host stand-in)";
    }
}

hipError_t hipMalloc(void **ptr, size_t size) { return alloc(ptr, size, Kind::Device); }
hipError_t hipFree(void *ptr) { return release(ptr, Kind::Device); }
hipError_t hipHostMalloc(void **ptr, size_t size, unsigned) { return alloc(ptr, size, Kind::Host); }
hipError_t hipHostFree(void *ptr) { return release(ptr, Kind::Host); }
hipError_t hipHostGetDevicePointer(void **dev, void *host, unsigned) {
    *dev = host;
    return hipSuccess;
}

hipError_t hipPointerGetAttributes(hipPointerAttribute_t *attr, const void *ptr) {
    std::memset(attr, 0, sizeof *attr);
    std::lock_guard<std::mutex> lock(mu);
    const auto *b = find_block(ptr);
    if (!b) return fail(hipErrorInvalidValue);  // unregistered host memory
    attr->type = b->second.kind == Kind::Device ? hipMemoryTypeDevice : hipMemoryTypeHost;
    attr->devicePointer = const_cast<void *>(ptr);
    attr->hostPointer = b->second.kind == Kind::Host ? const_cast<void *>(ptr) : nullptr;
    return hipSuccess;
}

hipError_t hipMemGetAddressRange(hipDeviceptr_t *base, size_t *size, hipDeviceptr_t ptr) {
    std::lock_guard<std::mutex> lock(mu);
    const auto *b = find_block(ptr);
    if (!b || b->second.kind != Kind::Device) return fail(hipErrorInvalidValue);
    *base = reinterpret_cast<void *>(b->first);
    *size = b->second.size;
    return hipSuccess;
}

hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t) {
    if (n && (!dst || !src)) return fail(hipErrorInvalidValue);
    if (n) std::memcpy(dst, src, n);
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
    *s = new ihipStream_t;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t e, unsigned) {
    return e ? hipSuccess : fail(hipErrorInvalidValue);
}
hipError_t hipEventCreate(hipEvent_t *e) { return hipEventCreateWithFlags(e, 0); }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags) {
    *e = new ihipEvent_t;
    (*e)->flags = flags;
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { return e ? hipSuccess : fail(hipErrorInvalidValue); }
hipError_t hipEventQuery(hipEvent_t e) { return e ? hipSuccess : fail(hipErrorInvalidValue); }
hipError_t hipEventSynchronize(hipEvent_t e) { return e ? hipSuccess : fail(hipErrorInvalidValue); }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
    if (!a || !b || (a->flags & hipEventDisableTiming) || (b->flags & hipEventDisableTiming))
        return fail(hipErrorInvalidValue);
    *ms = 0.0f;
    return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    if (!s || s->capturing) return fail(hipErrorInvalidValue);
    s->capturing = true;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
    if (!s || !s->capturing) return fail(hipErrorInvalidValue);
    s->capturing = false;
    *g = new ihipGraph;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *exec, hipGraph_t, hipGraphNode_t *, char *, size_t) {
    *exec = new hipGraphExec;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t) {
    return exec ? hipSuccess : fail(hipErrorInvalidValue);
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
    delete exec;
    return hipSuccess;
}
