// hip/hip_runtime.h stand-in for the host sanitizer build of the library's
// host code (batch planning, request loop, codec set-up, streaming, Leopard
// planners).  It declares the subset of the HIP runtime those files call;
// hip_host.cpp implements it on the host: "device" memory is host memory
// from a tracked allocator, copies run at once, streams and events are
// already complete, graphs record nothing.  Kernel launchers are stubbed in
// kernels_host.cpp.  Never used for the GPU build.
#pragma once

#include <cstddef>
#include <cstdint>

typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorNotSupported = 801,
} hipError_t;

typedef struct ihipStream_t *hipStream_t;
typedef struct ihipEvent_t *hipEvent_t;
typedef struct ihipGraph *hipGraph_t;
typedef struct hipGraphExec *hipGraphExec_t;
typedef struct hipGraphNode *hipGraphNode_t;
typedef void *hipDeviceptr_t;

typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4,
} hipMemcpyKind;

typedef enum hipMemoryType {
    hipMemoryTypeUnregistered = 0,
    hipMemoryTypeHost = 1,
    hipMemoryTypeDevice = 2,
    hipMemoryTypeManaged = 3,
    hipMemoryTypeArray = 10,
    hipMemoryTypeUnified = 11,
} hipMemoryType;

typedef struct hipPointerAttribute_t {
    hipMemoryType type;
    int device;
    void *devicePointer;
    void *hostPointer;
    int isManaged;
    unsigned allocationFlags;
} hipPointerAttribute_t;

typedef enum hipDeviceAttribute_t {
    hipDeviceAttributeMultiprocessorCount = 63,
} hipDeviceAttribute_t;

typedef enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamCaptureModeRelaxed = 2,
} hipStreamCaptureMode;

#define hipStreamNonBlocking 0x01
#define hipEventDisableTiming 0x2
#define hipHostMallocDefault 0x0

hipError_t hipGetDeviceCount(int *count);
hipError_t hipSetDevice(int device);
hipError_t hipDeviceGetAttribute(int *value, hipDeviceAttribute_t attr, int device);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);

hipError_t hipMalloc(void **ptr, size_t size);
hipError_t hipFree(void *ptr);
hipError_t hipHostMalloc(void **ptr, size_t size, unsigned flags = 0);
hipError_t hipHostFree(void *ptr);
hipError_t hipHostGetDevicePointer(void **dev, void *host, unsigned flags);
hipError_t hipPointerGetAttributes(hipPointerAttribute_t *attr, const void *ptr);
hipError_t hipMemGetAddressRange(hipDeviceptr_t *base, size_t *size, hipDeviceptr_t ptr);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind kind, hipStream_t s = nullptr);

hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t start, hipEvent_t stop);

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *exec, hipGraph_t g, hipGraphNode_t *err_node, char *log,
                               size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t exec);
