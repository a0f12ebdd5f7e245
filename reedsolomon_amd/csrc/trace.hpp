// trace.hpp — named host ranges for timeline tools: each C-ABI call, batch
// plan and launch, request-loop batch and stream block pushes a roctx range
// (rocprofv3 --marker-trace shows them beside the kernels).  With no tool
// attached a range costs two calls that return at once.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace rsx {

class TraceRange {
  public:
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

}  // namespace rsx

#define RSX_TRACE(name) ::rsx::TraceRange rsx_trace_range_(name)
