// rocprofiler-sdk-roctx/roctx.h stand-in for the host sanitizer build: the
// trace ranges of trace.hpp as no-ops (no profiler in that build).
#pragma once

inline int roctxRangePushA(const char *) { return 0; }
inline int roctxRangePop() { return 0; }
inline int roctxNameOsThread(const char *) { return 0; }
