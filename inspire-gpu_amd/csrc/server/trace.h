// Stage ranges for profilers (rocprofv3 --marker-trace): with
// INSPIRE_TRACE=1 in the environment, the server marks its host-side
// phases (preprocess steps, setup autotuning, each answer and each group
// chunk it enqueues, host-staged copies) as roctx ranges and names its
// streams, so a trace shows which launches belong to which stage. Off by
// default: a disabled range costs one branch. No counterpart in the
// reference beyond its env-gated phase prints (src/gpu_protocol.cu:1130-1179).
#pragma once
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdio>
#include <cstdlib>

namespace inspire {

inline bool trace_on() {
    static const bool on = [] {
        const char* t = std::getenv("INSPIRE_TRACE");
        return t && t[0] == '1';
    }();
    return on;
}

// A roctx range over a scope (printf-style name) when tracing is on.
class TraceRange {
public:
    template <class... A>
    explicit TraceRange(const char* fmt, A... a) : on_(trace_on()) {
        if (!on_) return;
        if constexpr (sizeof...(A) == 0) {
            roctxRangePushA(fmt);
        } else {
            char buf[160];
            std::snprintf(buf, sizeof buf, fmt, a...);
            roctxRangePushA(buf);
        }
    }
    ~TraceRange() {
        if (on_) roctxRangePop();
    }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;

private:
    bool on_;
};

inline void trace_name_stream(hipStream_t s, const char* name) {
    if (trace_on() && s) roctxNameHipStream(name, reinterpret_cast<const struct ihipStream_t*>(s));
}

}  // namespace inspire
