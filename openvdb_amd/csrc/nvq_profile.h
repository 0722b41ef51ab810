// Opt-in timing of the library's dominant kernels with HIP events on the launch stream.
// Off by default; when off a ProfScope costs one relaxed atomic load.
#pragma once
#include <hip/hip_runtime.h>

namespace nvq {

bool profilingOn();
void profileRecord(const char* name, hipEvent_t start, hipEvent_t stop);

// Brackets one kernel launch with HIP events when profiling is on.
struct ProfScope {
    const char* name;
    hipStream_t stream;
    hipEvent_t start = nullptr;

    ProfScope(const char* n, hipStream_t s) : name(n), stream(s)
    {
        if (!profilingOn()) return;
        if (hipEventCreate(&start) != hipSuccess) {
            start = nullptr;
            return;
        }
        if (hipEventRecord(start, stream) != hipSuccess) {
            (void)hipEventDestroy(start);
            start = nullptr;
        }
    }
    ~ProfScope()
    {
        if (!start) return;
        hipEvent_t stop = nullptr;
        if (hipEventCreate(&stop) == hipSuccess && hipEventRecord(stop, stream) == hipSuccess) {
            profileRecord(name, start, stop);
            return;
        }
        if (stop) (void)hipEventDestroy(stop);
        (void)hipEventDestroy(start);
    }
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;
};

} // namespace nvq
