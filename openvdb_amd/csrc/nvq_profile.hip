// Registry of the opt-in kernel timings (nvq_profile_* in nvq.h).
#include "nvq_profile.h"
#include "../../include/nvq.h"

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace nvq {
namespace {

struct Pending {
    const char* name;
    hipEvent_t start, stop;
};
struct Total {
    double ms = 0.0;
    uint64_t launches = 0;
};

std::atomic<int> gOn{0};
std::mutex gMu;
std::vector<Pending> gPending;
std::map<std::string, Total> gTotals;

// waits for every recorded launch and folds it into the totals (caller holds gMu)
void drainLocked()
{
    for (const Pending& p : gPending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.stop) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            Total& t = gTotals[p.name];
            t.ms += double(ms);
            t.launches += 1;
        }
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    gPending.clear();
}

} // namespace

bool profilingOn() { return gOn.load(std::memory_order_relaxed) != 0; }

void profileRecord(const char* name, hipEvent_t start, hipEvent_t stop)
{
    std::lock_guard<std::mutex> lock(gMu);
    gPending.push_back({name, start, stop});
    if (gPending.size() >= 4096) drainLocked(); // bounds the number of live events while profiling
}

} // namespace nvq

extern "C" {

int nvq_profile_enable(int enable)
{
    nvq::gOn.store(enable ? 1 : 0, std::memory_order_relaxed);
    return NVQ_OK;
}

int nvq_profile_read(const char* kernel, double* total_ms, uint64_t* launches)
{
    if (!kernel || !total_ms || !launches) return NVQ_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(nvq::gMu);
    nvq::drainLocked();
    const auto it = nvq::gTotals.find(kernel);
    *total_ms = it == nvq::gTotals.end() ? 0.0 : it->second.ms;
    *launches = it == nvq::gTotals.end() ? 0 : it->second.launches;
    return NVQ_OK;
}

int nvq_profile_reset(void)
{
    std::lock_guard<std::mutex> lock(nvq::gMu);
    nvq::drainLocked();
    nvq::gTotals.clear();
    return NVQ_OK;
}

} // extern "C"
