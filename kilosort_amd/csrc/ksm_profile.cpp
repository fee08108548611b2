// Optional per-kernel timing with HIP events.
//
// When a kernel's group bit is enabled, the entry points launch it with an
// event pair that the runtime stamps with the kernel's start and end;
// ksm_profile_query synchronises the stop events and returns the summed
// elapsed time of one kernel name.  Disabled by default, so the normal path
// records nothing.
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ksm_common.h"
#include "ksm_profile.h"

namespace {

struct Record {
    std::string name;
    hipEvent_t start, stop;
};

std::mutex g_mu;
unsigned g_mask = 0;
std::vector<Record> g_records;
std::vector<hipEvent_t> g_free;

hipEvent_t take_event() {
    if (!g_free.empty()) {
        hipEvent_t e = g_free.back();
        g_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

}  // namespace

namespace ksm {

ProfileScope::ProfileScope(const char *name, unsigned group, hipStream_t s)
    : name_(name), stream_(s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!(g_mask & group)) return;
    start_ = take_event();
    stop_ = take_event();
}

ProfileScope::~ProfileScope() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (start_ && stop_ && launched_) {
        g_records.push_back(Record{name_, start_, stop_});
        return;
    }
    if (start_) g_free.push_back(start_);
    if (stop_) g_free.push_back(stop_);
}

}  // namespace ksm

extern "C" {

int ksm_profile_enable(int mask) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_mask = (unsigned)mask;
    return KSM_OK;
}

int ksm_profile_reset(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &r : g_records) {
        hipEventSynchronize(r.stop);
        g_free.push_back(r.start);
        g_free.push_back(r.stop);
    }
    g_records.clear();
    return KSM_OK;
}

int ksm_profile_query(const char *name, double *total_ms, int *launches) {
    if (!name || !total_ms || !launches)
        return ksm::fail(KSM_ERR_INVALID, "ksm_profile_query: null argument");
    std::lock_guard<std::mutex> lk(g_mu);
    double sum = 0.0;
    int n = 0;
    for (auto &r : g_records) {
        if (r.name != name) continue;
        hipError_t e = hipEventSynchronize(r.stop);
        if (e != hipSuccess) return ksm::hip_fail(e, "ksm_profile_query");
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, r.start, r.stop);
        if (e != hipSuccess) return ksm::hip_fail(e, "ksm_profile_query");
        sum += ms;
        ++n;
    }
    *total_ms = sum;
    *launches = n;
    return KSM_OK;
}

int ksm_profile_launch(int i, char *name, int name_cap, double *start_ms, double *dur_ms) {
    if (!name || name_cap <= 0 || !start_ms || !dur_ms)
        return ksm::fail(KSM_ERR_INVALID, "ksm_profile_launch: null argument");
    std::lock_guard<std::mutex> lk(g_mu);
    if (i < 0 || (size_t)i >= g_records.size())
        return ksm::fail(KSM_ERR_INVALID, "ksm_profile_launch: launch %d of %zu", i,
                         g_records.size());
    const Record &r = g_records[(size_t)i];
    hipError_t e = hipEventSynchronize(r.stop);
    if (e != hipSuccess) return ksm::hip_fail(e, "ksm_profile_launch");
    float t0 = 0.f, d = 0.f;
    e = hipEventElapsedTime(&t0, g_records[0].start, r.start);
    if (e == hipSuccess) e = hipEventElapsedTime(&d, r.start, r.stop);
    if (e != hipSuccess) return ksm::hip_fail(e, "ksm_profile_launch");
    std::strncpy(name, r.name.c_str(), (size_t)name_cap - 1);
    name[name_cap - 1] = '\0';
    *start_ms = t0;
    *dur_ms = d;
    return KSM_OK;
}

}  // extern "C"
