// Launch timing records (see timing.hpp).  Host code only; the records are
// read back, and the events returned to a pool, by timing_read/timing_clear.
#include "timing.hpp"

#include <vector>

namespace glide {

namespace {

struct TimingRec {
    hipEvent_t start, stop;
    long units;
};

bool g_timing_on = false;
std::vector<TimingRec> g_timing[kTimedKernels];
std::vector<hipEvent_t> g_event_pool;

hipEvent_t pooled_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

}  // namespace

TimedLaunch::TimedLaunch(int id_, long units_, hipStream_t st_) : id(id_), units(units_), st(st_) {
    if (!g_timing_on) return;
    start = pooled_event();
    stop = pooled_event();
    if (start) hipEventRecord(start, st);
}

TimedLaunch::~TimedLaunch() {
    if (!start || !stop) return;
    hipEventRecord(stop, st);
    g_timing[id].push_back({start, stop, units});
}

void timing_enable(bool on) {
    timing_clear();
    g_timing_on = on;
}

void timing_clear() {
    for (auto& v : g_timing) {
        for (auto& r : v) {
            g_event_pool.push_back(r.start);
            g_event_pool.push_back(r.stop);
        }
        v.clear();
    }
}

hipError_t timing_read(int id, long* count, double* total_ms, double* total_units) {
    *count = 0;
    *total_ms = 0.0;
    *total_units = 0.0;
    if (id < 0 || id >= kTimedKernels) return hipErrorInvalidValue;
    for (auto& r : g_timing[id]) {
        hipError_t e = hipEventSynchronize(r.stop);
        if (e != hipSuccess) return e;
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, r.start, r.stop);
        if (e != hipSuccess) return e;
        *total_ms += ms;
        *total_units += (double)r.units;
        ++*count;
    }
    return hipSuccess;
}

}  // namespace glide
