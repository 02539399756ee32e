// Launch timing records (see timing.hpp).  Host code only.
#include "timing.hpp"

#include <map>

namespace glide {

namespace {

struct Totals {
    long count = 0;
    double ms = 0.0, units = 0.0;
};

bool g_timing_on = false;
std::vector<TimingPair> g_pending;  // launches not yet read
std::map<long, Totals> g_totals[kTimedKernels];  // per kernel, per launch unit count
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

hipError_t add_pair(const TimingPair& r) {
    hipError_t e = hipEventSynchronize(r.stop);
    if (e != hipSuccess) return e;
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, r.start, r.stop);
    if (e != hipSuccess) return e;
    Totals& t = g_totals[r.id][r.units];
    t.ms += ms;
    t.units += (double)r.units;
    ++t.count;
    return hipSuccess;
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
    g_pending.push_back({id, units, start, stop});
}

bool timing_on() { return g_timing_on; }

void timing_enable(bool on) {
    timing_clear();
    g_timing_on = on;
}

void timing_clear() {
    for (auto& r : g_pending) {
        g_event_pool.push_back(r.start);
        g_event_pool.push_back(r.stop);
    }
    g_pending.clear();
    for (auto& t : g_totals) t.clear();
}

hipError_t timing_read(int id, long* count, double* total_ms, double* total_units, long units_filter) {
    *count = 0;
    *total_ms = 0.0;
    *total_units = 0.0;
    if (id < 0 || id >= kTimedKernels) return hipErrorInvalidValue;
    for (auto& r : g_pending) {
        hipError_t e = add_pair(r);
        if (e != hipSuccess) return e;
        g_event_pool.push_back(r.start);
        g_event_pool.push_back(r.stop);
    }
    g_pending.clear();
    for (auto& kv : g_totals[id]) {
        if (units_filter > 0 && kv.first != units_filter) continue;
        *count += kv.second.count;
        *total_ms += kv.second.ms;
        *total_units += kv.second.units;
    }
    return hipSuccess;
}

}  // namespace glide
