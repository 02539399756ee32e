// Opt-in launch timing of the hot kernels: HIP event pairs recorded on the
// launch stream around each launch while enabled; each record keeps the
// launch's unit count (cells, or column nodes) so that per-launch
// algorithmic bytes can be computed by the caller.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace glide {

enum TimedKernel {
    kTimeVankaCell = 0,
    kTimeVankaApply,
    kTimeAdjointCell,
    kTimeVjp,
    kTimeResidual,
    kTimeAdjointApply,
    kTimeEnthColumn,
    kTimeEnthLayer,
    kTimeEnthResidual,
    kTimeAdjointFactor,
    kTimeVankaMc32,    // one colour phase of the red-black smoother, FP32 / FP64
    kTimeVankaMc64,
    kTimeAdjointMc32,  // one colour phase of the red-black adjoint pass, FP32 / FP64
    kTimeAdjointMc64,
    kTimedKernels
};

struct TimingPair {
    int id;
    long units;
    hipEvent_t start, stop;
};

bool timing_on();
void timing_enable(bool on);
void timing_clear();
// Totals of kernel `id` since the last enable; units_filter > 0 restricts them
// to launches of exactly that unit count (one grid level).
hipError_t timing_read(int id, long* count, double* total_ms, double* total_units, long units_filter = 0);
struct TimedLaunch {
    int id;
    long units;
    hipStream_t st;
    hipEvent_t start = nullptr, stop = nullptr;
    TimedLaunch(int id, long units, hipStream_t st);
    ~TimedLaunch();
};

}  // namespace glide
