// The allocation-free variant of the fused enthalpy column sweep.
//
// glide_enth_column_sweep takes its reduction partials and sweep status record
// from library scratch that the first call on a device allocates.  Callers
// that cannot allocate (code under stream capture, serving loops, harnesses
// that forbid hipMalloc) use glide_enth_column_sweep_noalloc instead, whose
// scratch is the file-scope device array below: one copy per GPU, held until
// the process ends.  Its device address is resolved once per device with
// hipGetSymbolAddress (the array's name is only a host-side stand-in) and the
// kernels receive that address.  All calls on one GPU share the array: the
// last call's contents stay in it, and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/glide_enthalpy.h"
#include "enthalpy_kernels.hpp"

namespace glide {
void set_last_error(const std::string& msg);
int enth_column_sweep_with_scratch(const glide_enth_column* C, double dt, int n_iter,
                                   const glide_enth_sweep_config* cfg, glide_enth_report* report, hipStream_t st,
                                   double* partials, EnthStatus* status);
}  // namespace glide

using namespace glide;

namespace {

// Layout: kEnthResBlocks reduction partials, then the status record.
constexpr size_t kPartialsBytes = sizeof(double) * kEnthResBlocks;
constexpr size_t kStatusOffset = (kPartialsBytes + 255) / 256 * 256;
constexpr size_t kScratchBytes = kStatusOffset + (sizeof(EnthStatus) + 255) / 256 * 256;
static_assert(GLIDE_NOALLOC_SCRATCH_BYTES >= kScratchBytes, "declared scratch size is too small");

static __device__ __align__(256) unsigned char scratch[GLIDE_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_ptr[kMaxDevices] = {};  // device address of `scratch` per device

// Resolves the scratch array's address on the current device (once).
int scratch_ptr(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_last_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
        return GLIDE_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        set_last_error("device index beyond the allocation-free scratch table");
        return GLIDE_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return GLIDE_ERR_HIP;
        }
        g_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_ptr[dev];
    return GLIDE_OK;
}

}  // namespace

namespace glide {
// Set-up from glide_init: resolve the current device's scratch address.
int noalloc_setup() {
    unsigned char* p = nullptr;
    return scratch_ptr(&p);
}
}  // namespace glide

extern "C" {

int glide_enth_column_sweep_noalloc(const glide_enth_column* C, double dt, int n_iter,
                                    const glide_enth_sweep_config* cfg, glide_enth_report* report, void* stream) {
    if (n_iter < 0 || n_iter > GLIDE_ENTH_MAX_SWEEPS) {
        set_last_error("n_iter must be in [0, GLIDE_ENTH_MAX_SWEEPS]: the scratch's sweep record holds "
                       "GLIDE_ENTH_MAX_SWEEPS norms");
        return GLIDE_ERR_INVALID_ARG;
    }
    unsigned char* p = nullptr;
    int rc = scratch_ptr(&p);
    if (rc != GLIDE_OK) return rc;
    // the sweep's first residual launch initialises the whole status record
    // and writes every partial it reduces, so nothing needs clearing here
    return enth_column_sweep_with_scratch(C, dt, n_iter, cfg, report, static_cast<hipStream_t>(stream),
                                          reinterpret_cast<double*>(p),
                                          reinterpret_cast<EnthStatus*>(p + kStatusOffset));
}

}  // extern "C"
