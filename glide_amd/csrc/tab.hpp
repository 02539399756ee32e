// Domain-table launches (no counterpart in the reference, which runs one
// domain per process: glide/model.py:25-39).  One launch runs the same stage
// of a V-cycle on many independent domains (an ensemble's levels of uneven
// sizes): job e of a table is one domain's arguments, and it owns blocks
// start[e] .. start[e + 1] - 1 of the launch, its own single-domain grid
// flattened x-fastest (gx blocks per row).  A block finds its job by binary
// search over start (the same for every thread of the block) and runs the
// single-domain kernel body with its block index inside that grid, so each
// domain's arithmetic is that of its own launch.
#pragma once

#include <hip/hip_runtime.h>

#include "ssa_kernels.hpp"

namespace glide {

// The job owning block b: the largest e < n with start[e] <= b.
__device__ __forceinline__ int tab_find(const int* __restrict__ start, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T>
struct TabResidual {
    SsaArgs<T> a;
    T *ou, *ov, *oH;
    int use_forcing, use_mask;
    T cr;
    int gx;
};
template <typename T>
struct TabEta {
    SsaArgs<T> a;
    T* eta;
    int gx;
};
template <typename T>
struct TabVankaMc {
    SsaArgs<T> a;
    const T* eta;
    T *u_dst, *v_dst, *H_dst, *mask;
    VankaArgs<T> vk;
    int gx;
};
template <typename T>
struct TabAdjFactor {
    SsaArgs<T> a;
    T* cache;
    VankaArgs<T> vk;
    int gx;
};
// one colour phase of the multicolour adjoint pass; omega is the signed
// update weight the kernel applies (the launcher of a single domain passes
// -omega)
template <typename T>
struct TabAdjPass {
    SsaArgs<T> a;
    const T* cache;
    const T *lu, *lv, *lH;
    T *lu_dst, *lv_dst, *lH_dst;
    const T *fu, *fv, *fH;
    T cr, omega;
    int gx;
};
template <typename T>
struct TabVjp {
    SsaArgs<T> a;
    const T *lu, *lv, *lH, *fu, *fv, *fH;
    T *ou, *ov, *oH;
    T cr;
    const T* eta_cache;
    int gx;
};
template <typename T>
struct TabGrad {
    SsaArgs<T> a;
    const T *lu, *lv;
    T *gbeta, *gbed;
    int exact;
    int gx;
};

// Element-wise jobs of either precision: the operations of EwBatch
// (axpby out = a x + b y (+ c z), copy, fill, correct), of MixedBatch
// (precision conversions and updates), the grounded fraction
// (out = phi, x = H, y = bed, a = c, b = relax) and zeroing where a mask is set
// (out = g, x = mask).  kTabEwPerBlock values per block.
enum TabEwOp {
    kTabAxpby32 = 0, kTabCopy32, kTabFill32, kTabCorrect32,
    kTabAxpby64, kTabCopy64, kTabFill64, kTabCorrect64,
    kTabTo32, kTabTo64, kTabDefect, kTabUpdate, kTabUpdateH, kTabAdd64, kTabNeg32,
    kTabGrounded32, kTabGrounded64, kTabZeroMasked64
};
struct TabEw {
    void* out;
    void* out2;
    const void *x, *y, *z;
    double a, b, c;
    long n;
    int op;
};
constexpr int kTabEwPerBlock = 1024;

// FP32 transfers and the pinning of multipliers: the arithmetic of
// k_restrict_* / k_prolong_* / k_pin_rows, one output row per block row.
// ny, nx: the coarse level for restrictions, the fine level for prolongations
// and the level itself for pinning (dst = u-facet plane, dv, dH, mask).
enum TabXferOp {
    kTabRestrictU = 0, kTabRestrictV, kTabRestrictMean, kTabRestrictMax,
    kTabProlongVfacet, kTabProlongHfacet, kTabProlongInject, kTabProlongBilinear, kTabPinRows
};
struct TabXfer {
    const float* src;
    float* dst;
    float *dv, *dH;
    const float* mask;
    int op;
    int ny, nx;
    int gx;
};

// Blocks of one job (its single-domain grid) and the launchers: jobs and
// start in device memory (start with n entries), `blocks` the total.
int tab_blocks_residual(int ny, int nx, int* gx);
int tab_blocks_eta(int ny, int nx, int* gx);
// the red-black phase's blocks hold `rows` rows of cell pairs (1, 2, 4 or 8)
bool tab_vanka_mc_rows_ok(int rows);
int tab_blocks_vanka_mc(int ny, int nx, int rows, int* gx);
int tab_blocks_adj_factor(int ny, int nx, int* gx);
// the multicolour adjoint phase tiles `rows` rows (tab_adj_pass_rows of the precision)
int tab_adj_pass_rows(bool fp32);
int tab_blocks_adj_pass(int ny, int nx, int rows, int* gx);
int tab_blocks_vjp(int ny, int nx, int* gx);
int tab_blocks_grad(int ny, int nx, int* gx);
int tab_blocks_xfer(int op, int ny, int nx, int* gx);
inline int tab_blocks_ew(long n) { return (int)((n + kTabEwPerBlock - 1) / kTabEwPerBlock); }

template <typename T>
hipError_t launch_tab_residual(const TabResidual<T>* jobs, const int* start, int n, int blocks, hipStream_t st);
template <typename T>
hipError_t launch_tab_eta(const TabEta<T>* jobs, const int* start, int n, int blocks, hipStream_t st);
template <typename T>
hipError_t launch_tab_vanka_mc(const TabVankaMc<T>* jobs, const int* start, int n, int blocks, int colour,
                               bool linear, int rows, hipStream_t st);
template <typename T>
hipError_t launch_tab_adj_factor(const TabAdjFactor<T>* jobs, const int* start, int n, int blocks, hipStream_t st);
template <typename T>
hipError_t launch_tab_adj_pass(const TabAdjPass<T>* jobs, const int* start, int n, int blocks, int colour,
                               hipStream_t st);
template <typename T>
hipError_t launch_tab_vjp(const TabVjp<T>* jobs, const int* start, int n, int blocks, hipStream_t st);
template <typename T>
hipError_t launch_tab_grad(const TabGrad<T>* jobs, const int* start, int n, int blocks, hipStream_t st);
hipError_t launch_tab_ew(const TabEw* jobs, const int* start, int n, int blocks, hipStream_t st);
hipError_t launch_tab_xfer(const TabXfer* jobs, const int* start, int n, int blocks, hipStream_t st);

}  // namespace glide
