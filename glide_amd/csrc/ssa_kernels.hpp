// Host-visible argument packs and launchers of the SSA kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "ssa_device.hpp"
#include "timing.hpp"

namespace glide {

// Read-only view of one level for the SSA kernels.
template <typename T>
struct SsaArgs {
    int ny, nx;
    T dxi, dt;
    Phys<T> P;
    const T *u, *v, *H, *phi, *mask, *bed, *B, *beta, *gamma;
    const T *f_u, *f_v, *f_H;
};

// Gather everything the patch of cell (i,j) reads besides its unknowns;
// eta_at(di, dj) gives the frozen viscosity of cell (i + di, j + dj).
template <typename T, class Ld = LdPlain, class EtaAt>
__device__ __forceinline__ void load_patch_with(Patch<T>& p, const SsaArgs<T>& a, EtaAt&& eta_at, int i, int j,
                                                const T* __restrict__ fu, const T* __restrict__ fv,
                                                const T* __restrict__ fH, const T* __restrict__ phi) {
    const int ny = a.ny, nx = a.nx;
    p.e_c = eta_at(0, 0); p.e_l = eta_at(0, -1); p.e_r = eta_at(0, 1);
    p.e_t = eta_at(-1, 0); p.e_b = eta_at(1, 0);
    p.e_tl = eta_at(-1, -1); p.e_tr = eta_at(-1, 1);
    p.e_bl = eta_at(1, -1); p.e_br = eta_at(1, 1);
    const T* __restrict__ H = a.H;
    p.H_l = Ld::get(H + offC(i, j - 1, ny, nx)); p.H_r = Ld::get(H + offC(i, j + 1, ny, nx));
    p.H_t = Ld::get(H + offC(i - 1, j, ny, nx)); p.H_b = Ld::get(H + offC(i + 1, j, ny, nx));
    p.H_tl = Ld::get(H + offC(i - 1, j - 1, ny, nx)); p.H_tr = Ld::get(H + offC(i - 1, j + 1, ny, nx));
    p.H_bl = Ld::get(H + offC(i + 1, j - 1, ny, nx)); p.H_br = Ld::get(H + offC(i + 1, j + 1, ny, nx));
    p.p_c = phi[offC(i, j, ny, nx)]; p.p_l = phi[offC(i, j - 1, ny, nx)]; p.p_r = phi[offC(i, j + 1, ny, nx)];
    p.p_t = phi[offC(i - 1, j, ny, nx)]; p.p_b = phi[offC(i + 1, j, ny, nx)];
    p.b_c = a.beta[offC(i, j, ny, nx)]; p.b_l = a.beta[offC(i, j - 1, ny, nx)]; p.b_r = a.beta[offC(i, j + 1, ny, nx)];
    p.b_t = a.beta[offC(i - 1, j, ny, nx)]; p.b_b = a.beta[offC(i + 1, j, ny, nx)];
    p.z_c = a.bed[offC(i, j, ny, nx)]; p.z_l = a.bed[offC(i, j - 1, ny, nx)]; p.z_r = a.bed[offC(i, j + 1, ny, nx)];
    p.z_t = a.bed[offC(i - 1, j, ny, nx)]; p.z_b = a.bed[offC(i + 1, j, ny, nx)];
    const T* __restrict__ u = a.u;
    const T* __restrict__ v = a.v;
    p.u_ll = Ld::get(u + offU(i, j - 1, ny, nx)); p.u_rr = Ld::get(u + offU(i, j + 2, ny, nx));
    p.u_tl = Ld::get(u + offU(i - 1, j, ny, nx)); p.u_tr = Ld::get(u + offU(i - 1, j + 1, ny, nx));
    p.u_bl = Ld::get(u + offU(i + 1, j, ny, nx)); p.u_br = Ld::get(u + offU(i + 1, j + 1, ny, nx));
    p.v_lt = Ld::get(v + offV(i, j - 1, ny, nx)); p.v_lb = Ld::get(v + offV(i + 1, j - 1, ny, nx));
    p.v_rt = Ld::get(v + offV(i, j + 1, ny, nx)); p.v_rb = Ld::get(v + offV(i + 1, j + 1, ny, nx));
    p.v_tt = Ld::get(v + offV(i - 1, j, ny, nx)); p.v_bb = Ld::get(v + offV(i + 2, j, ny, nx));
    if (fu) {
        p.f0 = fu[offU(i, j, ny, nx)]; p.f1 = fu[offU(i, j + 1, ny, nx)];
        p.f2 = fv[offV(i, j, ny, nx)]; p.f3 = fv[offV(i + 1, j, ny, nx)];
        p.f4 = fH[offC(i, j, ny, nx)];
    } else {
        p.f0 = p.f1 = p.f2 = p.f3 = p.f4 = T(0);
    }
    p.gam = a.gamma ? a.gamma[offC(i, j, ny, nx)] : T(0);
    p.in_l = j > 0; p.in_r = j + 1 < nx; p.in_t = i > 0; p.in_b = i + 1 < ny;
    p.cl = j > 0; p.cr = j + 1 < nx; p.ct = i > 0; p.cb = i + 1 < ny;
    p.vtl = i > 0 && j > 0; p.vtr = i > 0 && j + 1 < nx; p.vbl = i + 1 < ny && j > 0; p.vbr = i + 1 < ny && j + 1 < nx;
}


template <typename T>
struct VankaArgs {
    int newton_steps;
    T ssa_damping, mc_damping, calving_rate;
    int* iters;      // optional per-cell Newton iteration count
    int local_fp32;  // mixed form: patch Jacobian and its LU in FP32, residual and update in T
    // multicolour passes only (the additive pass fixes 0.5 and 1e-6): local
    // Newton relaxation and the stopping threshold on sum r^2
    T relax, tol;
};

template <typename T>
hipError_t launch_grounded(T* phi, const T* H, const T* bed, T c, T relax, long n, hipStream_t st);
template <typename T>
hipError_t launch_residual(const SsaArgs<T>& a, T* ou, T* ov, T* oH, int use_forcing, int use_mask, T calving_rate,
                           hipStream_t st);
template <typename T>
hipError_t launch_jvp(const SsaArgs<T>& a, const T* du, const T* dv, const T* dH, T* ou, T* ov, T* oH, int use_mask,
                      T calving_rate, hipStream_t st);
// One forward smoothing pass; phi_next != nullptr also refreshes the grounded
// fraction from the updated thickness for the next pass (relaxation relax_phi).
template <typename T>
hipError_t launch_vanka_pass(const SsaArgs<T>& a, T* u, T* v, T* H, T* work, T* mask, const VankaArgs<T>& vk, T omega,
                             T* phi_next, T relax_phi, hipStream_t st);
// One forward smoothing pass in one launch (patch solves of each tile row and
// of the cells above / left of it, then the update of the owned state): reads
// the state of a, writes the updated u, v, H to u_dst, v_dst, H_dst (buffers
// distinct from a's) and the active-set flags to mask.  vk.local_fp32 selects
// the mixed form (patch Jacobian and LU in FP32), as for launch_vanka_pass;
// for levels whose patch solves are latency-bound.
template <typename T>
hipError_t launch_vanka_fused(const SsaArgs<T>& a, T* u_dst, T* v_dst, T* H_dst, T* mask, const VankaArgs<T>& vk,
                              T omega, hipStream_t st);
#ifndef GLIDE_FUSED_VANKA_MAX_SIDE
#define GLIDE_FUSED_VANKA_MAX_SIDE 128
#endif
constexpr long kFusedVankaMaxCells = (long)GLIDE_FUSED_VANKA_MAX_SIDE * GLIDE_FUSED_VANKA_MAX_SIDE;
// Forward smoothing passes of levels at most 64 cells wide and
// kPersistentRows * kPersistentMaxBlocks rows high as one persistent launch
// (kPersistentRows rows per workgroup, a barrier between passes); sync holds
// three 64-bit words the launch zeroes first (arrival counter, timeout flag,
// XCD mask); force_fenced != 0 runs every barrier with the cache write-back
// and invalidation the workgroups need when they span XCDs.
#ifndef GLIDE_PERSISTENT_ROWS
#define GLIDE_PERSISTENT_ROWS 3
#endif
constexpr int kPersistentRows = GLIDE_PERSISTENT_ROWS;
#ifndef GLIDE_PERSISTENT_MAX_BLOCKS
#define GLIDE_PERSISTENT_MAX_BLOCKS 32
#endif
constexpr int kPersistentMaxBlocks = GLIDE_PERSISTENT_MAX_BLOCKS;
bool vanka_persistent_fits(int ny, int nx);
// Control words of a persistent launch: sync holds three 64-bit words the
// launch zeroes first (arrival counter, timeout flag, XCD mask); a timed-out
// barrier wait latches bit 0 of *latch (if non-null), which the host reads
// back at its next synchronisation; force_fenced != 0 runs every barrier
// with the cache maintenance that workgroups on several XCDs need;
// spin_limit bounds each barrier wait (polls; negative gives up at once).
struct PersistentCtl {
    unsigned long long* sync;
    unsigned int* latch;
    int force_fenced;
    int spin_limit;
};
constexpr int kGridSpinLimit = 1 << 20;
template <typename T>
hipError_t launch_vanka_persistent(const SsaArgs<T>& a, T* alt_u, T* alt_v, T* alt_H, T* mask,
                                   const VankaArgs<T>& vk, T omega, int n_iter, const PersistentCtl& ctl,
                                   hipStream_t st);
template <typename T>
hipError_t launch_vanka_dump(const SsaArgs<T>& a, T* J, T* r, T calving_rate, hipStream_t st);
template <typename T>
hipError_t launch_vanka_adjoint_pass(const SsaArgs<T>& a, const T* ru, const T* rv, const T* rH, T* lu, T* lv, T* lH,
                                     T* work, const VankaArgs<T>& vk, T omega, hipStream_t st);
// eta_cache: the cell viscosity cached by launch_adjoint_factor for the
// current state, or nullptr to compute it.
// One cached adjoint pass (transpose operator, patch solves and update) in one
// launch, reading the multipliers (lu, lv, lH) and writing the updated ones to
// separate buffers (dst); linear drag only.  The same results as launch_vjp +
// launch_vanka_adjoint_pass_cached; used on levels up to kAdjointPassFusedMaxCells.
template <typename T>
hipError_t launch_adjoint_pass_fused(const SsaArgs<T>& a, const T* cache, const T* lu, const T* lv, const T* lH,
                                     T* lu_dst, T* lv_dst, T* lH_dst, const T* fu, const T* fv, const T* fH,
                                     T calving_rate, T omega, hipStream_t st);
// One colour phase of the multicolour adjoint pass (no counterpart in the
// reference; linear drag): the fused pass's transpose operator and cached
// patch solves, but only cells with (i + j) % 2 == colour solve, and each
// multiplier they own moves by omega times the full solution (dst = l - omega d).
template <typename T>
hipError_t launch_adjoint_pass_mc(const SsaArgs<T>& a, const T* cache, const T* lu, const T* lv, const T* lH,
                                  T* lu_dst, T* lv_dst, T* lH_dst, const T* fu, const T* fv, const T* fH,
                                  T calving_rate, T omega, int colour, hipStream_t st);
#ifndef GLIDE_ADJOINT_PASS_FUSED_MAX_SIDE
#define GLIDE_ADJOINT_PASS_FUSED_MAX_SIDE 512
#endif
constexpr long kAdjointPassFusedMaxCells = (long)GLIDE_ADJOINT_PASS_FUSED_MAX_SIDE * GLIDE_ADJOINT_PASS_FUSED_MAX_SIDE;
// All adjoint passes of one call on a level at most 64 cells wide and
// 4 * kAdjointPersistentMaxBlocks rows high as one persistent launch (one tile
// row per workgroup, the barriers and sync words of launch_vanka_persistent);
// linear drag only, multipliers alternating with alt_u/v/H as for the fused
// passes.  At 64 x 64 (16 workgroups) one launch per pass measured faster.
constexpr int kAdjointPersistentMaxBlocks = 8;
bool adjoint_persistent_fits(int ny, int nx);
template <typename T>
hipError_t launch_adjoint_persistent(const SsaArgs<T>& a, const T* cache, T* lu, T* lv, T* lH, T* alt_u, T* alt_v,
                                     T* alt_H, const T* fu, const T* fv, const T* fH, T calving_rate, T omega,
                                     int n_iter, const PersistentCtl& ctl, hipStream_t st);
// All n_iter cached adjoint passes of a small level in one workgroup (linear
// drag, ny * nx <= kAdjointCoarseMaxCells, its LDS within the device limit:
// adjoint_coarse_fits); the same results as n_iter launch_vjp +
// launch_vanka_adjoint_pass_cached.  One compute unit does a pass in about
// 4 us + 15 ns per cell, a launch pair in about 13 us: the workgroup wins
// below about 24 x 24 cells.
constexpr long kAdjointCoarseMaxCells = 512;
template <typename T>
bool adjoint_coarse_fits(const SsaArgs<T>& a);
template <typename T>
hipError_t launch_adjoint_coarse(const SsaArgs<T>& a, const T* cache, T* lu, T* lv, T* lH, const T* fu, const T* fv,
                                 const T* fH, T calving_rate, T omega, int n_iter, hipStream_t st);
template <typename T>
hipError_t launch_vjp(const SsaArgs<T>& a, const T* lu, const T* lv, const T* lH, const T* fu, const T* fv,
                      const T* fH, T* ou, T* ov, T* oH, T calving_rate, const T* eta_cache, hipStream_t st);
// Adjoint smoother cache of one level and state: kAdjointCacheValues * ny * nx
// values (transposed patch LU factors and the cell viscosity).
constexpr int kAdjointCacheValues = 26;
constexpr int kAdjointCacheEta = 25;  // plane of the cell viscosity
template <typename T>
hipError_t launch_adjoint_factor(const SsaArgs<T>& a, T* cache, const VankaArgs<T>& vk, hipStream_t st);
// fused: one launch (patch solves of the tile and its left/top neighbours,
// then the in-place multiplier update); otherwise solve into work + apply.
template <typename T>
hipError_t launch_vanka_adjoint_pass_cached(int ny, int nx, const T* cache, const T* ru, const T* rv, const T* rH,
                                            T* lu, T* lv, T* lH, T* work, T omega, bool fused, hipStream_t st);
// exact: also the right / bottom cell terms of facets nx-1 and ny-1, which
// the reference's guards drop (grad.cu:62).
template <typename T>
hipError_t launch_grad_beta_bed(const SsaArgs<T>& a, const T* lu, const T* lv, T* gbeta, T* gbed, int exact,
                                hipStream_t st);
// g = 0 where mask > 0.5.
template <typename T>
hipError_t launch_zero_masked(T* g, const T* mask, long n, hipStream_t st);
// g_B = lambda^T dR/dB through the viscosity of the stresses.
template <typename T>
hipError_t launch_grad_B(const SsaArgs<T>& a, const T* lu, const T* lv, T* gB, hipStream_t st);


// Multicolour (red-black) Vanka smoothing (mc_kernels.hip; no counterpart in
// the reference).  launch_cell_eta fills eta (ny, nx) with the cell viscosity
// of the state in a.  launch_vanka_mc runs one colour phase: every cell with
// (i + j) % 2 == colour solves its 5x5 patch (local Newton with vk.relax,
// vk.tol, the Vanka dampings, viscosity from eta) against the state in a and
// writes its five unknowns to u_dst / v_dst / H_dst; the other cells copy their
// thickness and the domain-edge facets no cell of the colour owns, so the
// destination holds the complete new state; the active-set flag is written in
// place.  The destination buffers must differ from a's state.
template <typename T>
hipError_t launch_cell_eta(const SsaArgs<T>& a, T* eta, hipStream_t st);
// Zero the multipliers / corrections of the rows the patch solves pin: the
// domain-edge facets and the cells of the active set (mask > 0.5).
template <typename T>
hipError_t launch_pin_rows(T* du, T* dv, T* dH, const T* mask, int ny, int nx, hipStream_t st);
template <typename T>
hipError_t launch_vanka_mc(const SsaArgs<T>& a, const T* eta, T* u_dst, T* v_dst, T* H_dst, T* mask,
                           const VankaArgs<T>& vk, int colour, hipStream_t st);

// Element-wise jobs between FP64 and FP32 buffers of the mixed-precision
// solver (mc_kernels.hip), batched into one launch (the jobs must not read
// each other's outputs).
enum MixedOp {
    kMixTo32 = 0,     // out32 = (float) a64
    kMixTo64 = 1,     // out64 = (double) a32
    kMixDefect = 2,   // out32 = a32 - (float) b64
    kMixUpdate = 3,   // out64 += (double) a32 - (double)(float) out64
    kMixUpdateH = 4,  // kMixUpdate, then out64 = val where b32 (the active-set flag) > 0.5
    kMixAdd64 = 5,    // out64 += (double) a32
    kMixNeg32 = 6     // out32 = -(float) a64
};
struct MixedJob {
    void* out;
    const void *a, *b;
    double val;
    long n;
    int op;
};
constexpr int kMixedMaxJobs = 12;
struct MixedBatch {
    MixedJob job[kMixedMaxJobs];
    int count = 0;
    void add(int op, void* out, const void* a, const void* b, long n, double val = 0.0) {
        job[count++] = MixedJob{out, a, b, val, n, op};
    }
};
hipError_t launch_mixed_batch(const MixedBatch& b, hipStream_t st);

// Inner products and updates of the flexible GMRES outer iteration (flat FP64
// vectors): out[q] = <w, V_q> for q < nq (V_q at V + q * ld) and out[nq] =
// <w, w>, through kKrylovBlocks partials per value (partial holds
// (nq + 1) * kKrylovBlocks doubles); out += sign * sum_q c[q] V_q with c on
// the device.
constexpr int kKrylovMaxDots = 18;
constexpr int kKrylovBlocks = 512;
hipError_t launch_krylov_dots(const double* w, const double* V, long ld, long n, int nq, double* partial,
                              double* out, hipStream_t st);
hipError_t launch_krylov_axpy(double* out, const double* V, long ld, long n, int nq, const double* c, double sign,
                              hipStream_t st);

// transfers and element-wise helpers (transfer_kernels.hip)
template <typename T>
hipError_t launch_restrict_vfacet(const T* f, T* c, int nyc, int nxc, hipStream_t st);
template <typename T>
hipError_t launch_restrict_hfacet(const T* f, T* c, int nyc, int nxc, hipStream_t st);
template <typename T>
hipError_t launch_restrict_cell(const T* f, T* c, int nyc, int nxc, int use_max, hipStream_t st);
template <typename T>
hipError_t launch_prolong_vfacet_bilinear(const T* c, T* f, int nyf, int nxf, hipStream_t st);
template <typename T>
hipError_t launch_prolong_hfacet_bilinear(const T* c, T* f, int nyf, int nxf, hipStream_t st);
template <typename T>
hipError_t launch_prolong_cell(const T* c, T* f, int nyf, int nxf, int bilinear, hipStream_t st);

// Batched element-wise jobs (one launch; the jobs must not read each other's
// outputs): axpby out = a x + b y (+ c z), copy, fill out = a, and the
// correction step out = x + y (z += y in place when out == x), out2 = z + out.
enum EwOp { kEwAxpby = 0, kEwCopy = 1, kEwFill = 2, kEwCorrect = 3 };
template <typename T>
struct EwJob {
    T* out;
    T* out2;
    const T *x, *y, *z;
    T a, b, c;
    long n;
    int op;
};
constexpr int kEwMaxJobs = 8;
template <typename T>
struct EwBatch {
    EwJob<T> job[kEwMaxJobs];
    int count = 0;
    void axpby(T* o, T a_, const T* x_, T b_, const T* y_, long n_) { job[count++] = {o, nullptr, x_, y_, nullptr, a_, b_, T(0), n_, kEwAxpby}; }
    void copy(T* o, const T* x_, long n_) { job[count++] = {o, nullptr, x_, nullptr, nullptr, T(0), T(0), T(0), n_, kEwCopy}; }
    void fill(T* o, T v, long n_) { job[count++] = {o, nullptr, nullptr, nullptr, nullptr, v, T(0), T(0), n_, kEwFill}; }
    // z = z + y; x = w + z
    void correct(T* z, const T* y_, T* x_, const T* w, long n_) { job[count++] = {z, x_, z, y_, w, T(1), T(1), T(1), n_, kEwCorrect}; }
};
template <typename T>
hipError_t launch_ew_batch(const EwBatch<T>& b, hipStream_t st);

// Batched restrictions onto one coarse level (nyc, nxc).
enum RestrictKind { kRestrictU = 0, kRestrictV = 1, kRestrictCellMean = 2, kRestrictCellMax = 3 };
template <typename T>
struct RestrictJob {
    const T* f;
    T* c;
    int kind;
};
constexpr int kRestrictMaxJobs = 12;
template <typename T>
struct RestrictBatch {
    RestrictJob<T> job[kRestrictMaxJobs];
    int count = 0;
    void add(const T* f, T* c, int kind) { job[count++] = {f, c, kind}; }
};
template <typename T>
hipError_t launch_restrict_batch(const RestrictBatch<T>& b, int nyc, int nxc, hipStream_t st);

// out = a*x + b*y (+ c*z when z != nullptr), element-wise over n values.
template <typename T>
hipError_t launch_axpby(T* out, T a, const T* x, T b, const T* y, T c, const T* z, long n, hipStream_t st);
template <typename T>
hipError_t launch_copy(T* out, const T* x, long n, hipStream_t st);
template <typename T>
hipError_t launch_fill(T* out, T val, long n, hipStream_t st);
// partial sums of squares of up to three arrays into `partials` (nblk per
// array), then a single-block final pass into out3 (device).
template <typename T>
hipError_t launch_sumsq3(const T* a, long na, const T* b, long nb, const T* c, long nc, double* partials,
                         double* out3, hipStream_t st);
constexpr int kSumsqBlocks = 1024;

}  // namespace glide
