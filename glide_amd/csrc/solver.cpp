// C ABI of the MI355X SSA path and the native multigrid drivers.
//
// The FAS constrained-descent V-cycle (glide/multigrid.py:536-645) and the FAS
// adjoint V-cycle (glide/multigrid.py:905-973) run here in C++ and enqueue
// their kernels on the caller's stream; the only host synchronisation is the
// per-V-cycle residual norm the stopping test needs (none when both
// tolerances are non-positive: a fixed number of V-cycles, norms read once at
// the end, and the whole solve replayed from a cached HIP graph).  All field
// buffers belong to the caller; the library owns one small per-device
// reduction scratch with a failure latch (allocated by glide_init, or by the
// first call on a device that needs it) and the graph cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/glide_hip.h"
#include "../../include/glide_enthalpy.h"
#include "ssa_kernels.hpp"
#include "tab.hpp"

using namespace glide;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* where) {
    g_last_error = std::string(where) + ": " + hipGetErrorString(e);
    return GLIDE_ERR_HIP;
}

#define HIP_TRY(expr, where)                       \
    do {                                           \
        hipError_t _e = (expr);                    \
        if (_e != hipSuccess) return hip_fail(_e, where); \
    } while (0)

#define GLIDE_TRY(expr)              \
    do {                             \
        int _r = (expr);             \
        if (_r != GLIDE_OK) return _r; \
    } while (0)

// Pinned host doubles per device: the mirror of the 3 * 66 norm slots, then
// kKrylovHostSlots values of the flexible GMRES and Anderson iterations (up to
// seven rows of inner products).
constexpr int kKrylovHostSlots = 8 * (kKrylovMaxDots + 1);
constexpr int kHostSlots = 3 * 66 + kKrylovHostSlots;

struct DeviceScratch {
    double* partials = nullptr;  // 3 * kSumsqBlocks
    double* norms = nullptr;     // 3 * 66 slots
    double* host = nullptr;      // pinned: mirror of `norms`, then the Krylov slots
    unsigned* latch = nullptr;   // failure latch of the persistent launches (bit 0: a barrier timed out)
    unsigned* latch_host = nullptr;  // pinned: [0] mirror of `latch`, [1] zero (the value that clears it)
};

std::mutex g_scratch_mutex;
std::vector<DeviceScratch> g_scratch;

int scratch_for_current_device(DeviceScratch** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
    DeviceScratch& s = g_scratch[dev];
    if (!s.partials) {
        HIP_TRY(hipMalloc(&s.partials, sizeof(double) * 3 * kSumsqBlocks), "hipMalloc(reduction scratch)");
        HIP_TRY(hipMalloc(&s.norms, sizeof(double) * 3 * 66), "hipMalloc(norm slots)");
        HIP_TRY(hipHostMalloc(&s.host, sizeof(double) * kHostSlots, hipHostMallocDefault), "hipHostMalloc(norms)");
        HIP_TRY(hipMalloc(&s.latch, sizeof(unsigned)), "hipMalloc(latch)");
        HIP_TRY(hipMemset(s.latch, 0, sizeof(unsigned)), "hipMemset(latch)");
        HIP_TRY(hipHostMalloc(&s.latch_host, 2 * sizeof(unsigned), hipHostMallocDefault), "hipHostMalloc(latch)");
        s.latch_host[0] = s.latch_host[1] = 0u;
    }
    *out = &s;
    return GLIDE_OK;
}

// The failure latch: latch_enqueue copies it to the host and clears it on the
// stream; after the caller's synchronisation latch_check reports
// GLIDE_ERR_TIMEOUT if a persistent barrier timed out since the last check.
int latch_enqueue(DeviceScratch* s, hipStream_t st) {
    HIP_TRY(hipMemcpyAsync(s->latch_host, s->latch, sizeof(unsigned), hipMemcpyDeviceToHost, st), "latch D2H");
    HIP_TRY(hipMemcpyAsync(s->latch, s->latch_host + 1, sizeof(unsigned), hipMemcpyHostToDevice, st), "latch clear");
    return GLIDE_OK;
}
int latch_check(DeviceScratch* s) {
    if (s->latch_host[0] & 1u) {
        s->latch_host[0] = 0u;
        return fail(GLIDE_ERR_TIMEOUT,
                    "a persistent launch's barrier wait timed out: the results of that solve are invalid (NaN)");
    }
    return GLIDE_OK;
}

inline size_t nU(const glide_level& L) { return (size_t)L.ny * (L.nx + 1); }
inline size_t nV(const glide_level& L) { return (size_t)(L.ny + 1) * L.nx; }
inline size_t nC(const glide_level& L) { return (size_t)L.ny * L.nx; }

typedef double R;
inline R* P(void* p) { return static_cast<R*>(p); }
inline const R* CP(const void* p) { return static_cast<const R*>(p); }

Phys<R> make_phys(const glide_physics& g) {
    Phys<R> p;
    p.n = g.n; p.eps_reg = g.eps_reg; p.frd = g.flotation_reg_driving;
    p.m = g.m; p.u_reg = g.u_reg; p.water_drag = g.water_drag; p.frs = g.flotation_reg_sliding;
    p.calving_rate = g.calving_rate; p.frc = g.flotation_reg_calving; p.thklim = g.thklim;
    p.glen_p = (1.0 - g.n) / (2.0 * g.n);
    p.drag_q = (g.m - 1.0) / 2.0;
    p.linear_drag = (g.m == 1.0 && g.u_reg > 0.0) ? 1 : 0;
    p.glen_cube = g.n == 3.0 ? 1 : 0;
    return p;
}

SsaArgs<R> make_args(const glide_level& L, double dt, bool adjoint_forcing) {
    SsaArgs<R> a;
    a.ny = L.ny; a.nx = L.nx; a.dxi = 1.0 / L.dx; a.dt = dt;
    a.P = make_phys(L.phys);
    a.u = CP(L.u); a.v = CP(L.v); a.H = CP(L.H); a.phi = CP(L.phi); a.mask = CP(L.mask);
    a.bed = CP(L.bed); a.B = CP(L.B); a.beta = CP(L.beta); a.gamma = CP(L.gamma);
    a.f_u = CP(adjoint_forcing ? L.af_u : L.f_u);
    a.f_v = CP(adjoint_forcing ? L.af_v : L.f_v);
    a.f_H = CP(adjoint_forcing ? L.af_H : L.f_H);
    return a;
}

int check_level(const glide_level* L) {
    if (!L) return fail(GLIDE_ERR_INVALID_ARG, "level is NULL");
    if (L->ny < 2 || L->nx < 2) return fail(GLIDE_ERR_SHAPE, "level must have ny >= 2 and nx >= 2");
    if (!(L->dx > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dx must be positive");
    return GLIDE_OK;
}

int require(const void* p, const char* name) {
    if (!p) {
        g_last_error = std::string("required buffer is NULL: ") + name;
        return GLIDE_ERR_INVALID_ARG;
    }
    return GLIDE_OK;
}

#define REQ(ptr) GLIDE_TRY(require((ptr), #ptr))

int check_ssa_inputs(const glide_level& L) {
    REQ(L.u); REQ(L.v); REQ(L.H); REQ(L.phi); REQ(L.mask); REQ(L.bed); REQ(L.B); REQ(L.beta);
    return GLIDE_OK;
}

// ---------------------------------------------------------------- primitives
int grounded(const glide_level& L, double relax, hipStream_t st) {
    HIP_TRY(launch_grounded<R>(P(L.phi), CP(L.H), CP(L.bed), L.phys.flotation_reg_driving, relax, (long)nC(L), st),
            "grounded");
    return GLIDE_OK;
}

int residual(const glide_level& L, double dt, int flags, R* ou, R* ov, R* oH, hipStream_t st) {
    SsaArgs<R> a = make_args(L, dt, false);
    const double cr = (flags & GLIDE_FREEZE_CALVING) ? 0.0 : L.phys.calving_rate;
    HIP_TRY(launch_residual<R>(a, ou, ov, oH, (flags & GLIDE_USE_FORCING) ? 1 : 0, (flags & GLIDE_USE_MASK) ? 1 : 0,
                               cr, st),
            "residual");
    return GLIDE_OK;
}

int sumsq3_async(const R* a, size_t na, const R* b, size_t nb, const R* c, size_t nc, double* dev_out3,
                 hipStream_t st) {
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));
    HIP_TRY(launch_sumsq3<R>(a, (long)na, b, (long)nb, c, (long)nc, s->partials, dev_out3, st), "sumsq3");
    return GLIDE_OK;
}

// Adjoint passes from the cache run as one fused launch (default) or as
// solve + apply (glide_adjoint_fused_enable).
bool g_fused_adjoint = true;
// Forward passes of small levels as one fused launch (glide_forward_fused_enable).
bool g_fused_forward = true;
// Cached adjoint passes of levels of at most 512 cells as one workgroup
// (glide_adjoint_coarse_enable).
bool g_coarse_adjoint = true;
// Cached adjoint passes of levels up to kAdjointPassFusedMaxCells as one launch
// each, transpose operator included (glide_adjoint_pass_fused_enable).
bool g_pass_fused_adjoint = true;
// Forward passes of levels at most 64 cells wide and 96 rows high as one
// persistent launch per pass sequence (glide_forward_persistent_enable).
bool g_persistent_forward = true;
// Adjoint passes of the same levels likewise (glide_adjoint_persistent_enable).
bool g_persistent_adjoint = true;
// Persistent launches with fenced barriers even on one XCD (glide_persistent_fenced_enable).
bool g_persistent_fenced = false;
// Polls of a persistent barrier wait before it gives up (glide_persistent_spin_limit).
int g_spin_limit = kGridSpinLimit;

// Three 64-bit words at the end of a level's smoother workspace (5 ny nx
// values, of which the alternating state of the persistent and fused passes
// uses 3 ny nx + ny + nx): the arrival counter, timeout flag and XCD mask of a
// persistent launch.
unsigned long long* work_sync(const glide_level& L) {
    return reinterpret_cast<unsigned long long*>(P(L.work) + 5 * nC(L) - 3);
}

int persistent_ctl(const glide_level& L, PersistentCtl* ctl) {
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));
    ctl->sync = work_sync(L);
    ctl->latch = s->latch;
    ctl->force_fenced = g_persistent_fenced ? 1 : 0;
    ctl->spin_limit = g_spin_limit;
    return GLIDE_OK;
}

// Multicolour passes (glide_vanka_config.multicolour): per pass the cell
// viscosity of the current state into the workspace (behind the alternate
// state), then the red phase into the alternate buffers and the black phase
// back into the level's own; the grounded fraction is refreshed before each
// pass unless frozen.
int vanka_sweeps_mc(const glide_level& L, double dt, int n_iter, const VankaArgs<R>& vk, double relax_phi,
                    bool freeze_phi, hipStream_t st) {
    R* alt_u = P(L.work);
    R* alt_v = alt_u + nU(L);
    R* alt_H = alt_v + nV(L);
    R* eta = alt_H + nC(L);
    for (int it = 0; it < n_iter; ++it) {
        if (!freeze_phi) GLIDE_TRY(grounded(L, relax_phi, st));
        SsaArgs<R> a = make_args(L, dt, false);
        HIP_TRY(launch_cell_eta<R>(a, eta, st), "cell viscosity");
        HIP_TRY(launch_vanka_mc<R>(a, eta, alt_u, alt_v, alt_H, P(L.mask), vk, 0, st), "vanka red phase");
        SsaArgs<R> b = a;
        b.u = alt_u; b.v = alt_v; b.H = alt_H;
        HIP_TRY(launch_vanka_mc<R>(b, eta, P(L.u), P(L.v), P(L.H), P(L.mask), vk, 1, st), "vanka black phase");
    }
    return GLIDE_OK;
}

int vanka_sweeps(const glide_level& L, double dt, int n_iter, const glide_vanka_config& vc, int flags,
                 hipStream_t st) {
    const bool freeze_phi = (flags & GLIDE_FREEZE_PHI) != 0;
    VankaArgs<R> vk;
    vk.newton_steps = vc.newton_steps;
    vk.ssa_damping = vc.ssa_damping;
    vk.mc_damping = vc.mc_damping;
    vk.calving_rate = (flags & GLIDE_FREEZE_CALVING) ? 0.0 : L.phys.calving_rate;
    vk.iters = nullptr;
    vk.local_fp32 = vc.local_fp32;
    vk.relax = vc.relaxation;
    vk.tol = vc.newton_tol;
    if (vc.multicolour) return vanka_sweeps_mc(L, dt, n_iter, vk, vc.relax_phi, freeze_phi, st);
    // small levels with a frozen grounded fraction: one fused launch per pass,
    // the state alternating between the level's buffers and the workspace
    // (5 planes hold u, v and H); an odd pass count copies it back at the end
    if (freeze_phi && g_fused_forward && (long)L.ny * L.nx <= kFusedVankaMaxCells) {
        R* alt_u = P(L.work);
        R* alt_v = alt_u + nU(L);
        R* alt_H = alt_v + nV(L);
        if (g_persistent_forward && n_iter > 1 && vanka_persistent_fits(L.ny, L.nx)) {
            PersistentCtl ctl;
            GLIDE_TRY(persistent_ctl(L, &ctl));
            HIP_TRY(launch_vanka_persistent<R>(make_args(L, dt, false), alt_u, alt_v, alt_H, P(L.mask), vk, vc.omega,
                                               n_iter, ctl, st),
                    "vanka passes (persistent)");
        } else {
            for (int it = 0; it < n_iter; ++it) {
                const bool from_alt = (it & 1) != 0;
                SsaArgs<R> a = make_args(L, dt, false);
                if (from_alt) { a.u = alt_u; a.v = alt_v; a.H = alt_H; }
                HIP_TRY(launch_vanka_fused<R>(a, from_alt ? P(L.u) : alt_u, from_alt ? P(L.v) : alt_v,
                                              from_alt ? P(L.H) : alt_H, P(L.mask), vk, vc.omega, st),
                        "vanka fused");
            }
        }
        if (n_iter & 1) {
            HIP_TRY(launch_copy<R>(P(L.u), alt_u, (long)nU(L), st), "copy u");
            HIP_TRY(launch_copy<R>(P(L.v), alt_v, (long)nV(L), st), "copy v");
            HIP_TRY(launch_copy<R>(P(L.H), alt_H, (long)nC(L), st), "copy H");
        }
        return GLIDE_OK;
    }
    // the grounded fraction is refreshed before every pass: before the first
    // by itself, before the others by the previous pass's increment apply
    if (!freeze_phi && n_iter > 0) GLIDE_TRY(grounded(L, vc.relax_phi, st));
    for (int it = 0; it < n_iter; ++it) {
        SsaArgs<R> a = make_args(L, dt, false);
        R* phi_next = (!freeze_phi && it + 1 < n_iter) ? P(L.phi) : nullptr;
        HIP_TRY(launch_vanka_pass<R>(a, P(L.u), P(L.v), P(L.H), P(L.work), P(L.mask), vk, vc.omega, phi_next,
                                     vc.relax_phi, st),
                "vanka");
    }
    return GLIDE_OK;
}


// The adjoint smoother cache of a level (its transposed patch factors and
// cell viscosity) when the level has one and adjoint_prepare has filled it for
// the current state; nullptr otherwise.
const R* adj_cache(const glide_level& L, bool prepared) { return prepared ? CP(L.adj_cache) : nullptr; }

int vjp(const glide_level& L, double dt, int flags, const R* lu, const R* lv, const R* lH, R* ou, R* ov, R* oH,
        hipStream_t st, bool prepared = false) {
    SsaArgs<R> a = make_args(L, dt, true);
    const bool forcing = (flags & GLIDE_USE_FORCING) != 0;
    const double cr = (flags & GLIDE_FREEZE_CALVING) ? 0.0 : L.phys.calving_rate;
    const R* cache = adj_cache(L, prepared);
    HIP_TRY(launch_vjp<R>(a, lu, lv, lH, forcing ? CP(L.af_u) : nullptr, forcing ? CP(L.af_v) : nullptr,
                          forcing ? CP(L.af_H) : nullptr, ou, ov, oH, cr,
                          cache ? cache + (size_t)kAdjointCacheEta * nC(L) : nullptr, st),
            "vjp");
    return GLIDE_OK;
}

VankaArgs<R> adjoint_vanka_args(const glide_level& L, const glide_vanka_config& vc, int flags) {
    VankaArgs<R> vk;
    vk.newton_steps = 1;
    vk.ssa_damping = vc.ssa_damping;
    vk.mc_damping = vc.mc_damping;
    vk.calving_rate = (flags & GLIDE_FREEZE_CALVING) ? 0.0 : L.phys.calving_rate;
    vk.iters = nullptr;
    vk.local_fp32 = 0;
    vk.relax = 1.0;
    vk.tol = 0.0;
    return vk;
}

// Fills the level's adjoint smoother cache for its current state (no-op
// without a cache).  The state must not change until the adjoint passes that
// use it have run.
int adjoint_prepare(const glide_level& L, double dt, const glide_vanka_config& vc, int flags, hipStream_t st) {
    if (!L.adj_cache) return GLIDE_OK;
    SsaArgs<R> a = make_args(L, dt, true);
    HIP_TRY(launch_adjoint_factor<R>(a, P(L.adj_cache), adjoint_vanka_args(L, vc, flags), st), "adjoint factor");
    return GLIDE_OK;
}

int vanka_adjoint_sweeps(const glide_level& L, double dt, int n_iter, const glide_vanka_config& vc, int flags,
                         hipStream_t st, bool prepared = false) {
    const VankaArgs<R> vk = adjoint_vanka_args(L, vc, flags);
    const R* cache = adj_cache(L, prepared);
    if (cache && g_coarse_adjoint && n_iter > 0) {
        SsaArgs<R> a = make_args(L, dt, true);
        if (adjoint_coarse_fits<R>(a)) {
            HIP_TRY(launch_adjoint_coarse<R>(a, cache, P(L.lambda_u), P(L.lambda_v), P(L.lambda_H), CP(L.af_u),
                                             CP(L.af_v), CP(L.af_H), vk.calving_rate, vc.omega, n_iter, st),
                    "adjoint passes in one workgroup");
            return GLIDE_OK;
        }
    }
    if (cache && g_pass_fused_adjoint && (long)L.ny * L.nx <= kAdjointPassFusedMaxCells) {
        SsaArgs<R> a = make_args(L, dt, true);
        if (a.P.linear_drag) {
            int n_iter_done = 0;
            // the multipliers alternate between the level's buffers and its
            // 5-plane workspace (neighbouring tiles read them during a pass);
            // an odd pass count copies them back
            R* alt_u = P(L.work);
            R* alt_v = alt_u + nU(L);
            R* alt_H = alt_v + nV(L);
            if (g_persistent_adjoint && n_iter > 1 && adjoint_persistent_fits(L.ny, L.nx)) {
                PersistentCtl ctl;
                GLIDE_TRY(persistent_ctl(L, &ctl));
                HIP_TRY(launch_adjoint_persistent<R>(a, cache, P(L.lambda_u), P(L.lambda_v), P(L.lambda_H), alt_u,
                                                     alt_v, alt_H, CP(L.af_u), CP(L.af_v), CP(L.af_H),
                                                     vk.calving_rate, vc.omega, n_iter, ctl, st),
                        "adjoint passes (persistent)");
                n_iter_done = n_iter;
            }
            for (int it = n_iter_done; it < n_iter; ++it) {
                const bool from_alt = (it & 1) != 0;
                R* su = from_alt ? alt_u : P(L.lambda_u);
                R* sv = from_alt ? alt_v : P(L.lambda_v);
                R* sH = from_alt ? alt_H : P(L.lambda_H);
                R* du = from_alt ? P(L.lambda_u) : alt_u;
                R* dv = from_alt ? P(L.lambda_v) : alt_v;
                R* dH = from_alt ? P(L.lambda_H) : alt_H;
                HIP_TRY(launch_adjoint_pass_fused<R>(a, cache, su, sv, sH, du, dv, dH, CP(L.af_u), CP(L.af_v),
                                                     CP(L.af_H), vk.calving_rate, vc.omega, st),
                        "adjoint pass");
            }
            if (n_iter & 1) {
                HIP_TRY(launch_copy<R>(P(L.lambda_u), alt_u, (long)nU(L), st), "copy lambda_u");
                HIP_TRY(launch_copy<R>(P(L.lambda_v), alt_v, (long)nV(L), st), "copy lambda_v");
                HIP_TRY(launch_copy<R>(P(L.lambda_H), alt_H, (long)nC(L), st), "copy lambda_H");
            }
            return GLIDE_OK;
        }
    }
    for (int it = 0; it < n_iter; ++it) {
        GLIDE_TRY(vjp(L, dt, flags | GLIDE_USE_FORCING, CP(L.lambda_u), CP(L.lambda_v), CP(L.lambda_H), P(L.ar_u),
                      P(L.ar_v), P(L.ar_H), st, prepared));
        if (cache) {
            HIP_TRY(launch_vanka_adjoint_pass_cached<R>(L.ny, L.nx, cache, CP(L.ar_u), CP(L.ar_v), CP(L.ar_H),
                                                        P(L.lambda_u), P(L.lambda_v), P(L.lambda_H), P(L.work),
                                                        vc.omega, g_fused_adjoint, st),
                    "vanka adjoint");
        } else {
            SsaArgs<R> a = make_args(L, dt, true);
            HIP_TRY(launch_vanka_adjoint_pass<R>(a, CP(L.ar_u), CP(L.ar_v), CP(L.ar_H), P(L.lambda_u), P(L.lambda_v),
                                                 P(L.lambda_H), P(L.work), vk, vc.omega, st),
                    "vanka adjoint");
        }
    }
    return GLIDE_OK;
}

int restrict_c(const void* f, void* c, const glide_level& C, int use_max, hipStream_t st) {
    HIP_TRY(launch_restrict_cell<R>(CP(f), P(c), C.ny, C.nx, use_max, st), "restrict cell");
    return GLIDE_OK;
}
int axpby(void* out, double a, const void* x, double b, const void* y, size_t n, hipStream_t st) {
    HIP_TRY(launch_axpby<R>(P(out), a, CP(x), b, CP(y), 0.0, nullptr, (long)n, st), "axpby");
    return GLIDE_OK;
}
int copy(void* dst, const void* src, size_t n, hipStream_t st) {
    HIP_TRY(launch_copy<R>(P(dst), CP(src), (long)n, st), "copy");
    return GLIDE_OK;
}
int fill(void* dst, double val, size_t n, hipStream_t st) {
    HIP_TRY(launch_fill<R>(P(dst), val, (long)n, st), "fill");
    return GLIDE_OK;
}

// Restrict the forward state of level l to level l+1 (glide/multigrid.py:81-87).
int prolong_v(const void* c, void* f, const glide_level& F, int compat, hipStream_t st) {
    if (compat)
        HIP_TRY(launch_prolong_vfacet_bilinear<R>(CP(c), P(f), F.ny, F.nx, st), "prolong v (compat)");
    else
        HIP_TRY(launch_prolong_hfacet_bilinear<R>(CP(c), P(f), F.ny, F.nx, st), "prolong v");
    return GLIDE_OK;
}

struct FasCtx {
    const glide_level* L;
    int n_levels;
    double dt;
    glide_fas_config cfg;
    const glide_vanka_config* vc;
    hipStream_t st;
};

// FASCDSolver.vcycle (glide/multigrid.py:536-645).
int fas_vcycle(FasCtx& X, int l, bool finest) {
    const glide_level& L = X.L[l];
    const bool coarse = !finest;
    int sflags = 0;
    if (coarse && X.cfg.freeze_coarse_calving) sflags |= GLIDE_FREEZE_CALVING;
    if (coarse && X.cfg.freeze_coarse_phi) sflags |= GLIDE_FREEZE_PHI;
    const double thk = L.phys.thklim;
    hipStream_t st = X.st;
    // independent element-wise jobs of a stage run as one batched launch
    auto run = [&](const EwBatch<R>& b) -> int {
        HIP_TRY(launch_ew_batch<R>(b, st), "element-wise batch");
        return GLIDE_OK;
    };
    if (finest) {
        // w = x; chi = -H (then + thklim)
        EwBatch<R> b;
        b.copy(P(L.w_u), CP(L.u), (long)nU(L));
        b.copy(P(L.w_v), CP(L.v), (long)nV(L));
        b.copy(P(L.w_H), CP(L.H), (long)nC(L));
        b.axpby(P(L.chi), -1.0, CP(L.H), 0.0, CP(L.H), (long)nC(L));
        b.fill(P(L.gamma), thk, (long)nC(L));
        GLIDE_TRY(run(b));
        EwBatch<R> c;
        c.axpby(P(L.chi), 1.0, CP(L.chi), 1.0, CP(L.gamma), (long)nC(L));
        GLIDE_TRY(run(c));
    }
    if (l == X.n_levels - 1) {
        GLIDE_TRY(axpby(L.gamma, 1.0, L.w_H, 1.0, L.chi, nC(L), st));
        GLIDE_TRY(vanka_sweeps(L, X.dt, X.cfg.coarsest_steps, X.vc[l], sflags, st));
        GLIDE_TRY(fill(L.gamma, thk, nC(L), st));
        return GLIDE_OK;
    }
    const glide_level& N = X.L[l + 1];
    GLIDE_TRY(restrict_c(L.chi, N.chi, N, 1, st));
    // cphi = P_inj(-chi_c) + chi
    HIP_TRY(launch_prolong_cell<R>(CP(N.chi), P(L.cphi), L.ny, L.nx, 0, st), "prolong chi");
    GLIDE_TRY(axpby(L.cphi, -1.0, L.cphi, 1.0, L.chi, nC(L), st));
    GLIDE_TRY(axpby(L.gamma, 1.0, L.w_H, 1.0, L.cphi, nC(L), st));
    GLIDE_TRY(vanka_sweeps(L, X.dt, X.cfg.pre_steps, X.vc[l], sflags, st));
    {
        // gamma back to thklim; y = x - w
        EwBatch<R> b;
        b.fill(P(L.gamma), thk, (long)nC(L));
        b.axpby(P(L.y_u), 1.0, CP(L.u), -1.0, CP(L.w_u), (long)nU(L));
        b.axpby(P(L.y_v), 1.0, CP(L.v), -1.0, CP(L.w_v), (long)nV(L));
        b.axpby(P(L.y_H), 1.0, CP(L.H), -1.0, CP(L.w_H), (long)nC(L));
        GLIDE_TRY(run(b));
    }
    {
        // restrict the state
        RestrictBatch<R> rb;
        rb.add(CP(L.u), P(N.u), kRestrictU);
        rb.add(CP(L.v), P(N.v), kRestrictV);
        rb.add(CP(L.H), P(N.H), kRestrictCellMean);
        rb.add(CP(L.H_prev), P(N.H_prev), kRestrictCellMean);
        rb.add(CP(L.phi), P(N.phi), kRestrictCellMean);
        rb.add(CP(L.mask), P(N.mask), kRestrictCellMax);
        HIP_TRY(launch_restrict_batch<R>(rb, N.ny, N.nx, st), "restrict state");
    }
    {
        EwBatch<R> b;
        b.copy(P(N.w_u), CP(N.u), (long)nU(N));
        b.copy(P(N.w_v), CP(N.v), (long)nV(N));
        b.copy(P(N.w_H), CP(N.H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(residual(L, X.dt, GLIDE_USE_FORCING | (sflags & GLIDE_FREEZE_CALVING), P(L.r_u), P(L.r_v), P(L.r_H), st));
    {
        RestrictBatch<R> rb;
        rb.add(CP(L.r_u), P(N.r_u), kRestrictU);
        rb.add(CP(L.r_v), P(N.r_v), kRestrictV);
        rb.add(CP(L.r_H), P(N.r_H), kRestrictCellMean);
        HIP_TRY(launch_restrict_batch<R>(rb, N.ny, N.nx, st), "restrict residual");
    }
    if (!X.cfg.freeze_coarse_phi) GLIDE_TRY(grounded(N, 0.0, st));
    GLIDE_TRY(residual(N, X.dt, X.cfg.freeze_coarse_calving ? GLIDE_FREEZE_CALVING : 0, P(N.F_u), P(N.F_v), P(N.F_H),
                       st));
    {
        EwBatch<R> b;
        b.axpby(P(N.f_u), 1.0, CP(N.F_u), -1.0, CP(N.r_u), (long)nU(N));
        b.axpby(P(N.f_v), 1.0, CP(N.F_v), -1.0, CP(N.r_v), (long)nV(N));
        b.axpby(P(N.f_H), 1.0, CP(N.F_H), -1.0, CP(N.r_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(fas_vcycle(X, l + 1, false));
    {
        EwBatch<R> b;
        b.axpby(P(N.z_u), 1.0, CP(N.u), -1.0, CP(N.w_u), (long)nU(N));
        b.axpby(P(N.z_v), 1.0, CP(N.v), -1.0, CP(N.w_v), (long)nV(N));
        b.axpby(P(N.z_H), 1.0, CP(N.H), -1.0, CP(N.w_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    HIP_TRY(launch_prolong_vfacet_bilinear<R>(CP(N.z_u), P(L.z_u), L.ny, L.nx, st), "prolong z_u");
    GLIDE_TRY(prolong_v(N.z_v, L.z_v, L, X.cfg.compat_vprolong, st));
    HIP_TRY(launch_prolong_cell<R>(CP(N.z_H), P(L.z_H), L.ny, L.nx, 0, st), "prolong z_H");
    {
        // z += y; x = w + z; gamma = w_H + chi
        EwBatch<R> b;
        b.correct(P(L.z_u), CP(L.y_u), P(L.u), CP(L.w_u), (long)nU(L));
        b.correct(P(L.z_v), CP(L.y_v), P(L.v), CP(L.w_v), (long)nV(L));
        b.correct(P(L.z_H), CP(L.y_H), P(L.H), CP(L.w_H), (long)nC(L));
        b.axpby(P(L.gamma), 1.0, CP(L.w_H), 1.0, CP(L.chi), (long)nC(L));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(vanka_sweeps(L, X.dt, X.cfg.post_steps, X.vc[l], sflags, st));
    GLIDE_TRY(fill(L.gamma, thk, nC(L), st));
    if (finest) GLIDE_TRY(vanka_sweeps(L, X.dt, X.cfg.finest_steps, X.vc[l], 0, st));
    return GLIDE_OK;
}

// FASAdjointSolver.vcycle (glide/multigrid.py:905-973).
int adj_vcycle(FasCtx& X, int l, bool finest) {
    const glide_level& L = X.L[l];
    hipStream_t st = X.st;
    // every level's cache is filled for its current state before its first pass
    // (the start level by the caller, coarser levels after their restriction)
    if (l == X.n_levels - 1) return vanka_adjoint_sweeps(L, X.dt, X.cfg.coarsest_steps, X.vc[l], 0, st, true);
    const glide_level& N = X.L[l + 1];
    GLIDE_TRY(vanka_adjoint_sweeps(L, X.dt, X.cfg.pre_steps, X.vc[l], 0, st, true));
    auto run = [&](const EwBatch<R>& b) -> int {
        HIP_TRY(launch_ew_batch<R>(b, st), "element-wise batch");
        return GLIDE_OK;
    };
    {
        // restrict the state and the multipliers
        RestrictBatch<R> rb;
        rb.add(CP(L.u), P(N.u), kRestrictU);
        rb.add(CP(L.v), P(N.v), kRestrictV);
        rb.add(CP(L.H), P(N.H), kRestrictCellMean);
        rb.add(CP(L.H_prev), P(N.H_prev), kRestrictCellMean);
        rb.add(CP(L.phi), P(N.phi), kRestrictCellMean);
        rb.add(CP(L.mask), P(N.mask), kRestrictCellMax);
        rb.add(CP(L.lambda_u), P(N.lambda_u), kRestrictU);
        rb.add(CP(L.lambda_v), P(N.lambda_v), kRestrictV);
        rb.add(CP(L.lambda_H), P(N.lambda_H), kRestrictCellMean);
        HIP_TRY(launch_restrict_batch<R>(rb, N.ny, N.nx, st), "restrict state and multipliers");
    }
    GLIDE_TRY(adjoint_prepare(N, X.dt, X.vc[l + 1], 0, st));
    {
        EwBatch<R> b;
        b.copy(P(N.wl_u), CP(N.lambda_u), (long)nU(N));
        b.copy(P(N.wl_v), CP(N.lambda_v), (long)nV(N));
        b.copy(P(N.wl_H), CP(N.lambda_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(vjp(L, X.dt, GLIDE_USE_FORCING, CP(L.lambda_u), CP(L.lambda_v), CP(L.lambda_H), P(L.ar_u), P(L.ar_v),
                  P(L.ar_H), st, true));
    {
        RestrictBatch<R> rb;
        rb.add(CP(L.ar_u), P(N.ar_u), kRestrictU);
        rb.add(CP(L.ar_v), P(N.ar_v), kRestrictV);
        rb.add(CP(L.ar_H), P(N.ar_H), kRestrictCellMean);
        HIP_TRY(launch_restrict_batch<R>(rb, N.ny, N.nx, st), "restrict adjoint residual");
    }
    GLIDE_TRY(vjp(N, X.dt, 0, CP(N.lambda_u), CP(N.lambda_v), CP(N.lambda_H), P(N.vjp_u), P(N.vjp_v), P(N.vjp_H), st,
                  true));
    {
        EwBatch<R> b;
        b.axpby(P(N.af_u), 1.0, CP(N.vjp_u), -1.0, CP(N.ar_u), (long)nU(N));
        b.axpby(P(N.af_v), 1.0, CP(N.vjp_v), -1.0, CP(N.ar_v), (long)nV(N));
        b.axpby(P(N.af_H), 1.0, CP(N.vjp_H), -1.0, CP(N.ar_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(adj_vcycle(X, l + 1, false));
    {
        EwBatch<R> b;
        b.axpby(P(N.zl_u), 1.0, CP(N.lambda_u), -1.0, CP(N.wl_u), (long)nU(N));
        b.axpby(P(N.zl_v), 1.0, CP(N.lambda_v), -1.0, CP(N.wl_v), (long)nV(N));
        b.axpby(P(N.zl_H), 1.0, CP(N.lambda_H), -1.0, CP(N.wl_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    HIP_TRY(launch_prolong_vfacet_bilinear<R>(CP(N.zl_u), P(L.zl_u), L.ny, L.nx, st), "prolong zl_u");
    GLIDE_TRY(prolong_v(N.zl_v, L.zl_v, L, X.cfg.compat_vprolong, st));
    HIP_TRY(launch_prolong_cell<R>(CP(N.zl_H), P(L.zl_H), L.ny, L.nx, 1, st), "prolong zl_H");
    {
        EwBatch<R> b;
        b.axpby(P(L.lambda_u), 1.0, CP(L.lambda_u), 1.0, CP(L.zl_u), (long)nU(L));
        b.axpby(P(L.lambda_v), 1.0, CP(L.lambda_v), 1.0, CP(L.zl_v), (long)nV(L));
        b.axpby(P(L.lambda_H), 1.0, CP(L.lambda_H), 1.0, CP(L.zl_H), (long)nC(L));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(vanka_adjoint_sweeps(L, X.dt, X.cfg.post_steps, X.vc[l], 0, st, true));
    if (finest) GLIDE_TRY(vanka_adjoint_sweeps(L, X.dt, X.cfg.finest_steps, X.vc[l], 0, st, true));
    return GLIDE_OK;
}

int check_hierarchy(const glide_level* levels, int n_levels, int start, const glide_fas_config* cfg,
                    const glide_vanka_config* vc) {
    if (!levels || !cfg || !vc) return fail(GLIDE_ERR_INVALID_ARG, "levels, cfg and vc must be non-NULL");
    if (n_levels < 1 || start < 0 || start >= n_levels) return fail(GLIDE_ERR_INVALID_ARG, "bad level range");
    if (cfg->maximum_vcycles < 0 || cfg->maximum_vcycles > 64)
        return fail(GLIDE_ERR_INVALID_ARG, "maximum_vcycles must be in [0, 64]");
    for (int l = start; l < n_levels; ++l) {
        GLIDE_TRY(check_level(&levels[l]));
        GLIDE_TRY(check_ssa_inputs(levels[l]));
        if (!levels[l].work) return fail(GLIDE_ERR_INVALID_ARG, "level workspace is NULL");
        if (l > start) {
            if (levels[l].ny * 2 != levels[l - 1].ny || levels[l].nx * 2 != levels[l - 1].nx)
                return fail(GLIDE_ERR_SHAPE, "each coarse level must halve the finer level's ny and nx exactly");
        }
        if (cfg->compat_vprolong && n_levels - start > 1 && levels[l].ny != levels[l].nx)
            return fail(GLIDE_ERR_SHAPE, "compat_vprolong requires square levels");
        if (vc[l].newton_steps < 0) return fail(GLIDE_ERR_INVALID_ARG, "newton_steps must be >= 0");
        if (vc[l].multicolour && (long)levels[l].ny * levels[l].nx < (long)levels[l].ny + levels[l].nx)
            return fail(GLIDE_ERR_SHAPE, "multicolour smoothing needs ny * nx >= ny + nx on every level");
    }
    return GLIDE_OK;
}

// ---------------------------------------------------------------- graph replay
// A fixed-cycle solve enqueues the same few thousand launches every call.
// Those launches are captured once into a HIP graph on a private stream and
// replayed on the caller's stream afterwards, which removes most of the
// per-launch dispatch gap between the many short dependent kernels of the
// coarse levels.  The cache key is every byte the launches depend on (the
// level descriptors with their device pointers, the configs, dt, the timing
// switch); buffer contents are read at replay time as in eager execution.
// Captured graphs hold no memset nodes and no host work.  Elapsed times of
// events recorded inside a graph cannot be queried, so solves run eagerly
// while launch timing is on.
struct GraphEntry {
    std::string key;
    hipGraphExec_t exec = nullptr;
    unsigned long last_use = 0;
};
struct DeviceGraphs {
    hipStream_t cap = nullptr;
    std::vector<GraphEntry> entries;
    unsigned long clock = 0;
};
size_t g_max_graphs = 8;  // cached graphs per device (least recently used dropped)
bool g_graphs_on = true;
std::mutex g_graph_mutex;
std::vector<DeviceGraphs> g_graphs;

template <typename T>
void key_append(std::string& k, const T& v) {
    k.append(reinterpret_cast<const char*>(&v), sizeof(T));
}

std::string solve_key(char tag, const FasCtx& X, int start, bool norms) {
    std::string k(1, tag);
    key_append(k, norms);
    key_append(k, g_fused_forward);
    key_append(k, g_fused_adjoint);
    key_append(k, g_coarse_adjoint);
    key_append(k, g_pass_fused_adjoint);
    key_append(k, g_persistent_forward);
    key_append(k, g_persistent_adjoint);
    key_append(k, g_persistent_fenced);
    key_append(k, g_spin_limit);
    key_append(k, X.n_levels);
    key_append(k, start);
    key_append(k, X.dt);
    key_append(k, X.cfg);
    for (int l = start; l < X.n_levels; ++l) {
        key_append(k, X.L[l]);
        key_append(k, X.vc[l]);
    }
    return k;
}

void drop_entry(GraphEntry& e) {
    if (e.exec) hipGraphExecDestroy(e.exec);
    e.exec = nullptr;
}

// Replays (capturing on first use) the launches `body` enqueues on the stream
// it is given; on return the replay is enqueued on `st`.
template <typename Body>
int replay(const std::string& key, hipStream_t st, Body&& body) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_graph_mutex);
    if ((int)g_graphs.size() <= dev) g_graphs.resize(dev + 1);
    DeviceGraphs& G = g_graphs[dev];
    GraphEntry* e = nullptr;
    for (auto& x : G.entries)
        if (x.key == key) e = &x;
    if (!e) {
        if (!G.cap) HIP_TRY(hipStreamCreateWithFlags(&G.cap, hipStreamNonBlocking), "hipStreamCreate(capture)");
        HIP_TRY(hipStreamBeginCapture(G.cap, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        int rc = body(G.cap);
        hipGraph_t graph = nullptr;
        hipError_t ee = hipStreamEndCapture(G.cap, &graph);
        if (rc != GLIDE_OK || ee != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            if (rc != GLIDE_OK) return rc;
            return hip_fail(ee, "hipStreamEndCapture");
        }
        hipGraphExec_t exec = nullptr;
        hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate");
        while (!G.entries.empty() && G.entries.size() >= g_max_graphs) {
            size_t lru = 0;
            for (size_t q = 1; q < G.entries.size(); ++q)
                if (G.entries[q].last_use < G.entries[lru].last_use) lru = q;
            drop_entry(G.entries[lru]);
            G.entries.erase(G.entries.begin() + lru);
        }
        G.entries.push_back(GraphEntry{key, exec, 0});
        e = &G.entries.back();
    }
    e->last_use = ++G.clock;
    HIP_TRY(hipGraphLaunch(e->exec, st), "hipGraphLaunch");
    return GLIDE_OK;
}

// Shared solve loop of the forward and adjoint FAS drivers
// (glide/multigrid.py:502-533, 873-903).
template <typename Residual, typename Cycle>
int solve_loop(FasCtx& X, int start, glide_solve_report* rep, Residual&& resid, Cycle&& cycle, char tag) {
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));
    hipStream_t st = X.st;
    const bool fixed = !(X.cfg.relative_tolerance > 0.0) && !(X.cfg.absolute_tolerance > 0.0);
    double r0 = 0.0, absn = 0.0, rel = 1.0;
    double n3[3] = {0, 0, 0};
    // every synchronisation also reads the failure latch of the persistent
    // launches; a timed-out barrier or a non-finite norm fails the solve
    auto read_slots = [&](int nslots) -> int {
        HIP_TRY(hipMemcpyAsync(s->host, s->norms, sizeof(double) * 3 * nslots, hipMemcpyDeviceToHost, st), "norms D2H");
        GLIDE_TRY(latch_enqueue(s, st));
        HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
        GLIDE_TRY(latch_check(s));
        for (int q = 0; q < 3 * nslots; ++q)
            if (!std::isfinite(s->host[q]))
                return fail(GLIDE_ERR_NONFINITE, "non-finite residual norm in the multigrid solve");
        return GLIDE_OK;
    };
    int it = 0;
    if (fixed) {
        // initial residual (slot 0), the cycles, the residual after each
        // (slots 1..); without a report the norms are not computed at all
        const bool norms = rep != nullptr;
        auto body = [&](hipStream_t cs) -> int {
            X.st = cs;
            int rc = norms ? resid(true, s->norms) : GLIDE_OK;
            for (int k = 0; rc == GLIDE_OK && k < X.cfg.maximum_vcycles; ++k) {
                rc = cycle();
                if (rc == GLIDE_OK && norms) rc = resid(false, s->norms + 3 * (k + 1));
            }
            X.st = st;
            return rc;
        };
        it = X.cfg.maximum_vcycles;
        if (g_graphs_on && !timing_on()) GLIDE_TRY(replay(solve_key(tag, X, start, norms), st, body));
        else GLIDE_TRY(body(st));
        if (!rep) return GLIDE_OK;  // nothing to report: no host synchronisation
        GLIDE_TRY(read_slots(it + 1));
        r0 = std::sqrt(s->host[0] + s->host[1] + s->host[2]);
        absn = r0;
        for (int k = 0; k < it; ++k) {
            const double* q = s->host + 3 * (k + 1);
            absn = std::sqrt(q[0] + q[1] + q[2]);
            if (rep) rep->norms[k] = absn;
        }
        const double* q = s->host + 3 * it;
        n3[0] = q[0]; n3[1] = q[1]; n3[2] = q[2];
    } else {
        GLIDE_TRY(resid(true, s->norms));
        GLIDE_TRY(read_slots(1));
        r0 = std::sqrt(s->host[0] + s->host[1] + s->host[2]);
        n3[0] = s->host[0]; n3[1] = s->host[1]; n3[2] = s->host[2];
        absn = r0;
        while (rel > X.cfg.relative_tolerance && absn > X.cfg.absolute_tolerance && it < X.cfg.maximum_vcycles) {
            GLIDE_TRY(cycle());
            GLIDE_TRY(resid(false, s->norms));
            GLIDE_TRY(read_slots(1));
            n3[0] = s->host[0]; n3[1] = s->host[1]; n3[2] = s->host[2];
            absn = std::sqrt(n3[0] + n3[1] + n3[2]);
            rel = absn / r0;
            if (rep) rep->norms[it] = absn;
            ++it;
        }
    }
    if (rep) {
        rep->vcycles = it;
        rep->initial_norm = r0;
        rep->final_norm = absn;
        rep->norms3[0] = std::sqrt(n3[0]); rep->norms3[1] = std::sqrt(n3[1]); rep->norms3[2] = std::sqrt(n3[2]);
    }
    return GLIDE_OK;
}


// ====================================================== mixed precision
// The mixed-precision form (no counterpart in the reference): an FP64
// defect-correction loop around FP32 FAS-CD V-cycles.  Per outer iteration:
// the FP64 residual r(x) of the finest level, the FP32 copy x32 = fl(x), the
// FP32 forcing f32 = F32(x32) - fl(r(x)), FP32 V-cycles with the red-black
// smoother (the coarsest level in FP64) from x32, and the FP64 update
// x += y32 - fl(x).  The FP32 operator's own rounding cancels in the
// correction, so the loop converges to the FP64 solution; the FP64 residual
// decides when to stop, and the FP64 state owns the active set (masked cells
// are set to the thickness limit exactly).  An outer iteration whose FP64
// residual grows is undone and replaced by one FP64 V-cycle of the reference
// form (additive smoother, default counts).

template <typename T>
inline T* PT(void* p) { return static_cast<T*>(p); }
template <typename T>
inline const T* CPT(const void* p) { return static_cast<const T*>(p); }

template <typename T>
SsaArgs<T> make_args_t(const glide_level& L, double dt, bool adjoint_forcing) {
    const Phys<R> q = make_phys(L.phys);
    SsaArgs<T> a;
    a.ny = L.ny; a.nx = L.nx; a.dxi = T(1.0 / L.dx); a.dt = T(dt);
    a.P.n = T(q.n); a.P.eps_reg = T(q.eps_reg); a.P.frd = T(q.frd); a.P.m = T(q.m); a.P.u_reg = T(q.u_reg);
    a.P.water_drag = T(q.water_drag); a.P.frs = T(q.frs); a.P.calving_rate = T(q.calving_rate); a.P.frc = T(q.frc);
    a.P.thklim = T(q.thklim); a.P.glen_p = T(q.glen_p); a.P.drag_q = T(q.drag_q);
    a.P.linear_drag = q.linear_drag; a.P.glen_cube = q.glen_cube;
    a.u = CPT<T>(L.u); a.v = CPT<T>(L.v); a.H = CPT<T>(L.H); a.phi = CPT<T>(L.phi); a.mask = CPT<T>(L.mask);
    a.bed = CPT<T>(L.bed); a.B = CPT<T>(L.B); a.beta = CPT<T>(L.beta); a.gamma = CPT<T>(L.gamma);
    a.f_u = CPT<T>(adjoint_forcing ? L.af_u : L.f_u);
    a.f_v = CPT<T>(adjoint_forcing ? L.af_v : L.f_v);
    a.f_H = CPT<T>(adjoint_forcing ? L.af_H : L.f_H);
    return a;
}

template <typename T>
int residual_t(const glide_level& L, double dt, int flags, T* ou, T* ov, T* oH, hipStream_t st) {
    SsaArgs<T> a = make_args_t<T>(L, dt, false);
    const T cr = T((flags & GLIDE_FREEZE_CALVING) ? 0.0 : L.phys.calving_rate);
    HIP_TRY(launch_residual<T>(a, ou, ov, oH, (flags & GLIDE_USE_FORCING) ? 1 : 0, (flags & GLIDE_USE_MASK) ? 1 : 0,
                               cr, st),
            "residual");
    return GLIDE_OK;
}

// Multicolour passes on a level of element type T (see vanka_sweeps_mc).
template <typename T>
int mc_sweeps_t(const glide_level& L, double dt, int n_iter, const glide_vanka_config& vc, int flags,
                hipStream_t st) {
    const bool freeze_phi = (flags & GLIDE_FREEZE_PHI) != 0;
    VankaArgs<T> vk;
    vk.newton_steps = vc.newton_steps;
    vk.ssa_damping = T(vc.ssa_damping);
    vk.mc_damping = T(vc.mc_damping);
    vk.calving_rate = T((flags & GLIDE_FREEZE_CALVING) ? 0.0 : L.phys.calving_rate);
    vk.iters = nullptr;
    vk.local_fp32 = 0;
    vk.relax = T(vc.relaxation);
    vk.tol = T(vc.newton_tol);
    T* alt_u = PT<T>(L.work);
    T* alt_v = alt_u + nU(L);
    T* alt_H = alt_v + nV(L);
    T* eta = alt_H + nC(L);
    for (int it = 0; it < n_iter; ++it) {
        if (!freeze_phi)
            HIP_TRY(launch_grounded<T>(PT<T>(L.phi), CPT<T>(L.H), CPT<T>(L.bed), T(L.phys.flotation_reg_driving),
                                       T(vc.relax_phi), (long)nC(L), st),
                    "grounded");
        SsaArgs<T> a = make_args_t<T>(L, dt, false);
        HIP_TRY(launch_cell_eta<T>(a, eta, st), "cell viscosity");
        HIP_TRY(launch_vanka_mc<T>(a, eta, alt_u, alt_v, alt_H, PT<T>(L.mask), vk, 0, st), "vanka red phase");
        SsaArgs<T> b = a;
        b.u = alt_u; b.v = alt_v; b.H = alt_H;
        HIP_TRY(launch_vanka_mc<T>(b, eta, PT<T>(L.u), PT<T>(L.v), PT<T>(L.H), PT<T>(L.mask), vk, 1, st),
                "vanka black phase");
    }
    return GLIDE_OK;
}

struct MixedCtx {
    const glide_level* L64;  // FP64 hierarchy (the coarsest level's passes, the outer residual)
    const glide_level* L32;  // FP32 hierarchy of the inner V-cycles
    int n_levels;
    double dt;
    glide_fas_config cfg;
    const glide_vanka_config* vc;
    bool fp64_coarsest;
    hipStream_t st;
};

int run_mixed(const MixedBatch& b, hipStream_t st) {
    HIP_TRY(launch_mixed_batch(b, st), "precision conversion");
    return GLIDE_OK;
}

// The coarsest level's passes of an FP32 V-cycle, run on the FP64 copy of the
// level: state, forcing and thickness bounds in, state and active set out.
int coarsest_fp64(MixedCtx& X, int l, int n_iter, int flags) {
    const glide_level& A = X.L32[l];
    const glide_level& D = X.L64[l];
    MixedBatch in;
    in.add(kMixTo64, D.u, A.u, nullptr, (long)nU(A));
    in.add(kMixTo64, D.v, A.v, nullptr, (long)nV(A));
    in.add(kMixTo64, D.H, A.H, nullptr, (long)nC(A));
    in.add(kMixTo64, D.mask, A.mask, nullptr, (long)nC(A));
    in.add(kMixTo64, D.phi, A.phi, nullptr, (long)nC(A));
    in.add(kMixTo64, D.f_u, A.f_u, nullptr, (long)nU(A));
    in.add(kMixTo64, D.f_v, A.f_v, nullptr, (long)nV(A));
    in.add(kMixTo64, D.f_H, A.f_H, nullptr, (long)nC(A));
    in.add(kMixTo64, D.gamma, A.gamma, nullptr, (long)nC(A));
    GLIDE_TRY(run_mixed(in, X.st));
    GLIDE_TRY(mc_sweeps_t<double>(D, X.dt, n_iter, X.vc[l], flags, X.st));
    MixedBatch out;
    out.add(kMixTo32, A.u, D.u, nullptr, (long)nU(A));
    out.add(kMixTo32, A.v, D.v, nullptr, (long)nV(A));
    out.add(kMixTo32, A.H, D.H, nullptr, (long)nC(A));
    out.add(kMixTo32, A.mask, D.mask, nullptr, (long)nC(A));
    return run_mixed(out, X.st);
}

// FASCDSolver.vcycle (glide/multigrid.py:536-645) on the FP32 hierarchy with
// the multicolour smoother; the same stages as fas_vcycle.
int mixed_vcycle(MixedCtx& X, int l, bool finest) {
    typedef float F;
    const glide_level& L = X.L32[l];
    const bool coarse = !finest;
    int sflags = 0;
    if (coarse && X.cfg.freeze_coarse_calving) sflags |= GLIDE_FREEZE_CALVING;
    if (coarse && X.cfg.freeze_coarse_phi) sflags |= GLIDE_FREEZE_PHI;
    const F thk = F(L.phys.thklim);
    hipStream_t st = X.st;
    auto run = [&](const EwBatch<F>& b) -> int {
        HIP_TRY(launch_ew_batch<F>(b, st), "element-wise batch");
        return GLIDE_OK;
    };
    auto smooth = [&](const glide_level& A, int n, int flags) -> int {
        if (n <= 0) return GLIDE_OK;
        return mc_sweeps_t<F>(A, X.dt, n, X.vc[l], flags, st);
    };
    if (finest) {
        EwBatch<F> b;
        b.copy(PT<F>(L.w_u), CPT<F>(L.u), (long)nU(L));
        b.copy(PT<F>(L.w_v), CPT<F>(L.v), (long)nV(L));
        b.copy(PT<F>(L.w_H), CPT<F>(L.H), (long)nC(L));
        b.axpby(PT<F>(L.chi), F(-1), CPT<F>(L.H), F(0), CPT<F>(L.H), (long)nC(L));
        b.fill(PT<F>(L.gamma), thk, (long)nC(L));
        GLIDE_TRY(run(b));
        EwBatch<F> c;
        c.axpby(PT<F>(L.chi), F(1), CPT<F>(L.chi), F(1), CPT<F>(L.gamma), (long)nC(L));
        GLIDE_TRY(run(c));
    }
    if (l == X.n_levels - 1) {
        EwBatch<F> b;
        b.axpby(PT<F>(L.gamma), F(1), CPT<F>(L.w_H), F(1), CPT<F>(L.chi), (long)nC(L));
        GLIDE_TRY(run(b));
        if (X.fp64_coarsest) GLIDE_TRY(coarsest_fp64(X, l, X.cfg.coarsest_steps, sflags));
        else GLIDE_TRY(smooth(L, X.cfg.coarsest_steps, sflags));
        EwBatch<F> c;
        c.fill(PT<F>(L.gamma), thk, (long)nC(L));
        return run(c);
    }
    const glide_level& N = X.L32[l + 1];
    HIP_TRY(launch_restrict_cell<F>(CPT<F>(L.chi), PT<F>(N.chi), N.ny, N.nx, 1, st), "restrict chi");
    HIP_TRY(launch_prolong_cell<F>(CPT<F>(N.chi), PT<F>(L.cphi), L.ny, L.nx, 0, st), "prolong chi");
    {
        EwBatch<F> b;
        b.axpby(PT<F>(L.cphi), F(-1), CPT<F>(L.cphi), F(1), CPT<F>(L.chi), (long)nC(L));
        GLIDE_TRY(run(b));
        EwBatch<F> c;
        c.axpby(PT<F>(L.gamma), F(1), CPT<F>(L.w_H), F(1), CPT<F>(L.cphi), (long)nC(L));
        GLIDE_TRY(run(c));
    }
    GLIDE_TRY(smooth(L, X.cfg.pre_steps, sflags));
    {
        EwBatch<F> b;
        b.fill(PT<F>(L.gamma), thk, (long)nC(L));
        b.axpby(PT<F>(L.y_u), F(1), CPT<F>(L.u), F(-1), CPT<F>(L.w_u), (long)nU(L));
        b.axpby(PT<F>(L.y_v), F(1), CPT<F>(L.v), F(-1), CPT<F>(L.w_v), (long)nV(L));
        b.axpby(PT<F>(L.y_H), F(1), CPT<F>(L.H), F(-1), CPT<F>(L.w_H), (long)nC(L));
        GLIDE_TRY(run(b));
    }
    {
        RestrictBatch<F> rb;
        rb.add(CPT<F>(L.u), PT<F>(N.u), kRestrictU);
        rb.add(CPT<F>(L.v), PT<F>(N.v), kRestrictV);
        rb.add(CPT<F>(L.H), PT<F>(N.H), kRestrictCellMean);
        rb.add(CPT<F>(L.phi), PT<F>(N.phi), kRestrictCellMean);
        rb.add(CPT<F>(L.mask), PT<F>(N.mask), kRestrictCellMax);
        HIP_TRY(launch_restrict_batch<F>(rb, N.ny, N.nx, st), "restrict state");
    }
    {
        EwBatch<F> b;
        b.copy(PT<F>(N.w_u), CPT<F>(N.u), (long)nU(N));
        b.copy(PT<F>(N.w_v), CPT<F>(N.v), (long)nV(N));
        b.copy(PT<F>(N.w_H), CPT<F>(N.H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(residual_t<F>(L, X.dt, GLIDE_USE_FORCING | (sflags & GLIDE_FREEZE_CALVING), PT<F>(L.r_u), PT<F>(L.r_v),
                            PT<F>(L.r_H), st));
    {
        RestrictBatch<F> rb;
        rb.add(CPT<F>(L.r_u), PT<F>(N.r_u), kRestrictU);
        rb.add(CPT<F>(L.r_v), PT<F>(N.r_v), kRestrictV);
        rb.add(CPT<F>(L.r_H), PT<F>(N.r_H), kRestrictCellMean);
        HIP_TRY(launch_restrict_batch<F>(rb, N.ny, N.nx, st), "restrict residual");
    }
    if (!X.cfg.freeze_coarse_phi)
        HIP_TRY(launch_grounded<F>(PT<F>(N.phi), CPT<F>(N.H), CPT<F>(N.bed), F(N.phys.flotation_reg_driving), F(0),
                                   (long)nC(N), st),
                "grounded");
    GLIDE_TRY(residual_t<F>(N, X.dt, X.cfg.freeze_coarse_calving ? GLIDE_FREEZE_CALVING : 0, PT<F>(N.F_u),
                            PT<F>(N.F_v), PT<F>(N.F_H), st));
    {
        EwBatch<F> b;
        b.axpby(PT<F>(N.f_u), F(1), CPT<F>(N.F_u), F(-1), CPT<F>(N.r_u), (long)nU(N));
        b.axpby(PT<F>(N.f_v), F(1), CPT<F>(N.F_v), F(-1), CPT<F>(N.r_v), (long)nV(N));
        b.axpby(PT<F>(N.f_H), F(1), CPT<F>(N.F_H), F(-1), CPT<F>(N.r_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(mixed_vcycle(X, l + 1, false));
    {
        EwBatch<F> b;
        b.axpby(PT<F>(N.z_u), F(1), CPT<F>(N.u), F(-1), CPT<F>(N.w_u), (long)nU(N));
        b.axpby(PT<F>(N.z_v), F(1), CPT<F>(N.v), F(-1), CPT<F>(N.w_v), (long)nV(N));
        b.axpby(PT<F>(N.z_H), F(1), CPT<F>(N.H), F(-1), CPT<F>(N.w_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    HIP_TRY(launch_prolong_vfacet_bilinear<F>(CPT<F>(N.z_u), PT<F>(L.z_u), L.ny, L.nx, st), "prolong z_u");
    if (X.cfg.compat_vprolong)
        HIP_TRY(launch_prolong_vfacet_bilinear<F>(CPT<F>(N.z_v), PT<F>(L.z_v), L.ny, L.nx, st), "prolong z_v");
    else
        HIP_TRY(launch_prolong_hfacet_bilinear<F>(CPT<F>(N.z_v), PT<F>(L.z_v), L.ny, L.nx, st), "prolong z_v");
    HIP_TRY(launch_prolong_cell<F>(CPT<F>(N.z_H), PT<F>(L.z_H), L.ny, L.nx, 0, st), "prolong z_H");
    {
        EwBatch<F> b;
        b.correct(PT<F>(L.z_u), CPT<F>(L.y_u), PT<F>(L.u), CPT<F>(L.w_u), (long)nU(L));
        b.correct(PT<F>(L.z_v), CPT<F>(L.y_v), PT<F>(L.v), CPT<F>(L.w_v), (long)nV(L));
        b.correct(PT<F>(L.z_H), CPT<F>(L.y_H), PT<F>(L.H), CPT<F>(L.w_H), (long)nC(L));
        b.axpby(PT<F>(L.gamma), F(1), CPT<F>(L.w_H), F(1), CPT<F>(L.chi), (long)nC(L));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(smooth(L, X.cfg.post_steps, sflags));
    {
        EwBatch<F> b;
        b.fill(PT<F>(L.gamma), thk, (long)nC(L));
        GLIDE_TRY(run(b));
    }
    if (finest) GLIDE_TRY(smooth(L, X.cfg.finest_steps, 0));
    return GLIDE_OK;
}

int check_level32(const glide_level& L, bool adjoint) {
    REQ(L.u); REQ(L.v); REQ(L.H); REQ(L.phi); REQ(L.mask); REQ(L.bed); REQ(L.B); REQ(L.beta); REQ(L.work);
    if (!adjoint) {
        REQ(L.f_u); REQ(L.f_v); REQ(L.f_H); REQ(L.r_u); REQ(L.r_v); REQ(L.r_H); REQ(L.F_u); REQ(L.F_v); REQ(L.F_H);
        REQ(L.gamma); REQ(L.w_u); REQ(L.w_v); REQ(L.w_H); REQ(L.y_u); REQ(L.y_v); REQ(L.y_H); REQ(L.z_u); REQ(L.z_v);
        REQ(L.z_H); REQ(L.chi); REQ(L.cphi);
    } else {
        REQ(L.lambda_u); REQ(L.lambda_v); REQ(L.lambda_H); REQ(L.af_u); REQ(L.af_v); REQ(L.af_H); REQ(L.ar_u);
        REQ(L.ar_v); REQ(L.ar_H); REQ(L.vjp_u); REQ(L.vjp_v); REQ(L.vjp_H); REQ(L.wl_u); REQ(L.wl_v); REQ(L.wl_H);
        REQ(L.zl_u); REQ(L.zl_v); REQ(L.zl_H); REQ(L.adj_cache);
    }
    return GLIDE_OK;
}

int check_mixed(const glide_level* levels64, const glide_level* levels32, int n_levels, int start,
                const glide_fas_config* cfg, const glide_vanka_config* vc, const glide_mixed_config* mc,
                bool adjoint) {
    GLIDE_TRY(check_hierarchy(levels64, n_levels, start, cfg, vc));
    if (!levels32 || !mc) return fail(GLIDE_ERR_INVALID_ARG, "levels32 and the mixed config must be non-NULL");
    if (mc->max_outer < 0 || mc->max_outer > 64) return fail(GLIDE_ERR_INVALID_ARG, "max_outer must be in [0, 64]");
    if (mc->inner_vcycles < 1) return fail(GLIDE_ERR_INVALID_ARG, "inner_vcycles must be >= 1");
    if (mc->krylov_dim < 0 || mc->krylov_dim > 16) return fail(GLIDE_ERR_INVALID_ARG, "krylov_dim must be in [0, 16]");
    if (mc->fp64_newton_steps < 0 || mc->fp64_newton_steps > 64 || !(mc->fp64_relaxation >= 0.0))
        return fail(GLIDE_ERR_INVALID_ARG, "fp64_newton_steps must be in [0, 64] and fp64_relaxation >= 0");
    if (mc->krylov_dim > 0 && !mc->krylov_work) return fail(GLIDE_ERR_INVALID_ARG, "krylov_dim > 0 needs krylov_work");
    for (int l = start; l < n_levels; ++l) {
        const glide_level& A = levels32[l];
        if (A.ny != levels64[l].ny || A.nx != levels64[l].nx)
            return fail(GLIDE_ERR_SHAPE, "the FP32 and FP64 hierarchies must have the same level shapes");
        GLIDE_TRY(check_level32(A, adjoint));
        if (!vc[l].multicolour) return fail(GLIDE_ERR_INVALID_ARG, "the mixed form runs multicolour smoothing only");
        if ((long)A.ny * A.nx < (long)A.ny + A.nx)
            return fail(GLIDE_ERR_SHAPE, "multicolour smoothing needs ny * nx >= ny + nx on every level");
    }
    return GLIDE_OK;
}

// The reference form's smoother and cycle counts (glide/operators.py:314-327,
// glide/multigrid.py:671-682): the fallback when a mixed outer iteration fails.
std::vector<glide_vanka_config> reference_vanka_configs(int n) {
    glide_vanka_config v{};
    v.omega = 0.5; v.relax_phi = 0.0; v.newton_steps = 30; v.relaxation = 0.5;
    v.ssa_damping = 0.01; v.mc_damping = 1.0; v.local_fp32 = 0; v.multicolour = 0; v.newton_tol = 1e-6;
    return std::vector<glide_vanka_config>((size_t)n, v);
}

// The outer iteration of both mixed solves: `resid` enqueues the FP64
// residual norm into a slot, `step(stream, phase)` one outer correction
// (phase 1; with two_phase, phase 2 is an FP64 V-cycle of the same form), `undo`
// restores the state saved by `step`, `fallback` runs one reference-form
// cycle.  A phase-1 iteration that reduces the norm by less than a factor
// kPhaseSwitch, or fails, moves the loop to phase 2 (a failed iteration is
// undone and redone in phase 2); an iteration that fails in the last phase is
// undone and replaced by the fallback.  stats (optional): fallbacks, phase-1
// iterations, phase-2 iterations.
constexpr double kPhaseSwitch = 0.6;

// accel(stream, phase, norms3) runs after every accepted iteration of the
// given phase (0 after a fallback): it may move the state and then replaces
// norms3 by the squared norms of its new residual.
template <typename Resid, typename Step, typename Undo, typename Fallback, typename Key, typename Prep,
          typename Accel>
int mixed_outer_loop(const glide_mixed_config& mc, hipStream_t st, Key&& key, bool two_phase,
                     glide_solve_report* rep, int* stats, Resid&& resid, Step&& step, Undo&& undo,
                     Fallback&& fallback, Prep&& prep, Accel&& accel) {
    if (!rep && !(mc.relative_tolerance > 0.0) && !(mc.absolute_tolerance > 0.0)) {
        // fixed count without a report: max_outer FP32 iterations enqueued
        // without a host wait (prep forms what an iteration reads from the
        // last residual); no norms, no undo or fallback
        auto body = [&](hipStream_t cs) -> int {
            for (int it = 0; it < mc.max_outer; ++it) {
                GLIDE_TRY(prep(cs));
                GLIDE_TRY(step(cs, 1));
            }
            return GLIDE_OK;
        };
        if (g_graphs_on && !timing_on()) GLIDE_TRY(replay(key(0), st, body));
        else GLIDE_TRY(body(st));
        if (stats) {
            stats[0] = 0;
            stats[1] = mc.max_outer;
            stats[2] = 0;
        }
        return GLIDE_OK;
    }
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));
    auto read_norm = [&](double* n3) -> int {
        HIP_TRY(hipMemcpyAsync(s->host, s->norms, sizeof(double) * 3, hipMemcpyDeviceToHost, st), "norms D2H");
        GLIDE_TRY(latch_enqueue(s, st));
        HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
        GLIDE_TRY(latch_check(s));
        n3[0] = s->host[0]; n3[1] = s->host[1]; n3[2] = s->host[2];
        return GLIDE_OK;
    };
    auto attempt = [&](int phase, double* m3) -> int {
        auto body = [&](hipStream_t cs) -> int {
            GLIDE_TRY(step(cs, phase));
            return resid(cs, s->norms);
        };
        if (g_graphs_on && !timing_on()) GLIDE_TRY(replay(key(phase), st, body));
        else GLIDE_TRY(body(st));
        return read_norm(m3);
    };
    double n3[3];
    GLIDE_TRY(resid(st, s->norms));
    GLIDE_TRY(read_norm(n3));
    const double r0 = std::sqrt(n3[0] + n3[1] + n3[2]);
    if (!std::isfinite(r0)) return fail(GLIDE_ERR_NONFINITE, "non-finite residual norm at the start of the mixed solve");
    double r = r0;
    int it = 0, nfall = 0, n1 = 0, n2 = 0;
    int phase = 1;
    const int last_phase = two_phase ? 2 : 1;
    while (it < mc.max_outer && r > mc.absolute_tolerance && (r0 <= 0.0 || r / r0 > mc.relative_tolerance)) {
        double m3[3];
        GLIDE_TRY(attempt(phase, m3));
        (phase == 1 ? n1 : n2)++;
        double rn = std::sqrt(m3[0] + m3[1] + m3[2]);
        if (!(rn < r) && phase < last_phase) {  // failed before the last phase: redo in the next one
            GLIDE_TRY(undo(st));
            phase = last_phase;
            GLIDE_TRY(attempt(phase, m3));
            ++n2;
            rn = std::sqrt(m3[0] + m3[1] + m3[2]);
        }
        if (!(rn < r)) {  // failed in the last phase: undo, one reference-form cycle instead
            GLIDE_TRY(undo(st));
            GLIDE_TRY(fallback(st));
            GLIDE_TRY(resid(st, s->norms));
            GLIDE_TRY(read_norm(m3));
            GLIDE_TRY(accel(st, 0, m3));
            ++nfall;
        } else {
            GLIDE_TRY(accel(st, phase, m3));
            if (phase < last_phase && rn > kPhaseSwitch * r) phase = last_phase;
        }
        std::memcpy(n3, m3, sizeof(n3));
        r = std::sqrt(n3[0] + n3[1] + n3[2]);
        if (!std::isfinite(r)) return fail(GLIDE_ERR_NONFINITE, "non-finite residual norm in the mixed solve");
        if (rep && it < 64) rep->norms[it] = r;
        ++it;
    }
    if (rep) {
        rep->vcycles = it;
        rep->initial_norm = r0;
        rep->final_norm = r;
        rep->norms3[0] = std::sqrt(n3[0]); rep->norms3[1] = std::sqrt(n3[1]); rep->norms3[2] = std::sqrt(n3[2]);
    }
    if (stats) {
        stats[0] = nfall;
        stats[1] = n1;
        stats[2] = n2;
    }
    return GLIDE_OK;
}

// Flexible GMRES(m) outer iteration of the mixed adjoint solve on flat FP64
// vectors of the start level's multipliers (u, v, H one after another): the
// basis V_0..V_m and the preconditioned vectors Z_0..Z_{m-1}, then the
// reduction partials, two sets of inner products and the update coefficients.
struct KrylovLayout {
    long n, nu, nv, nc;
    double *V, *Z, *partial, *dots, *coef;
};
long krylov_vector_len(int ny, int nx) { return (long)ny * (nx + 1) + (long)(ny + 1) * nx + (long)ny * nx; }
int64_t krylov_doubles(int ny, int nx, int m) {
    const long n = krylov_vector_len(ny, nx);
    const long vectors = std::max(2L * m + 1, 2L * std::min(m, 6) + 4);
    return vectors * n + (long)(kKrylovMaxDots + 1) * kKrylovBlocks + 8L * (kKrylovMaxDots + 1) + m;
}
KrylovLayout krylov_layout(void* work, int ny, int nx, int m) {
    KrylovLayout K;
    K.nu = (long)ny * (nx + 1);
    K.nv = (long)(ny + 1) * nx;
    K.nc = (long)ny * nx;
    K.n = K.nu + K.nv + K.nc;
    K.V = static_cast<double*>(work);
    K.Z = K.V + (long)(m + 1) * K.n;
    K.partial = K.Z + (long)m * K.n;
    K.dots = K.partial + (long)(kKrylovMaxDots + 1) * kKrylovBlocks;
    K.coef = K.dots + 2 * (kKrylovMaxDots + 1);
    return K;
}

// Anderson mixing of the mixed forward solve's FP64 cycles in the same
// workspace: depth min(m, kAndersonMax); the differences dF, dG of the last
// depth iterations, f and g of this one and the last, the reduction partials,
// inner products and coefficients.
constexpr int kAndersonMax = 6;
int anderson_depth(int m) { return std::min(m, kAndersonMax); }
struct AndersonLayout {
    long n;
    double *dF, *dG, *f, *g, *f_prev, *g_prev, *partial, *dots, *coef;
};
AndersonLayout anderson_layout(void* work, int ny, int nx, int depth) {
    AndersonLayout W;
    W.n = krylov_vector_len(ny, nx);
    W.dF = static_cast<double*>(work);
    W.dG = W.dF + (long)depth * W.n;
    W.f = W.dG + (long)depth * W.n;
    W.g = W.f + W.n;
    W.f_prev = W.g + W.n;
    W.g_prev = W.f_prev + W.n;
    W.partial = W.g_prev + W.n;
    W.dots = W.partial + (long)(kKrylovMaxDots + 1) * kKrylovBlocks;
    W.coef = W.dots + (long)(kAndersonMax + 1) * (kKrylovMaxDots + 1);
    return W;
}

// Returned by krylov_outer_loop when a restart cycle does not reduce the
// true residual: the caller continues with the defect-correction loop.
constexpr int kKrylovStalled = 1;

// resid(stream, slot) forms ar = J^T lambda - g in ar_{u,v,H} and its squared
// norms; precond(stream, v, z) z = M v (the FP32 correction V-cycles);
// apply(stream, z, w) w = J^T z.  The residual r = g - J^T lambda = -ar.
template <typename Resid, typename Precond, typename Apply>
int krylov_outer_loop(const glide_mixed_config& mc, hipStream_t st, const KrylovLayout& K, int m,
                      const glide_level& D, glide_solve_report* rep, int* stats, int* iterations, Resid&& resid,
                      Precond&& precond, Apply&& apply) {
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));
    double* hk = s->host + 3 * 66;
    auto read_norm = [&](double* n3) -> int {
        HIP_TRY(hipMemcpyAsync(s->host, s->norms, sizeof(double) * 3, hipMemcpyDeviceToHost, st), "norms D2H");
        GLIDE_TRY(latch_enqueue(s, st));
        HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
        GLIDE_TRY(latch_check(s));
        n3[0] = s->host[0]; n3[1] = s->host[1]; n3[2] = s->host[2];
        return GLIDE_OK;
    };
    R* lam[3] = {P(D.lambda_u), P(D.lambda_v), P(D.lambda_H)};
    const R* ar[3] = {CP(D.ar_u), CP(D.ar_v), CP(D.ar_H)};
    const long len[3] = {K.nu, K.nv, K.nc}, off[3] = {0, K.nu, K.nu + K.nv};
    double n3[3];
    GLIDE_TRY(resid(st, s->norms));
    GLIDE_TRY(read_norm(n3));
    const double r0 = std::sqrt(n3[0] + n3[1] + n3[2]);
    if (!std::isfinite(r0)) return fail(GLIDE_ERR_NONFINITE, "non-finite residual norm at the start of the mixed solve");
    auto converged = [&](double r) {
        return !(r > mc.absolute_tolerance && (r0 <= 0.0 || r / r0 > mc.relative_tolerance));
    };
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    auto h = [&](int i, int j) -> double& { return H[(size_t)i * m + j]; };
    double r = r0;
    int it = 0, rc = GLIDE_OK;
    while (it < mc.max_outer && !converged(r)) {
        EwBatch<R> b0;
        for (int q = 0; q < 3; ++q) b0.axpby(K.V + off[q], -1.0 / r, ar[q], 0.0, ar[q], len[q]);
        HIP_TRY(launch_ew_batch<R>(b0, st), "Krylov start vector");
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = r;
        int j = 0;
        while (j < m && it < mc.max_outer) {
            double* Vj = K.V + (long)j * K.n;
            double* Zj = K.Z + (long)j * K.n;
            double* W = Vj + K.n;
            GLIDE_TRY(precond(st, Vj, Zj));
            GLIDE_TRY(apply(st, Zj, W));
            // classical Gram-Schmidt, twice
            double* d1 = K.dots;
            double* d2 = K.dots + (kKrylovMaxDots + 1);
            HIP_TRY(launch_krylov_dots(W, K.V, K.n, K.n, j + 1, K.partial, d1, st), "Krylov inner products");
            HIP_TRY(launch_krylov_axpy(W, K.V, K.n, K.n, j + 1, d1, -1.0, st), "Krylov orthogonalisation");
            HIP_TRY(launch_krylov_dots(W, K.V, K.n, K.n, j + 1, K.partial, d2, st), "Krylov inner products");
            HIP_TRY(launch_krylov_axpy(W, K.V, K.n, K.n, j + 1, d2, -1.0, st), "Krylov orthogonalisation");
            HIP_TRY(hipMemcpyAsync(hk, d1, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, st), "Krylov D2H");
            HIP_TRY(hipMemcpyAsync(hk + (j + 1), d2, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, st),
                    "Krylov D2H");
            GLIDE_TRY(latch_enqueue(s, st));
            HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
            GLIDE_TRY(latch_check(s));
            // |W|^2 after the second pass from the first pass's norm
            double s2 = 0.0;
            for (int i = 0; i <= j; ++i) {
                h(i, j) = hk[i] + hk[j + 1 + i];
                s2 += hk[j + 1 + i] * hk[j + 1 + i];
            }
            const double hn = std::sqrt(std::max(hk[2 * j + 2] - s2, 0.0));
            if (!std::isfinite(hn)) return fail(GLIDE_ERR_NONFINITE, "non-finite Krylov vector in the mixed solve");
            if (hn > 0.0) {
                EwBatch<R> bs;
                bs.axpby(W, 1.0 / hn, W, 0.0, W, K.n);
                HIP_TRY(launch_ew_batch<R>(bs, st), "Krylov normalisation");
            }
            // Givens rotations of the Hessenberg column
            for (int i = 0; i < j; ++i) {
                const double a = h(i, j), c = h(i + 1, j);
                h(i, j) = cs[i] * a + sn[i] * c;
                h(i + 1, j) = -sn[i] * a + cs[i] * c;
            }
            const double rho = std::hypot(h(j, j), hn);
            cs[j] = rho > 0.0 ? h(j, j) / rho : 1.0;
            sn[j] = rho > 0.0 ? hn / rho : 0.0;
            h(j, j) = rho;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            const double est = std::fabs(g[j + 1]);
            if (rep && it < 64) rep->norms[it] = est;
            ++it;
            ++j;
            if (!std::isfinite(est)) return fail(GLIDE_ERR_NONFINITE, "non-finite residual estimate in the mixed solve");
            if (converged(est) || hn == 0.0) break;
        }
        // x += Z y with H y = g (upper triangular, j x j)
        for (int i = j - 1; i >= 0; --i) {
            double t = g[i];
            for (int k = i + 1; k < j; ++k) t -= h(i, k) * y[k];
            y[i] = h(i, i) != 0.0 ? t / h(i, i) : 0.0;
        }
        for (int i = 0; i < j; ++i) hk[i] = y[i];
        HIP_TRY(hipMemcpyAsync(K.coef, hk, sizeof(double) * j, hipMemcpyHostToDevice, st), "Krylov H2D");
        for (int q = 0; q < 3; ++q)
            HIP_TRY(launch_krylov_axpy(lam[q], K.Z + off[q], K.n, len[q], j, K.coef, 1.0, st), "Krylov update");
        GLIDE_TRY(resid(st, s->norms));
        GLIDE_TRY(read_norm(n3));
        const double rn = std::sqrt(n3[0] + n3[1] + n3[2]);
        if (!std::isfinite(rn)) return fail(GLIDE_ERR_NONFINITE, "non-finite residual norm in the mixed solve");
        const bool stalled = !(rn < r);
        r = rn;
        if (stalled) {
            rc = kKrylovStalled;
            break;
        }
    }
    if (rep) {
        rep->vcycles = it;
        rep->initial_norm = r0;
        rep->final_norm = r;
        rep->norms3[0] = std::sqrt(n3[0]); rep->norms3[1] = std::sqrt(n3[1]); rep->norms3[2] = std::sqrt(n3[2]);
    }
    if (stats) {
        stats[0] = 0;
        stats[1] = it;
        stats[2] = 0;
    }
    *iterations = it;
    return rc;
}

std::string mixed_key(char tag, const glide_level* l64, const glide_level* l32, int n_levels, int start, double dt,
                      const glide_fas_config& cfg, const glide_vanka_config* vc, const glide_mixed_config& mc) {
    std::string k(1, tag);
    key_append(k, n_levels);
    key_append(k, start);
    key_append(k, dt);
    key_append(k, cfg);
    key_append(k, mc.inner_vcycles);
    key_append(k, mc.fp64_coarsest);
    key_append(k, mc.max_outer);
    key_append(k, mc.fp64_newton_steps);
    key_append(k, mc.fp64_relaxation);
    key_append(k, g_spin_limit);
    key_append(k, g_persistent_fenced);
    for (int l = start; l < n_levels; ++l) {
        key_append(k, l64[l]);
        key_append(k, l32[l]);
        key_append(k, vc[l]);
    }
    return k;
}


// Adjoint side of the mixed form: the multicolour adjoint passes on a level of
// element type T (linear drag), with the level's cache prepared for its state.
template <typename T>
VankaArgs<T> adjoint_args_t(const glide_level& L, const glide_vanka_config& vc) {
    VankaArgs<T> vk;
    vk.newton_steps = 1;
    vk.ssa_damping = T(vc.ssa_damping);
    vk.mc_damping = T(vc.mc_damping);
    vk.calving_rate = T(L.phys.calving_rate);
    vk.iters = nullptr;
    vk.local_fp32 = 0;
    vk.relax = T(1);
    vk.tol = T(0);
    return vk;
}

template <typename T>
int adjoint_prepare_t(const glide_level& L, double dt, const glide_vanka_config& vc, hipStream_t st) {
    SsaArgs<T> a = make_args_t<T>(L, dt, true);
    HIP_TRY(launch_adjoint_factor<T>(a, PT<T>(L.adj_cache), adjoint_args_t<T>(L, vc), st), "adjoint factor");
    return GLIDE_OK;
}

template <typename T>
int mc_adjoint_sweeps_t(const glide_level& L, double dt, int n_iter, const glide_vanka_config& vc, hipStream_t st) {
    SsaArgs<T> a = make_args_t<T>(L, dt, true);
    if (!a.P.linear_drag) return fail(GLIDE_ERR_INVALID_ARG, "the multicolour adjoint passes need linear drag");
    T* alt_u = PT<T>(L.work);
    T* alt_v = alt_u + nU(L);
    T* alt_H = alt_v + nV(L);
    const T cr = T(L.phys.calving_rate);
    for (int it = 0; it < n_iter; ++it) {
        HIP_TRY(launch_adjoint_pass_mc<T>(a, CPT<T>(L.adj_cache), CPT<T>(L.lambda_u), CPT<T>(L.lambda_v),
                                          CPT<T>(L.lambda_H), alt_u, alt_v, alt_H, CPT<T>(L.af_u), CPT<T>(L.af_v),
                                          CPT<T>(L.af_H), cr, T(vc.omega), 0, st),
                "adjoint red phase");
        HIP_TRY(launch_adjoint_pass_mc<T>(a, CPT<T>(L.adj_cache), alt_u, alt_v, alt_H, PT<T>(L.lambda_u),
                                          PT<T>(L.lambda_v), PT<T>(L.lambda_H), CPT<T>(L.af_u), CPT<T>(L.af_v),
                                          CPT<T>(L.af_H), cr, T(vc.omega), 1, st),
                "adjoint black phase");
    }
    return GLIDE_OK;
}

template <typename T>
int vjp_t(const glide_level& L, double dt, bool forcing, const void* lu, const void* lv, const void* lH, void* ou,
          void* ov, void* oH, bool cached, hipStream_t st) {
    SsaArgs<T> a = make_args_t<T>(L, dt, true);
    HIP_TRY(launch_vjp<T>(a, CPT<T>(lu), CPT<T>(lv), CPT<T>(lH), forcing ? CPT<T>(L.af_u) : nullptr,
                          forcing ? CPT<T>(L.af_v) : nullptr, forcing ? CPT<T>(L.af_H) : nullptr, PT<T>(ou), PT<T>(ov),
                          PT<T>(oH), T(L.phys.calving_rate),
                          cached ? CPT<T>(L.adj_cache) + (size_t)kAdjointCacheEta * nC(L) : nullptr, st),
            "vjp");
    return GLIDE_OK;
}

// The coarsest level's adjoint passes of an FP32 V-cycle on its FP64 copy.
int coarsest_adjoint_fp64(MixedCtx& X, int l, int n_iter) {
    const glide_level& A = X.L32[l];
    const glide_level& D = X.L64[l];
    MixedBatch in;
    in.add(kMixTo64, D.u, A.u, nullptr, (long)nU(A));
    in.add(kMixTo64, D.v, A.v, nullptr, (long)nV(A));
    in.add(kMixTo64, D.H, A.H, nullptr, (long)nC(A));
    in.add(kMixTo64, D.phi, A.phi, nullptr, (long)nC(A));
    in.add(kMixTo64, D.mask, A.mask, nullptr, (long)nC(A));
    in.add(kMixTo64, D.lambda_u, A.lambda_u, nullptr, (long)nU(A));
    in.add(kMixTo64, D.lambda_v, A.lambda_v, nullptr, (long)nV(A));
    in.add(kMixTo64, D.lambda_H, A.lambda_H, nullptr, (long)nC(A));
    in.add(kMixTo64, D.af_u, A.af_u, nullptr, (long)nU(A));
    in.add(kMixTo64, D.af_v, A.af_v, nullptr, (long)nV(A));
    in.add(kMixTo64, D.af_H, A.af_H, nullptr, (long)nC(A));
    GLIDE_TRY(run_mixed(in, X.st));
    GLIDE_TRY(adjoint_prepare_t<double>(D, X.dt, X.vc[l], X.st));
    GLIDE_TRY(mc_adjoint_sweeps_t<double>(D, X.dt, n_iter, X.vc[l], X.st));
    MixedBatch out;
    out.add(kMixTo32, A.lambda_u, D.lambda_u, nullptr, (long)nU(A));
    out.add(kMixTo32, A.lambda_v, D.lambda_v, nullptr, (long)nV(A));
    out.add(kMixTo32, A.lambda_H, D.lambda_H, nullptr, (long)nC(A));
    return run_mixed(out, X.st);
}

// FASAdjointSolver.vcycle (glide/multigrid.py:905-973) on the FP32 hierarchy
// with the multicolour adjoint passes; the same stages as adj_vcycle.  The
// start level's cache is prepared by the caller.
int mixed_adj_vcycle(MixedCtx& X, int l, bool finest) {
    typedef float F;
    const glide_level& L = X.L32[l];
    hipStream_t st = X.st;
    auto smooth = [&](int n) -> int { return n > 0 ? mc_adjoint_sweeps_t<F>(L, X.dt, n, X.vc[l], st) : GLIDE_OK; };
    if (l == X.n_levels - 1) {
        if (X.fp64_coarsest) return coarsest_adjoint_fp64(X, l, X.cfg.coarsest_steps);
        return smooth(X.cfg.coarsest_steps);
    }
    const glide_level& N = X.L32[l + 1];
    GLIDE_TRY(smooth(X.cfg.pre_steps));
    {
        RestrictBatch<F> rb;
        rb.add(CPT<F>(L.u), PT<F>(N.u), kRestrictU);
        rb.add(CPT<F>(L.v), PT<F>(N.v), kRestrictV);
        rb.add(CPT<F>(L.H), PT<F>(N.H), kRestrictCellMean);
        rb.add(CPT<F>(L.phi), PT<F>(N.phi), kRestrictCellMean);
        rb.add(CPT<F>(L.mask), PT<F>(N.mask), kRestrictCellMax);
        rb.add(CPT<F>(L.lambda_u), PT<F>(N.lambda_u), kRestrictU);
        rb.add(CPT<F>(L.lambda_v), PT<F>(N.lambda_v), kRestrictV);
        rb.add(CPT<F>(L.lambda_H), PT<F>(N.lambda_H), kRestrictCellMean);
        HIP_TRY(launch_restrict_batch<F>(rb, N.ny, N.nx, st), "restrict state and multipliers");
    }
    HIP_TRY(launch_pin_rows<F>(PT<F>(N.lambda_u), PT<F>(N.lambda_v), PT<F>(N.lambda_H), CPT<F>(N.mask), N.ny, N.nx,
                               st),
            "pin rows");
    if (!(X.fp64_coarsest && l + 1 == X.n_levels - 1)) GLIDE_TRY(adjoint_prepare_t<F>(N, X.dt, X.vc[l + 1], st));
    auto run = [&](const EwBatch<F>& b) -> int {
        HIP_TRY(launch_ew_batch<F>(b, st), "element-wise batch");
        return GLIDE_OK;
    };
    {
        EwBatch<F> b;
        b.copy(PT<F>(N.wl_u), CPT<F>(N.lambda_u), (long)nU(N));
        b.copy(PT<F>(N.wl_v), CPT<F>(N.lambda_v), (long)nV(N));
        b.copy(PT<F>(N.wl_H), CPT<F>(N.lambda_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(vjp_t<F>(L, X.dt, true, L.lambda_u, L.lambda_v, L.lambda_H, L.ar_u, L.ar_v, L.ar_H, true, st));
    {
        RestrictBatch<F> rb;
        rb.add(CPT<F>(L.ar_u), PT<F>(N.ar_u), kRestrictU);
        rb.add(CPT<F>(L.ar_v), PT<F>(N.ar_v), kRestrictV);
        rb.add(CPT<F>(L.ar_H), PT<F>(N.ar_H), kRestrictCellMean);
        HIP_TRY(launch_restrict_batch<F>(rb, N.ny, N.nx, st), "restrict adjoint residual");
    }
    const bool n_cached = !(X.fp64_coarsest && l + 1 == X.n_levels - 1);
    GLIDE_TRY(vjp_t<F>(N, X.dt, false, N.lambda_u, N.lambda_v, N.lambda_H, N.vjp_u, N.vjp_v, N.vjp_H, n_cached, st));
    {
        EwBatch<F> b;
        b.axpby(PT<F>(N.af_u), F(1), CPT<F>(N.vjp_u), F(-1), CPT<F>(N.ar_u), (long)nU(N));
        b.axpby(PT<F>(N.af_v), F(1), CPT<F>(N.vjp_v), F(-1), CPT<F>(N.ar_v), (long)nV(N));
        b.axpby(PT<F>(N.af_H), F(1), CPT<F>(N.vjp_H), F(-1), CPT<F>(N.ar_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    GLIDE_TRY(mixed_adj_vcycle(X, l + 1, false));
    {
        EwBatch<F> b;
        b.axpby(PT<F>(N.zl_u), F(1), CPT<F>(N.lambda_u), F(-1), CPT<F>(N.wl_u), (long)nU(N));
        b.axpby(PT<F>(N.zl_v), F(1), CPT<F>(N.lambda_v), F(-1), CPT<F>(N.wl_v), (long)nV(N));
        b.axpby(PT<F>(N.zl_H), F(1), CPT<F>(N.lambda_H), F(-1), CPT<F>(N.wl_H), (long)nC(N));
        GLIDE_TRY(run(b));
    }
    HIP_TRY(launch_prolong_vfacet_bilinear<F>(CPT<F>(N.zl_u), PT<F>(L.zl_u), L.ny, L.nx, st), "prolong zl_u");
    if (X.cfg.compat_vprolong)
        HIP_TRY(launch_prolong_vfacet_bilinear<F>(CPT<F>(N.zl_v), PT<F>(L.zl_v), L.ny, L.nx, st), "prolong zl_v");
    else
        HIP_TRY(launch_prolong_hfacet_bilinear<F>(CPT<F>(N.zl_v), PT<F>(L.zl_v), L.ny, L.nx, st), "prolong zl_v");
    HIP_TRY(launch_prolong_cell<F>(CPT<F>(N.zl_H), PT<F>(L.zl_H), L.ny, L.nx, 1, st), "prolong zl_H");
    {
        EwBatch<F> b;
        b.axpby(PT<F>(L.lambda_u), F(1), CPT<F>(L.lambda_u), F(1), CPT<F>(L.zl_u), (long)nU(L));
        b.axpby(PT<F>(L.lambda_v), F(1), CPT<F>(L.lambda_v), F(1), CPT<F>(L.zl_v), (long)nV(L));
        b.axpby(PT<F>(L.lambda_H), F(1), CPT<F>(L.lambda_H), F(1), CPT<F>(L.zl_H), (long)nC(L));
        GLIDE_TRY(run(b));
    }
    HIP_TRY(launch_pin_rows<F>(PT<F>(L.lambda_u), PT<F>(L.lambda_v), PT<F>(L.lambda_H), CPT<F>(L.mask), L.ny, L.nx,
                               st),
            "pin rows");
    GLIDE_TRY(smooth(X.cfg.post_steps));
    if (finest) GLIDE_TRY(smooth(X.cfg.finest_steps));
    return GLIDE_OK;
}


// ====================================================== batched ensemble step
// An ensemble of independent domains of uneven sizes (no counterpart in the
// reference, which runs one domain per process: glide/model.py:25-39) stepped
// by domain-table launches (tab.hpp): every stage of the mixed forward solve,
// of the mixed adjoint solve and of the gradients is one launch over all the
// domains that take part in it.  The V-cycles run level-synchronously: the
// descent of level l covers every domain deeper than l + 1, then the
// coarsest-level passes of all domains (each at its own coarsest level) run
// together, then the ascents from the deepest level up.  Each domain sees
// exactly the stages, in the same order and with the same arithmetic, of its
// own fixed-count asynchronous mixed solves (glide_fascd_solve_mixed and
// glide_fas_adjoint_solve_mixed with both tolerances zero and no report);
// only the independent work of other domains is interleaved.  The launch
// sequence is recorded once per time step length into a device-resident
// table arena and replayed as one HIP graph.

enum TabKind {
    kKEw, kKXfer, kKRes32, kKRes64, kKEta32, kKEta64, kKMc32, kKMc64, kKFac32, kKFac64, kKPass32, kKPass64,
    kKVjp32, kKVjp64, kKGrad64
};

struct TabOp {
    int kind;
    size_t jobs, start;  // byte offsets in the arena
    int n, blocks, colour, linear, rows;
};

struct TabProgram {
    std::vector<unsigned char> arena;
    std::vector<TabOp> ops;
    size_t put(const void* p, size_t bytes, size_t align) {
        size_t off = (arena.size() + align - 1) / align * align;
        arena.resize(off + bytes);
        std::memcpy(arena.data() + off, p, bytes);
        return off;
    }
    template <typename Job>
    void add(int kind, const std::vector<Job>& jobs, const std::vector<int>& blocks, int colour = 0,
             int linear = 0, int rows = 0) {
        if (jobs.empty()) return;
        TabOp op{};
        op.kind = kind;
        op.n = (int)jobs.size();
        op.colour = colour;
        op.linear = linear;
        op.rows = rows;
        op.jobs = put(jobs.data(), sizeof(Job) * jobs.size(), 64);
        std::vector<int> start(jobs.size());
        int acc = 0;
        for (size_t e = 0; e < jobs.size(); ++e) {
            start[e] = acc;
            acc += blocks[e];
        }
        op.start = put(start.data(), sizeof(int) * start.size(), 16);
        op.blocks = acc;
        ops.push_back(op);
    }
};

// Element-wise jobs of one stage.
struct EwList {
    std::vector<TabEw> j;
    std::vector<int> b;
    void add(int op, void* out, const void* x, const void* y, const void* z, double a, double bb, double c, long n,
             void* out2 = nullptr) {
        if (n <= 0) return;
        j.push_back(TabEw{out, out2, x, y, z, a, bb, c, n, op});
        b.push_back(tab_blocks_ew(n));
    }
    void axpby32(void* o, float a, const void* x, float bb, const void* y, size_t n) {
        add(kTabAxpby32, o, x, y, nullptr, a, bb, 0.0, (long)n);
    }
    void copy32(void* o, const void* x, size_t n) { add(kTabCopy32, o, x, nullptr, nullptr, 0, 0, 0, (long)n); }
    void fill32(void* o, float v, size_t n) { add(kTabFill32, o, nullptr, nullptr, nullptr, v, 0, 0, (long)n); }
    // z = z + y; x = w + z
    void correct32(void* z, const void* y, void* x, const void* w, size_t n) {
        add(kTabCorrect32, z, z, y, w, 1, 1, 1, (long)n, x);
    }
    void axpby64(void* o, double a, const void* x, double bb, const void* y, size_t n) {
        add(kTabAxpby64, o, x, y, nullptr, a, bb, 0.0, (long)n);
    }
    void copy64(void* o, const void* x, size_t n) { add(kTabCopy64, o, x, nullptr, nullptr, 0, 0, 0, (long)n); }
    void fill64(void* o, double v, size_t n) { add(kTabFill64, o, nullptr, nullptr, nullptr, v, 0, 0, (long)n); }
    void mix(int op, void* o, const void* x, const void* y, size_t n, double val = 0.0) {
        add(op, o, x, y, nullptr, val, 0, 0, (long)n);
    }
};

struct XferList {
    std::vector<TabXfer> j;
    std::vector<int> b;
    // ny, nx: the coarse level (restrictions) or the fine level (prolongations)
    void add(int op, const void* src, void* dst, int ny, int nx) {
        TabXfer q{static_cast<const float*>(src), static_cast<float*>(dst), nullptr, nullptr, nullptr, op, ny, nx, 0};
        b.push_back(tab_blocks_xfer(op, ny, nx, &q.gx));
        j.push_back(q);
    }
    void pin(const glide_level& L) {
        TabXfer q{nullptr, PT<float>(L.lambda_u), PT<float>(L.lambda_v), PT<float>(L.lambda_H), CPT<float>(L.mask),
                  kTabPinRows, L.ny, L.nx, 0};
        b.push_back(tab_blocks_xfer(kTabPinRows, L.ny, L.nx, &q.gx));
        j.push_back(q);
    }
};

struct EnsDomain {
    std::vector<glide_level> L64, L32;
    int nl;
};

struct EnsRef {  // one domain's level in a stage, with the stage flags of that level
    int d, l, flags;
};

struct EnsRec {
    const std::vector<EnsDomain>& D;
    const glide_fas_config& cfg;   // forward V-cycles
    const glide_fas_config& acfg;  // adjoint V-cycles
    const glide_vanka_config& vf;
    const glide_vanka_config& va;
    bool fp64_coarsest;
    int inner;
    double dt;
    TabProgram& prog;
    int lmax;
    const std::vector<int>& rows;  // red-black phase tile (rows per block) of level l's stages; [lmax]: coarsest

    const glide_level& l64(int d, int l) const { return D[d].L64[l]; }
    const glide_level& l32(int d, int l) const { return D[d].L32[l]; }
    int sflags(int l) const {
        if (l == 0) return 0;
        return (cfg.freeze_coarse_calving ? GLIDE_FREEZE_CALVING : 0) | (cfg.freeze_coarse_phi ? GLIDE_FREEZE_PHI : 0);
    }
    std::vector<int> deeper_than(int l) const {  // domains with a level below l
        std::vector<int> s;
        for (int d = 0; d < (int)D.size(); ++d)
            if (D[d].nl > l + 1) s.push_back(d);
        return s;
    }
    void ew(const EwList& e) { prog.add(kKEw, e.j, e.b); }
    void xfer(const XferList& x) { prog.add(kKXfer, x.j, x.b); }

    template <typename T>
    int residual(const std::vector<const glide_level*>& Ls, const std::vector<int>& flags, const std::vector<void*>& out,
                 int stride) {
        std::vector<TabResidual<T>> jobs;
        std::vector<int> blocks;
        for (size_t q = 0; q < Ls.size(); ++q) {
            const glide_level& L = *Ls[q];
            TabResidual<T> J{};
            J.a = make_args_t<T>(L, dt, false);
            J.ou = PT<T>(out[q * stride + 0]);
            J.ov = PT<T>(out[q * stride + 1]);
            J.oH = PT<T>(out[q * stride + 2]);
            J.use_forcing = (flags[q] & GLIDE_USE_FORCING) ? 1 : 0;
            J.use_mask = (flags[q] & GLIDE_USE_MASK) ? 1 : 0;
            J.cr = T((flags[q] & GLIDE_FREEZE_CALVING) ? 0.0 : L.phys.calving_rate);
            blocks.push_back(tab_blocks_residual(L.ny, L.nx, &J.gx));
            jobs.push_back(J);
        }
        prog.add(sizeof(T) == 4 ? kKRes32 : kKRes64, jobs, blocks);
        return GLIDE_OK;
    }

    // Multicolour passes (mc_sweeps_t) on the given levels of type T.
    template <typename T>
    int mc_sweeps(const std::vector<const glide_level*>& Ls, const std::vector<int>& flags, int n_iter, int slot) {
        if (Ls.empty() || n_iter <= 0) return GLIDE_OK;
        const int linear = make_phys(Ls[0]->phys).linear_drag;
        const int tile = rows[slot];
        std::vector<TabEta<T>> eta;
        std::vector<TabVankaMc<T>> red, black;
        std::vector<int> be, bm;
        EwList g;
        for (size_t q = 0; q < Ls.size(); ++q) {
            const glide_level& L = *Ls[q];
            if (make_phys(L.phys).linear_drag != linear)
                return fail(GLIDE_ERR_INVALID_ARG, "the ensemble's domains must share the drag law");
            VankaArgs<T> vk;
            vk.newton_steps = vf.newton_steps;
            vk.ssa_damping = T(vf.ssa_damping);
            vk.mc_damping = T(vf.mc_damping);
            vk.calving_rate = T((flags[q] & GLIDE_FREEZE_CALVING) ? 0.0 : L.phys.calving_rate);
            vk.iters = nullptr;
            vk.local_fp32 = 0;
            vk.relax = T(vf.relaxation);
            vk.tol = T(vf.newton_tol);
            T* alt_u = PT<T>(L.work);
            T* alt_v = alt_u + nU(L);
            T* alt_H = alt_v + nV(L);
            T* e = alt_H + nC(L);
            if (!(flags[q] & GLIDE_FREEZE_PHI))
                g.add(sizeof(T) == 4 ? kTabGrounded32 : kTabGrounded64, L.phi, L.H, L.bed, nullptr,
                      (double)T(L.phys.flotation_reg_driving), (double)T(vf.relax_phi), 0.0, (long)nC(L));
            TabEta<T> E{};
            E.a = make_args_t<T>(L, dt, false);
            E.eta = e;
            be.push_back(tab_blocks_eta(L.ny, L.nx, &E.gx));
            eta.push_back(E);
            TabVankaMc<T> R{};
            R.a = E.a;
            R.eta = e;
            R.u_dst = alt_u; R.v_dst = alt_v; R.H_dst = alt_H;
            R.mask = PT<T>(L.mask);
            R.vk = vk;
            bm.push_back(tab_blocks_vanka_mc(L.ny, L.nx, tile, &R.gx));
            red.push_back(R);
            TabVankaMc<T> B = R;
            B.a.u = alt_u; B.a.v = alt_v; B.a.H = alt_H;
            B.u_dst = PT<T>(L.u); B.v_dst = PT<T>(L.v); B.H_dst = PT<T>(L.H);
            black.push_back(B);
        }
        const bool f32 = sizeof(T) == 4;
        for (int it = 0; it < n_iter; ++it) {
            ew(g);
            prog.add(f32 ? kKEta32 : kKEta64, eta, be);
            prog.add(f32 ? kKMc32 : kKMc64, red, bm, 0, linear, tile);
            prog.add(f32 ? kKMc32 : kKMc64, black, bm, 1, linear, tile);
        }
        return GLIDE_OK;
    }

    int smooth32(const std::vector<int>& S, int l, int n, int flags) {
        std::vector<const glide_level*> Ls;
        for (int d : S) Ls.push_back(&l32(d, l));
        return mc_sweeps<float>(Ls, std::vector<int>(Ls.size(), flags), n, l);
    }

    // ------------------------------------------------ forward (mixed_vcycle)
    int fwd_down(int l) {
        typedef float F;
        const std::vector<int> S = deeper_than(l);
        if (S.empty()) return GLIDE_OK;
        const int sf = sflags(l);
        {
            XferList x;
            for (int d : S) x.add(kTabRestrictMax, l32(d, l).chi, l32(d, l + 1).chi, l32(d, l + 1).ny, l32(d, l + 1).nx);
            xfer(x);
            XferList y;
            for (int d : S) y.add(kTabProlongInject, l32(d, l + 1).chi, l32(d, l).cphi, l32(d, l).ny, l32(d, l).nx);
            xfer(y);
        }
        {
            EwList a, b;
            for (int d : S) {
                const glide_level& L = l32(d, l);
                a.axpby32(L.cphi, F(-1), L.cphi, F(1), L.chi, nC(L));
                b.axpby32(L.gamma, F(1), L.w_H, F(1), L.cphi, nC(L));
            }
            ew(a);
            ew(b);
        }
        GLIDE_TRY(smooth32(S, l, cfg.pre_steps, sf));
        {
            EwList a;
            for (int d : S) {
                const glide_level& L = l32(d, l);
                a.fill32(L.gamma, F(L.phys.thklim), nC(L));
                a.axpby32(L.y_u, F(1), L.u, F(-1), L.w_u, nU(L));
                a.axpby32(L.y_v, F(1), L.v, F(-1), L.w_v, nV(L));
                a.axpby32(L.y_H, F(1), L.H, F(-1), L.w_H, nC(L));
            }
            ew(a);
        }
        {
            XferList x;
            for (int d : S) {
                const glide_level &L = l32(d, l), &N = l32(d, l + 1);
                x.add(kTabRestrictU, L.u, N.u, N.ny, N.nx);
                x.add(kTabRestrictV, L.v, N.v, N.ny, N.nx);
                x.add(kTabRestrictMean, L.H, N.H, N.ny, N.nx);
                x.add(kTabRestrictMean, L.phi, N.phi, N.ny, N.nx);
                x.add(kTabRestrictMax, L.mask, N.mask, N.ny, N.nx);
            }
            xfer(x);
        }
        {
            EwList a;
            for (int d : S) {
                const glide_level& N = l32(d, l + 1);
                a.copy32(N.w_u, N.u, nU(N));
                a.copy32(N.w_v, N.v, nV(N));
                a.copy32(N.w_H, N.H, nC(N));
            }
            ew(a);
        }
        {
            std::vector<const glide_level*> Ls;
            std::vector<int> fl;
            std::vector<void*> out;
            for (int d : S) {
                const glide_level& L = l32(d, l);
                Ls.push_back(&L);
                fl.push_back(GLIDE_USE_FORCING | (sf & GLIDE_FREEZE_CALVING));
                out.insert(out.end(), {L.r_u, L.r_v, L.r_H});
            }
            GLIDE_TRY(residual<F>(Ls, fl, out, 3));
        }
        {
            XferList x;
            for (int d : S) {
                const glide_level &L = l32(d, l), &N = l32(d, l + 1);
                x.add(kTabRestrictU, L.r_u, N.r_u, N.ny, N.nx);
                x.add(kTabRestrictV, L.r_v, N.r_v, N.ny, N.nx);
                x.add(kTabRestrictMean, L.r_H, N.r_H, N.ny, N.nx);
            }
            xfer(x);
        }
        if (!cfg.freeze_coarse_phi) {
            EwList g;
            for (int d : S) {
                const glide_level& N = l32(d, l + 1);
                g.add(kTabGrounded32, N.phi, N.H, N.bed, nullptr, (double)F(N.phys.flotation_reg_driving), 0.0, 0.0,
                      (long)nC(N));
            }
            ew(g);
        }
        {
            std::vector<const glide_level*> Ls;
            std::vector<int> fl;
            std::vector<void*> out;
            for (int d : S) {
                const glide_level& N = l32(d, l + 1);
                Ls.push_back(&N);
                fl.push_back(cfg.freeze_coarse_calving ? GLIDE_FREEZE_CALVING : 0);
                out.insert(out.end(), {N.F_u, N.F_v, N.F_H});
            }
            GLIDE_TRY(residual<F>(Ls, fl, out, 3));
        }
        EwList a;
        for (int d : S) {
            const glide_level& N = l32(d, l + 1);
            a.axpby32(N.f_u, F(1), N.F_u, F(-1), N.r_u, nU(N));
            a.axpby32(N.f_v, F(1), N.F_v, F(-1), N.r_v, nV(N));
            a.axpby32(N.f_H, F(1), N.F_H, F(-1), N.r_H, nC(N));
        }
        ew(a);
        return GLIDE_OK;
    }

    int fwd_coarsest() {
        typedef float F;
        EwList a;
        for (size_t d = 0; d < D.size(); ++d) {
            const glide_level& L = l32((int)d, D[d].nl - 1);
            a.axpby32(L.gamma, F(1), L.w_H, F(1), L.chi, nC(L));
        }
        ew(a);
        if (fp64_coarsest) {
            EwList in, out;
            std::vector<const glide_level*> Ls;
            std::vector<int> fl;
            for (size_t d = 0; d < D.size(); ++d) {
                const int c = D[d].nl - 1;
                const glide_level &A = l32((int)d, c), &Dc = l64((int)d, c);
                in.mix(kTabTo64, Dc.u, A.u, nullptr, nU(A));
                in.mix(kTabTo64, Dc.v, A.v, nullptr, nV(A));
                in.mix(kTabTo64, Dc.H, A.H, nullptr, nC(A));
                in.mix(kTabTo64, Dc.mask, A.mask, nullptr, nC(A));
                in.mix(kTabTo64, Dc.phi, A.phi, nullptr, nC(A));
                in.mix(kTabTo64, Dc.f_u, A.f_u, nullptr, nU(A));
                in.mix(kTabTo64, Dc.f_v, A.f_v, nullptr, nV(A));
                in.mix(kTabTo64, Dc.f_H, A.f_H, nullptr, nC(A));
                in.mix(kTabTo64, Dc.gamma, A.gamma, nullptr, nC(A));
                out.mix(kTabTo32, A.u, Dc.u, nullptr, nU(A));
                out.mix(kTabTo32, A.v, Dc.v, nullptr, nV(A));
                out.mix(kTabTo32, A.H, Dc.H, nullptr, nC(A));
                out.mix(kTabTo32, A.mask, Dc.mask, nullptr, nC(A));
                Ls.push_back(&Dc);
                fl.push_back(sflags(c));
            }
            ew(in);
            GLIDE_TRY(mc_sweeps<double>(Ls, fl, cfg.coarsest_steps, lmax));
            ew(out);
        } else {
            std::vector<const glide_level*> Ls;
            std::vector<int> fl;
            for (size_t d = 0; d < D.size(); ++d) {
                const int c = D[d].nl - 1;
                Ls.push_back(&l32((int)d, c));
                fl.push_back(sflags(c));
            }
            GLIDE_TRY(mc_sweeps<float>(Ls, fl, cfg.coarsest_steps, lmax));
        }
        EwList b;
        for (size_t d = 0; d < D.size(); ++d) {
            const glide_level& L = l32((int)d, D[d].nl - 1);
            b.fill32(L.gamma, F(L.phys.thklim), nC(L));
        }
        ew(b);
        return GLIDE_OK;
    }

    int fwd_up(int l) {
        typedef float F;
        const std::vector<int> S = deeper_than(l);
        if (S.empty()) return GLIDE_OK;
        {
            EwList a;
            for (int d : S) {
                const glide_level& N = l32(d, l + 1);
                a.axpby32(N.z_u, F(1), N.u, F(-1), N.w_u, nU(N));
                a.axpby32(N.z_v, F(1), N.v, F(-1), N.w_v, nV(N));
                a.axpby32(N.z_H, F(1), N.H, F(-1), N.w_H, nC(N));
            }
            ew(a);
        }
        {
            XferList x;
            for (int d : S) {
                const glide_level &L = l32(d, l), &N = l32(d, l + 1);
                x.add(kTabProlongVfacet, N.z_u, L.z_u, L.ny, L.nx);
                x.add(cfg.compat_vprolong ? kTabProlongVfacet : kTabProlongHfacet, N.z_v, L.z_v, L.ny, L.nx);
                x.add(kTabProlongInject, N.z_H, L.z_H, L.ny, L.nx);
            }
            xfer(x);
        }
        {
            EwList a;
            for (int d : S) {
                const glide_level& L = l32(d, l);
                a.correct32(L.z_u, L.y_u, L.u, L.w_u, nU(L));
                a.correct32(L.z_v, L.y_v, L.v, L.w_v, nV(L));
                a.correct32(L.z_H, L.y_H, L.H, L.w_H, nC(L));
                a.axpby32(L.gamma, F(1), L.w_H, F(1), L.chi, nC(L));
            }
            ew(a);
        }
        GLIDE_TRY(smooth32(S, l, cfg.post_steps, sflags(l)));
        {
            EwList a;
            for (int d : S) a.fill32(l32(d, l).gamma, F(l32(d, l).phys.thklim), nC(l32(d, l)));
            ew(a);
        }
        if (l == 0) GLIDE_TRY(smooth32(S, 0, cfg.finest_steps, 0));
        return GLIDE_OK;
    }

    int fwd_vcycle() {
        typedef float F;
        EwList a, b;
        for (size_t d = 0; d < D.size(); ++d) {
            const glide_level& L = l32((int)d, 0);
            a.copy32(L.w_u, L.u, nU(L));
            a.copy32(L.w_v, L.v, nV(L));
            a.copy32(L.w_H, L.H, nC(L));
            a.axpby32(L.chi, F(-1), L.H, F(0), L.H, nC(L));
            a.fill32(L.gamma, F(L.phys.thklim), nC(L));
            b.axpby32(L.chi, F(1), L.chi, F(1), L.gamma, nC(L));
        }
        ew(a);
        ew(b);
        for (int l = 0; l + 1 < lmax; ++l) GLIDE_TRY(fwd_down(l));
        GLIDE_TRY(fwd_coarsest());
        for (int l = lmax - 2; l >= 0; --l) GLIDE_TRY(fwd_up(l));
        return GLIDE_OK;
    }

    // glide_fascd_solve_mixed, fixed count without a report
    int forward(int iterations) {
        {
            EwList a;
            for (size_t d = 0; d < D.size(); ++d) {
                const glide_level& L = l64((int)d, 0);
                a.axpby64(L.f_H, 1.0 / dt, L.H_prev, 1.0, L.smb, nC(L));
                a.fill64(L.gamma, L.phys.thklim, nC(L));
                for (int l = 0; l < D[d].nl; ++l) {
                    const glide_level &E = l64((int)d, l), &G = l32((int)d, l);
                    a.mix(kTabTo32, G.bed, E.bed, nullptr, nC(E));
                    a.mix(kTabTo32, G.B, E.B, nullptr, nC(E));
                    a.mix(kTabTo32, G.beta, E.beta, nullptr, nC(E));
                }
            }
            ew(a);
        }
        for (int it = 0; it < iterations; ++it) {
            EwList g, to32, defect, upd;
            std::vector<const glide_level*> L64s, L32s;
            std::vector<void*> out64, out32;
            for (size_t d = 0; d < D.size(); ++d) {
                const glide_level &Dd = l64((int)d, 0), &A = l32((int)d, 0);
                g.add(kTabGrounded64, Dd.phi, Dd.H, Dd.bed, nullptr, Dd.phys.flotation_reg_driving, 0.0, 0.0,
                      (long)nC(Dd));
                L64s.push_back(&Dd);
                out64.insert(out64.end(), {Dd.r_u, Dd.r_v, Dd.r_H});
                to32.mix(kTabTo32, A.u, Dd.u, nullptr, nU(Dd));
                to32.mix(kTabTo32, A.v, Dd.v, nullptr, nV(Dd));
                to32.mix(kTabTo32, A.H, Dd.H, nullptr, nC(Dd));
                to32.mix(kTabTo32, A.mask, Dd.mask, nullptr, nC(Dd));
                to32.mix(kTabTo32, A.phi, Dd.phi, nullptr, nC(Dd));
                L32s.push_back(&A);
                out32.insert(out32.end(), {A.F_u, A.F_v, A.F_H});
                defect.mix(kTabDefect, A.f_u, A.F_u, Dd.r_u, nU(Dd));
                defect.mix(kTabDefect, A.f_v, A.F_v, Dd.r_v, nV(Dd));
                defect.mix(kTabDefect, A.f_H, A.F_H, Dd.r_H, nC(Dd));
                upd.mix(kTabUpdate, Dd.u, A.u, nullptr, nU(Dd));
                upd.mix(kTabUpdate, Dd.v, A.v, nullptr, nV(Dd));
                upd.mix(kTabUpdateH, Dd.H, A.H, A.mask, nC(Dd), Dd.phys.thklim);
                upd.mix(kTabTo64, Dd.mask, A.mask, nullptr, nC(Dd));
            }
            ew(g);
            GLIDE_TRY(residual<double>(L64s, std::vector<int>(L64s.size(), GLIDE_USE_FORCING), out64, 3));
            ew(to32);
            GLIDE_TRY(residual<float>(L32s, std::vector<int>(L32s.size(), 0), out32, 3));
            ew(defect);
            for (int q = 0; q < inner; ++q) GLIDE_TRY(fwd_vcycle());
            ew(upd);
        }
        return GLIDE_OK;
    }

    // ------------------------------------------------ adjoint (mixed_adj_vcycle)
    template <typename T>
    int factor(const std::vector<const glide_level*>& Ls) {
        std::vector<TabAdjFactor<T>> jobs;
        std::vector<int> blocks;
        for (const glide_level* p : Ls) {
            TabAdjFactor<T> J{};
            J.a = make_args_t<T>(*p, dt, true);
            J.cache = PT<T>(p->adj_cache);
            J.vk = adjoint_args_t<T>(*p, va);
            blocks.push_back(tab_blocks_adj_factor(p->ny, p->nx, &J.gx));
            jobs.push_back(J);
        }
        prog.add(sizeof(T) == 4 ? kKFac32 : kKFac64, jobs, blocks);
        return GLIDE_OK;
    }

    template <typename T>
    int adj_sweeps(const std::vector<const glide_level*>& Ls, int n_iter) {
        if (Ls.empty() || n_iter <= 0) return GLIDE_OK;
        std::vector<TabAdjPass<T>> red, black;
        std::vector<int> blocks;
        for (const glide_level* p : Ls) {
            const glide_level& L = *p;
            if (!make_phys(L.phys).linear_drag)
                return fail(GLIDE_ERR_INVALID_ARG, "the multicolour adjoint passes need linear drag");
            T* alt_u = PT<T>(L.work);
            T* alt_v = alt_u + nU(L);
            T* alt_H = alt_v + nV(L);
            TabAdjPass<T> R{};
            R.a = make_args_t<T>(L, dt, true);
            R.cache = CPT<T>(L.adj_cache);
            R.lu = CPT<T>(L.lambda_u); R.lv = CPT<T>(L.lambda_v); R.lH = CPT<T>(L.lambda_H);
            R.lu_dst = alt_u; R.lv_dst = alt_v; R.lH_dst = alt_H;
            R.fu = CPT<T>(L.af_u); R.fv = CPT<T>(L.af_v); R.fH = CPT<T>(L.af_H);
            R.cr = T(L.phys.calving_rate);
            R.omega = -T(va.omega);
            blocks.push_back(tab_blocks_adj_pass(L.ny, L.nx, tab_adj_pass_rows(sizeof(T) == 4), &R.gx));
            red.push_back(R);
            TabAdjPass<T> B = R;
            B.lu = alt_u; B.lv = alt_v; B.lH = alt_H;
            B.lu_dst = PT<T>(L.lambda_u); B.lv_dst = PT<T>(L.lambda_v); B.lH_dst = PT<T>(L.lambda_H);
            black.push_back(B);
        }
        const int k = sizeof(T) == 4 ? kKPass32 : kKPass64;
        for (int it = 0; it < n_iter; ++it) {
            prog.add(k, red, blocks, 0);
            prog.add(k, black, blocks, 1);
        }
        return GLIDE_OK;
    }

    template <typename T>
    int vjp_stage(const std::vector<const glide_level*>& Ls, const std::vector<int>& forcing,
                  const std::vector<int>& cached, const std::vector<void*>& in, const std::vector<void*>& out) {
        std::vector<TabVjp<T>> jobs;
        std::vector<int> blocks;
        for (size_t q = 0; q < Ls.size(); ++q) {
            const glide_level& L = *Ls[q];
            TabVjp<T> J{};
            J.a = make_args_t<T>(L, dt, true);
            J.lu = CPT<T>(in[3 * q]); J.lv = CPT<T>(in[3 * q + 1]); J.lH = CPT<T>(in[3 * q + 2]);
            J.fu = forcing[q] ? CPT<T>(L.af_u) : nullptr;
            J.fv = forcing[q] ? CPT<T>(L.af_v) : nullptr;
            J.fH = forcing[q] ? CPT<T>(L.af_H) : nullptr;
            J.ou = PT<T>(out[3 * q]); J.ov = PT<T>(out[3 * q + 1]); J.oH = PT<T>(out[3 * q + 2]);
            J.cr = T(L.phys.calving_rate);
            J.eta_cache = cached[q] ? CPT<T>(L.adj_cache) + (size_t)kAdjointCacheEta * nC(L) : nullptr;
            blocks.push_back(tab_blocks_vjp(L.ny, L.nx, &J.gx));
            jobs.push_back(J);
        }
        prog.add(sizeof(T) == 4 ? kKVjp32 : kKVjp64, jobs, blocks);
        return GLIDE_OK;
    }

    int adj_smooth32(const std::vector<int>& S, int l, int n) {
        std::vector<const glide_level*> Ls;
        for (int d : S) Ls.push_back(&l32(d, l));
        return adj_sweeps<float>(Ls, n);
    }

    int adj_down(int l) {
        typedef float F;
        const std::vector<int> S = deeper_than(l);
        if (S.empty()) return GLIDE_OK;
        GLIDE_TRY(adj_smooth32(S, l, acfg.pre_steps));
        {
            XferList x;
            for (int d : S) {
                const glide_level &L = l32(d, l), &N = l32(d, l + 1);
                x.add(kTabRestrictU, L.u, N.u, N.ny, N.nx);
                x.add(kTabRestrictV, L.v, N.v, N.ny, N.nx);
                x.add(kTabRestrictMean, L.H, N.H, N.ny, N.nx);
                x.add(kTabRestrictMean, L.phi, N.phi, N.ny, N.nx);
                x.add(kTabRestrictMax, L.mask, N.mask, N.ny, N.nx);
                x.add(kTabRestrictU, L.lambda_u, N.lambda_u, N.ny, N.nx);
                x.add(kTabRestrictV, L.lambda_v, N.lambda_v, N.ny, N.nx);
                x.add(kTabRestrictMean, L.lambda_H, N.lambda_H, N.ny, N.nx);
            }
            xfer(x);
            XferList p;
            for (int d : S) p.pin(l32(d, l + 1));
            xfer(p);
        }
        {
            std::vector<const glide_level*> Ls;
            for (int d : S)
                if (!(fp64_coarsest && l + 1 == D[d].nl - 1)) Ls.push_back(&l32(d, l + 1));
            GLIDE_TRY(factor<F>(Ls));
        }
        {
            EwList a;
            for (int d : S) {
                const glide_level& N = l32(d, l + 1);
                a.copy32(N.wl_u, N.lambda_u, nU(N));
                a.copy32(N.wl_v, N.lambda_v, nV(N));
                a.copy32(N.wl_H, N.lambda_H, nC(N));
            }
            ew(a);
        }
        {
            std::vector<const glide_level*> Ls;
            std::vector<void*> in, out;
            for (int d : S) {
                const glide_level& L = l32(d, l);
                Ls.push_back(&L);
                in.insert(in.end(), {L.lambda_u, L.lambda_v, L.lambda_H});
                out.insert(out.end(), {L.ar_u, L.ar_v, L.ar_H});
            }
            GLIDE_TRY(vjp_stage<F>(Ls, std::vector<int>(Ls.size(), 1), std::vector<int>(Ls.size(), 1), in, out));
        }
        {
            XferList x;
            for (int d : S) {
                const glide_level &L = l32(d, l), &N = l32(d, l + 1);
                x.add(kTabRestrictU, L.ar_u, N.ar_u, N.ny, N.nx);
                x.add(kTabRestrictV, L.ar_v, N.ar_v, N.ny, N.nx);
                x.add(kTabRestrictMean, L.ar_H, N.ar_H, N.ny, N.nx);
            }
            xfer(x);
        }
        {
            std::vector<const glide_level*> Ls;
            std::vector<int> cached;
            std::vector<void*> in, out;
            for (int d : S) {
                const glide_level& N = l32(d, l + 1);
                Ls.push_back(&N);
                cached.push_back(!(fp64_coarsest && l + 1 == D[d].nl - 1) ? 1 : 0);
                in.insert(in.end(), {N.lambda_u, N.lambda_v, N.lambda_H});
                out.insert(out.end(), {N.vjp_u, N.vjp_v, N.vjp_H});
            }
            GLIDE_TRY(vjp_stage<F>(Ls, std::vector<int>(Ls.size(), 0), cached, in, out));
        }
        EwList a;
        for (int d : S) {
            const glide_level& N = l32(d, l + 1);
            a.axpby32(N.af_u, F(1), N.vjp_u, F(-1), N.ar_u, nU(N));
            a.axpby32(N.af_v, F(1), N.vjp_v, F(-1), N.ar_v, nV(N));
            a.axpby32(N.af_H, F(1), N.vjp_H, F(-1), N.ar_H, nC(N));
        }
        ew(a);
        return GLIDE_OK;
    }

    int adj_coarsest() {
        if (!fp64_coarsest) {
            std::vector<const glide_level*> Ls;
            for (size_t d = 0; d < D.size(); ++d) Ls.push_back(&l32((int)d, D[d].nl - 1));
            return adj_sweeps<float>(Ls, acfg.coarsest_steps);
        }
        EwList in, out;
        std::vector<const glide_level*> Ls;
        for (size_t d = 0; d < D.size(); ++d) {
            const int c = D[d].nl - 1;
            const glide_level &A = l32((int)d, c), &Dc = l64((int)d, c);
            in.mix(kTabTo64, Dc.u, A.u, nullptr, nU(A));
            in.mix(kTabTo64, Dc.v, A.v, nullptr, nV(A));
            in.mix(kTabTo64, Dc.H, A.H, nullptr, nC(A));
            in.mix(kTabTo64, Dc.phi, A.phi, nullptr, nC(A));
            in.mix(kTabTo64, Dc.mask, A.mask, nullptr, nC(A));
            in.mix(kTabTo64, Dc.lambda_u, A.lambda_u, nullptr, nU(A));
            in.mix(kTabTo64, Dc.lambda_v, A.lambda_v, nullptr, nV(A));
            in.mix(kTabTo64, Dc.lambda_H, A.lambda_H, nullptr, nC(A));
            in.mix(kTabTo64, Dc.af_u, A.af_u, nullptr, nU(A));
            in.mix(kTabTo64, Dc.af_v, A.af_v, nullptr, nV(A));
            in.mix(kTabTo64, Dc.af_H, A.af_H, nullptr, nC(A));
            out.mix(kTabTo32, A.lambda_u, Dc.lambda_u, nullptr, nU(A));
            out.mix(kTabTo32, A.lambda_v, Dc.lambda_v, nullptr, nV(A));
            out.mix(kTabTo32, A.lambda_H, Dc.lambda_H, nullptr, nC(A));
            Ls.push_back(&Dc);
        }
        ew(in);
        GLIDE_TRY(factor<double>(Ls));
        GLIDE_TRY(adj_sweeps<double>(Ls, acfg.coarsest_steps));
        ew(out);
        return GLIDE_OK;
    }

    int adj_up(int l) {
        typedef float F;
        const std::vector<int> S = deeper_than(l);
        if (S.empty()) return GLIDE_OK;
        {
            EwList a;
            for (int d : S) {
                const glide_level& N = l32(d, l + 1);
                a.axpby32(N.zl_u, F(1), N.lambda_u, F(-1), N.wl_u, nU(N));
                a.axpby32(N.zl_v, F(1), N.lambda_v, F(-1), N.wl_v, nV(N));
                a.axpby32(N.zl_H, F(1), N.lambda_H, F(-1), N.wl_H, nC(N));
            }
            ew(a);
        }
        {
            XferList x;
            for (int d : S) {
                const glide_level &L = l32(d, l), &N = l32(d, l + 1);
                x.add(kTabProlongVfacet, N.zl_u, L.zl_u, L.ny, L.nx);
                x.add(acfg.compat_vprolong ? kTabProlongVfacet : kTabProlongHfacet, N.zl_v, L.zl_v, L.ny, L.nx);
                x.add(kTabProlongBilinear, N.zl_H, L.zl_H, L.ny, L.nx);
            }
            xfer(x);
        }
        {
            EwList a;
            for (int d : S) {
                const glide_level& L = l32(d, l);
                a.axpby32(L.lambda_u, F(1), L.lambda_u, F(1), L.zl_u, nU(L));
                a.axpby32(L.lambda_v, F(1), L.lambda_v, F(1), L.zl_v, nV(L));
                a.axpby32(L.lambda_H, F(1), L.lambda_H, F(1), L.zl_H, nC(L));
            }
            ew(a);
            XferList p;
            for (int d : S) p.pin(l32(d, l));
            xfer(p);
        }
        GLIDE_TRY(adj_smooth32(S, l, acfg.post_steps));
        if (l == 0) GLIDE_TRY(adj_smooth32(S, 0, acfg.finest_steps));
        return GLIDE_OK;
    }

    int adj_vcycle() {
        for (int l = 0; l + 1 < lmax; ++l) GLIDE_TRY(adj_down(l));
        GLIDE_TRY(adj_coarsest());
        for (int l = lmax - 2; l >= 0; --l) GLIDE_TRY(adj_up(l));
        return GLIDE_OK;
    }

    // glide_fas_adjoint_solve_mixed, fixed count without a report, then the
    // gradients `which` (glide_gradients)
    int adjoint(int iterations, int which) {
        {
            EwList a;
            std::vector<const glide_level*> A0;
            for (size_t d = 0; d < D.size(); ++d) {
                const glide_level &Dd = l64((int)d, 0), &A = l32((int)d, 0);
                a.mix(kTabTo32, A.u, Dd.u, nullptr, nU(Dd));
                a.mix(kTabTo32, A.v, Dd.v, nullptr, nV(Dd));
                a.mix(kTabTo32, A.H, Dd.H, nullptr, nC(Dd));
                a.mix(kTabTo32, A.phi, Dd.phi, nullptr, nC(Dd));
                a.mix(kTabTo32, A.mask, Dd.mask, nullptr, nC(Dd));
                for (int l = 0; l < D[d].nl; ++l) {
                    const glide_level &E = l64((int)d, l), &G = l32((int)d, l);
                    a.mix(kTabTo32, G.bed, E.bed, nullptr, nC(E));
                    a.mix(kTabTo32, G.B, E.B, nullptr, nC(E));
                    a.mix(kTabTo32, G.beta, E.beta, nullptr, nC(E));
                }
                A0.push_back(&A);
            }
            ew(a);
            GLIDE_TRY(factor<float>(A0));
        }
        for (int it = 0; it < iterations; ++it) {
            std::vector<const glide_level*> L64s;
            std::vector<void*> in, out;
            EwList neg, add;
            XferList pin;
            for (size_t d = 0; d < D.size(); ++d) {
                const glide_level &Dd = l64((int)d, 0), &A = l32((int)d, 0);
                L64s.push_back(&Dd);
                in.insert(in.end(), {Dd.lambda_u, Dd.lambda_v, Dd.lambda_H});
                out.insert(out.end(), {Dd.ar_u, Dd.ar_v, Dd.ar_H});
                neg.mix(kTabNeg32, A.af_u, Dd.ar_u, nullptr, nU(Dd));
                neg.mix(kTabNeg32, A.af_v, Dd.ar_v, nullptr, nV(Dd));
                neg.mix(kTabNeg32, A.af_H, Dd.ar_H, nullptr, nC(Dd));
                neg.fill32(A.lambda_u, 0.0f, nU(A));
                neg.fill32(A.lambda_v, 0.0f, nV(A));
                neg.fill32(A.lambda_H, 0.0f, nC(A));
                pin.pin(A);
                add.mix(kTabAdd64, Dd.lambda_u, A.lambda_u, nullptr, nU(Dd));
                add.mix(kTabAdd64, Dd.lambda_v, A.lambda_v, nullptr, nV(Dd));
                add.mix(kTabAdd64, Dd.lambda_H, A.lambda_H, nullptr, nC(Dd));
            }
            GLIDE_TRY(vjp_stage<double>(L64s, std::vector<int>(L64s.size(), 1), std::vector<int>(L64s.size(), 0), in,
                                        out));
            ew(neg);
            for (int q = 0; q < inner; ++q) GLIDE_TRY(adj_vcycle());
            xfer(pin);
            ew(add);
        }
        if (which & 3) {
            std::vector<TabGrad<double>> jobs;
            std::vector<int> blocks;
            for (size_t d = 0; d < D.size(); ++d) {
                const glide_level& L = l64((int)d, 0);
                TabGrad<double> J{};
                J.a = make_args(L, dt, true);
                J.lu = CP(L.lambda_u);
                J.lv = CP(L.lambda_v);
                J.gbeta = (which & 1) ? P(L.g_beta) : nullptr;
                J.gbed = (which & 2) ? P(L.g_bed) : nullptr;
                J.exact = (which & GLIDE_GRAD_EXACT) ? 1 : 0;
                blocks.push_back(tab_blocks_grad(L.ny, L.nx, &J.gx));
                jobs.push_back(J);
            }
            prog.add(kKGrad64, jobs, blocks);
        }
        EwList g, z;
        for (size_t d = 0; d < D.size(); ++d) {
            const glide_level& L = l64((int)d, 0);
            if (which & 4) {
                g.axpby64(L.g_H_prev, -1.0 / dt, L.lambda_H, 0.0, L.lambda_H, nC(L));
                if (which & GLIDE_GRAD_EXACT) z.add(kTabZeroMasked64, L.g_H_prev, L.mask, nullptr, nullptr, 0, 0, 0, (long)nC(L));
            }
            if (which & 8) {
                g.axpby64(L.g_smb, -1.0, L.lambda_H, 0.0, L.lambda_H, nC(L));
                if (which & GLIDE_GRAD_EXACT) z.add(kTabZeroMasked64, L.g_smb, L.mask, nullptr, nullptr, 0, 0, 0, (long)nC(L));
            }
        }
        ew(g);
        ew(z);
        return GLIDE_OK;
    }

    // The ensemble step: forward solve, H_prev <- H, adjoint seeds
    // dJ/du = u, dJ/dv = v (forcing -u, -v, 0), adjoint solve and gradients.
    int step(int iterations, int which) {
        GLIDE_TRY(forward(iterations));
        EwList a;
        for (size_t d = 0; d < D.size(); ++d) {
            const glide_level& L = l64((int)d, 0);
            a.copy64(L.H_prev, L.H, nC(L));
            a.axpby64(L.af_u, -1.0, L.u, 0.0, L.u, nU(L));
            a.axpby64(L.af_v, -1.0, L.v, 0.0, L.v, nV(L));
            a.fill64(L.af_H, 0.0, nC(L));
        }
        ew(a);
        return adjoint(iterations, which);
    }
};

int run_tab_op(const TabOp& op, const unsigned char* dev, hipStream_t st) {
    const int* start = reinterpret_cast<const int*>(dev + op.start);
    const void* jobs = dev + op.jobs;
    hipError_t e = hipSuccess;
    switch (op.kind) {
        case kKEw: e = launch_tab_ew(static_cast<const TabEw*>(jobs), start, op.n, op.blocks, st); break;
        case kKXfer: e = launch_tab_xfer(static_cast<const TabXfer*>(jobs), start, op.n, op.blocks, st); break;
        case kKRes32:
            e = launch_tab_residual<float>(static_cast<const TabResidual<float>*>(jobs), start, op.n, op.blocks, st);
            break;
        case kKRes64:
            e = launch_tab_residual<double>(static_cast<const TabResidual<double>*>(jobs), start, op.n, op.blocks, st);
            break;
        case kKEta32: e = launch_tab_eta<float>(static_cast<const TabEta<float>*>(jobs), start, op.n, op.blocks, st); break;
        case kKEta64: e = launch_tab_eta<double>(static_cast<const TabEta<double>*>(jobs), start, op.n, op.blocks, st); break;
        case kKMc32:
            e = launch_tab_vanka_mc<float>(static_cast<const TabVankaMc<float>*>(jobs), start, op.n, op.blocks,
                                           op.colour, op.linear != 0, op.rows, st);
            break;
        case kKMc64:
            e = launch_tab_vanka_mc<double>(static_cast<const TabVankaMc<double>*>(jobs), start, op.n, op.blocks,
                                            op.colour, op.linear != 0, op.rows, st);
            break;
        case kKFac32:
            e = launch_tab_adj_factor<float>(static_cast<const TabAdjFactor<float>*>(jobs), start, op.n, op.blocks, st);
            break;
        case kKFac64:
            e = launch_tab_adj_factor<double>(static_cast<const TabAdjFactor<double>*>(jobs), start, op.n, op.blocks,
                                              st);
            break;
        case kKPass32:
            e = launch_tab_adj_pass<float>(static_cast<const TabAdjPass<float>*>(jobs), start, op.n, op.blocks,
                                           op.colour, st);
            break;
        case kKPass64:
            e = launch_tab_adj_pass<double>(static_cast<const TabAdjPass<double>*>(jobs), start, op.n, op.blocks,
                                            op.colour, st);
            break;
        case kKVjp32: e = launch_tab_vjp<float>(static_cast<const TabVjp<float>*>(jobs), start, op.n, op.blocks, st); break;
        case kKVjp64: e = launch_tab_vjp<double>(static_cast<const TabVjp<double>*>(jobs), start, op.n, op.blocks, st); break;
        case kKGrad64:
            e = launch_tab_grad<double>(static_cast<const TabGrad<double>*>(jobs), start, op.n, op.blocks, st);
            break;
        default: return fail(GLIDE_ERR_INVALID_ARG, "unknown domain-table launch");
    }
    HIP_TRY(e, "domain-table launch");
    return GLIDE_OK;
}

}  // namespace

struct glide_ensemble {
    std::vector<EnsDomain> dom;
    glide_fas_config cfg, acfg;
    glide_vanka_config vf, va;
    glide_mixed_config mc;
    int which = 0;
    int lmax = 0;
    std::vector<int> rows;  // red-black tiles: levels 0 .. lmax - 1, then the coarsest stage
    double dt = 0.0;
    TabProgram prog;
    unsigned char* dev = nullptr;
    size_t dev_bytes = 0;
    int device = 0;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
};

namespace {

void ensemble_drop(glide_ensemble* e) {
    if (e->exec) hipGraphExecDestroy(e->exec);
    e->exec = nullptr;
    if (e->dev) hipFree(e->dev);
    e->dev = nullptr;
    e->dev_bytes = 0;
    e->prog = TabProgram{};
    e->dt = 0.0;
}

int ensemble_record(glide_ensemble* e, double dt) {
    ensemble_drop(e);
    EnsRec R{e->dom, e->cfg, e->acfg, e->vf, e->va, e->mc.fp64_coarsest != 0, e->mc.inner_vcycles, dt, e->prog, e->lmax,
             e->rows};
    GLIDE_TRY(R.step(e->mc.max_outer, e->which));
    if (!e->prog.arena.empty()) {
        HIP_TRY(hipMalloc(&e->dev, e->prog.arena.size()), "hipMalloc(domain tables)");
        e->dev_bytes = e->prog.arena.size();
        HIP_TRY(hipMemcpy(e->dev, e->prog.arena.data(), e->prog.arena.size(), hipMemcpyHostToDevice),
                "hipMemcpy(domain tables)");
    }
    e->dt = dt;
    return GLIDE_OK;
}

int ensemble_enqueue(glide_ensemble* e, hipStream_t st) {
    for (const TabOp& op : e->prog.ops) GLIDE_TRY(run_tab_op(op, e->dev, st));
    return GLIDE_OK;
}

}  // namespace

namespace glide {
void set_last_error(const std::string& msg) { g_last_error = msg; }
int noalloc_setup();
}  // namespace glide

// ====================================================================== C ABI
extern "C" {

const char* glide_version(void) { return "glide-mi355x 0.2.0 (gfx950, FP64 and mixed FP32/FP64)"; }

int glide_init(void) {
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));
    GLIDE_TRY(glide_enth_init());
    return noalloc_setup();
}

int glide_check_status(void* stream) {
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));
    hipStream_t st = (hipStream_t)stream;
    GLIDE_TRY(latch_enqueue(s, st));
    HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
    return latch_check(s);
}

int glide_persistent_spin_limit(int polls) {
    g_spin_limit = polls;
    return GLIDE_OK;
}

int glide_last_error(char* buf, size_t len) {
    const std::string& e = g_last_error;
    if (buf && len > 0) {
        size_t n = e.size() < len - 1 ? e.size() : len - 1;
        std::memcpy(buf, e.data(), n);
        buf[n] = '\0';
    }
    return (int)e.size();
}

int glide_compute_grounded(const glide_level* L, double relaxation, void* stream) {
    GLIDE_TRY(check_level(L));
    REQ(L->phi); REQ(L->H); REQ(L->bed);
    return grounded(*L, relaxation, (hipStream_t)stream);
}

int glide_ssa_residual(const glide_level* L, double dt, int flags, void* out_u, void* out_v, void* out_H,
                       void* stream) {
    GLIDE_TRY(check_level(L));
    GLIDE_TRY(check_ssa_inputs(*L));
    REQ(out_u); REQ(out_v); REQ(out_H);
    if ((flags & GLIDE_USE_FORCING)) { REQ(L->f_u); REQ(L->f_v); REQ(L->f_H); }
    if ((flags & GLIDE_USE_MASK)) { REQ(L->gamma); }
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    return residual(*L, dt, flags, P(out_u), P(out_v), P(out_H), (hipStream_t)stream);
}

int glide_ssa_jvp(const glide_level* L, double dt, int flags, const void* d_u, const void* d_v, const void* d_H,
                  void* out_u, void* out_v, void* out_H, void* stream) {
    GLIDE_TRY(check_level(L));
    GLIDE_TRY(check_ssa_inputs(*L));
    REQ(d_u); REQ(d_v); REQ(d_H); REQ(out_u); REQ(out_v); REQ(out_H);
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    SsaArgs<R> a = make_args(*L, dt, false);
    const double cr = (flags & GLIDE_FREEZE_CALVING) ? 0.0 : L->phys.calving_rate;
    HIP_TRY(launch_jvp<R>(a, CP(d_u), CP(d_v), CP(d_H), P(out_u), P(out_v), P(out_H), (flags & GLIDE_USE_MASK) ? 1 : 0,
                          cr, (hipStream_t)stream),
            "jvp");
    return GLIDE_OK;
}

int glide_norms3(const void* a, size_t na, const void* b, size_t nb, const void* c, size_t nc, double* out3,
                 void* stream) {
    if (!out3) return fail(GLIDE_ERR_INVALID_ARG, "out3 is NULL");
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));
    hipStream_t st = (hipStream_t)stream;
    GLIDE_TRY(sumsq3_async(CP(a), a ? na : 0, CP(b), b ? nb : 0, CP(c), c ? nc : 0, s->norms + 3 * 65, st));
    HIP_TRY(hipMemcpyAsync(s->host + 3 * 65, s->norms + 3 * 65, 3 * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (int k = 0; k < 3; ++k) out3[k] = std::sqrt(s->host[3 * 65 + k]);
    return GLIDE_OK;
}

int glide_vanka_sweeps(const glide_level* L, double dt, int n_iter, const glide_vanka_config* vc, int flags,
                       void* stream) {
    GLIDE_TRY(check_level(L));
    GLIDE_TRY(check_ssa_inputs(*L));
    if (!vc) return fail(GLIDE_ERR_INVALID_ARG, "vc is NULL");
    REQ(L->f_u); REQ(L->f_v); REQ(L->f_H); REQ(L->gamma); REQ(L->work);
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    if (n_iter < 0) return fail(GLIDE_ERR_INVALID_ARG, "n_iter must be >= 0");
    return vanka_sweeps(*L, dt, n_iter, *vc, flags, (hipStream_t)stream);
}

int glide_vanka_dump(const glide_level* L, double dt, void* J, void* r, void* stream) {
    GLIDE_TRY(check_level(L));
    GLIDE_TRY(check_ssa_inputs(*L));
    REQ(J); REQ(r); REQ(L->f_u); REQ(L->f_v); REQ(L->f_H); REQ(L->gamma);
    SsaArgs<R> a = make_args(*L, dt, false);
    HIP_TRY(launch_vanka_dump<R>(a, P(J), P(r), L->phys.calving_rate, (hipStream_t)stream), "vanka dump");
    return GLIDE_OK;
}

int glide_restrict_vfacet(const void* fine, void* coarse, int ny_coarse, int nx_coarse, void* stream) {
    if (!fine || !coarse || ny_coarse < 1 || nx_coarse < 1) return fail(GLIDE_ERR_INVALID_ARG, "bad restrict args");
    HIP_TRY(launch_restrict_vfacet<R>(CP(fine), P(coarse), ny_coarse, nx_coarse, (hipStream_t)stream), "restrict");
    return GLIDE_OK;
}
int glide_restrict_hfacet(const void* fine, void* coarse, int ny_coarse, int nx_coarse, void* stream) {
    if (!fine || !coarse || ny_coarse < 1 || nx_coarse < 1) return fail(GLIDE_ERR_INVALID_ARG, "bad restrict args");
    HIP_TRY(launch_restrict_hfacet<R>(CP(fine), P(coarse), ny_coarse, nx_coarse, (hipStream_t)stream), "restrict");
    return GLIDE_OK;
}
int glide_restrict_cell(const void* fine, void* coarse, int ny_coarse, int nx_coarse, int use_max, void* stream) {
    if (!fine || !coarse || ny_coarse < 1 || nx_coarse < 1) return fail(GLIDE_ERR_INVALID_ARG, "bad restrict args");
    HIP_TRY(launch_restrict_cell<R>(CP(fine), P(coarse), ny_coarse, nx_coarse, use_max, (hipStream_t)stream),
            "restrict");
    return GLIDE_OK;
}
int glide_prolong_vfacet_bilinear(const void* coarse, void* fine, int ny_fine, int nx_fine, void* stream) {
    if (!fine || !coarse || ny_fine < 2 || nx_fine < 2 || (ny_fine & 1) || (nx_fine & 1))
        return fail(GLIDE_ERR_INVALID_ARG, "bad prolong args");
    HIP_TRY(launch_prolong_vfacet_bilinear<R>(CP(coarse), P(fine), ny_fine, nx_fine, (hipStream_t)stream), "prolong");
    return GLIDE_OK;
}
int glide_prolong_hfacet_bilinear(const void* coarse, void* fine, int ny_fine, int nx_fine, void* stream) {
    if (!fine || !coarse || ny_fine < 2 || nx_fine < 2 || (ny_fine & 1) || (nx_fine & 1))
        return fail(GLIDE_ERR_INVALID_ARG, "bad prolong args");
    HIP_TRY(launch_prolong_hfacet_bilinear<R>(CP(coarse), P(fine), ny_fine, nx_fine, (hipStream_t)stream), "prolong");
    return GLIDE_OK;
}
int glide_prolong_cell(const void* coarse, void* fine, int ny_fine, int nx_fine, int bilinear, void* stream) {
    if (!fine || !coarse || ny_fine < 2 || nx_fine < 2 || (ny_fine & 1) || (nx_fine & 1))
        return fail(GLIDE_ERR_INVALID_ARG, "bad prolong args");
    HIP_TRY(launch_prolong_cell<R>(CP(coarse), P(fine), ny_fine, nx_fine, bilinear, (hipStream_t)stream), "prolong");
    return GLIDE_OK;
}

int glide_fascd_solve(const glide_level* levels, int n_levels, int start_level, double dt,
                      const glide_fas_config* cfg, const glide_vanka_config* vc, glide_solve_report* report,
                      void* stream) {
    GLIDE_TRY(check_hierarchy(levels, n_levels, start_level, cfg, vc));
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    FasCtx X{levels, n_levels, dt, *cfg, vc, (hipStream_t)stream};
    const glide_level& L = levels[start_level];
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));  // allocated before any capture
    // set_rhs: f_H = H_prev/dt + smb (glide/operators.py:275-276)
    HIP_TRY(launch_axpby<R>(P(L.f_H), 1.0 / dt, CP(L.H_prev), 1.0, CP(L.smb), 0.0, nullptr, (long)nC(L), X.st),
            "rhs");
    auto resid = [&](bool initial, double* slot) -> int {
        if (initial) GLIDE_TRY(grounded(L, 0.0, X.st));
        GLIDE_TRY(residual(L, dt, GLIDE_USE_FORCING | GLIDE_USE_MASK, P(L.r_u), P(L.r_v), P(L.r_H), X.st));
        return sumsq3_async(CP(L.r_u), nU(L), CP(L.r_v), nV(L), CP(L.r_H), nC(L), slot, X.st);
    };
    auto cycle = [&]() -> int { return fas_vcycle(X, start_level, true); };
    return solve_loop(X, start_level, report, resid, cycle, 'f');
}

int glide_fascd_vcycle(const glide_level* levels, int n_levels, int level, int finest, double dt,
                       const glide_fas_config* cfg, const glide_vanka_config* vc, void* stream) {
    GLIDE_TRY(check_hierarchy(levels, n_levels, level, cfg, vc));
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    FasCtx X{levels, n_levels, dt, *cfg, vc, (hipStream_t)stream};
    return fas_vcycle(X, level, finest != 0);
}

int glide_fas_adjoint_vcycle(const glide_level* levels, int n_levels, int level, int finest, double dt,
                             const glide_fas_config* cfg, const glide_vanka_config* vc, void* stream) {
    GLIDE_TRY(check_hierarchy(levels, n_levels, level, cfg, vc));
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    FasCtx X{levels, n_levels, dt, *cfg, vc, (hipStream_t)stream};
    GLIDE_TRY(adjoint_prepare(levels[level], dt, vc[level], 0, X.st));
    return adj_vcycle(X, level, finest != 0);
}

int glide_ssa_vjp(const glide_level* L, double dt, int flags, const void* lambda_u, const void* lambda_v,
                  const void* lambda_H, void* out_u, void* out_v, void* out_H, void* stream) {
    GLIDE_TRY(check_level(L));
    GLIDE_TRY(check_ssa_inputs(*L));
    REQ(lambda_u); REQ(lambda_v); REQ(lambda_H); REQ(out_u); REQ(out_v); REQ(out_H);
    if (flags & GLIDE_USE_FORCING) { REQ(L->af_u); REQ(L->af_v); REQ(L->af_H); }
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    return vjp(*L, dt, flags, CP(lambda_u), CP(lambda_v), CP(lambda_H), P(out_u), P(out_v), P(out_H),
               (hipStream_t)stream);
}

int glide_vanka_adjoint_sweeps(const glide_level* L, double dt, int n_iter, const glide_vanka_config* vc, int flags,
                               void* stream) {
    GLIDE_TRY(check_level(L));
    GLIDE_TRY(check_ssa_inputs(*L));
    if (!vc) return fail(GLIDE_ERR_INVALID_ARG, "vc is NULL");
    REQ(L->lambda_u); REQ(L->lambda_v); REQ(L->lambda_H); REQ(L->af_u); REQ(L->af_v); REQ(L->af_H);
    REQ(L->ar_u); REQ(L->ar_v); REQ(L->ar_H); REQ(L->work);
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    GLIDE_TRY(adjoint_prepare(*L, dt, *vc, flags, (hipStream_t)stream));
    return vanka_adjoint_sweeps(*L, dt, n_iter, *vc, flags, (hipStream_t)stream, true);
}

int glide_fas_adjoint_solve(const glide_level* levels, int n_levels, int start_level, double dt,
                            const glide_fas_config* cfg, const glide_vanka_config* vc, glide_solve_report* report,
                            void* stream) {
    GLIDE_TRY(check_hierarchy(levels, n_levels, start_level, cfg, vc));
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    FasCtx X{levels, n_levels, dt, *cfg, vc, (hipStream_t)stream};
    const glide_level& L = levels[start_level];
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));  // allocated before any capture
    GLIDE_TRY(adjoint_prepare(L, dt, vc[start_level], 0, X.st));
    auto resid = [&](bool, double* slot) -> int {
        GLIDE_TRY(vjp(L, dt, GLIDE_USE_FORCING, CP(L.lambda_u), CP(L.lambda_v), CP(L.lambda_H), P(L.ar_u), P(L.ar_v),
                      P(L.ar_H), X.st, true));
        return sumsq3_async(CP(L.ar_u), nU(L), CP(L.ar_v), nV(L), CP(L.ar_H), nC(L), slot, X.st);
    };
    auto cycle = [&]() -> int { return adj_vcycle(X, start_level, true); };
    return solve_loop(X, start_level, report, resid, cycle, 'a');
}

int glide_fascd_solve_mixed(const glide_level* levels64, const glide_level* levels32, int n_levels, int start_level,
                            double dt, const glide_fas_config* cfg, const glide_vanka_config* vc,
                            const glide_mixed_config* mc, glide_solve_report* report, int* stats, void* stream) {
    GLIDE_TRY(check_mixed(levels64, levels32, n_levels, start_level, cfg, vc, mc, false));
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    const glide_level& D = levels64[start_level];
    const glide_level& A = levels32[start_level];
    REQ(D.f_H); REQ(D.r_u); REQ(D.r_v); REQ(D.r_H); REQ(D.H_prev); REQ(D.smb); REQ(D.gamma);
    REQ(D.w_u); REQ(D.w_v); REQ(D.w_H); REQ(D.cphi);
    hipStream_t st = (hipStream_t)stream;
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));  // allocated before any capture
    // set_rhs (glide/operators.py:275-276) in FP64, the active-set bound, and
    // the parameters of every level in FP32
    HIP_TRY(launch_axpby<R>(P(D.f_H), 1.0 / dt, CP(D.H_prev), 1.0, CP(D.smb), 0.0, nullptr, (long)nC(D), st), "rhs");
    GLIDE_TRY(fill(D.gamma, D.phys.thklim, nC(D), st));
    for (int l = start_level; l < n_levels; ++l) {
        const glide_level& E = levels64[l];
        const glide_level& G = levels32[l];
        MixedBatch pb;
        pb.add(kMixTo32, G.bed, E.bed, nullptr, (long)nC(E));
        pb.add(kMixTo32, G.B, E.B, nullptr, (long)nC(E));
        pb.add(kMixTo32, G.beta, E.beta, nullptr, (long)nC(E));
        GLIDE_TRY(run_mixed(pb, st));
    }
    MixedCtx X{levels64, levels32, n_levels, dt, *cfg, vc, mc->fp64_coarsest != 0, st};
    std::vector<glide_vanka_config> vc64(vc, vc + n_levels);
    for (auto& q : vc64) {
        if (mc->fp64_newton_steps > 0) q.newton_steps = mc->fp64_newton_steps;
        if (mc->fp64_relaxation > 0.0) q.relaxation = mc->fp64_relaxation;
    }
    auto resid = [&](hipStream_t cs, double* slot) -> int {
        GLIDE_TRY(grounded(D, 0.0, cs));
        GLIDE_TRY(residual(D, dt, GLIDE_USE_FORCING | GLIDE_USE_MASK, P(D.r_u), P(D.r_v), P(D.r_H), cs));
        return sumsq3_async(CP(D.r_u), nU(D), CP(D.r_v), nV(D), CP(D.r_H), nC(D), slot, cs);
    };
    auto step = [&](hipStream_t cs, int phase) -> int {
        X.st = cs;
        EwBatch<R> save;
        save.copy(P(D.w_u), CP(D.u), (long)nU(D));
        save.copy(P(D.w_v), CP(D.v), (long)nV(D));
        save.copy(P(D.w_H), CP(D.H), (long)nC(D));
        save.copy(P(D.cphi), CP(D.mask), (long)nC(D));
        HIP_TRY(launch_ew_batch<R>(save, cs), "save state");
        if (phase == 2) {
            // one FP64 V-cycle with the same red-black smoother and counts
            // (its local Newton steps and relaxation from mc if set)
            FasCtx F{levels64, n_levels, dt, *cfg, vc64.data(), cs};
            GLIDE_TRY(fas_vcycle(F, start_level, true));
            X.st = st;
            return GLIDE_OK;
        }
        // the defect without the active set (the inner solve applies the bound)
        GLIDE_TRY(residual(D, dt, GLIDE_USE_FORCING, P(D.r_u), P(D.r_v), P(D.r_H), cs));
        MixedBatch to32;
        to32.add(kMixTo32, A.u, D.u, nullptr, (long)nU(D));
        to32.add(kMixTo32, A.v, D.v, nullptr, (long)nV(D));
        to32.add(kMixTo32, A.H, D.H, nullptr, (long)nC(D));
        to32.add(kMixTo32, A.mask, D.mask, nullptr, (long)nC(D));
        to32.add(kMixTo32, A.phi, D.phi, nullptr, (long)nC(D));
        GLIDE_TRY(run_mixed(to32, cs));
        GLIDE_TRY(residual_t<float>(A, dt, 0, PT<float>(A.F_u), PT<float>(A.F_v), PT<float>(A.F_H), cs));
        MixedBatch defect;
        defect.add(kMixDefect, A.f_u, A.F_u, D.r_u, (long)nU(D));
        defect.add(kMixDefect, A.f_v, A.F_v, D.r_v, (long)nV(D));
        defect.add(kMixDefect, A.f_H, A.F_H, D.r_H, (long)nC(D));
        GLIDE_TRY(run_mixed(defect, cs));
        for (int q = 0; q < mc->inner_vcycles; ++q) GLIDE_TRY(mixed_vcycle(X, start_level, true));
        MixedBatch upd;
        upd.add(kMixUpdate, D.u, A.u, nullptr, (long)nU(D));
        upd.add(kMixUpdate, D.v, A.v, nullptr, (long)nV(D));
        upd.add(kMixUpdateH, D.H, A.H, A.mask, (long)nC(D), D.phys.thklim);
        upd.add(kMixTo64, D.mask, A.mask, nullptr, (long)nC(D));
        GLIDE_TRY(run_mixed(upd, cs));
        X.st = st;
        return GLIDE_OK;
    };
    auto undo = [&](hipStream_t cs) -> int {
        EwBatch<R> b;
        b.copy(P(D.u), CP(D.w_u), (long)nU(D));
        b.copy(P(D.v), CP(D.w_v), (long)nV(D));
        b.copy(P(D.H), CP(D.w_H), (long)nC(D));
        b.copy(P(D.mask), CP(D.cphi), (long)nC(D));
        HIP_TRY(launch_ew_batch<R>(b, cs), "restore state");
        return GLIDE_OK;
    };
    std::vector<glide_vanka_config> ref = reference_vanka_configs(n_levels);
    auto fallback = [&](hipStream_t cs) -> int {
        glide_fas_config fc = *cfg;
        fc.coarsest_steps = 200; fc.pre_steps = 10; fc.post_steps = 20; fc.finest_steps = 50;
        FasCtx F{levels64, n_levels, dt, fc, ref.data(), cs};
        return fas_vcycle(F, start_level, true);
    };
    auto key = [&](int phase) {
        return mixed_key(phase == 0 ? 'a' : phase == 1 ? 'm' : 'f', levels64, levels32, n_levels, start_level, dt,
                         *cfg, vc, *mc);
    };
    auto prep = [&](hipStream_t cs) -> int { return grounded(D, 0.0, cs); };
    const bool fixed = !(mc->relative_tolerance > 0.0) && !(mc->absolute_tolerance > 0.0);
    const int aa_m = (mc->krylov_dim > 0 && (report || !fixed)) ? anderson_depth(mc->krylov_dim) : 0;
    AndersonLayout W{};
    if (aa_m > 0) W = anderson_layout(mc->krylov_work, D.ny, D.nx, aa_m);
    int aa_h = 0, aa_next = 0;
    bool aa_prev = false;
    R* st3[3] = {P(D.u), P(D.v), P(D.H)};
    const R* saved3[3] = {CP(D.w_u), CP(D.w_v), CP(D.w_H)};
    const long len3[3] = {(long)nU(D), (long)nV(D), (long)nC(D)}, off3[3] = {0, (long)nU(D), (long)(nU(D) + nV(D))};
    double* hk = s->host + 3 * 66;
    // Anderson mixing of the FP64 finishing cycles x -> G(x): with f = G(x) - x,
    // x_new = G(x_k) - sum_i gamma_i (G_i - G_{i-1}), gamma minimising
    // |f_k - sum_i gamma_i (f_i - f_{i-1})| over the last aa_m differences;
    // kept only if it lowers the FP64 residual norm, else G(x_k) and a
    // fresh history.
    auto accel = [&](hipStream_t cs, int phase, double* m3) -> int {
        if (aa_m == 0) return GLIDE_OK;
        if (phase != 2) {
            aa_h = aa_next = 0;
            aa_prev = false;
            return GLIDE_OK;
        }
        EwBatch<R> b;
        for (int q = 0; q < 3; ++q) {
            b.copy(W.g + off3[q], st3[q], len3[q]);                              // g_k = G(x_k)
            b.axpby(W.f + off3[q], 1.0, st3[q], -1.0, saved3[q], len3[q]);      // f_k = g_k - x_k
        }
        HIP_TRY(launch_ew_batch<R>(b, cs), "Anderson residual");
        if (aa_prev) {
            const int slot = aa_next % aa_m;
            EwBatch<R> d;
            d.axpby(W.dF + (long)slot * W.n, 1.0, W.f, -1.0, W.f_prev, W.n);
            d.axpby(W.dG + (long)slot * W.n, 1.0, W.g, -1.0, W.g_prev, W.n);
            HIP_TRY(launch_ew_batch<R>(d, cs), "Anderson differences");
            aa_next++;
            aa_h = std::min(aa_h + 1, aa_m);
        }
        {
            EwBatch<R> c;
            c.copy(W.f_prev, W.f, W.n);
            c.copy(W.g_prev, W.g, W.n);
            HIP_TRY(launch_ew_batch<R>(c, cs), "Anderson history");
        }
        aa_prev = true;
        if (aa_h == 0) return GLIDE_OK;
        const int h = aa_h, row = kKrylovMaxDots + 1;
        for (int i = 0; i < h; ++i)
            HIP_TRY(launch_krylov_dots(W.dF + (long)i * W.n, W.dF, W.n, W.n, h, W.partial, W.dots + i * row, cs),
                    "Anderson inner products");
        HIP_TRY(launch_krylov_dots(W.f, W.dF, W.n, W.n, h, W.partial, W.dots + h * row, cs), "Anderson inner products");
        HIP_TRY(hipMemcpyAsync(hk, W.dots, sizeof(double) * (h + 1) * row, hipMemcpyDeviceToHost, cs), "Anderson D2H");
        HIP_TRY(hipStreamSynchronize(cs), "hipStreamSynchronize");
        // (dF^T dF + delta I) gamma = dF^T f, Gaussian elimination with partial pivoting
        double M[kAndersonMax][kAndersonMax + 1];
        double dmax = 0.0;
        for (int i = 0; i < h; ++i) dmax = std::max(dmax, hk[i * row + i]);
        for (int i = 0; i < h; ++i) {
            for (int j = 0; j < h; ++j) M[i][j] = hk[i * row + j] + (i == j ? 1e-12 * dmax : 0.0);
            M[i][h] = hk[h * row + i];
        }
        for (int c = 0; c < h; ++c) {
            int piv = c;
            for (int i = c + 1; i < h; ++i)
                if (std::fabs(M[i][c]) > std::fabs(M[piv][c])) piv = i;
            if (piv != c)
                for (int j = 0; j <= h; ++j) std::swap(M[c][j], M[piv][j]);
            if (M[c][c] == 0.0) return GLIDE_OK;  // no usable direction: keep G(x_k)
            for (int i = c + 1; i < h; ++i) {
                const double f = M[i][c] / M[c][c];
                for (int j = c; j <= h; ++j) M[i][j] -= f * M[c][j];
            }
        }
        double gam[kAndersonMax];
        for (int i = h - 1; i >= 0; --i) {
            double t = M[i][h];
            for (int j = i + 1; j < h; ++j) t -= M[i][j] * gam[j];
            gam[i] = t / M[i][i];
            if (!std::isfinite(gam[i])) return GLIDE_OK;
        }
        for (int i = 0; i < h; ++i) hk[i] = gam[i];
        HIP_TRY(hipMemcpyAsync(W.coef, hk, sizeof(double) * h, hipMemcpyHostToDevice, cs), "Anderson H2D");
        for (int q = 0; q < 3; ++q)
            HIP_TRY(launch_krylov_axpy(st3[q], W.dG + off3[q], W.n, len3[q], h, W.coef, -1.0, cs), "Anderson update");
        double a3[3];
        GLIDE_TRY(resid(cs, s->norms));
        HIP_TRY(hipMemcpyAsync(s->host, s->norms, sizeof(double) * 3, hipMemcpyDeviceToHost, cs), "norms D2H");
        HIP_TRY(hipStreamSynchronize(cs), "hipStreamSynchronize");
        a3[0] = s->host[0]; a3[1] = s->host[1]; a3[2] = s->host[2];
        const double ra = std::sqrt(a3[0] + a3[1] + a3[2]), rg = std::sqrt(m3[0] + m3[1] + m3[2]);
        if (ra < rg) {
            std::memcpy(m3, a3, sizeof(a3));
            return GLIDE_OK;
        }
        // not better: back to G(x_k), a fresh history
        EwBatch<R> back;
        for (int q = 0; q < 3; ++q) back.copy(st3[q], W.g + off3[q], len3[q]);
        HIP_TRY(launch_ew_batch<R>(back, cs), "Anderson revert");
        GLIDE_TRY(resid(cs, s->norms));  // the residual buffers and phi of G(x_k) again
        aa_h = aa_next = 0;
        return GLIDE_OK;
    };
    return mixed_outer_loop(*mc, st, key, true, report, stats, resid, step, undo, fallback, prep, accel);
}

int glide_fas_adjoint_solve_mixed(const glide_level* levels64, const glide_level* levels32, int n_levels,
                                  int start_level, double dt, const glide_fas_config* cfg,
                                  const glide_vanka_config* vc, const glide_mixed_config* mc,
                                  glide_solve_report* report, int* stats, void* stream) {
    GLIDE_TRY(check_mixed(levels64, levels32, n_levels, start_level, cfg, vc, mc, true));
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    const glide_level& D = levels64[start_level];
    const glide_level& A = levels32[start_level];
    REQ(D.lambda_u); REQ(D.lambda_v); REQ(D.lambda_H); REQ(D.af_u); REQ(D.af_v); REQ(D.af_H);
    REQ(D.ar_u); REQ(D.ar_v); REQ(D.ar_H); REQ(D.wl_u); REQ(D.wl_v); REQ(D.wl_H);
    for (int l = start_level; l < n_levels; ++l) {
        if (!levels64[l].adj_cache)
            return fail(GLIDE_ERR_INVALID_ARG, "the mixed adjoint solve needs an adjoint cache on every FP64 level");
        if (!(levels64[l].phys.m == 1.0 && levels64[l].phys.u_reg > 0.0))
            return fail(GLIDE_ERR_INVALID_ARG, "the mixed adjoint solve needs linear drag (m = 1)");
    }
    hipStream_t st = (hipStream_t)stream;
    DeviceScratch* s = nullptr;
    GLIDE_TRY(scratch_for_current_device(&s));
    // the forward state and the parameters in FP32, the start level's cache
    {
        MixedBatch b;
        b.add(kMixTo32, A.u, D.u, nullptr, (long)nU(D));
        b.add(kMixTo32, A.v, D.v, nullptr, (long)nV(D));
        b.add(kMixTo32, A.H, D.H, nullptr, (long)nC(D));
        b.add(kMixTo32, A.phi, D.phi, nullptr, (long)nC(D));
        b.add(kMixTo32, A.mask, D.mask, nullptr, (long)nC(D));
        GLIDE_TRY(run_mixed(b, st));
    }
    for (int l = start_level; l < n_levels; ++l) {
        const glide_level& E = levels64[l];
        const glide_level& G = levels32[l];
        MixedBatch pb;
        pb.add(kMixTo32, G.bed, E.bed, nullptr, (long)nC(E));
        pb.add(kMixTo32, G.B, E.B, nullptr, (long)nC(E));
        pb.add(kMixTo32, G.beta, E.beta, nullptr, (long)nC(E));
        GLIDE_TRY(run_mixed(pb, st));
    }
    GLIDE_TRY(adjoint_prepare_t<float>(A, dt, vc[start_level], st));
    MixedCtx X{levels64, levels32, n_levels, dt, *cfg, vc, mc->fp64_coarsest != 0, st};
    auto resid = [&](hipStream_t cs, double* slot) -> int {
        GLIDE_TRY(vjp(D, dt, GLIDE_USE_FORCING, CP(D.lambda_u), CP(D.lambda_v), CP(D.lambda_H), P(D.ar_u), P(D.ar_v),
                      P(D.ar_H), cs));
        return sumsq3_async(CP(D.ar_u), nU(D), CP(D.ar_v), nV(D), CP(D.ar_H), nC(D), slot, cs);
    };
    auto step = [&](hipStream_t cs, int) -> int {
        X.st = cs;
        EwBatch<R> save;
        save.copy(P(D.wl_u), CP(D.lambda_u), (long)nU(D));
        save.copy(P(D.wl_v), CP(D.lambda_v), (long)nV(D));
        save.copy(P(D.wl_H), CP(D.lambda_H), (long)nC(D));
        HIP_TRY(launch_ew_batch<R>(save, cs), "save multipliers");
        // the correction solve J^T d = -(J^T lambda - g) from d = 0
        MixedBatch b;
        b.add(kMixNeg32, A.af_u, D.ar_u, nullptr, (long)nU(D));
        b.add(kMixNeg32, A.af_v, D.ar_v, nullptr, (long)nV(D));
        b.add(kMixNeg32, A.af_H, D.ar_H, nullptr, (long)nC(D));
        GLIDE_TRY(run_mixed(b, cs));
        EwBatch<float> z;
        z.fill(PT<float>(A.lambda_u), 0.0f, (long)nU(A));
        z.fill(PT<float>(A.lambda_v), 0.0f, (long)nV(A));
        z.fill(PT<float>(A.lambda_H), 0.0f, (long)nC(A));
        HIP_TRY(launch_ew_batch<float>(z, cs), "zero correction");
        for (int q = 0; q < mc->inner_vcycles; ++q) GLIDE_TRY(mixed_adj_vcycle(X, start_level, true));
        // the multipliers of pinned rows (domain-edge facets, active-set
        // cells) stay zero: the transpose operator does not determine them
        HIP_TRY(launch_pin_rows<float>(PT<float>(A.lambda_u), PT<float>(A.lambda_v), PT<float>(A.lambda_H),
                                       CPT<float>(A.mask), A.ny, A.nx, cs),
                "pin rows");
        MixedBatch u;
        u.add(kMixAdd64, D.lambda_u, A.lambda_u, nullptr, (long)nU(D));
        u.add(kMixAdd64, D.lambda_v, A.lambda_v, nullptr, (long)nV(D));
        u.add(kMixAdd64, D.lambda_H, A.lambda_H, nullptr, (long)nC(D));
        GLIDE_TRY(run_mixed(u, cs));
        X.st = st;
        return GLIDE_OK;
    };
    auto undo = [&](hipStream_t cs) -> int {
        EwBatch<R> b;
        b.copy(P(D.lambda_u), CP(D.wl_u), (long)nU(D));
        b.copy(P(D.lambda_v), CP(D.wl_v), (long)nV(D));
        b.copy(P(D.lambda_H), CP(D.wl_H), (long)nC(D));
        HIP_TRY(launch_ew_batch<R>(b, cs), "restore multipliers");
        return GLIDE_OK;
    };
    std::vector<glide_vanka_config> ref = reference_vanka_configs(n_levels);
    auto fallback = [&](hipStream_t cs) -> int {
        glide_fas_config fc = *cfg;
        fc.coarsest_steps = 200; fc.pre_steps = 10; fc.post_steps = 20; fc.finest_steps = 50;
        FasCtx F{levels64, n_levels, dt, fc, ref.data(), cs};
        GLIDE_TRY(adjoint_prepare(D, dt, ref[start_level], 0, cs));
        return adj_vcycle(F, start_level, true);
    };
    auto key = [&](int phase) {
        return mixed_key(phase == 0 ? 'b' : 'n', levels64, levels32, n_levels, start_level, dt, *cfg, vc, *mc);
    };
    auto prep = [&](hipStream_t cs) -> int {
        return vjp(D, dt, GLIDE_USE_FORCING, CP(D.lambda_u), CP(D.lambda_v), CP(D.lambda_H), P(D.ar_u), P(D.ar_v),
                   P(D.ar_H), cs);
    };
    auto no_accel = [](hipStream_t, int, double*) -> int { return GLIDE_OK; };
    const bool fixed = !(mc->relative_tolerance > 0.0) && !(mc->absolute_tolerance > 0.0);
    if (mc->krylov_dim > 0 && (report || !fixed)) {
        const KrylovLayout K = krylov_layout(mc->krylov_work, D.ny, D.nx, mc->krylov_dim);
        // z = M v: the FP32 correction V-cycles for J^T d = v from d = 0 (one
        // cached graph), the pinned rows zero
        auto precond = [&](hipStream_t cs, const double* v, double* z) -> int {
            MixedBatch b;
            b.add(kMixTo32, A.af_u, v, nullptr, (long)nU(D));
            b.add(kMixTo32, A.af_v, v + K.nu, nullptr, (long)nV(D));
            b.add(kMixTo32, A.af_H, v + K.nu + K.nv, nullptr, (long)nC(D));
            GLIDE_TRY(run_mixed(b, cs));
            auto body = [&](hipStream_t gs) -> int {
                X.st = gs;
                EwBatch<float> zf;
                zf.fill(PT<float>(A.lambda_u), 0.0f, (long)nU(A));
                zf.fill(PT<float>(A.lambda_v), 0.0f, (long)nV(A));
                zf.fill(PT<float>(A.lambda_H), 0.0f, (long)nC(A));
                HIP_TRY(launch_ew_batch<float>(zf, gs), "zero correction");
                for (int q = 0; q < mc->inner_vcycles; ++q) GLIDE_TRY(mixed_adj_vcycle(X, start_level, true));
                HIP_TRY(launch_pin_rows<float>(PT<float>(A.lambda_u), PT<float>(A.lambda_v), PT<float>(A.lambda_H),
                                               CPT<float>(A.mask), A.ny, A.nx, gs),
                        "pin rows");
                X.st = st;
                return GLIDE_OK;
            };
            if (g_graphs_on && !timing_on())
                GLIDE_TRY(replay(mixed_key('p', levels64, levels32, n_levels, start_level, dt, *cfg, vc, *mc), cs,
                                 body));
            else
                GLIDE_TRY(body(cs));
            MixedBatch c;
            c.add(kMixTo64, z, A.lambda_u, nullptr, (long)nU(D));
            c.add(kMixTo64, z + K.nu, A.lambda_v, nullptr, (long)nV(D));
            c.add(kMixTo64, z + K.nu + K.nv, A.lambda_H, nullptr, (long)nC(D));
            return run_mixed(c, cs);
        };
        auto apply = [&](hipStream_t cs, const double* zv, double* w) -> int {
            return vjp(D, dt, 0, zv, zv + K.nu, zv + K.nu + K.nv, w, w + K.nu, w + K.nu + K.nv, cs);
        };
        int its = 0;
        const int rc = krylov_outer_loop(*mc, st, K, mc->krylov_dim, D, report, stats, &its, resid, precond, apply);
        if (rc != kKrylovStalled) return rc;
        // a restart cycle that did not reduce the residual: the remaining
        // iterations as defect correction (with its undo and fallback)
        glide_mixed_config rest = *mc;
        rest.max_outer = std::max(0, mc->max_outer - its);
        const int rc2 = mixed_outer_loop(rest, st, key, false, report, stats, resid, step, undo, fallback, prep,
                                         no_accel);
        if (stats) stats[1] += its;
        return rc2;
    }
    return mixed_outer_loop(*mc, st, key, false, report, stats, resid, step, undo, fallback, prep, no_accel);
}

int64_t glide_krylov_work_doubles(int ny, int nx, int m) {
    if (ny < 1 || nx < 1 || m < 1 || m > 16) return -1;
    return krylov_doubles(ny, nx, m);
}

int glide_adjoint_fused_enable(int on) {
    g_fused_adjoint = on != 0;
    return GLIDE_OK;
}

int glide_forward_fused_enable(int on) {
    g_fused_forward = on != 0;
    return GLIDE_OK;
}

int glide_adjoint_coarse_enable(int on) {
    g_coarse_adjoint = on != 0;
    return GLIDE_OK;
}

int glide_forward_persistent_enable(int on) {
    g_persistent_forward = on != 0;
    return GLIDE_OK;
}

int glide_adjoint_persistent_enable(int on) {
    g_persistent_adjoint = on != 0;
    return GLIDE_OK;
}

int glide_persistent_fenced_enable(int on) {
    g_persistent_fenced = on != 0;
    return GLIDE_OK;
}

int glide_adjoint_pass_fused_enable(int on) {
    g_pass_fused_adjoint = on != 0;
    return GLIDE_OK;
}

int glide_graphs_enable(int on) {
    g_graphs_on = on != 0;
    return GLIDE_OK;
}

int glide_graphs_capacity(int n) {
    if (n < 1 || n > 4096) return fail(GLIDE_ERR_INVALID_ARG, "the graph cache holds 1 to 4096 graphs");
    std::lock_guard<std::mutex> lk(g_graph_mutex);
    g_max_graphs = (size_t)n;
    return GLIDE_OK;
}

int glide_graphs_clear(void) {
    std::lock_guard<std::mutex> lk(g_graph_mutex);
    for (auto& G : g_graphs) {
        for (auto& e : G.entries) drop_entry(e);
        G.entries.clear();
    }
    return GLIDE_OK;
}

int glide_timing_enable(int on) {
    timing_enable(on != 0);
    return GLIDE_OK;
}

int glide_timing_read(int kernel, long* count, double* total_ms, double* total_cells) {
    if (!count || !total_ms || !total_cells) return fail(GLIDE_ERR_INVALID_ARG, "NULL output");
    HIP_TRY(timing_read(kernel, count, total_ms, total_cells), "timing_read");
    return GLIDE_OK;
}

int glide_timing_read_units(int kernel, long units, long* count, double* total_ms) {
    if (!count || !total_ms) return fail(GLIDE_ERR_INVALID_ARG, "NULL output");
    double u = 0.0;
    HIP_TRY(timing_read(kernel, count, total_ms, &u, units), "timing_read");
    return GLIDE_OK;
}

int glide_gradient_rate_factor(const glide_level* L, double dt, void* g_B, void* stream) {
    GLIDE_TRY(check_level(L));
    GLIDE_TRY(check_ssa_inputs(*L));
    REQ(L->lambda_u); REQ(L->lambda_v); REQ(g_B);
    SsaArgs<R> a = make_args(*L, dt, true);
    HIP_TRY(launch_grad_B<R>(a, CP(L->lambda_u), CP(L->lambda_v), P(g_B), (hipStream_t)stream), "grad B");
    return GLIDE_OK;
}

int glide_gradients(const glide_level* L, double dt, int which, void* stream) {
    GLIDE_TRY(check_level(L));
    const int exact = (which & GLIDE_GRAD_EXACT) ? 1 : 0;
    hipStream_t st = (hipStream_t)stream;
    REQ(L->lambda_H);
    if (which & 3) {
        GLIDE_TRY(check_ssa_inputs(*L));
        REQ(L->lambda_u); REQ(L->lambda_v);
        if ((which & 1)) REQ(L->g_beta);
        if ((which & 2)) REQ(L->g_bed);
        SsaArgs<R> a = make_args(*L, dt, true);
        HIP_TRY(launch_grad_beta_bed<R>(a, CP(L->lambda_u), CP(L->lambda_v), (which & 1) ? P(L->g_beta) : nullptr,
                                        (which & 2) ? P(L->g_bed) : nullptr, exact, st),
                "gradients");
    }
    if (exact && (which & 12)) REQ(L->mask);
    if (which & 4) {
        REQ(L->g_H_prev);
        if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
        HIP_TRY(launch_axpby<R>(P(L->g_H_prev), -1.0 / dt, CP(L->lambda_H), 0.0, CP(L->lambda_H), 0.0, nullptr,
                                (long)nC(*L), st),
                "grad H_prev");
        if (exact) HIP_TRY(launch_zero_masked<R>(P(L->g_H_prev), CP(L->mask), (long)nC(*L), st), "grad H_prev mask");
    }
    if (which & 8) {
        REQ(L->g_smb);
        HIP_TRY(launch_axpby<R>(P(L->g_smb), -1.0, CP(L->lambda_H), 0.0, CP(L->lambda_H), 0.0, nullptr, (long)nC(*L),
                                st),
                "grad smb");
        if (exact) HIP_TRY(launch_zero_masked<R>(P(L->g_smb), CP(L->mask), (long)nC(*L), st), "grad smb mask");
    }
    return GLIDE_OK;
}


// ---------------------------------------------------------------- ensemble
static int ensemble_check_level(const glide_level& L, bool fp64, int l, int nl) {
    GLIDE_TRY(check_level(&L));
    REQ(L.u); REQ(L.v); REQ(L.H); REQ(L.phi); REQ(L.mask); REQ(L.bed); REQ(L.B); REQ(L.beta); REQ(L.work);
    REQ(L.f_u); REQ(L.f_v); REQ(L.f_H); REQ(L.r_u); REQ(L.r_v); REQ(L.r_H); REQ(L.gamma);
    REQ(L.lambda_u); REQ(L.lambda_v); REQ(L.lambda_H); REQ(L.af_u); REQ(L.af_v); REQ(L.af_H);
    if (!fp64) {
        REQ(L.F_u); REQ(L.F_v); REQ(L.F_H); REQ(L.w_u); REQ(L.w_v); REQ(L.w_H); REQ(L.y_u); REQ(L.y_v); REQ(L.y_H);
        REQ(L.z_u); REQ(L.z_v); REQ(L.z_H); REQ(L.chi); REQ(L.cphi); REQ(L.ar_u); REQ(L.ar_v); REQ(L.ar_H);
        REQ(L.vjp_u); REQ(L.vjp_v); REQ(L.vjp_H); REQ(L.wl_u); REQ(L.wl_v); REQ(L.wl_H); REQ(L.zl_u); REQ(L.zl_v);
        REQ(L.zl_H); REQ(L.adj_cache);
    } else {
        if (l == 0) { REQ(L.H_prev); REQ(L.smb); REQ(L.ar_u); REQ(L.ar_v); REQ(L.ar_H); }
        if (l == nl - 1) REQ(L.adj_cache);
    }
    if (!(L.phys.m == 1.0 && L.phys.u_reg > 0.0))
        return fail(GLIDE_ERR_INVALID_ARG, "the ensemble's mixed adjoint solves need linear drag (m = 1)");
    return GLIDE_OK;
}

int glide_ensemble_create(int32_t n_domains, const int32_t* n_levels, const glide_level* const* levels64,
                          const glide_level* const* levels32, const glide_fas_config* cfg,
                          const glide_fas_config* adjoint_cfg, const glide_vanka_config* forward_vc,
                          const glide_vanka_config* adjoint_vc, const glide_mixed_config* mc, int32_t grad_which,
                          glide_ensemble** out) {
    if (!out) return fail(GLIDE_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (n_domains < 1 || !n_levels || !levels64 || !levels32 || !cfg || !adjoint_cfg || !forward_vc || !adjoint_vc ||
        !mc)
        return fail(GLIDE_ERR_INVALID_ARG, "the ensemble needs at least one domain and every argument");
    if (mc->max_outer < 0 || mc->max_outer > 64 || mc->inner_vcycles < 1)
        return fail(GLIDE_ERR_INVALID_ARG, "max_outer must be in [0, 64] and inner_vcycles >= 1");
    for (const glide_fas_config* c : {cfg, adjoint_cfg})
        if (c->coarsest_steps < 0 || c->pre_steps < 0 || c->post_steps < 0 || c->finest_steps < 0)
            return fail(GLIDE_ERR_INVALID_ARG, "negative smoothing counts");
    if (grad_which & ~(15 | GLIDE_GRAD_EXACT)) return fail(GLIDE_ERR_INVALID_ARG, "unknown gradient flags");
    glide_ensemble* e = new glide_ensemble();
    for (int d = 0; d < n_domains; ++d) {
        const int nl = n_levels[d];
        if (nl < 1 || nl > 16 || !levels64[d] || !levels32[d]) {
            delete e;
            return fail(GLIDE_ERR_INVALID_ARG, "each domain needs 1 to 16 levels");
        }
        EnsDomain D;
        D.nl = nl;
        D.L64.assign(levels64[d], levels64[d] + nl);
        D.L32.assign(levels32[d], levels32[d] + nl);
        for (int l = 0; l < nl; ++l) {
            int rc = ensemble_check_level(D.L64[l], true, l, nl);
            if (rc == GLIDE_OK) rc = ensemble_check_level(D.L32[l], false, l, nl);
            if (rc == GLIDE_OK && (D.L32[l].ny != D.L64[l].ny || D.L32[l].nx != D.L64[l].nx))
                rc = fail(GLIDE_ERR_SHAPE, "FP32 and FP64 levels differ in shape");
            if (rc == GLIDE_OK && l > 0 &&
                (D.L64[l - 1].ny != 2 * D.L64[l].ny || D.L64[l - 1].nx != 2 * D.L64[l].nx))
                rc = fail(GLIDE_ERR_SHAPE, "each level must halve the one above it");
            if (rc == GLIDE_OK && (grad_which & 1) && l == 0 && !D.L64[0].g_beta) rc = fail(GLIDE_ERR_INVALID_ARG, "g_beta is NULL");
            if (rc == GLIDE_OK && (grad_which & 2) && l == 0 && !D.L64[0].g_bed) rc = fail(GLIDE_ERR_INVALID_ARG, "g_bed is NULL");
            if (rc == GLIDE_OK && (grad_which & 4) && l == 0 && !D.L64[0].g_H_prev) rc = fail(GLIDE_ERR_INVALID_ARG, "g_H_prev is NULL");
            if (rc == GLIDE_OK && (grad_which & 8) && l == 0 && !D.L64[0].g_smb) rc = fail(GLIDE_ERR_INVALID_ARG, "g_smb is NULL");
            if (rc != GLIDE_OK) {
                delete e;
                return rc;
            }
        }
        e->lmax = std::max(e->lmax, nl);
        e->dom.push_back(std::move(D));
    }
    e->cfg = *cfg;
    e->acfg = *adjoint_cfg;
    e->vf = *forward_vc;
    e->va = *adjoint_vc;
    e->mc = *mc;
    e->which = grad_which;
    e->rows.assign(e->lmax + 1, 4);
    if (hipGetDevice(&e->device) != hipSuccess) {
        delete e;
        return fail(GLIDE_ERR_HIP, "hipGetDevice");
    }
    *out = e;
    return GLIDE_OK;
}

int glide_ensemble_step(glide_ensemble* e, double dt, void* stream) {
    if (!e) return fail(GLIDE_ERR_INVALID_ARG, "ensemble is NULL");
    if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev != e->device) return fail(GLIDE_ERR_INVALID_ARG, "the ensemble belongs to another device");
    hipStream_t st = (hipStream_t)stream;
    if (e->dt != dt) {
        // a replay of the previous tables may still be running (on any stream
        // an earlier step was given)
        if (e->exec || e->dev) HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
        GLIDE_TRY(ensemble_record(e, dt));
    }
    if (!g_graphs_on || timing_on()) return ensemble_enqueue(e, st);
    if (!e->exec) {
        if (!e->cap) HIP_TRY(hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking), "hipStreamCreate(capture)");
        HIP_TRY(hipStreamBeginCapture(e->cap, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        const int rc = ensemble_enqueue(e, e->cap);
        hipGraph_t graph = nullptr;
        const hipError_t ee = hipStreamEndCapture(e->cap, &graph);
        if (rc != GLIDE_OK || ee != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            if (rc != GLIDE_OK) return rc;
            return hip_fail(ee, "hipStreamEndCapture");
        }
        const hipError_t ei = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            e->exec = nullptr;
            return hip_fail(ei, "hipGraphInstantiate");
        }
    }
    HIP_TRY(hipGraphLaunch(e->exec, st), "hipGraphLaunch");
    return GLIDE_OK;
}

int32_t glide_ensemble_launches(const glide_ensemble* e) { return e ? (int32_t)e->prog.ops.size() : -1; }

int32_t glide_ensemble_levels(const glide_ensemble* e) { return e ? (int32_t)e->lmax : -1; }

int glide_ensemble_set_tiles(glide_ensemble* e, const int32_t* rows, int32_t n) {
    if (!e || !rows) return fail(GLIDE_ERR_INVALID_ARG, "ensemble or rows is NULL");
    if (n != e->lmax + 1) return fail(GLIDE_ERR_INVALID_ARG, "rows needs one entry per level and one for the coarsest stage");
    for (int q = 0; q < n; ++q)
        if (!tab_vanka_mc_rows_ok(rows[q])) return fail(GLIDE_ERR_INVALID_ARG, "rows per block must be 1, 2, 4 or 8");
    e->rows.assign(rows, rows + n);
    e->dt = 0.0;  // recorded again (after the device drains) at the next step
    return GLIDE_OK;
}

int glide_ensemble_destroy(glide_ensemble* e) {
    if (!e) return GLIDE_OK;
    if (e->exec || e->dev) HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");  // no replay in flight
    ensemble_drop(e);
    if (e->cap) hipStreamDestroy(e->cap);
    delete e;
    return GLIDE_OK;
}

}  // extern "C"
