// C ABI of the enthalpy column solver and the fused thermal step.
//
// EnthalpyOperators.column_sweep (glide/enthalpy.py:426-471) runs here as one
// stream-ordered sequence: the initial residual, then per iteration the
// column smoother (E -> E_work), the layer smoother (E_work -> E) and the
// residual, each followed by a one-block reduction that records the norm and
// raises a device flag once the stopping test passes; every later launch
// reads the flag and returns at once.  The host never waits inside the
// sweep.  Buffers belong to the caller; the library owns a small per-device
// reduction scratch and status record (allocated by glide_enth_init, or by
// the first call on a device that needs them).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/glide_enthalpy.h"
#include "enthalpy_kernels.hpp"

namespace glide {
void set_last_error(const std::string& msg);
}

using namespace glide;

namespace {

constexpr double kSecPerYear = 365.25 * 86400.0;
constexpr double kGravity = 9.81;

int fail(int code, const std::string& msg) {
    set_last_error(msg);
    return code;
}

#define HIP_TRY(expr, where)                                                                  \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return fail(GLIDE_ERR_HIP, std::string(where) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define GLIDE_TRY(expr)                \
    do {                               \
        int _r = (expr);               \
        if (_r != GLIDE_OK) return _r; \
    } while (0)

struct EnthScratch {
    double* partials = nullptr;   // kEnthResBlocks
    double* max_out = nullptr;    // 1
    unsigned long long* acc = nullptr;  // adjoint passes' maximum and block count, kept zero between uses
    unsigned long long* rmax = nullptr;  // kEnthRmaxSlots: the column pass's max-|r| accumulators
    EnthStatus* status = nullptr;
    EnthStatus* host = nullptr;   // pinned mirror
};

std::mutex g_mutex;
std::vector<EnthScratch> g_scratch;

int scratch(EnthScratch** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_mutex);
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
    EnthScratch& s = g_scratch[dev];
    if (!s.partials) {
        HIP_TRY(hipMalloc(&s.partials, sizeof(double) * kEnthResBlocks), "hipMalloc(enthalpy partials)");
        HIP_TRY(hipMalloc(&s.max_out, sizeof(double)), "hipMalloc(enthalpy max)");
        HIP_TRY(hipMalloc(&s.acc, 2 * sizeof(unsigned long long)), "hipMalloc(enthalpy accumulator)");
        HIP_TRY(hipMemset(s.acc, 0, 2 * sizeof(unsigned long long)), "hipMemset(enthalpy accumulator)");
        HIP_TRY(hipMalloc(&s.status, sizeof(EnthStatus)), "hipMalloc(enthalpy status)");
        HIP_TRY(hipMalloc(&s.rmax, sizeof(unsigned long long) * kEnthRmaxSlots), "hipMalloc(enthalpy max slots)");
        HIP_TRY(hipHostMalloc(&s.host, sizeof(EnthStatus), hipHostMallocDefault), "hipHostMalloc(enthalpy status)");
    }
    *out = &s;
    return GLIDE_OK;
}

// A reported sweep whose residual norm went non-finite (NaN propagates through
// the max reductions and never passes the stopping test) is an error.
int check_report_finite(const glide_enth_report& rep) {
    if (!std::isfinite(rep.r0) || !std::isfinite(rep.r))
        return fail(GLIDE_ERR_NONFINITE, "non-finite enthalpy residual norm (the report holds the norms)");
    return GLIDE_OK;
}

typedef double R;
inline R* P(void* p) { return static_cast<R*>(p); }
inline const R* CP(const void* p) { return static_cast<const R*>(p); }
inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

int require(const void* p, const char* name) {
    if (!p) return fail(GLIDE_ERR_INVALID_ARG, std::string("required buffer is NULL: ") + name);
    return GLIDE_OK;
}
#define REQ(ptr) GLIDE_TRY(require((ptr), #ptr))

int check_column(const glide_enth_column* C) {
    if (!C) return fail(GLIDE_ERR_INVALID_ARG, "column grid is NULL");
    if (C->ny < 1 || C->nx < 1) return fail(GLIDE_ERR_SHAPE, "enthalpy grid must have ny >= 1 and nx >= 1");
    if (C->nz < 2 || C->nz > 64) return fail(GLIDE_ERR_SHAPE, "enthalpy columns need 2 <= nz <= 64");
    if (!(C->dx > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dx must be positive");
    if (C->term_flags & ~7) return fail(GLIDE_ERR_INVALID_ARG, "unknown enthalpy term flag");
    return GLIDE_OK;
}

// Operator arguments; `needs_velocity` for kernels that read u3d/v3d.
int make_args(const glide_enth_column* C, double dt, bool needs_state, EnthArgs<R>* a) {
    GLIDE_TRY(check_column(C));
    REQ(C->H);
    if (needs_state) {
        if (!(dt > 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "dt must be positive");
        REQ(C->E);
        REQ(C->E_prev);
        REQ(C->f_E);
        REQ(C->H_prev);
        REQ(C->E_surface);
        if (C->term_flags & GLIDE_ENTH_HORIZONTAL) {
            REQ(C->u3d);
            REQ(C->v3d);
        }
        if (C->term_flags & GLIDE_ENTH_OMEGA) REQ(C->omega);
    }
    a->ny = C->ny;
    a->nx = C->nx;
    a->nz = C->nz;
    a->flags = C->term_flags;
    a->dxi = 1.0 / C->dx;
    a->dt = dt;
    a->drain_rate = C->drain_rate;
    a->h_thin = C->h_thin;
    a->lf_sq = std::sqrt(1.0 + C->lf_c);
    a->E = CP(C->E);
    a->E_prev = CP(C->E_prev);
    a->f_E = CP(C->f_E);
    a->u3d = CP(C->u3d);
    a->v3d = CP(C->v3d);
    a->omega = (C->term_flags & GLIDE_ENTH_OMEGA) ? CP(C->omega) : nullptr;
    a->H = CP(C->H);
    a->H_prev = CP(C->H_prev);
    a->E_surface = CP(C->E_surface);
    a->Q_geo = CP(C->Q_geo);
    a->phi_strain = CP(C->phi_strain);
    a->skip = nullptr;
    a->u2d = nullptr;
    a->v2d = nullptr;
    a->vel_div = 1.0;
    return GLIDE_OK;
}

// The fused sweep with its reduction partials (kEnthResBlocks values) and
// status record in caller-chosen device memory; `host` is a pinned mirror of
// the record for the report, or nullptr (then the report is read through a
// stack copy).
int sweep_core(const glide_enth_column* C, double dt, int n_iter, const glide_enth_sweep_config* cfg,
               glide_enth_report* report, hipStream_t st, double* partials, EnthStatus* status, EnthStatus* host,
               const R* u2d, const R* v2d, double vel_div, unsigned long long* rmax = nullptr) {
    if (!cfg) return fail(GLIDE_ERR_INVALID_ARG, "sweep config is NULL");
    if (n_iter < 0 || n_iter > GLIDE_ENTH_MAX_SWEEPS)
        return fail(GLIDE_ERR_INVALID_ARG, "n_iter must be in [0, GLIDE_ENTH_MAX_SWEEPS]");
    if (cfg->n_newton < 0) return fail(GLIDE_ERR_INVALID_ARG, "n_newton must be non-negative");
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, dt, true, &A));
    A.u2d = u2d;
    A.v2d = v2d;
    A.vel_div = vel_div;
    REQ(C->r_E);
    if (cfg->alternating) REQ(C->E_work);
    else if (!cfg->multicolour) REQ(C->delta_E);
    const double rtol = cfg->relative_tolerance, atol = cfg->absolute_tolerance;
    if (C->sweep_work && u2d && rmax && cfg->alternating && !cfg->multicolour && cfg->n_newton >= 1 && n_iter >= 1 &&
        enth_lm_supported(C->nz)) {
        // The same sweep on level-major copies (sweep_work: E, E_prev, f_E,
        // omega planes; E_work holds the column pass's output in that layout).
        const long ncol = (long)C->ny * C->nx, plane = ncol * C->nz;
        R* wk = P(C->sweep_work);
        R *Elm = wk, *Eplm = wk + plane, *fElm = wk + 2 * plane, *omlm = wk + 3 * plane, *Wlm = P(C->E_work);
        const int nz = C->nz;
        HIP_TRY(launch_enth_lm_layout<R>(CP(C->E), Elm, ncol, nz, true, st), "enthalpy level-major copy");
        HIP_TRY(launch_enth_lm_layout<R>(CP(C->E_prev), Eplm, ncol, nz, true, st), "enthalpy level-major copy");
        HIP_TRY(launch_enth_lm_layout<R>(CP(C->f_E), fElm, ncol, nz, true, st), "enthalpy level-major copy");
        EnthArgs<R> L = A;
        L.E = Elm;
        L.E_prev = Eplm;
        L.f_E = fElm;
        if (A.omega) {
            HIP_TRY(launch_enth_lm_layout<R>(A.omega, omlm, ncol, nz, true, st), "enthalpy level-major copy");
            L.omega = omlm;
        }
        HIP_TRY(launch_enth_residual_lm<R>(L, nullptr, partials, status, 0, rtol, atol, st), "enthalpy residual");
        L.skip = &status->done;
        EnthArgs<R> Lw = L;
        Lw.E = Wlm;
        HIP_TRY(launch_enth_reduce_slots(rmax, status, 0, rtol, atol, st), "enthalpy max slots");
        for (int it = 0; it < n_iter; ++it) {
            HIP_TRY(launch_enth_column_lm<R>(L, Wlm, cfg->omega, cfg->n_newton, cfg->relaxation,
                                             it > 0 ? rmax : nullptr, st),
                    "enthalpy column smoother");
            if (it > 0) HIP_TRY(launch_enth_reduce_slots(rmax, status, 1, rtol, atol, st), "enthalpy sweep norm");
            HIP_TRY(launch_enth_layer_lm<R>(Lw, Elm, cfg->omega, st), "enthalpy layer smoother");
        }
        // the final iterate's residual into r_E even when the sweep stopped
        // early (the bookkeeping closes nothing then)
        EnthArgs<R> Lf = L;
        Lf.skip = nullptr;
        HIP_TRY(launch_enth_residual_lm<R>(Lf, P(C->r_E), partials, status, 1, rtol, atol, st), "enthalpy residual");
        HIP_TRY(launch_enth_lm_layout<R>(Elm, P(C->E), ncol, nz, false, st), "enthalpy layout");
    } else {
        HIP_TRY(launch_enth_residual<R>(A, P(C->r_E), 1, partials, nullptr, status, 0, rtol, atol, st),
                "enthalpy residual");
        EnthArgs<R> Ak = A;
        Ak.skip = &status->done;
        if (rmax && cfg->alternating && !cfg->multicolour && cfg->n_newton >= 1) {
            // The reference's sweep with the residual of each iterate taken from
            // the next column pass's first Newton rows (the same values the
            // residual launch computes): sweep it's norm is closed after pass
            // it + 1's column launch, whose layer launch then skips if the
            // sweep stopped, leaving E at the iterate the reference stops at.
            HIP_TRY(launch_enth_reduce_slots(rmax, status, 0, rtol, atol, st), "enthalpy max slots");
            EnthArgs<R> Aw = Ak;
            Aw.E = CP(C->E_work);
            for (int it = 0; it < n_iter; ++it) {
                HIP_TRY(launch_enth_column_smooth<R>(Ak, nullptr, P(C->E_work), cfg->omega, cfg->n_newton,
                                                     cfg->relaxation, st, -1, nullptr, it > 0 ? rmax : nullptr,
                                                     P(C->r_E)),
                        "enthalpy column smoother");
                if (it > 0) HIP_TRY(launch_enth_reduce_slots(rmax, status, 1, rtol, atol, st), "enthalpy sweep norm");
                HIP_TRY(launch_enth_layer_smooth<R>(Aw, nullptr, P(C->E), cfg->omega, st), "enthalpy layer smoother");
            }
            if (n_iter > 0)
                HIP_TRY(launch_enth_residual<R>(Ak, P(C->r_E), 1, partials, nullptr, status, 1, rtol, atol, st),
                        "enthalpy residual");
        } else
        for (int it = 0; it < n_iter; ++it) {
            if (cfg->multicolour && !cfg->alternating) {  // both colour phases in place on E
                for (int colour = 0; colour < 2; ++colour)
                    HIP_TRY(launch_enth_column_smooth<R>(Ak, nullptr, nullptr, cfg->omega, cfg->n_newton, cfg->relaxation,
                                                         st, colour, P(C->E)),
                            "enthalpy column smoother (red-black)");
            } else if (cfg->multicolour) {
                HIP_TRY(launch_enth_column_smooth<R>(Ak, nullptr, nullptr, cfg->omega, cfg->n_newton, cfg->relaxation, st,
                                                     0, P(C->E_work)),
                        "enthalpy column smoother (even columns)");
                EnthArgs<R> Aw = Ak;
                Aw.E = CP(C->E_work);
                HIP_TRY(launch_enth_column_smooth<R>(Aw, nullptr, nullptr, cfg->omega, cfg->n_newton, cfg->relaxation, st,
                                                     1, P(C->E_work)),
                        "enthalpy column smoother (odd columns)");
                HIP_TRY(launch_enth_layer_smooth<R>(Aw, nullptr, P(C->E), cfg->omega, st), "enthalpy layer smoother");
            } else if (cfg->alternating) {
                HIP_TRY(launch_enth_column_smooth<R>(Ak, nullptr, P(C->E_work), cfg->omega, cfg->n_newton,
                                                     cfg->relaxation, st),
                        "enthalpy column smoother");
                EnthArgs<R> Aw = Ak;
                Aw.E = CP(C->E_work);
                HIP_TRY(launch_enth_layer_smooth<R>(Aw, nullptr, P(C->E), cfg->omega, st), "enthalpy layer smoother");
            } else {
                HIP_TRY(launch_enth_column_smooth<R>(Ak, P(C->delta_E), nullptr, 0.0, cfg->n_newton, cfg->relaxation,
                                                     st),
                        "enthalpy column smoother");
                HIP_TRY(launch_enth_apply<R>(Ak, P(C->E), CP(C->delta_E), cfg->omega, st), "enthalpy update");
            }
            HIP_TRY(launch_enth_residual<R>(Ak, P(C->r_E), 1, partials, nullptr, status, 1, rtol, atol, st),
                    "enthalpy residual");
        }
    }
    if (report) {
        EnthStatus local;
        EnthStatus* h = host ? host : &local;
        HIP_TRY(hipMemcpyAsync(h, status, sizeof(EnthStatus), hipMemcpyDeviceToHost, st), "copy status");
        HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
        report->sweeps = h->sweeps;
        report->pad = 0;
        report->r0 = h->r0;
        report->r = h->r;
        std::memset(report->norms, 0, sizeof(report->norms));
        for (int q = 0; q < h->sweeps && q < GLIDE_ENTH_MAX_SWEEPS; ++q) report->norms[q] = h->norms[q];
        GLIDE_TRY(check_report_finite(*report));
    }
    return GLIDE_OK;
}

int sweep_impl(const glide_enth_column* C, double dt, int n_iter, const glide_enth_sweep_config* cfg,
               glide_enth_report* report, hipStream_t st, const R* u2d = nullptr, const R* v2d = nullptr,
               double vel_div = 1.0) {
    // argument errors are reported before any device work (or allocation)
    if (!cfg) return fail(GLIDE_ERR_INVALID_ARG, "sweep config is NULL");
    if (n_iter < 0 || n_iter > GLIDE_ENTH_MAX_SWEEPS)
        return fail(GLIDE_ERR_INVALID_ARG, "n_iter must be in [0, GLIDE_ENTH_MAX_SWEEPS]");
    EnthScratch* s = nullptr;
    GLIDE_TRY(scratch(&s));
    return sweep_core(C, dt, n_iter, cfg, report, st, s->partials, s->status, s->host, u2d, v2d, vel_div, s->rmax);
}

}  // namespace

namespace glide {
int enth_column_sweep_with_scratch(const glide_enth_column* C, double dt, int n_iter,
                                   const glide_enth_sweep_config* cfg, glide_enth_report* report, hipStream_t st,
                                   double* partials, EnthStatus* status) {
    return sweep_core(C, dt, n_iter, cfg, report, st, partials, status, nullptr, nullptr, nullptr, 1.0);
}
}  // namespace glide

extern "C" {

int glide_enth_init(void) {
    EnthScratch* s = nullptr;
    return scratch(&s);
}

int glide_enth_residual(const glide_enth_column* C, double dt, int use_forcing, double* max_abs, void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, dt, true, &A));
    REQ(C->r_E);
    EnthScratch* s = nullptr;
    GLIDE_TRY(scratch(&s));
    hipStream_t st = S(stream);
    HIP_TRY(launch_enth_residual<R>(A, P(C->r_E), use_forcing ? 1 : 0, s->partials, max_abs ? s->max_out : nullptr,
                                    nullptr, 0, 0.0, 0.0, st),
            "enthalpy residual");
    if (max_abs) {
        HIP_TRY(hipMemcpyAsync(&s->host->r, s->max_out, sizeof(double), hipMemcpyDeviceToHost, st), "copy max");
        HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
        *max_abs = s->host->r;
    }
    return GLIDE_OK;
}

int glide_enth_column_smooth(const glide_enth_column* C, double dt, int n_newton, double relaxation, void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, dt, true, &A));
    REQ(C->delta_E);
    if (n_newton < 0) return fail(GLIDE_ERR_INVALID_ARG, "n_newton must be non-negative");
    HIP_TRY(launch_enth_column_smooth<R>(A, P(C->delta_E), nullptr, 0.0, n_newton, relaxation, S(stream)),
            "enthalpy column smoother");
    return GLIDE_OK;
}

int glide_enth_layer_smooth(const glide_enth_column* C, double dt, void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, dt, true, &A));
    REQ(C->delta_E);
    HIP_TRY(launch_enth_layer_smooth<R>(A, P(C->delta_E), nullptr, 0.0, S(stream)), "enthalpy layer smoother");
    return GLIDE_OK;
}

int glide_enth_apply(const glide_enth_column* C, double omega, void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, 1.0, false, &A));
    REQ(C->E);
    REQ(C->delta_E);
    HIP_TRY(launch_enth_apply<R>(A, P(C->E), CP(C->delta_E), omega, S(stream)), "enthalpy update");
    return GLIDE_OK;
}

int glide_enth_elementary(const void* x, int64_t n, void* exp_out, void* log_out, void* stream) {
    REQ(x);
    if (n < 0) return fail(GLIDE_ERR_INVALID_ARG, "n must be non-negative");
    if (n == 0) return GLIDE_OK;
    HIP_TRY(launch_enth_elementary(CP(x), (long)n, P(exp_out), P(log_out), S(stream)), "enthalpy elementary functions");
    return GLIDE_OK;
}

int glide_enth_column_sweep(const glide_enth_column* C, double dt, int n_iter, const glide_enth_sweep_config* cfg,
                            glide_enth_report* report, void* stream) {
    return sweep_impl(C, dt, n_iter, cfg, report, S(stream));
}

int glide_enth_omega(const glide_enth_column* C, const void* dh_dt, const void* bmb, double divide_by, void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, 1.0, false, &A));
    REQ(C->u3d);
    REQ(C->v3d);
    REQ(C->omega);
    REQ(dh_dt);
    if (divide_by == 0.0) return fail(GLIDE_ERR_INVALID_ARG, "divide_by must be non-zero");
    HIP_TRY(launch_enth_omega<R>(A, CP(dh_dt), CP(bmb), divide_by, P(C->omega), S(stream)), "enthalpy omega");
    return GLIDE_OK;
}

int glide_enth_broadcast_velocity(const glide_enth_column* C, const void* u2d, const void* v2d, double divide_by,
                                  void* stream) {
    GLIDE_TRY(check_column(C));
    REQ(C->u3d);
    REQ(C->v3d);
    REQ(u2d);
    REQ(v2d);
    if (divide_by == 0.0) return fail(GLIDE_ERR_INVALID_ARG, "divide_by must be non-zero");
    const long nu = (long)C->ny * (C->nx + 1), nv = (long)(C->ny + 1) * C->nx;
    HIP_TRY(launch_enth_broadcast<R>(CP(u2d), P(C->u3d), nu, C->nz, divide_by, S(stream)), "broadcast u");
    HIP_TRY(launch_enth_broadcast<R>(CP(v2d), P(C->v3d), nv, C->nz, divide_by, S(stream)), "broadcast v");
    return GLIDE_OK;
}

int glide_enth_frictional_heating(const glide_enth_column* C, const void* beta, double m, double u_reg,
                                  void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, 1.0, false, &A));
    REQ(C->u3d);
    REQ(C->v3d);
    REQ(C->Q_fh);
    REQ(beta);
    HIP_TRY(launch_enth_rhs<R>(A, CP(beta), m, u_reg, 1, P(C->Q_fh), nullptr, S(stream)), "frictional heating");
    return GLIDE_OK;
}

int glide_enth_strain_heating(const glide_enth_column* C, const void* u2d, const void* v2d, const void* B,
                              double n_glen, double eps_reg, double rho_i, void* stream) {
    GLIDE_TRY(check_column(C));
    REQ(C->phi_strain);
    REQ(u2d);
    REQ(v2d);
    REQ(B);
    if (!(n_glen > 0.0) || !(eps_reg >= 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "n_glen > 0 and eps_reg >= 0 required");
    HIP_TRY(launch_enth_strain_heating<R>(CP(u2d), CP(v2d), CP(B), C->ny, C->nx, C->nz, 1.0 / C->dx, n_glen, eps_reg,
                                          rho_i * kGravity, P(C->phi_strain), S(stream)),
            "strain heating");
    return GLIDE_OK;
}

int glide_enth_set_rhs(const glide_enth_column* C, void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, 1.0, false, &A));
    REQ(C->f_E);
    REQ(C->Q_geo);
    REQ(C->Q_fh);
    HIP_TRY(launch_enth_rhs<R>(A, nullptr, 0.0, 0.0, 0, P(C->Q_fh), P(C->f_E), S(stream)), "enthalpy forcing");
    return GLIDE_OK;
}

int glide_enth_arrhenius(const glide_enth_column* C, double n_glen, double scale, void* B, void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, 1.0, false, &A));
    REQ(C->E);
    REQ(B);
    if (!(n_glen > 0.0) || scale == 0.0) return fail(GLIDE_ERR_INVALID_ARG, "n_glen must be positive, scale non-zero");
    HIP_TRY(launch_enth_arrhenius<R>(A, n_glen, scale, P(B), S(stream)), "rate factor");
    return GLIDE_OK;
}

int glide_enth_diagnostics(const glide_enth_column* C, void* temperature, void* water, void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, 1.0, false, &A));
    REQ(C->E);
    HIP_TRY(launch_enth_diagnostics<R>(A, P(temperature), P(water), S(stream)), "enthalpy diagnostics");
    return GLIDE_OK;
}

int glide_enth_from_temperature(const glide_enth_column* C, const void* temperature, void* E_out, void* stream) {
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, 1.0, false, &A));
    REQ(temperature);
    REQ(E_out);
    HIP_TRY(launch_enth_from_temperature<R>(A, CP(temperature), P(E_out), S(stream)), "enthalpy from temperature");
    return GLIDE_OK;
}

int glide_enth_surface_from_temperature(const glide_enth_column* C, const void* T_surface, void* E_surface_out,
                                        void* stream) {
    GLIDE_TRY(check_column(C));
    REQ(T_surface);
    REQ(E_surface_out);
    HIP_TRY(launch_enth_surface_from_temperature<R>(CP(T_surface), P(E_surface_out), (long)C->ny * C->nx, S(stream)),
            "surface enthalpy");
    return GLIDE_OK;
}

int glide_enth_adjoint_solve(const glide_enth_column* C, const glide_enth_adjoint* X, double dt, const void* dL_dE,
                             const void* dL_dB, double n_glen, double B_scale, int n_iter,
                             const glide_enth_sweep_config* cfg, glide_enth_report* report, void* stream) {
    if (!X || !cfg) return fail(GLIDE_ERR_INVALID_ARG, "adjoint buffers or config is NULL");
    if (n_iter < 0 || n_iter > GLIDE_ENTH_MAX_SWEEPS)
        return fail(GLIDE_ERR_INVALID_ARG, "n_iter must be in [0, GLIDE_ENTH_MAX_SWEEPS]");
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, dt, true, &A));
    REQ(X->rows);
    REQ(X->g);
    REQ(X->lambda);
    REQ(X->lambda_work);
    if (dL_dB && (!(n_glen > 0.0) || B_scale == 0.0))
        return fail(GLIDE_ERR_INVALID_ARG, "n_glen must be positive and B_scale non-zero");
    EnthScratch* s = nullptr;
    GLIDE_TRY(scratch(&s));
    hipStream_t st = S(stream);
    const int ny = C->ny, nx = C->nx, nz = C->nz;
    const bool flat = X->u2d || X->v2d;
    if (flat) {
        REQ(X->u2d);
        REQ(X->v2d);
        if (!(X->vel_div != 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "vel_div must be non-zero");
        A.u2d = CP(X->u2d);
        A.v2d = CP(X->v2d);
        A.vel_div = X->vel_div;
    }
    // factors, right-hand side and multipliers in the checkerboard layout
    HIP_TRY(launch_enth_adj_factor<R>(A, P(X->rows), st), "adjoint factors");
    HIP_TRY(launch_enth_adj_rhs<R>(A, CP(dL_dE), CP(dL_dB), n_glen, B_scale, P(X->g), st), "adjoint rhs");
    HIP_TRY(launch_enth_adj_layout<R>(ny, nx, nz, CP(X->lambda), P(X->lambda_work), true, st), "adjoint layout");
    const double rtol = cfg->relative_tolerance, atol = cfg->absolute_tolerance;
    HIP_TRY(launch_enth_adj_residual<R>(nullptr, A, flat, CP(X->rows), CP(X->g), CP(X->lambda_work), s->acc,
                                        s->status, 0, rtol, atol, st),
            "adjoint residual");
    const int* skip = &s->status->done;
    // red-black column Gauss-Seidel sweeps; each records the residual of the
    // columns before their update and stops (device-side) at the tolerance
    for (int it = 0; it < n_iter; ++it)
        HIP_TRY(launch_enth_adj_sweep_rb<R>(skip, A, flat, CP(X->rows), CP(X->g), P(X->lambda_work), cfg->omega,
                                            s->acc, s->status, rtol, atol, st),
                "adjoint sweep");
    HIP_TRY(launch_enth_adj_layout<R>(ny, nx, nz, CP(X->lambda_work), P(X->lambda), false, st), "adjoint layout");
    HIP_TRY(launch_enth_adj_grads<R>(A, CP(X->lambda), P(X->g_E_prev), P(X->g_phi_strain), P(X->g_Q_geo),
                                     P(X->g_E_surface), st),
            "adjoint gradients");
    if (report) {
        HIP_TRY(hipMemcpyAsync(s->host, s->status, sizeof(EnthStatus), hipMemcpyDeviceToHost, st), "copy status");
        HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
        report->sweeps = s->host->sweeps;
        report->pad = 0;
        report->r0 = s->host->r0;
        report->r = s->host->r;
        std::memset(report->norms, 0, sizeof(report->norms));
        for (int q = 0; q < s->host->sweeps && q < GLIDE_ENTH_MAX_SWEEPS; ++q) report->norms[q] = s->host->norms[q];
        GLIDE_TRY(check_report_finite(*report));
    }
    return GLIDE_OK;
}

int glide_thermal_step(const glide_enth_column* C, const void* u2d, const void* v2d, const void* beta, void* B,
                       const glide_thermal_config* tc, const glide_enth_sweep_config* sc, glide_enth_report* report,
                       void* stream) {
    if (!tc || !sc) return fail(GLIDE_ERR_INVALID_ARG, "thermal or sweep config is NULL");
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, tc->dt, true, &A));
    REQ(C->u3d);
    REQ(C->v3d);
    REQ(C->omega);
    REQ(C->dh_dt);
    REQ(C->Q_geo);
    REQ(C->Q_fh);
    if (tc->frictional_heating) REQ(beta);
    if (tc->update_rheology || tc->strain_heating) {
        REQ(B);
        if (!(tc->n_glen > 0.0) || !(tc->rho_i > 0.0))
            return fail(GLIDE_ERR_INVALID_ARG, "n_glen and rho_i must be positive");
    }
    hipStream_t st = S(stream);
    // glide/model.py:180-197: layer velocities in m/a, omega from the realised
    // thickness change in m/a, then both converted to m/s.  The velocity is
    // the same on every level, so the kernels read the SSA planes (divided as
    // the broadcast is); u3d / v3d are still filled in m/s for the caller.
    const double dt_yr = tc->dt / kSecPerYear;
    HIP_TRY(launch_enth_tendency<R>(CP(C->H), CP(C->H_prev), dt_yr, P(C->dh_dt), (long)C->ny * C->nx, st),
            "thickness tendency");
    EnthArgs<R> A2 = A;
    A2.u2d = CP(u2d);
    A2.v2d = CP(v2d);
    A2.vel_div = 1.0;
    HIP_TRY(launch_enth_omega<R>(A2, CP(C->dh_dt), nullptr, kSecPerYear, P(C->omega), st), "enthalpy omega");
    GLIDE_TRY(glide_enth_broadcast_velocity(C, u2d, v2d, kSecPerYear, stream));
    A2.vel_div = kSecPerYear;
    if (tc->strain_heating)
        GLIDE_TRY(glide_enth_strain_heating(C, u2d, v2d, B, tc->n_glen, tc->eps_reg, tc->rho_i, stream));
    // glide/model.py:199-203, glide/enthalpy.py:473-519
    HIP_TRY(launch_enth_rhs<R>(A2, CP(beta), tc->m, tc->u_reg, tc->frictional_heating ? 1 : 0, P(C->Q_fh),
                               P(C->f_E), st),
            "enthalpy forcing");
    GLIDE_TRY(sweep_impl(C, tc->dt, tc->n_smooth, sc, report, st, CP(u2d), CP(v2d), kSecPerYear));
    // glide/model.py:205-213
    if (tc->update_rheology) {
        const double scale = tc->rho_i * kGravity * std::pow(kSecPerYear, 1.0 / tc->n_glen);
        HIP_TRY(launch_enth_arrhenius<R>(A, tc->n_glen, scale, P(B), st), "rate factor");
    }
    return GLIDE_OK;
}

int64_t glide_thermal_vjp_work_doubles(int ny, int nx) {
    if (ny < 1 || nx < 1) return -1;
    return (int64_t)kThermalVjpPlanes * ny * nx;
}

int glide_thermal_step_vjp(const glide_enth_column* C, const void* u2d, const void* v2d, const void* beta,
                           const void* B, const glide_thermal_config* tc, const glide_thermal_vjp* V, void* stream) {
    if (!tc || !V) return fail(GLIDE_ERR_INVALID_ARG, "thermal config or gradient buffers is NULL");
    EnthArgs<R> A;
    GLIDE_TRY(make_args(C, tc->dt, true, &A));
    REQ(u2d);
    REQ(v2d);
    REQ(V->lambda);
    REQ(V->work);
    REQ(V->g_u);
    REQ(V->g_v);
    REQ(V->g_H);
    if (tc->frictional_heating) {
        REQ(beta);
        REQ(V->g_Q);
    }
    if (tc->strain_heating) {
        REQ(B);
        REQ(V->g_phi_strain);
        REQ(C->phi_strain);
    }
    if (V->dL_dB && !tc->update_rheology)
        return fail(GLIDE_ERR_INVALID_ARG, "dL_dB needs update_rheology (the step returns no rate factor)");
    if ((V->dL_dB || tc->strain_heating) && (!(tc->n_glen > 0.0) || !(tc->rho_i > 0.0)))
        return fail(GLIDE_ERR_INVALID_ARG, "n_glen and rho_i must be positive");
    if (tc->strain_heating && !(tc->eps_reg >= 0.0)) return fail(GLIDE_ERR_INVALID_ARG, "eps_reg must be >= 0");
    A.u2d = CP(u2d);
    A.v2d = CP(v2d);
    A.vel_div = kSecPerYear;
    ThermalVjpArgs<R> W{};
    W.lam = CP(V->lambda);
    W.g_phi = tc->strain_heating ? CP(V->g_phi_strain) : nullptr;
    W.g_Q = tc->frictional_heating ? CP(V->g_Q) : nullptr;
    W.gB = (tc->update_rheology && V->dL_dB) ? CP(V->dL_dB) : nullptr;
    W.B_in = tc->strain_heating ? CP(B) : nullptr;
    W.beta = tc->frictional_heating ? CP(beta) : nullptr;
    W.n_glen = tc->n_glen;
    W.B_scale = tc->rho_i * kGravity * std::pow(kSecPerYear, 1.0 / tc->n_glen);
    W.glen_p = (1.0 - tc->n_glen) / (2.0 * tc->n_glen);
    W.eps_reg = tc->eps_reg;
    W.rho_g = tc->rho_i * kGravity;
    W.m = tc->m;
    W.u_reg = tc->u_reg;
    W.dt_yr = tc->dt / kSecPerYear;
    W.work = P(V->work);
    W.g_u = P(V->g_u);
    W.g_v = P(V->g_v);
    W.g_H = P(V->g_H);
    W.g_H_prev = P(V->g_H_prev);
    W.g_beta = tc->frictional_heating ? P(V->g_beta) : nullptr;
    W.g_B = tc->strain_heating ? P(V->g_B) : nullptr;
    HIP_TRY(launch_thermal_vjp<R>(A, W, S(stream)), "thermal step input gradients");
    return GLIDE_OK;
}

}  // extern "C"
