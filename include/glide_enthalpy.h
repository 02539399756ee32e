/*
 * glide_enthalpy.h — C ABI of the MI355X (gfx950) enthalpy column solver
 * coupled to the SSA step: residual, column Newton smoother, layer smoother,
 * the fused column sweep, the omega integration, forcing, the rate-factor
 * feedback to the momentum solver, and one fused thermal step.
 *
 * Same conventions as glide_hip.h: caller-owned FP64 device buffers,
 * row-major and contiguous; stream passed as void*; 0 on success, a negative
 * glide_status on failure (text from glide_last_error()).  Layouts:
 *   E, E_prev, f_E, omega, phi_strain, r_E, delta_E, E_work: (ny, nx, nz)
 *   u3d (nz, ny, nx+1), v3d (nz, ny+1, nx)
 *   H, H_prev, E_surface, Q_geo, Q_fh, dh_dt: (ny, nx)
 * Enthalpy is scaled by E_SCALE = C_I (T_MELT - T_REF); 2 <= nz <= 64.
 *
 * Reference interfaces replaced (JacobDowns/glide):
 *   glide_enth_residual          <- EnthalpyOperators.compute_residual / kernel enthalpy_compute_residual
 *                                   (glide/enthalpy.py:341-364, glide/cuda/enthalpy.cu:783-905)
 *   glide_enth_column_smooth     <- EnthalpyOperators.column_smooth / kernel enthalpy_column_smooth
 *                                   (glide/enthalpy.py:366-393, glide/cuda/enthalpy.cu:932-1112)
 *   glide_enth_layer_smooth      <- EnthalpyOperators.layer_smooth / kernel enthalpy_layer_smooth
 *                                   (glide/enthalpy.py:395-424, glide/cuda/enthalpy.cu:1128-1255)
 *   glide_enth_apply             <- the update E += omega delta_E in column_sweep (glide/enthalpy.py:452-457)
 *   glide_enth_column_sweep      <- EnthalpyOperators.column_sweep (glide/enthalpy.py:426-471)
 *   glide_enth_omega             <- EnthalpyOperators.compute_omega / kernel compute_omega
 *                                   (glide/enthalpy.py:303-339, glide/cuda/enthalpy.cu:1276-1372)
 *   glide_enth_broadcast_velocity <- EnthalpyOperators.broadcast_velocity and the m/a -> m/s
 *                                   conversion of ThermalModel.step (glide/enthalpy.py:289-301,
 *                                   glide/model.py:180-197)
 *   glide_enth_frictional_heating <- ThermalModel._compute_frictional_heating (glide/model.py:215-228)
 *   glide_enth_set_rhs           <- EnthalpyOperators.set_rhs after its optional snapshot
 *                                   (glide/enthalpy.py:473-519)
 *   glide_enth_arrhenius         <- EnthalpyOperators.get_arrhenius_factor and the head-unit scale of
 *                                   ThermalModel.step (glide/enthalpy.py:600-643, glide/model.py:205-213)
 *   glide_enth_diagnostics       <- EnthalpyOperators.get_temperature / get_water_content
 *                                   (glide/enthalpy.py:565-598)
 *   glide_enth_from_temperature  <- EnthalpyOperators.initialize_from_temperature (glide/enthalpy.py:534-563)
 *   glide_enth_surface_from_temperature
 *                                <- EnthalpyOperators.set_surface_enthalpy_from_temperature
 *                                   (glide/enthalpy.py:521-532)
 *   glide_enth_strain_heating    <- no counterpart (the reference's phi_strain is caller-supplied,
 *                                   glide/enthalpy.py:108-131); SSA dissipation for the coupled step
 *   glide_thermal_step           <- ThermalModel.step (glide/model.py:167-213)
 *   glide_enth_adjoint_solve     <- no counterpart: adjoint of the converged enthalpy solve
 */
#ifndef GLIDE_ENTHALPY_H
#define GLIDE_ENTHALPY_H

#include "glide_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GLIDE_ENTH_HORIZONTAL 1 /* horizontal Lax-Friedrichs advection */
#define GLIDE_ENTH_OMEGA 2      /* sigma advection by omega */
#define GLIDE_ENTH_DRAINAGE 4   /* temperate-ice drainage */
#define GLIDE_ENTH_MAX_SWEEPS 256

/* One grid of columns (EnthalpyOperators state, forcing and work arrays). */
typedef struct glide_enth_column {
    int32_t ny, nx, nz;
    int32_t term_flags; /* GLIDE_ENTH_* bits (EnthalpyTermFlags.bitmask) */
    double dx;          /* m */
    double drain_rate;  /* s^-1 (EnthalpyForcing.drain_rate) */
    double h_thin;      /* m: thinner columns are clamped to the surface enthalpy */
    double lf_c;        /* Lax-Friedrichs dissipation parameter (ColumnSmootherConfig.lf_c) */
    void *E, *E_prev, *f_E, *omega, *phi_strain, *r_E, *delta_E, *E_work;
    void *u3d, *v3d;
    void *H, *H_prev, *E_surface, *Q_geo, *Q_fh, *dh_dt;
    /* Optional (no counterpart in the reference), NULL for none: 4 nz ny nx
     * doubles of workspace.  With it, a fused alternating sweep with a
     * depth-constant velocity (glide_thermal_step) and 9 <= nz <= 64 runs on
     * level-major copies of E, E_prev, f_E and omega (E_work then holds the
     * column pass's output in that layout); results equal the column-major
     * sweep's to rounding. */
    void *sweep_work;
} glide_enth_column;

/* ColumnSmootherConfig (glide/enthalpy.py:183-198) minus the Python hook. */
typedef struct glide_enth_sweep_config {
    double omega;      /* relaxation of each sweep's update */
    double relaxation; /* Newton step relaxation */
    double relative_tolerance, absolute_tolerance;
    int32_t n_newton;
    int32_t alternating; /* follow each column sweep by a layer sweep */
    /* (no counterpart in the reference) red-black column ordering: the
     * columns with i + j even are smoothed against the previous iterate, then
     * the odd ones against the even ones' new values (block Gauss-Seidel over
     * columns; fewer sweeps), followed by the layer sweep when alternating is
     * set, in place on E (no E_work) when it is not.  Near the pressure-
     * melting point the discrete system can have more than one solution, and
     * this order can settle on a different one than the reference's. */
    int32_t multicolour;
} glide_enth_sweep_config;

typedef struct glide_enth_report {
    int32_t sweeps; /* sweeps taken */
    int32_t pad;
    double r0, r;   /* max |r| before the first and after the last sweep */
    double norms[GLIDE_ENTH_MAX_SWEEPS];
} glide_enth_report;

/* ThermalModel settings (glide/model.py:97-132); strain_heating and eps_reg
 * have no counterpart in the reference (see glide_enth_strain_heating). */
typedef struct glide_thermal_config {
    double dt;       /* s */
    double m, u_reg; /* sliding exponent and regularisation for Q_fh */
    double rho_i;    /* ice density of the B unit conversion and strain heating */
    double n_glen;
    double eps_reg;  /* strain-rate regularisation of the momentum solver's viscosity (a^-2) */
    int32_t n_smooth;
    int32_t frictional_heating;
    int32_t update_rheology;
    int32_t strain_heating; /* phi_strain from the SSA velocities and B before the sweep */
} glide_thermal_config;

/* Allocates this library's enthalpy scratch (reduction partials and the sweep
 * status record) on the current device, so that later calls allocate nothing
 * (no counterpart in the reference; glide_init calls it). */
int glide_enth_init(void);
/* r_E = F(E) - f_E (use_forcing != 0) or F(E); when max_abs is non-NULL the
 * stream is synchronised and max |r_E| is stored there. */
int glide_enth_residual(const glide_enth_column* C, double dt, int use_forcing, double* max_abs, void* stream);
/* delta_E = column Newton solve (n_newton steps, relaxation) minus E. */
int glide_enth_column_smooth(const glide_enth_column* C, double dt, int n_newton, double relaxation, void* stream);
/* delta_E = -r / J_kk. */
int glide_enth_layer_smooth(const glide_enth_column* C, double dt, void* stream);
/* E += omega delta_E. */
int glide_enth_apply(const glide_enth_column* C, double omega, void* stream);
/* Up to n_iter sweeps (column smoother, then layer smoother when alternating),
 * stopping after the first sweep whose max |r| is below the relative or the
 * absolute tolerance.  The stopping test runs on the device: no host
 * synchronisation unless `report` is non-NULL (then one, at the end).  A NaN
 * residual is carried through the max-norm, never passes the stopping test
 * and, with a report, returns GLIDE_ERR_NONFINITE after filling the report.
 * Uses r_E and E_work (alternating) or delta_E (not alternating). */
int glide_enth_column_sweep(const glide_enth_column* C, double dt, int n_iter, const glide_enth_sweep_config* cfg,
                            glide_enth_report* report, void* stream);
/* glide_enth_column_sweep without any device allocation.  It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  Its reduction partials and sweep record live in a file-scope
 * device array of GLIDE_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends; n_iter beyond the record's GLIDE_ENTH_MAX_SWEEPS norm slots is
 * rejected.  Every call on one GPU shares that array (calls that overlap on
 * one GPU race); the results equal glide_enth_column_sweep's.  The report,
 * when requested, is read through a synchronous copy into `report`. */
#define GLIDE_NOALLOC_SCRATCH_BYTES 20480
int glide_enth_column_sweep_noalloc(const glide_enth_column* C, double dt, int n_iter,
                                    const glide_enth_sweep_config* cfg, glide_enth_report* report, void* stream);
/* omega = H sigma_dot from dh_dt and the layer velocities (same units), with
 * basal melt bmb (NULL = 0), divided by divide_by. */
int glide_enth_omega(const glide_enth_column* C, const void* dh_dt, const void* bmb, double divide_by, void* stream);
/* u3d[k] = u2d / divide_by, v3d[k] = v2d / divide_by for every level k. */
int glide_enth_broadcast_velocity(const glide_enth_column* C, const void* u2d, const void* v2d, double divide_by,
                                  void* stream);
/* Q_fh = beta (|u_b|^2 + u_reg^2)^((m+1)/2) from the bed-level velocities. */
int glide_enth_frictional_heating(const glide_enth_column* C, const void* beta, double m, double u_reg,
                                  void* stream);
/* phi_strain[k] = 4 eta eps_e^2 rho_i g / a (W/m^3) on every level from the
 * depth-averaged velocities u2d, v2d (m/a) and B (head units), with eps_e^2
 * and eta = B/2 (eps_e^2 + eps_reg)^((1-n)/(2n)) the momentum solver's cell
 * strain-rate invariant and viscosity.  No counterpart in the reference,
 * which leaves phi_strain to the caller. */
int glide_enth_strain_heating(const glide_enth_column* C, const void* u2d, const void* v2d, const void* B,
                              double n_glen, double eps_reg, double rho_i, void* stream);
/* f_E from phi_strain (NULL = 0), Q_geo and Q_fh. */
int glide_enth_set_rhs(const glide_enth_column* C, void* stream);
/* B = (mean over levels of A(T)(1 + 181.25 w))^(-1/n_glen) / scale. */
int glide_enth_arrhenius(const glide_enth_column* C, double n_glen, double scale, void* B, void* stream);
/* Temperature and/or water content (either may be NULL), (ny, nx, nz). */
int glide_enth_diagnostics(const glide_enth_column* C, void* temperature, void* water, void* stream);
/* E_out = scaled enthalpy of the (ny, nx, nz) temperature capped at T_pmp. */
int glide_enth_from_temperature(const glide_enth_column* C, const void* temperature, void* E_out, void* stream);
/* E_surface_out = scaled enthalpy of the (ny, nx) temperature capped at T_MELT. */
int glide_enth_surface_from_temperature(const glide_enth_column* C, const void* T_surface, void* E_surface_out,
                                        void* stream);
/* Adjoint of the converged solve (no counterpart in the reference): buffers.
 * The workspaces hold checkerboard planes of GLIDE_ENTH_ADJ_PLANE doubles:
 * node (i, j, k) of colour c = (i + j) % 2 at ((c * nz + k) * ny + i) *
 * ((nx + 1) / 2) + j / 2. */
#define GLIDE_ENTH_ADJ_PLANE(ny, nx, nz) (2L * (long)(nz) * (long)(ny) * (long)(((nx) + 1) / 2))
typedef struct glide_enth_adjoint {
    void* rows;        /* workspace, 7 planes: Thomas factors of the transposed column blocks, couplings */
    void* g;           /* workspace, 1 plane: right-hand side */
    void* lambda;      /* (ny, nx, nz): initial guess in, adjoint state out */
    void* lambda_work; /* workspace, 1 plane: the adjoint state during the sweeps */
    void* r;           /* reserved, may be NULL */
    void *g_E_prev, *g_phi_strain; /* (ny, nx, nz) gradients, either may be NULL */
    void *g_Q_geo, *g_E_surface;   /* (ny, nx) gradients (g_Q_geo is also dL/dQ_fh), either may be NULL */
    /* Optional depth-constant velocity (no counterpart in the reference): when
     * u2d (ny, nx+1) and v2d (ny+1, nx) are set, level k's velocity is
     * u2d / vel_div (the values u3d[k] holds in the coupled step; u3d and v3d
     * are then not read) and the horizontal couplings are stored once per
     * column instead of per level. */
    const void *u2d, *v2d;
    double vel_div;
} glide_enth_adjoint;

/* Adjoint of the thermal solve at the converged E of C: solves J^T lambda = g
 * with J the exact Jacobian of the system the sweeps solve (interior rows,
 * Dirichlet surface node, clamped thin columns) and g = dL/dE + (dB/dE)^T
 * dL/dB (dL_dB may be NULL; B = rate factor / B_scale), by red-black column
 * Gauss-Seidel sweeps (each column's transposed tridiagonal block solved
 * exactly by one thread from factors computed once, relaxation cfg->omega) for up to n_iter sweeps with the device-side
 * stopping test on max |g - J^T lambda| of the columns before their update
 * (cfg tolerances); then writes the parameter gradients. */
int glide_enth_adjoint_solve(const glide_enth_column* C, const glide_enth_adjoint* X, double dt, const void* dL_dE,
                             const void* dL_dB, double n_glen, double B_scale, int n_iter,
                             const glide_enth_sweep_config* cfg, glide_enth_report* report, void* stream);

/* No counterpart in the reference: the exponential and logarithm the
 * enthalpy rows use (written for the rows' argument ranges), over n doubles:
 * exp_out = e^x, log_out = log x for x > 0 (NaN otherwise); either output may
 * be NULL.  For tests of their accuracy against the general routines. */
int glide_enth_elementary(const void* x, int64_t n, void* exp_out, void* log_out, void* stream);

/* ThermalModel.step after pre_momentum: velocities u2d, v2d in m/a and H the
 * new thickness; fills u3d, v3d (m/s), dh_dt, omega, Q_fh (when
 * frictional_heating), phi_strain (when strain_heating, from the B given),
 * f_E, runs the column sweep and, when update_rheology, overwrites B with
 * the new rate factor (head units). */
int glide_thermal_step(const glide_enth_column* C, const void* u2d, const void* v2d, const void* beta, void* B,
                       const glide_thermal_config* tc, const glide_enth_sweep_config* sc, glide_enth_report* report,
                       void* stream);

/* Gradients of the converged thermal step with respect to the inputs the
 * momentum step hands it (no counterpart in the reference, whose enthalpy
 * solve has no adjoint; with the momentum adjoint it closes the adjoint of
 * the coupled step of glide/model.py:156-213): dL/dp = -lambda^T dG/dp for
 * the velocities u2d (ny, nx+1) and v2d (ny+1, nx) in m/a, the new and
 * previous thicknesses H and H_prev, the sliding coefficient beta (through the
 * frictional heating) and the rate factor B the strain heating read, with
 * lambda and the phi_strain / Q_fh gradients from glide_enth_adjoint_solve at
 * the same state.  Every input reaches G directly (time term, 1/H and the
 * pressure-melting enthalpy in the diffusion and drainage, H phi_strain
 * forcing, facet mass fluxes of the horizontal advection) and through omega
 * (thickness tendency and flux divergence), Q_fh and phi_strain; dL_dB (the
 * seed of the new rate factor, when update_rheology) adds its dependence on
 * H through the pressure-melting temperature.  Buffers: */
typedef struct glide_thermal_vjp {
    const void* lambda;       /* (ny, nx, nz) adjoint state */
    const void* g_phi_strain; /* (ny, nx, nz) from the same adjoint solve; read when tc->strain_heating */
    const void* g_Q;          /* (ny, nx) g_Q_geo of the same solve (= dL/dQ_fh); read when tc->frictional_heating */
    const void* dL_dB;        /* (ny, nx) the solve's rate-factor seed, or NULL */
    void* work;               /* glide_thermal_vjp_work_doubles(ny, nx) doubles */
    void *g_u, *g_v, *g_H;    /* outputs (ny, nx+1), (ny+1, nx), (ny, nx) */
    void *g_H_prev;           /* output (ny, nx), may be NULL */
    void *g_beta;             /* output (ny, nx) when tc->frictional_heating, may be NULL */
    void *g_B;                /* output (ny, nx) when tc->strain_heating, may be NULL */
} glide_thermal_vjp;
/* Work doubles glide_thermal_step_vjp needs at (ny, nx), or -1 for invalid sizes. */
int64_t glide_thermal_vjp_work_doubles(int ny, int nx);
/* C holds the converged state of glide_thermal_step (E, E_prev, H, H_prev,
 * omega, phi_strain) with the velocities u2d, v2d, beta and the rate factor B
 * that step read (before its update), and tc its configuration. */
int glide_thermal_step_vjp(const glide_enth_column* C, const void* u2d, const void* v2d, const void* beta,
                           const void* B, const glide_thermal_config* tc, const glide_thermal_vjp* V, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GLIDE_ENTHALPY_H */
