/*
 * glide_hip.h — C ABI of the MI355X (gfx950) implementation of GLIDE's coupled
 * SSA momentum + mass step (FAS constrained-descent multigrid with additive 5x5
 * Vanka smoothing), its adjoint, and the parameter gradients.
 *
 * Every buffer is caller-owned device memory (row-major, contiguous, FP64):
 *   u-facets (ny, nx+1), v-facets (ny+1, nx), cells (ny, nx).
 * Pointers are plain device addresses; `stream` is a hipStream_t passed as
 * void* (NULL = the null stream).  Every entry point returns 0 on success and
 * a negative glide_status on failure; glide_last_error() gives the text.
 * No entry point keeps a pointer past its return.  Calls are stream-ordered
 * and do not synchronise the host unless stated.
 *
 * Reference interfaces replaced (JacobDowns/glide):
 *   glide_compute_grounded     <- ForwardOperators.compute_phi / kernel compute_grounded
 *                                 (glide/operators.py:173-192, glide/cuda/viscosity.cu:16-36)
 *   glide_ssa_residual         <- ForwardOperators.compute_residual / kernel compute_residual
 *                                 (glide/operators.py:71-127, glide/cuda/residuals.cu:5-333)
 *   glide_ssa_jvp              <- ForwardOperators.compute_jvp / kernel compute_jvp
 *                                 (glide/operators.py:129-171, glide/cuda/residuals.cu:340-660)
 *   glide_vanka_sweeps         <- ForwardOperators.vanka_sweep / kernel vanka_smooth
 *                                 (glide/operators.py:194-251, glide/cuda/vanka.cu:477-660)
 *   glide_vanka_dump           <- ForwardOperators.vanka_dump (glide/operators.py:253-273,
 *                                 glide/cuda/vanka.cu:800-908)
 *   glide_restrict_* / glide_prolong_*
 *                              <- Multigrid.restrict_* / prolongate_* (glide/multigrid.py:118-259,
 *                                 glide/cuda/transfer.cu)
 *   glide_fascd_solve          <- FASCDSolver.solve / vcycle (glide/multigrid.py:502-645)
 *   glide_ssa_vjp              <- AdjointOperators.compute_residual / compute_vjp / kernel compute_vjp
 *                                 (glide/operators.py:379-472, glide/cuda/residuals.cu:666-1163)
 *   glide_vanka_adjoint_sweeps <- AdjointOperators.vanka_sweep / kernel vanka_smooth_adjoint
 *                                 (glide/operators.py:475-524, glide/cuda/vanka.cu:662-797)
 *   glide_fas_adjoint_solve    <- FASAdjointSolver.solve / vcycle (glide/multigrid.py:873-973)
 *   glide_gradients            <- AdjointOperators.compute_gradient_{beta,bed,H_prev,smb}
 *                                 (glide/operators.py:526-593, glide/cuda/grad.cu)
 */
#ifndef GLIDE_HIP_H
#define GLIDE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum glide_status {
    GLIDE_OK = 0,
    GLIDE_ERR_INVALID_ARG = -1,
    GLIDE_ERR_SHAPE = -2,
    GLIDE_ERR_HIP = -3,
    GLIDE_ERR_NONFINITE = -4, /* a residual norm the call computed is NaN or infinite */
    GLIDE_ERR_TIMEOUT = -5    /* a persistent launch's bounded barrier wait gave up */
} glide_status;

/* Flags for the residual / VJP entry points. */
#define GLIDE_USE_FORCING 1   /* subtract the level's forcing f (residual) / adjoint forcing (VJP) */
#define GLIDE_USE_MASK 2      /* active-set rows H - gamma where mask is set (residual) */
#define GLIDE_FREEZE_CALVING 4 /* calving rate treated as zero */
#define GLIDE_FREEZE_PHI 8     /* do not refresh the grounded fraction before smoothing */

/* Scalar parameters of one level (glide/grid.py:29-145). */
typedef struct glide_physics {
    double n, eps_reg, flotation_reg_driving;
    double m, u_reg, water_drag, flotation_reg_sliding;
    double calving_rate, flotation_reg_calving;
    double thklim;
} glide_physics;

/* One multigrid level: shape, spacing, and the device buffers the solver reads
 * and writes (glide/grid.py:147-372, glide/operators.py:31-56,
 * glide/multigrid.py:647-664,975-985).  Unused pointers may be NULL for entry
 * points that do not touch them. */
typedef struct glide_level {
    int32_t ny, nx;
    double dx;
    glide_physics phys;
    /* state */
    void *u, *v, *H, *H_prev, *phi, *mask;
    /* parameters and forcing */
    void *bed, *B, *beta, *smb;
    /* forward operator buffers */
    void *f_u, *f_v, *f_H, *r_u, *r_v, *r_H, *F_u, *F_v, *F_H, *gamma;
    /* FAS-CD scratch */
    void *w_u, *w_v, *w_H, *y_u, *y_v, *y_H, *z_u, *z_v, *z_H, *chi, *cphi;
    /* adjoint state and buffers */
    void *lambda_u, *lambda_v, *lambda_H;
    void *af_u, *af_v, *af_H, *ar_u, *ar_v, *ar_H, *vjp_u, *vjp_v, *vjp_H;
    void *wl_u, *wl_v, *wl_H, *zl_u, *zl_v, *zl_H;
    /* gradients */
    void *g_beta, *g_bed, *g_H_prev, *g_smb;
    /* smoother workspace: 5 * ny * nx values */
    void *work;
    /* Optional adjoint smoother cache, 26 * ny * nx values (no counterpart in
     * the reference).  When set, adjoint solves and sweeps factorise each
     * cell's transposed patch Jacobian once per level and state (the forward
     * state is fixed during an adjoint solve) and reuse the factors and the
     * cell viscosity in every pass; NULL recomputes them in every pass. */
    void *adj_cache;
} glide_level;

/* VankaConfig + NewtonConfig (glide/operators.py:314-327).  The forward
 * smoother uses a local Newton relaxation of 0.5, as the reference kernel does
 * (glide/cuda/vanka.cu:615); `relaxation` is accepted for API parity. */
typedef struct glide_vanka_config {
    double omega, relax_phi;
    int32_t newton_steps;
    double relaxation, ssa_damping, mc_damping;
    /* Mixed form (no counterpart in the reference): 1 evaluates and factorises
     * each cell's patch Jacobian in FP32 while the residual, the stopping test
     * and the Newton update stay FP64; 0 is the plain FP64 form. */
    int32_t local_fp32;
    /* Multicolour smoothing (no counterpart in the reference): 1 replaces the
     * additive pass by a red-black pass (red cells, then black cells, each
     * writing its five unknowns at full weight; `omega` unused), whose local
     * Newton iteration uses `relaxation` and stops once sum r^2 <= newton_tol
     * (tested before each update, at most newton_steps updates); the adjoint
     * passes likewise.  The fixed point is the additive smoother's. */
    int32_t multicolour;
    double newton_tol;
} glide_vanka_config;

/* FASCDConfig / FASAdjointConfig (glide/multigrid.py:671-682, 992-1001).
 * compat_vprolong = 1 reproduces the reference's v-correction prolongation
 * through the u-facet routine (glide/multigrid.py:212-235; square grids only);
 * 0 uses the v-facet bilinear operator. */
typedef struct glide_fas_config {
    int32_t freeze_coarse_calving, freeze_coarse_phi;
    int32_t coarsest_steps, pre_steps, post_steps, finest_steps, maximum_vcycles;
    double relative_tolerance, absolute_tolerance;
    int32_t compat_vprolong;
} glide_fas_config;

typedef struct glide_solve_report {
    int32_t vcycles;
    double initial_norm, final_norm;
    double norms[64];     /* |r| after each V-cycle */
    double norms3[3];     /* |r_u|, |r_v|, |r_H| at exit */
} glide_solve_report;

const char* glide_version(void);
/* Allocates the library's per-device scratch (reduction partials, norm slots,
 * the failure latch, the enthalpy status record) on the current device, so
 * that no later call allocates device memory; calls made inside a caller's
 * stream capture need it first (no counterpart in the reference). */
int glide_init(void);
/* Synchronises `stream` and reports (then clears) the failure latch: GLIDE_OK,
 * or GLIDE_ERR_TIMEOUT if a persistent launch's barrier wait gave up since
 * the last check (its solve's results are NaN).  Solves that synchronise for
 * a report check the latch themselves (no counterpart in the reference). */
int glide_check_status(void* stream);
/* Copies the calling thread's last error message into buf (NUL-terminated);
 * returns its full length. */
int glide_last_error(char* buf, size_t len);

int glide_compute_grounded(const glide_level* L, double relaxation, void* stream);
int glide_ssa_residual(const glide_level* L, double dt, int flags, void* out_u, void* out_v, void* out_H,
                       void* stream);
/* J(x) d of the residual along d = (d_u, d_v, d_H) with phi, mask, bed, B and
 * beta frozen.  As in the reference: Dirichlet facet rows are zero and the
 * thickness row is (1 - mask) times the mass-balance derivative when
 * GLIDE_USE_MASK is set; GLIDE_FREEZE_CALVING drops the calving term. */
int glide_ssa_jvp(const glide_level* L, double dt, int flags, const void* d_u, const void* d_v, const void* d_H,
                  void* out_u, void* out_v, void* out_H, void* stream);
/* |a|_2, |b|_2, |c|_2 into host memory out3 (synchronises the stream). */
int glide_norms3(const void* a, size_t na, const void* b, size_t nb, const void* c, size_t nc, double* out3,
                 void* stream);
int glide_vanka_sweeps(const glide_level* L, double dt, int n_iter, const glide_vanka_config* vc, int flags,
                       void* stream);
/* Local 5x5 systems at the current state: J (ny*nx, 25), r (ny*nx, 5). */
int glide_vanka_dump(const glide_level* L, double dt, void* J, void* r, void* stream);

int glide_restrict_vfacet(const void* fine, void* coarse, int ny_coarse, int nx_coarse, void* stream);
int glide_restrict_hfacet(const void* fine, void* coarse, int ny_coarse, int nx_coarse, void* stream);
int glide_restrict_cell(const void* fine, void* coarse, int ny_coarse, int nx_coarse, int use_max, void* stream);
int glide_prolong_vfacet_bilinear(const void* coarse, void* fine, int ny_fine, int nx_fine, void* stream);
int glide_prolong_hfacet_bilinear(const void* coarse, void* fine, int ny_fine, int nx_fine, void* stream);
int glide_prolong_cell(const void* coarse, void* fine, int ny_fine, int nx_fine, int bilinear, void* stream);

/* Full FAS-CD solve on levels[start_level..n_levels-1]; vc holds one config
 * per level; 0 <= maximum_vcycles <= 64.  A non-finite residual norm
 * returns GLIDE_ERR_NONFINITE, a timed-out persistent barrier
 * GLIDE_ERR_TIMEOUT (both only where the solve synchronises).  Synchronises the stream once per V-cycle for the stopping test
 * unless both tolerances are <= 0 (then it runs exactly maximum_vcycles
 * V-cycles and synchronises once at the end to fill the report; with report
 * NULL it neither synchronises nor computes the per-cycle residual norms, and
 * the level's residual buffers are not refreshed after the last cycle;
 * glide_fas_adjoint_solve alike). */
int glide_fascd_solve(const glide_level* levels, int n_levels, int start_level, double dt,
                      const glide_fas_config* cfg, const glide_vanka_config* vc, glide_solve_report* report,
                      void* stream);

/* The mixed-precision form of glide_fascd_solve (no counterpart in the
 * reference).  An FP64 defect-correction loop: per outer iteration the FP64
 * residual r(x) of the start level (GLIDE_USE_FORCING, the active set), the
 * FP32 copy x32 of the state, the FP32 forcing F32(x32) - r(x), inner_vcycles
 * FP32 FAS-CD V-cycles with the red-black smoother (vc[l].multicolour must be
 * 1; the counts, freeze flags and compat_vprolong come from cfg, whose
 * tolerances and maximum_vcycles are not used) whose coarsest level runs in
 * FP64 when fp64_coarsest is set, then x += y32 - fl32(x) in FP64 with the
 * inner active set (masked cells set to thklim exactly).  It stops when the
 * FP64 norm meets relative_tolerance or absolute_tolerance, or after
 * max_outer iterations (<= 64).  The FP32 copy of the state rounds every
 * correction to FP32 steps of the state's size (~6e-8 relative), so once an
 * outer iteration reduces the norm by less than 0.6 (or fails), the loop
 * switches to FP64 FAS-CD V-cycles with the same red-black smoother, counts
 * and flags on levels64 (the finish to the FP64 tolerance).  An outer
 * iteration whose FP64 norm does not decrease in that phase is undone and
 * replaced by one FP64 V-cycle of the reference form (additive smoother,
 * default counts).  stats (optional, 3 ints): fallbacks, FP32 iterations,
 * FP64 red-black V-cycles.
 * levels32 mirrors levels64 in FP32 (state, phi, mask, bed, B, beta, f/r/F,
 * gamma, the FAS scratch and the 5-plane workspace); levels64 needs the FP64
 * buffers of glide_fascd_solve on every level.  The
 * outer iteration is
 * replayed from a cached HIP graph and synchronises once per iteration for the
 * norm; the report holds the FP64 norm after each outer iteration.  With both
 * tolerances <= 0 and report NULL the solve runs exactly max_outer FP32
 * iterations, enqueued (from one cached graph) without any host wait, norm,
 * undo or fallback; glide_fas_adjoint_solve_mixed alike. */
typedef struct glide_mixed_config {
    int32_t max_outer;
    int32_t inner_vcycles;
    double relative_tolerance, absolute_tolerance;
    int32_t fp64_coarsest;
    /* krylov_dim m > 0 (with a report or a tolerance): the adjoint solve's
     * outer iteration is flexible GMRES(m) (restarted every m iterations)
     * with the FP32 correction V-cycles as the preconditioner, and the
     * forward solve's FP64 finishing cycles are mixed by Anderson
     * acceleration of depth min(m, 6) (an extrapolated state is kept only if
     * it lowers the FP64 residual norm), in krylov_work (device,
     * glide_krylov_work_doubles(ny, nx, m) doubles for the start level);
     * 0 keeps the plain loops. */
    int32_t krylov_dim;
    void* krylov_work;
    /* The forward solve's FP64 finishing V-cycles: local Newton steps and
     * relaxation of their red-black passes (0 keeps the vc values).  Near
     * the solution one full step (1, 1.0) suffices: at 512^2 the finishing
     * needs fewer cycles and each pass a third of the work. */
    int32_t fp64_newton_steps;
    double fp64_relaxation;
} glide_mixed_config;
/* Doubles of glide_mixed_config.krylov_work for an ny x nx start level and
 * restart length m (1 <= m <= 16); -1 for invalid arguments. */
int64_t glide_krylov_work_doubles(int ny, int nx, int m);
int glide_fascd_solve_mixed(const glide_level* levels64, const glide_level* levels32, int n_levels, int start_level,
                            double dt, const glide_fas_config* cfg, const glide_vanka_config* vc,
                            const glide_mixed_config* mc, glide_solve_report* report, int* stats, void* stream);

/* The mixed-precision form of glide_fas_adjoint_solve (no counterpart in the
 * reference; linear drag).  With mc->krylov_dim = m > 0 (and a report or a
 * tolerance) the outer iteration is flexible GMRES(m) on J^T lambda = g,
 * right-preconditioned by the FP32 correction V-cycles below, with the FP64
 * transpose operator, classical Gram-Schmidt with one reorthogonalisation and
 * the true residual formed at every restart; the report's norms are the
 * iteration's residual estimates.  Otherwise, per outer iteration the FP64 adjoint residual
 * rho = J^T lambda - g of the start level, then inner_vcycles FP32 FAS adjoint
 * V-cycles for the correction J^T d = -rho from d = 0 with the red-black
 * adjoint passes (vc[l].multicolour = 1; each colour's patch solves move
 * their multipliers by omega times the full solution; the cached patch
 * factors use vc's dampings), the coarsest level in FP64 when fp64_coarsest
 * is set, and lambda += d in FP64.  Stopping, undo-and-fallback (one FP64
 * reference adjoint V-cycle), graph replay and the report as for
 * glide_fascd_solve_mixed (one phase).  The multipliers of the rows the
 * patch solves pin (domain-edge facets, active-set cells) are kept at zero.
 * levels32 needs the FP32 state, multiplier,
 * adjoint-operator and adjoint-scratch buffers, the workspace and adj_cache;
 * every FP64 level needs adj_cache. */
int glide_fas_adjoint_solve_mixed(const glide_level* levels64, const glide_level* levels32, int n_levels,
                                  int start_level, double dt, const glide_fas_config* cfg,
                                  const glide_vanka_config* vc, const glide_mixed_config* mc,
                                  glide_solve_report* report, int* stats, void* stream);

/* One V-cycle from `level` down (FASCDSolver.vcycle, glide/multigrid.py:536-645);
 * finest != 0 stores w and chi first and adds the finest-level sweeps. */
int glide_fascd_vcycle(const glide_level* levels, int n_levels, int level, int finest, double dt,
                       const glide_fas_config* cfg, const glide_vanka_config* vc, void* stream);

int glide_ssa_vjp(const glide_level* L, double dt, int flags, const void* lambda_u, const void* lambda_v,
                  const void* lambda_H, void* out_u, void* out_v, void* out_H, void* stream);
int glide_vanka_adjoint_sweeps(const glide_level* L, double dt, int n_iter, const glide_vanka_config* vc, int flags,
                               void* stream);
int glide_fas_adjoint_solve(const glide_level* levels, int n_levels, int start_level, double dt,
                            const glide_fas_config* cfg, const glide_vanka_config* vc, glide_solve_report* report,
                            void* stream);
/* One adjoint V-cycle (FASAdjointSolver.vcycle, glide/multigrid.py:905-973). */
int glide_fas_adjoint_vcycle(const glide_level* levels, int n_levels, int level, int finest, double dt,
                             const glide_fas_config* cfg, const glide_vanka_config* vc, void* stream);
/* which: bit 0 beta, bit 1 bed, bit 2 H_prev, bit 3 smb; with
 * GLIDE_GRAD_EXACT (no counterpart in the reference) the gradient of the
 * discrete problem: the right / bottom cell terms of facets nx-1 and ny-1
 * that the reference's guards drop (glide/cuda/grad.cu:62) are kept, and the
 * H_prev and smb gradients are zero in active-set cells (whose rows do not
 * depend on them; the reference's multipliers there are not determined). */
#define GLIDE_GRAD_EXACT 16
int glide_gradients(const glide_level* L, double dt, int which, void* stream);
/* No counterpart in the reference (its gradients are those of
 * glide_gradients): g_B = lambda^T dR/dB, the gradient with respect to the
 * rate factor B through the cell viscosity eta = B/2 (...)^((1-n)/(2n)) of the
 * normal stresses and the four corners' shear stresses, over the rows of
 * interior facets (the multipliers of L; every cell, including active-set
 * cells, whose facets' stress rows still hold their viscosity). */
int glide_gradient_rate_factor(const glide_level* L, double dt, void* g_B, void* stream);

/* Graph replay of fixed-cycle solves (no counterpart in the reference; the
 * reference launches every kernel from Python): when both tolerances are <= 0,
 * glide_fascd_solve / glide_fas_adjoint_solve capture their launches into a
 * HIP graph on first use and replay it on later calls with identical level
 * descriptors, configs and dt (buffer contents may change freely).  On by
 * default; glide_graphs_clear drops the cached graphs (call it before freeing
 * buffers that a cached graph refers to if the addresses could be reused by
 * unrelated data). */
int glide_graphs_enable(int on);
int glide_graphs_clear(void);
/* Number of graphs kept per device (default 8; the least recently used is
 * dropped), e.g. two per domain when an ensemble of domains is stepped. */
int glide_graphs_capacity(int n);
/* Adjoint passes of levels with an adj_cache (no counterpart in the
 * reference): 1 (default) runs each pass after the transpose operator as one
 * launch (patch solves of the tile and its left/top neighbours, then the
 * in-place multiplier update); 0 solves into the workspace and applies in a
 * second launch.  Both give the same multipliers. */
int glide_adjoint_fused_enable(int on);
/* Forward smoothing of levels of at most 128 x 128 cells with a frozen
 * grounded fraction (no counterpart in the reference): 1 (default) runs each
 * pass as one launch (the patch solves of each tile row and of the cells
 * above and left of it, then the owned state written to the other of two
 * buffers, the level's own and its workspace); 0 solves into the workspace
 * and applies in a second launch.  Both give the same state. */
int glide_forward_fused_enable(int on);
/* Forward smoothing of levels at most 64 cells wide and 96 rows high with a
 * frozen grounded fraction (no counterpart in the reference): 1 (default) runs
 * all passes of one call as one persistent launch of one workgroup per three
 * rows, a device-wide barrier between passes (it uses the last two values of
 * the level's workspace as its arrival counter and timeout flag; a timed-out
 * barrier leaves NaN thicknesses); 0 launches each pass as above.  Both give
 * the same state. */
int glide_forward_persistent_enable(int on);
/* Cached adjoint smoothing of levels of at most 512 cells with linear drag
 * (no counterpart in the reference): 1 (default) runs all passes of one call
 * in a single workgroup (transpose operator, patch solves and update per pass,
 * separated by workgroup barriers), when the device's LDS per block holds the
 * level; 0 launches the transpose operator and the adjoint pass per pass.
 * Both give the same multipliers. */
int glide_adjoint_coarse_enable(int on);
/* Cached adjoint passes of levels of at most 512 x 512 cells with linear drag
 * (no counterpart in the reference): 1 (default) runs each pass, the
 * transpose operator included, as one launch (a tile recomputes its left and
 * top neighbours' residuals); 0 launches the transpose operator and the
 * adjoint pass separately.  Both give the same multipliers. */
int glide_adjoint_pass_fused_enable(int on);
/* Cached adjoint smoothing of levels at most 64 cells wide and 32 rows high
 * with linear drag, when the one-workgroup form does not apply (no
 * counterpart in the reference): 1 (default) runs all passes of one call as
 * one persistent launch of one workgroup per four rows, with the barriers of
 * glide_forward_persistent_enable (same workspace words; a timed-out barrier
 * leaves NaN multipliers); 0 launches one fused pass at a time.  Both give
 * the same multipliers. */
int glide_adjoint_persistent_enable(int on);
/* Persistent launches (the two switches above; no counterpart in the
 * reference): 1 runs every barrier with the cache write-back and invalidation
 * that workgroups on different XCDs need, also when the launch finds them all
 * on one XCD (where the default 0 skips it); for testing that path.  Both give
 * the same results. */
int glide_persistent_fenced_enable(int on);
/* Polls a persistent barrier wait makes before it gives up (default 2^20);
 * a negative value gives up at the first barrier, which exercises the
 * failure path (NaN results, GLIDE_ERR_TIMEOUT from the next synchronising
 * call) without any kernel waiting (testing switch; no counterpart in the
 * reference). */
int glide_persistent_spin_limit(int polls);

/* Measurement support (no counterpart in the reference): when enabled, HIP
 * event pairs are recorded around every launch of the hot kernels
 * (0 forward patch solve, 1 patch increment apply, 2 adjoint patch solve,
 * 3 transpose operator, 4 residual, 5 adjoint increment apply, 6 enthalpy
 * column smoother, 7 enthalpy layer smoother, 8 enthalpy residual, 9 adjoint
 * patch factorisation into the cache, 10 / 11 one colour phase of the
 * red-black smoother in FP32 / FP64, 12 / 13 one colour phase of the red-black
 * adjoint pass in FP32 / FP64; the red-black launches count every cell).
 * glide_timing_read synchronises those events and returns the launch count,
 * summed duration and summed unit count (cells; column nodes for 6-8) of
 * kernel `kernel` since the last enable. */
int glide_timing_enable(int on);
int glide_timing_read(int kernel, long* count, double* total_ms, double* total_cells);
/* The same restricted to launches over exactly `units` cells (one grid level). */
int glide_timing_read_units(int kernel, long units, long* count, double* total_ms);

/* Batched ensemble step (no counterpart in the reference, which builds one
 * hierarchy per domain and process: glide/model.py:25-39,
 * glide/multigrid.py:8-44).  An ensemble of independent domains of uneven
 * sizes and level counts takes one forward + adjoint step per
 * glide_ensemble_step: per domain exactly the work of
 * glide_fascd_solve_mixed (max_outer fixed FP32 iterations, no report, start
 * level 0), H_prev = H, the adjoint seeds dJ/du = u, dJ/dv = v (forcing -u, -v,
 * 0), glide_fas_adjoint_solve_mixed (the same counts) and glide_gradients
 * (grad_which), but every stage is one launch over all the domains taking
 * part in it, from a table of per-domain arguments: the V-cycles run level by
 * level (the descent of a level covers every domain deeper than it, all the
 * domains' coarsest-level passes run together, then the ascents).  Per domain
 * the results are bitwise those of its own asynchronous mixed solves.
 * levels64[d] / levels32[d] point at domain d's n_levels[d] FP64 / FP32
 * level descriptors (the buffers of glide_fascd_solve_mixed and
 * glide_fas_adjoint_solve_mixed together, each level half the one above;
 * cfg / adjoint_cfg the V-cycle counts of the forward / adjoint solves;
 * linear drag); the descriptors are copied, the buffers must outlive the
 * ensemble.  The launch sequence and its tables (device memory the ensemble
 * owns) are recorded at the first step of a time step length and replayed as
 * one HIP graph (no allocation or host wait in later steps of that length;
 * recording again and glide_ensemble_destroy first synchronise the device).
 * glide_ensemble_launches gives the launches of one recorded step. */
typedef struct glide_ensemble glide_ensemble;
int glide_ensemble_create(int32_t n_domains, const int32_t* n_levels, const glide_level* const* levels64,
                          const glide_level* const* levels32, const glide_fas_config* cfg,
                          const glide_fas_config* adjoint_cfg, const glide_vanka_config* forward_vc,
                          const glide_vanka_config* adjoint_vc, const glide_mixed_config* mc, int32_t grad_which,
                          glide_ensemble** out);
int glide_ensemble_step(glide_ensemble* e, double dt, void* stream);
int32_t glide_ensemble_launches(const glide_ensemble* e);
/* Tile shapes of the red-black smoothing phases (no counterpart in the
 * reference): rows[l] cell-pair rows per block (1, 2, 4 or 8; default 4) for
 * the stages of level l, l < glide_ensemble_levels(e) (the most levels of any
 * domain), and rows[levels] for the coarsest-level passes; n = levels + 1.
 * The results do not depend on the tiles (each cell's arithmetic is the
 * same); the tables are recorded again at the next step.  Chosen by
 * glide_amd.ensemble.autotune_tiles from timed steps. */
int32_t glide_ensemble_levels(const glide_ensemble* e);
int glide_ensemble_set_tiles(glide_ensemble* e, const int32_t* rows, int32_t n);
int glide_ensemble_destroy(glide_ensemble* e);

#ifdef __cplusplus
}
#endif

#endif /* GLIDE_HIP_H */
