// Launch interface of the enthalpy kernels (enthalpy_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "timing.hpp"

namespace glide {

// Everything a column kernel reads.  Layouts: E, E_prev, f_E, omega,
// phi_strain (ny, nx, nz); u3d (nz, ny, nx+1); v3d (nz, ny+1, nx); H, H_prev,
// E_surface, Q_geo (ny, nx).
template <typename T>
struct EnthArgs {
    int ny, nx, nz, flags;  // flags: 1 horizontal advection, 2 sigma advection, 4 drainage
    T dxi, dt, drain_rate, h_thin;
    T lf_sq;  // sqrt(1 + lf_c), the Lax-Friedrichs dissipation factor
    const T *E, *E_prev, *f_E, *u3d, *v3d, *omega, *H, *H_prev, *E_surface, *Q_geo, *phi_strain;
    const int* skip;  // nullable: a launch does nothing while *skip != 0
    // Depth-constant velocity (the coupled step's broadcast SSA velocity):
    // when u2d is set, level k's velocity is u2d / vel_div, the same value the
    // broadcast u3d[k] holds, read from one (ny, nx+1) / (ny+1, nx) plane.
    const T *u2d, *v2d;
    T vel_div;
};

// Device record of a fused column sweep (EnthalpyOperators.column_sweep).
constexpr int kEnthMaxSweeps = 256;
struct EnthStatus {
    double r0, r;
    int sweeps, done;  // `done` is the skip flag of the remaining launches
    double norms[kEnthMaxSweeps];
};

constexpr int kEnthResBlocks = 2048;  // residual grid (grid-stride over columns)
constexpr int kEnthRmaxSlots = 256;   // max-|r| accumulators of the column pass (bit patterns of doubles)

// max |r| of the residual over the grid, into `max_out` (device), or, with
// `status`, the sweep bookkeeping: mode 0 starts a sweep (r0), mode 1 ends an
// iteration (count, norm, stopping test).
template <typename T>
hipError_t launch_enth_residual(const EnthArgs<T>& A, T* r_out, int use_forcing, double* partials, double* max_out,
                                EnthStatus* status, int status_mode, double rtol, double atol, hipStream_t st);
// Column Newton smoother; writes delta = E_col - E (if non-null) and/or
// E_out = E + omega_relax * delta (if non-null).  colour >= 0: one phase of
// the red-black column ordering through E_mc (delta and E_out null): phase 0
// smooths the even columns ((i + j) even) from A.E into E_mc and copies the
// odd ones (nothing to copy when E_mc == A.E), phase 1 smooths the odd
// columns in place (A.E == E_mc).
// rmax (optional, kEnthRmaxSlots values, zero on entry): the column pass
// also forms max |r| of the state it starts from (its first Newton
// iteration's rows, the residual launch_enth_residual computes) and folds it
// into the slots (64-bit atomic max of the double's bit pattern, one slot per
// wavefront modulo kEnthRmaxSlots; a NaN stays), and writes that residual to
// r_out when given.
template <typename T>
hipError_t launch_enth_column_smooth(const EnthArgs<T>& A, T* delta, T* E_out, T omega_relax, int n_newton,
                                     T relaxation, hipStream_t st, int colour = -1, T* E_mc = nullptr,
                                     unsigned long long* rmax = nullptr, T* r_out = nullptr);
// The maximum over the slots of a column pass, zeroing them: mode 0 only
// zeroes, mode 1 also closes one sweep iteration in the status record (as
// launch_enth_residual's status mode 1; nothing once the sweep stopped).
hipError_t launch_enth_reduce_slots(unsigned long long* slots, EnthStatus* status, int mode, double rtol, double atol,
                                    hipStream_t st);
// Pointwise Jacobi layer smoother; delta = -r / J_kk and/or E_out = E + omega_relax * delta.
template <typename T>
hipError_t launch_enth_layer_smooth(const EnthArgs<T>& A, T* delta, T* E_out, T omega_relax, hipStream_t st);
// E += omega_relax * delta (skips with A.skip).
template <typename T>
hipError_t launch_enth_apply(const EnthArgs<T>& A, T* E, const T* delta, T omega_relax, hipStream_t st);
// omega = H sigma_dot from the sigma-space continuity equation, divided by out_div.
template <typename T>
hipError_t launch_enth_omega(const EnthArgs<T>& A, const T* dhdt, const T* bmb, T out_div, T* omega, hipStream_t st);
// dst[k] = src / div for every layer k of a (nz, n2d) array.
template <typename T>
hipError_t launch_enth_broadcast(const T* src, T* dst, long n2d, int nz, T div, hipStream_t st);
// out = (a - b) / div elementwise (the column thickness tendency).
template <typename T>
hipError_t launch_enth_tendency(const T* a, const T* b, T div, T* out, long n, hipStream_t st);
// Basal frictional heating (if `frictional`) and the forcing f_E.
template <typename T>
hipError_t launch_enth_rhs(const EnthArgs<T>& A, const T* beta, T m, T u_reg, int frictional, T* Q_fh, T* f_E,
                           hipStream_t st);
// B = mean_k(A_k (1 + 181.25 w_k))^(-1/n) / scale.
template <typename T>
hipError_t launch_enth_arrhenius(const EnthArgs<T>& A, T n_glen, T scale, T* B, hipStream_t st);
// SSA strain heating 4 eta eps_e^2 rho_g / SEC_PER_YR written to every level.
template <typename T>
hipError_t launch_enth_strain_heating(const T* u, const T* v, const T* B, int ny, int nx, int nz, T dxi, T n_glen,
                                     T eps_reg, T rho_g, T* phi, hipStream_t st);
// Temperature (clipped at the pressure melting point) and water content.
template <typename T>
hipError_t launch_enth_diagnostics(const EnthArgs<T>& A, T* temperature, T* water, hipStream_t st);
// E = C_I (min(T, T_pmp) - T_REF) / E_SCALE for a (ny, nx, nz) temperature.
template <typename T>
hipError_t launch_enth_from_temperature(const EnthArgs<T>& A, const T* temperature, T* E, hipStream_t st);
// E_surface = C_I (min(T_s, T_MELT) - T_REF) / E_SCALE.
template <typename T>
hipError_t launch_enth_surface_from_temperature(const T* Ts, T* Es, long n, hipStream_t st);

// Adjoint of the converged solve: factors of the exact Jacobian's transposed
// column blocks, right-hand side with the rate-factor term, red-black column
// sweeps, residual with the sweep bookkeeping, gradients.
// Workspaces of the adjoint in the checkerboard layout (doubles per plane:
// 2 * nz * ny * ((nx + 1) / 2)); rows holds seven planes: the Thomas factors
// (aT, w, cp) of every column's transposed block and the transposed
// horizontal couplings (per level, or on level 0 with A.u2d set).
template <typename T>
hipError_t launch_enth_adj_factor(const EnthArgs<T>& A, T* rows, hipStream_t st);
template <typename T>
hipError_t launch_enth_adj_rhs(const EnthArgs<T>& A, const T* gE, const T* gB, T n_glen, T scale, T* g,
                               hipStream_t st);
// lam between the (ny, nx, nz) and the checkerboard layouts.
template <typename T>
hipError_t launch_enth_adj_layout(int ny, int nx, int nz, const T* src, T* dst, bool to_cb, hipStream_t st);
// One adjoint sweep: red and black column Gauss-Seidel passes (in place) and
// the stopping test on the max residual of the columns before their update
// (status mode 1).  flat: the couplings were stored for a depth-constant
// velocity.
template <typename T>
hipError_t launch_enth_adj_sweep_rb(const int* skip, const EnthArgs<T>& A, bool flat, const T* rows, const T* g,
                                    T* lam, T omega_relax, unsigned long long* acc, EnthStatus* status, double rtol,
                                    double atol, hipStream_t st);
// max |g - J^T lam| into the sweep bookkeeping (status mode 0 starts it).
// acc: two counters (maximum, finished blocks), zero on entry and left zero.
template <typename T>
hipError_t launch_enth_adj_residual(const int* skip, const EnthArgs<T>& A, bool flat, const T* rows, const T* g,
                                    const T* lam, unsigned long long* acc, EnthStatus* status, int status_mode,
                                    double rtol, double atol, hipStream_t st);
template <typename T>
hipError_t launch_enth_adj_grads(const EnthArgs<T>& A, const T* lam, T* g_E_prev, T* g_phi, T* g_Q, T* g_Es,
                                 hipStream_t st);

// Gradients of the thermal step with respect to the inputs the momentum step
// hands it (velocities, thicknesses, sliding coefficient, the rate factor the
// strain heating read), from the adjoint state of the converged solve.
constexpr int kThermalVjpPlanes = 14;  // (ny, nx) work planes
template <typename T>
struct ThermalVjpArgs {
    const T* lam;    // (ny, nx, nz) adjoint state
    const T* g_phi;  // (ny, nx, nz) dL/dphi_strain, read when B_in is set
    const T* g_Q;    // (ny, nx) dL/dQ_fh, read when beta is set
    const T* gB;     // (ny, nx) seed of the new rate factor, or null
    const T* B_in;   // rate factor the strain heating read, or null (no strain heating)
    const T* beta;   // sliding coefficient, or null (no frictional heating)
    T n_glen, B_scale, glen_p, eps_reg, rho_g, m, u_reg, dt_yr;
    T* work;         // kThermalVjpPlanes planes of (ny, nx)
    T *g_u, *g_v, *g_H, *g_H_prev, *g_beta, *g_B;
};
// A.u2d / A.v2d: the step's velocities in m/a.
template <typename T>
hipError_t launch_thermal_vjp(const EnthArgs<T>& A, const ThermalVjpArgs<T>& V, hipStream_t st);

// exp and log as the enthalpy rows compute them, over n values (for tests).
hipError_t launch_enth_elementary(const double* x, long n, double* e_out, double* l_out, hipStream_t st);

// Level-major form of the fused alternating sweep (thermal step with a
// depth-constant velocity, A.u2d set, 9 <= nz <= 64): A.E, A.E_prev, A.f_E and
// A.omega point to (nz, ny nx) planes.
bool enth_lm_supported(int nz);
// (ny nx, nz) -> (nz, ny nx) when to_lm, else back.
template <typename T>
hipError_t launch_enth_lm_layout(const T* src, T* dst, long ncol, int nz, bool to_lm, hipStream_t st);
// Column Newton smoother: W = E + omega_relax (E_col - E), with rmax as
// launch_enth_column_smooth's (no residual output).
template <typename T>
hipError_t launch_enth_column_lm(const EnthArgs<T>& A, T* W, T omega_relax, int n_newton, T relaxation,
                                 unsigned long long* rmax, hipStream_t st);
// Layer smoother: E_out = E + omega_relax delta.
template <typename T>
hipError_t launch_enth_layer_lm(const EnthArgs<T>& A, T* E_out, T omega_relax, hipStream_t st);
// Residual with forcing into the sweep bookkeeping (status mode as
// launch_enth_residual's) and, when r_out is given, written in the
// (ny, nx, nz) layout.
template <typename T>
hipError_t launch_enth_residual_lm(const EnthArgs<T>& A, T* r_out, double* partials, EnthStatus* status,
                                   int status_mode, double rtol, double atol, hipStream_t st);

}  // namespace glide
