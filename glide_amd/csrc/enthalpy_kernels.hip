// gfx950 kernels of the conservative sigma-coordinate enthalpy solver.
//
// Discretisation follows glide/cuda/enthalpy.cu (time term, vertical
// diffusion with smoothed K(E), upwind sigma advection, Lax-Friedrichs
// horizontal advection with the momentum solver's mass flux, drainage) and
// glide/enthalpy.py (forcing, diagnostics, rate factor).
//
// Work decomposition: one wavefront per column, lane k = sigma level k
// (nz <= 64).  In the (ny, nx, nz) layout a column is contiguous, so every
// per-level load and store of a wavefront is one coalesced segment, and the
// horizontal neighbours of a level are contiguous columns too.  Vertical
// neighbours come from cross-lane shuffles.  The column's tridiagonal Newton
// system is solved with parallel cyclic reduction across the 64 lanes (rows
// >= nz are identity rows) instead of a serial per-thread Thomas sweep, so no
// per-thread arrays of length nz are needed.  The adjoint sweeps, whose
// column blocks are fixed, instead factor them once and solve one column per
// thread in a level-major checkerboard layout (see the adjoint layout below).
#include <hip/hip_runtime.h>

#include "enthalpy_kernels.hpp"
#include "ssa_device.hpp"

namespace glide {

// glide/enthalpy.py:46-65
constexpr double kRhoI = 910.0, kRhoW = 1000.0, kCI = 2009.0, kKI = 2.1, kLHeat = 3.34e5;
constexpr double kTRef = 223.15, kTMelt = 273.15, kBetaCC = 7.9e-8, kGrav = 9.81;
constexpr double kKCold = kKI / kCI, kKTempFactor = 0.1;
constexpr double kEScale = kCI * (kTMelt - kTRef);
constexpr double kSecPerYr = 365.25 * 86400.0;
constexpr double kMassFluxReg = 10.0 / (kSecPerYr * kSecPerYr);
constexpr double kMassFluxRegYr = 10.0;
constexpr double kDrainSharpness = 0.01;
constexpr double kClausius = kBetaCC * kRhoI * kGrav;  // K per m of ice
constexpr int kColsPerBlock = 4;                        // wavefronts (columns) per 256-thread block

template <typename T>
__device__ __forceinline__ T t_pmp(T sigma, T H) {
    return T(kTMelt) - T(kClausius) * ((T(1) - sigma) * H);
}
// Elementary functions of the rows, written for the arguments the rows pass
// (finite, far from overflow): without the general routines' special cases
// each takes well under half their instructions, and each stays within a few
// ulp of the correctly rounded value (tests compare them with the general
// routines over the rows' argument ranges).
//
// e^x for finite |x| <= 700: Cody-Waite reduction x = n ln2 + r with ln2 in
// two parts (the reduction by fused multiply-adds), |r| <= ln2 / 2, then e^r
// by its Taylor polynomial to degree 13 (the first omitted term is below
// 6e-18 relative) in Horner form, scaled by 2^n.
__device__ __forceinline__ double exp_rows(double x) {
    constexpr double kLog2e = 1.4426950408889634, kLn2Hi = 0x1.62e42fefa39efp-1, kLn2Lo = 0x1.abc9e3b39803fp-56;
    const double n = rint(x * kLog2e);
    const double r = fma(-n, kLn2Lo, fma(-n, kLn2Hi, x));
    double p = 1.0 / 6227020800.0;  // 1 / 13!
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}
// log(u) for finite normal u > 0: u = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m - 1) / (m + 1) (m - 1 exact), |s| <= 0.1716, by
// the atanh series 2 s (1 + s^2/3 + ... + s^18/19) (the first omitted term is
// below 3e-17 relative), plus e ln2 in two parts.
__device__ __forceinline__ double log_rows(double u) {
    constexpr double kLn2Hi = 0x1.62e42fefa39efp-1, kLn2Lo = 0x1.abc9e3b39803fp-56;
    double m = __builtin_amdgcn_frexp_mant(u);  // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(u);
    if (m < 0.7071067811865476) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) * fast_rcp(m + 1.0), s2 = s * s;
    double p = 1.0 / 19.0;
    p = fma(p, s2, 1.0 / 17.0);
    p = fma(p, s2, 1.0 / 15.0);
    p = fma(p, s2, 1.0 / 13.0);
    p = fma(p, s2, 1.0 / 11.0);
    p = fma(p, s2, 1.0 / 9.0);
    p = fma(p, s2, 1.0 / 7.0);
    p = fma(p, s2, 1.0 / 5.0);
    p = fma(p, s2, 1.0 / 3.0);
    p = fma(p, s2, 1.0);
    const double de = (double)e;
    return fma(de, kLn2Hi, fma(de, kLn2Lo, 2.0 * s * p));
}

// The constant quotients below are formed at compile time, so the hot
// helpers multiply instead of dividing (within an ulp or two of the quotient).
template <typename T>
__device__ __forceinline__ T e_pmp(T sigma, T H) {
    return (t_pmp(sigma, H) - T(kTRef)) * T(kCI / kEScale);
}
// glide/cuda/enthalpy.cu:48-98
template <typename T>
__device__ __forceinline__ T diffusivity(T E, T Ep) {
    T z = (E - Ep) * T(kEScale) * T(0.01);
    z = fmin(fmax(z, T(-20)), T(20));
    T s = fast_rcp(T(1) + exp_rows(-z));
    return T(kKCold) * (T(1) - s + T(kKTempFactor) * s);
}
// Drained water content (glide/cuda/enthalpy.cu:48-98): a softplus of the
// enthalpy above the pressure-melting value, log(1 + exp(a)) / s with
// a = s (E - Ep) E_SCALE clamped to [-20, 20] (the excess itself beyond
// a = 10), over the latent heat; its slope is the logistic factor
// exp(a) / (1 + exp(a)) over the latent heat (the reference writes it as
// 1 / (1 + exp(-a)), equal up to rounding; it only enters the Newton
// Jacobian).  Both come from one exponential.
template <typename T>
__device__ __forceinline__ void water_and_slope(T E, T Ep, T& w, T& dw) {
    const T x = (E - Ep) * T(kEScale);
    const T ax = fmin(fmax(T(kDrainSharpness) * x, T(-20)), T(20));
    const T ez = exp_rows(ax);
    w = ((ax > T(10)) ? x : log_rows(T(1) + ez) * T(1.0 / kDrainSharpness)) * T(1.0 / kLHeat);
    dw = (ez * fast_rcp(T(1) + ez)) * T(1.0 / kLHeat);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

template <typename T>
__device__ __forceinline__ T lane_get(T x, int src) {
    return __shfl(x, src, 64);
}

__device__ __forceinline__ bool skipped(const int* skip) { return skip && *skip; }

// Maximum that keeps a NaN operand (fmax returns the other operand), so that
// a non-finite residual reaches the reported norm and fails the stopping test.
__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

// Horizontal advection pieces of one level that do not depend on the
// level's own enthalpy: face mass fluxes (zero on the domain edge) and the
// neighbouring columns' enthalpy (glide/cuda/enthalpy.cu:671-763).
template <typename T>
struct Horiz {
    T Hu_l, Hu_r, Hv_t, Hv_b;
    T E_xm, E_xp, E_ym, E_yp;
};

// Lax-Friedrichs thickness flux through an x facet between columns A (left)
// and B (right), and through a y facet between A (top) and B (bottom).
template <typename T>
__device__ __forceinline__ T lf_flux_x(T u, T HA, T HB) {
    return T(0.5) * (HA + HB) * u - T(0.5) * sqrt(u * u + T(kMassFluxReg)) * (HB - HA);
}
template <typename T>
__device__ __forceinline__ T lf_flux_y(T v, T HA, T HB) {
    return T(0.5) * (HA + HB) * v - T(0.5) * sqrt(v * v + T(kMassFluxReg)) * (HA - HB);
}

// Face mass fluxes of column (i, j) at level k (zero on the domain edge);
// the neighbour enthalpies are left zero.
template <typename T>
__device__ __forceinline__ Horiz<T> face_fluxes(const EnthArgs<T>& A, int i, int j, int k) {
    const int ny = A.ny, nx = A.nx;
    Horiz<T> h{};
    const T Hh = A.H[(long)i * nx + j];
    const T* u = A.u2d ? A.u2d : A.u3d + (long)k * ny * (nx + 1);
    const T* v = A.u2d ? A.v2d : A.v3d + (long)k * (ny + 1) * nx;
    const T div = A.u2d ? A.vel_div : T(1);
    if (j > 0) h.Hu_l = lf_flux_x(u[(long)i * (nx + 1) + j] / div, A.H[(long)i * nx + j - 1], Hh);
    if (j < nx - 1) h.Hu_r = lf_flux_x(u[(long)i * (nx + 1) + j + 1] / div, Hh, A.H[(long)i * nx + j + 1]);
    if (i > 0) h.Hv_t = lf_flux_y(v[(long)i * nx + j] / div, A.H[(long)(i - 1) * nx + j], Hh);
    if (i < ny - 1) h.Hv_b = lf_flux_y(v[(long)(i + 1) * nx + j] / div, Hh, A.H[(long)(i + 1) * nx + j]);
    return h;
}

template <typename T>
__device__ __forceinline__ Horiz<T> load_horiz(const EnthArgs<T>& A, int i, int j, int k) {
    const int ny = A.ny, nx = A.nx, nz = A.nz;
    Horiz<T> h = face_fluxes(A, i, j, k);
    auto col = [&](int ii, int jj) {
        ii = min(max(ii, 0), ny - 1);
        jj = min(max(jj, 0), nx - 1);
        return A.E[((long)ii * nx + jj) * nz + k];
    };
    h.E_xm = col(i, j - 1);
    h.E_xp = col(i, j + 1);
    h.E_ym = col(i - 1, j);
    h.E_yp = col(i + 1, j);
    return h;
}

// Lax-Friedrichs enthalpy flux 0.5 U (El + Er) - 0.5 |U| sq (Er - El) and its
// derivatives with respect to El and Er.
template <typename T>
__device__ __forceinline__ T lf_flux(T U, T El, T Er, T sq, T& dl, T& dr) {
    T al = fabs(U) * sq;
    dl = T(0.5) * (U + al);
    dr = T(0.5) * (U - al);
    return T(0.5) * U * (El + Er) - T(0.5) * al * (Er - El);
}

// rho_i div(H u E) at a level whose own enthalpy is Eh: residual and diagonal.
template <typename T>
__device__ __forceinline__ void horiz_adv(const Horiz<T>& h, T Eh, T sq, T dxi, T& res, T& diag) {
    T dll, dlr, drl, drr, dtl, dtr, dbl, dbr;
    T fl = lf_flux(h.Hu_l, h.E_xm, Eh, sq, dll, dlr);
    T fr = lf_flux(h.Hu_r, Eh, h.E_xp, sq, drl, drr);
    T ft = lf_flux(h.Hv_t, Eh, h.E_ym, sq, dtl, dtr);
    T fb = lf_flux(h.Hv_b, h.E_yp, Eh, sq, dbl, dbr);
    res = T(kRhoI) * ((fr - fl) + (ft - fb)) * dxi;
    diag = T(kRhoI) * ((drl - dlr) + (dtl - dbr)) * dxi;
}

// One level's residual (without forcing) and tridiagonal Jacobian row
// (glide/cuda/enthalpy.cu:783-905).  Ekm1/Ekp1 and om_m/om_p are the
// vertical neighbours (meaningless where the stencil does not use them).
template <typename T>
struct Row {
    T r, a, b, c;
};

// Diffusivity at the interface above level k (between k and k + 1; lane k
// holds Ek, Ekp1 the level above) and, from lane k - 1, at the interface
// below: each interface's value is formed once per column.  Every lane of
// the wavefront must call it (the lower value is a cross-lane read).
template <typename T>
__device__ __forceinline__ void interface_K(int k, int nz, T h, T Ek, T Ekp1, T& Ku, T& Kl) {
    const T dsig = T(1) / T(nz - 1);
    const T Ep_k = e_pmp(T(k) * dsig, h), Ep_p = e_pmp(T(min(k + 1, nz - 1)) * dsig, h);
    Ku = diffusivity(T(0.5) * (Ek + Ekp1), T(0.5) * (Ep_k + Ep_p));
    Kl = lane_get(Ku, max(k - 1, 0));
}

// Per-column constants of the rows: functions of nz, dt and the column's
// thickness only, formed once per column so that the rows divide by nothing.
template <typename T>
struct ColConst {
    T dsig, di, hdi, d2, idd;  // 1/(nz-1), 1/dsig, 1/(dsig/2), 1/dsig^2, 1/(dsig dsig/2)
    T hinv, invdt;             // 1/H, 1/dt
};
template <typename T>
__device__ __forceinline__ ColConst<T> col_const(const EnthArgs<T>& A, T h) {
    ColConst<T> c;
    c.dsig = T(1) / T(A.nz - 1);
    c.di = T(1) / c.dsig;
    c.hdi = T(1) / (T(0.5) * c.dsig);
    c.d2 = T(1) / (c.dsig * c.dsig);
    c.idd = T(1) / (c.dsig * (T(0.5) * c.dsig));
    c.hinv = T(1) / h;
    c.invdt = T(1) / A.dt;
    return c;
}

// The column's own terms of level k's row (sigma advection, diffusion with
// the interface diffusivities Ku / Kl from interface_K, drainage), added to
// w, which holds the time and horizontal terms.
template <typename T>
__device__ __forceinline__ void column_terms(const EnthArgs<T>& A, const ColConst<T>& c, int k, T h, T Ek, T Ekm1,
                                             T Ekp1, T om_m, T om_k, T om_p, T Ku, T Kl, Row<T>& w) {
    const T Ep_k = e_pmp(T(k) * c.dsig, h);
    if (k == 0) {
        if (A.flags & 2) {
            T wh = T(0.5) * (om_k + om_p);
            T wp = fmax(wh, T(0)), wn = fmin(wh, T(0));
            w.r += T(kRhoI) * ((wp * Ek + wn * Ekp1) - om_k * Ek) * c.hdi;
            w.b += T(kRhoI) * (wp - om_k) * c.hdi;
            w.c += T(kRhoI) * wn * c.hdi;
        }
        // half-cell diffusion flux at the bed: -K (E_1 - E_0) / (H dsig) over dsig / 2
        const T co = c.hinv * Ku * c.idd;
        w.r += -co * (Ekp1 - Ek);
        w.b += co;
        w.c -= co;
    } else {
        if (A.flags & 2) {
            T wu = T(0.5) * (om_k + om_p), wl = T(0.5) * (om_m + om_k);
            T up = fmax(wu, T(0)), un = fmin(wu, T(0)), lp = fmax(wl, T(0)), ln = fmin(wl, T(0));
            w.r += T(kRhoI) * ((up * Ek + un * Ekp1) - (lp * Ekm1 + ln * Ek)) * c.di;
            w.a -= T(kRhoI) * lp * c.di;
            w.c += T(kRhoI) * un * c.di;
            w.b += T(kRhoI) * (up - ln) * c.di;
        }
        w.r += -c.hinv * c.d2 * (Ku * (Ekp1 - Ek) - Kl * (Ek - Ekm1));
        T dl = c.hinv * c.d2 * Kl, du = c.hinv * c.d2 * Ku;
        w.a -= dl;
        w.c -= du;
        w.b += dl + du;
    }
    if (A.flags & 4) {
        T wc, dwc;
        water_and_slope(Ek, Ep_k, wc, dwc);
        w.r += h * T(kRhoW) * T(kLHeat) * (A.drain_rate * wc) * T(1.0 / kEScale);
        w.b += h * T(kRhoW) * T(kLHeat) * A.drain_rate * dwc;
    }
}

// One level's residual (without forcing) and tridiagonal Jacobian row.
template <typename T>
__device__ __forceinline__ Row<T> level_row(const EnthArgs<T>& A, const ColConst<T>& c, int k, T h, T hp, T Ek,
                                            T Ekm1, T Ekp1, T Epk, T om_m, T om_k, T om_p, const Horiz<T>& hz, T Ku,
                                            T Kl) {
    Row<T> w;
    w.r = T(kRhoI) * (h * Ek - hp * Epk) * c.invdt;
    w.b = T(kRhoI) * h * c.invdt;
    w.a = T(0);
    w.c = T(0);
    if (A.flags & 1) {
        T hr, hd;
        horiz_adv(hz, Ek, A.lf_sq, A.dxi, hr, hd);
        w.r += hr;
        w.b += hd;
    }
    column_terms(A, c, k, h, Ek, Ekm1, Ekp1, om_m, om_k, om_p, Ku, Kl, w);
    return w;
}

// Parallel cyclic reduction over the wavefront: lane k holds row k,
// a x_{k-1} + b x_k + c x_{k+1} = d, with identity rows beyond the column;
// returns x_k.  Six doubling steps cover 64 rows.  Each lane forms the
// reciprocal of its own pivot once per step (refined hardware estimate, within
// an ulp or two of the IEEE quotient) and the neighbours read that instead of
// dividing by the shuffled pivot.
template <typename T>
__device__ __forceinline__ T pcr_solve(T a, T b, T c, T d) {
    const int k = lane_id();
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const bool lo = k - s >= 0, hi = k + s < 64;
        const int km = lo ? k - s : k, kp = hi ? k + s : k;
        const T rb = fast_rcp(b);
        T am = lane_get(a, km), rm = lane_get(rb, km), cm = lane_get(c, km), dm = lane_get(d, km);
        T ap = lane_get(a, kp), rp = lane_get(rb, kp), cp = lane_get(c, kp), dp = lane_get(d, kp);
        T al = lo ? -a * rm : T(0);
        T ga = hi ? -c * rp : T(0);
        T na = al * am, nc = ga * cp;
        T nb = b + al * cm + ga * ap;
        T nd = d + al * dm + ga * dp;
        a = na;
        b = nb;
        c = nc;
        d = nd;
    }
    return d * fast_rcp(b);
}

// Column handled by a wavefront.
struct ColIdx {
    long col;
    int i, j, k;
    bool valid;
};

__device__ __forceinline__ ColIdx col_index(long col, int nx, long ncol) {
    ColIdx c;
    c.col = col;
    c.valid = col < ncol;
    c.i = c.valid ? (int)(col / nx) : 0;
    c.j = c.valid ? (int)(col % nx) : 0;
    c.k = lane_id();
    return c;
}
__device__ __forceinline__ long wave_column() { return (long)blockIdx.x * kColsPerBlock + (threadIdx.x >> 6); }

// Residual r = F(E) - f_E (glide/cuda/enthalpy.cu:783-905) on a grid-stride
// loop over columns, and per-block max |r| into partials[blockIdx.x].
template <typename T>
__global__ __launch_bounds__(256) void k_enth_residual(EnthArgs<T> A, T* __restrict__ r_out, int use_forcing,
                                                       double* __restrict__ partials) {
    __shared__ double smax[kColsPerBlock];
    if (skipped(A.skip)) return;
    const int nz = A.nz, k = lane_id();
    const long ncol = (long)A.ny * A.nx;
    const bool in = k < nz;
    double m = 0.0;
    for (long col = wave_column(); col < ncol; col += (long)gridDim.x * kColsPerBlock) {
        const ColIdx q = col_index(col, A.nx, ncol);
        const long base = col * nz;
        const T h = A.H[col];
        const T Ek = in ? A.E[base + k] : T(0);
        const T omk = (in && A.omega) ? A.omega[base + k] : T(0);
        const T Ekm1 = lane_get(Ek, max(k - 1, 0)), Ekp1 = lane_get(Ek, min(k + 1, 63));
        const T omm = lane_get(omk, max(k - 1, 0)), omp = lane_get(omk, min(k + 1, nz - 1));
        T Ku, Kl;
        interface_K(k, nz, h, Ek, Ekp1, Ku, Kl);
        T r = T(0);
        if (in && k < nz - 1 && !(h < A.h_thin)) {
            Horiz<T> hz{};
            if (A.flags & 1) hz = load_horiz(A, q.i, q.j, k);
            Row<T> w = level_row(A, col_const(A, h), k, h, A.H_prev[col], Ek, Ekm1, Ekp1, A.E_prev[base + k], omm, omk,
                                 omp, hz, Ku, Kl);
            r = use_forcing ? w.r - A.f_E[base + k] : w.r;
        }
        if (in && r_out) r_out[base + k] = r;
        m = nan_max(m, fabs((double)r));
    }
    for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_down(m, off, 64));
    if (k == 0) smax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = nan_max(nan_max(smax[0], smax[1]), nan_max(smax[2], smax[3]));
}

// Sweep bookkeeping: mode 0 starts a sweep (r0, zero count, clear the skip
// flag); mode 1 closes one sweep iteration (count, norm, stopping test of
// glide/enthalpy.py:426-471).
__device__ __forceinline__ void record_sweep(EnthStatus* status, int mode, double r, double rtol, double atol) {
    if (mode == 0) {
        status->r0 = r;
        status->r = r;
        status->sweeps = 0;
        status->done = 0;
    } else {
        const int it = status->sweeps;
        if (it < kEnthMaxSweeps) status->norms[it] = r;
        status->sweeps = it + 1;
        status->r = r;
        const double r0 = status->r0;
        // a NaN norm (now or at the start) never stops the sweep
        if (r == r && r0 == r0) {
            const double rel = r0 > 0.0 ? r / r0 : 0.0;
            if (rel < rtol || r < atol) status->done = 1;
        }
    }
}

// Final max over the partials.  With a status record: mode 0 starts a sweep
// (r0, zero count, clear the skip flag); mode 1 closes one sweep iteration
// (count, norm, stopping test of glide/enthalpy.py:426-471) unless the
// sweep already stopped.
__global__ __launch_bounds__(256) void k_enth_reduce(const double* __restrict__ part, int n, double* __restrict__ out,
                                                     EnthStatus* status, int mode, double rtol, double atol) {
    __shared__ double sh[4];
    if (status && mode == 1 && status->done) return;
    double m = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) m = nan_max(m, part[q]);
    for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double r = nan_max(nan_max(sh[0], sh[1]), nan_max(sh[2], sh[3]));
    if (out) *out = r;
    if (status) record_sweep(status, mode, r, rtol, atol);
}

// Maximum over the column pass's max-|r| slots (bit patterns of
// non-negative doubles, or NaN), zeroing them; mode 1 closes one sweep
// iteration as k_enth_reduce does, mode 0 only zeroes.
__global__ __launch_bounds__(256) void k_enth_reduce_slots(unsigned long long* __restrict__ slots, EnthStatus* status,
                                                           int mode, double rtol, double atol) {
    __shared__ double sh[4];
    if (status && mode == 1 && status->done) return;
    double m = 0.0;
    for (int q = threadIdx.x; q < kEnthRmaxSlots; q += blockDim.x) {
        m = nan_max(m, __longlong_as_double((long long)slots[q]));
        slots[q] = 0ull;
    }
    if (mode != 1) return;
    for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double r = nan_max(nan_max(sh[0], sh[1]), nan_max(sh[2], sh[3]));
    if (status) record_sweep(status, 1, r, rtol, atol);
}

// Column Newton smoother (glide/cuda/enthalpy.cu:932-1112): Newton on the
// column with horizontal neighbours frozen at E; writes delta = E_col - E
// and/or E_out = E + omega_relax * delta.
// colour >= 0 (no counterpart in the reference): one phase of the red-black
// column ordering, writing through E_mc instead of E_out, one wavefront per
// column pair of a row.  Phase 0 smooths the columns with (i + j) even and
// copies the others (E_mc = E) unless it runs in place (E_mc == A.E); phase 1
// runs in place on E_mc and smooths the odd columns, whose horizontal
// neighbours are all even, so that they see the new values (a block
// Gauss-Seidel step over columns) and no column of a phase reads another
// column's writes.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_column_smooth(EnthArgs<T> A, T* __restrict__ delta, T* __restrict__ E_out,
                                                            T omega_relax, int n_newton, T relaxation, int colour,
                                                            T* E_mc, unsigned long long* __restrict__ rmax,
                                                            T* __restrict__ r_out) {
    if (skipped(A.skip)) return;
    const long ncol = (long)A.ny * A.nx;
    ColIdx q;
    if (colour < 0) {
        q = col_index(wave_column(), A.nx, ncol);
        if (!q.valid) return;
    } else {
        // wave w: row i, column pair m; its column of the phase's colour and,
        // in phase 0, the copy of the other column of the pair
        const int P = (A.nx + 1) >> 1;
        const long w = wave_column();
        if (w >= (long)A.ny * P) return;
        const int i = (int)(w / P), m = (int)(w % P), p = (i + colour) & 1;
        const int j = 2 * m + p, jo = 2 * m + 1 - p;
        if (colour == 0 && E_mc != A.E && jo < A.nx && lane_id() < A.nz) {
            const long b = ((long)i * A.nx + jo) * A.nz + lane_id();
            E_mc[b] = A.E[b];
        }
        if (j >= A.nx) return;
        q.col = (long)i * A.nx + j;
        q.i = i;
        q.j = j;
        q.k = lane_id();
        q.valid = true;
        E_out = nullptr;
    }
    const int nz = A.nz, k = q.k;
    const bool in = k < nz;
    const long base = q.col * nz;
    const T h = A.H[q.col];
    const T E0 = in ? A.E[base + k] : T(0);
    const T Es = A.E_surface[q.col];
    T El = E0;
    if (h < A.h_thin) {
        El = fmin(Es, e_pmp(T(k) * (T(1) / T(nz - 1)), h));
        if (r_out && in) r_out[base + k] = T(0);
    } else {
        const ColConst<T> cst = col_const(A, h);
        const T hp = A.H_prev[q.col];
        const bool interior = in && k < nz - 1;
        const T Epk = in ? A.E_prev[base + k] : T(0);
        const T fEk = in ? A.f_E[base + k] : T(0);
        const T omk = (in && A.omega) ? A.omega[base + k] : T(0);
        const T omm = lane_get(omk, max(k - 1, 0)), omp = lane_get(omk, min(k + 1, nz - 1));
        Horiz<T> hz{};
        if ((A.flags & 1) && interior) hz = load_horiz(A, q.i, q.j, k);
        double r_first = 0.0;  // |r| of the starting state (first Newton iteration's rows)
        for (int it = 0; it < n_newton; ++it) {
            const T Ekm1 = lane_get(El, max(k - 1, 0)), Ekp1 = lane_get(El, min(k + 1, 63));
            T Ku, Kl;
            interface_K(k, nz, h, El, Ekp1, Ku, Kl);
            T a = T(0), b = T(1), c = T(0), d = T(0);
            if (interior) {
                Row<T> w = level_row(A, cst, k, h, hp, El, Ekm1, Ekp1, Epk, omm, omk, omp, hz, Ku, Kl);
                a = k == 0 ? T(0) : w.a;
                b = w.b;
                c = w.c;
                d = -(w.r - fEk);
                if (it == 0) {
                    r_first = fabs((double)d);
                    if (r_out) r_out[base + k] = -d;
                }
            } else {
                if (it == 0 && r_out && in) r_out[base + k] = T(0);
                if (k == nz - 1) d = -(El - Es);  // Dirichlet surface row
            }
            const T dx = pcr_solve(a, b, c, d);
            if (in) El += relaxation * dx;
        }
        if (rmax && n_newton > 0) {  // every lane of the wavefront is here (h is the column's)
            double m = r_first;
            for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_down(m, off, 64));
            if (k == 0)
                atomicMax(rmax + (wave_column() & (kEnthRmaxSlots - 1)), (unsigned long long)__double_as_longlong(m));
        }
    }
    if (!in) return;
    const T dE = El - E0;
    if (delta) delta[base + k] = dE;
    if (E_out) E_out[base + k] = E0 + omega_relax * dE;
    if (colour >= 0) E_mc[base + k] = E0 + omega_relax * dE;
}

// Pointwise Jacobi layer smoother (glide/cuda/enthalpy.cu:1128-1255):
// delta = -r / J_kk at every interior node, zero on the surface row and in
// thin columns.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_layer_smooth(EnthArgs<T> A, T* __restrict__ delta, T* __restrict__ E_out,
                                                           T omega_relax) {
    if (skipped(A.skip)) return;
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid) return;
    const int nz = A.nz, k = q.k;
    const bool in = k < nz;
    const long base = q.col * nz;
    const T h = A.H[q.col];
    const T Ek = in ? A.E[base + k] : T(0);
    const T omk = (in && A.omega) ? A.omega[base + k] : T(0);
    const T Ekm1 = lane_get(Ek, max(k - 1, 0)), Ekp1 = lane_get(Ek, min(k + 1, 63));
    const T omm = lane_get(omk, max(k - 1, 0)), omp = lane_get(omk, min(k + 1, nz - 1));
    T Ku, Kl;
    interface_K(k, nz, h, Ek, Ekp1, Ku, Kl);
    if (!in) return;
    T out = T(0);
    if (k < nz - 1 && !(h < A.h_thin)) {
        Horiz<T> hz{};
        if (A.flags & 1) hz = load_horiz(A, q.i, q.j, k);
        Row<T> w = level_row(A, col_const(A, h), k, h, A.H_prev[q.col], Ek, Ekm1, Ekp1, A.E_prev[base + k], omm, omk, omp,
                             hz, Ku, Kl);
        const T r = w.r - A.f_E[base + k];
        out = (fabs(w.b) > T(1e-30)) ? -r / w.b : T(0);
    }
    if (delta) delta[base + k] = out;
    if (E_out) E_out[base + k] = Ek + omega_relax * out;
}

template <typename T>
__global__ __launch_bounds__(256) void k_enth_apply(const int* __restrict__ skip, T* __restrict__ E,
                                                    const T* __restrict__ delta, T omega_relax, long n) {
    if (skipped(skip)) return;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        E[q] += omega_relax * delta[q];
}

// omega = H sigma_dot by upward integration of the sigma-space continuity
// equation (glide/cuda/enthalpy.cu:1276-1372): lanes form the per-level
// flux divergence, lane 0 integrates in level order; the result is divided
// by out_div.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_omega(EnthArgs<T> A, const T* __restrict__ dhdt,
                                                    const T* __restrict__ bmb, T out_div, T* __restrict__ omega) {
    __shared__ T sdiv[kColsPerBlock][64];
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid) return;
    const int nz = A.nz, k = q.k, w = threadIdx.x >> 6;
    const int ny = A.ny, nx = A.nx, i = q.i, j = q.j;
    const long base = q.col * nz;
    const T h = A.H[q.col];
    if (h < T(1e-3)) {
        if (k < nz) omega[base + k] = T(0);
        return;
    }
    if (k < nz) {
        const T* u = A.u2d ? A.u2d : A.u3d + (long)k * ny * (nx + 1);
        const T* v = A.u2d ? A.v2d : A.v3d + (long)k * (ny + 1) * nx;
        T fr = 0, fl = 0, fb = 0, ft = 0;
        if (j < nx - 1) {
            T ur = u[(long)i * (nx + 1) + j + 1], Hp = A.H[q.col + 1];
            fr = T(0.5) * (h + Hp) * ur - T(0.5) * sqrt(ur * ur + T(kMassFluxRegYr)) * (Hp - h);
        }
        if (j > 0) {
            T ul = u[(long)i * (nx + 1) + j], Hm = A.H[q.col - 1];
            fl = T(0.5) * (Hm + h) * ul - T(0.5) * sqrt(ul * ul + T(kMassFluxRegYr)) * (h - Hm);
        }
        if (i < ny - 1) {
            T vb = v[(long)(i + 1) * nx + j], Hp = A.H[q.col + nx];
            fb = T(0.5) * (h + Hp) * vb - T(0.5) * sqrt(vb * vb + T(kMassFluxRegYr)) * (h - Hp);
        }
        if (i > 0) {
            T vt = v[(long)i * nx + j], Hm = A.H[q.col - nx];
            ft = T(0.5) * (Hm + h) * vt - T(0.5) * sqrt(vt * vt + T(kMassFluxRegYr)) * (Hm - h);
        }
        sdiv[w][k] = (fr - fl + ft - fb) * A.dxi;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (k == 0) {  // the upward sum in level order, in place over the divergences
        const T dsig = T(1) / T(nz - 1), dh = dhdt[q.col];
        T o = bmb ? -bmb[q.col] : T(0);
        T prev = sdiv[w][0];
        sdiv[w][0] = o;
        for (int m = 1; m < nz; ++m) {
            const T cur = sdiv[w][m];
            o = o - dsig * (dh + T(0.5) * (prev + cur));
            sdiv[w][m] = o;
            prev = cur;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (k < nz) omega[base + k] = sdiv[w][k] / out_div;
}

template <typename T>
__global__ __launch_bounds__(256) void k_enth_broadcast(const T* __restrict__ src, T* __restrict__ dst, long n2d,
                                                        long n, T div) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        dst[q] = src[q % n2d] / div;
}

template <typename T>
__global__ __launch_bounds__(256) void k_enth_tendency(const T* __restrict__ a, const T* __restrict__ b, T div,
                                                       T* __restrict__ out, long n) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        out[q] = (a[q] - b[q]) / div;
}

// Basal frictional heating Q_fh = beta |u_b|^(m+1) from the bed-layer
// velocities averaged to cell centres (glide/model.py:215-228), then the
// forcing f_E of glide/enthalpy.py:473-519.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_rhs(EnthArgs<T> A, const T* __restrict__ beta, T m, T u_reg,
                                                  int frictional, T* __restrict__ Q_fh, T* __restrict__ f_E) {
    // frictional: Q_fh from the velocities (else read); f_E == nullptr: Q_fh only.
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid) return;
    const int nz = A.nz, k = q.k, nx = A.nx;
    const long base = q.col * nz;
    T qfh;
    if (frictional) {
        const T* ub = A.u2d ? A.u2d : A.u3d;
        const T* vb = A.u2d ? A.v2d : A.v3d;
        const T div = A.u2d ? A.vel_div : T(1);
        T uc = T(0.5) * (ub[(long)q.i * (nx + 1) + q.j + 1] / div + ub[(long)q.i * (nx + 1) + q.j] / div);
        T vc = T(0.5) * (vb[(long)(q.i + 1) * nx + q.j] / div + vb[(long)q.i * nx + q.j] / div);
        T speed = sqrt(uc * uc + vc * vc + u_reg * u_reg);
        qfh = beta[q.col] * pow(speed, m + T(1));
        if (k == 0) Q_fh[q.col] = qfh;
    } else {
        qfh = Q_fh[q.col];
    }
    if (k >= nz || !f_E) return;
    const T h = A.H[q.col];
    const T dsig_half = T(0.5) * (T(1) / T(nz - 1));
    T f = A.phi_strain ? h * A.phi_strain[base + k] / T(kEScale) : T(0);
    if (k == 0) f += (A.Q_geo[q.col] + qfh) / (dsig_half * T(kEScale));
    if (k == nz - 1) f = T(0);
    f_E[base + k] = f;
}

// SSA strain heating, uniform over the column (no counterpart in the
// reference, which leaves phi_strain to the caller): 4 eta eps_e^2 rho_i g /
// SEC_PER_YR in W/m^3, with eps_e^2 the momentum solver's cell strain-rate
// invariant of the velocities (m/a) and eta = B/2 (eps_e^2 + eps_reg)^p its
// Glen viscosity (B in head units).
template <typename T>
__global__ __launch_bounds__(256) void k_enth_strain_heating(const T* __restrict__ u, const T* __restrict__ v,
                                                             const T* __restrict__ B, int ny, int nx, int nz, T dxi,
                                                             T glen_p, T eps_reg, T rho_g, T* __restrict__ phi) {
    const long ncol = (long)ny * nx;
    const ColIdx q = col_index(wave_column(), nx, ncol);
    if (!q.valid || q.k >= nz) return;
    const CellVel<T> c = load_cell_vel(u, v, q.i, q.j, ny, nx);
    const T e2 = strain_invariant(c, q.i, q.j, ny, nx, dxi, T(0));
    const T eta = T(0.5) * B[q.col] * pow(e2 + eps_reg, glen_p);
    phi[q.col * nz + q.k] = T(4) * eta * e2 * rho_g / T(kSecPerYr);
}

// Paterson-Budd rate factor A(T) (1 + 181.25 w) averaged over the levels and
// returned as B = <A>^(-1/n) / scale (glide/enthalpy.py:600-643,
// glide/model.py:209-213).
template <typename T>
__global__ __launch_bounds__(256) void k_enth_arrhenius(EnthArgs<T> A, T n_glen, T scale, T* __restrict__ B) {
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid) return;
    const int nz = A.nz, k = q.k;
    T val = T(0);
    if (k < nz) {
        const T h = A.H[q.col];
        const T Tp = t_pmp(T(k) * (T(1) / T(nz - 1)), h);
        const T Ek = A.E[q.col * nz + k];
        const T Tk = fmin(Ek * T(kEScale) / T(kCI) + T(kTRef), Tp);
        const T Af = Tk < T(263.15) ? T(3.985e-13) : T(1.916e3);
        const T Qf = Tk < T(263.15) ? T(60e3) : T(139e3);
        const T Ak = Af * exp(-Qf / (T(8.314) * fmax(Tk, T(1))));
        const T wk = fmax(Ek * T(kEScale) - T(kCI) * (Tp - T(kTRef)), T(0)) / T(kLHeat);
        val = Ak * (T(1) + T(181.25) * wk);
    }
    for (int off = 32; off > 0; off >>= 1) val += __shfl_down(val, off, 64);
    if (k == 0) B[q.col] = pow(val / T(nz), T(-1) / n_glen) / scale;
}

// Temperature min(E E_SCALE / C_I + T_REF, T_pmp) and water content
// max(E E_SCALE - E_pmp, 0) / L (glide/enthalpy.py:565-597).
template <typename T>
__global__ __launch_bounds__(256) void k_enth_diagnostics(EnthArgs<T> A, T* __restrict__ temperature,
                                                          T* __restrict__ water) {
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid || q.k >= A.nz) return;
    const int nz = A.nz, k = q.k;
    const T Tp = t_pmp(T(k) * (T(1) / T(nz - 1)), A.H[q.col]);
    const T Ek = A.E[q.col * nz + k];
    if (temperature) temperature[q.col * nz + k] = fmin(Ek * T(kEScale) / T(kCI) + T(kTRef), Tp);
    if (water) water[q.col * nz + k] = fmax(Ek * T(kEScale) - T(kCI) * (Tp - T(kTRef)), T(0)) / T(kLHeat);
}

// glide/enthalpy.py:534-563
template <typename T>
__global__ __launch_bounds__(256) void k_enth_from_temperature(EnthArgs<T> A, const T* __restrict__ temperature,
                                                               T* __restrict__ E) {
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid || q.k >= A.nz) return;
    const int nz = A.nz, k = q.k;
    const T Tp = t_pmp(T(k) * (T(1) / T(nz - 1)), A.H[q.col]);
    const T Tk = fmin(temperature[q.col * nz + k], Tp);
    E[q.col * nz + k] = T(kCI) * (Tk - T(kTRef)) / T(kEScale);
}

// glide/enthalpy.py:521-532
template <typename T>
__global__ __launch_bounds__(256) void k_enth_surface_from_temperature(const T* __restrict__ Ts, T* __restrict__ Es,
                                                                       long n) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        Es[q] = T(kCI) * (fmin(Ts[q], T(kTMelt)) - T(kTRef)) / T(kEScale);
}

// ------------------------------------------------------------------ adjoint
// Adjoint of the converged solve (no counterpart in the reference): the
// system G(E) = 0 solved by the sweeps has rows R (interior nodes of thick
// columns), E - E_surface (their surface node) and E - min(E_surface, E_pmp)
// (thin columns).  Its exact Jacobian differentiates every term, including
// dK/dE of the smoothed diffusivity and the drainage softplus branches.
template <typename T>
__device__ __forceinline__ T dK_dE(T E, T Ep) {
    T zr = (E - Ep) * T(kEScale) * T(0.01);
    if (zr < T(-20) || zr > T(20)) return T(0);
    T s = T(1) / (T(1) + exp(-zr));
    return T(-0.9) * T(kKCold) * s * (T(1) - s) * T(kEScale) * T(0.01);
}
template <typename T>
__device__ __forceinline__ T dsoftplus_dx(T E, T Ep) {
    T ax = T(kDrainSharpness) * (E - Ep) * T(kEScale);
    if (ax > T(10)) return T(1);
    if (ax < T(-20)) return T(0);
    return T(1) / (T(1) + exp(-ax));
}

template <typename T>
struct JRow {
    T a, b, c, xm, xp, ym, yp;
};

// Exact Jacobian row of an interior node of a thick column; hz the column's
// face fluxes at this level, Ekm1 / Ekp1 the column neighbours' enthalpy.
template <typename T>
__device__ __forceinline__ JRow<T> exact_row(const EnthArgs<T>& A, const Horiz<T>& hz, int k, T h, T Ek, T Ekm1,
                                             T Ekp1, T om_m, T om_k, T om_p) {
    const int nz = A.nz;
    const T dsig = T(1) / T(nz - 1);
    const T hinv = T(1) / h;
    JRow<T> w{T(0), T(kRhoI) * h / A.dt, T(0), T(0), T(0), T(0), T(0)};
    if (A.flags & 1) {
        const T sq = A.lf_sq, f = T(kRhoI) * A.dxi;
        auto dl = [&](T U) { return T(0.5) * (U + fabs(U) * sq); };
        auto dr = [&](T U) { return T(0.5) * (U - fabs(U) * sq); };
        w.xm = -f * dl(hz.Hu_l);
        w.xp = f * dr(hz.Hu_r);
        w.ym = f * dr(hz.Hv_t);
        w.yp = -f * dl(hz.Hv_b);
        w.b += f * ((dl(hz.Hu_r) - dr(hz.Hu_l)) + (dl(hz.Hv_t) - dr(hz.Hv_b)));
    }
    const T Ep_k = e_pmp(T(k) * dsig, h);
    if (k == 0) {
        if (A.flags & 2) {
            T wh = T(0.5) * (om_k + om_p);
            T wp = fmax(wh, T(0)), wn = fmin(wh, T(0));
            T hdi = T(1) / (T(0.5) * dsig);
            w.b += T(kRhoI) * (wp - om_k) * hdi;
            w.c += T(kRhoI) * wn * hdi;
        }
        const T Em = T(0.5) * (Ek + Ekp1), Epm = T(0.5) * (Ep_k + e_pmp(dsig, h));
        const T Kh = diffusivity(Em, Epm), dKh = dK_dE(Em, Epm);
        const T f = -hinv / (dsig * (T(0.5) * dsig));
        w.b += f * (dKh * T(0.5) * (Ekp1 - Ek) - Kh);
        w.c += f * (dKh * T(0.5) * (Ekp1 - Ek) + Kh);
    } else {
        if (A.flags & 2) {
            T wu = T(0.5) * (om_k + om_p), wl = T(0.5) * (om_m + om_k);
            T up = fmax(wu, T(0)), un = fmin(wu, T(0)), lp = fmax(wl, T(0)), ln = fmin(wl, T(0));
            T di = T(1) / dsig;
            w.a -= T(kRhoI) * lp * di;
            w.c += T(kRhoI) * un * di;
            w.b += T(kRhoI) * (up - ln) * di;
        }
        const T Ep_m = e_pmp(T(k - 1) * dsig, h), Ep_p = e_pmp(T(min(k + 1, nz - 1)) * dsig, h);
        const T Eu = T(0.5) * (Ek + Ekp1), Epu = T(0.5) * (Ep_k + Ep_p);
        const T El = T(0.5) * (Ek + Ekm1), Epl = T(0.5) * (Ep_k + Ep_m);
        const T Ku = diffusivity(Eu, Epu), dKu = dK_dE(Eu, Epu), Kl = diffusivity(El, Epl), dKl = dK_dE(El, Epl);
        const T f = -hinv / (dsig * dsig);
        w.c += f * (dKu * T(0.5) * (Ekp1 - Ek) + Ku);
        w.b += f * (dKu * T(0.5) * (Ekp1 - Ek) - Ku - dKl * T(0.5) * (Ek - Ekm1) - Kl);
        w.a += f * (-dKl * T(0.5) * (Ek - Ekm1) + Kl);
    }
    if (A.flags & 4) w.b += h * T(kRhoW) * A.drain_rate * dsoftplus_dx(Ek, Ep_k);
    return w;
}

// ------------------------------------------------------------ adjoint layout
// The adjoint sweeps work one thread per column (a serial Thomas solve down
// the column with precomputed factors), so their planes are kept level-major
// in a checkerboard layout: node (i, j, k) of colour c = (i + j) & 1 sits at
// ((c * nz + k) * ny + i) * hw + (j >> 1), hw = (nx + 1) / 2.  A colour's
// columns are then contiguous along a row at every level, and so are the
// other colour's left, right, upper and lower neighbours of a run of them.
__device__ __forceinline__ long cb_base(int c, int i, int t, int ny, int nz, int hw) {
    return ((long)c * nz * ny + i) * hw + t;
}

// Virtual block vb of nvb: blocks go round-robin over the eight XCDs, so each
// XCD is given one contiguous range of rows and finds the neighbouring rows a
// pass rereads in its own L2.
__device__ __forceinline__ long xcd_block(long vb, long nvb) {
    const long q = nvb / 8, r = nvb % 8, x = vb % 8, l = vb / 8;
    return x * q + (x < r ? x : r) + l;
}

// Transposed horizontal couplings of node (i, j): the coefficient the
// neighbour's row holds for this node's enthalpy (its xp / xm / yp / ym in
// exact_row), from the shared facet's flux, which the node's own face_fluxes
// hold; zero where the neighbour's row is an identity row (thin column; the
// surface level is the caller's).
template <typename T>
struct AdjHoriz {
    T l, r, t, b;
};
template <typename T>
__device__ __forceinline__ AdjHoriz<T> transposed_couplings(const EnthArgs<T>& A, const Horiz<T>& hz, int i, int j) {
    AdjHoriz<T> c{T(0), T(0), T(0), T(0)};
    if (!(A.flags & 1)) return c;
    const int ny = A.ny, nx = A.nx;
    const T sq = A.lf_sq, f = T(kRhoI) * A.dxi;
    auto dl = [&](T U) { return T(0.5) * (U + fabs(U) * sq); };
    auto dr = [&](T U) { return T(0.5) * (U - fabs(U) * sq); };
    auto thick = [&](int ii, int jj) { return !(A.H[(long)ii * nx + jj] < A.h_thin); };
    if (j > 0 && thick(i, j - 1)) c.l = f * dr(hz.Hu_l);
    if (j < nx - 1 && thick(i, j + 1)) c.r = -f * dl(hz.Hu_r);
    if (i > 0 && thick(i - 1, j)) c.t = -f * dl(hz.Hv_t);
    if (i < ny - 1 && thick(i + 1, j)) c.b = f * dr(hz.Hv_b);
    return c;
}

// Factors of every column's transposed tridiagonal block J^T at the converged
// E, one thread per column walking down it: plane 0 the sub-diagonal aT_k =
// J_{k-1,k}, plane 1 the Thomas pivot w_k = b_k - aT_k cp_{k-1}, plane 2 cp_k =
// J_{k+1,k} / w_k; planes 3-6 the transposed horizontal couplings (left,
// right, upper, lower neighbour) per level, or, with a depth-constant
// velocity, once per column on level 0.  Identity rows on surface nodes and
// thin columns.  (Staging the columns through LDS, as k_enth_adj_layout does,
// measured slower here: the per-level Jacobian work wants more waves per CU
// than the staging space allows.)
template <typename T>
__global__ __launch_bounds__(256) void k_enth_adj_factor(EnthArgs<T> A, T* __restrict__ rows) {
    const int ny = A.ny, nx = A.nx, nz = A.nz, hw = (nx + 1) / 2;
    const long ncol = (long)ny * nx, LS = (long)ny * hw, P = 2 * nz * LS;
    const bool flat = A.u2d != nullptr;
    for (long col = (long)blockIdx.x * 256 + threadIdx.x; col < ncol; col += (long)gridDim.x * 256) {
        const int i = (int)(col / nx), j = (int)(col % nx);
        const long o = cb_base((i + j) & 1, i, j >> 1, ny, nz, hw), base = col * nz;
        const T h = A.H[col];
        const bool exact = !(h < A.h_thin);
        Horiz<T> hz{};
        if (A.flags & 1) hz = face_fluxes(A, i, j, 0);
        if (flat) {
            const AdjHoriz<T> c = transposed_couplings(A, hz, i, j);
            rows[3 * P + o] = c.l;
            rows[4 * P + o] = c.r;
            rows[5 * P + o] = c.t;
            rows[6 * P + o] = c.b;
        }
        T Em = T(0), E0 = A.E[base], Ep = nz > 1 ? A.E[base + 1] : T(0);
        T om_m = T(0), om0 = A.omega ? A.omega[base] : T(0), om_p = (A.omega && nz > 1) ? A.omega[base + 1] : T(0);
        T c_prev = T(0), w_prev = T(1), cp_prev = T(0);
        for (int k = 0; k < nz; ++k) {
            if (!flat && k > 0 && (A.flags & 1)) hz = face_fluxes(A, i, j, k);
            JRow<T> w{T(0), T(1), T(0), T(0), T(0), T(0), T(0)};
            if (k < nz - 1 && exact) w = exact_row(A, hz, k, h, E0, Em, Ep, om_m, om0, om_p);
            const long ok = o + k * LS;
            if (k > 0) {  // J_{k,k-1} completes column entry k - 1
                cp_prev = w.a / w_prev;
                rows[2 * P + ok - LS] = cp_prev;
            }
            const T aT = k > 0 ? c_prev : T(0);
            const T wk = w.b - aT * cp_prev;
            rows[ok] = aT;
            rows[P + ok] = wk;
            c_prev = w.c;
            w_prev = wk;
            if (!flat) {
                AdjHoriz<T> c{T(0), T(0), T(0), T(0)};
                if (k < nz - 1) c = transposed_couplings(A, hz, i, j);
                rows[3 * P + ok] = c.l;
                rows[4 * P + ok] = c.r;
                rows[5 * P + ok] = c.t;
                rows[6 * P + ok] = c.b;
            }
            Em = E0;
            E0 = Ep;
            Ep = k + 2 < nz ? A.E[base + k + 2] : T(0);
            om_m = om0;
            om0 = om_p;
            om_p = (A.omega && k + 2 < nz) ? A.omega[base + k + 2] : om0;
        }
        rows[2 * P + o + (nz - 1) * LS] = T(0);
    }
}

// Load at a byte offset from a wave-uniform base: the address is then a
// scalar base plus a 32-bit per-lane offset, and the level stride moves into
// the scalar base.
template <typename T>
__device__ __forceinline__ T ld_at(const T* __restrict__ base, unsigned byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

constexpr int kAdjChunk = 4;  // levels whose loads a thread issues together

// One colour's columns against J^T lam = g (lam in the checkerboard layout):
// the residual g - J^T lam of each column before its update (max |.| into
// *acc) and, with SOLVE, the column's transposed block solved
// by the stored Thomas factors against g minus the horizontal coupling of the
// other colour's multipliers, which this pass does not change; lam = lam +
// omega_relax (lam_col - lam).  NZB >= nz bounds the forward elimination kept
// in registers.  A thread issues the loads of kAdjChunk levels at once
// (branch-free: a missing neighbour reads the thread's own node against a zero
// coupling), so that several memory latencies overlap down the column.  Byte
// offsets within one checkerboard plane fit 32 bits (checked by the host).
// acc[0] collects the maximum (bit patterns of non-negative doubles, atomic
// max), acc[1] counts finished blocks; with close_mode >= 0 the last block
// records the sweep (record_sweep) and clears both.
template <typename T, int NZB, bool SOLVE>
__global__ __launch_bounds__(256) void k_enth_adj_pass(const int* __restrict__ skip, int ny, int nx, int nz, int flat,
                                                       const T* __restrict__ rows, const T* __restrict__ g,
                                                       T* __restrict__ lam, T omega_relax, int color,
                                                       unsigned long long* acc, EnthStatus* status,
                                                       int close_mode, double rtol, double atol) {
    constexpr int KC = kAdjChunk;
    __shared__ double smax[4];
    if (skipped(skip)) return;
    const int hw = (nx + 1) / 2;
    const long LS = (long)ny * hw, P = 2 * nz * LS, n = LS;
    const T* __restrict__ rA = rows;
    const T* __restrict__ rW = rows + P;
    const T* __restrict__ rC = rows + 2 * P;
    double m = 0.0;
    const long w = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const int i = (int)(w / hw), t = (int)(w % hw);
    const int j = 2 * t + ((i + color) & 1);
    if (w < n && j < nx) {
        const long own = cb_base(color, i, t, ny, nz, hw), oth = cb_base(color ^ 1, i, 0, ny, nz, hw);
        const unsigned b_own = (unsigned)own * sizeof(T);
        const unsigned b_l = (unsigned)(j > 0 ? oth + ((j - 1) >> 1) : own) * sizeof(T);
        const unsigned b_r = (unsigned)(j + 1 < nx ? oth + ((j + 1) >> 1) : own) * sizeof(T);
        const unsigned b_u = (unsigned)(i > 0 ? oth - hw + t : own) * sizeof(T);
        const unsigned b_d = (unsigned)(i + 1 < ny ? oth + hw + t : own) * sizeof(T);
        T cl = T(0), cr = T(0), ct = T(0), cb = T(0);
        if (flat) {  // zero on the domain edge, where the offsets above fall back to the own node
            cl = ld_at(rows + 3 * P, b_own);
            cr = ld_at(rows + 4 * P, b_own);
            ct = ld_at(rows + 5 * P, b_own);
            cb = ld_at(rows + 6 * P, b_own);
        }
        T dp[NZB];
        T lm = T(0), l0 = ld_at(lam, b_own), cpm = T(0), dprev = T(0);
#pragma unroll
        for (int k0 = 0; k0 < NZB; k0 += KC) {
            if (k0 < nz) {
                T a[KC], wv[KC], c[KC], gv[KC], lp[KC], xl[KC], xr[KC], xu[KC], xd[KC];
                T ql[KC], qr[KC], qu[KC], qd[KC];
#pragma unroll
                for (int q = 0; q < KC; ++q) {
                    const long lk = (long)min(k0 + q, nz - 1) * LS, lp1 = (long)min(k0 + q + 1, nz - 1) * LS;
                    a[q] = ld_at(rA + lk, b_own);
                    wv[q] = ld_at(rW + lk, b_own);
                    c[q] = ld_at(rC + lk, b_own);
                    gv[q] = ld_at(g + lk, b_own);
                    lp[q] = ld_at(lam + lp1, b_own);
                    xl[q] = ld_at(lam + lk, b_l);
                    xr[q] = ld_at(lam + lk, b_r);
                    xu[q] = ld_at(lam + lk, b_u);
                    xd[q] = ld_at(lam + lk, b_d);
                    if (!flat) {
                        ql[q] = ld_at(rows + 3 * P + lk, b_own);
                        qr[q] = ld_at(rows + 4 * P + lk, b_own);
                        qu[q] = ld_at(rows + 5 * P + lk, b_own);
                        qd[q] = ld_at(rows + 6 * P + lk, b_own);
                    }
                }
#pragma unroll
                for (int q = 0; q < KC; ++q) {
                    const int k = k0 + q;
                    if (k < nz) {
                        const T lpk = k + 1 < nz ? lp[q] : T(0);
                        T hs = T(0);
                        if (k < nz - 1) {  // every neighbour's surface row is an identity row
                            if (flat) hs = cl * xl[q] + cr * xr[q] + ct * xu[q] + cb * xd[q];
                            else hs = ql[q] * xl[q] + qr[q] * xr[q] + qu[q] * xu[q] + qd[q] * xd[q];
                        }
                        const T rhs = gv[q] - hs;
                        const T b = wv[q] + a[q] * cpm, cT = c[q] * wv[q];
                        m = nan_max(m, fabs((double)(rhs - (a[q] * lm + b * l0 + cT * lpk))));
                        if (SOLVE) {
                            dprev = (rhs - a[q] * dprev) / wv[q];
                            dp[k] = dprev;
                        }
                        lm = l0;
                        l0 = lpk;
                        cpm = c[q];
                    }
                }
            }
        }
        if (SOLVE) {
            T x = T(0);
#pragma unroll
            for (int k1 = NZB - 1; k1 >= 0; k1 -= KC) {
                if (k1 - KC + 1 < nz) {
                    T c[KC], lo[KC];
#pragma unroll
                    for (int q = 0; q < KC; ++q) {
                        const long lk = (long)max(min(k1 - q, nz - 1), 0) * LS;
                        c[q] = ld_at(rC + lk, b_own);
                        lo[q] = ld_at(lam + lk, b_own);
                    }
#pragma unroll
                    for (int q = 0; q < KC; ++q) {
                        const int k = k1 - q;
                        if (k >= 0 && k < nz) {
                            x = dp[k] - c[q] * x;
                            *reinterpret_cast<T*>(reinterpret_cast<char*>(lam + (long)k * LS) + b_own) =
                                lo[q] + omega_relax * (x - lo[q]);
                        }
                    }
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    // a NaN maximum (positive quiet NaN from fabs) also wins the unsigned
    // comparison of bit patterns, so it survives the accumulation
    const unsigned long long prev = atomicMax(
        acc, (unsigned long long)__double_as_longlong(nan_max(nan_max(smax[0], smax[1]), nan_max(smax[2], smax[3]))));
    if (close_mode < 0) return;
    // waiting for the returned value orders this block's maximum before its
    // count; all of them are device-scope atomics on the same two words
    asm volatile("" ::"v"(prev));
    if (atomicAdd(acc + 1, 1ull) != (unsigned long long)gridDim.x - 1) return;
    const double r = __longlong_as_double((long long)atomicExch(acc, 0ull));
    atomicExch(acc + 1, 0ull);
    record_sweep(status, close_mode, r, rtol, atol);
}

// Threads (columns) per block of k_enth_adj_layout: the block's columns are
// one contiguous chunk of the (ny, nx, nz) layout, staged through LDS by
// coalesced loads.
constexpr int kStageCols = 64;

// Columns [c0, c0 + ncols) of a (ny, nx, nz) array into s[q * nz + k].
template <typename T>
__device__ __forceinline__ void stage_columns(T* s, const T* __restrict__ src, long c0, int ncols, int nz) {
    const long base = c0 * nz;
    const int n = ncols * nz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) s[q] = src[base + q];
}

// lam between the (ny, nx, nz) layout and the checkerboard layout, one thread
// per column, the (ny, nx, nz) side staged through LDS (kStageCols * nz
// doubles of dynamic LDS).
template <typename T>
__global__ __launch_bounds__(kStageCols) void k_enth_adj_layout(int ny, int nx, int nz, const T* __restrict__ src,
                                                                T* __restrict__ dst, int to_cb) {
    extern __shared__ double lds[];
    T* sc = reinterpret_cast<T*>(lds);
    const int hw = (nx + 1) / 2;
    const long ncol = (long)ny * nx, LS = (long)ny * hw;
    const long c0 = (long)blockIdx.x * kStageCols;
    const int ncols = (int)(ncol - c0 < kStageCols ? ncol - c0 : kStageCols);
    const long col = c0 + threadIdx.x;
    const bool live = (int)threadIdx.x < ncols;
    const int i = (int)(col / nx), j = (int)(col % nx);
    const long o = cb_base((i + j) & 1, i, j >> 1, ny, nz, hw);
    if (to_cb) {
        stage_columns(sc, src, c0, ncols, nz);
        __syncthreads();
        if (live)
            for (int k = 0; k < nz; ++k) dst[o + k * LS] = sc[threadIdx.x * nz + k];
    } else {
        if (live)
            for (int k = 0; k < nz; ++k) sc[threadIdx.x * nz + k] = src[o + k * LS];
        __syncthreads();
        const long base = c0 * nz;
        for (int q = threadIdx.x; q < ncols * nz; q += blockDim.x) dst[base + q] = sc[q];
    }
}

// Rate-factor term of level k: A_k (1 + 181.25 w_k) and its derivative with
// respect to E_k.
template <typename T>
struct RateTerm {
    T val, dA;
};
template <typename T>
__device__ __forceinline__ RateTerm<T> rate_term(T Ek, T h, int k, int nz) {
    const T Tp = t_pmp(T(k) * (T(1) / T(nz - 1)), h);
    const T Tu = Ek * T(kEScale) / T(kCI) + T(kTRef);
    const T Tk = fmin(Tu, Tp);
    const T Af = Tk < T(263.15) ? T(3.985e-13) : T(1.916e3);
    const T Qf = Tk < T(263.15) ? T(60e3) : T(139e3);
    const T Ts = fmax(Tk, T(1));
    const T Ak = Af * exp(-Qf / (T(8.314) * Ts));
    const T wraw = Ek * T(kEScale) - T(kCI) * (Tp - T(kTRef));
    const T wk = fmax(wraw, T(0)) / T(kLHeat);
    const T dT = (Tu < Tp && Tk > T(1)) ? T(kEScale) / T(kCI) : T(0);
    const T dw = wraw > T(0) ? T(kEScale) / T(kLHeat) : T(0);
    RateTerm<T> r;
    r.val = Ak * (T(1) + T(181.25) * wk);
    r.dA = Ak * Qf / (T(8.314) * Ts * Ts) * dT * (T(1) + T(181.25) * wk) + Ak * T(181.25) * dw;
    return r;
}

// Right-hand side g = dL/dE + (dB/dE)^T dL/dB of the rate factor
// B = mean_k(A_k (1 + 181.25 w_k))^(-1/n) / scale, into the checkerboard
// layout, one thread per column.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_adj_rhs(EnthArgs<T> A, const T* __restrict__ gE,
                                                      const T* __restrict__ gB, T n_glen, T scale,
                                                      T* __restrict__ g) {
    const int ny = A.ny, nx = A.nx, nz = A.nz, hw = (nx + 1) / 2;
    const long ncol = (long)ny * nx, LS = (long)ny * hw;
    for (long col = (long)blockIdx.x * 256 + threadIdx.x; col < ncol; col += (long)gridDim.x * 256) {
        const int i = (int)(col / nx), j = (int)(col % nx);
        const long o = cb_base((i + j) & 1, i, j >> 1, ny, nz, hw), base = col * nz;
        const T h = A.H[col];
        T fac = T(0);
        if (gB) {
            T sum = T(0);
            for (int k = 0; k < nz; ++k) sum += rate_term(A.E[base + k], h, k, nz).val;
            const T mean = sum / T(nz);
            const T dBdmean = (T(-1) / n_glen) * pow(mean, T(-1) / n_glen - T(1)) / scale;
            fac = gB[col] * dBdmean / T(nz);
        }
        for (int k = 0; k < nz; ++k) {
            T out = gE ? gE[base + k] : T(0);
            if (gB) out += fac * rate_term(A.E[base + k], h, k, nz).dA;
            g[o + k * LS] = out;
        }
    }
}

// Parameter gradients dL/dp = -lam^T dG/dp: E_prev, phi_strain (3-D), Q_geo
// (equal to the Q_fh gradient), E_surface (2-D).
template <typename T>
__global__ __launch_bounds__(256) void k_enth_adj_grads(EnthArgs<T> A, const T* __restrict__ lam,
                                                        T* __restrict__ g_E_prev, T* __restrict__ g_phi,
                                                        T* __restrict__ g_Q, T* __restrict__ g_Es) {
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid) return;
    const int nz = A.nz, k = q.k;
    const bool in = k < nz;
    const long n = q.col * nz + k;
    const T h = A.H[q.col];
    const bool thin = h < A.h_thin;
    const T l = in ? lam[n] : T(0);
    const bool live = in && !thin && k < nz - 1;
    if (in && g_E_prev) g_E_prev[n] = live ? l * T(kRhoI) * A.H_prev[q.col] / A.dt : T(0);
    if (in && g_phi) g_phi[n] = live ? l * h / T(kEScale) : T(0);
    const T l0 = lane_get(l, 0), ltop = lane_get(l, nz - 1);
    T contrib = T(0);
    if (in && thin) {
        const T Es = A.E_surface[q.col];
        contrib = (Es < e_pmp(T(k) * (T(1) / T(nz - 1)), h)) ? l : T(0);
    }
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off, 64);
    if (k == 0) {
        if (g_Q) g_Q[q.col] = thin ? T(0) : l0 / (T(0.5) * (T(1) / T(nz - 1)) * T(kEScale));
        if (g_Es) g_Es[q.col] = thin ? contrib : ltop;
    }
}

// ------------------------------------------------- thermal-step input gradients
// dL/dp = -lam^T dG/dp for the inputs the momentum step hands the thermal
// step (glide_thermal_step): the depth-constant velocities u, v (m/a), the new
// and previous thicknesses H, H_prev, the sliding coefficient beta (through
// the frictional heating) and the rate factor the strain heating read.  Each
// input reaches G directly (time term, diffusion through 1/H and the
// pressure-melting enthalpy, drainage, h phi forcing, facet mass fluxes of the
// horizontal advection) and through omega (upward integration of the
// thickness tendency and the facet-flux divergence), Q_fh and phi_strain; the
// new rate factor depends on H through the pressure-melting temperature.
//
// Two passes, both gathers (no atomics, bitwise reproducible): one wavefront
// per column reduces its levels' partials to per-column values (direct H
// terms, the omega tendency adjoint, the adjoint of each of its four facet
// fluxes, strain-rate and basal-speed weights) in kThermalVjpPlanes (ny, nx)
// planes; then one thread per facet / cell gathers from the cells around it.
template <typename T>
__device__ __forceinline__ T depmp_dh(int k, int nz) {
    return -T(kClausius) * (T(1) - T(k) * (T(1) / T(nz - 1))) * T(kCI / kEScale);
}

// Derivative of level k's rate-factor term A_k (1 + 181.25 w_k) with respect
// to the column thickness (through the pressure-melting temperature).
template <typename T>
__device__ __forceinline__ T rate_term_dh(T Ek, T h, int k, int nz) {
    const T sig = T(k) * (T(1) / T(nz - 1));
    const T Tp = t_pmp(sig, h), dTp = -T(kClausius) * (T(1) - sig);
    const T Tu = Ek * T(kEScale) / T(kCI) + T(kTRef);
    const T Tk = fmin(Tu, Tp);
    const T Af = Tk < T(263.15) ? T(3.985e-13) : T(1.916e3);
    const T Qf = Tk < T(263.15) ? T(60e3) : T(139e3);
    const T Ts = fmax(Tk, T(1));
    const T Ak = Af * exp(-Qf / (T(8.314) * Ts));
    const T dTs = (!(Tu < Tp) && Tk > T(1)) ? dTp : T(0);
    const T wraw = Ek * T(kEScale) - T(kCI) * (Tp - T(kTRef));
    const T wk = fmax(wraw, T(0)) / T(kLHeat);
    const T dwk = wraw > T(0) ? -T(kCI) * dTp / T(kLHeat) : T(0);
    return Ak * Qf / (T(8.314) * Ts * Ts) * dTs * (T(1) + T(181.25) * wk) + Ak * T(181.25) * dwk;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// d/dU of the Lax-Friedrichs flux 0.5 U (El + Er) - 0.5 |U| sq (Er - El).
template <typename T>
__device__ __forceinline__ T lf_flux_dU(T U, T El, T Er, T sq) {
    const T sgn = U > T(0) ? T(1) : (U < T(0) ? T(-1) : T(0));
    return T(0.5) * (El + Er) - T(0.5) * sgn * sq * (Er - El);
}

template <typename T>
__global__ __launch_bounds__(256) void k_thermal_vjp_columns(EnthArgs<T> A, ThermalVjpArgs<T> V) {
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid) return;
    const int nz = A.nz, k = q.k, ny = A.ny, nx = A.nx, i = q.i, j = q.j;
    const bool in = k < nz;
    const long base = q.col * nz, col = q.col;
    const T h = A.H[col];
    const bool thin = h < A.h_thin;
    const bool live = in && !thin && k < nz - 1;
    const T dsig = T(1) / T(nz - 1);
    const T l = in ? V.lam[base + k] : T(0);
    const T Ek = in ? A.E[base + k] : T(0);
    const T omk = (in && A.omega) ? A.omega[base + k] : T(0);
    const T Ekm1 = lane_get(Ek, max(k - 1, 0)), Ekp1 = lane_get(Ek, min(k + 1, 63));
    const T omm = lane_get(omk, max(k - 1, 0)), omp = lane_get(omk, min(k + 1, nz - 1));
    T Ku, Kl;
    interface_K(k, nz, h, Ek, Ekp1, Ku, Kl);
    // interface diffusivities' thickness derivatives (through E_pmp)
    const int kp = min(k + 1, nz - 1), kc = min(k, nz - 1);
    const T Ep_k = e_pmp(T(kc) * dsig, h), Ep_p = e_pmp(T(kp) * dsig, h);
    const T dEp_k = depmp_dh<T>(kc, nz), dEp_p = depmp_dh<T>(kp, nz);
    const T dKu = -dK_dE(T(0.5) * (Ek + Ekp1), T(0.5) * (Ep_k + Ep_p)) * T(0.5) * (dEp_k + dEp_p);
    const T dKl = lane_get(dKu, max(k - 1, 0));
    // lam_k times the row's partials: H, H_prev, omega_k / omega_k+1 /
    // omega_k-1, and the four facet fluxes
    T gh = T(0), ghp = T(0), cs = T(0), cu = T(0), cd = T(0), fl = T(0), fr = T(0), ft = T(0), fb = T(0);
    if (live) {
        const T hinv = T(1) / h;
        T drdh = T(kRhoI) * Ek / A.dt;
        if (k == 0) {
            const T idd = T(1) / (dsig * (T(0.5) * dsig));
            drdh += hinv * hinv * idd * Ku * (Ekp1 - Ek) - hinv * idd * dKu * (Ekp1 - Ek);
        } else {
            const T d2 = T(1) / (dsig * dsig);
            drdh += hinv * hinv * d2 * (Ku * (Ekp1 - Ek) - Kl * (Ek - Ekm1)) -
                    hinv * d2 * (dKu * (Ekp1 - Ek) - dKl * (Ek - Ekm1));
        }
        if (A.flags & 4) {
            T wc, dwc;
            water_and_slope(Ek, Ep_k, wc, dwc);
            drdh += T(kRhoW) * T(kLHeat) * A.drain_rate * wc * T(1.0 / kEScale) -
                    h * T(kRhoW) * A.drain_rate * dsoftplus_dx(Ek, Ep_k) * dEp_k;
        }
        if (A.phi_strain) drdh -= A.phi_strain[base + k] * T(1.0 / kEScale);
        gh = -l * drdh;
        ghp = l * T(kRhoI) * A.E_prev[base + k] / A.dt;
        if (A.flags & 2) {
            if (k == 0) {
                const T hdi = T(1) / (T(0.5) * dsig);
                const T D = (T(0.5) * (omk + omp) > T(0)) ? Ek : Ekp1;
                cs = l * T(kRhoI) * hdi * (T(0.5) * D - Ek);
                cu = l * T(kRhoI) * hdi * T(0.5) * D;
            } else {
                const T di = T(1) / dsig;
                const T Du = (T(0.5) * (omk + omp) > T(0)) ? Ek : Ekp1;
                const T Dl = (T(0.5) * (omm + omk) > T(0)) ? Ekm1 : Ek;
                cs = l * T(kRhoI) * di * T(0.5) * (Du - Dl);
                cu = l * T(kRhoI) * di * T(0.5) * Du;
                cd = -l * T(kRhoI) * di * T(0.5) * Dl;
            }
        }
        if (A.flags & 1) {
            const Horiz<T> hz = load_horiz(A, i, j, k);
            const T f = l * T(kRhoI) * A.dxi, sq = A.lf_sq;
            fl = -f * lf_flux_dU(hz.Hu_l, hz.E_xm, Ek, sq);
            fr = f * lf_flux_dU(hz.Hu_r, Ek, hz.E_xp, sq);
            ft = f * lf_flux_dU(hz.Hv_t, Ek, hz.E_ym, sq);
            fb = -f * lf_flux_dU(hz.Hv_b, hz.E_yp, Ek, sq);
        }
    }
    // clamped thin column: G = E - min(E_surface, E_pmp(h))
    if (in && thin && !(A.E_surface[col] < Ep_k)) gh = l * dEp_k;
    // omega_m = -m dsig (dh/dt + div) / SEC_PER_YR (no basal melt): the
    // tendency and the divergence share one adjoint
    const T cu_below = lane_get(cu, max(k - 1, 0)), cd_above = lane_get(cd, min(k + 1, 63));
    const T gom = -(cs + (k > 0 ? cu_below : T(0)) + cd_above);
    const T gdh = (in && !(h < T(1e-3))) ? gom * (-T(k) * dsig / T(kSecPerYr)) : T(0);
    T rv = T(0), rdh = T(0);
    if (V.gB && in) {
        rv = rate_term(Ek, h, k, nz).val;
        rdh = rate_term_dh(Ek, h, k, nz);
    }
    const T Sgh = wave_sum(gh), Sghp = wave_sum(ghp), Sgdh = wave_sum(gdh);
    const T Sfl = wave_sum(fl), Sfr = wave_sum(fr), Sft = wave_sum(ft), Sfb = wave_sum(fb);
    const T Srv = wave_sum(rv), Srdh = wave_sum(rdh);
    const T Sphi = wave_sum((V.B_in && V.g_phi && in) ? V.g_phi[base + k] : T(0));
    if (k != 0) return;
    const long P = ncol;
    T* W = V.work;
    T ghd = Sgh;
    if (V.gB) {
        const T mean = Srv / T(nz);
        const T dBdmean = (T(-1) / V.n_glen) * pow(mean, T(-1) / V.n_glen - T(1)) / V.B_scale;
        ghd += V.gB[col] * dBdmean / T(nz) * Srdh;
    }
    W[col] = ghd;
    W[P + col] = Sgdh;
    // facet adjoints in the units of the m/a mass flux the omega divergence
    // uses; the advection's flux is that flux over SEC_PER_YR
    const T vs = T(1) / T(kSecPerYr), dxi = A.dxi;
    W[2 * P + col] = -Sfl * vs - Sgdh * dxi;
    W[3 * P + col] = -Sfr * vs + Sgdh * dxi;
    W[4 * P + col] = -Sft * vs + Sgdh * dxi;
    W[5 * P + col] = -Sfb * vs - Sgdh * dxi;
    if (V.g_H_prev) V.g_H_prev[col] = Sghp - Sgdh / V.dt_yr;
    if (V.B_in) {  // phi = 4 eta eps_e^2 rho_i g / SEC_PER_YR, eta = B/2 (eps_e^2 + eps_reg)^p
        const CellVel<T> c = load_cell_vel(A.u2d, A.v2d, i, j, ny, nx);
        const T X = (c.ur - c.ul) * dxi, Y = (c.vt - c.vb) * dxi;
        T stl = T(0), str = T(0), sbl = T(0), sbr = T(0);
        if (i > 0 && j > 0) stl = T(0.5) * ((c.utl - c.ul) * dxi + (c.vt - c.vlt) * dxi);
        if (i > 0 && j < nx - 1) str = T(0.5) * ((c.utr - c.ur) * dxi + (c.vrt - c.vt) * dxi);
        if (i < ny - 1 && j > 0) sbl = T(0.5) * ((c.ul - c.ubl) * dxi + (c.vb - c.vlb) * dxi);
        if (i < ny - 1 && j < nx - 1) sbr = T(0.5) * ((c.ur - c.ubr) * dxi + (c.vrb - c.vb) * dxi);
        const T e2 = strain_invariant(c, i, j, ny, nx, dxi, T(0));
        const T pw = pow(e2 + V.eps_reg, V.glen_p);
        const T coef = T(4) * V.rho_g / T(kSecPerYr);
        const T G = Sphi * coef * T(0.5) * V.B_in[col] * (pw + V.glen_p * e2 * pw / (e2 + V.eps_reg));
        W[6 * P + col] = G * dxi * (T(2) * X + Y);
        W[7 * P + col] = G * dxi * (T(2) * Y + X);
        W[8 * P + col] = G * T(0.25) * dxi * stl;
        W[9 * P + col] = G * T(0.25) * dxi * str;
        W[10 * P + col] = G * T(0.25) * dxi * sbl;
        W[11 * P + col] = G * T(0.25) * dxi * sbr;
        if (V.g_B) V.g_B[col] = Sphi * coef * T(0.5) * pw * e2;
    }
    if (V.beta) {  // Q_fh = beta |u_b|^(m+1), u_b the cell-centred velocity in m/s
        const T gq = V.g_Q[col], yr = T(kSecPerYr);
        const T uc = T(0.5) * (A.u2d[(long)i * (nx + 1) + j + 1] / yr + A.u2d[(long)i * (nx + 1) + j] / yr);
        const T vc = T(0.5) * (A.v2d[(long)(i + 1) * nx + j] / yr + A.v2d[(long)i * nx + j] / yr);
        const T speed = sqrt(uc * uc + vc * vc + V.u_reg * V.u_reg);
        if (V.g_beta) V.g_beta[col] = gq * pow(speed, V.m + T(1));
        const T FQ = gq * V.beta[col] * (V.m + T(1)) * pow(speed, V.m - T(1));
        W[12 * P + col] = FQ * uc * T(0.5) / yr;
        W[13 * P + col] = FQ * vc * T(0.5) / yr;
    }
}

// One thread per facet / cell over the (ny + 1, nx + 1) corners: x facet
// (i, j) for i < ny, y facet (i, j) for j < nx, cell (i, j) for both.
template <typename T>
__global__ __launch_bounds__(256) void k_thermal_vjp_gather(EnthArgs<T> A, ThermalVjpArgs<T> V, int strain,
                                                            int frictional) {
    const int ny = A.ny, nx = A.nx;
    const long P = (long)ny * nx, n = (long)(ny + 1) * (nx + 1);
    const T* W = V.work;
    const T reg = T(kMassFluxRegYr);
    auto w = [&](int p, int ii, int jj) -> T {
        return (ii >= 0 && ii < ny && jj >= 0 && jj < nx) ? W[p * P + (long)ii * nx + jj] : T(0);
    };
    auto Hc = [&](int ii, int jj) { return A.H[(long)ii * nx + jj]; };
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t / (nx + 1)), j = (int)(t % (nx + 1));
        if (i < ny) {  // x facet between cells (i, j - 1) and (i, j)
            T g = T(0);
            if (j > 0 && j < nx) {
                const T U = A.u2d[(long)i * (nx + 1) + j], s = sqrt(U * U + reg);
                const T HA = Hc(i, j - 1), HB = Hc(i, j);
                g += (w(3, i, j - 1) + w(2, i, j)) * (T(0.5) * (HA + HB) - T(0.5) * U / s * (HB - HA));
            }
            if (strain)
                g += -w(6, i, j) - w(8, i, j) + w(10, i, j) + w(6, i, j - 1) - w(9, i, j - 1) + w(11, i, j - 1) +
                     w(8, i + 1, j) + w(9, i + 1, j - 1) - w(10, i - 1, j) - w(11, i - 1, j - 1);
            if (frictional) g += w(12, i, j - 1) + w(12, i, j);
            V.g_u[(long)i * (nx + 1) + j] = g;
        }
        if (j < nx) {  // y facet between cells (i - 1, j) and (i, j)
            T g = T(0);
            if (i > 0 && i < ny) {
                const T Vv = A.v2d[(long)i * nx + j], s = sqrt(Vv * Vv + reg);
                const T HA = Hc(i - 1, j), HB = Hc(i, j);
                g += (w(5, i - 1, j) + w(4, i, j)) * (T(0.5) * (HA + HB) - T(0.5) * Vv / s * (HA - HB));
            }
            if (strain)
                g += w(7, i, j) + w(8, i, j) - w(9, i, j) - w(7, i - 1, j) + w(10, i - 1, j) - w(11, i - 1, j) -
                     w(8, i, j + 1) + w(9, i, j - 1) - w(10, i - 1, j + 1) + w(11, i - 1, j - 1);
            if (frictional) g += w(13, i - 1, j) + w(13, i, j);
            V.g_v[(long)i * nx + j] = g;
        }
        if (i < ny && j < nx) {  // the cell's thickness: direct terms, tendency, its four facet fluxes
            T g = w(0, i, j) + w(1, i, j) / V.dt_yr;
            if (j > 0) {
                const T U = A.u2d[(long)i * (nx + 1) + j], s = sqrt(U * U + reg);
                g += (w(3, i, j - 1) + w(2, i, j)) * (T(0.5) * U - T(0.5) * s);
            }
            if (j < nx - 1) {
                const T U = A.u2d[(long)i * (nx + 1) + j + 1], s = sqrt(U * U + reg);
                g += (w(3, i, j) + w(2, i, j + 1)) * (T(0.5) * U + T(0.5) * s);
            }
            if (i > 0) {
                const T Vv = A.v2d[(long)i * nx + j], s = sqrt(Vv * Vv + reg);
                g += (w(5, i - 1, j) + w(4, i, j)) * (T(0.5) * Vv + T(0.5) * s);
            }
            if (i < ny - 1) {
                const T Vv = A.v2d[(long)(i + 1) * nx + j], s = sqrt(Vv * Vv + reg);
                g += (w(5, i, j) + w(4, i + 1, j)) * (T(0.5) * Vv - T(0.5) * s);
            }
            V.g_H[(long)i * nx + j] = g;
        }
    }
}

// The rows' exponential and logarithm over an array (for tests of their
// accuracy): e_out = exp_rows(x), l_out = log_rows(x) where x > 0 (else NaN).
__global__ __launch_bounds__(256) void k_enth_elementary(const double* __restrict__ x, long n,
                                                         double* __restrict__ e_out, double* __restrict__ l_out) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const double v = x[q];
        if (e_out) e_out[q] = exp_rows(v);
        if (l_out) l_out[q] = v > 0.0 ? log_rows(v) : __builtin_nan("");
    }
}
hipError_t launch_enth_elementary(const double* x, long n, double* e_out, double* l_out, hipStream_t st) {
    const long b = (n + 255) / 256;
    k_enth_elementary<<<(unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096), 256, 0, st>>>(x, n, e_out, l_out);
    return hipGetLastError();
}

// ------------------------------------------------ level-major fused sweep
// (no counterpart in the reference, whose sweep works on the (ny, nx, nz)
// arrays, glide/enthalpy.py:426-471).  The sweep of a thermal step runs on
// (nz, ny * nx) copies of E, E_prev, f_E and omega, in which level k of a run
// of columns is contiguous.  The column pass gives a column P consecutive
// lanes (4, 8 or 16), lane q holding levels [q MB, (q + 1) MB), MB = 3 or 4
// (levels from nz on are identity rows), and solves the column's tridiagonal
// Newton system by a partitioned Thomas algorithm: each lane eliminates its
// block down to two rows coupling the block's first and last unknowns to the
// neighbouring blocks', the first rows are eliminated into one row per lane
// in the blocks' last unknowns, which cyclic reduction across the P lanes
// solves, and each lane substitutes back.  Neighbouring lanes exchange values
// through DPP row shifts, not LDS.  The layer pass and the residual take one
// thread per column and stream up the levels.  The rows are level_row's, with
// the terms that do not depend on the column's own enthalpy (the previous
// enthalpy, the forcing and the neighbouring columns' advection) gathered
// into one value per level.

// Row shifts of a double across the lanes of each 16-lane row (DPP): lane
// i reads lane i - S (row_shr) or i + S (row_shl) of its row, 0 where that
// lane is outside the row.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int S>
__device__ __forceinline__ double from_below(double x) {  // lane i - S
    static_assert(S >= 1 && S <= 15, "row shift");
    return dpp_f64<0x110 + S>(x);
}
template <int S>
__device__ __forceinline__ double from_above(double x) {  // lane i + S
    static_assert(S >= 1 && S <= 15, "row shift");
    return dpp_f64<0x100 + S>(x);
}

constexpr int kLmCols = 64;  // columns per block of the layout transposes

// (ny nx, nz) -> (nz, ny nx) (to_lm) or back, kLmCols columns per block
// through LDS; both sides are read and written in contiguous runs.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_lm_layout(const T* __restrict__ src, T* __restrict__ dst, long ncol,
                                                        int nz, int to_lm) {
    __shared__ T s[kLmCols * 65];
    const long c0 = (long)blockIdx.x * kLmCols;
    const int nc = (int)min((long)kLmCols, ncol - c0);
    if (to_lm) {
        for (int q = threadIdx.x; q < nc * nz; q += blockDim.x) s[(q / nz) * 65 + q % nz] = src[c0 * nz + q];
        __syncthreads();
        for (int q = threadIdx.x; q < kLmCols * nz; q += blockDim.x) {
            const int k = q / kLmCols, c = q % kLmCols;
            if (c < nc) dst[(long)k * ncol + c0 + c] = s[c * 65 + k];
        }
    } else {
        for (int q = threadIdx.x; q < kLmCols * nz; q += blockDim.x) {
            const int k = q / kLmCols, c = q % kLmCols;
            if (c < nc) s[c * 65 + k] = src[(long)k * ncol + c0 + c];
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nc * nz; q += blockDim.x) dst[c0 * nz + q] = s[(q / nz) * 65 + q % nz];
    }
}

// Horizontal Lax-Friedrichs advection of a column split by its unknowns
// (horiz_adv with the flux derivatives as coefficients): the coefficient of
// the column's own enthalpy (horiz_adv's diagonal) and those of the
// neighbours' E_xp, E_xm, E_ym, E_yp.  Depth-constant velocity (A.u2d).
template <typename T>
struct LmHoriz {
    T self, xp, xm, ym, yp;
};
template <typename T>
__device__ __forceinline__ LmHoriz<T> lm_horiz(const EnthArgs<T>& A, int i, int j) {
    LmHoriz<T> c{T(0), T(0), T(0), T(0), T(0)};
    if (!(A.flags & 1)) return c;
    const Horiz<T> hz = face_fluxes(A, i, j, 0);
    const T sq = A.lf_sq, f = T(kRhoI) * A.dxi;
    auto dl = [&](T U) { return T(0.5) * (U + fabs(U) * sq); };
    auto dr = [&](T U) { return T(0.5) * (U - fabs(U) * sq); };
    c.self = f * ((dl(hz.Hu_r) - dr(hz.Hu_l)) + (dl(hz.Hv_t) - dr(hz.Hv_b)));
    c.xp = f * dr(hz.Hu_r);
    c.xm = -f * dl(hz.Hu_l);
    c.ym = f * dr(hz.Hv_t);
    c.yp = -f * dl(hz.Hv_b);
    return c;
}

// The E-independent part of level k's row on the (nz, ny nx) planes:
// -rho H_prev E_prev / dt - f_E plus the neighbours' horizontal advection.
template <typename T>
__device__ __forceinline__ T lm_constant(const EnthArgs<T>& A, const LmHoriz<T>& hc, long ncol, long col, int k, int i,
                                         int j, T hp, T invdt) {
    const long o = (long)k * ncol;
    T g = -(T(kRhoI) * hp * A.E_prev[o + col] * invdt) - A.f_E[o + col];
    if (A.flags & 1) {
        const int nx = A.nx, im = max(i - 1, 0), ip = min(i + 1, A.ny - 1), jm = max(j - 1, 0),
                  jp = min(j + 1, nx - 1);
        const T* E = A.E + o;
        g += (hc.xp * E[(long)i * nx + jp] + hc.xm * E[(long)i * nx + jm]) +
             (hc.ym * E[(long)im * nx + j] + hc.yp * E[(long)ip * nx + j]);
    }
    return g;
}

// One stage of parallel cyclic reduction over the P lanes of a column
// group, one row a z_{q-1} + b z_q + c z_{q+1} = d per lane: a = 0 on the
// group's first S lanes and c = 0 on its last S lanes (true of the initial
// rows at S = 1 and kept by every stage), so values read across the group's
// edge are multiplied by zero.
template <typename T, int S>
__device__ __forceinline__ void pcr_stage(T& a, T& b, T& c, T& d) {
    const T rb = fast_rcp(b);
    const T an = a * rb, cn = c * rb, dn = d * rb;
    const T am = from_below<S>(an), cm = from_below<S>(cn), dm = from_below<S>(dn);
    const T ap = from_above<S>(an), cp = from_above<S>(cn), dp = from_above<S>(dn);
    const T na = -a * am, nc = -c * cp;
    b = b - a * cm - c * ap;
    d = d - a * dm - c * dp;
    a = na;
    c = nc;
}

// Column Newton smoother on the level-major planes (A.E etc. in (nz, ny nx)
// layout): W = E + omega_relax (E_col - E), and with rmax the max |r| of the
// starting state (first Newton iteration's rows) folded into the slots as
// k_enth_column_smooth does.  P lanes per column (4 or 8, consecutive lanes
// of one 16-lane row), MB >= 3 levels per lane, P MB >= nz.  Left at the
// compiler's register allocation (143 VGPRs, 3 waves per SIMD, no spills):
// held to 4 or 5 waves per SIMD the spills cost more (at 512 x 512 x 41 a
// sweep measured 425 us unconstrained, 503 us at 4 waves, 807 us at 5).
template <typename T, int P, int MB>
__global__ __launch_bounds__(256) void k_enth_column_lm(EnthArgs<T> A, T* __restrict__ W, T omega_relax,
                                                        int n_newton, T relaxation,
                                                        unsigned long long* __restrict__ rmax) {
    static_assert(MB >= 3, "the partitioned solve needs three levels per lane");
    static_assert(P == 4 || P == 8 || P == 16, "4, 8 or 16 lanes per column (within a 16-lane DPP row)");
    constexpr int kCols = 256 / P;  // columns per block
    if (skipped(A.skip)) return;
    const int nz = A.nz, nx = A.nx;
    const long ncol = (long)A.ny * nx;
    const long nvb = (ncol + kCols - 1) / kCols;
    const long col = xcd_block(blockIdx.x, nvb) * kCols + (threadIdx.x / P);
    const bool valid = col < ncol;  // the same for the P lanes of a column
    const long cc = valid ? col : ncol - 1;
    const int q = threadIdx.x % P, k0 = q * MB;
    const int i = (int)(cc / nx), j = (int)(cc % nx);
    const T h = A.H[cc], Es = A.E_surface[cc];
    const T dsig = T(1) / T(nz - 1);
    T El[MB];
#pragma unroll
    for (int m = 0; m < MB; ++m) El[m] = (k0 + m < nz) ? A.E[(long)(k0 + m) * ncol + cc] : T(0);
    double r_first = 0.0;
    if (h < A.h_thin) {
#pragma unroll
        for (int m = 0; m < MB; ++m)
            if (k0 + m < nz) El[m] = fmin(Es, e_pmp(T(k0 + m) * dsig, h));
    } else {
        const ColConst<T> cst = col_const(A, h);
        const T hp = A.H_prev[cc], invdt = cst.invdt;
        const LmHoriz<T> hc = lm_horiz(A, i, j);
        const T B0 = T(kRhoI) * h * invdt + hc.self;
        T G[MB], om[MB];
#pragma unroll
        for (int m = 0; m < MB; ++m) {
            const int k = k0 + m;
            G[m] = (k < nz - 1) ? lm_constant(A, hc, ncol, cc, k, i, j, hp, invdt) : T(0);
            om[m] = (A.omega && k < nz) ? A.omega[(long)k * ncol + cc] : T(0);
        }
        // omega one level below and above the block (clamped as the column kernel's)
        const T om_lo = (A.omega && k0 < nz) ? A.omega[(long)max(k0 - 1, 0) * ncol + cc] : T(0);
        const T om_hi = A.omega ? A.omega[(long)min(k0 + MB, nz - 1) * ncol + cc] : T(0);
        for (int it = 0; it < n_newton; ++it) {
            // The per-level terms that do not change between iterations
            // (pressure-melting enthalpies, advection coefficients) are formed
            // again in each one: held across iterations they would take more
            // registers than the rows and the elimination together.  The empty
            // statements only hide these inputs' values from that hoisting.
            T hh = h;
            asm volatile("" : "+v"(hh));
#pragma unroll
            for (int m = 0; m < MB; ++m) asm volatile("" : "+v"(om[m]));
            const T Eb = from_below<1>(El[MB - 1]);  // level k0 - 1
            const T Ea = from_above<1>(El[0]);       // level k0 + MB
            // interface diffusivities above the block's levels; the one below
            // level k0 is the lane below's last
            T Ku[MB];
#pragma unroll
            for (int m = 0; m < MB; ++m) {
                const int k = k0 + m;
                Ku[m] = T(0);
                if (k < nz - 1) {
                    const T Ep_k = e_pmp(T(k) * dsig, hh), Ep_p = e_pmp(T(k + 1) * dsig, hh);
                    Ku[m] = diffusivity(T(0.5) * (El[m] + (m < MB - 1 ? El[m + 1] : Ea)), T(0.5) * (Ep_k + Ep_p));
                }
            }
            const T Kl0 = from_below<1>(Ku[MB - 1]);
            // rows and forward elimination: row m becomes
            // Ac[m] x_0 + x_m + Cc[m] x_{m+1} = Dc[m] (row 0: Ac[0] multiplies
            // the lane below's last unknown)
            T Ac[MB], Cc[MB], Dc[MB];
#pragma unroll
            for (int m = 0; m < MB; ++m) {
                const int k = k0 + m;
                const T Ekm1 = m > 0 ? El[m - 1] : Eb;
                const T Ekp1 = m < MB - 1 ? El[m + 1] : Ea;
                T a = T(0), b = T(1), c = T(0), d = T(0);
                if (k < nz - 1) {
                    Row<T> w{B0 * El[m] + G[m], T(0), B0, T(0)};
                    column_terms(A, cst, k, hh, El[m], Ekm1, Ekp1, m > 0 ? om[m - 1] : om_lo, om[m],
                                 m < MB - 1 ? om[m + 1] : om_hi, Ku[m], m > 0 ? Ku[m - 1] : Kl0, w);
                    a = k == 0 ? T(0) : w.a;
                    b = w.b;
                    c = w.c;
                    d = -w.r;
                    if (it == 0) r_first = nan_max(r_first, fabs((double)d));
                } else if (k == nz - 1) {
                    d = -(El[m] - Es);  // Dirichlet surface row
                }
                if (m < 2) {
                    const T rb = fast_rcp(b);
                    Ac[m] = a * rb;
                    Cc[m] = c * rb;
                    Dc[m] = d * rb;
                } else {
                    const T rb = fast_rcp(b - a * Cc[m - 1]);
                    Ac[m] = -a * Ac[m - 1] * rb;
                    Cc[m] = c * rb;
                    Dc[m] = (d - a * Dc[m - 1]) * rb;
                }
            }
            // backward elimination: rows 1 .. MB-2 in terms of x_0 and x_{MB-1}
#pragma unroll
            for (int m = MB - 3; m >= 1; --m) {
                Dc[m] = Dc[m] - Cc[m] * Dc[m + 1];
                Ac[m] = Ac[m] - Cc[m] * Ac[m + 1];
                Cc[m] = -Cc[m] * Cc[m + 1];
            }
            // the block's first row against the lane below's last unknown and
            // its own last one: fa xl_{q-1} + x_0 + fc xl_q = fd; its last row:
            // la x_0 + xl_q + lc x0_{q+1} = ld
            const T s0 = fast_rcp(T(1) - Cc[0] * Ac[1]);
            const T fa = Ac[0] * s0, fc = -Cc[0] * Cc[1] * s0, fd = (Dc[0] - Cc[0] * Dc[1]) * s0;
            const T la = Ac[MB - 1], lc = Cc[MB - 1], ld = Dc[MB - 1];
            // x_0 eliminated through the first rows: one row per lane in the
            // last unknowns z_q = xl_q, solved by cyclic reduction over the group
            const T fa_n = from_above<1>(fa), fc_n = from_above<1>(fc), fd_n = from_above<1>(fd);
            T za = -la * fa, zb = T(1) - la * fc - lc * fa_n, zc = -lc * fc_n, zd = ld - la * fd - lc * fd_n;
            pcr_stage<T, 1>(za, zb, zc, zd);
            pcr_stage<T, 2>(za, zb, zc, zd);
            if (P >= 8) pcr_stage<T, 4>(za, zb, zc, zd);
            if (P == 16) pcr_stage<T, 8>(za, zb, zc, zd);
            const T xl = zd * fast_rcp(zb);
            const T x0 = fd - fa * from_below<1>(xl) - fc * xl;
            Dc[0] = x0;
            Dc[MB - 1] = xl;
#pragma unroll
            for (int m = 1; m < MB - 1; ++m) Dc[m] = Dc[m] - Ac[m] * x0 - Cc[m] * xl;
#pragma unroll
            for (int m = 0; m < MB; ++m)
                if (k0 + m < nz) El[m] += relaxation * Dc[m];
        }
    }
    if (valid) {
#pragma unroll
        for (int m = 0; m < MB; ++m) {
            const int k = k0 + m;
            if (k < nz) {
                const long o = (long)k * ncol + cc;
                const T E0 = A.E[o];
                W[o] = E0 + omega_relax * (El[m] - E0);
            }
        }
    }
    if (rmax && n_newton > 0) {  // every lane of the wavefront is here
        double mx = valid ? r_first : 0.0;
        for (int off = 32; off > 0; off >>= 1) mx = nan_max(mx, __shfl_down(mx, off, 64));
        if ((threadIdx.x & 63) == 0)
            atomicMax(rmax + (((long)blockIdx.x * 4 + (threadIdx.x >> 6)) & (kEnthRmaxSlots - 1)),
                      (unsigned long long)__double_as_longlong(mx));
    }
}

// One column's rows on the level-major planes, streamed up the levels:
// calls f(k, w) with level k's row (residual with forcing in w.r) for
// k = 0 .. nz - 2 of a thick column.
template <typename T, typename F>
__device__ __forceinline__ void lm_stream_rows(const EnthArgs<T>& A, long ncol, long col, F&& f) {
    const int nz = A.nz, nx = A.nx, i = (int)(col / nx), j = (int)(col % nx);
    const T h = A.H[col], hp = A.H_prev[col];
    const ColConst<T> cst = col_const(A, h);
    const T invdt = cst.invdt, dsig = cst.dsig;
    const LmHoriz<T> hc = lm_horiz(A, i, j);
    const T B0 = T(kRhoI) * h * invdt + hc.self;
    const T* E = A.E;
    const T* om = A.omega;
    T Em = E[col], Ek = E[col], Ep = E[ncol + col];
    T om_m = om ? om[col] : T(0), om_k = om_m, om_p = om ? om[(long)min(1, nz - 1) * ncol + col] : T(0);
    T Kl = T(0);
    for (int k = 0; k < nz - 1; ++k) {
        const T Ep_k = e_pmp(T(k) * dsig, h), Ep_p = e_pmp(T(k + 1) * dsig, h);
        const T Ku = diffusivity(T(0.5) * (Ek + Ep), T(0.5) * (Ep_k + Ep_p));
        Row<T> w{B0 * Ek + lm_constant(A, hc, ncol, col, k, i, j, hp, invdt), T(0), B0, T(0)};
        column_terms(A, cst, k, h, Ek, Em, Ep, om_m, om_k, om_p, Ku, Kl, w);
        f(k, Ek, w);
        Kl = Ku;
        Em = Ek;
        Ek = Ep;
        Ep = (k + 2 < nz) ? E[(long)(k + 2) * ncol + col] : T(0);
        om_m = om_k;
        om_k = om_p;
        om_p = om ? om[(long)min(k + 2, nz - 1) * ncol + col] : T(0);
    }
}

// Pointwise Jacobi layer smoother on the level-major planes:
// E_out = E + omega_relax (-r / J_kk), E_out = E on surface rows and thin columns.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_layer_lm(EnthArgs<T> A, T* __restrict__ E_out, T omega_relax) {
    if (skipped(A.skip)) return;
    const long ncol = (long)A.ny * A.nx, nvb = (ncol + 255) / 256;
    const long col = xcd_block(blockIdx.x, nvb) * 256 + threadIdx.x;
    if (col >= ncol) return;
    const int nz = A.nz;
    if (A.H[col] < A.h_thin) {
        for (int k = 0; k < nz; ++k) E_out[(long)k * ncol + col] = A.E[(long)k * ncol + col];
        return;
    }
    lm_stream_rows(A, ncol, col, [&](int k, T Ek, const Row<T>& w) {
        const T out = (fabs(w.b) > T(1e-30)) ? -w.r / w.b : T(0);
        E_out[(long)k * ncol + col] = Ek + omega_relax * out;
    });
    E_out[(long)(nz - 1) * ncol + col] = A.E[(long)(nz - 1) * ncol + col];
}

// Residual (with forcing) on the level-major planes: per-block max |r| into
// partials[blockIdx.x] (grid-stride over columns) and, when r_out is given,
// r written in the (ny, nx, nz) layout (zero on surface rows and thin columns).
template <typename T>
__global__ __launch_bounds__(256) void k_enth_residual_lm(EnthArgs<T> A, T* __restrict__ r_out,
                                                          double* __restrict__ partials) {
    __shared__ double smax[4];
    if (skipped(A.skip)) return;
    const int nz = A.nz;
    const long ncol = (long)A.ny * A.nx;
    double mx = 0.0;
    for (long col = (long)blockIdx.x * 256 + threadIdx.x; col < ncol; col += (long)gridDim.x * 256) {
        const long base = col * nz;
        if (A.H[col] < A.h_thin) {
            if (r_out)
                for (int k = 0; k < nz; ++k) r_out[base + k] = T(0);
            continue;
        }
        lm_stream_rows(A, ncol, col, [&](int k, T, const Row<T>& w) {
            if (r_out) r_out[base + k] = w.r;
            mx = nan_max(mx, fabs((double)w.r));
        });
        if (r_out) r_out[base + nz - 1] = T(0);
    }
    for (int off = 32; off > 0; off >>= 1) mx = nan_max(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = nan_max(nan_max(smax[0], smax[1]), nan_max(smax[2], smax[3]));
}

// ------------------------------------------------------------------ launchers
static inline unsigned col_blocks(int ny, int nx) {
    return (unsigned)(((long)ny * nx + kColsPerBlock - 1) / kColsPerBlock);
}
static inline unsigned flat_blocks(long n) {
    long b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

template <typename T>
hipError_t launch_enth_residual(const EnthArgs<T>& A, T* r_out, int use_forcing, double* partials, double* max_out,
                                EnthStatus* status, int status_mode, double rtol, double atol, hipStream_t st) {
    unsigned g = col_blocks(A.ny, A.nx);
    if (g > (unsigned)kEnthResBlocks) g = kEnthResBlocks;
    {
        TimedLaunch tl(kTimeEnthResidual, (long)A.ny * A.nx * A.nz, st);
        k_enth_residual<T><<<g, 256, 0, st>>>(A, r_out, use_forcing, partials);
    }
    if (max_out || status) k_enth_reduce<<<1, 256, 0, st>>>(partials, (int)g, max_out, status, status_mode, rtol, atol);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_column_smooth(const EnthArgs<T>& A, T* delta, T* E_out, T omega_relax, int n_newton,
                                     T relaxation, hipStream_t st, int colour, T* E_mc, unsigned long long* rmax,
                                     T* r_out) {
    if (colour >= 0 && (!E_mc || delta || E_out)) return hipErrorInvalidValue;
    if (colour >= 0 && rmax) return hipErrorInvalidValue;
    if (colour == 1 && A.E != E_mc) return hipErrorInvalidValue;
    TimedLaunch tl(kTimeEnthColumn, (long)A.ny * A.nx * A.nz, st);
    // one wavefront per column, or per column pair of a row in a colour phase
    const unsigned blocks = colour < 0 ? col_blocks(A.ny, A.nx) : col_blocks(A.ny, (A.nx + 1) / 2);
    k_enth_column_smooth<T><<<blocks, 256, 0, st>>>(A, delta, E_out, omega_relax, n_newton, relaxation, colour, E_mc,
                                                    rmax, rmax ? r_out : nullptr);
    return hipGetLastError();
}

hipError_t launch_enth_reduce_slots(unsigned long long* slots, EnthStatus* status, int mode, double rtol, double atol,
                                    hipStream_t st) {
    k_enth_reduce_slots<<<1, 256, 0, st>>>(slots, status, mode, rtol, atol);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_layer_smooth(const EnthArgs<T>& A, T* delta, T* E_out, T omega_relax, hipStream_t st) {
    TimedLaunch tl(kTimeEnthLayer, (long)A.ny * A.nx * A.nz, st);
    k_enth_layer_smooth<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, delta, E_out, omega_relax);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_apply(const EnthArgs<T>& A, T* E, const T* delta, T omega_relax, hipStream_t st) {
    long n = (long)A.ny * A.nx * A.nz;
    k_enth_apply<T><<<flat_blocks(n), 256, 0, st>>>(A.skip, E, delta, omega_relax, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_omega(const EnthArgs<T>& A, const T* dhdt, const T* bmb, T out_div, T* omega, hipStream_t st) {
    k_enth_omega<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, dhdt, bmb, out_div, omega);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_broadcast(const T* src, T* dst, long n2d, int nz, T div, hipStream_t st) {
    long n = n2d * nz;
    k_enth_broadcast<T><<<flat_blocks(n), 256, 0, st>>>(src, dst, n2d, n, div);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_tendency(const T* a, const T* b, T div, T* out, long n, hipStream_t st) {
    k_enth_tendency<T><<<flat_blocks(n), 256, 0, st>>>(a, b, div, out, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_rhs(const EnthArgs<T>& A, const T* beta, T m, T u_reg, int frictional, T* Q_fh, T* f_E,
                           hipStream_t st) {
    k_enth_rhs<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, beta, m, u_reg, frictional, Q_fh, f_E);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_strain_heating(const T* u, const T* v, const T* B, int ny, int nx, int nz, T dxi, T n_glen,
                                     T eps_reg, T rho_g, T* phi, hipStream_t st) {
    const T glen_p = (T(1) - n_glen) / (T(2) * n_glen);
    k_enth_strain_heating<T><<<col_blocks(ny, nx), 256, 0, st>>>(u, v, B, ny, nx, nz, dxi, glen_p, eps_reg, rho_g, phi);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_arrhenius(const EnthArgs<T>& A, T n_glen, T scale, T* B, hipStream_t st) {
    k_enth_arrhenius<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, n_glen, scale, B);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_diagnostics(const EnthArgs<T>& A, T* temperature, T* water, hipStream_t st) {
    k_enth_diagnostics<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, temperature, water);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_from_temperature(const EnthArgs<T>& A, const T* temperature, T* E, hipStream_t st) {
    k_enth_from_temperature<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, temperature, E);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_surface_from_temperature(const T* Ts, T* Es, long n, hipStream_t st) {
    k_enth_surface_from_temperature<T><<<flat_blocks(n), 256, 0, st>>>(Ts, Es, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_enth_adj_factor(const EnthArgs<T>& A, T* rows, hipStream_t st) {
    k_enth_adj_factor<T><<<flat_blocks((long)A.ny * A.nx), 256, 0, st>>>(A, rows);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_rhs(const EnthArgs<T>& A, const T* gE, const T* gB, T n_glen, T scale, T* g,
                               hipStream_t st) {
    k_enth_adj_rhs<T><<<flat_blocks((long)A.ny * A.nx), 256, 0, st>>>(A, gE, gB, n_glen, scale, g);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_layout(int ny, int nx, int nz, const T* src, T* dst, bool to_cb, hipStream_t st) {
    const unsigned nb = (unsigned)(((long)ny * nx + kStageCols - 1) / kStageCols);
    k_enth_adj_layout<T><<<nb, kStageCols, kStageCols * nz * sizeof(T), st>>>(ny, nx, nz, src, dst, to_cb ? 1 : 0);
    return hipGetLastError();
}

template <typename T, bool SOLVE>
static void adj_pass(const int* skip, const EnthArgs<T>& A, bool flat, const T* rows, const T* g, T* lam,
                     T omega_relax, int color, unsigned long long* acc, EnthStatus* status, int close_mode,
                     double rtol, double atol, hipStream_t st) {
    const unsigned nb = (unsigned)(((long)A.ny * ((A.nx + 1) / 2) + 255) / 256);
    const int ny = A.ny, nx = A.nx, nz = A.nz, f = flat ? 1 : 0;
#define GLIDE_ADJ_PASS(NZB)                                                                                        \
    k_enth_adj_pass<T, NZB, SOLVE><<<nb, 256, 0, st>>>(skip, ny, nx, nz, f, rows, g, lam, omega_relax, color, acc, \
                                                        status, close_mode, rtol, atol)
    if (nz <= 16) GLIDE_ADJ_PASS(16);
    else if (nz <= 32) GLIDE_ADJ_PASS(32);
    else if (nz <= 48) GLIDE_ADJ_PASS(48);
    else GLIDE_ADJ_PASS(64);
#undef GLIDE_ADJ_PASS
}
// The passes address a checkerboard plane with 32-bit byte offsets.
template <typename T>
static bool adj_shape_ok(const EnthArgs<T>& A) {
    return A.nz <= 64 && 2L * A.nz * A.ny * ((A.nx + 1) / 2) * (long)sizeof(T) < (1L << 32);
}
template <typename T>
hipError_t launch_enth_adj_sweep_rb(const int* skip, const EnthArgs<T>& A, bool flat, const T* rows, const T* g,
                                    T* lam, T omega_relax, unsigned long long* acc, EnthStatus* status, double rtol,
                                    double atol, hipStream_t st) {
    if (!adj_shape_ok(A)) return hipErrorInvalidValue;
    const long half = (long)A.ny * A.nx * A.nz / 2;
    {
        TimedLaunch tl(kTimeEnthColumn, half, st);
        adj_pass<T, true>(skip, A, flat, rows, g, lam, omega_relax, 0, acc, status, -1, rtol, atol, st);
    }
    {
        TimedLaunch tl(kTimeEnthColumn, half, st);
        adj_pass<T, true>(skip, A, flat, rows, g, lam, omega_relax, 1, acc, status, 1, rtol, atol, st);
    }
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_residual(const int* skip, const EnthArgs<T>& A, bool flat, const T* rows, const T* g,
                                    const T* lam, unsigned long long* acc, EnthStatus* status, int status_mode,
                                    double rtol, double atol, hipStream_t st) {
    if (!adj_shape_ok(A)) return hipErrorInvalidValue;
    T* l = const_cast<T*>(lam);  // the residual pass does not write lam
    adj_pass<T, false>(skip, A, flat, rows, g, l, T(1), 0, acc, status, -1, rtol, atol, st);
    adj_pass<T, false>(skip, A, flat, rows, g, l, T(1), 1, acc, status, status_mode, rtol, atol, st);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_grads(const EnthArgs<T>& A, const T* lam, T* g_E_prev, T* g_phi, T* g_Q, T* g_Es,
                                 hipStream_t st) {
    k_enth_adj_grads<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, lam, g_E_prev, g_phi, g_Q, g_Es);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_thermal_vjp(const EnthArgs<T>& A, const ThermalVjpArgs<T>& V, hipStream_t st) {
    if (A.nz < 2 || A.nz > 64 || !A.u2d || !A.v2d) return hipErrorInvalidValue;
    k_thermal_vjp_columns<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, V);
    const long n = (long)(A.ny + 1) * (A.nx + 1);
    const long nb = (n + 255) / 256;
    const unsigned g = (unsigned)(nb < 65535 ? nb : 65535);
    k_thermal_vjp_gather<T><<<g, 256, 0, st>>>(A, V, V.B_in ? 1 : 0, V.beta ? 1 : 0);
    return hipGetLastError();
}

// Level-major sweep launchers.
template <typename T>
hipError_t launch_enth_lm_layout(const T* src, T* dst, long ncol, int nz, bool to_lm, hipStream_t st) {
    if (nz > 64) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((ncol + kLmCols - 1) / kLmCols);
    k_enth_lm_layout<T><<<nb, 256, 0, st>>>(src, dst, ncol, nz, to_lm ? 1 : 0);
    return hipGetLastError();
}
bool enth_lm_supported(int nz) { return nz >= 9 && nz <= 64; }
template <typename T>
hipError_t launch_enth_column_lm(const EnthArgs<T>& A, T* W, T omega_relax, int n_newton, T relaxation,
                                 unsigned long long* rmax, hipStream_t st) {
    if (!enth_lm_supported(A.nz) || !A.u2d) return hipErrorInvalidValue;
    const long ncol = (long)A.ny * A.nx;
    // three or four levels per lane: 4 lanes per column up to 16 levels, 8 up
    // to 32, 16 above (measured at 41 levels: 16 lanes 6% faster than 8)
    const int P = A.nz <= 16 ? 4 : A.nz <= 32 ? 8 : 16, mb = (A.nz + P - 1) / P;
    const unsigned nb = (unsigned)((ncol + 256 / P - 1) / (256 / P));
    TimedLaunch tl(kTimeEnthColumn, ncol * A.nz, st);
#define GLIDE_LM_COLUMN(PP, MB)                                                                                   \
    if (P == PP && mb == MB) {                                                                                    \
        k_enth_column_lm<T, PP, MB><<<nb, 256, 0, st>>>(A, W, omega_relax, n_newton, relaxation, rmax);           \
        return hipGetLastError();                                                                                 \
    }
    GLIDE_LM_COLUMN(4, 3)
    GLIDE_LM_COLUMN(4, 4)
    GLIDE_LM_COLUMN(8, 3)
    GLIDE_LM_COLUMN(8, 4)
    GLIDE_LM_COLUMN(16, 3)
    GLIDE_LM_COLUMN(16, 4)
#undef GLIDE_LM_COLUMN
    return hipErrorInvalidValue;
}
template <typename T>
hipError_t launch_enth_layer_lm(const EnthArgs<T>& A, T* E_out, T omega_relax, hipStream_t st) {
    if (!enth_lm_supported(A.nz) || !A.u2d) return hipErrorInvalidValue;
    const long ncol = (long)A.ny * A.nx;
    TimedLaunch tl(kTimeEnthLayer, ncol * A.nz, st);
    k_enth_layer_lm<T><<<(unsigned)((ncol + 255) / 256), 256, 0, st>>>(A, E_out, omega_relax);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_residual_lm(const EnthArgs<T>& A, T* r_out, double* partials, EnthStatus* status,
                                   int status_mode, double rtol, double atol, hipStream_t st) {
    if (!enth_lm_supported(A.nz) || !A.u2d) return hipErrorInvalidValue;
    const long ncol = (long)A.ny * A.nx;
    long g = (ncol + 255) / 256;
    if (g > kEnthResBlocks) g = kEnthResBlocks;
    {
        TimedLaunch tl(kTimeEnthResidual, ncol * A.nz, st);
        k_enth_residual_lm<T><<<(unsigned)g, 256, 0, st>>>(A, r_out, partials);
    }
    k_enth_reduce<<<1, 256, 0, st>>>(partials, (int)g, nullptr, status, status_mode, rtol, atol);
    return hipGetLastError();
}

#define INST(T)                                                                                                      \
    template hipError_t launch_enth_residual<T>(const EnthArgs<T>&, T*, int, double*, double*, EnthStatus*, int,     \
                                                double, double, hipStream_t);                                        \
    template hipError_t launch_enth_column_smooth<T>(const EnthArgs<T>&, T*, T*, T, int, T, hipStream_t, int, T*,   \
                                                     unsigned long long*, T*);                                      \
    template hipError_t launch_enth_layer_smooth<T>(const EnthArgs<T>&, T*, T*, T, hipStream_t);                     \
    template hipError_t launch_enth_apply<T>(const EnthArgs<T>&, T*, const T*, T, hipStream_t);                      \
    template hipError_t launch_enth_omega<T>(const EnthArgs<T>&, const T*, const T*, T, T*, hipStream_t);            \
    template hipError_t launch_enth_broadcast<T>(const T*, T*, long, int, T, hipStream_t);                           \
    template hipError_t launch_enth_tendency<T>(const T*, const T*, T, T*, long, hipStream_t);                       \
    template hipError_t launch_enth_rhs<T>(const EnthArgs<T>&, const T*, T, T, int, T*, T*, hipStream_t);            \
    template hipError_t launch_enth_arrhenius<T>(const EnthArgs<T>&, T, T, T*, hipStream_t);                         \
    template hipError_t launch_enth_strain_heating<T>(const T*, const T*, const T*, int, int, int, T, T, T, T, T*,   \
                                                      hipStream_t);                                                  \
    template hipError_t launch_enth_diagnostics<T>(const EnthArgs<T>&, T*, T*, hipStream_t);                         \
    template hipError_t launch_enth_from_temperature<T>(const EnthArgs<T>&, const T*, T*, hipStream_t);              \
    template hipError_t launch_enth_surface_from_temperature<T>(const T*, T*, long, hipStream_t);              \
    template hipError_t launch_enth_adj_factor<T>(const EnthArgs<T>&, T*, hipStream_t);                              \
    template hipError_t launch_enth_adj_rhs<T>(const EnthArgs<T>&, const T*, const T*, T, T, T*, hipStream_t);       \
    template hipError_t launch_enth_adj_layout<T>(int, int, int, const T*, T*, bool, hipStream_t);                   \
    template hipError_t launch_enth_adj_sweep_rb<T>(const int*, const EnthArgs<T>&, bool, const T*, const T*, T*, T, \
                                                    unsigned long long*, EnthStatus*, double, double, hipStream_t);  \
    template hipError_t launch_enth_adj_residual<T>(const int*, const EnthArgs<T>&, bool, const T*, const T*,        \
                                                    const T*, unsigned long long*, EnthStatus*, int, double, double, \
                                                    hipStream_t);                                                    \
    template hipError_t launch_enth_adj_grads<T>(const EnthArgs<T>&, const T*, T*, T*, T*, T*, hipStream_t);      \
    template hipError_t launch_thermal_vjp<T>(const EnthArgs<T>&, const ThermalVjpArgs<T>&, hipStream_t);           \
    template hipError_t launch_enth_lm_layout<T>(const T*, T*, long, int, bool, hipStream_t);                        \
    template hipError_t launch_enth_column_lm<T>(const EnthArgs<T>&, T*, T, int, T, unsigned long long*,            \
                                                 hipStream_t);                                                       \
    template hipError_t launch_enth_layer_lm<T>(const EnthArgs<T>&, T*, T, hipStream_t);                             \
    template hipError_t launch_enth_residual_lm<T>(const EnthArgs<T>&, T*, double*, EnthStatus*, int, double, double, \
                                                   hipStream_t);
INST(double)
#undef INST

}  // namespace glide
