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
// per-thread arrays of length nz are needed.
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
template <typename T>
__device__ __forceinline__ T e_pmp(T sigma, T H) {
    return T(kCI) * (t_pmp(sigma, H) - T(kTRef)) / T(kEScale);
}
// glide/cuda/enthalpy.cu:48-98
template <typename T>
__device__ __forceinline__ T diffusivity(T E, T Ep) {
    T z = (E - Ep) * T(kEScale) * T(0.01);
    z = fmin(fmax(z, T(-20)), T(20));
    T s = T(1) / (T(1) + exp(-z));
    return T(kKCold) * (T(1) - s + T(kKTempFactor) * s);
}
template <typename T>
__device__ __forceinline__ T water_content(T E, T Ep) {
    T x = (E - Ep) * T(kEScale);
    T ax = fmin(fmax(T(kDrainSharpness) * x, T(-20)), T(20));
    T sp = (ax > T(10)) ? x : log(T(1) + exp(ax)) / T(kDrainSharpness);
    return sp / T(kLHeat);
}
template <typename T>
__device__ __forceinline__ T dwater(T E, T Ep) {
    T z = fmin(fmax(T(kDrainSharpness) * (E - Ep) * T(kEScale), T(-20)), T(20));
    return (T(1) / (T(1) + exp(-z))) / T(kLHeat);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

template <typename T>
__device__ __forceinline__ T lane_get(T x, int src) {
    return __shfl(x, src, 64);
}

__device__ __forceinline__ bool skipped(const int* skip) { return skip && *skip; }

// Horizontal advection pieces of one level that do not depend on the
// level's own enthalpy: face mass fluxes (zero on the domain edge) and the
// neighbouring columns' enthalpy (glide/cuda/enthalpy.cu:671-763).
template <typename T>
struct Horiz {
    T Hu_l, Hu_r, Hv_t, Hv_b;
    T E_xm, E_xp, E_ym, E_yp;
};

template <typename T>
__device__ __forceinline__ Horiz<T> load_horiz(const EnthArgs<T>& A, int i, int j, int k) {
    const int ny = A.ny, nx = A.nx, nz = A.nz;
    Horiz<T> h;
    const T Hh = A.H[(long)i * nx + j];
    const T* u = A.u2d ? A.u2d : A.u3d + (long)k * ny * (nx + 1);
    const T* v = A.u2d ? A.v2d : A.v3d + (long)k * (ny + 1) * nx;
    const T div = A.u2d ? A.vel_div : T(1);
    h.Hu_l = h.Hu_r = h.Hv_t = h.Hv_b = T(0);
    if (j > 0) {
        T ul = u[(long)i * (nx + 1) + j] / div, Hm = A.H[(long)i * nx + j - 1];
        h.Hu_l = T(0.5) * (Hm + Hh) * ul - T(0.5) * sqrt(ul * ul + T(kMassFluxReg)) * (Hh - Hm);
    }
    if (j < nx - 1) {
        T ur = u[(long)i * (nx + 1) + j + 1] / div, Hp = A.H[(long)i * nx + j + 1];
        h.Hu_r = T(0.5) * (Hh + Hp) * ur - T(0.5) * sqrt(ur * ur + T(kMassFluxReg)) * (Hp - Hh);
    }
    if (i > 0) {
        T vt = v[(long)i * nx + j] / div, Hm = A.H[(long)(i - 1) * nx + j];
        h.Hv_t = T(0.5) * (Hm + Hh) * vt - T(0.5) * sqrt(vt * vt + T(kMassFluxReg)) * (Hm - Hh);
    }
    if (i < ny - 1) {
        T vb = v[(long)(i + 1) * nx + j] / div, Hp = A.H[(long)(i + 1) * nx + j];
        h.Hv_b = T(0.5) * (Hh + Hp) * vb - T(0.5) * sqrt(vb * vb + T(kMassFluxReg)) * (Hh - Hp);
    }
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

template <typename T>
__device__ __forceinline__ Row<T> level_row(const EnthArgs<T>& A, int k, T h, T hinv, T hp, T Ek, T Ekm1, T Ekp1,
                                            T Epk, T om_m, T om_k, T om_p, const Horiz<T>& hz) {
    const int nz = A.nz;
    const T dsig = T(1) / T(nz - 1);
    const T invdt = T(1) / A.dt;
    Row<T> w;
    w.r = T(kRhoI) * (h * Ek - hp * Epk) * invdt;
    w.b = T(kRhoI) * h * invdt;
    w.a = T(0);
    w.c = T(0);
    if (A.flags & 1) {
        T hr, hd;
        horiz_adv(hz, Ek, A.lf_sq, A.dxi, hr, hd);
        w.r += hr;
        w.b += hd;
    }
    const T Ep_k = e_pmp(T(k) * dsig, h);
    if (k == 0) {
        if (A.flags & 2) {
            T wh = T(0.5) * (om_k + om_p);
            T wp = fmax(wh, T(0)), wn = fmin(wh, T(0));
            T hdi = T(1) / (T(0.5) * dsig);
            w.r += T(kRhoI) * ((wp * Ek + wn * Ekp1) - om_k * Ek) * hdi;
            w.b += T(kRhoI) * (wp - om_k) * hdi;
            w.c += T(kRhoI) * wn * hdi;
        }
        T Kh = diffusivity(T(0.5) * (Ek + Ekp1), T(0.5) * (Ep_k + e_pmp(dsig, h)));
        T dh = T(0.5) * dsig;
        w.r += (-hinv * Kh * (Ekp1 - Ek) / dsig) / dh;
        T co = hinv * Kh / (dsig * dh);
        w.b += co;
        w.c -= co;
    } else {
        if (A.flags & 2) {
            T wu = T(0.5) * (om_k + om_p), wl = T(0.5) * (om_m + om_k);
            T up = fmax(wu, T(0)), un = fmin(wu, T(0)), lp = fmax(wl, T(0)), ln = fmin(wl, T(0));
            T di = T(1) / dsig;
            w.r += T(kRhoI) * ((up * Ek + un * Ekp1) - (lp * Ekm1 + ln * Ek)) * di;
            w.a -= T(kRhoI) * lp * di;
            w.c += T(kRhoI) * un * di;
            w.b += T(kRhoI) * (up - ln) * di;
        }
        T Ep_m = e_pmp(T(k - 1) * dsig, h), Ep_p = e_pmp(T(min(k + 1, nz - 1)) * dsig, h);
        T Ku = diffusivity(T(0.5) * (Ek + Ekp1), T(0.5) * (Ep_k + Ep_p));
        T Kl = diffusivity(T(0.5) * (Ek + Ekm1), T(0.5) * (Ep_k + Ep_m));
        T d2 = T(1) / (dsig * dsig);
        w.r += -hinv * d2 * (Ku * (Ekp1 - Ek) - Kl * (Ek - Ekm1));
        T dl = hinv * d2 * Kl, du = hinv * d2 * Ku;
        w.a -= dl;
        w.c -= du;
        w.b += dl + du;
    }
    if (A.flags & 4) {
        w.r += h * T(kRhoW) * T(kLHeat) * (A.drain_rate * water_content(Ek, Ep_k)) / T(kEScale);
        w.b += h * T(kRhoW) * T(kLHeat) * A.drain_rate * dwater(Ek, Ep_k);
    }
    return w;
}

// Parallel cyclic reduction over the wavefront: lane k holds row k,
// a x_{k-1} + b x_k + c x_{k+1} = d, with identity rows beyond the column;
// returns x_k.  Six doubling steps cover 64 rows.
template <typename T>
__device__ __forceinline__ T pcr_solve(T a, T b, T c, T d) {
    const int k = lane_id();
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const bool lo = k - s >= 0, hi = k + s < 64;
        const int km = lo ? k - s : k, kp = hi ? k + s : k;
        T am = lane_get(a, km), bm = lane_get(b, km), cm = lane_get(c, km), dm = lane_get(d, km);
        T ap = lane_get(a, kp), bp = lane_get(b, kp), cp = lane_get(c, kp), dp = lane_get(d, kp);
        T al = lo ? -a / bm : T(0);
        T ga = hi ? -c / bp : T(0);
        T na = al * am, nc = ga * cp;
        T nb = b + al * cm + ga * ap;
        T nd = d + al * dm + ga * dp;
        a = na;
        b = nb;
        c = nc;
        d = nd;
    }
    return d / b;
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
        T r = T(0);
        if (in && k < nz - 1 && !(h < A.h_thin)) {
            Horiz<T> hz{};
            if (A.flags & 1) hz = load_horiz(A, q.i, q.j, k);
            Row<T> w = level_row(A, k, h, T(1) / h, A.H_prev[col], Ek, Ekm1, Ekp1, A.E_prev[base + k], omm, omk,
                                 omp, hz);
            r = use_forcing ? w.r - A.f_E[base + k] : w.r;
        }
        if (in && r_out) r_out[base + k] = r;
        m = fmax(m, fabs((double)r));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if (k == 0) smax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
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
    for (int q = threadIdx.x; q < n; q += blockDim.x) m = fmax(m, part[q]);
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    if (out) *out = r;
    if (!status) return;
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
        const double rel = r0 > 0.0 ? r / r0 : 0.0;
        if (rel < rtol || r < atol) status->done = 1;
    }
}

// Column Newton smoother (glide/cuda/enthalpy.cu:932-1112): Newton on the
// column with horizontal neighbours frozen at E; writes delta = E_col - E
// and/or E_out = E + omega_relax * delta.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_column_smooth(EnthArgs<T> A, T* __restrict__ delta, T* __restrict__ E_out,
                                                            T omega_relax, int n_newton, T relaxation) {
    if (skipped(A.skip)) return;
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid) return;
    const int nz = A.nz, k = q.k;
    const bool in = k < nz;
    const long base = q.col * nz;
    const T h = A.H[q.col];
    const T E0 = in ? A.E[base + k] : T(0);
    const T Es = A.E_surface[q.col];
    T El = E0;
    if (h < A.h_thin) {
        El = fmin(Es, e_pmp(T(k) * (T(1) / T(nz - 1)), h));
    } else {
        const T hinv = T(1) / h, hp = A.H_prev[q.col];
        const bool interior = in && k < nz - 1;
        const T Epk = in ? A.E_prev[base + k] : T(0);
        const T fEk = in ? A.f_E[base + k] : T(0);
        const T omk = (in && A.omega) ? A.omega[base + k] : T(0);
        const T omm = lane_get(omk, max(k - 1, 0)), omp = lane_get(omk, min(k + 1, nz - 1));
        Horiz<T> hz{};
        if ((A.flags & 1) && interior) hz = load_horiz(A, q.i, q.j, k);
        for (int it = 0; it < n_newton; ++it) {
            const T Ekm1 = lane_get(El, max(k - 1, 0)), Ekp1 = lane_get(El, min(k + 1, 63));
            T a = T(0), b = T(1), c = T(0), d = T(0);
            if (interior) {
                Row<T> w = level_row(A, k, h, hinv, hp, El, Ekm1, Ekp1, Epk, omm, omk, omp, hz);
                a = k == 0 ? T(0) : w.a;
                b = w.b;
                c = w.c;
                d = -(w.r - fEk);
            } else if (k == nz - 1) {
                d = -(El - Es);  // Dirichlet surface row
            }
            const T dx = pcr_solve(a, b, c, d);
            if (in) El += relaxation * dx;
        }
    }
    if (!in) return;
    const T dE = El - E0;
    if (delta) delta[base + k] = dE;
    if (E_out) E_out[base + k] = E0 + omega_relax * dE;
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
    if (!in) return;
    T out = T(0);
    if (k < nz - 1 && !(h < A.h_thin)) {
        Horiz<T> hz{};
        if (A.flags & 1) hz = load_horiz(A, q.i, q.j, k);
        Row<T> w = level_row(A, k, h, T(1) / h, A.H_prev[q.col], Ek, Ekm1, Ekp1, A.E_prev[base + k], omm, omk, omp,
                             hz);
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

// Exact Jacobian row of an interior node of a thick column; Ekm1 / Ekp1 the
// column neighbours' enthalpy.
template <typename T>
__device__ __forceinline__ JRow<T> exact_row(const EnthArgs<T>& A, int i, int j, int k, T h, T Ek, T Ekm1, T Ekp1,
                                             T om_m, T om_k, T om_p) {
    const int nz = A.nz;
    const T dsig = T(1) / T(nz - 1);
    const T hinv = T(1) / h;
    JRow<T> w{T(0), T(kRhoI) * h / A.dt, T(0), T(0), T(0), T(0), T(0)};
    if (A.flags & 1) {
        const Horiz<T> hz = load_horiz(A, i, j, k);
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

// Rows (a, b, c, xm, xp, ym, yp) at the converged E into seven (ny, nx, nz)
// planes; identity rows on surface nodes and thin columns.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_adj_rows(EnthArgs<T> A, T* __restrict__ rows) {
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid) return;
    const int nz = A.nz, k = q.k;
    const bool in = k < nz;
    const long base = q.col * nz, n3 = ncol * nz;
    const T h = A.H[q.col];
    const T Ek = in ? A.E[base + k] : T(0);
    const T omk = (in && A.omega) ? A.omega[base + k] : T(0);
    const T Ekm1 = lane_get(Ek, max(k - 1, 0)), Ekp1 = lane_get(Ek, min(k + 1, 63));
    const T omm = lane_get(omk, max(k - 1, 0)), omp = lane_get(omk, min(k + 1, nz - 1));
    if (!in) return;
    JRow<T> w{T(0), T(1), T(0), T(0), T(0), T(0), T(0)};
    if (k < nz - 1 && !(h < A.h_thin)) w = exact_row(A, q.i, q.j, k, h, Ek, Ekm1, Ekp1, omm, omk, omp);
    const long n = base + k;
    rows[n] = w.a;
    rows[n3 + n] = w.b;
    rows[2 * n3 + n] = w.c;
    rows[3 * n3 + n] = w.xm;
    rows[4 * n3 + n] = w.xp;
    rows[5 * n3 + n] = w.ym;
    rows[6 * n3 + n] = w.yp;
}

// Horizontal part of (J^T lam) at node (i, j, k): the neighbouring rows'
// couplings to this node's enthalpy.
template <typename T>
__device__ __forceinline__ T jt_horizontal(const T* __restrict__ rows, const T* __restrict__ lam, int ny, int nx,
                                           int nz, long n3, int i, int j, int k) {
    T s = T(0);
    if (j > 0) { const long m = ((long)i * nx + j - 1) * nz + k; s += rows[4 * n3 + m] * lam[m]; }
    if (j < nx - 1) { const long m = ((long)i * nx + j + 1) * nz + k; s += rows[3 * n3 + m] * lam[m]; }
    if (i > 0) { const long m = ((long)(i - 1) * nx + j) * nz + k; s += rows[6 * n3 + m] * lam[m]; }
    if (i < ny - 1) { const long m = ((long)(i + 1) * nx + j) * nz + k; s += rows[5 * n3 + m] * lam[m]; }
    return s;
}

// Column pass on J^T lam = g: each column's transposed tridiagonal block
// solved by PCR against g minus the horizontal coupling of lam_in;
// lam_out = lam_in + omega_relax (lam_col - lam_in).
template <typename T>
__global__ __launch_bounds__(256) void k_enth_adj_column(const int* __restrict__ skip, int ny, int nx, int nz,
                                                         const T* __restrict__ rows, const T* __restrict__ g,
                                                         const T* __restrict__ lam_in, T* __restrict__ lam_out,
                                                         T omega_relax) {
    if (skipped(skip)) return;
    const long ncol = (long)ny * nx;
    const ColIdx q = col_index(wave_column(), nx, ncol);
    if (!q.valid) return;
    const int k = q.k;
    const bool in = k < nz;
    const long n3 = ncol * nz, n = q.col * nz + k;
    const T a = in ? rows[n] : T(0), b = in ? rows[n3 + n] : T(1), c = in ? rows[2 * n3 + n] : T(0);
    const T cm = lane_get(c, max(k - 1, 0)), ap = lane_get(a, min(k + 1, 63));
    const T aT = k > 0 ? cm : T(0);        // (J^T)_{k,k-1} = J_{k-1,k}
    const T cT = (k + 1 < nz) ? ap : T(0);  // (J^T)_{k,k+1} = J_{k+1,k}
    T rhs = T(0), l0 = T(0);
    if (in) {
        l0 = lam_in[n];
        rhs = g[n] - jt_horizontal(rows, lam_in, ny, nx, nz, n3, q.i, q.j, k);
    }
    const T lc = pcr_solve(in ? aT : T(0), b, in ? cT : T(0), rhs);
    if (in) lam_out[n] = l0 + omega_relax * (lc - l0);
}

// Column of colour `color` ((i + j) % 2 == color) with colour index w:
// rows alternate between the counts ce (even rows) and co (odd rows).
__device__ __forceinline__ ColIdx color_col_index(long w, int color, int ny, int nx) {
    const int ce = (nx + 1 - (color & 1)) / 2;  // columns of this colour in an even row
    const long p = w / nx, r = w % nx;          // a pair of rows holds nx of them
    const int i = (int)(r < ce ? 2 * p : 2 * p + 1);
    const int t = (int)(r < ce ? r : r - ce);
    const int j = 2 * t + ((i + color) & 1);
    ColIdx c;
    c.valid = i < ny && j < nx;
    c.i = c.valid ? i : 0;
    c.j = c.valid ? j : 0;
    c.col = (long)c.i * nx + c.j;
    c.k = lane_id();
    return c;
}

// Red-black column Gauss-Seidel pass on J^T lam = g, in place: the columns of
// one colour solve their transposed tridiagonal block (PCR) against g minus
// the horizontal coupling of the other colour's current multipliers, which
// this pass does not change; lam = lam + omega_relax (lam_col - lam).  The
// max |g - J^T lam| of the columns before their update goes to
// partials[blockIdx.x] (the residual the sweep's stopping test reads).
template <typename T>
__global__ __launch_bounds__(256) void k_enth_adj_column_rb(const int* __restrict__ skip, int ny, int nx, int nz,
                                                            const T* __restrict__ rows, const T* __restrict__ g,
                                                            T* __restrict__ lam, T omega_relax, int color,
                                                            double* __restrict__ partials) {
    __shared__ double smax[kColsPerBlock];
    if (skipped(skip)) return;
    const long ncol = (long)ny * nx, n3 = ncol * nz;
    const long nc = (long)((ny + 1) / 2) * nx;  // colour indices covered (pairs of rows)
    const int k = lane_id();
    const bool in = k < nz;
    double m = 0.0;
    for (long w = wave_column(); w < nc; w += (long)gridDim.x * kColsPerBlock) {
        const ColIdx q = color_col_index(w, color, ny, nx);
        if (!q.valid) continue;
        const long n = q.col * nz + k;
        const T a = in ? rows[n] : T(0), b = in ? rows[n3 + n] : T(1), c = in ? rows[2 * n3 + n] : T(0);
        const T cm = lane_get(c, max(k - 1, 0)), ap = lane_get(a, min(k + 1, 63));
        const T aT = (in && k > 0) ? cm : T(0);        // (J^T)_{k,k-1} = J_{k-1,k}
        const T cT = (in && k + 1 < nz) ? ap : T(0);   // (J^T)_{k,k+1} = J_{k+1,k}
        T rhs = T(0), l0 = T(0);
        if (in) {
            l0 = lam[n];
            rhs = g[n] - jt_horizontal(rows, lam, ny, nx, nz, n3, q.i, q.j, k);
        }
        const T lm = lane_get(l0, max(k - 1, 0)), lp = lane_get(l0, min(k + 1, 63));
        const T res = in ? rhs - (aT * lm + b * l0 + cT * lp) : T(0);
        m = fmax(m, fabs((double)res));
        const T lc = pcr_solve(aT, b, cT, rhs);
        if (in) lam[n] = l0 + omega_relax * (lc - l0);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if (k == 0) smax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
}

// (J^T lam)_n at one node.
template <typename T>
__device__ __forceinline__ T jt_apply(const T* __restrict__ rows, const T* __restrict__ lam, int ny, int nx, int nz,
                                      long n3, int i, int j, int k, long n) {
    T s = rows[n3 + n] * lam[n];
    if (k > 0) s += rows[2 * n3 + n - 1] * lam[n - 1];
    if (k < nz - 1) s += rows[n + 1] * lam[n + 1];
    return s + jt_horizontal(rows, lam, ny, nx, nz, n3, i, j, k);
}

// Layer pass: lam_out = lam_in + omega_relax (g - J^T lam_in) / J_nn.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_adj_layer(const int* __restrict__ skip, int ny, int nx, int nz,
                                                        const T* __restrict__ rows, const T* __restrict__ g,
                                                        const T* __restrict__ lam_in, T* __restrict__ lam_out,
                                                        T omega_relax) {
    if (skipped(skip)) return;
    const long n3 = (long)ny * nx * nz;
    for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < n3; n += (long)gridDim.x * blockDim.x) {
        const int k = (int)(n % nz);
        const long col = n / nz;
        const int i = (int)(col / nx), j = (int)(col % nx);
        const T r = g[n] - jt_apply(rows, lam_in, ny, nx, nz, n3, i, j, k, n);
        const T d = rows[n3 + n];
        lam_out[n] = lam_in[n] + (fabs(d) > T(1e-30) ? omega_relax * r / d : T(0));
    }
}

// Adjoint residual g - J^T lam and per-block max |.|.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_adj_residual(const int* __restrict__ skip, int ny, int nx, int nz,
                                                           const T* __restrict__ rows, const T* __restrict__ g,
                                                           const T* __restrict__ lam, T* __restrict__ r_out,
                                                           double* __restrict__ partials) {
    __shared__ double smax[4];
    if (skipped(skip)) return;
    const long n3 = (long)ny * nx * nz;
    double m = 0.0;
    for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < n3; n += (long)gridDim.x * blockDim.x) {
        const int k = (int)(n % nz);
        const long col = n / nz;
        const T r = g[n] - jt_apply(rows, lam, ny, nx, nz, n3, (int)(col / nx), (int)(col % nx), k, n);
        if (r_out) r_out[n] = r;
        m = fmax(m, fabs((double)r));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
}

// Right-hand side g = dL/dE + (dB/dE)^T dL/dB of the rate factor
// B = mean_k(A_k (1 + 181.25 w_k))^(-1/n) / scale.
template <typename T>
__global__ __launch_bounds__(256) void k_enth_adj_rhs(EnthArgs<T> A, const T* __restrict__ gE,
                                                      const T* __restrict__ gB, T n_glen, T scale,
                                                      T* __restrict__ g) {
    const long ncol = (long)A.ny * A.nx;
    const ColIdx q = col_index(wave_column(), A.nx, ncol);
    if (!q.valid) return;
    const int nz = A.nz, k = q.k;
    const bool in = k < nz;
    const long n = q.col * nz + k;
    T val = T(0), dA = T(0);
    if (in && gB) {
        const T h = A.H[q.col];
        const T Tp = t_pmp(T(k) * (T(1) / T(nz - 1)), h);
        const T Ek = A.E[n];
        const T Tu = Ek * T(kEScale) / T(kCI) + T(kTRef);
        const T Tk = fmin(Tu, Tp);
        const T Af = Tk < T(263.15) ? T(3.985e-13) : T(1.916e3);
        const T Qf = Tk < T(263.15) ? T(60e3) : T(139e3);
        const T Ts = fmax(Tk, T(1));
        const T Ak = Af * exp(-Qf / (T(8.314) * Ts));
        const T wraw = Ek * T(kEScale) - T(kCI) * (Tp - T(kTRef));
        const T wk = fmax(wraw, T(0)) / T(kLHeat);
        val = Ak * (T(1) + T(181.25) * wk);
        const T dT = (Tu < Tp && Tk > T(1)) ? T(kEScale) / T(kCI) : T(0);
        const T dw = wraw > T(0) ? T(kEScale) / T(kLHeat) : T(0);
        dA = Ak * Qf / (T(8.314) * Ts * Ts) * dT * (T(1) + T(181.25) * wk) + Ak * T(181.25) * dw;
    }
    T sum = val;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if (!in) return;
    T out = gE ? gE[n] : T(0);
    if (gB) {
        const T mean = sum / T(nz);
        const T dBdmean = (T(-1) / n_glen) * pow(mean, T(-1) / n_glen - T(1)) / scale;
        out += gB[q.col] * dBdmean * dA / T(nz);
    }
    g[n] = out;
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
                                     T relaxation, hipStream_t st) {
    TimedLaunch tl(kTimeEnthColumn, (long)A.ny * A.nx * A.nz, st);
    k_enth_column_smooth<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, delta, E_out, omega_relax, n_newton, relaxation);
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
hipError_t launch_enth_adj_rows(const EnthArgs<T>& A, T* rows, hipStream_t st) {
    k_enth_adj_rows<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, rows);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_rhs(const EnthArgs<T>& A, const T* gE, const T* gB, T n_glen, T scale, T* g,
                               hipStream_t st) {
    k_enth_adj_rhs<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, gE, gB, n_glen, scale, g);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_column(const int* skip, int ny, int nx, int nz, const T* rows, const T* g,
                                  const T* lam_in, T* lam_out, T omega_relax, hipStream_t st) {
    k_enth_adj_column<T><<<col_blocks(ny, nx), 256, 0, st>>>(skip, ny, nx, nz, rows, g, lam_in, lam_out,
                                                              omega_relax);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_sweep_rb(const int* skip, int ny, int nx, int nz, const T* rows, const T* g, T* lam,
                                    T omega_relax, double* partials, EnthStatus* status, double rtol, double atol,
                                    hipStream_t st) {
    const long nc = (long)((ny + 1) / 2) * nx;
    long nb = (nc + kColsPerBlock - 1) / kColsPerBlock;
    const unsigned g_ = (unsigned)(nb < kEnthResBlocks / 2 ? (nb > 0 ? nb : 1) : kEnthResBlocks / 2);
    {
        TimedLaunch tl(kTimeEnthColumn, (long)ny * nx * nz / 2, st);
        k_enth_adj_column_rb<T><<<g_, 256, 0, st>>>(skip, ny, nx, nz, rows, g, lam, omega_relax, 0, partials);
    }
    {
        TimedLaunch tl(kTimeEnthColumn, (long)ny * nx * nz / 2, st);
        k_enth_adj_column_rb<T><<<g_, 256, 0, st>>>(skip, ny, nx, nz, rows, g, lam, omega_relax, 1, partials + g_);
    }
    k_enth_reduce<<<1, 256, 0, st>>>(partials, (int)(2 * g_), nullptr, status, 1, rtol, atol);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_layer(const int* skip, int ny, int nx, int nz, const T* rows, const T* g,
                                 const T* lam_in, T* lam_out, T omega_relax, hipStream_t st) {
    k_enth_adj_layer<T><<<flat_blocks((long)ny * nx * nz), 256, 0, st>>>(skip, ny, nx, nz, rows, g, lam_in, lam_out,
                                                                          omega_relax);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_residual(const int* skip, int ny, int nx, int nz, const T* rows, const T* g,
                                    const T* lam, T* r_out, double* partials, EnthStatus* status, int status_mode,
                                    double rtol, double atol, hipStream_t st) {
    long nb = ((long)ny * nx * nz + 255) / 256;
    unsigned g_ = (unsigned)(nb < kEnthResBlocks ? (nb > 0 ? nb : 1) : kEnthResBlocks);
    k_enth_adj_residual<T><<<g_, 256, 0, st>>>(skip, ny, nx, nz, rows, g, lam, r_out, partials);
    k_enth_reduce<<<1, 256, 0, st>>>(partials, (int)g_, nullptr, status, status_mode, rtol, atol);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_enth_adj_grads(const EnthArgs<T>& A, const T* lam, T* g_E_prev, T* g_phi, T* g_Q, T* g_Es,
                                 hipStream_t st) {
    k_enth_adj_grads<T><<<col_blocks(A.ny, A.nx), 256, 0, st>>>(A, lam, g_E_prev, g_phi, g_Q, g_Es);
    return hipGetLastError();
}

#define INST(T)                                                                                                      \
    template hipError_t launch_enth_residual<T>(const EnthArgs<T>&, T*, int, double*, double*, EnthStatus*, int,     \
                                                double, double, hipStream_t);                                        \
    template hipError_t launch_enth_column_smooth<T>(const EnthArgs<T>&, T*, T*, T, int, T, hipStream_t);            \
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
    template hipError_t launch_enth_adj_rows<T>(const EnthArgs<T>&, T*, hipStream_t);                                \
    template hipError_t launch_enth_adj_rhs<T>(const EnthArgs<T>&, const T*, const T*, T, T, T*, hipStream_t);       \
    template hipError_t launch_enth_adj_column<T>(const int*, int, int, int, const T*, const T*, const T*, T*, T,    \
                                                  hipStream_t);                                                      \
    template hipError_t launch_enth_adj_sweep_rb<T>(const int*, int, int, int, const T*, const T*, T*, T, double*,   \
                                                    EnthStatus*, double, double, hipStream_t);                       \
    template hipError_t launch_enth_adj_layer<T>(const int*, int, int, int, const T*, const T*, const T*, T*, T,     \
                                                 hipStream_t);                                                       \
    template hipError_t launch_enth_adj_residual<T>(const int*, int, int, int, const T*, const T*, const T*, T*,     \
                                                    double*, EnthStatus*, int, double, double, hipStream_t);         \
    template hipError_t launch_enth_adj_grads<T>(const EnthArgs<T>&, const T*, T*, T*, T*, T*, hipStream_t);
INST(double)
#undef INST

}  // namespace glide
