// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of the reference's conservative sigma-coordinate
// enthalpy solver: pressure-melting enthalpy, smoothed diffusivity and water
// content, the time / diffusion / sigma-advection / horizontal Lax-Friedrichs
// advection / drainage terms, the full residual, the column Newton-Thomas
// smoother, the pointwise layer smoother, the omega integration and the
// Paterson-Budd rate factor.  Written from glide/cuda/enthalpy.cu and
// glide/enthalpy.py read as text; the reference computes in FP32 with
// fast-math intrinsics, this restatement in IEEE FP64.
//
// Layouts: E, E_prev, f_E, omega (ny, nx, nz); u3d (nz, ny, nx+1);
// v3d (nz, ny+1, nx); cell fields (ny, nx).  Enthalpy is scaled by E_SCALE.
//
// Pins (tests/test_oracle_enthalpy.py): the reference's analytic tests
// (cold column, Pe = -5 advection-diffusion, temperate-bed flux, two-zone
// Stefan profile, horizontal advection symmetry/boundedness/conservation),
// Thomas vs scipy.linalg.solve_banded, omega against the closed form for
// uniform velocity divergence, zero residual for a uniform state with
// consistent omega.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

// glide/enthalpy.py:46-65
constexpr double RHO_I = 910.0, RHO_W = 1000.0, C_I = 2009.0, K_I = 2.1, L_HEAT = 3.34e5;
constexpr double T_REF = 223.15, T_MELT = 273.15, BETA_CC = 7.9e-8, GRAVITY = 9.81;
constexpr double K_COLD = K_I / C_I, K_TEMP_FACTOR = 0.1;
constexpr double E_SCALE = C_I * (T_MELT - T_REF);
constexpr double SEC_PER_YR = 365.25 * 86400.0;
constexpr double MASS_FLUX_REG = 10.0 / (SEC_PER_YR * SEC_PER_YR);
constexpr double MASS_FLUX_REG_YR = 10.0;
constexpr double DRAIN_SHARPNESS = 0.01;
constexpr int TERM_HORIZ = 1, TERM_SIGMA = 2, TERM_DRAIN = 4;

inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

// glide/cuda/enthalpy.cu:48-98
inline double E_pmp(double sigma, double H) {
    double depth = (1.0 - sigma) * H;
    double Tp = T_MELT - BETA_CC * RHO_I * GRAVITY * depth;
    return C_I * (Tp - T_REF) / E_SCALE;
}
inline double K_of(double E, double Ep) {
    double z = (E - Ep) * E_SCALE * 0.01;
    z = std::fmin(std::fmax(z, -20.0), 20.0);
    double s = 1.0 / (1.0 + std::exp(-z));
    return K_COLD * (1.0 - s + K_TEMP_FACTOR * s);
}
inline double water(double E, double Ep) {
    double x = (E - Ep) * E_SCALE;
    double ax = DRAIN_SHARPNESS * x;
    ax = std::fmin(std::fmax(ax, -20.0), 20.0);
    double sp = (ax > 10.0) ? x : std::log(1.0 + std::exp(ax)) / DRAIN_SHARPNESS;
    return sp / L_HEAT;
}
inline double dwater_dE(double E, double Ep) {
    double z = DRAIN_SHARPNESS * (E - Ep) * E_SCALE;
    z = std::fmin(std::fmax(z, -20.0), 20.0);
    double s = 1.0 / (1.0 + std::exp(-z));
    return s / L_HEAT;
}

struct Col {
    int ny, nx, nz;
    double E3(const double* E, int i, int j, int k) const {
        return E[((size_t)clampi(i, 0, ny - 1) * nx + clampi(j, 0, nx - 1)) * nz + clampi(k, 0, nz - 1)];
    }
    double U3(const double* u, int k, int i, int j) const {
        return u[((size_t)clampi(k, 0, nz - 1) * ny + clampi(i, 0, ny - 1)) * (nx + 1) + clampi(j, 0, nx)];
    }
    double V3(const double* v, int k, int i, int j) const {
        return v[((size_t)clampi(k, 0, nz - 1) * (ny + 1) + clampi(i, 0, ny)) * nx + clampi(j, 0, nx - 1)];
    }
    double C(const double* a, int i, int j) const { return a[(size_t)clampi(i, 0, ny - 1) * nx + clampi(j, 0, nx - 1)]; }
};

struct LF { double res, dl, dr; };
inline LF lf_flux(double u, double El, double Er, double lf_c) {
    double alpha = std::fabs(u) * std::sqrt(1.0 + lf_c);
    return {0.5 * u * (El + Er) - 0.5 * alpha * (Er - El), 0.5 * (u + alpha), 0.5 * (u - alpha)};
}

// glide/cuda/enthalpy.cu:671-763
void horiz_adv(const Col& G, const double* E, const double* u3d, const double* v3d, const double* H, double E_here,
               int i, int j, int k, double dxi, double lf_c, double* res, double* diag) {
    const int ny = G.ny, nx = G.nx;
    double Hh = H[(size_t)i * nx + j];
    double ul = G.U3(u3d, k, i, j), ur = G.U3(u3d, k, i, j + 1);
    double Exm = G.E3(E, i, j - 1, k), Exp = G.E3(E, i, j + 1, k);
    LF fl{0, 0, 0}, fr{0, 0, 0}, ft{0, 0, 0}, fb{0, 0, 0};
    if (j > 0) {
        double Hxm = G.C(H, i, j - 1);
        double Hu = 0.5 * (Hxm + Hh) * ul - 0.5 * std::sqrt(ul * ul + MASS_FLUX_REG) * (Hh - Hxm);
        fl = lf_flux(Hu, Exm, E_here, lf_c);
    } else {
        fl = lf_flux(0.0, E_here, E_here, lf_c);
    }
    if (j < nx - 1) {
        double Hxp = G.C(H, i, j + 1);
        double Hu = 0.5 * (Hh + Hxp) * ur - 0.5 * std::sqrt(ur * ur + MASS_FLUX_REG) * (Hxp - Hh);
        fr = lf_flux(Hu, E_here, Exp, lf_c);
    } else {
        fr = lf_flux(0.0, E_here, E_here, lf_c);
    }
    double vt = G.V3(v3d, k, i, j), vb = G.V3(v3d, k, i + 1, j);
    double Eym = G.E3(E, i - 1, j, k), Eyp = G.E3(E, i + 1, j, k);
    if (i > 0) {
        double Hym = G.C(H, i - 1, j);
        double Hv = 0.5 * (Hym + Hh) * vt - 0.5 * std::sqrt(vt * vt + MASS_FLUX_REG) * (Hym - Hh);
        ft = lf_flux(Hv, E_here, Eym, lf_c);
    } else {
        ft = lf_flux(0.0, E_here, E_here, lf_c);
    }
    if (i < ny - 1) {
        double Hyp = G.C(H, i + 1, j);
        double Hv = 0.5 * (Hh + Hyp) * vb - 0.5 * std::sqrt(vb * vb + MASS_FLUX_REG) * (Hh - Hyp);
        fb = lf_flux(Hv, Eyp, E_here, lf_c);
    } else {
        fb = lf_flux(0.0, E_here, E_here, lf_c);
    }
    *res = RHO_I * ((fr.res - fl.res) + (ft.res - fb.res)) * dxi;
    *diag = RHO_I * ((fr.dl - fl.dr) + (ft.dl - fb.dr)) * dxi;
}

struct Params {
    double dx, dt, drain_rate, h_thin, lf_c;
    int flags;
};

// Residual row k of column (i,j) with vertical neighbours from `Ecol` (the
// column's current values) and horizontal neighbours from the global E;
// returns r (without forcing) and the tridiagonal row (a, b, c).
void node_row(const Col& G, const double* E, const double* Ecol, double Es, const double* Ep, const double* u3d,
              const double* v3d, const double* om, const double* H, const double* Hp, const Params& P, int i, int j,
              int k, double* r_out, double* a_out, double* b_out, double* c_out) {
    const int nx = G.nx, nz = G.nz;
    const size_t col = ((size_t)i * nx + j) * nz;
    const double h = H[(size_t)i * nx + j], hinv = 1.0 / h, hp = Hp[(size_t)i * nx + j];
    const double dsig = 1.0 / (double)(nz - 1);
    const double Ek = Ecol[k];
    const double sig = k * dsig;
    const double Epk = E_pmp(sig, h);
    const double invdt = 1.0 / P.dt;
    double r = RHO_I * (h * Ek - hp * Ep[col + k]) * invdt;
    double b = RHO_I * h * invdt, a = 0.0, c = 0.0;
    if (P.flags & TERM_HORIZ) {
        double hr, hd;
        horiz_adv(G, E, u3d, v3d, H, Ek, i, j, k, 1.0 / P.dx, P.lf_c, &hr, &hd);
        r += hr;
        b += hd;
    }
    if (k == 0) {
        const double Epb = E_pmp(0.0, h);
        const double Ekp1 = Ecol[1];
        if (P.flags & TERM_SIGMA) {
            double w0 = om[col], w1 = om[col + 1];
            double wh = 0.5 * (w0 + w1);
            double wp = std::fmax(wh, 0.0), wn = std::fmin(wh, 0.0);
            double Fu = wp * Ek + wn * Ekp1, Fb = w0 * Ek;
            double hdi = 1.0 / (0.5 * dsig);
            r += RHO_I * (Fu - Fb) * hdi;
            b += RHO_I * (wp - w0) * hdi;
            c += RHO_I * wn * hdi;
        }
        {
            double Kh = K_of(0.5 * (Ek + Ekp1), 0.5 * (Epb + E_pmp(dsig, h)));
            double dh = 0.5 * dsig;
            r += (-hinv * Kh * (Ekp1 - Ek) / dsig - 0.0) / dh;
            double co = hinv * Kh / (dsig * dh);
            b += co;
            c += -co;
        }
        if (P.flags & TERM_DRAIN) {
            r += h * RHO_W * L_HEAT * (P.drain_rate * water(Ek, Epb)) / E_SCALE;
            b += h * RHO_W * L_HEAT * P.drain_rate * dwater_dE(Ek, Epb);
        }
    } else {
        const double Ekm1 = Ecol[k - 1];
        const double Ekp1 = (k < nz - 1) ? Ecol[k + 1] : Es;
        const double Epm = E_pmp(std::max(k - 1, 0) * dsig, h), Epp = E_pmp(std::min(k + 1, nz - 1) * dsig, h);
        if (P.flags & TERM_SIGMA) {
            double wm = om[col + k - 1], w0 = om[col + k], wp_ = om[col + std::min(k + 1, nz - 1)];
            double wu = 0.5 * (w0 + wp_), wl = 0.5 * (wm + w0);
            double up = std::fmax(wu, 0.0), un = std::fmin(wu, 0.0);
            double lp = std::fmax(wl, 0.0), ln = std::fmin(wl, 0.0);
            double Fu = up * Ek + un * Ekp1, Fl = lp * Ekm1 + ln * Ek;
            double di = 1.0 / dsig;
            r += RHO_I * (Fu - Fl) * di;
            a += -RHO_I * lp * di;
            c += RHO_I * un * di;
            b += RHO_I * (up - ln) * di;
        }
        {
            double Ku = K_of(0.5 * (Ek + Ekp1), 0.5 * (Epk + Epp));
            double Kl = K_of(0.5 * (Ek + Ekm1), 0.5 * (Epk + Epm));
            double d2 = 1.0 / (dsig * dsig);
            r += -hinv * d2 * (Ku * (Ekp1 - Ek) - Kl * (Ek - Ekm1));
            double dl = hinv * d2 * Kl, du = hinv * d2 * Ku;
            a += -dl;
            c += -du;
            b += dl + du;
        }
        if (P.flags & TERM_DRAIN) {
            r += h * RHO_W * L_HEAT * (P.drain_rate * water(Ek, Epk)) / E_SCALE;
            b += h * RHO_W * L_HEAT * P.drain_rate * dwater_dE(Ek, Epk);
        }
    }
    *r_out = r;
    *a_out = a;
    *b_out = b;
    *c_out = c;
}

// ------------------------------------------------------------ exact Jacobian
// Derivatives used by the adjoint of the converged enthalpy solve (no
// counterpart in the reference).  Unlike the Newton rows above (frozen K, the
// drainage sigmoid), these differentiate every term: dK/dE of the smoothed
// diffusivity and the exact derivative of the drainage softplus branches.
inline double dK_dE(double E, double Ep) {
    double zr = (E - Ep) * E_SCALE * 0.01;
    if (zr < -20.0 || zr > 20.0) return 0.0;
    double s = 1.0 / (1.0 + std::exp(-zr));
    return -0.9 * K_COLD * s * (1.0 - s) * E_SCALE * 0.01;
}
inline double dsoftplus_dx(double E, double Ep) {
    double ax = DRAIN_SHARPNESS * (E - Ep) * E_SCALE;
    if (ax > 10.0) return 1.0;
    if (ax < -20.0) return 0.0;
    return 1.0 / (1.0 + std::exp(-ax));
}

// Row of the exact Jacobian of node (i,j,k) (interior, non-thin): column
// coefficients a (E_{k-1}), b (E_k), c (E_{k+1}) and horizontal couplings
// xm, xp, ym, yp (E of the neighbouring columns at level k).
void exact_row(const Col& G, const double* E, const double* u3d, const double* v3d, const double* om,
               const double* H, const double* Hp, const Params& P, int i, int j, int k, double* row) {
    const int nx = G.nx, nz = G.nz;
    const size_t col = ((size_t)i * nx + j) * nz;
    const double h = H[(size_t)i * nx + j], hinv = 1.0 / h;
    (void)Hp;
    const double dsig = 1.0 / (double)(nz - 1);
    const double Ek = E[col + k];
    const double Epk = E_pmp(k * dsig, h);
    double a = 0, b = RHO_I * h / P.dt, c = 0, xm = 0, xp = 0, ym = 0, yp = 0;
    if (P.flags & TERM_HORIZ) {
        const int ny = G.ny;
        const double sq = std::sqrt(1.0 + P.lf_c), dxi = 1.0 / P.dx;
        double Hh = h;
        auto U3 = [&](int kk, int ii, int jj) { return G.U3(u3d, kk, ii, jj); };
        auto V3 = [&](int kk, int ii, int jj) { return G.V3(v3d, kk, ii, jj); };
        auto dl = [&](double U) { return 0.5 * (U + std::fabs(U) * sq); };
        auto dr = [&](double U) { return 0.5 * (U - std::fabs(U) * sq); };
        double Ul = 0, Ur = 0, Ut = 0, Ub = 0;
        if (j > 0) {
            double ul = U3(k, i, j), Hm = G.C(H, i, j - 1);
            Ul = 0.5 * (Hm + Hh) * ul - 0.5 * std::sqrt(ul * ul + MASS_FLUX_REG) * (Hh - Hm);
        }
        if (j < nx - 1) {
            double ur = U3(k, i, j + 1), Hq = G.C(H, i, j + 1);
            Ur = 0.5 * (Hh + Hq) * ur - 0.5 * std::sqrt(ur * ur + MASS_FLUX_REG) * (Hq - Hh);
        }
        if (i > 0) {
            double vt = V3(k, i, j), Hm = G.C(H, i - 1, j);
            Ut = 0.5 * (Hm + Hh) * vt - 0.5 * std::sqrt(vt * vt + MASS_FLUX_REG) * (Hm - Hh);
        }
        if (i < ny - 1) {
            double vb = V3(k, i + 1, j), Hq = G.C(H, i + 1, j);
            Ub = 0.5 * (Hh + Hq) * vb - 0.5 * std::sqrt(vb * vb + MASS_FLUX_REG) * (Hh - Hq);
        }
        xm = -RHO_I * dxi * dl(Ul);
        xp = RHO_I * dxi * dr(Ur);
        ym = RHO_I * dxi * dr(Ut);
        yp = -RHO_I * dxi * dl(Ub);
        b += RHO_I * dxi * ((dl(Ur) - dr(Ul)) + (dl(Ut) - dr(Ub)));
    }
    if (k == 0) {
        const double E1 = E[col + 1];
        if (P.flags & TERM_SIGMA) {
            double w0 = om[col], w1 = om[col + 1];
            double wh = 0.5 * (w0 + w1);
            double wp = std::fmax(wh, 0.0), wn = std::fmin(wh, 0.0);
            double hdi = 1.0 / (0.5 * dsig);
            b += RHO_I * (wp - w0) * hdi;
            c += RHO_I * wn * hdi;
        }
        const double Em = 0.5 * (Ek + E1), Epm = 0.5 * (Epk + E_pmp(dsig, h));
        const double Kh = K_of(Em, Epm), dKh = dK_dE(Em, Epm);
        const double f = -hinv / (dsig * 0.5 * dsig);
        b += f * (dKh * 0.5 * (E1 - Ek) - Kh);
        c += f * (dKh * 0.5 * (E1 - Ek) + Kh);
    } else {
        const double Ekm1 = E[col + k - 1], Ekp1 = E[col + k + 1];
        if (P.flags & TERM_SIGMA) {
            double wm = om[col + k - 1], w0 = om[col + k], wp_ = om[col + std::min(k + 1, nz - 1)];
            double wu = 0.5 * (w0 + wp_), wl = 0.5 * (wm + w0);
            double up = std::fmax(wu, 0.0), un = std::fmin(wu, 0.0);
            double lp = std::fmax(wl, 0.0), ln = std::fmin(wl, 0.0);
            double di = 1.0 / dsig;
            a += -RHO_I * lp * di;
            c += RHO_I * un * di;
            b += RHO_I * (up - ln) * di;
        }
        const double Epm = E_pmp(std::max(k - 1, 0) * dsig, h), Epp = E_pmp(std::min(k + 1, nz - 1) * dsig, h);
        const double Eu = 0.5 * (Ek + Ekp1), Epu = 0.5 * (Epk + Epp);
        const double El = 0.5 * (Ek + Ekm1), Epl = 0.5 * (Epk + Epm);
        const double Ku = K_of(Eu, Epu), dKu = dK_dE(Eu, Epu), Kl = K_of(El, Epl), dKl = dK_dE(El, Epl);
        const double f = -hinv / (dsig * dsig);
        c += f * (dKu * 0.5 * (Ekp1 - Ek) + Ku);
        b += f * (dKu * 0.5 * (Ekp1 - Ek) - Ku - dKl * 0.5 * (Ek - Ekm1) - Kl);
        a += f * (-dKl * 0.5 * (Ek - Ekm1) + Kl);
    }
    if (P.flags & TERM_DRAIN) b += h * RHO_W * P.drain_rate * dsoftplus_dx(Ek, Epk);
    row[0] = a; row[1] = b; row[2] = c; row[3] = xm; row[4] = xp; row[5] = ym; row[6] = yp;
}

}  // namespace

extern "C" {

typedef struct orc_enth_params {
    double dx, dt, drain_rate, h_thin, lf_c;
    int term_flags;
} orc_enth_params;

double orc_E_pmp(double sigma, double H) { return E_pmp(sigma, H); }
double orc_K(double E, double Ep) { return K_of(E, Ep); }
double orc_water(double E, double Ep) { return water(E, Ep); }

// glide/cuda/enthalpy.cu:783-905; returns max |out|.
double orc_enth_residual(double* out, const double* E, const double* Ep, const double* fE, const double* u3d,
                         const double* v3d, const double* om, const double* H, const double* Hp, const double* Es,
                         const orc_enth_params* pp, int use_forcing, int ny, int nx, int nz) {
    Col G{ny, nx, nz};
    Params P{pp->dx, pp->dt, pp->drain_rate, pp->h_thin, pp->lf_c, pp->term_flags};
    double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int c = 0; c < ny * nx; ++c) {
        int i = c / nx, j = c % nx;
        size_t col = (size_t)c * nz;
        if (H[c] < P.h_thin) {
            for (int k = 0; k < nz; ++k) out[col + k] = 0.0;
            continue;
        }
        for (int k = 0; k < nz; ++k) {
            if (k == nz - 1) { out[col + k] = 0.0; continue; }
            double r, a, b, cc;
            node_row(G, E, E + col, Es[c], Ep, u3d, v3d, om, H, Hp, P, i, j, k, &r, &a, &b, &cc);
            double v = use_forcing ? (r - fE[col + k]) : r;
            out[col + k] = v;
            mx = std::fmax(mx, std::fabs(v));
        }
    }
    return mx;
}

// glide/cuda/enthalpy.cu:932-1112
void orc_enth_column_smooth(double* dE, const double* E, const double* Ep, const double* fE, const double* u3d,
                            const double* v3d, const double* om, const double* H, const double* Hp, const double* Es,
                            const orc_enth_params* pp, int n_newton, double relaxation, int ny, int nx, int nz) {
    Col G{ny, nx, nz};
    Params P{pp->dx, pp->dt, pp->drain_rate, pp->h_thin, pp->lf_c, pp->term_flags};
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ny * nx; ++c) {
        int i = c / nx, j = c % nx;
        size_t col = (size_t)c * nz;
        double h = H[c];
        if (h < P.h_thin) {
            double dsig = 1.0 / (double)(nz - 1);
            for (int k = 0; k < nz; ++k) dE[col + k] = std::fmin(Es[c], E_pmp(k * dsig, h)) - E[col + k];
            continue;
        }
        std::vector<double> El(E + col, E + col + nz), a(nz), b(nz), cc(nz), rhs(nz), d(nz);
        for (int it = 0; it < n_newton; ++it) {
            for (int k = 0; k < nz; ++k) {
                if (k == nz - 1) {
                    a[k] = 0.0; b[k] = 1.0; cc[k] = 0.0; rhs[k] = -(El[k] - Es[c]);
                    continue;
                }
                double r;
                node_row(G, E, El.data(), Es[c], Ep, u3d, v3d, om, H, Hp, P, i, j, k, &r, &a[k], &b[k], &cc[k]);
                r -= fE[col + k];
                if (k == 0) a[k] = 0.0;
                rhs[k] = -r;
            }
            for (int k = 1; k < nz; ++k) {
                if (std::fabs(b[k - 1]) < 1e-30) continue;
                double w = a[k] / b[k - 1];
                b[k] -= w * cc[k - 1];
                rhs[k] -= w * rhs[k - 1];
            }
            d[nz - 1] = rhs[nz - 1] / b[nz - 1];
            for (int k = nz - 2; k >= 0; --k) d[k] = (rhs[k] - cc[k] * d[k + 1]) / b[k];
            for (int k = 0; k < nz; ++k) El[k] += relaxation * d[k];
        }
        for (int k = 0; k < nz; ++k) dE[col + k] = El[k] - E[col + k];
    }
}

// glide/cuda/enthalpy.cu:1128-1255
void orc_enth_layer_smooth(double* dE, const double* E, const double* Ep, const double* fE, const double* u3d,
                           const double* v3d, const double* om, const double* H, const double* Hp, const double* Es,
                           const orc_enth_params* pp, int ny, int nx, int nz) {
    Col G{ny, nx, nz};
    Params P{pp->dx, pp->dt, pp->drain_rate, pp->h_thin, pp->lf_c, pp->term_flags};
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ny * nx; ++c) {
        int i = c / nx, j = c % nx;
        size_t col = (size_t)c * nz;
        for (int k = 0; k < nz; ++k) {
            if (k == nz - 1 || H[c] < P.h_thin) { dE[col + k] = 0.0; continue; }
            double r, a, b, cc;
            node_row(G, E, E + col, Es[c], Ep, u3d, v3d, om, H, Hp, P, i, j, k, &r, &a, &b, &cc);
            r -= fE[col + k];
            dE[col + k] = (std::fabs(b) > 1e-30) ? (-r / b) : 0.0;
        }
    }
}

// glide/cuda/enthalpy.cu:1276-1372 (velocities and dH/dt in m/a)
void orc_enth_omega(double* om, const double* u3d, const double* v3d, const double* H, const double* dhdt,
                    const double* bmb, double dx, int ny, int nx, int nz) {
    Col G{ny, nx, nz};
    const double dxi = 1.0 / dx, dsig = 1.0 / (double)(nz - 1);
    for (int c = 0; c < ny * nx; ++c) {
        int i = c / nx, j = c % nx;
        size_t col = (size_t)c * nz;
        double h = H[c];
        if (h < 1e-3) {
            for (int k = 0; k < nz; ++k) om[col + k] = 0.0;
            continue;
        }
        double Hxm = G.C(H, i, j - 1), Hxp = G.C(H, i, j + 1), Hym = G.C(H, i - 1, j), Hyp = G.C(H, i + 1, j);
        std::vector<double> div(nz);
        for (int k = 0; k < nz; ++k) {
            double ul = G.U3(u3d, k, i, j), ur = G.U3(u3d, k, i, j + 1);
            double vt = G.V3(v3d, k, i, j), vb = G.V3(v3d, k, i + 1, j);
            double fr = 0, fl = 0, fb = 0, ft = 0;
            if (j < nx - 1) fr = 0.5 * (h + Hxp) * ur - 0.5 * std::sqrt(ur * ur + MASS_FLUX_REG_YR) * (Hxp - h);
            if (j > 0) fl = 0.5 * (Hxm + h) * ul - 0.5 * std::sqrt(ul * ul + MASS_FLUX_REG_YR) * (h - Hxm);
            if (i < ny - 1) fb = 0.5 * (h + Hyp) * vb - 0.5 * std::sqrt(vb * vb + MASS_FLUX_REG_YR) * (h - Hyp);
            if (i > 0) ft = 0.5 * (Hym + h) * vt - 0.5 * std::sqrt(vt * vt + MASS_FLUX_REG_YR) * (Hym - h);
            div[k] = (fr - fl + ft - fb) * dxi;
        }
        double w = -bmb[c];
        om[col] = w;
        for (int k = 1; k < nz; ++k) {
            w = w - dsig * (dhdt[c] + 0.5 * (div[k - 1] + div[k]));
            om[col + k] = w;
        }
    }
}

// EnthalpyOperators.column_sweep (glide/enthalpy.py:426-471); E updated in
// place; returns the number of sweeps taken, r0 and the final max |r|.
int orc_enth_column_sweep(double* E, const double* Ep, const double* fE, const double* u3d, const double* v3d,
                          const double* om, const double* H, const double* Hp, const double* Es,
                          const orc_enth_params* pp, int n_iter, int alternating, double omega_relax, int n_newton,
                          double relaxation, double rtol, double atol, double* r0_out, double* r_out, int ny, int nx,
                          int nz) {
    const size_t n = (size_t)ny * nx * nz;
    std::vector<double> r(n), d(n);
    double r0 = orc_enth_residual(r.data(), E, Ep, fE, u3d, v3d, om, H, Hp, Es, pp, 1, ny, nx, nz);
    double rr = r0;
    int it = 0;
    for (; it < n_iter;) {
        orc_enth_column_smooth(d.data(), E, Ep, fE, u3d, v3d, om, H, Hp, Es, pp, n_newton, relaxation, ny, nx, nz);
        for (size_t q = 0; q < n; ++q) E[q] += omega_relax * d[q];
        if (alternating) {
            orc_enth_layer_smooth(d.data(), E, Ep, fE, u3d, v3d, om, H, Hp, Es, pp, ny, nx, nz);
            for (size_t q = 0; q < n; ++q) E[q] += omega_relax * d[q];
        }
        rr = orc_enth_residual(r.data(), E, Ep, fE, u3d, v3d, om, H, Hp, Es, pp, 1, ny, nx, nz);
        ++it;
        double rel = r0 > 0 ? rr / r0 : 0.0;
        if (rel < rtol || rr < atol) break;
    }
    if (r0_out) *r0_out = r0;
    if (r_out) *r_out = rr;
    return it;
}

// EnthalpyOperators.get_arrhenius_factor (glide/enthalpy.py:600-643): depth
// mean of the Paterson-Budd A(T) (1 + 181.25 w), returned as B = A^(-1/n).
void orc_enth_arrhenius(double* B, const double* E, const double* H, double n_glen, int ny, int nx, int nz) {
    const double dsig = 1.0 / (double)(nz - 1);
    for (int c = 0; c < ny * nx; ++c) {
        double s = 0.0;
        for (int k = 0; k < nz; ++k) {
            double sig = k * dsig;
            double depth = (1.0 - sig) * H[c];
            double Tp = T_MELT - BETA_CC * RHO_I * GRAVITY * depth;
            double Ek = E[(size_t)c * nz + k];
            double T = std::fmin(Ek * E_SCALE / C_I + T_REF, Tp);
            double Af = T < 263.15 ? 3.985e-13 : 1.916e3;
            double Qf = T < 263.15 ? 60e3 : 139e3;
            double A = Af * std::exp(-Qf / (8.314 * std::fmax(T, 1.0)));
            double Epmp = C_I * (Tp - T_REF);
            double w = std::fmax(Ek * E_SCALE - Epmp, 0.0) / L_HEAT;
            s += A * (1.0 + 181.25 * w);
        }
        B[c] = std::pow(s / nz, -1.0 / n_glen);
    }
}

// System G(E) = 0 whose solution the column sweep converges to: the residual
// on interior nodes of non-thin columns, E - E_surface on their surface node,
// E - min(E_surface, E_pmp) on thin columns (no counterpart in the reference;
// the adjoint's system).
void orc_enth_system_residual(double* out, const double* E, const double* Ep, const double* fE, const double* u3d,
                              const double* v3d, const double* om, const double* H, const double* Hp,
                              const double* Es, const orc_enth_params* pp, int ny, int nx, int nz) {
    orc_enth_residual(out, E, Ep, fE, u3d, v3d, om, H, Hp, Es, pp, 1, ny, nx, nz);
    const double dsig = 1.0 / (double)(nz - 1);
    for (int c = 0; c < ny * nx; ++c) {
        size_t col = (size_t)c * nz;
        if (H[c] < pp->h_thin) {
            for (int k = 0; k < nz; ++k) out[col + k] = E[col + k] - std::fmin(Es[c], E_pmp(k * dsig, H[c]));
        } else {
            out[col + nz - 1] = E[col + nz - 1] - Es[c];
        }
    }
}

// Rows (a, b, c, xm, xp, ym, yp) of dG/dE at E, (ny, nx, nz, 7); identity
// rows on surface and thin-column nodes.
void orc_enth_jacobian_rows(double* rows, const double* E, const double* u3d, const double* v3d, const double* om,
                            const double* H, const double* Hp, const orc_enth_params* pp, int ny, int nx, int nz) {
    Col G{ny, nx, nz};
    Params P{pp->dx, pp->dt, pp->drain_rate, pp->h_thin, pp->lf_c, pp->term_flags};
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ny * nx; ++c) {
        int i = c / nx, j = c % nx;
        for (int k = 0; k < nz; ++k) {
            double* row = rows + ((size_t)c * nz + k) * 7;
            if (H[c] < P.h_thin || k == nz - 1) {
                for (int q = 0; q < 7; ++q) row[q] = 0.0;
                row[1] = 1.0;
                continue;
            }
            exact_row(G, E, u3d, v3d, om, H, Hp, P, i, j, k, row);
        }
    }
}

// dB/dE^T gB: the rate factor's derivative (B as orc_enth_arrhenius divided by
// `scale`) applied to the cotangent gB (ny, nx), into gE (ny, nx, nz).
void orc_enth_arrhenius_vjp(double* gE, const double* gB, const double* E, const double* H, double n_glen,
                            double scale, int ny, int nx, int nz) {
    const double dsig = 1.0 / (double)(nz - 1);
    for (int c = 0; c < ny * nx; ++c) {
        double s = 0.0;
        std::vector<double> dA(nz);
        for (int k = 0; k < nz; ++k) {
            double depth = (1.0 - k * dsig) * H[c];
            double Tp = T_MELT - BETA_CC * RHO_I * GRAVITY * depth;
            double Ek = E[(size_t)c * nz + k];
            double Tu = Ek * E_SCALE / C_I + T_REF;
            double T = std::fmin(Tu, Tp);
            double Af = T < 263.15 ? 3.985e-13 : 1.916e3;
            double Qf = T < 263.15 ? 60e3 : 139e3;
            double Ts = std::fmax(T, 1.0);
            double A = Af * std::exp(-Qf / (8.314 * Ts));
            double Epmp = C_I * (Tp - T_REF);
            double wraw = Ek * E_SCALE - Epmp;
            double w = std::fmax(wraw, 0.0) / L_HEAT;
            s += A * (1.0 + 181.25 * w);
            double dT = (Tu < Tp && T > 1.0) ? E_SCALE / C_I : 0.0;
            double dAdE = A * Qf / (8.314 * Ts * Ts) * dT;
            double dwdE = wraw > 0.0 ? E_SCALE / L_HEAT : 0.0;
            dA[k] = dAdE * (1.0 + 181.25 * w) + A * 181.25 * dwdE;
        }
        double mean = s / nz;
        double dBdmean = (-1.0 / n_glen) * std::pow(mean, -1.0 / n_glen - 1.0) / scale;
        for (int k = 0; k < nz; ++k) gE[(size_t)c * nz + k] = gB[c] * dBdmean * dA[k] / nz;
    }
}

}  // extern "C"
