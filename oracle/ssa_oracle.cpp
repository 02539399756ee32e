// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of the reference's coupled SSA momentum + mass
// discretisation, its additive 5x5 Vanka smoother, the multigrid transfers,
// the FAS constrained-descent V-cycle, the transpose (VJP) operator, the
// adjoint Vanka smoother, the FAS adjoint V-cycle and the parameter
// gradients.  It exists to check the HIP product path: only the tests, the
// smoke check and the CPU-baseline leg of the benchmark load it.  Nothing in
// the product links, imports or falls back to it.
//
// Written from the reference's CUDA/Python sources read as text:
//   glide/cuda/common.cu, viscosity.cu, stress.cu, flux.cu, residuals.cu,
//   vanka.cu, grad.cu, transfer.cu; glide/operators.py, glide/multigrid.py,
//   glide/model.py.
// The reference computes in FP32 with fast-math intrinsics; this restatement
// computes every operation in IEEE FP64 with libm pow/exp/sqrt.
//
// Pins (tests/test_oracle_ssa.py): closed-form flux/stress/viscosity scalars,
// residual == zero for a consistent manufactured state, mass telescoping,
// central-FD derivative checks of the local Vanka Jacobian entries the
// reference assembles, a dot-product (Jᵀ vs FD-J) test of the VJP, numpy
// re-computations of every transfer, numpy.linalg.solve of the 5x5 LU, and
// FD checks of the beta/bed gradients against the residual.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// Physical constants of one level (glide/grid.py:29-145 defaults).
typedef struct orc_phys {
    double n, eps_reg, flotation_reg_driving;
    double m, u_reg, water_drag, flotation_reg_sliding;
    double calving_rate, flotation_reg_calving;
    double thklim;
} orc_phys;

}  // extern "C"

namespace {

inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

// Clamp-to-edge accessors (glide/cuda/common.cu:95-142).
struct Acc {
    int ny, nx;
    double U(const double* a, int i, int j) const { return a[clampi(i, 0, ny - 1) * (nx + 1) + clampi(j, 0, nx)]; }
    double V(const double* a, int i, int j) const { return a[clampi(i, 0, ny) * nx + clampi(j, 0, nx - 1)]; }
    double C(const double* a, int i, int j) const { return a[clampi(i, 0, ny - 1) * nx + clampi(j, 0, nx - 1)]; }
};

// ------------------------------------------------------------------ physics
// Sigmoid grounded fraction (glide/cuda/common.cu:66-82).
inline double grounded_fraction(double H, double bed, double c) {
    double z = c * (bed + 0.917 * H);
    z = std::fmin(std::fmax(z, -20.0), 20.0);
    return 1.0 / (1.0 + std::exp(-z));
}

// Cell viscosity and its directional derivative along (du, dv)
// (glide/cuda/viscosity.cu:62-168).
struct EtaD { double v, d; };

EtaD cell_viscosity(const Acc& A, const double* u, const double* v, const double* du, const double* dv,
                    const double* B, int i, int j, double n, double eps_reg, double dx) {
    const double dxi = 1.0 / dx;
    const double p = (1.0 - n) / (2.0 * n);
    const int ny = A.ny, nx = A.nx;
    double ul = A.U(u, i, j), ur = A.U(u, i, j + 1), vt = A.V(v, i, j), vb = A.V(v, i + 1, j);
    double exx = (ur - ul) * dxi, eyy = (vt - vb) * dxi;
    double exx_d = 0, eyy_d = 0;
    double dul = 0, dur = 0, dvt = 0, dvb = 0;
    if (du) {
        dul = A.U(du, i, j); dur = A.U(du, i, j + 1); dvt = A.V(dv, i, j); dvb = A.V(dv, i + 1, j);
        exx_d = (dur - dul) * dxi; eyy_d = (dvt - dvb) * dxi;
    }
    double s2 = 0, s2_d = 0;
    // four corner shear rates, each masked at the domain's boundary vertices
    {
        double e = 0, ed = 0;
        if (i > 0 && j > 0) {
            e = 0.5 * ((A.U(u, i - 1, j) - ul) * dxi + (vt - A.V(v, i, j - 1)) * dxi);
            if (du) ed = 0.5 * ((A.U(du, i - 1, j) - dul) * dxi + (dvt - A.V(dv, i, j - 1)) * dxi);
        }
        s2 += e * e; s2_d += 2 * e * ed;
    }
    {
        double e = 0, ed = 0;
        if (i > 0 && j < nx - 1) {
            e = 0.5 * ((A.U(u, i - 1, j + 1) - ur) * dxi + (A.V(v, i, j + 1) - vt) * dxi);
            if (du) ed = 0.5 * ((A.U(du, i - 1, j + 1) - dur) * dxi + (A.V(dv, i, j + 1) - dvt) * dxi);
        }
        s2 += e * e; s2_d += 2 * e * ed;
    }
    {
        double e = 0, ed = 0;
        if (i < ny - 1 && j > 0) {
            e = 0.5 * ((ul - A.U(u, i + 1, j)) * dxi + (vb - A.V(v, i + 1, j - 1)) * dxi);
            if (du) ed = 0.5 * ((dul - A.U(du, i + 1, j)) * dxi + (dvb - A.V(dv, i + 1, j - 1)) * dxi);
        }
        s2 += e * e; s2_d += 2 * e * ed;
    }
    {
        double e = 0, ed = 0;
        if (i < ny - 1 && j < nx - 1) {
            e = 0.5 * ((ur - A.U(u, i + 1, j + 1)) * dxi + (A.V(v, i + 1, j + 1) - vb) * dxi);
            if (du) ed = 0.5 * ((dur - A.U(du, i + 1, j + 1)) * dxi + (A.V(dv, i + 1, j + 1) - dvb) * dxi);
        }
        s2 += e * e; s2_d += 2 * e * ed;
    }
    double eII = exx * exx + eyy * eyy + exx * eyy + 0.25 * s2 + eps_reg;
    double eII_d = 2 * exx * exx_d + 2 * eyy * eyy_d + exx * eyy_d + exx_d * eyy + 0.25 * s2_d;
    double b = A.C(B, i, j);
    EtaD r;
    r.v = 0.5 * b * std::pow(eII, p);
    r.d = du ? 0.5 * b * p * std::pow(eII, p - 1.0) * eII_d : 0.0;
    return r;
}

// Strain-rate invariant of cell (i,j) without regularisation: the same
// terms, in the same order, as cell_viscosity's eII.
double strain_invariant0(const Acc& A, const double* u, const double* v, int i, int j, double dx) {
    const double dxi = 1.0 / dx;
    const int ny = A.ny, nx = A.nx;
    double ul = A.U(u, i, j), ur = A.U(u, i, j + 1), vt = A.V(v, i, j), vb = A.V(v, i + 1, j);
    double exx = (ur - ul) * dxi, eyy = (vt - vb) * dxi;
    double s2 = 0;
    double e = (i > 0 && j > 0) ? 0.5 * ((A.U(u, i - 1, j) - ul) * dxi + (vt - A.V(v, i, j - 1)) * dxi) : 0.0;
    s2 += e * e;
    e = (i > 0 && j < nx - 1) ? 0.5 * ((A.U(u, i - 1, j + 1) - ur) * dxi + (A.V(v, i, j + 1) - vt) * dxi) : 0.0;
    s2 += e * e;
    e = (i < ny - 1 && j > 0) ? 0.5 * ((ul - A.U(u, i + 1, j)) * dxi + (vb - A.V(v, i + 1, j - 1)) * dxi) : 0.0;
    s2 += e * e;
    e = (i < ny - 1 && j < nx - 1) ? 0.5 * ((ur - A.U(u, i + 1, j + 1)) * dxi + (A.V(v, i + 1, j + 1) - vb) * dxi)
                                   : 0.0;
    s2 += e * e;
    return exx * exx + eyy * eyy + exx * eyy + 0.25 * s2;
}

// Normal stresses on cell ci/cj (glide/cuda/stress.cu:43-112).
struct Normal { double res, d_ul, d_ur, d_vt, d_vb, d_etaH; };
Normal sigma_xx(double ul, double ur, double vt, double vb, double etaH, double dxi, int cj, int nx) {
    Normal s{0, 0, 0, 0, 0, 0};
    if (cj < 0 || cj >= nx) return s;
    double e = 2.0 * (ur - ul) * dxi + (vt - vb) * dxi;
    double k = 2.0 * etaH * dxi;
    s.res = 2.0 * etaH * e;
    s.d_ul = -2.0 * k; s.d_ur = 2.0 * k; s.d_vt = k; s.d_vb = -k; s.d_etaH = 2.0 * e;
    return s;
}
Normal sigma_yy(double ul, double ur, double vt, double vb, double etaH, double dxi, int ci, int ny) {
    Normal s{0, 0, 0, 0, 0, 0};
    if (ci < 0 || ci >= ny) return s;
    double e = (ur - ul) * dxi + 2.0 * (vt - vb) * dxi;
    double k = 2.0 * etaH * dxi;
    s.res = 2.0 * etaH * e;
    s.d_ul = -k; s.d_ur = k; s.d_vt = 2.0 * k; s.d_vb = -2.0 * k; s.d_etaH = 2.0 * e;
    return s;
}
// Shear stress on vertex (vi, vj); zero on the domain's boundary vertices
// (glide/cuda/stress.cu:159-183).
struct Shear { double res, d_ut, d_ub, d_vl, d_vr, d_etaH; };
Shear sigma_xy(double ut, double ub, double vl, double vr, double etaH, double dxi, int vi, int vj, int ny, int nx) {
    Shear s{0, 0, 0, 0, 0, 0};
    if (vi <= 0 || vi >= ny || vj <= 0 || vj >= nx) return s;
    double e = 0.5 * ((ut - ub) * dxi + (vr - vl) * dxi);
    double k = etaH * dxi;
    s.res = 2.0 * etaH * e;
    s.d_ut = k; s.d_ub = -k; s.d_vl = -k; s.d_vr = k; s.d_etaH = 2.0 * e;
    return s;
}
// Basal drag on a facet (glide/cuda/stress.cu:253-369); x and y forms share
// the arithmetic: w is the facet velocity, c0..c3 the four cross-component
// neighbours, (betaA, phiA) / (betaB, phiB) the two adjacent cells.
struct Drag { double res, d_w, d_c[4], d_betaA, d_betaB; };
Drag basal_drag(double w, const double c[4], double phiA, double phiB, double betaA, double betaB, const orc_phys& P) {
    Drag t;
    double beff = 0.5 * (betaA * phiA + betaB * phiB);
    double s2 = w * w + 0.25 * (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + c[3] * c[3]);
    double q = (P.m - 1.0) / 2.0;
    double pw = std::pow(s2 + P.u_reg, q);
    double dpw = q * std::pow(s2 + P.u_reg, q - 1.0);
    t.res = -(beff * pw + P.water_drag) * w;
    t.d_w = -(beff * (2.0 * dpw * w * w + pw) + P.water_drag);
    for (int k = 0; k < 4; ++k) t.d_c[k] = -beff * (0.5 * dpw * w * c[k]);
    t.d_betaA = -0.5 * phiA * pw * w;
    t.d_betaB = -0.5 * phiB * pw * w;
    return t;
}
// Driving stress across a facet between cell A (left/top) and cell B
// (right/bottom): res = Havg * (S_B - S_A) / dx for x, (S_A - S_B)/dx for y
// (glide/cuda/stress.cu:418-538).  `inner` is false on the domain edge.
struct Drive { double res, d_HA, d_HB, d_bedA, d_bedB; };
Drive driving_x(double Hl, double Hr, double bl, double br, double pl, double pr, double dxi, bool inner) {
    Drive d{0, 0, 0, 0, 0};
    if (!inner) return d;
    double Havg = 0.5 * (Hl + Hr);
    double basel = pl * bl - (1.0 - pl) * 0.917 * Hl;
    double baser = pr * br - (1.0 - pr) * 0.917 * Hr;
    double dbl = -(1.0 - pl) * 0.917, dbr = -(1.0 - pr) * 0.917;
    double Sl = basel + Hl, Sr = baser + Hr;
    d.res = Havg * (Sr - Sl) * dxi;
    d.d_HA = 0.5 * (Sr - Sl) * dxi - Havg * (1.0 + dbl) * dxi;
    d.d_HB = 0.5 * (Sr - Sl) * dxi + Havg * (1.0 + dbr) * dxi;
    d.d_bedA = -Havg * pl * dxi;
    d.d_bedB = Havg * pr * dxi;
    return d;
}
Drive driving_y(double Ht, double Hb, double bt, double bb, double pt, double pb, double dxi, bool inner) {
    Drive d{0, 0, 0, 0, 0};
    if (!inner) return d;
    double Havg = 0.5 * (Ht + Hb);
    double baset = pt * bt - (1.0 - pt) * 0.917 * Ht;
    double baseb = pb * bb - (1.0 - pb) * 0.917 * Hb;
    double dbt = -(1.0 - pt) * 0.917, dbb = -(1.0 - pb) * 0.917;
    double St = baset + Ht, Sb = baseb + Hb;
    d.res = Havg * (St - Sb) * dxi;
    d.d_HA = 0.5 * (St - Sb) * dxi + Havg * (1.0 + dbt) * dxi;
    d.d_HB = 0.5 * (St - Sb) * dxi - Havg * (1.0 + dbb) * dxi;
    d.d_bedA = Havg * pt * dxi;
    d.d_bedB = -Havg * pb * dxi;
    return d;
}
// Lax-Friedrichs mass flux through an x facet (H_A left, H_B right) or a
// y facet (H_A top, H_B bottom, positive velocity points to the top);
// zero on the domain edge (glide/cuda/flux.cu:39-135).
struct Flux { double res, d_w, d_HA, d_HB; };
Flux flux_x(double w, double HA, double HB, bool inner) {
    Flux f{0, 0, 0, 0};
    if (!inner) return f;
    double Havg = 0.5 * (HA + HB);
    double mag = std::sqrt(w * w + 10.0);
    f.res = Havg * w - 0.5 * mag * (HB - HA);
    f.d_HA = 0.5 * (w + mag);
    f.d_HB = 0.5 * (w - mag);
    f.d_w = Havg - 0.5 * (w / mag) * (HB - HA);
    return f;
}
Flux flux_y(double w, double HA, double HB, bool inner) {
    Flux f{0, 0, 0, 0};
    if (!inner) return f;
    double Havg = 0.5 * (HA + HB);
    double mag = std::sqrt(w * w + 10.0);
    f.res = Havg * w - 0.5 * mag * (HA - HB);
    f.d_HA = 0.5 * (w - mag);
    f.d_HB = 0.5 * (w + mag);
    f.d_w = Havg - 0.5 * (w / mag) * (HA - HB);
    return f;
}
// Facet calving outflow (glide/cuda/flux.cu:244-273): no boundary test.
inline double calving(double Hthis, double phithis, double phiother, double rate, double* dH) {
    double k = (1.0 - phithis) * (1.0 - phiother) * rate;
    *dH = k;
    return k * Hthis;
}

// ------------------------------------------------------------------ viscosity field
void viscosity_field(const Acc& A, const double* u, const double* v, const double* du, const double* dv,
                     const double* B, double n, double eps, double dx, std::vector<double>& eta,
                     std::vector<double>* etad) {
    const int ny = A.ny, nx = A.nx;
    eta.assign((size_t)ny * nx, 0.0);
    if (etad) etad->assign((size_t)ny * nx, 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            EtaD e = cell_viscosity(A, u, v, du, dv, B, i, j, n, eps, dx);
            eta[(size_t)i * nx + j] = e.v;
            if (etad) (*etad)[(size_t)i * nx + j] = e.d;
        }
}

struct Fields {
    const double *u, *v, *H, *phi, *mask, *f_u, *f_v, *f_H, *bed, *B, *beta, *gamma;
};

// ------------------------------------------------------------------ residual
// Coupled residual (glide/cuda/residuals.cu:5-333).
void residual(const Fields& F, double* r_u, double* r_v, double* r_H, bool use_forcing, bool use_mask,
              const orc_phys& P, double dx, double dt, int ny, int nx) {
    Acc A{ny, nx};
    const double dxi = 1.0 / dx;
    std::vector<double> eta;
    viscosity_field(A, F.u, F.v, nullptr, nullptr, F.B, P.n, P.eps_reg, dx, eta, nullptr);
    auto ETA = [&](int i, int j) { return eta[(size_t)clampi(i, 0, ny - 1) * nx + clampi(j, 0, nx - 1)]; };
    auto etaHv = [&](int vi, int vj) {  // vertex (vi,vj): cells (vi-1,vj-1),(vi-1,vj),(vi,vj-1),(vi,vj)
        return 0.25 * (ETA(vi - 1, vj - 1) * A.C(F.H, vi - 1, vj - 1) + ETA(vi - 1, vj) * A.C(F.H, vi - 1, vj) +
                       ETA(vi, vj - 1) * A.C(F.H, vi, vj - 1) + ETA(vi, vj) * A.C(F.H, vi, vj));
    };
    const double cr = P.calving_rate;
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= ny; ++i) {
        for (int j = 0; j <= nx; ++j) {
            if (i < ny && j < nx) {
                double Hc = A.C(F.H, i, j), pc = A.C(F.phi, i, j), dH;
                double rH = Hc / dt;
                if (use_forcing) rH -= A.C(F.f_H, i, j);
                double Hl = A.C(F.H, i, j - 1), Hr = A.C(F.H, i, j + 1), Ht = A.C(F.H, i - 1, j), Hb = A.C(F.H, i + 1, j);
                rH -= flux_x(A.U(F.u, i, j), Hl, Hc, j > 0 && j < nx).res * dxi;
                rH += calving(Hc, pc, A.C(F.phi, i, j - 1), cr, &dH) * dxi;
                rH += flux_x(A.U(F.u, i, j + 1), Hc, Hr, j + 1 > 0 && j + 1 < nx).res * dxi;
                rH += calving(Hc, pc, A.C(F.phi, i, j + 1), cr, &dH) * dxi;
                rH += flux_y(A.V(F.v, i, j), Ht, Hc, i > 0 && i < ny).res * dxi;
                rH += calving(Hc, pc, A.C(F.phi, i - 1, j), cr, &dH) * dxi;
                rH -= flux_y(A.V(F.v, i + 1, j), Hc, Hb, i + 1 > 0 && i + 1 < ny).res * dxi;
                rH += calving(Hc, pc, A.C(F.phi, i + 1, j), cr, &dH) * dxi;
                double mk = use_mask ? A.C(F.mask, i, j) : 0.0;
                r_H[(size_t)i * nx + j] = (1.0 - mk) * rH + mk * (Hc - A.C(F.gamma, i, j));
            }
            if (i < ny) {  // u facet (i,j)
                double ru = use_forcing ? -A.U(F.f_u, i, j) : 0.0;
                if (j < nx)
                    ru += sigma_xx(A.U(F.u, i, j), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i + 1, j),
                                   ETA(i, j) * A.C(F.H, i, j), dxi, j, nx).res * dxi;
                if (j - 1 >= 0)
                    ru -= sigma_xx(A.U(F.u, i, j - 1), A.U(F.u, i, j), A.V(F.v, i, j - 1), A.V(F.v, i + 1, j - 1),
                                   ETA(i, j - 1) * A.C(F.H, i, j - 1), dxi, j - 1, nx).res * dxi;
                if (i > 0 && i < ny && j > 0 && j < nx)
                    ru += sigma_xy(A.U(F.u, i - 1, j), A.U(F.u, i, j), A.V(F.v, i, j - 1), A.V(F.v, i, j),
                                   etaHv(i, j), dxi, i, j, ny, nx).res * dxi;
                if (i + 1 > 0 && i + 1 < ny && j > 0 && j < nx)
                    ru -= sigma_xy(A.U(F.u, i, j), A.U(F.u, i + 1, j), A.V(F.v, i + 1, j - 1), A.V(F.v, i + 1, j),
                                   etaHv(i + 1, j), dxi, i + 1, j, ny, nx).res * dxi;
                double c[4] = {A.V(F.v, i, j - 1), A.V(F.v, i, j), A.V(F.v, i + 1, j - 1), A.V(F.v, i + 1, j)};
                ru += basal_drag(A.U(F.u, i, j), c, A.C(F.phi, i, j - 1), A.C(F.phi, i, j), A.C(F.beta, i, j - 1),
                                 A.C(F.beta, i, j), P).res;
                ru -= driving_x(A.C(F.H, i, j - 1), A.C(F.H, i, j), A.C(F.bed, i, j - 1), A.C(F.bed, i, j),
                                A.C(F.phi, i, j - 1), A.C(F.phi, i, j), dxi, j > 0 && j < nx).res;
                if (j == 0 || j == nx) ru = A.U(F.u, i, j);
                r_u[(size_t)i * (nx + 1) + j] = ru;
            }
            if (j < nx) {  // v facet (i,j)
                double rv = use_forcing ? -A.V(F.f_v, i, j) : 0.0;
                if (i - 1 >= 0)
                    rv += sigma_yy(A.U(F.u, i - 1, j), A.U(F.u, i - 1, j + 1), A.V(F.v, i - 1, j), A.V(F.v, i, j),
                                   ETA(i - 1, j) * A.C(F.H, i - 1, j), dxi, i - 1, ny).res * dxi;
                if (i < ny)
                    rv -= sigma_yy(A.U(F.u, i, j), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i + 1, j),
                                   ETA(i, j) * A.C(F.H, i, j), dxi, i, ny).res * dxi;
                if (i > 0 && i < ny && j > 0 && j < nx)
                    rv -= sigma_xy(A.U(F.u, i - 1, j), A.U(F.u, i, j), A.V(F.v, i, j - 1), A.V(F.v, i, j),
                                   etaHv(i, j), dxi, i, j, ny, nx).res * dxi;
                if (i > 0 && i < ny && j + 1 > 0 && j + 1 < nx)
                    rv += sigma_xy(A.U(F.u, i - 1, j + 1), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i, j + 1),
                                   etaHv(i, j + 1), dxi, i, j + 1, ny, nx).res * dxi;
                double c[4] = {A.U(F.u, i - 1, j), A.U(F.u, i - 1, j + 1), A.U(F.u, i, j), A.U(F.u, i, j + 1)};
                rv += basal_drag(A.V(F.v, i, j), c, A.C(F.phi, i - 1, j), A.C(F.phi, i, j), A.C(F.beta, i - 1, j),
                                 A.C(F.beta, i, j), P).res;
                rv -= driving_y(A.C(F.H, i - 1, j), A.C(F.H, i, j), A.C(F.bed, i - 1, j), A.C(F.bed, i, j),
                                A.C(F.phi, i - 1, j), A.C(F.phi, i, j), dxi, i > 0 && i < ny).res;
                if (i == 0 || i == ny) rv = A.V(F.v, i, j);
                r_v[(size_t)i * nx + j] = rv;
            }
        }
    }
}

// ------------------------------------------------------------------ JVP
// Directional derivative J(x) d of the coupled residual along d = (du, dv, dH)
// with the grounded fraction, the mask, bed, B and beta frozen
// (glide/cuda/residuals.cu:340-660).  Term by term the derivative of
// residual() above, with two reference conventions: rows of Dirichlet facets
// are zero (not d itself, residuals.cu:567,659), and the thickness row is
// (1 - mask) times the derivative of the mass balance, without the mask's
// H - gamma term (residuals.cu:437).
void jvp(const Fields& F, const double* du, const double* dv, const double* dH, double* o_u, double* o_v,
         double* o_H, bool use_mask, const orc_phys& P, double dx, double dt, int ny, int nx) {
    Acc A{ny, nx};
    const double dxi = 1.0 / dx;
    std::vector<double> eta, etd;
    viscosity_field(A, F.u, F.v, du, dv, F.B, P.n, P.eps_reg, dx, eta, &etd);
    auto ETA = [&](int i, int j) { return eta[(size_t)clampi(i, 0, ny - 1) * nx + clampi(j, 0, nx - 1)]; };
    auto ETD = [&](int i, int j) { return etd[(size_t)clampi(i, 0, ny - 1) * nx + clampi(j, 0, nx - 1)]; };
    // eta H of a cell and its derivative
    auto eH = [&](int i, int j) { return ETA(i, j) * A.C(F.H, i, j); };
    auto deH = [&](int i, int j) { return ETD(i, j) * A.C(F.H, i, j) + ETA(i, j) * A.C(dH, i, j); };
    // vertex (vi,vj) average over cells (vi-1,vj-1),(vi-1,vj),(vi,vj-1),(vi,vj)
    auto eHv = [&](int vi, int vj) {
        return 0.25 * (eH(vi - 1, vj - 1) + eH(vi - 1, vj) + eH(vi, vj - 1) + eH(vi, vj));
    };
    auto deHv = [&](int vi, int vj) {
        return 0.25 * (deH(vi - 1, vj - 1) + deH(vi - 1, vj) + deH(vi, vj - 1) + deH(vi, vj));
    };
    const double cr = P.calving_rate;
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= ny; ++i) {
        for (int j = 0; j <= nx; ++j) {
            if (i < ny && j < nx) {
                double Hc = A.C(F.H, i, j), pc = A.C(F.phi, i, j), k;
                double dHc = A.C(dH, i, j);
                double d = dHc / dt;
                Flux f = flux_x(A.U(F.u, i, j), A.C(F.H, i, j - 1), Hc, j > 0 && j < nx);
                d -= (f.d_w * A.U(du, i, j) + f.d_HA * A.C(dH, i, j - 1) + f.d_HB * dHc) * dxi;
                calving(Hc, pc, A.C(F.phi, i, j - 1), cr, &k);
                d += k * dHc * dxi;
                f = flux_x(A.U(F.u, i, j + 1), Hc, A.C(F.H, i, j + 1), j + 1 > 0 && j + 1 < nx);
                d += (f.d_w * A.U(du, i, j + 1) + f.d_HA * dHc + f.d_HB * A.C(dH, i, j + 1)) * dxi;
                calving(Hc, pc, A.C(F.phi, i, j + 1), cr, &k);
                d += k * dHc * dxi;
                f = flux_y(A.V(F.v, i, j), A.C(F.H, i - 1, j), Hc, i > 0 && i < ny);
                d += (f.d_w * A.V(dv, i, j) + f.d_HA * A.C(dH, i - 1, j) + f.d_HB * dHc) * dxi;
                calving(Hc, pc, A.C(F.phi, i - 1, j), cr, &k);
                d += k * dHc * dxi;
                f = flux_y(A.V(F.v, i + 1, j), Hc, A.C(F.H, i + 1, j), i + 1 > 0 && i + 1 < ny);
                d -= (f.d_w * A.V(dv, i + 1, j) + f.d_HA * dHc + f.d_HB * A.C(dH, i + 1, j)) * dxi;
                calving(Hc, pc, A.C(F.phi, i + 1, j), cr, &k);
                d += k * dHc * dxi;
                double mk = use_mask ? A.C(F.mask, i, j) : 0.0;
                o_H[(size_t)i * nx + j] = (1.0 - mk) * d;
            }
            if (i < ny) {  // u facet (i,j)
                double d = 0.0;
                if (j > 0 && j < nx) {
                    Normal s = sigma_xx(A.U(F.u, i, j), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i + 1, j), eH(i, j),
                                        dxi, j, nx);
                    d += (s.d_ul * A.U(du, i, j) + s.d_ur * A.U(du, i, j + 1) + s.d_vt * A.V(dv, i, j) +
                          s.d_vb * A.V(dv, i + 1, j) + s.d_etaH * deH(i, j)) * dxi;
                    s = sigma_xx(A.U(F.u, i, j - 1), A.U(F.u, i, j), A.V(F.v, i, j - 1), A.V(F.v, i + 1, j - 1),
                                 eH(i, j - 1), dxi, j - 1, nx);
                    d -= (s.d_ul * A.U(du, i, j - 1) + s.d_ur * A.U(du, i, j) + s.d_vt * A.V(dv, i, j - 1) +
                          s.d_vb * A.V(dv, i + 1, j - 1) + s.d_etaH * deH(i, j - 1)) * dxi;
                    Shear t = sigma_xy(A.U(F.u, i - 1, j), A.U(F.u, i, j), A.V(F.v, i, j - 1), A.V(F.v, i, j), eHv(i, j),
                                       dxi, i, j, ny, nx);
                    d += (t.d_ut * A.U(du, i - 1, j) + t.d_ub * A.U(du, i, j) + t.d_vl * A.V(dv, i, j - 1) +
                          t.d_vr * A.V(dv, i, j) + t.d_etaH * deHv(i, j)) * dxi;
                    t = sigma_xy(A.U(F.u, i, j), A.U(F.u, i + 1, j), A.V(F.v, i + 1, j - 1), A.V(F.v, i + 1, j),
                                 eHv(i + 1, j), dxi, i + 1, j, ny, nx);
                    d -= (t.d_ut * A.U(du, i, j) + t.d_ub * A.U(du, i + 1, j) + t.d_vl * A.V(dv, i + 1, j - 1) +
                          t.d_vr * A.V(dv, i + 1, j) + t.d_etaH * deHv(i + 1, j)) * dxi;
                    double c[4] = {A.V(F.v, i, j - 1), A.V(F.v, i, j), A.V(F.v, i + 1, j - 1), A.V(F.v, i + 1, j)};
                    double dc[4] = {A.V(dv, i, j - 1), A.V(dv, i, j), A.V(dv, i + 1, j - 1), A.V(dv, i + 1, j)};
                    Drag g = basal_drag(A.U(F.u, i, j), c, A.C(F.phi, i, j - 1), A.C(F.phi, i, j),
                                        A.C(F.beta, i, j - 1), A.C(F.beta, i, j), P);
                    d += g.d_w * A.U(du, i, j) + g.d_c[0] * dc[0] + g.d_c[1] * dc[1] + g.d_c[2] * dc[2] +
                         g.d_c[3] * dc[3];
                    Drive r = driving_x(A.C(F.H, i, j - 1), A.C(F.H, i, j), A.C(F.bed, i, j - 1), A.C(F.bed, i, j),
                                        A.C(F.phi, i, j - 1), A.C(F.phi, i, j), dxi, true);
                    d -= r.d_HA * A.C(dH, i, j - 1) + r.d_HB * A.C(dH, i, j);
                }
                o_u[(size_t)i * (nx + 1) + j] = d;
            }
            if (j < nx) {  // v facet (i,j)
                double d = 0.0;
                if (i > 0 && i < ny) {
                    Normal s = sigma_yy(A.U(F.u, i - 1, j), A.U(F.u, i - 1, j + 1), A.V(F.v, i - 1, j), A.V(F.v, i, j),
                                        eH(i - 1, j), dxi, i - 1, ny);
                    d += (s.d_ul * A.U(du, i - 1, j) + s.d_ur * A.U(du, i - 1, j + 1) + s.d_vt * A.V(dv, i - 1, j) +
                          s.d_vb * A.V(dv, i, j) + s.d_etaH * deH(i - 1, j)) * dxi;
                    s = sigma_yy(A.U(F.u, i, j), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i + 1, j), eH(i, j), dxi,
                                 i, ny);
                    d -= (s.d_ul * A.U(du, i, j) + s.d_ur * A.U(du, i, j + 1) + s.d_vt * A.V(dv, i, j) +
                          s.d_vb * A.V(dv, i + 1, j) + s.d_etaH * deH(i, j)) * dxi;
                    Shear t = sigma_xy(A.U(F.u, i - 1, j), A.U(F.u, i, j), A.V(F.v, i, j - 1), A.V(F.v, i, j), eHv(i, j),
                                       dxi, i, j, ny, nx);
                    d -= (t.d_ut * A.U(du, i - 1, j) + t.d_ub * A.U(du, i, j) + t.d_vl * A.V(dv, i, j - 1) +
                          t.d_vr * A.V(dv, i, j) + t.d_etaH * deHv(i, j)) * dxi;
                    t = sigma_xy(A.U(F.u, i - 1, j + 1), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i, j + 1),
                                 eHv(i, j + 1), dxi, i, j + 1, ny, nx);
                    d += (t.d_ut * A.U(du, i - 1, j + 1) + t.d_ub * A.U(du, i, j + 1) + t.d_vl * A.V(dv, i, j) +
                          t.d_vr * A.V(dv, i, j + 1) + t.d_etaH * deHv(i, j + 1)) * dxi;
                    double c[4] = {A.U(F.u, i - 1, j), A.U(F.u, i - 1, j + 1), A.U(F.u, i, j), A.U(F.u, i, j + 1)};
                    double dc[4] = {A.U(du, i - 1, j), A.U(du, i - 1, j + 1), A.U(du, i, j), A.U(du, i, j + 1)};
                    Drag g = basal_drag(A.V(F.v, i, j), c, A.C(F.phi, i - 1, j), A.C(F.phi, i, j),
                                        A.C(F.beta, i - 1, j), A.C(F.beta, i, j), P);
                    d += g.d_w * A.V(dv, i, j) + g.d_c[0] * dc[0] + g.d_c[1] * dc[1] + g.d_c[2] * dc[2] +
                         g.d_c[3] * dc[3];
                    Drive r = driving_y(A.C(F.H, i - 1, j), A.C(F.H, i, j), A.C(F.bed, i - 1, j), A.C(F.bed, i, j),
                                        A.C(F.phi, i - 1, j), A.C(F.phi, i, j), dxi, true);
                    d -= r.d_HA * A.C(dH, i - 1, j) + r.d_HB * A.C(dH, i, j);
                }
                o_v[(size_t)i * nx + j] = d;
            }
        }
    }
}

// ------------------------------------------------------------------ local 5x5
// Local patch system of cell (i,j) for the unknowns
// x = (u_l, u_r, v_t, v_b, H_c) with viscosity frozen at `eta`
// (glide/cuda/vanka.cu:90-475).  Neighbour values come from the arrays.
void build_local(double J[25], double r[5], const double x[5], const Fields& F, const std::vector<double>& eta,
                 const Acc& A, const orc_phys& P, double calving_rate, double dx, double dt, int i, int j) {
    const int ny = A.ny, nx = A.nx;
    const double dxi = 1.0 / dx;
    for (int k = 0; k < 25; ++k) J[k] = 0;
    for (int k = 0; k < 5; ++k) r[k] = 0;
    const double ul = x[0], ur = x[1], vt = x[2], vb = x[3], Hc = x[4];
    auto ETA = [&](int a, int b) { return eta[(size_t)clampi(a, 0, ny - 1) * nx + clampi(b, 0, nx - 1)]; };
    const double pc = A.C(F.phi, i, j), pl = A.C(F.phi, i, j - 1), pr = A.C(F.phi, i, j + 1);
    const double pt = A.C(F.phi, i - 1, j), pb = A.C(F.phi, i + 1, j);
    const double Hl = A.C(F.H, i, j - 1), Hr = A.C(F.H, i, j + 1), Ht = A.C(F.H, i - 1, j), Hb = A.C(F.H, i + 1, j);
    double dH;
    // mass conservation row
    {
        J[24] = 1.0 / dt;
        r[4] += Hc / dt;
        Flux fl = flux_x(ul, Hl, Hc, j > 0 && j < nx);
        J[20] -= fl.d_w * dxi; J[24] -= fl.d_HB * dxi; r[4] -= fl.res * dxi;
        double c = calving(Hc, pc, pl, calving_rate, &dH); J[24] += dH * dxi; r[4] += c * dxi;
        Flux fr = flux_x(ur, Hc, Hr, j + 1 > 0 && j + 1 < nx);
        J[21] += fr.d_w * dxi; J[24] += fr.d_HA * dxi; r[4] += fr.res * dxi;
        c = calving(Hc, pc, pr, calving_rate, &dH); J[24] += dH * dxi; r[4] += c * dxi;
        Flux ft = flux_y(vt, Ht, Hc, i > 0 && i < ny);
        J[22] += ft.d_w * dxi; J[24] += ft.d_HB * dxi; r[4] += ft.res * dxi;
        c = calving(Hc, pc, pt, calving_rate, &dH); J[24] += dH * dxi; r[4] += c * dxi;
        Flux fb = flux_y(vb, Hc, Hb, i + 1 > 0 && i + 1 < ny);
        J[23] -= fb.d_w * dxi; J[24] -= fb.d_HA * dxi; r[4] -= fb.res * dxi;
        c = calving(Hc, pc, pb, calving_rate, &dH); J[24] += dH * dxi; r[4] += c * dxi;
    }
    // centre-cell normal stresses feed all four momentum rows
    {
        double ec = ETA(i, j);
        Normal sx = sigma_xx(ul, ur, vt, vb, ec * Hc, dxi, j, nx);
        r[0] += sx.res * dxi;
        J[0] += sx.d_ul * dxi; J[1] += sx.d_ur * dxi; J[2] += sx.d_vt * dxi; J[3] += sx.d_vb * dxi;
        J[4] += sx.d_etaH * ec * dxi;
        r[1] -= sx.res * dxi;
        J[5] -= sx.d_ul * dxi; J[6] -= sx.d_ur * dxi; J[7] -= sx.d_vt * dxi; J[8] -= sx.d_vb * dxi;
        J[9] -= sx.d_etaH * ec * dxi;
        Normal sy = sigma_yy(ul, ur, vt, vb, ec * Hc, dxi, i, ny);
        r[2] -= sy.res * dxi;
        J[10] -= sy.d_ul * dxi; J[11] -= sy.d_ur * dxi; J[12] -= sy.d_vt * dxi; J[13] -= sy.d_vb * dxi;
        J[14] -= sy.d_etaH * ec * dxi;
        r[3] += sy.res * dxi;
        J[15] += sy.d_ul * dxi; J[16] += sy.d_ur * dxi; J[17] += sy.d_vt * dxi; J[18] += sy.d_vb * dxi;
        J[19] += sy.d_etaH * ec * dxi;
    }
    {  // left cell sigma_xx -> u_l row
        Normal s = sigma_xx(A.U(F.u, i, j - 1), ul, A.V(F.v, i, j - 1), A.V(F.v, i + 1, j - 1), ETA(i, j - 1) * Hl, dxi,
                            j - 1, nx);
        r[0] -= s.res * dxi; J[0] -= s.d_ur * dxi;
    }
    {  // right cell sigma_xx -> u_r row
        Normal s = sigma_xx(ur, A.U(F.u, i, j + 2), A.V(F.v, i, j + 1), A.V(F.v, i + 1, j + 1), ETA(i, j + 1) * Hr, dxi,
                            j + 1, nx);
        r[1] += s.res * dxi; J[6] += s.d_ul * dxi;
    }
    {  // top cell sigma_yy -> v_t row
        Normal s = sigma_yy(A.U(F.u, i - 1, j), A.U(F.u, i - 1, j + 1), A.V(F.v, i - 1, j), vt, ETA(i - 1, j) * Ht, dxi,
                            i - 1, ny);
        r[2] += s.res * dxi; J[12] += s.d_vb * dxi;
    }
    {  // bottom cell sigma_yy -> v_b row
        Normal s = sigma_yy(A.U(F.u, i + 1, j), A.U(F.u, i + 1, j + 1), vb, A.V(F.v, i + 2, j), ETA(i + 1, j) * Hb, dxi,
                            i + 1, ny);
        r[3] -= s.res * dxi; J[18] -= s.d_vt * dxi;
    }
    {  // top-left vertex
        double etaHv = 0.25 * (ETA(i - 1, j - 1) * A.C(F.H, i - 1, j - 1) + ETA(i - 1, j) * Ht + ETA(i, j - 1) * Hl +
                               ETA(i, j) * Hc);
        double dHc = 0.25 * ETA(i, j);
        Shear s = sigma_xy(A.U(F.u, i - 1, j), ul, A.V(F.v, i, j - 1), vt, etaHv, dxi, i, j, ny, nx);
        r[0] += s.res * dxi; J[0] += s.d_ub * dxi; J[4] += s.d_etaH * dHc * dxi;
        r[2] -= s.res * dxi; J[12] -= s.d_vr * dxi; J[14] -= s.d_etaH * dHc * dxi;
    }
    {  // top-right vertex
        double etaHv = 0.25 * (ETA(i - 1, j) * Ht + ETA(i - 1, j + 1) * A.C(F.H, i - 1, j + 1) + ETA(i, j) * Hc +
                               ETA(i, j + 1) * Hr);
        double dHc = 0.25 * ETA(i, j);
        Shear s = sigma_xy(A.U(F.u, i - 1, j + 1), ur, vt, A.V(F.v, i, j + 1), etaHv, dxi, i, j + 1, ny, nx);
        r[1] += s.res * dxi; J[6] += s.d_ub * dxi; J[9] += s.d_etaH * dHc * dxi;
        r[2] += s.res * dxi; J[12] += s.d_vl * dxi; J[14] += s.d_etaH * dHc * dxi;
    }
    {  // bottom-left vertex
        double etaHv = 0.25 * (ETA(i, j - 1) * Hl + ETA(i, j) * Hc + ETA(i + 1, j - 1) * A.C(F.H, i + 1, j - 1) +
                               ETA(i + 1, j) * Hb);
        double dHc = 0.25 * ETA(i, j);
        Shear s = sigma_xy(ul, A.U(F.u, i + 1, j), A.V(F.v, i + 1, j - 1), vb, etaHv, dxi, i + 1, j, ny, nx);
        r[0] -= s.res * dxi; J[0] -= s.d_ut * dxi; J[4] -= s.d_etaH * dHc * dxi;
        r[3] -= s.res * dxi; J[18] -= s.d_vr * dxi; J[19] -= s.d_etaH * dHc * dxi;
    }
    {  // bottom-right vertex
        double etaHv = 0.25 * (ETA(i, j) * Hc + ETA(i, j + 1) * Hr + ETA(i + 1, j) * Hb +
                               ETA(i + 1, j + 1) * A.C(F.H, i + 1, j + 1));
        double dHc = 0.25 * ETA(i, j);
        Shear s = sigma_xy(ur, A.U(F.u, i + 1, j + 1), vb, A.V(F.v, i + 1, j + 1), etaHv, dxi, i + 1, j + 1, ny, nx);
        r[1] -= s.res * dxi; J[6] -= s.d_ut * dxi; J[9] -= s.d_etaH * dHc * dxi;
        r[3] += s.res * dxi; J[18] += s.d_vl * dxi; J[19] += s.d_etaH * dHc * dxi;
    }
    const double bc = A.C(F.beta, i, j);
    {  // basal drag, left u facet
        double c[4] = {A.V(F.v, i, j - 1), vt, A.V(F.v, i + 1, j - 1), vb};
        Drag t = basal_drag(ul, c, pl, pc, A.C(F.beta, i, j - 1), bc, P);
        r[0] += t.res; J[0] += t.d_w; J[2] += t.d_c[1]; J[3] += t.d_c[3];
    }
    {  // right u facet
        double c[4] = {vt, A.V(F.v, i, j + 1), vb, A.V(F.v, i + 1, j + 1)};
        Drag t = basal_drag(ur, c, pc, pr, bc, A.C(F.beta, i, j + 1), P);
        r[1] += t.res; J[6] += t.d_w; J[7] += t.d_c[0]; J[8] += t.d_c[2];
    }
    {  // top v facet
        double c[4] = {A.U(F.u, i - 1, j), A.U(F.u, i - 1, j + 1), ul, ur};
        Drag t = basal_drag(vt, c, pt, pc, A.C(F.beta, i - 1, j), bc, P);
        r[2] += t.res; J[12] += t.d_w; J[10] += t.d_c[2]; J[11] += t.d_c[3];
    }
    {  // bottom v facet
        double c[4] = {ul, ur, A.U(F.u, i + 1, j), A.U(F.u, i + 1, j + 1)};
        Drag t = basal_drag(vb, c, pc, pb, bc, A.C(F.beta, i + 1, j), P);
        r[3] += t.res; J[18] += t.d_w; J[15] += t.d_c[0]; J[16] += t.d_c[1];
    }
    const double bedc = A.C(F.bed, i, j);
    {
        Drive d = driving_x(Hl, Hc, A.C(F.bed, i, j - 1), bedc, pl, pc, dxi, j > 0 && j < nx);
        r[0] -= d.res; J[4] -= d.d_HB;
    }
    {
        Drive d = driving_x(Hc, Hr, bedc, A.C(F.bed, i, j + 1), pc, pr, dxi, j + 1 > 0 && j + 1 < nx);
        r[1] -= d.res; J[9] -= d.d_HA;
    }
    {
        Drive d = driving_y(Ht, Hc, A.C(F.bed, i - 1, j), bedc, pt, pc, dxi, i > 0 && i < ny);
        r[2] -= d.res; J[14] -= d.d_HB;
    }
    {
        Drive d = driving_y(Hc, Hb, bedc, A.C(F.bed, i + 1, j), pc, pb, dxi, i + 1 > 0 && i + 1 < ny);
        r[3] -= d.res; J[19] -= d.d_HA;
    }
}

// No-pivot Doolittle LU solve of a 5x5 system (glide/cuda/vanka.cu:5-48).
void lu5(const double* Ain, const double* b, double* x) {
    double M[5][5];
    for (int a = 0; a < 5; ++a)
        for (int c = 0; c < 5; ++c) M[a][c] = Ain[a * 5 + c];
    for (int k = 0; k < 5; ++k) {
        double inv = 1.0 / M[k][k];
        for (int a = k + 1; a < 5; ++a) {
            M[a][k] *= inv;
            for (int c = k + 1; c < 5; ++c) M[a][c] -= M[a][k] * M[k][c];
        }
    }
    double y[5];
    for (int a = 0; a < 5; ++a) {
        double s = b[a];
        for (int c = 0; c < a; ++c) s -= M[a][c] * y[c];
        y[a] = s;
    }
    for (int a = 4; a >= 0; --a) {
        double s = y[a];
        for (int c = a + 1; c < 5; ++c) s -= M[a][c] * x[c];
        x[a] = s / M[a][a];
    }
}

inline void dirichlet_row(double J[25], int row) {
    for (int k = 0; k < 5; ++k) J[row * 5 + k] = 0.0;
    for (int k = 0; k < 5; ++k) J[k * 5 + row] = 0.0;
    J[row * 5 + row] = 1.0;
}

// One additive Vanka smoothing pass (glide/cuda/vanka.cu:477-660 plus the
// damped update of glide/operators.py:243-251).  Every cell solves its patch
// against the state at the start of the pass.
void vanka_pass(double* u, double* v, double* H, double* mask, const Fields& F0, const orc_phys& P,
                double calving_rate, double dx, double dt, int ny, int nx, int newton_steps, double ssa_damping,
                double mc_damping, double omega, std::vector<double>* dbg_newton_iters) {
    Acc A{ny, nx};
    std::vector<double> eta;
    viscosity_field(A, F0.u, F0.v, nullptr, nullptr, F0.B, P.n, P.eps_reg, dx, eta, nullptr);
    const size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx, nh = (size_t)ny * nx;
    std::vector<double> du(nu, 0.0), dvv(nv, 0.0), dH(nh, 0.0), dl(nh), dr(nh), dt_(nh), db(nh), newmask(nh);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < ny; ++i) {
        for (int j = 0; j < nx; ++j) {
            const double thk = A.C(F0.gamma, i, j);
            double x[5] = {A.U(F0.u, i, j), A.U(F0.u, i, j + 1), A.V(F0.v, i, j), A.V(F0.v, i + 1, j), A.C(F0.H, i, j)};
            double cmp[5] = {0, 0, 0, 0, 0};
            double masked = A.C(mask, i, j);
            double rnorm = 1.0;
            const double tol = 1e-6;
            int k = 0;
            while (k < newton_steps && rnorm > tol) {
                double J[25], r[5];
                build_local(J, r, x, F0, eta, A, P, calving_rate, dx, dt, i, j);
                r[0] -= A.U(F0.f_u, i, j);
                r[1] -= A.U(F0.f_u, i, j + 1);
                r[2] -= A.V(F0.f_v, i, j);
                r[3] -= A.V(F0.f_v, i + 1, j);
                r[4] -= A.C(F0.f_H, i, j);
                J[0] -= ssa_damping; J[6] -= ssa_damping; J[12] -= ssa_damping; J[18] -= ssa_damping;
                J[24] += mc_damping;
                if (j == 0) { dirichlet_row(J, 0); r[0] = x[0]; }
                if (j == nx - 1) { dirichlet_row(J, 1); r[1] = x[1]; }
                if (i == 0) { dirichlet_row(J, 2); r[2] = x[2]; }
                if (i == ny - 1) { dirichlet_row(J, 3); r[3] = x[3]; }
                if ((x[4] - dt * r[4]) <= thk) {
                    masked = 1.0;
                    dirichlet_row(J, 4);
                    r[4] = x[4] - thk;
                } else {
                    masked = 0.0;
                }
                double d[5];
                lu5(J, r, d);
                rnorm = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3] + r[4] * r[4];
                const double relax = 0.5;  // fixed in the reference kernel (vanka.cu:615)
                for (int a = 0; a < 5; ++a) {  // compensated update (vanka.cu:617-640)
                    double y = -relax * d[a] - cmp[a];
                    double t = x[a] + y;
                    cmp[a] = (t - x[a]) - y;
                    x[a] = t;
                }
                x[4] = std::fmax(x[4], thk);
                ++k;
            }
            size_t c = (size_t)i * nx + j;
            dl[c] = 0.5 * (x[0] - A.U(F0.u, i, j));
            dr[c] = 0.5 * (x[1] - A.U(F0.u, i, j + 1));
            dt_[c] = 0.5 * (x[2] - A.V(F0.v, i, j));
            db[c] = 0.5 * (x[3] - A.V(F0.v, i + 1, j));
            dH[c] = x[4] - A.C(F0.H, i, j);
            newmask[c] = masked;
            if (dbg_newton_iters) (*dbg_newton_iters)[c] = k;
        }
    }
    // two contributions per interior facet: a sum of two terms is order-independent
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            size_t c = (size_t)i * nx + j;
            du[(size_t)i * (nx + 1) + j] += dl[c];
            du[(size_t)i * (nx + 1) + j + 1] += dr[c];
            dvv[(size_t)i * nx + j] += dt_[c];
            dvv[(size_t)(i + 1) * nx + j] += db[c];
        }
    for (size_t k = 0; k < nu; ++k) u[k] += omega * du[k];
    for (size_t k = 0; k < nv; ++k) v[k] += omega * dvv[k];
    for (size_t k = 0; k < nh; ++k) H[k] += omega * dH[k];
    std::memcpy(mask, newmask.data(), nh * sizeof(double));
}

// Multicolour Vanka pass (no counterpart in the reference; the smoother of
// the mixed-precision solver): two phases, red cells ((i + j) even) then
// black cells.  In a phase every cell of the colour runs a local Newton
// iteration on the reference's 5x5 patch system (build_local, viscosity frozen
// at the state the pass starts from) against the state at the start of the
// phase, with relaxation `relax` and the Vanka dampings, and writes its five
// unknowns directly: each facet belongs to exactly one cell of a colour, so
// there is no averaging and no damping of the update.  Cells of one colour
// read each other's old values (a phase is a block-Jacobi step over
// non-overlapping patches), which keeps the result independent of the order.
// Local stopping: sum r^2 < tol, tested before the update as in the additive
// pass.  Fixed point: the same as the additive pass (r = 0 with the same
// active set).
void mc_pass(double* u, double* v, double* H, double* mask, const Fields& F0, const orc_phys& P, double calving_rate,
             double dx, double dt, int ny, int nx, int newton_steps, double relax, double tol, double ssa_damping,
             double mc_damping, std::vector<double>* dbg_newton_iters) {
    Acc A{ny, nx};
    const size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx, nh = (size_t)ny * nx;
    std::vector<double> eta;
    viscosity_field(A, u, v, nullptr, nullptr, F0.B, P.n, P.eps_reg, dx, eta, nullptr);
    std::vector<double> su(nu), sv(nv), sH(nh);
    for (int colour = 0; colour < 2; ++colour) {
        std::memcpy(su.data(), u, nu * sizeof(double));
        std::memcpy(sv.data(), v, nv * sizeof(double));
        std::memcpy(sH.data(), H, nh * sizeof(double));
        Fields S = F0;
        S.u = su.data(); S.v = sv.data(); S.H = sH.data();
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < ny; ++i) {
            for (int j = (i + colour) & 1; j < nx; j += 2) {
                const double thk = A.C(S.gamma, i, j);
                double x[5] = {A.U(S.u, i, j), A.U(S.u, i, j + 1), A.V(S.v, i, j), A.V(S.v, i + 1, j), A.C(S.H, i, j)};
                double masked = A.C(mask, i, j);
                double rnorm = 1.0;
                int k = 0;
                while (k < newton_steps && rnorm > tol) {
                    double J[25], r[5];
                    build_local(J, r, x, S, eta, A, P, calving_rate, dx, dt, i, j);
                    r[0] -= A.U(S.f_u, i, j);
                    r[1] -= A.U(S.f_u, i, j + 1);
                    r[2] -= A.V(S.f_v, i, j);
                    r[3] -= A.V(S.f_v, i + 1, j);
                    r[4] -= A.C(S.f_H, i, j);
                    J[0] -= ssa_damping; J[6] -= ssa_damping; J[12] -= ssa_damping; J[18] -= ssa_damping;
                    J[24] += mc_damping;
                    if (j == 0) { dirichlet_row(J, 0); r[0] = x[0]; }
                    if (j == nx - 1) { dirichlet_row(J, 1); r[1] = x[1]; }
                    if (i == 0) { dirichlet_row(J, 2); r[2] = x[2]; }
                    if (i == ny - 1) { dirichlet_row(J, 3); r[3] = x[3]; }
                    if ((x[4] - dt * r[4]) <= thk) {
                        masked = 1.0;
                        dirichlet_row(J, 4);
                        r[4] = x[4] - thk;
                    } else {
                        masked = 0.0;
                    }
                    double d[5];
                    lu5(J, r, d);
                    rnorm = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3] + r[4] * r[4];
                    for (int a = 0; a < 5; ++a) x[a] -= relax * d[a];
                    x[4] = std::fmax(x[4], thk);
                    ++k;
                }
                u[(size_t)i * (nx + 1) + j] = x[0];
                u[(size_t)i * (nx + 1) + j + 1] = x[1];
                v[(size_t)i * nx + j] = x[2];
                v[(size_t)(i + 1) * nx + j] = x[3];
                H[(size_t)i * nx + j] = x[4];
                mask[(size_t)i * nx + j] = masked;
                if (dbg_newton_iters) (*dbg_newton_iters)[(size_t)i * nx + j] = k;
            }
        }
    }
}

// ------------------------------------------------------------------ transpose operator
// Jᵀλ (minus f when use_forcing), scatter form of glide/cuda/residuals.cu:666-1163.
void vjp(const Fields& F, const double* lu, const double* lv, const double* lH, double* out_u, double* out_v,
         double* out_H, bool use_forcing, const orc_phys& P, double calving_rate, double dx, double dt, int ny,
         int nx) {
    Acc A{ny, nx};
    const double dxi = 1.0 / dx;
    std::vector<double> eta, etad;
    viscosity_field(A, F.u, F.v, lu, lv, F.B, P.n, P.eps_reg, dx, eta, &etad);
    const size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx, nh = (size_t)ny * nx;
    std::vector<double> au(nu, 0.0), av(nv, 0.0), aH(nh, 0.0);
    auto addU = [&](int i, int j, double x) { if (i >= 0 && i < ny && j >= 0 && j <= nx) au[(size_t)i * (nx + 1) + j] += x; };
    auto addV = [&](int i, int j, double x) { if (i >= 0 && i <= ny && j >= 0 && j < nx) av[(size_t)i * nx + j] += x; };
    auto addH = [&](int i, int j, double x) { if (i >= 0 && i < ny && j >= 0 && j < nx) aH[(size_t)i * nx + j] += x; };
    auto ETA = [&](int a, int b) { return eta[(size_t)clampi(a, 0, ny - 1) * nx + clampi(b, 0, nx - 1)]; };
    auto ETAD = [&](int a, int b) { return etad[(size_t)clampi(a, 0, ny - 1) * nx + clampi(b, 0, nx - 1)]; };
    auto Hc_ = [&](int a, int b) { return A.C(F.H, a, b); };
    // Rows scatter into rows i-1..i+1 only: rows of one colour (i mod 3) never
    // touch the same targets, so each colour runs in parallel and the
    // summation order is fixed.
    for (int colour = 0; colour < 3; ++colour)
#pragma omp parallel for schedule(static)
    for (int i = colour; i <= ny; i += 3) {
        for (int j = 0; j <= nx; ++j) {
            if (i < ny && j < nx) {
                double Hc = Hc_(i, j), pc = A.C(F.phi, i, j), dH;
                double lHc = A.C(lH, i, j) * (1.0 - A.C(F.mask, i, j));
                addH(i, j, lHc / dt);
                Flux fl = flux_x(A.U(F.u, i, j), Hc_(i, j - 1), Hc, j > 0 && j < nx);
                addH(i, j, -lHc * fl.d_HB * dxi); addH(i, j - 1, -lHc * fl.d_HA * dxi); addU(i, j, -lHc * fl.d_w * dxi);
                calving(Hc, pc, A.C(F.phi, i, j - 1), calving_rate, &dH); addH(i, j, lHc * dH * dxi);
                Flux fr = flux_x(A.U(F.u, i, j + 1), Hc, Hc_(i, j + 1), j + 1 > 0 && j + 1 < nx);
                addH(i, j, lHc * fr.d_HA * dxi); addH(i, j + 1, lHc * fr.d_HB * dxi); addU(i, j + 1, lHc * fr.d_w * dxi);
                calving(Hc, pc, A.C(F.phi, i, j + 1), calving_rate, &dH); addH(i, j, lHc * dH * dxi);
                Flux ft = flux_y(A.V(F.v, i, j), Hc_(i - 1, j), Hc, i > 0 && i < ny);
                addH(i, j, lHc * ft.d_HB * dxi); addH(i - 1, j, lHc * ft.d_HA * dxi); addV(i, j, lHc * ft.d_w * dxi);
                calving(Hc, pc, A.C(F.phi, i - 1, j), calving_rate, &dH); addH(i, j, lHc * dH * dxi);
                Flux fb = flux_y(A.V(F.v, i + 1, j), Hc, Hc_(i + 1, j), i + 1 > 0 && i + 1 < ny);
                addH(i, j, -lHc * fb.d_HA * dxi); addH(i + 1, j, -lHc * fb.d_HB * dxi); addV(i + 1, j, -lHc * fb.d_w * dxi);
                calving(Hc, pc, A.C(F.phi, i + 1, j), calving_rate, &dH); addH(i, j, lHc * dH * dxi);
                if (use_forcing) addH(i, j, -A.C(F.f_H, i, j));
            }
            if (i < ny) {  // rows of the u facet (i,j), applied by symmetry of the velocity block
                const double lul = A.U(lu, i, j);
                if (j < nx) {
                    Normal s = sigma_xx(A.U(F.u, i, j), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i + 1, j),
                                        ETA(i, j) * Hc_(i, j), dxi, j, nx);
                    double lam = s.d_ul * A.U(lu, i, j) + s.d_ur * A.U(lu, i, j + 1) + s.d_vt * A.V(lv, i, j) +
                                 s.d_vb * A.V(lv, i + 1, j) + s.d_etaH * (Hc_(i, j) * ETAD(i, j));
                    addU(i, j, lam * dxi);
                    addH(i, j, lul * s.d_etaH * ETA(i, j) * dxi);
                }
                if (j - 1 >= 0) {
                    Normal s = sigma_xx(A.U(F.u, i, j - 1), A.U(F.u, i, j), A.V(F.v, i, j - 1), A.V(F.v, i + 1, j - 1),
                                        ETA(i, j - 1) * Hc_(i, j - 1), dxi, j - 1, nx);
                    double lam = s.d_ul * A.U(lu, i, j - 1) + s.d_ur * A.U(lu, i, j) + s.d_vt * A.V(lv, i, j - 1) +
                                 s.d_vb * A.V(lv, i + 1, j - 1) + s.d_etaH * (Hc_(i, j - 1) * ETAD(i, j - 1));
                    addU(i, j, -lam * dxi);
                    addH(i, j - 1, -lul * s.d_etaH * ETA(i, j - 1) * dxi);
                }
                if (i > 0 && i < ny && j > 0 && j < nx) {  // vertex (i,j)
                    double eHv = 0.25 * (ETA(i - 1, j - 1) * Hc_(i - 1, j - 1) + ETA(i - 1, j) * Hc_(i - 1, j) +
                                         ETA(i, j - 1) * Hc_(i, j - 1) + ETA(i, j) * Hc_(i, j));
                    double eHvd = 0.25 * (ETAD(i - 1, j - 1) * Hc_(i - 1, j - 1) + ETAD(i - 1, j) * Hc_(i - 1, j) +
                                          ETAD(i, j - 1) * Hc_(i, j - 1) + ETAD(i, j) * Hc_(i, j));
                    Shear s = sigma_xy(A.U(F.u, i - 1, j), A.U(F.u, i, j), A.V(F.v, i, j - 1), A.V(F.v, i, j), eHv, dxi,
                                       i, j, ny, nx);
                    double lam = s.d_ut * A.U(lu, i - 1, j) + s.d_ub * A.U(lu, i, j) + s.d_vl * A.V(lv, i, j - 1) +
                                 s.d_vr * A.V(lv, i, j) + s.d_etaH * eHvd;
                    addU(i, j, lam * dxi);
                    double k = lul * s.d_etaH * 0.25 * dxi;
                    addH(i - 1, j - 1, k * ETA(i - 1, j - 1)); addH(i - 1, j, k * ETA(i - 1, j));
                    addH(i, j - 1, k * ETA(i, j - 1)); addH(i, j, k * ETA(i, j));
                }
                if (i + 1 > 0 && i + 1 < ny && j > 0 && j < nx) {  // vertex (i+1,j)
                    double eHv = 0.25 * (ETA(i, j - 1) * Hc_(i, j - 1) + ETA(i, j) * Hc_(i, j) +
                                         ETA(i + 1, j - 1) * Hc_(i + 1, j - 1) + ETA(i + 1, j) * Hc_(i + 1, j));
                    double eHvd = 0.25 * (ETAD(i, j - 1) * Hc_(i, j - 1) + ETAD(i, j) * Hc_(i, j) +
                                          ETAD(i + 1, j - 1) * Hc_(i + 1, j - 1) + ETAD(i + 1, j) * Hc_(i + 1, j));
                    Shear s = sigma_xy(A.U(F.u, i, j), A.U(F.u, i + 1, j), A.V(F.v, i + 1, j - 1), A.V(F.v, i + 1, j),
                                       eHv, dxi, i + 1, j, ny, nx);
                    double lam = s.d_ut * A.U(lu, i, j) + s.d_ub * A.U(lu, i + 1, j) + s.d_vl * A.V(lv, i + 1, j - 1) +
                                 s.d_vr * A.V(lv, i + 1, j) + s.d_etaH * eHvd;
                    addU(i, j, -lam * dxi);
                    double k = -lul * s.d_etaH * 0.25 * dxi;
                    addH(i, j - 1, k * ETA(i, j - 1)); addH(i, j, k * ETA(i, j));
                    addH(i + 1, j - 1, k * ETA(i + 1, j - 1)); addH(i + 1, j, k * ETA(i + 1, j));
                }
                {
                    double c[4] = {A.V(F.v, i, j - 1), A.V(F.v, i, j), A.V(F.v, i + 1, j - 1), A.V(F.v, i + 1, j)};
                    Drag t = basal_drag(A.U(F.u, i, j), c, A.C(F.phi, i, j - 1), A.C(F.phi, i, j), A.C(F.beta, i, j - 1),
                                        A.C(F.beta, i, j), P);
                    addU(i, j, lul * t.d_w);
                    addV(i, j - 1, lul * t.d_c[0]); addV(i, j, lul * t.d_c[1]);
                    addV(i + 1, j - 1, lul * t.d_c[2]); addV(i + 1, j, lul * t.d_c[3]);
                }
                {
                    Drive d = driving_x(Hc_(i, j - 1), Hc_(i, j), A.C(F.bed, i, j - 1), A.C(F.bed, i, j),
                                        A.C(F.phi, i, j - 1), A.C(F.phi, i, j), dxi, j > 0 && j < nx);
                    addH(i, j - 1, -lul * d.d_HA); addH(i, j, -lul * d.d_HB);
                }
                if (use_forcing) addU(i, j, -A.U(F.f_u, i, j));
            }
            if (j < nx) {  // rows of the v facet (i,j)
                const double lvt = A.V(lv, i, j);
                if (i - 1 >= 0) {
                    Normal s = sigma_yy(A.U(F.u, i - 1, j), A.U(F.u, i - 1, j + 1), A.V(F.v, i - 1, j), A.V(F.v, i, j),
                                        ETA(i - 1, j) * Hc_(i - 1, j), dxi, i - 1, ny);
                    double lam = s.d_ul * A.U(lu, i - 1, j) + s.d_ur * A.U(lu, i - 1, j + 1) + s.d_vt * A.V(lv, i - 1, j) +
                                 s.d_vb * A.V(lv, i, j) + s.d_etaH * (Hc_(i - 1, j) * ETAD(i - 1, j));
                    addV(i, j, lam * dxi);
                    addH(i - 1, j, lvt * s.d_etaH * ETA(i - 1, j) * dxi);
                }
                if (i < ny) {
                    Normal s = sigma_yy(A.U(F.u, i, j), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i + 1, j),
                                        ETA(i, j) * Hc_(i, j), dxi, i, ny);
                    double lam = s.d_ul * A.U(lu, i, j) + s.d_ur * A.U(lu, i, j + 1) + s.d_vt * A.V(lv, i, j) +
                                 s.d_vb * A.V(lv, i + 1, j) + s.d_etaH * (Hc_(i, j) * ETAD(i, j));
                    addV(i, j, -lam * dxi);
                    addH(i, j, -lvt * s.d_etaH * ETA(i, j) * dxi);
                }
                if (i > 0 && i < ny && j > 0 && j < nx) {  // vertex (i,j)
                    double eHv = 0.25 * (ETA(i - 1, j - 1) * Hc_(i - 1, j - 1) + ETA(i - 1, j) * Hc_(i - 1, j) +
                                         ETA(i, j - 1) * Hc_(i, j - 1) + ETA(i, j) * Hc_(i, j));
                    double eHvd = 0.25 * (ETAD(i - 1, j - 1) * Hc_(i - 1, j - 1) + ETAD(i - 1, j) * Hc_(i - 1, j) +
                                          ETAD(i, j - 1) * Hc_(i, j - 1) + ETAD(i, j) * Hc_(i, j));
                    Shear s = sigma_xy(A.U(F.u, i - 1, j), A.U(F.u, i, j), A.V(F.v, i, j - 1), A.V(F.v, i, j), eHv, dxi,
                                       i, j, ny, nx);
                    double lam = s.d_ut * A.U(lu, i - 1, j) + s.d_ub * A.U(lu, i, j) + s.d_vl * A.V(lv, i, j - 1) +
                                 s.d_vr * A.V(lv, i, j) + s.d_etaH * eHvd;
                    addV(i, j, -lam * dxi);
                    double k = -lvt * s.d_etaH * 0.25 * dxi;
                    addH(i - 1, j - 1, k * ETA(i - 1, j - 1)); addH(i - 1, j, k * ETA(i - 1, j));
                    addH(i, j - 1, k * ETA(i, j - 1)); addH(i, j, k * ETA(i, j));
                }
                if (i > 0 && i < ny && j + 1 > 0 && j + 1 < nx) {  // vertex (i,j+1)
                    double eHv = 0.25 * (ETA(i - 1, j) * Hc_(i - 1, j) + ETA(i - 1, j + 1) * Hc_(i - 1, j + 1) +
                                         ETA(i, j) * Hc_(i, j) + ETA(i, j + 1) * Hc_(i, j + 1));
                    double eHvd = 0.25 * (ETAD(i - 1, j) * Hc_(i - 1, j) + ETAD(i - 1, j + 1) * Hc_(i - 1, j + 1) +
                                          ETAD(i, j) * Hc_(i, j) + ETAD(i, j + 1) * Hc_(i, j + 1));
                    Shear s = sigma_xy(A.U(F.u, i - 1, j + 1), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i, j + 1),
                                       eHv, dxi, i, j + 1, ny, nx);
                    double lam = s.d_ut * A.U(lu, i - 1, j + 1) + s.d_ub * A.U(lu, i, j + 1) + s.d_vl * A.V(lv, i, j) +
                                 s.d_vr * A.V(lv, i, j + 1) + s.d_etaH * eHvd;
                    addV(i, j, lam * dxi);
                    double k = lvt * s.d_etaH * 0.25 * dxi;
                    addH(i - 1, j, k * ETA(i - 1, j)); addH(i - 1, j + 1, k * ETA(i - 1, j + 1));
                    addH(i, j, k * ETA(i, j)); addH(i, j + 1, k * ETA(i, j + 1));
                }
                {
                    double c[4] = {A.U(F.u, i - 1, j), A.U(F.u, i - 1, j + 1), A.U(F.u, i, j), A.U(F.u, i, j + 1)};
                    Drag t = basal_drag(A.V(F.v, i, j), c, A.C(F.phi, i - 1, j), A.C(F.phi, i, j), A.C(F.beta, i - 1, j),
                                        A.C(F.beta, i, j), P);
                    addV(i, j, lvt * t.d_w);
                    addU(i - 1, j, lvt * t.d_c[0]); addU(i - 1, j + 1, lvt * t.d_c[1]);
                    addU(i, j, lvt * t.d_c[2]); addU(i, j + 1, lvt * t.d_c[3]);
                }
                {
                    Drive d = driving_y(Hc_(i - 1, j), Hc_(i, j), A.C(F.bed, i - 1, j), A.C(F.bed, i, j),
                                        A.C(F.phi, i - 1, j), A.C(F.phi, i, j), dxi, i > 0 && i < ny);
                    addH(i - 1, j, -lvt * d.d_HA); addH(i, j, -lvt * d.d_HB);
                }
                if (use_forcing) addV(i, j, -A.V(F.f_v, i, j));
            }
        }
    }
    // Dirichlet facets and masked cells receive nothing (residuals.cu:1122-1160).
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j <= nx; ++j) out_u[(size_t)i * (nx + 1) + j] = (j > 0 && j < nx) ? au[(size_t)i * (nx + 1) + j] : 0.0;
    for (int i = 0; i <= ny; ++i)
        for (int j = 0; j < nx; ++j) out_v[(size_t)i * nx + j] = (i > 0 && i < ny) ? av[(size_t)i * nx + j] : 0.0;
    for (size_t k = 0; k < nh; ++k) out_H[k] = F.mask[k] > 0.0 ? 0.0 : aH[k];
}

// Adjoint Vanka pass (glide/cuda/vanka.cu:662-797, glide/operators.py:517-524):
// lambda -= omega * delta, delta from J_locᵀ delta = r_adj.
void vanka_adjoint_pass(double* lu, double* lv, double* lH, const Fields& F, const double* ru, const double* rv,
                        const double* rH, const orc_phys& P, double calving_rate, double dx, double dt, int ny, int nx,
                        double ssa_damping, double mc_damping, double omega) {
    Acc A{ny, nx};
    std::vector<double> eta;
    viscosity_field(A, F.u, F.v, nullptr, nullptr, F.B, P.n, P.eps_reg, dx, eta, nullptr);
    const size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx, nh = (size_t)ny * nx;
    std::vector<double> dl(nh), dr(nh), dtp(nh), db(nh), dH(nh);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            double x[5] = {A.U(F.u, i, j), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i + 1, j), A.C(F.H, i, j)};
            double J[25], rhs[5];
            build_local(J, rhs, x, F, eta, A, P, calving_rate, dx, dt, i, j);
            J[0] -= ssa_damping; J[6] -= ssa_damping; J[12] -= ssa_damping; J[18] -= ssa_damping;
            J[24] += mc_damping;
            rhs[0] = A.U(ru, i, j); rhs[1] = A.U(ru, i, j + 1); rhs[2] = A.V(rv, i, j); rhs[3] = A.V(rv, i + 1, j);
            rhs[4] = A.C(rH, i, j);
            if (j == 0) { dirichlet_row(J, 0); rhs[0] = 0; }
            if (j == nx - 1) { dirichlet_row(J, 1); rhs[1] = 0; }
            if (i == 0) { dirichlet_row(J, 2); rhs[2] = 0; }
            if (i == ny - 1) { dirichlet_row(J, 3); rhs[3] = 0; }
            if (A.C(F.mask, i, j) > 0.5) { dirichlet_row(J, 4); rhs[4] = 0; }
            double JT[25];
            for (int a = 0; a < 5; ++a)
                for (int b = 0; b < 5; ++b) JT[a * 5 + b] = J[b * 5 + a];
            double d[5];
            lu5(JT, rhs, d);
            size_t c = (size_t)i * nx + j;
            dl[c] = 0.5 * d[0]; dr[c] = 0.5 * d[1]; dtp[c] = 0.5 * d[2]; db[c] = 0.5 * d[3]; dH[c] = d[4];
        }
    std::vector<double> du(nu, 0.0), dv(nv, 0.0);
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            size_t c = (size_t)i * nx + j;
            du[(size_t)i * (nx + 1) + j] += dl[c];
            du[(size_t)i * (nx + 1) + j + 1] += dr[c];
            dv[(size_t)i * nx + j] += dtp[c];
            dv[(size_t)(i + 1) * nx + j] += db[c];
        }
    for (size_t k = 0; k < nu; ++k) lu[k] -= omega * du[k];
    for (size_t k = 0; k < nv; ++k) lv[k] -= omega * dv[k];
    for (size_t k = 0; k < nh; ++k) lH[k] -= omega * dH[k];
}

// Multicolour adjoint pass (no counterpart in the reference; the adjoint
// smoother of the mixed-precision solver): per colour, the residual
// Jᵀλ - f at the current multipliers, then every cell of the colour solves its
// transposed, damped and pinned patch system and subtracts omega times the
// solution from its five multipliers (each unknown belongs to one cell of a
// colour, so no averaging).
void mc_adjoint_pass(double* lu, double* lv, double* lH, const Fields& F, const orc_phys& P, double calving_rate,
                     double dx, double dt, int ny, int nx, double ssa_damping, double mc_damping, double omega) {
    Acc A{ny, nx};
    std::vector<double> eta;
    viscosity_field(A, F.u, F.v, nullptr, nullptr, F.B, P.n, P.eps_reg, dx, eta, nullptr);
    const size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx, nh = (size_t)ny * nx;
    std::vector<double> ru(nu), rv(nv), rH(nh);
    for (int colour = 0; colour < 2; ++colour) {
        vjp(F, lu, lv, lH, ru.data(), rv.data(), rH.data(), true, P, calving_rate, dx, dt, ny, nx);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < ny; ++i)
            for (int j = (i + colour) & 1; j < nx; j += 2) {
                double x[5] = {A.U(F.u, i, j), A.U(F.u, i, j + 1), A.V(F.v, i, j), A.V(F.v, i + 1, j),
                               A.C(F.H, i, j)};
                double J[25], rhs[5];
                build_local(J, rhs, x, F, eta, A, P, calving_rate, dx, dt, i, j);
                J[0] -= ssa_damping; J[6] -= ssa_damping; J[12] -= ssa_damping; J[18] -= ssa_damping;
                J[24] += mc_damping;
                rhs[0] = A.U(ru.data(), i, j); rhs[1] = A.U(ru.data(), i, j + 1);
                rhs[2] = A.V(rv.data(), i, j); rhs[3] = A.V(rv.data(), i + 1, j);
                rhs[4] = A.C(rH.data(), i, j);
                if (j == 0) { dirichlet_row(J, 0); rhs[0] = 0; }
                if (j == nx - 1) { dirichlet_row(J, 1); rhs[1] = 0; }
                if (i == 0) { dirichlet_row(J, 2); rhs[2] = 0; }
                if (i == ny - 1) { dirichlet_row(J, 3); rhs[3] = 0; }
                if (A.C(F.mask, i, j) > 0.5) { dirichlet_row(J, 4); rhs[4] = 0; }
                double JT[25];
                for (int a = 0; a < 5; ++a)
                    for (int b = 0; b < 5; ++b) JT[a * 5 + b] = J[b * 5 + a];
                double d[5];
                lu5(JT, rhs, d);
                lu[(size_t)i * (nx + 1) + j] -= omega * d[0];
                lu[(size_t)i * (nx + 1) + j + 1] -= omega * d[1];
                lv[(size_t)i * nx + j] -= omega * d[2];
                lv[(size_t)(i + 1) * nx + j] -= omega * d[3];
                lH[(size_t)i * nx + j] -= omega * d[4];
            }
    }
}

// Parameter gradients (glide/cuda/grad.cu:1-169); the reference's j<nx-1 /
// i<ny-1 guards are kept.
void grad_beta(double* g, const Fields& F, const double* lu, const double* lv, const orc_phys& P, int ny, int nx) {
    Acc A{ny, nx};
    std::fill(g, g + (size_t)ny * nx, 0.0);
    for (int i = 0; i <= ny; ++i)
        for (int j = 0; j <= nx; ++j) {
            if (i < ny) {
                double c[4] = {A.V(F.v, i, j - 1), A.V(F.v, i, j), A.V(F.v, i + 1, j - 1), A.V(F.v, i + 1, j)};
                Drag t = basal_drag(A.U(F.u, i, j), c, A.C(F.phi, i, j - 1), A.C(F.phi, i, j), A.C(F.beta, i, j - 1),
                                    A.C(F.beta, i, j), P);
                double l = A.U(lu, i, j);
                if (j > 0) g[(size_t)i * nx + j - 1] += l * t.d_betaA;
                if (j < nx - 1) g[(size_t)i * nx + j] += l * t.d_betaB;
            }
            if (j < nx) {
                double c[4] = {A.U(F.u, i - 1, j), A.U(F.u, i - 1, j + 1), A.U(F.u, i, j), A.U(F.u, i, j + 1)};
                Drag t = basal_drag(A.V(F.v, i, j), c, A.C(F.phi, i - 1, j), A.C(F.phi, i, j), A.C(F.beta, i - 1, j),
                                    A.C(F.beta, i, j), P);
                double l = A.V(lv, i, j);
                if (i > 0) g[(size_t)(i - 1) * nx + j] += l * t.d_betaA;
                if (i < ny - 1) g[(size_t)i * nx + j] += l * t.d_betaB;
            }
        }
}

void grad_bed(double* g, const Fields& F, const double* lu, const double* lv, double dx, int ny, int nx) {
    Acc A{ny, nx};
    const double dxi = 1.0 / dx;
    std::fill(g, g + (size_t)ny * nx, 0.0);
    for (int i = 0; i <= ny; ++i)
        for (int j = 0; j <= nx; ++j) {
            if (i < ny) {
                Drive d = driving_x(A.C(F.H, i, j - 1), A.C(F.H, i, j), A.C(F.bed, i, j - 1), A.C(F.bed, i, j),
                                    A.C(F.phi, i, j - 1), A.C(F.phi, i, j), dxi, j > 0 && j < nx);
                double l = A.U(lu, i, j);
                if (j > 0) g[(size_t)i * nx + j - 1] += -l * d.d_bedA;
                if (j < nx - 1) g[(size_t)i * nx + j] += -l * d.d_bedB;
            }
            if (j < nx) {
                Drive d = driving_y(A.C(F.H, i - 1, j), A.C(F.H, i, j), A.C(F.bed, i - 1, j), A.C(F.bed, i, j),
                                    A.C(F.phi, i - 1, j), A.C(F.phi, i, j), dxi, i > 0 && i < ny);
                double l = A.V(lv, i, j);
                if (i > 0) g[(size_t)(i - 1) * nx + j] += -l * d.d_bedA;
                if (i < ny - 1) g[(size_t)i * nx + j] += -l * d.d_bedB;
            }
        }
}

// ------------------------------------------------------------------ transfers
// glide/cuda/transfer.cu:1-86,141-228,325-395.
void restrict_vfacet(const double* f, double* c, int nyc, int nxc) {
    for (int J = 0; J < nyc; ++J)
        for (int I = 0; I <= nxc; ++I)
            c[J * (nxc + 1) + I] = 0.5 * (f[2 * J * (2 * nxc + 1) + 2 * I] + f[(2 * J + 1) * (2 * nxc + 1) + 2 * I]);
}
void restrict_hfacet(const double* f, double* c, int nyc, int nxc) {
    for (int J = 0; J <= nyc; ++J)
        for (int I = 0; I < nxc; ++I)
            c[J * nxc + I] = 0.5 * (f[2 * J * (2 * nxc) + 2 * I] + f[2 * J * (2 * nxc) + 2 * I + 1]);
}
void restrict_cell(const double* f, double* c, int nyc, int nxc, bool use_max) {
    const int nxf = 2 * nxc;
    for (int J = 0; J < nyc; ++J)
        for (int I = 0; I < nxc; ++I) {
            double a = f[2 * J * nxf + 2 * I], b = f[2 * J * nxf + 2 * I + 1];
            double cc = f[(2 * J + 1) * nxf + 2 * I], d = f[(2 * J + 1) * nxf + 2 * I + 1];
            c[J * nxc + I] = use_max ? std::fmax(std::fmax(a, b), std::fmax(cc, d)) : 0.25 * (a + b + cc + d);
        }
}
// Bilinear u-facet prolongation onto an (nyf, nxf+1) array from (nyf/2, nxf/2+1).
void prolong_vfacet_bilinear(const double* c, double* f, int nyf, int nxf) {
    const int nyc = nyf / 2, nxc = nxf / 2;
    for (int j = 0; j < nyf; ++j)
        for (int i = 0; i <= nxf; ++i) {
            double Jf = (j - 0.5) * 0.5, If = i * 0.5;
            Jf = std::fmax(0.0, std::fmin(Jf, (double)(nyc - 1)));
            If = std::fmax(0.0, std::fmin(If, (double)nxc));
            int Jl = (int)Jf, Il = (int)If;
            int Jh = std::min(Jl + 1, nyc - 1), Ih = std::min(Il + 1, nxc);
            double ty = Jf - Jl, tx = If - Il;
            const int s = nxc + 1;
            double v00 = c[Jl * s + Il], v01 = c[Jl * s + Ih], v10 = c[Jh * s + Il], v11 = c[Jh * s + Ih];
            f[j * (nxf + 1) + i] = (1.0 - ty) * ((1.0 - tx) * v00 + tx * v01) + ty * ((1.0 - tx) * v10 + tx * v11);
        }
}
// Bilinear v-facet prolongation onto an (nyf+1, nxf) array (the reference's
// unused hfacet kernel, transfer.cu:268-321); used in the corrected mode.
void prolong_hfacet_bilinear(const double* c, double* f, int nyf, int nxf) {
    const int nyc = nyf / 2, nxc = nxf / 2;
    for (int j = 0; j <= nyf; ++j)
        for (int i = 0; i < nxf; ++i) {
            double Jf = j * 0.5, If = (i - 0.5) * 0.5;
            Jf = std::fmax(0.0, std::fmin(Jf, (double)nyc));
            If = std::fmax(0.0, std::fmin(If, (double)(nxc - 1)));
            int Jl = (int)Jf, Il = (int)If;
            int Jh = std::min(Jl + 1, nyc), Ih = std::min(Il + 1, nxc - 1);
            double ty = Jf - Jl, tx = If - Il;
            double v00 = c[Jl * nxc + Il], v01 = c[Jl * nxc + Ih], v10 = c[Jh * nxc + Il], v11 = c[Jh * nxc + Ih];
            f[j * nxf + i] = (1.0 - ty) * ((1.0 - tx) * v00 + tx * v01) + ty * ((1.0 - tx) * v10 + tx * v11);
        }
}
void prolong_cell_injection(const double* c, double* f, int nyf, int nxf) {
    for (int j = 0; j < nyf; ++j)
        for (int i = 0; i < nxf; ++i) f[j * nxf + i] = c[(j / 2) * (nxf / 2) + i / 2];
}
void prolong_cell_bilinear(const double* c, double* f, int nyf, int nxf) {
    const int nyc = nyf / 2, nxc = nxf / 2;
    for (int j = 0; j < nyf; ++j)
        for (int i = 0; i < nxf; ++i) {
            double Jf = (j - 0.5) * 0.5, If = (i - 0.5) * 0.5;
            Jf = std::fmax(0.0, std::fmin(Jf, (double)(nyc - 1)));
            If = std::fmax(0.0, std::fmin(If, (double)(nxc - 1)));
            int Jl = (int)Jf, Il = (int)If;
            int Jh = std::min(Jl + 1, nyc - 1), Ih = std::min(Il + 1, nxc - 1);
            double ty = Jf - Jl, tx = If - Il;
            double v00 = c[Jl * nxc + Il], v01 = c[Jl * nxc + Ih], v10 = c[Jh * nxc + Il], v11 = c[Jh * nxc + Ih];
            f[j * nxf + i] = (1.0 - ty) * ((1.0 - tx) * v00 + tx * v01) + ty * ((1.0 - tx) * v10 + tx * v11);
        }
}

void grounded_field(double* phi, const double* H, const double* bed, double c, double relax, int n) {
    for (int k = 0; k < n; ++k) phi[k] = (1.0 - relax) * grounded_fraction(H[k], bed[k], c) + relax * phi[k];
}

}  // namespace

// ====================================================================== multigrid driver
namespace {

enum FieldId {
    F_U = 0, F_V, F_H, F_HPREV, F_PHI, F_MASK, F_BED, F_B, F_BETA, F_SMB,
    F_LU, F_LV, F_LH,
    F_FU, F_FV, F_FH,          // forward forcing
    F_AFU, F_AFV, F_AFH,       // adjoint forcing
    F_GBETA, F_GBED, F_GHPREV, F_GSMB,
    F_RU, F_RV, F_RH,          // forward residual
    F_ARU, F_ARV, F_ARH,       // adjoint residual
    F_MASKLAST,
    F_COUNT
};

enum Kind { K_U, K_V, K_C };

Kind kind_of(int id) {
    switch (id) {
        case F_U: case F_LU: case F_FU: case F_AFU: case F_RU: case F_ARU: return K_U;
        case F_V: case F_LV: case F_FV: case F_AFV: case F_RV: case F_ARV: return K_V;
        default: return K_C;
    }
}

struct Level {
    int ny, nx;
    double dx;
    orc_phys P;
    std::vector<double> f[F_COUNT];
    // forward operator scratch
    std::vector<double> F_u, F_v, F_H, gamma;
    std::vector<double> w_u, w_v, w_H, y_u, y_v, y_H, z_u, z_v, z_H, chi, cphi;
    // adjoint scratch
    std::vector<double> vjp_u, vjp_v, vjp_H, wl_u, wl_v, wl_H, zl_u, zl_v, zl_H;
    size_t size(Kind k) const {
        return k == K_U ? (size_t)ny * (nx + 1) : k == K_V ? (size_t)(ny + 1) * nx : (size_t)ny * nx;
    }
};

}  // namespace

extern "C" {

typedef struct orc_fas_config {
    int freeze_coarse_calving, freeze_coarse_phi;
    int coarsest_steps, pre_steps, post_steps, finest_steps, maximum_vcycles;
    double relative_tolerance, absolute_tolerance;
    int compat_hfacet;  // 1: v prolongation through the u-facet routine (reference behaviour)
    // 1: keep the multipliers of pinned rows (domain-edge facets, active-set
    // cells) at zero in adjoint V-cycles; the reference lets the coarse-grid
    // correction write into them (0, its behaviour).  Not a reference option.
    int project_pinned;
} orc_fas_config;

typedef struct orc_vanka_config {
    double omega, relax_phi;
    int newton_steps;
    double ssa_damping, mc_damping;
    int multicolour;      // 0: the reference's additive pass; 1: mc_pass / mc_adjoint_pass
    double relaxation;    // local Newton relaxation of mc_pass (the additive pass fixes 0.5)
    double newton_tol;    // local stopping threshold on sum r^2 of mc_pass (the additive pass fixes 1e-6)
} orc_vanka_config;

typedef struct orc_report {
    int vcycles;
    double initial_norm, final_norm;
    double norms[64];
} orc_report;

struct orc_mg {
    std::vector<Level> L;
};

}  // extern "C"

namespace {

Fields fields_of(Level& l, bool adjoint) {
    Fields F;
    F.u = l.f[F_U].data(); F.v = l.f[F_V].data(); F.H = l.f[F_H].data(); F.phi = l.f[F_PHI].data();
    F.mask = l.f[F_MASK].data();
    F.f_u = adjoint ? l.f[F_AFU].data() : l.f[F_FU].data();
    F.f_v = adjoint ? l.f[F_AFV].data() : l.f[F_FV].data();
    F.f_H = adjoint ? l.f[F_AFH].data() : l.f[F_FH].data();
    F.bed = l.f[F_BED].data(); F.B = l.f[F_B].data(); F.beta = l.f[F_BETA].data(); F.gamma = l.gamma.data();
    return F;
}

void restrict_field(int id, Level& fine, Level& coarse) {
    Kind k = kind_of(id);
    const double* f = fine.f[id].data();
    double* c = coarse.f[id].data();
    if (k == K_U) restrict_vfacet(f, c, coarse.ny, coarse.nx);
    else if (k == K_V) restrict_hfacet(f, c, coarse.ny, coarse.nx);
    else restrict_cell(f, c, coarse.ny, coarse.nx, id == F_MASK);
}

struct Solver {
    orc_mg* mg;
    orc_fas_config cfg;
    const orc_vanka_config* vc;  // per level
    double dt;

    // ForwardOperators.compute_residual (glide/operators.py:71-127)
    double residual_level(int l, bool use_mask, bool operator_only, bool freeze_calving, bool freeze_phi,
                          double* norms3) {
        Level& L = mg->L[l];
        if (!freeze_phi)
            grounded_field(L.f[F_PHI].data(), L.f[F_H].data(), L.f[F_BED].data(), L.P.flotation_reg_driving, 0.0,
                           L.ny * L.nx);
        orc_phys P = L.P;
        if (freeze_calving) P.calving_rate = 0.0;
        Fields F = fields_of(L, false);
        double *ou, *ov, *oH;
        if (operator_only) { ou = L.F_u.data(); ov = L.F_v.data(); oH = L.F_H.data(); }
        else { ou = L.f[F_RU].data(); ov = L.f[F_RV].data(); oH = L.f[F_RH].data(); }
        residual(F, ou, ov, oH, !operator_only, use_mask, P, L.dx, dt, L.ny, L.nx);
        if (norms3) {
            double su = 0, sv = 0, sh = 0;
            for (size_t k = 0; k < L.size(K_U); ++k) su += ou[k] * ou[k];
            for (size_t k = 0; k < L.size(K_V); ++k) sv += ov[k] * ov[k];
            for (size_t k = 0; k < L.size(K_C); ++k) sh += oH[k] * oH[k];
            norms3[0] = std::sqrt(su); norms3[1] = std::sqrt(sv); norms3[2] = std::sqrt(sh);
            return std::sqrt(su + sv + sh);
        }
        return 0.0;
    }

    // ForwardOperators.vanka_sweep (glide/operators.py:194-251)
    void sweep(int l, int n_iter, bool freeze_calving, bool freeze_phi) {
        Level& L = mg->L[l];
        const orc_vanka_config& V = vc[l];
        for (int it = 0; it < n_iter; ++it) {
            if (!freeze_phi)
                grounded_field(L.f[F_PHI].data(), L.f[F_H].data(), L.f[F_BED].data(), L.P.flotation_reg_driving,
                               V.relax_phi, L.ny * L.nx);
            Fields F = fields_of(L, false);
            if (V.multicolour)
                mc_pass(L.f[F_U].data(), L.f[F_V].data(), L.f[F_H].data(), L.f[F_MASK].data(), F, L.P,
                        freeze_calving ? 0.0 : L.P.calving_rate, L.dx, dt, L.ny, L.nx, V.newton_steps, V.relaxation,
                        V.newton_tol, V.ssa_damping, V.mc_damping, nullptr);
            else
                vanka_pass(L.f[F_U].data(), L.f[F_V].data(), L.f[F_H].data(), L.f[F_MASK].data(), F, L.P,
                           freeze_calving ? 0.0 : L.P.calving_rate, L.dx, dt, L.ny, L.nx, V.newton_steps,
                           V.ssa_damping, V.mc_damping, V.omega, nullptr);
        }
    }

    void prolong_v(const double* c, double* f, int nyf, int nxf) {
        if (cfg.compat_hfacet) prolong_vfacet_bilinear(c, f, nyf, nxf);  // reference behaviour (multigrid.py:212-235)
        else prolong_hfacet_bilinear(c, f, nyf, nxf);
    }

    // FASCDSolver.vcycle (glide/multigrid.py:536-645)
    void vcycle(int l, bool finest) {
        const bool coarse = !finest;
        Level& L = mg->L[l];
        const size_t nu = L.size(K_U), nv = L.size(K_V), nh = L.size(K_C);
        const bool fcal = coarse && cfg.freeze_coarse_calving, fphi = coarse && cfg.freeze_coarse_phi;
        if (finest) {
            L.w_u = L.f[F_U]; L.w_v = L.f[F_V]; L.w_H = L.f[F_H];
            for (size_t k = 0; k < nh; ++k) L.chi[k] = L.P.thklim - L.f[F_H][k];
        }
        if (l == (int)mg->L.size() - 1) {
            for (size_t k = 0; k < nh; ++k) L.gamma[k] = L.w_H[k] + L.chi[k];
            sweep(l, cfg.coarsest_steps, fcal, fphi);
            std::fill(L.gamma.begin(), L.gamma.end(), L.P.thklim);
            return;
        }
        Level& N = mg->L[l + 1];
        restrict_cell(L.chi.data(), N.chi.data(), N.ny, N.nx, true);
        std::vector<double> negchi(N.chi.size());
        for (size_t k = 0; k < negchi.size(); ++k) negchi[k] = -N.chi[k];
        prolong_cell_injection(negchi.data(), L.cphi.data(), L.ny, L.nx);
        for (size_t k = 0; k < nh; ++k) L.cphi[k] += L.chi[k];
        for (size_t k = 0; k < nh; ++k) L.gamma[k] = L.w_H[k] + L.cphi[k];
        sweep(l, cfg.pre_steps, fcal, fphi);
        std::fill(L.gamma.begin(), L.gamma.end(), L.P.thklim);
        for (size_t k = 0; k < nu; ++k) L.y_u[k] = L.f[F_U][k] - L.w_u[k];
        for (size_t k = 0; k < nv; ++k) L.y_v[k] = L.f[F_V][k] - L.w_v[k];
        for (size_t k = 0; k < nh; ++k) L.y_H[k] = L.f[F_H][k] - L.w_H[k];
        for (int id : {F_U, F_V, F_H, F_HPREV, F_PHI, F_MASK}) restrict_field(id, L, N);
        N.w_u = N.f[F_U]; N.w_v = N.f[F_V]; N.w_H = N.f[F_H];
        residual_level(l, false, false, fcal, true, nullptr);
        for (int id : {F_RU, F_RV, F_RH}) restrict_field(id, L, N);
        residual_level(l + 1, false, true, cfg.freeze_coarse_calving, cfg.freeze_coarse_phi, nullptr);
        for (size_t k = 0; k < N.size(K_U); ++k) N.f[F_FU][k] = N.F_u[k] - N.f[F_RU][k];
        for (size_t k = 0; k < N.size(K_V); ++k) N.f[F_FV][k] = N.F_v[k] - N.f[F_RV][k];
        for (size_t k = 0; k < N.size(K_C); ++k) N.f[F_FH][k] = N.F_H[k] - N.f[F_RH][k];
        vcycle(l + 1, false);
        for (size_t k = 0; k < N.size(K_U); ++k) N.z_u[k] = N.f[F_U][k] - N.w_u[k];
        for (size_t k = 0; k < N.size(K_V); ++k) N.z_v[k] = N.f[F_V][k] - N.w_v[k];
        for (size_t k = 0; k < N.size(K_C); ++k) N.z_H[k] = N.f[F_H][k] - N.w_H[k];
        prolong_vfacet_bilinear(N.z_u.data(), L.z_u.data(), L.ny, L.nx);
        prolong_v(N.z_v.data(), L.z_v.data(), L.ny, L.nx);
        prolong_cell_injection(N.z_H.data(), L.z_H.data(), L.ny, L.nx);
        for (size_t k = 0; k < nu; ++k) { L.z_u[k] += L.y_u[k]; L.f[F_U][k] = L.w_u[k] + L.z_u[k]; }
        for (size_t k = 0; k < nv; ++k) { L.z_v[k] += L.y_v[k]; L.f[F_V][k] = L.w_v[k] + L.z_v[k]; }
        for (size_t k = 0; k < nh; ++k) { L.z_H[k] += L.y_H[k]; L.f[F_H][k] = L.w_H[k] + L.z_H[k]; }
        for (size_t k = 0; k < nh; ++k) L.gamma[k] = L.w_H[k] + L.chi[k];
        sweep(l, cfg.post_steps, fcal, fphi);
        std::fill(L.gamma.begin(), L.gamma.end(), L.P.thklim);
        if (finest) sweep(l, cfg.finest_steps, false, false);
    }

    // FASCDSolver.solve (glide/multigrid.py:502-533)
    void solve(int start, orc_report* rep) {
        Level& L = mg->L[start];
        for (size_t k = 0; k < L.size(K_C); ++k) L.f[F_FH][k] = L.f[F_HPREV][k] / dt + L.f[F_SMB][k];
        double n3[3];
        double r0 = residual_level(start, true, false, false, false, n3);
        double rel = 1.0, absn = r0;
        int it = 0;
        if (rep) { rep->initial_norm = r0; }
        while (rel > cfg.relative_tolerance && absn > cfg.absolute_tolerance && it < cfg.maximum_vcycles) {
            vcycle(start, true);
            absn = residual_level(start, true, false, false, true, n3);
            rel = absn / r0;
            if (rep && it < 64) rep->norms[it] = absn;
            ++it;
        }
        if (rep) { rep->vcycles = it; rep->final_norm = absn; }
    }

    // AdjointOperators.compute_residual / compute_vjp (glide/operators.py:379-472)
    double adj_residual(int l, bool use_forcing, bool freeze_calving, double* out_u, double* out_v, double* out_H,
                        double* norms3) {
        Level& L = mg->L[l];
        Fields F = fields_of(L, true);
        vjp(F, L.f[F_LU].data(), L.f[F_LV].data(), L.f[F_LH].data(), out_u, out_v, out_H, use_forcing, L.P,
            freeze_calving ? 0.0 : L.P.calving_rate, L.dx, dt, L.ny, L.nx);
        if (norms3) {
            double su = 0, sv = 0, sh = 0;
            for (size_t k = 0; k < L.size(K_U); ++k) su += out_u[k] * out_u[k];
            for (size_t k = 0; k < L.size(K_V); ++k) sv += out_v[k] * out_v[k];
            for (size_t k = 0; k < L.size(K_C); ++k) sh += out_H[k] * out_H[k];
            norms3[0] = std::sqrt(su); norms3[1] = std::sqrt(sv); norms3[2] = std::sqrt(sh);
            return std::sqrt(su + sv + sh);
        }
        return 0;
    }

    void adj_sweep(int l, int n_iter) {
        Level& L = mg->L[l];
        const orc_vanka_config& V = vc[l];
        for (int it = 0; it < n_iter; ++it) {
            if (V.multicolour) {
                Fields F = fields_of(L, true);
                mc_adjoint_pass(L.f[F_LU].data(), L.f[F_LV].data(), L.f[F_LH].data(), F, L.P, L.P.calving_rate,
                                L.dx, dt, L.ny, L.nx, V.ssa_damping, V.mc_damping, V.omega);
                continue;
            }
            adj_residual(l, true, false, L.f[F_ARU].data(), L.f[F_ARV].data(), L.f[F_ARH].data(), nullptr);
            Fields F = fields_of(L, true);
            vanka_adjoint_pass(L.f[F_LU].data(), L.f[F_LV].data(), L.f[F_LH].data(), F, L.f[F_ARU].data(),
                               L.f[F_ARV].data(), L.f[F_ARH].data(), L.P, L.P.calving_rate, L.dx, dt, L.ny, L.nx,
                               V.ssa_damping, V.mc_damping, V.omega);
        }
    }

    // Zero the multipliers the transpose operator leaves undetermined.
    void pin_multipliers(Level& L) {
        if (!cfg.project_pinned) return;
        for (int i = 0; i < L.ny; ++i) {
            L.f[F_LU][(size_t)i * (L.nx + 1)] = 0.0;
            L.f[F_LU][(size_t)i * (L.nx + 1) + L.nx] = 0.0;
        }
        for (int j = 0; j < L.nx; ++j) {
            L.f[F_LV][j] = 0.0;
            L.f[F_LV][(size_t)L.ny * L.nx + j] = 0.0;
        }
        for (size_t k = 0; k < L.size(K_C); ++k)
            if (L.f[F_MASK][k] > 0.5) L.f[F_LH][k] = 0.0;
    }

    // FASAdjointSolver.vcycle (glide/multigrid.py:905-973)
    void adj_vcycle(int l, bool finest) {
        Level& L = mg->L[l];
        if (l == (int)mg->L.size() - 1) { adj_sweep(l, cfg.coarsest_steps); return; }
        Level& N = mg->L[l + 1];
        adj_sweep(l, cfg.pre_steps);
        for (int id : {F_U, F_V, F_H, F_HPREV, F_PHI, F_MASK}) restrict_field(id, L, N);
        for (int id : {F_LU, F_LV, F_LH}) restrict_field(id, L, N);
        pin_multipliers(N);
        N.wl_u = N.f[F_LU]; N.wl_v = N.f[F_LV]; N.wl_H = N.f[F_LH];
        adj_residual(l, true, false, L.f[F_ARU].data(), L.f[F_ARV].data(), L.f[F_ARH].data(), nullptr);
        for (int id : {F_ARU, F_ARV, F_ARH}) restrict_field(id, L, N);
        adj_residual(l + 1, false, false, N.vjp_u.data(), N.vjp_v.data(), N.vjp_H.data(), nullptr);
        for (size_t k = 0; k < N.size(K_U); ++k) N.f[F_AFU][k] = N.vjp_u[k] - N.f[F_ARU][k];
        for (size_t k = 0; k < N.size(K_V); ++k) N.f[F_AFV][k] = N.vjp_v[k] - N.f[F_ARV][k];
        for (size_t k = 0; k < N.size(K_C); ++k) N.f[F_AFH][k] = N.vjp_H[k] - N.f[F_ARH][k];
        adj_vcycle(l + 1, false);
        for (size_t k = 0; k < N.size(K_U); ++k) N.zl_u[k] = N.f[F_LU][k] - N.wl_u[k];
        for (size_t k = 0; k < N.size(K_V); ++k) N.zl_v[k] = N.f[F_LV][k] - N.wl_v[k];
        for (size_t k = 0; k < N.size(K_C); ++k) N.zl_H[k] = N.f[F_LH][k] - N.wl_H[k];
        prolong_vfacet_bilinear(N.zl_u.data(), L.zl_u.data(), L.ny, L.nx);
        prolong_v(N.zl_v.data(), L.zl_v.data(), L.ny, L.nx);
        prolong_cell_bilinear(N.zl_H.data(), L.zl_H.data(), L.ny, L.nx);
        for (size_t k = 0; k < L.size(K_U); ++k) L.f[F_LU][k] += L.zl_u[k];
        for (size_t k = 0; k < L.size(K_V); ++k) L.f[F_LV][k] += L.zl_v[k];
        for (size_t k = 0; k < L.size(K_C); ++k) L.f[F_LH][k] += L.zl_H[k];
        pin_multipliers(L);
        adj_sweep(l, cfg.post_steps);
        if (finest) adj_sweep(l, cfg.finest_steps);
    }

    // FASAdjointSolver.solve (glide/multigrid.py:873-903)
    void adj_solve(int start, orc_report* rep) {
        Level& L = mg->L[start];
        double n3[3];
        double r0 = adj_residual(start, true, false, L.f[F_ARU].data(), L.f[F_ARV].data(), L.f[F_ARH].data(), n3);
        double rel = 1.0, absn = r0;
        int it = 0;
        if (rep) rep->initial_norm = r0;
        while (rel > cfg.relative_tolerance && absn > cfg.absolute_tolerance && it < cfg.maximum_vcycles) {
            adj_vcycle(start, true);
            absn = adj_residual(start, true, false, L.f[F_ARU].data(), L.f[F_ARV].data(), L.f[F_ARH].data(), n3);
            rel = absn / r0;
            if (rep && it < 64) rep->norms[it] = absn;
            ++it;
        }
        if (rep) { rep->vcycles = it; rep->final_norm = absn; }
    }
};

}  // namespace

// ====================================================================== C API
extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void orc_grounded(double* phi, const double* H, const double* bed, double c, double relax, int ny, int nx) {
    grounded_field(phi, H, bed, c, relax, ny * nx);
}

void orc_residual(double* r_u, double* r_v, double* r_H, const double* u, const double* v, const double* H,
                  const double* phi, const double* mask, const double* f_u, const double* f_v, const double* f_H,
                  const double* bed, const double* B, const double* beta, const double* gamma, int use_forcing,
                  int use_mask, const orc_phys* P, double dx, double dt, int ny, int nx) {
    Fields F{u, v, H, phi, mask, f_u, f_v, f_H, bed, B, beta, gamma};
    residual(F, r_u, r_v, r_H, use_forcing != 0, use_mask != 0, *P, dx, dt, ny, nx);
}

void orc_jvp(double* o_u, double* o_v, double* o_H, const double* u, const double* v, const double* H,
             const double* du, const double* dv, const double* dH, const double* phi, const double* mask,
             const double* bed, const double* B, const double* beta, int use_mask, const orc_phys* P, double dx,
             double dt, int ny, int nx) {
    Fields F{u, v, H, phi, mask, nullptr, nullptr, nullptr, bed, B, beta, nullptr};
    jvp(F, du, dv, dH, o_u, o_v, o_H, use_mask != 0, *P, dx, dt, ny, nx);
}

// SSA strain heating (no counterpart in the reference, which leaves
// phi_strain to the caller): Phi = 4 eta eps_e^2 rho_i g / SEC_PER_YR in
// W/m^3 from the depth-averaged velocities (m/a) and B (head units), with
// eta = 0.5 B (eps_e^2 + eps_reg)^((1-n)/(2n)) the momentum solver's cell
// viscosity and eps_e^2 its strain-rate invariant without regularisation.
void orc_strain_heating(double* phi, const double* u, const double* v, const double* B, double n, double eps_reg,
                        double dx, double rho_g, int ny, int nx) {
    Acc A{ny, nx};
    const double sec = 365.25 * 86400.0;
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            double e2 = strain_invariant0(A, u, v, i, j, dx);
            double eta = cell_viscosity(A, u, v, nullptr, nullptr, B, i, j, n, eps_reg, dx).v;
            phi[(size_t)i * nx + j] = 4.0 * eta * e2 * rho_g / sec;
        }
}

void orc_viscosity(double* eta, double* etad, const double* u, const double* v, const double* du, const double* dv,
                   const double* B, double n, double eps_reg, double dx, int ny, int nx) {
    Acc A{ny, nx};
    std::vector<double> e, ed;
    viscosity_field(A, u, v, du, dv, B, n, eps_reg, dx, e, du ? &ed : nullptr);
    std::memcpy(eta, e.data(), e.size() * sizeof(double));
    if (etad && du) std::memcpy(etad, ed.data(), ed.size() * sizeof(double));
}

// Local 5x5 system of every cell at the current state, as the reference's
// vanka_dump kernel (vanka.cu:800-908): J (ny*nx,25) and r (ny*nx,5).
void orc_vanka_dump(double* Jout, double* rout, const double* u, const double* v, const double* H, const double* phi,
                    const double* f_u, const double* f_v, const double* f_H, const double* bed, const double* B,
                    const double* beta, const double* gamma, const orc_phys* P, double dx, double dt, int ny, int nx) {
    Acc A{ny, nx};
    Fields F{u, v, H, phi, nullptr, f_u, f_v, f_H, bed, B, beta, gamma};
    std::vector<double> eta;
    viscosity_field(A, u, v, nullptr, nullptr, B, P->n, P->eps_reg, dx, eta, nullptr);
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            double x[5] = {A.U(u, i, j), A.U(u, i, j + 1), A.V(v, i, j), A.V(v, i + 1, j), A.C(H, i, j)};
            double J[25], r[5];
            build_local(J, r, x, F, eta, A, *P, P->calving_rate, dx, dt, i, j);
            r[0] -= A.U(f_u, i, j); r[1] -= A.U(f_u, i, j + 1); r[2] -= A.V(f_v, i, j); r[3] -= A.V(f_v, i + 1, j);
            r[4] -= A.C(f_H, i, j);
            if (j == 0) { dirichlet_row(J, 0); r[0] = x[0]; }
            if (j == nx - 1) { dirichlet_row(J, 1); r[1] = x[1]; }
            if (i == 0) { dirichlet_row(J, 2); r[2] = x[2]; }
            if (i == ny - 1) { dirichlet_row(J, 3); r[3] = x[3]; }
            if ((x[4] - dt * r[4]) <= A.C(gamma, i, j)) { dirichlet_row(J, 4); r[4] = x[4] - A.C(gamma, i, j); }
            size_t c = (size_t)i * nx + j;
            std::memcpy(Jout + 25 * c, J, sizeof(J));
            std::memcpy(rout + 5 * c, r, sizeof(r));
        }
}

// Local residual/Jacobian of one cell at arbitrary local values x[5], for FD
// checks of the assembled Jacobian entries (no Dirichlet/mask rows, no damping).
void orc_local_system(double* Jout, double* rout, const double* x, const double* u, const double* v, const double* H,
                      const double* phi, const double* bed, const double* B, const double* beta, const orc_phys* P,
                      double dx, double dt, int ny, int nx, int i, int j) {
    Acc A{ny, nx};
    Fields F{u, v, H, phi, nullptr, nullptr, nullptr, nullptr, bed, B, beta, nullptr};
    std::vector<double> eta;
    viscosity_field(A, u, v, nullptr, nullptr, B, P->n, P->eps_reg, dx, eta, nullptr);
    build_local(Jout, rout, x, F, eta, A, *P, P->calving_rate, dx, dt, i, j);
}

void orc_lu5(const double* A, const double* b, double* x) { lu5(A, b, x); }

void orc_vanka_sweep(double* u, double* v, double* H, double* mask, const double* phi, const double* f_u,
                     const double* f_v, const double* f_H, const double* bed, const double* B, const double* beta,
                     const double* gamma, const orc_phys* P, double dx, double dt, int ny, int nx, int newton_steps,
                     double ssa_damping, double mc_damping, double omega, double* newton_iters) {
    const size_t nu = (size_t)ny * (nx + 1), nv = (size_t)(ny + 1) * nx, nh = (size_t)ny * nx;
    std::vector<double> u0(u, u + nu), v0(v, v + nv), H0(H, H + nh);
    Fields F{u0.data(), v0.data(), H0.data(), phi, mask, f_u, f_v, f_H, bed, B, beta, gamma};
    std::vector<double> it(nh);
    vanka_pass(u, v, H, mask, F, *P, P->calving_rate, dx, dt, ny, nx, newton_steps, ssa_damping, mc_damping, omega,
               &it);
    if (newton_iters) std::memcpy(newton_iters, it.data(), nh * sizeof(double));
}

void orc_mc_sweep(double* u, double* v, double* H, double* mask, const double* phi, const double* f_u, const double* f_v,
                  const double* f_H, const double* bed, const double* B, const double* beta, const double* gamma,
                  const orc_phys* P, double dx, double dt, int ny, int nx, int newton_steps, double relax, double tol,
                  double ssa_damping, double mc_damping, double* newton_iters) {
    const size_t nh = (size_t)ny * nx;
    Fields F{nullptr, nullptr, nullptr, phi, mask, f_u, f_v, f_H, bed, B, beta, gamma};
    std::vector<double> it(nh);
    mc_pass(u, v, H, mask, F, *P, P->calving_rate, dx, dt, ny, nx, newton_steps, relax, tol, ssa_damping, mc_damping,
            &it);
    if (newton_iters) std::memcpy(newton_iters, it.data(), nh * sizeof(double));
}

void orc_mc_adjoint_sweep(double* lu, double* lv, double* lH, const double* u, const double* v, const double* H,
                          const double* phi, const double* mask, const double* af_u, const double* af_v,
                          const double* af_H, const double* bed, const double* B, const double* beta,
                          const orc_phys* P, double dx, double dt, int ny, int nx, double ssa_damping,
                          double mc_damping, double omega) {
    Fields F{u, v, H, phi, mask, af_u, af_v, af_H, bed, B, beta, nullptr};
    mc_adjoint_pass(lu, lv, lH, F, *P, P->calving_rate, dx, dt, ny, nx, ssa_damping, mc_damping, omega);
}

void orc_vjp(double* out_u, double* out_v, double* out_H, const double* u, const double* v, const double* H,
             const double* lu, const double* lv, const double* lH, const double* phi, const double* mask,
             const double* f_u, const double* f_v, const double* f_H, const double* bed, const double* B,
             const double* beta, int use_forcing, const orc_phys* P, double dx, double dt, int ny, int nx) {
    Fields F{u, v, H, phi, mask, f_u, f_v, f_H, bed, B, beta, nullptr};
    vjp(F, lu, lv, lH, out_u, out_v, out_H, use_forcing != 0, *P, P->calving_rate, dx, dt, ny, nx);
}

void orc_vanka_adjoint_sweep(double* lu, double* lv, double* lH, const double* ru, const double* rv, const double* rH,
                             const double* u, const double* v, const double* H, const double* phi, const double* mask,
                             const double* bed, const double* B, const double* beta, const orc_phys* P, double dx,
                             double dt, int ny, int nx, double ssa_damping, double mc_damping, double omega) {
    Fields F{u, v, H, phi, mask, nullptr, nullptr, nullptr, bed, B, beta, nullptr};
    vanka_adjoint_pass(lu, lv, lH, F, ru, rv, rH, *P, P->calving_rate, dx, dt, ny, nx, ssa_damping, mc_damping, omega);
}

void orc_grad_beta(double* g, const double* u, const double* v, const double* H, const double* phi, const double* beta,
                   const double* lu, const double* lv, const orc_phys* P, int ny, int nx) {
    Fields F{u, v, H, phi, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, beta, nullptr};
    grad_beta(g, F, lu, lv, *P, ny, nx);
}

void orc_grad_bed(double* g, const double* H, const double* phi, const double* bed, const double* lu,
                  const double* lv, double dx, int ny, int nx) {
    Fields F{nullptr, nullptr, H, phi, nullptr, nullptr, nullptr, nullptr, bed, nullptr, nullptr, nullptr};
    grad_bed(g, F, lu, lv, dx, ny, nx);
}

void orc_restrict_vfacet(const double* f, double* c, int nyc, int nxc) { restrict_vfacet(f, c, nyc, nxc); }
void orc_restrict_hfacet(const double* f, double* c, int nyc, int nxc) { restrict_hfacet(f, c, nyc, nxc); }
void orc_restrict_cell(const double* f, double* c, int nyc, int nxc, int use_max) {
    restrict_cell(f, c, nyc, nxc, use_max != 0);
}
void orc_prolong_vfacet_bilinear(const double* c, double* f, int nyf, int nxf) { prolong_vfacet_bilinear(c, f, nyf, nxf); }
void orc_prolong_hfacet_bilinear(const double* c, double* f, int nyf, int nxf) { prolong_hfacet_bilinear(c, f, nyf, nxf); }
void orc_prolong_cell_injection(const double* c, double* f, int nyf, int nxf) { prolong_cell_injection(c, f, nyf, nxf); }
void orc_prolong_cell_bilinear(const double* c, double* f, int nyf, int nxf) { prolong_cell_bilinear(c, f, nyf, nxf); }

// ---- multigrid hierarchy (glide/multigrid.py:8-87, 264-281)
orc_mg* orc_mg_create(int ny, int nx, double dx, int n_levels, const orc_phys* P) {
    orc_mg* mg = new orc_mg;
    int cy = ny, cx = nx;
    double cdx = dx;
    for (int l = 0; l < n_levels; ++l) {
        Level L;
        L.ny = cy; L.nx = cx; L.dx = cdx; L.P = *P;
        for (int id = 0; id < F_COUNT; ++id) L.f[id].assign(L.size(kind_of(id)), 0.0);
        L.F_u.assign(L.size(K_U), 0); L.F_v.assign(L.size(K_V), 0); L.F_H.assign(L.size(K_C), 0);
        L.gamma.assign(L.size(K_C), P->thklim);
        for (auto* s : {&L.w_u, &L.y_u, &L.z_u, &L.vjp_u, &L.wl_u, &L.zl_u}) s->assign(L.size(K_U), 0);
        for (auto* s : {&L.w_v, &L.y_v, &L.z_v, &L.vjp_v, &L.wl_v, &L.zl_v}) s->assign(L.size(K_V), 0);
        for (auto* s : {&L.w_H, &L.y_H, &L.z_H, &L.chi, &L.cphi, &L.vjp_H, &L.wl_H, &L.zl_H}) s->assign(L.size(K_C), 0);
        mg->L.push_back(std::move(L));
        cy /= 2; cx /= 2; cdx *= 2;
    }
    return mg;
}

void orc_mg_destroy(orc_mg* mg) { delete mg; }

int orc_mg_levels(orc_mg* mg) { return (int)mg->L.size(); }

// HierarchyFieldManager.set: write `data` at `level` and restrict it down.
void orc_mg_set(orc_mg* mg, int id, int level, const double* data) {
    Level& L = mg->L[level];
    std::memcpy(L.f[id].data(), data, L.f[id].size() * sizeof(double));
    for (int l = level; l + 1 < (int)mg->L.size(); ++l) restrict_field(id, mg->L[l], mg->L[l + 1]);
}

// Write one level only (no restriction).
void orc_mg_put(orc_mg* mg, int id, int level, const double* data) {
    Level& L = mg->L[level];
    std::memcpy(L.f[id].data(), data, L.f[id].size() * sizeof(double));
}

void orc_mg_get(orc_mg* mg, int id, int level, double* out) {
    Level& L = mg->L[level];
    std::memcpy(out, L.f[id].data(), L.f[id].size() * sizeof(double));
}

void orc_fascd_solve(orc_mg* mg, double dt, int start_level, const orc_fas_config* cfg, const orc_vanka_config* vc,
                     orc_report* rep) {
    Solver s{mg, *cfg, vc, dt};
    s.solve(start_level, rep);
}

void orc_fas_adjoint_solve(orc_mg* mg, double dt, int start_level, const orc_fas_config* cfg,
                           const orc_vanka_config* vc, orc_report* rep) {
    Solver s{mg, *cfg, vc, dt};
    s.adj_solve(start_level, rep);
}

// IceDynamics.backward's gradient tail (glide/model.py:91-94).
void orc_gradients(orc_mg* mg, int level, double dt) {
    Level& L = mg->L[level];
    Fields F = fields_of(L, true);
    grad_beta(L.f[F_GBETA].data(), F, L.f[F_LU].data(), L.f[F_LV].data(), L.P, L.ny, L.nx);
    grad_bed(L.f[F_GBED].data(), F, L.f[F_LU].data(), L.f[F_LV].data(), L.dx, L.ny, L.nx);
    for (size_t k = 0; k < L.size(K_C); ++k) {
        L.f[F_GHPREV][k] = -L.f[F_LH][k] / dt;
        L.f[F_GSMB][k] = -L.f[F_LH][k];
    }
}

}  // extern "C"
