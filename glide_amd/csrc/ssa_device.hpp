// Device-side physics of the coupled SSA momentum + mass discretisation on the
// MAC grid: clamped accessors, Glen-law viscosity, stress / drag / flux terms
// and the 5x5 cell-patch system of the additive Vanka smoother.
//
// Discretisation follows the reference (glide/cuda/common.cu, viscosity.cu,
// stress.cu, flux.cu, vanka.cu); the code is organised for gfx950: every
// neighbour value a cell's local Newton iteration needs is gathered once into
// a register-resident Patch before the iteration starts, so the inner loop is
// pure VALU work with no memory traffic.
#pragma once

#include <hip/hip_runtime.h>

namespace glide {

template <typename T>
struct Phys {
    T n, eps_reg, frd, m, u_reg, water_drag, frs, calving_rate, frc, thklim;
    T glen_p;        // (1-n)/(2n)
    T drag_q;        // (m-1)/2
    int linear_drag; // m == 1 and u_reg > 0: the drag power is exactly 1 and its derivative 0
};

__device__ __forceinline__ int clampi(int a, int lo, int hi) { return min(max(a, lo), hi); }

// Clamp-to-edge element offsets for u-facets (ny, nx+1), v-facets (ny+1, nx)
// and cells (ny, nx).
__device__ __forceinline__ long offU(int i, int j, int ny, int nx) {
    return (long)clampi(i, 0, ny - 1) * (nx + 1) + clampi(j, 0, nx);
}
__device__ __forceinline__ long offV(int i, int j, int ny, int nx) {
    return (long)clampi(i, 0, ny) * nx + clampi(j, 0, nx - 1);
}
__device__ __forceinline__ long offC(int i, int j, int ny, int nx) {
    return (long)clampi(i, 0, ny - 1) * nx + clampi(j, 0, nx - 1);
}

template <typename T>
__device__ __forceinline__ T grounded_fraction(T H, T bed, T c) {
    T z = c * (bed + T(0.917) * H);
    z = fmin(fmax(z, T(-20)), T(20));
    return T(1) / (T(1) + exp(-z));
}

// Velocities around cell (i,j) that its strain-rate invariant reads.
template <typename T>
struct CellVel {
    T ul, ur, vt, vb;          // own faces
    T utl, vlt, utr, vrt;      // top corners: U(i-1,j), V(i,j-1), U(i-1,j+1), V(i,j+1)
    T ubl, vlb, ubr, vrb;      // bottom corners: U(i+1,j), V(i+1,j-1), U(i+1,j+1), V(i+1,j+1)
};

template <typename T>
__device__ __forceinline__ CellVel<T> load_cell_vel(const T* __restrict__ u, const T* __restrict__ v, int i, int j,
                                                    int ny, int nx) {
    CellVel<T> c;
    c.ul = u[offU(i, j, ny, nx)];
    c.ur = u[offU(i, j + 1, ny, nx)];
    c.vt = v[offV(i, j, ny, nx)];
    c.vb = v[offV(i + 1, j, ny, nx)];
    c.utl = u[offU(i - 1, j, ny, nx)];
    c.vlt = v[offV(i, j - 1, ny, nx)];
    c.utr = u[offU(i - 1, j + 1, ny, nx)];
    c.vrt = v[offV(i, j + 1, ny, nx)];
    c.ubl = u[offU(i + 1, j, ny, nx)];
    c.vlb = v[offV(i + 1, j - 1, ny, nx)];
    c.ubr = u[offU(i + 1, j + 1, ny, nx)];
    c.vrb = v[offV(i + 1, j + 1, ny, nx)];
    return c;
}

// Strain-rate invariant of cell (i,j): normal rates on the faces plus the
// mean of the squared shear rates on its four corners, each corner masked on
// the domain boundary (glide/cuda/viscosity.cu:118-168).
template <typename T>
__device__ __forceinline__ T strain_invariant(const CellVel<T>& c, int i, int j, int ny, int nx, T dxi, T eps) {
    T exx = (c.ur - c.ul) * dxi, eyy = (c.vt - c.vb) * dxi;
    T s = T(0);
    if (i > 0 && j > 0) { T e = T(0.5) * ((c.utl - c.ul) * dxi + (c.vt - c.vlt) * dxi); s += e * e; }
    if (i > 0 && j < nx - 1) { T e = T(0.5) * ((c.utr - c.ur) * dxi + (c.vrt - c.vt) * dxi); s += e * e; }
    if (i < ny - 1 && j > 0) { T e = T(0.5) * ((c.ul - c.ubl) * dxi + (c.vb - c.vlb) * dxi); s += e * e; }
    if (i < ny - 1 && j < nx - 1) { T e = T(0.5) * ((c.ur - c.ubr) * dxi + (c.vrb - c.vb) * dxi); s += e * e; }
    return exx * exx + eyy * eyy + exx * eyy + T(0.25) * s + eps;
}

// Directional derivative of the invariant along the velocity direction d.
template <typename T>
__device__ __forceinline__ T strain_invariant_dir(const CellVel<T>& c, const CellVel<T>& d, int i, int j, int ny,
                                                  int nx, T dxi) {
    T exx = (c.ur - c.ul) * dxi, eyy = (c.vt - c.vb) * dxi;
    T dxx = (d.ur - d.ul) * dxi, dyy = (d.vt - d.vb) * dxi;
    T s = T(0);
    if (i > 0 && j > 0) {
        T e = T(0.5) * ((c.utl - c.ul) * dxi + (c.vt - c.vlt) * dxi);
        T de = T(0.5) * ((d.utl - d.ul) * dxi + (d.vt - d.vlt) * dxi);
        s += T(2) * e * de;
    }
    if (i > 0 && j < nx - 1) {
        T e = T(0.5) * ((c.utr - c.ur) * dxi + (c.vrt - c.vt) * dxi);
        T de = T(0.5) * ((d.utr - d.ur) * dxi + (d.vrt - d.vt) * dxi);
        s += T(2) * e * de;
    }
    if (i < ny - 1 && j > 0) {
        T e = T(0.5) * ((c.ul - c.ubl) * dxi + (c.vb - c.vlb) * dxi);
        T de = T(0.5) * ((d.ul - d.ubl) * dxi + (d.vb - d.vlb) * dxi);
        s += T(2) * e * de;
    }
    if (i < ny - 1 && j < nx - 1) {
        T e = T(0.5) * ((c.ur - c.ubr) * dxi + (c.vrb - c.vb) * dxi);
        T de = T(0.5) * ((d.ur - d.ubr) * dxi + (d.vrb - d.vb) * dxi);
        s += T(2) * e * de;
    }
    return T(2) * exx * dxx + T(2) * eyy * dyy + exx * dyy + dxx * eyy + T(0.25) * s;
}

// ---------------------------------------------------------------- stress terms
// sigma_xx / sigma_yy on a cell: value and derivatives with respect to the
// four face velocities and the viscosity-thickness product.
template <typename T>
struct NormalStress { T res, d_ul, d_ur, d_vt, d_vb, d_eH; };

template <typename T>
__device__ __forceinline__ NormalStress<T> sxx(T ul, T ur, T vt, T vb, T eH, T dxi, bool on) {
    NormalStress<T> s{T(0), T(0), T(0), T(0), T(0), T(0)};
    if (!on) return s;
    T e = T(2) * (ur - ul) * dxi + (vt - vb) * dxi;
    T k = T(2) * eH * dxi;
    s.res = T(2) * eH * e;
    s.d_ul = T(-2) * k; s.d_ur = T(2) * k; s.d_vt = k; s.d_vb = -k; s.d_eH = T(2) * e;
    return s;
}
template <typename T>
__device__ __forceinline__ NormalStress<T> syy(T ul, T ur, T vt, T vb, T eH, T dxi, bool on) {
    NormalStress<T> s{T(0), T(0), T(0), T(0), T(0), T(0)};
    if (!on) return s;
    T e = (ur - ul) * dxi + T(2) * (vt - vb) * dxi;
    T k = T(2) * eH * dxi;
    s.res = T(2) * eH * e;
    s.d_ul = -k; s.d_ur = k; s.d_vt = T(2) * k; s.d_vb = T(-2) * k; s.d_eH = T(2) * e;
    return s;
}
// sigma_xy on an interior vertex (zero on boundary vertices).
template <typename T>
struct ShearStress { T res, d_ut, d_ub, d_vl, d_vr, d_eH; };
template <typename T>
__device__ __forceinline__ ShearStress<T> sxy(T ut, T ub, T vl, T vr, T eH, T dxi, bool on) {
    ShearStress<T> s{T(0), T(0), T(0), T(0), T(0), T(0)};
    if (!on) return s;
    T e = T(0.5) * ((ut - ub) * dxi + (vr - vl) * dxi);
    T k = eH * dxi;
    s.res = T(2) * eH * e;
    s.d_ut = k; s.d_ub = -k; s.d_vl = -k; s.d_vr = k; s.d_eH = T(2) * e;
    return s;
}

// Weertman drag with the 4-neighbour cross-velocity (stress.cu:253-369).
template <typename T>
struct Drag { T res, d_w, c0, c1, c2, c3, d_bA, d_bB; };
template <typename T>
__device__ __forceinline__ Drag<T> basal_drag(T w, T c0, T c1, T c2, T c3, T phA, T phB, T bA, T bB,
                                              const Phys<T>& P) {
    Drag<T> t;
    T beff = T(0.5) * (bA * phA + bB * phB);
    T pw, dpw;
    if (P.linear_drag) {
        pw = T(1);
        dpw = T(0);
    } else {
        T s2 = w * w + T(0.25) * (c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3);
        pw = pow(s2 + P.u_reg, P.drag_q);
        dpw = P.drag_q * pow(s2 + P.u_reg, P.drag_q - T(1));
    }
    t.res = -(beff * pw + P.water_drag) * w;
    t.d_w = -(beff * (T(2) * dpw * w * w + pw) + P.water_drag);
    T k = -beff * T(0.5) * dpw * w;
    t.c0 = k * c0; t.c1 = k * c1; t.c2 = k * c2; t.c3 = k * c3;
    t.d_bA = T(-0.5) * phA * pw * w;
    t.d_bB = T(-0.5) * phB * pw * w;
    return t;
}

// Driving stress across a facet; A = left/top cell, B = right/bottom cell.
template <typename T>
struct Drive { T res, d_HA, d_HB, d_zA, d_zB; };
template <typename T>
__device__ __forceinline__ Drive<T> drive_x(T HA, T HB, T zA, T zB, T pA, T pB, T dxi, bool on) {
    Drive<T> d{T(0), T(0), T(0), T(0), T(0)};
    if (!on) return d;
    T Hm = T(0.5) * (HA + HB);
    T SA = pA * zA - (T(1) - pA) * T(0.917) * HA + HA;
    T SB = pB * zB - (T(1) - pB) * T(0.917) * HB + HB;
    T gA = T(1) - (T(1) - pA) * T(0.917), gB = T(1) - (T(1) - pB) * T(0.917);
    d.res = Hm * (SB - SA) * dxi;
    d.d_HA = T(0.5) * (SB - SA) * dxi - Hm * gA * dxi;
    d.d_HB = T(0.5) * (SB - SA) * dxi + Hm * gB * dxi;
    d.d_zA = -Hm * pA * dxi;
    d.d_zB = Hm * pB * dxi;
    return d;
}
template <typename T>
__device__ __forceinline__ Drive<T> drive_y(T HA, T HB, T zA, T zB, T pA, T pB, T dxi, bool on) {
    Drive<T> d{T(0), T(0), T(0), T(0), T(0)};
    if (!on) return d;
    T Hm = T(0.5) * (HA + HB);
    T SA = pA * zA - (T(1) - pA) * T(0.917) * HA + HA;
    T SB = pB * zB - (T(1) - pB) * T(0.917) * HB + HB;
    T gA = T(1) - (T(1) - pA) * T(0.917), gB = T(1) - (T(1) - pB) * T(0.917);
    d.res = Hm * (SA - SB) * dxi;
    d.d_HA = T(0.5) * (SA - SB) * dxi + Hm * gA * dxi;
    d.d_HB = T(0.5) * (SA - SB) * dxi - Hm * gB * dxi;
    d.d_zA = Hm * pA * dxi;
    d.d_zB = -Hm * pB * dxi;
    return d;
}

// Lax-Friedrichs thickness flux with regularised speed sqrt(w^2 + 10);
// x facets: A left, B right; y facets: A top, B bottom, w positive towards
// the top.
template <typename T>
struct Flux { T res, d_w, d_HA, d_HB; };
template <typename T>
__device__ __forceinline__ Flux<T> flux_x(T w, T HA, T HB, bool on) {
    Flux<T> f{T(0), T(0), T(0), T(0)};
    if (!on) return f;
    T Hm = T(0.5) * (HA + HB);
    T a = sqrt(w * w + T(10));
    f.res = Hm * w - T(0.5) * a * (HB - HA);
    f.d_HA = T(0.5) * (w + a);
    f.d_HB = T(0.5) * (w - a);
    f.d_w = Hm - T(0.5) * (w / a) * (HB - HA);
    return f;
}
template <typename T>
__device__ __forceinline__ Flux<T> flux_y(T w, T HA, T HB, bool on) {
    Flux<T> f{T(0), T(0), T(0), T(0)};
    if (!on) return f;
    T Hm = T(0.5) * (HA + HB);
    T a = sqrt(w * w + T(10));
    f.res = Hm * w - T(0.5) * a * (HA - HB);
    f.d_HA = T(0.5) * (w - a);
    f.d_HB = T(0.5) * (w + a);
    f.d_w = Hm - T(0.5) * (w / a) * (HA - HB);
    return f;
}

// ---------------------------------------------------------------- Vanka patch
// Everything cell (i,j)'s local system reads besides its five unknowns
// x = (u_l, u_r, v_t, v_b, H_c).  Viscosity is frozen at the state the
// smoothing pass started from.
template <typename T>
struct Patch {
    T e_c, e_l, e_r, e_t, e_b, e_tl, e_tr, e_bl, e_br;
    T H_l, H_r, H_t, H_b, H_tl, H_tr, H_bl, H_br;
    T p_c, p_l, p_r, p_t, p_b;
    T b_c, b_l, b_r, b_t, b_b;
    T z_c, z_l, z_r, z_t, z_b;
    T u_ll, u_rr, u_tl, u_tr, u_bl, u_br;
    T v_lt, v_lb, v_rt, v_rb, v_tt, v_bb;
    T f0, f1, f2, f3, f4;
    T gam;
    // geometry flags
    bool in_l, in_r, in_t, in_b;         // facet j, j+1, i, i+1 strictly inside the domain
    bool cl, cr, ct, cb;                 // neighbour cells exist
    bool vtl, vtr, vbl, vbr;             // corner vertices interior
};

// Local residual r(x) and the reference's local Jacobian J(x) (row-major);
// viscosity frozen (glide/cuda/vanka.cu:90-475).
template <typename T>
__device__ __forceinline__ void local_system(const Patch<T>& p, const T x[5], T J[25], T r[5], const Phys<T>& P,
                                             T calving_rate, T dxi, T dt) {
#pragma unroll
    for (int k = 0; k < 25; ++k) J[k] = T(0);
#pragma unroll
    for (int k = 0; k < 5; ++k) r[k] = T(0);
    const T ul = x[0], ur = x[1], vt = x[2], vb = x[3], Hc = x[4];

    // thickness row: dH/dt + div(q) + calving
    J[24] = T(1) / dt;
    r[4] += Hc / dt;
    {
        Flux<T> f = flux_x(ul, p.H_l, Hc, p.in_l);
        J[20] -= f.d_w * dxi; J[24] -= f.d_HB * dxi; r[4] -= f.res * dxi;
    }
    {
        T k = (T(1) - p.p_c) * (T(1) - p.p_l) * calving_rate;
        J[24] += k * dxi; r[4] += k * Hc * dxi;
    }
    {
        Flux<T> f = flux_x(ur, Hc, p.H_r, p.in_r);
        J[21] += f.d_w * dxi; J[24] += f.d_HA * dxi; r[4] += f.res * dxi;
    }
    {
        T k = (T(1) - p.p_c) * (T(1) - p.p_r) * calving_rate;
        J[24] += k * dxi; r[4] += k * Hc * dxi;
    }
    {
        Flux<T> f = flux_y(vt, p.H_t, Hc, p.in_t);
        J[22] += f.d_w * dxi; J[24] += f.d_HB * dxi; r[4] += f.res * dxi;
    }
    {
        T k = (T(1) - p.p_c) * (T(1) - p.p_t) * calving_rate;
        J[24] += k * dxi; r[4] += k * Hc * dxi;
    }
    {
        Flux<T> f = flux_y(vb, Hc, p.H_b, p.in_b);
        J[23] -= f.d_w * dxi; J[24] -= f.d_HA * dxi; r[4] -= f.res * dxi;
    }
    {
        T k = (T(1) - p.p_c) * (T(1) - p.p_b) * calving_rate;
        J[24] += k * dxi; r[4] += k * Hc * dxi;
    }

    // centre cell normal stresses: feed all four velocity rows
    {
        NormalStress<T> s = sxx(ul, ur, vt, vb, p.e_c * Hc, dxi, true);
        T a = s.d_eH * p.e_c * dxi;
        r[0] += s.res * dxi; J[0] += s.d_ul * dxi; J[1] += s.d_ur * dxi; J[2] += s.d_vt * dxi; J[3] += s.d_vb * dxi; J[4] += a;
        r[1] -= s.res * dxi; J[5] -= s.d_ul * dxi; J[6] -= s.d_ur * dxi; J[7] -= s.d_vt * dxi; J[8] -= s.d_vb * dxi; J[9] -= a;
        NormalStress<T> q = syy(ul, ur, vt, vb, p.e_c * Hc, dxi, true);
        T b = q.d_eH * p.e_c * dxi;
        r[2] -= q.res * dxi; J[10] -= q.d_ul * dxi; J[11] -= q.d_ur * dxi; J[12] -= q.d_vt * dxi; J[13] -= q.d_vb * dxi; J[14] -= b;
        r[3] += q.res * dxi; J[15] += q.d_ul * dxi; J[16] += q.d_ur * dxi; J[17] += q.d_vt * dxi; J[18] += q.d_vb * dxi; J[19] += b;
    }
    {
        NormalStress<T> s = sxx(p.u_ll, ul, p.v_lt, p.v_lb, p.e_l * p.H_l, dxi, p.cl);
        r[0] -= s.res * dxi; J[0] -= s.d_ur * dxi;
    }
    {
        NormalStress<T> s = sxx(ur, p.u_rr, p.v_rt, p.v_rb, p.e_r * p.H_r, dxi, p.cr);
        r[1] += s.res * dxi; J[6] += s.d_ul * dxi;
    }
    {
        NormalStress<T> s = syy(p.u_tl, p.u_tr, p.v_tt, vt, p.e_t * p.H_t, dxi, p.ct);
        r[2] += s.res * dxi; J[12] += s.d_vb * dxi;
    }
    {
        NormalStress<T> s = syy(p.u_bl, p.u_br, vb, p.v_bb, p.e_b * p.H_b, dxi, p.cb);
        r[3] -= s.res * dxi; J[18] -= s.d_vt * dxi;
    }
    const T qc = T(0.25) * p.e_c;  // d(vertex eta*H)/dH_c
    {
        T eH = T(0.25) * (p.e_tl * p.H_tl + p.e_t * p.H_t + p.e_l * p.H_l + p.e_c * Hc);
        ShearStress<T> s = sxy(p.u_tl, ul, p.v_lt, vt, eH, dxi, p.vtl);
        T a = s.d_eH * qc * dxi;
        r[0] += s.res * dxi; J[0] += s.d_ub * dxi; J[4] += a;
        r[2] -= s.res * dxi; J[12] -= s.d_vr * dxi; J[14] -= a;
    }
    {
        T eH = T(0.25) * (p.e_t * p.H_t + p.e_tr * p.H_tr + p.e_c * Hc + p.e_r * p.H_r);
        ShearStress<T> s = sxy(p.u_tr, ur, vt, p.v_rt, eH, dxi, p.vtr);
        T a = s.d_eH * qc * dxi;
        r[1] += s.res * dxi; J[6] += s.d_ub * dxi; J[9] += a;
        r[2] += s.res * dxi; J[12] += s.d_vl * dxi; J[14] += a;
    }
    {
        T eH = T(0.25) * (p.e_l * p.H_l + p.e_c * Hc + p.e_bl * p.H_bl + p.e_b * p.H_b);
        ShearStress<T> s = sxy(ul, p.u_bl, p.v_lb, vb, eH, dxi, p.vbl);
        T a = s.d_eH * qc * dxi;
        r[0] -= s.res * dxi; J[0] -= s.d_ut * dxi; J[4] -= a;
        r[3] -= s.res * dxi; J[18] -= s.d_vr * dxi; J[19] -= a;
    }
    {
        T eH = T(0.25) * (p.e_c * Hc + p.e_r * p.H_r + p.e_b * p.H_b + p.e_br * p.H_br);
        ShearStress<T> s = sxy(ur, p.u_br, vb, p.v_rb, eH, dxi, p.vbr);
        T a = s.d_eH * qc * dxi;
        r[1] -= s.res * dxi; J[6] -= s.d_ut * dxi; J[9] -= a;
        r[3] += s.res * dxi; J[18] += s.d_vl * dxi; J[19] += a;
    }
    {
        Drag<T> t = basal_drag(ul, p.v_lt, vt, p.v_lb, vb, p.p_l, p.p_c, p.b_l, p.b_c, P);
        r[0] += t.res; J[0] += t.d_w; J[2] += t.c1; J[3] += t.c3;
    }
    {
        Drag<T> t = basal_drag(ur, vt, p.v_rt, vb, p.v_rb, p.p_c, p.p_r, p.b_c, p.b_r, P);
        r[1] += t.res; J[6] += t.d_w; J[7] += t.c0; J[8] += t.c2;
    }
    {
        Drag<T> t = basal_drag(vt, p.u_tl, p.u_tr, ul, ur, p.p_t, p.p_c, p.b_t, p.b_c, P);
        r[2] += t.res; J[12] += t.d_w; J[10] += t.c2; J[11] += t.c3;
    }
    {
        Drag<T> t = basal_drag(vb, ul, ur, p.u_bl, p.u_br, p.p_c, p.p_b, p.b_c, p.b_b, P);
        r[3] += t.res; J[18] += t.d_w; J[15] += t.c0; J[16] += t.c1;
    }
    {
        Drive<T> d = drive_x(p.H_l, Hc, p.z_l, p.z_c, p.p_l, p.p_c, dxi, p.in_l);
        r[0] -= d.res; J[4] -= d.d_HB;
    }
    {
        Drive<T> d = drive_x(Hc, p.H_r, p.z_c, p.z_r, p.p_c, p.p_r, dxi, p.in_r);
        r[1] -= d.res; J[9] -= d.d_HA;
    }
    {
        Drive<T> d = drive_y(p.H_t, Hc, p.z_t, p.z_c, p.p_t, p.p_c, dxi, p.in_t);
        r[2] -= d.res; J[14] -= d.d_HB;
    }
    {
        Drive<T> d = drive_y(Hc, p.H_b, p.z_c, p.z_b, p.p_c, p.p_b, dxi, p.in_b);
        r[3] -= d.res; J[19] -= d.d_HA;
    }
}

template <typename T>
__device__ __forceinline__ void pin_row(T J[25], int row) {
#pragma unroll
    for (int k = 0; k < 5; ++k) { J[row * 5 + k] = T(0); J[k * 5 + row] = T(0); }
    J[row * 5 + row] = T(1);
}

// No-pivot Doolittle LU solve of A x = b, fully unrolled so that A stays in
// VGPRs.  A is overwritten.
template <typename T>
__device__ __forceinline__ void lu5_solve(T A[25], const T b[5], T x[5]) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        T inv = T(1) / A[k * 5 + k];
#pragma unroll
        for (int i = k + 1; i < 5; ++i) {
            A[i * 5 + k] *= inv;
#pragma unroll
            for (int j = k + 1; j < 5; ++j) A[i * 5 + j] -= A[i * 5 + k] * A[k * 5 + j];
        }
    }
    T y[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        T s = b[i];
#pragma unroll
        for (int j = 0; j < i; ++j) s -= A[i * 5 + j] * y[j];
        y[i] = s;
    }
#pragma unroll
    for (int i = 4; i >= 0; --i) {
        T s = y[i];
#pragma unroll
        for (int j = i + 1; j < 5; ++j) s -= A[i * 5 + j] * x[j];
        x[i] = s / A[i * 5 + i];
    }
}

}  // namespace glide
