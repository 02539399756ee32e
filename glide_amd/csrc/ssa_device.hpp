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
    int glen_cube;   // n == 3: the viscosity power e2^((1-n)/(2n)) is e2^(-1/3)
};

// e2^(-1/3) for the n = 3 Glen law without the FP64 pow: an FP32 log2/exp2
// estimate (relative error ~1e-6) refined by two division-free Newton steps
// on g(y) = y^-3 - e2, y <- y (4 - e2 y^3) / 3, whose relative error maps
// e -> -2 e^2 (1e-6 -> 2e-12 -> below FP64 rounding).  Outside the range the
// FP32 estimate covers the library pow is used.
__device__ __forceinline__ double inv_cbrt(double x) {
    if (!(x > 1e-30 && x < 1e30)) return pow(x, -1.0 / 3.0);
    double y = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)x) * (-1.0f / 3.0f));
    y = y * (4.0 - x * (y * y * y)) * (1.0 / 3.0);
    y = y * (4.0 - x * (y * y * y)) * (1.0 / 3.0);
    return y;
}
__device__ __forceinline__ float inv_cbrt(float x) { return pow(x, -1.0f / 3.0f); }

// The Glen-law viscosity power e2^((1-n)/(2n)).
template <typename T>
__device__ __forceinline__ T glen_power(T e2, const Phys<T>& P) {
    return P.glen_cube ? inv_cbrt(e2) : pow(e2, P.glen_p);
}

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

// Loads of the state: plain, or device-scope relaxed loads (served by the
// L2, never by a possibly stale L1 line) for kernels whose workgroups read
// what other workgroups of the same launch wrote.
struct LdPlain {
    template <typename T>
    __device__ static __forceinline__ T get(const T* p) { return *p; }
};
struct LdCoherent {
    template <typename T>
    __device__ static __forceinline__ T get(const T* p) {
        return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// Velocities around cell (i,j) that its strain-rate invariant reads.
template <typename T>
struct CellVel {
    T ul, ur, vt, vb;          // own faces
    T utl, vlt, utr, vrt;      // top corners: U(i-1,j), V(i,j-1), U(i-1,j+1), V(i,j+1)
    T ubl, vlb, ubr, vrb;      // bottom corners: U(i+1,j), V(i+1,j-1), U(i+1,j+1), V(i+1,j+1)
};

template <typename T, class Ld = LdPlain>
__device__ __forceinline__ CellVel<T> load_cell_vel(const T* __restrict__ u, const T* __restrict__ v, int i, int j,
                                                    int ny, int nx) {
    CellVel<T> c;
    c.ul = Ld::get(u + offU(i, j, ny, nx));
    c.ur = Ld::get(u + offU(i, j + 1, ny, nx));
    c.vt = Ld::get(v + offV(i, j, ny, nx));
    c.vb = Ld::get(v + offV(i + 1, j, ny, nx));
    c.utl = Ld::get(u + offU(i - 1, j, ny, nx));
    c.vlt = Ld::get(v + offV(i, j - 1, ny, nx));
    c.utr = Ld::get(u + offU(i - 1, j + 1, ny, nx));
    c.vrt = Ld::get(v + offV(i, j + 1, ny, nx));
    c.ubl = Ld::get(u + offU(i + 1, j, ny, nx));
    c.vlb = Ld::get(v + offV(i + 1, j - 1, ny, nx));
    c.ubr = Ld::get(u + offU(i + 1, j + 1, ny, nx));
    c.vrb = Ld::get(v + offV(i + 1, j + 1, ny, nx));
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

// ----------------------------------------------------------- folded cell system
// During a smoothing pass the viscosity, the parameters and every neighbour
// value of a cell's patch are frozen, so the patch system of local_system is
// a low-order polynomial in the cell's unknowns x = (u_l, u_r, v_t, v_b, H_c)
// apart from the four facet flux speeds sqrt(w^2 + 10) and, for m != 1, the
// drag power.  CellSys holds the patch folded into those polynomial
// coefficients once per cell, and cell_system evaluates J(x) and r(x) - f
// (with the Vanka dampings) from it: the same matrix and residual as
// local_system term by term, at about a quarter of the FP64 instructions per
// Newton iteration.  Boundary facets, cells and vertices enter as zero
// coefficients instead of branches.
template <typename T>
struct CellSys {
    T D, hD;                            // 1/dx, 0.5/dx
    // thickness row
    T H_l, H_r, H_t, H_b;               // neighbour thickness of the four facet fluxes
    T dL, dR, dT, dB;                   // D on interior facets, 0 on the domain edge
    T r4h;                              // 1/dt + calving coefficient * D (coefficient of H_c in r4)
    T j24c;                             // r4h + mc_damping
    T f4;
    // velocity block: centre stresses K = kH * H_c in a fixed +-1/+-2 pattern,
    // diagonal a_k + h_k * H_c
    T kH;                               // 2 e_c D^2
    T a0, a1, a2, a3, h0, h1, h2, h3;
    // residual rows 0-3: constant parts (neighbour stresses minus forcing) and
    // coefficients of the row's own velocity (neighbour normal stress + linear drag)
    T c0, c1, c2, c3, o0, o1, o2, o3;
    T eD;                               // 2 e_c D (centre strain -> stress)
    // shear vertices tl, tr, bl, br: eta*H = C + q H_c, strain = s + hD * (own velocities)
    T C_tl, C_tr, C_bl, C_br, q_tl, q_tr, q_bl, q_br;
    T s_tl, s_tr, s_bl, s_br;
    // driving stress: surface S_c = Zc + gc H_c against the neighbours' S
    T S_l, S_r, S_t, S_b, Zc, gc;
    // nonlinear drag: effective beta of the four facets and the cross velocities
    T be_l, be_r, be_t, be_b;
    T u_tl, u_tr, u_bl, u_br, v_lt, v_lb, v_rt, v_rb;
};

template <typename T>
__device__ __forceinline__ void make_cell_sys(CellSys<T>& s, const Patch<T>& p, const Phys<T>& P, T calving_rate,
                                              T dxi, T dt, T ssa_damping, T mc_damping) {
    const T D = dxi, D2 = dxi * dxi;
    s.D = D;
    s.hD = T(0.5) * D;
    s.H_l = p.H_l; s.H_r = p.H_r; s.H_t = p.H_t; s.H_b = p.H_b;
    s.dL = p.in_l ? D : T(0); s.dR = p.in_r ? D : T(0); s.dT = p.in_t ? D : T(0); s.dB = p.in_b ? D : T(0);
    const T kcal = (T(1) - p.p_c) * calving_rate *
                   ((T(1) - p.p_l) + (T(1) - p.p_r) + (T(1) - p.p_t) + (T(1) - p.p_b));
    s.r4h = T(1) / dt + kcal * D;
    s.j24c = s.r4h + mc_damping;
    s.f4 = p.f4;
    s.kH = T(2) * p.e_c * D2;
    // shear vertices (glide/cuda/stress.cu sigma_xy): eta*H averaged over the four cells
    const T qc = T(0.25) * p.e_c;
    s.C_tl = p.vtl ? T(0.25) * (p.e_tl * p.H_tl + p.e_t * p.H_t + p.e_l * p.H_l) : T(0);
    s.C_tr = p.vtr ? T(0.25) * (p.e_t * p.H_t + p.e_tr * p.H_tr + p.e_r * p.H_r) : T(0);
    s.C_bl = p.vbl ? T(0.25) * (p.e_l * p.H_l + p.e_bl * p.H_bl + p.e_b * p.H_b) : T(0);
    s.C_br = p.vbr ? T(0.25) * (p.e_r * p.H_r + p.e_b * p.H_b + p.e_br * p.H_br) : T(0);
    s.q_tl = p.vtl ? qc : T(0); s.q_tr = p.vtr ? qc : T(0); s.q_bl = p.vbl ? qc : T(0); s.q_br = p.vbr ? qc : T(0);
    s.s_tl = s.hD * (p.u_tl - p.v_lt);
    s.s_tr = s.hD * (p.u_tr + p.v_rt);
    s.s_bl = -s.hD * (p.u_bl + p.v_lb);
    s.s_br = s.hD * (p.v_rb - p.u_br);
    // neighbour normal stresses (sigma_xx of the left/right cell, sigma_yy of the top/bottom cell)
    const T nl = p.cl ? T(2) * p.e_l * p.H_l * D2 : T(0);
    const T nr = p.cr ? T(2) * p.e_r * p.H_r * D2 : T(0);
    const T nt = p.ct ? T(2) * p.e_t * p.H_t * D2 : T(0);
    const T nb = p.cb ? T(2) * p.e_b * p.H_b * D2 : T(0);
    s.c0 = -nl * (p.v_lt - p.v_lb - T(2) * p.u_ll) - p.f0;
    s.c1 = nr * (T(2) * p.u_rr + p.v_rt - p.v_rb) - p.f1;
    s.c2 = nt * (p.u_tr - p.u_tl + T(2) * p.v_tt) - p.f2;
    s.c3 = -nb * (p.u_br - p.u_bl - T(2) * p.v_bb) - p.f3;
    // drag: effective beta per facet; the linear law contributes -(beta_eff + water_drag) w
    s.be_l = T(0.5) * (p.b_l * p.p_l + p.b_c * p.p_c);
    s.be_r = T(0.5) * (p.b_c * p.p_c + p.b_r * p.p_r);
    s.be_t = T(0.5) * (p.b_t * p.p_t + p.b_c * p.p_c);
    s.be_b = T(0.5) * (p.b_c * p.p_c + p.b_b * p.p_b);
    const T lin = P.linear_drag ? T(1) : T(0);
    s.o0 = -T(2) * nl - lin * (s.be_l + P.water_drag);
    s.o1 = -T(2) * nr - lin * (s.be_r + P.water_drag);
    s.o2 = -T(2) * nt - lin * (s.be_t + P.water_drag);
    s.o3 = -T(2) * nb - lin * (s.be_b + P.water_drag);
    // diagonal: neighbour normal stress, the two shear vertices touching the facet, linear drag, damping
    s.a0 = s.o0 - (s.C_tl + s.C_bl) * D2 - ssa_damping;
    s.a1 = s.o1 - (s.C_tr + s.C_br) * D2 - ssa_damping;
    s.a2 = s.o2 - (s.C_tl + s.C_tr) * D2 - ssa_damping;
    s.a3 = s.o3 - (s.C_bl + s.C_br) * D2 - ssa_damping;
    s.h0 = -(s.q_tl + s.q_bl) * D2;
    s.h1 = -(s.q_tr + s.q_br) * D2;
    s.h2 = -(s.q_tl + s.q_tr) * D2;
    s.h3 = -(s.q_bl + s.q_br) * D2;
    s.eD = T(2) * p.e_c * D;
    // driving stress surfaces (glide/cuda/stress.cu driving terms)
    s.S_l = p.p_l * p.z_l - (T(1) - p.p_l) * T(0.917) * p.H_l + p.H_l;
    s.S_r = p.p_r * p.z_r - (T(1) - p.p_r) * T(0.917) * p.H_r + p.H_r;
    s.S_t = p.p_t * p.z_t - (T(1) - p.p_t) * T(0.917) * p.H_t + p.H_t;
    s.S_b = p.p_b * p.z_b - (T(1) - p.p_b) * T(0.917) * p.H_b + p.H_b;
    s.Zc = p.p_c * p.z_c;
    s.gc = T(1) - (T(1) - p.p_c) * T(0.917);
    s.u_tl = p.u_tl; s.u_tr = p.u_tr; s.u_bl = p.u_bl; s.u_br = p.u_br;
    s.v_lt = p.v_lt; s.v_lb = p.v_lb; s.v_rt = p.v_rt; s.v_rb = p.v_rb;
}

// 1/sqrt(s) and 1/a to within an ulp or two: the hardware estimates refined
// by two Newton steps each (for the flux speeds, s >= 10, and the LU pivots,
// which are diagonally dominated and far from denormal).
__device__ __forceinline__ double fast_rsqrt(double s) {
    double y = __builtin_amdgcn_rsq(s);
    const double h = 0.5 * s;
    y = y * (1.5 - h * y * y);
    y = y * (1.5 - h * y * y);
    return y;
}
__device__ __forceinline__ double fast_rcp(double a) {
    double y = __builtin_amdgcn_rcp(a);
    y = fma(y, fma(-a, y, 1.0), y);
    y = fma(y, fma(-a, y, 1.0), y);
    return y;
}
// FP32: the hardware estimates are already within an ulp or so of FP32.
__device__ __forceinline__ float fast_rsqrt(float s) { return __builtin_amdgcn_rsqf(s); }
__device__ __forceinline__ float fast_rcp(float a) { return __builtin_amdgcn_rcpf(a); }

// Drag power of the nonlinear law at facet velocity w with cross velocities c0..c3.
template <typename T>
__device__ __forceinline__ void drag_power(T w, T c0, T c1, T c2, T c3, const Phys<T>& P, T& pw, T& dpw) {
    T s2 = w * w + T(0.25) * (c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3);
    pw = pow(s2 + P.u_reg, P.drag_q);
    dpw = P.drag_q * pow(s2 + P.u_reg, P.drag_q - T(1));
}

// The coefficients of a cell system in another precision (the Jacobian of
// the mixed form is evaluated and factorised in FP32).
template <typename TJ, typename T>
__device__ __forceinline__ void cast_cell_sys(CellSys<TJ>& q, const CellSys<T>& s) {
    const T* src = &s.D;
    TJ* dst = &q.D;
    constexpr int n = (int)(sizeof(CellSys<T>) / sizeof(T));
#pragma unroll
    for (int k = 0; k < n; ++k) dst[k] = TJ(src[k]);
}

// r(x) - f (precision T) and J(x) (precision TJ) of the cell with the Vanka
// dampings, before boundary / active-set pinning.  q holds the coefficients
// of s in TJ (the same object when TJ == T).  LINEAR: the drag law is
// linear (m == 1, u_reg > 0).
template <typename T, typename TJ, bool LINEAR>
__device__ __forceinline__ void cell_system(const CellSys<T>& s, const CellSys<TJ>& q, const T x[5], TJ J[25],
                                            T r[5], const Phys<T>& P) {
    // ------------------------------------------------------------ residual
    {
        const T ul = x[0], ur = x[1], vt = x[2], vb = x[3], Hc = x[4];
        const T D = s.D;
        const T du = ur - ul, dv = vt - vb;
        const T exx = (T(2) * du + dv) * D, eyy = (du + T(2) * dv) * D;
        const T sx = s.eD * Hc * exx, sy = s.eD * Hc * eyy;    // 2 eta_c H_c e D
        const T eH_tl = s.C_tl + s.q_tl * Hc, eH_tr = s.C_tr + s.q_tr * Hc;
        const T eH_bl = s.C_bl + s.q_bl * Hc, eH_br = s.C_br + s.q_br * Hc;
        const T e_tl = s.s_tl + s.hD * (vt - ul), e_tr = s.s_tr - s.hD * (ur + vt);
        const T e_bl = s.s_bl + s.hD * (ul + vb), e_br = s.s_br + s.hD * (ur - vb);
        const T t_tl = T(2) * D * eH_tl * e_tl, t_tr = T(2) * D * eH_tr * e_tr;
        const T t_bl = T(2) * D * eH_bl * e_bl, t_br = T(2) * D * eH_br * e_br;
        const T Sc = s.Zc + s.gc * Hc;
        const T Hm_l = T(0.5) * (s.H_l + Hc), Hm_r = T(0.5) * (Hc + s.H_r);
        const T Hm_t = T(0.5) * (s.H_t + Hc), Hm_b = T(0.5) * (Hc + s.H_b);
        r[0] = sx + s.c0 + s.o0 * ul + t_tl - t_bl - Hm_l * (Sc - s.S_l) * s.dL;
        r[1] = -sx + s.c1 + s.o1 * ur + t_tr - t_br - Hm_r * (s.S_r - Sc) * s.dR;
        r[2] = -sy + s.c2 + s.o2 * vt - t_tl + t_tr - Hm_t * (s.S_t - Sc) * s.dT;
        r[3] = sy + s.c3 + s.o3 * vb - t_bl + t_br - Hm_b * (Sc - s.S_b) * s.dB;
        const T sl = ul * ul + T(10), sr = ur * ur + T(10), st = vt * vt + T(10), sb = vb * vb + T(10);
        const T a_l = sl * fast_rsqrt(sl), a_r = sr * fast_rsqrt(sr);   // sqrt(w^2 + 10)
        const T a_t = st * fast_rsqrt(st), a_b = sb * fast_rsqrt(sb);
        const T Fl = Hm_l * ul - T(0.5) * a_l * (Hc - s.H_l);
        const T Fr = Hm_r * ur - T(0.5) * a_r * (s.H_r - Hc);
        const T Ft = Hm_t * vt - T(0.5) * a_t * (s.H_t - Hc);
        const T Fb = Hm_b * vb - T(0.5) * a_b * (Hc - s.H_b);
        r[4] = s.r4h * Hc + (Fr * s.dR - Fl * s.dL) + (Ft * s.dT - Fb * s.dB) - s.f4;
    }
    // ------------------------------------------------------------ Jacobian
    const TJ ul = TJ(x[0]), ur = TJ(x[1]), vt = TJ(x[2]), vb = TJ(x[3]), Hc = TJ(x[4]);
    const TJ D = q.D;
    const TJ K = q.kH * Hc;
    const TJ du = ur - ul, dv = vt - vb;
    const TJ jx = q.eD * ((TJ(2) * du + dv) * D), jy = q.eD * ((du + TJ(2) * dv) * D);
    const TJ e_tl = q.s_tl + q.hD * (vt - ul), e_tr = q.s_tr - q.hD * (ur + vt);
    const TJ e_bl = q.s_bl + q.hD * (ul + vb), e_br = q.s_br + q.hD * (ur - vb);
    const TJ g_tl = TJ(2) * D * q.q_tl * e_tl, g_tr = TJ(2) * D * q.q_tr * e_tr;
    const TJ g_bl = TJ(2) * D * q.q_bl * e_bl, g_br = TJ(2) * D * q.q_br * e_br;
    const TJ Sc = q.Zc + q.gc * Hc;
    const TJ Hm_l = TJ(0.5) * (q.H_l + Hc), Hm_r = TJ(0.5) * (Hc + q.H_r);
    const TJ Hm_t = TJ(0.5) * (q.H_t + Hc), Hm_b = TJ(0.5) * (Hc + q.H_b);
    const TJ dS_l = Sc - q.S_l, dS_r = q.S_r - Sc, dS_t = q.S_t - Sc, dS_b = Sc - q.S_b;
    const TJ K2 = TJ(2) * K;
    J[0] = q.a0 + q.h0 * Hc - K2; J[1] = K2;  J[2] = K;   J[3] = -K;
    J[5] = K2;  J[6] = q.a1 + q.h1 * Hc - K2; J[7] = -K;  J[8] = K;
    J[10] = K;  J[11] = -K; J[12] = q.a2 + q.h2 * Hc - K2; J[13] = K2;
    J[15] = -K; J[16] = K;  J[17] = K2; J[18] = q.a3 + q.h3 * Hc - K2;
    J[4] = jx + g_tl - g_bl - (TJ(0.5) * dS_l + Hm_l * q.gc) * q.dL;
    J[9] = -jx + g_tr - g_br - (TJ(0.5) * dS_r - Hm_r * q.gc) * q.dR;
    J[14] = -jy - g_tl + g_tr - (TJ(0.5) * dS_t - Hm_t * q.gc) * q.dT;
    J[19] = jy - g_bl + g_br - (TJ(0.5) * dS_b + Hm_b * q.gc) * q.dB;
    const TJ sl = ul * ul + TJ(10), sr = ur * ur + TJ(10), st = vt * vt + TJ(10), sb = vb * vb + TJ(10);
    const TJ il = fast_rsqrt(sl), ir = fast_rsqrt(sr), it = fast_rsqrt(st), ib = fast_rsqrt(sb);
    const TJ a_l = sl * il, a_r = sr * ir, a_t = st * it, a_b = sb * ib;
    J[20] = -(Hm_l - TJ(0.5) * (ul * il) * (Hc - q.H_l)) * q.dL;
    J[21] = (Hm_r - TJ(0.5) * (ur * ir) * (q.H_r - Hc)) * q.dR;
    J[22] = (Hm_t - TJ(0.5) * (vt * it) * (q.H_t - Hc)) * q.dT;
    J[23] = -(Hm_b - TJ(0.5) * (vb * ib) * (Hc - q.H_b)) * q.dB;
    J[24] = q.j24c + TJ(0.5) * ((ur + a_r) * q.dR - (ul - a_l) * q.dL + (vt + a_t) * q.dT - (vb - a_b) * q.dB);

    if (!LINEAR) {
        // glide/cuda/stress.cu basal drag with the 4-neighbour cross velocity:
        // residual terms in T, their derivatives in TJ
        const T xl = x[0], xr = x[1], xt = x[2], xb = x[3];
        T pw, dpw;
        drag_power(xl, s.v_lt, xt, s.v_lb, xb, P, pw, dpw);
        r[0] -= (s.be_l * pw + P.water_drag) * xl;
        J[0] -= TJ(s.be_l * (T(2) * dpw * xl * xl + pw) + P.water_drag);
        TJ k = TJ(-s.be_l * T(0.5) * dpw * xl);
        J[2] += k * vt; J[3] += k * vb;
        drag_power(xr, xt, s.v_rt, xb, s.v_rb, P, pw, dpw);
        r[1] -= (s.be_r * pw + P.water_drag) * xr;
        J[6] -= TJ(s.be_r * (T(2) * dpw * xr * xr + pw) + P.water_drag);
        k = TJ(-s.be_r * T(0.5) * dpw * xr);
        J[7] += k * vt; J[8] += k * vb;
        drag_power(xt, s.u_tl, s.u_tr, xl, xr, P, pw, dpw);
        r[2] -= (s.be_t * pw + P.water_drag) * xt;
        J[12] -= TJ(s.be_t * (T(2) * dpw * xt * xt + pw) + P.water_drag);
        k = TJ(-s.be_t * T(0.5) * dpw * xt);
        J[10] += k * ul; J[11] += k * ur;
        drag_power(xb, xl, xr, s.u_bl, s.u_br, P, pw, dpw);
        r[3] -= (s.be_b * pw + P.water_drag) * xb;
        J[18] -= TJ(s.be_b * (T(2) * dpw * xb * xb + pw) + P.water_drag);
        k = TJ(-s.be_b * T(0.5) * dpw * xb);
        J[15] += k * ul; J[16] += k * ur;
    }
}

// lu5_solve with refined hardware reciprocals of the pivots, used for both
// the elimination and the back substitution (no divisions).
template <typename T>
__device__ __forceinline__ void lu5_solve_rcp(T A[25], const T b[5], T x[5]) {
    T inv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        inv[k] = fast_rcp(A[k * 5 + k]);
#pragma unroll
        for (int i = k + 1; i < 5; ++i) {
            A[i * 5 + k] *= inv[k];
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
        x[i] = s * inv[i];
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
