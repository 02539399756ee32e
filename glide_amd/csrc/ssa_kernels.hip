// gfx950 kernels of the coupled SSA momentum + mass operator:
//   grounded fraction, residual, additive Vanka smoothing (forward and
//   adjoint), the transpose operator (VJP) and the parameter gradients.
//
// Layout (row-major, as the Python API exposes it): u-facets (ny, nx+1),
// v-facets (ny+1, nx), cells (ny, nx).  One launch index (i, j) covers the
// cell (i,j), the u-facet (i,j) and the v-facet (i,j); the index range is
// [0, ny] x [0, nx] so the last facet column/row is included.
//
// Tiles are 64 wide in x so that one wavefront reads one contiguous row
// segment of every array.  The viscosity of the tile's cells plus a one-cell
// ring is computed once per block into LDS.  The transpose operator is a
// deterministic two-phase gather (per-entity partial products in LDS, then one
// owner per output): no floating-point atomics, so repeated runs are bitwise
// identical.
#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "ssa_device.hpp"
#include "ssa_kernels.hpp"
#include "tab.hpp"

namespace glide {

constexpr int TX = 64;  // tile width (one wavefront per row)
constexpr int TY = 4;   // tile height
constexpr int NT = TX * TY;
// Tile height of the forward patch solve (a compile-time tuning constant)
#ifndef GLIDE_VANKA_ROWS
#define GLIDE_VANKA_ROWS 1
#endif
constexpr int kVankaRows = GLIDE_VANKA_ROWS;
// Tile height of the transpose operator
#ifndef GLIDE_VJP_ROWS
#define GLIDE_VJP_ROWS 4
#endif
constexpr int kVjpRows = GLIDE_VJP_ROWS;

template <typename T, class Ld = LdPlain>
__device__ __forceinline__ T cell_viscosity(const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ B,
                                            int ci, int cj, int ny, int nx, T dxi, const Phys<T>& P) {
    CellVel<T> c = load_cell_vel<T, Ld>(u, v, ci, cj, ny, nx);
    T e2 = strain_invariant(c, ci, cj, ny, nx, dxi, P.eps_reg);
    return T(0.5) * B[(long)ci * nx + cj] * glen_power(e2, P);
}

// ------------------------------------------------------------------ grounded
template <typename T>
__global__ __launch_bounds__(256) void k_grounded(T* __restrict__ phi, const T* __restrict__ H,
                                                  const T* __restrict__ bed, T c, T relax, long n) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    T g = grounded_fraction(H[k], bed[k], c);
    phi[k] = relax == T(0) ? g : (T(1) - relax) * g + relax * phi[k];
}

// ------------------------------------------------------------------ residual
// Kernel bodies take their block index (bx, by) as arguments, so that the
// domain-table launches (one launch over many domains) run them unchanged.
template <typename T>
__device__ __forceinline__ void residual_body(int bx, int by, const SsaArgs<T>& a, T* __restrict__ out_u,
                                              T* __restrict__ out_v, T* __restrict__ out_H, int use_forcing,
                                              int use_mask, T calving_rate) {
    __shared__ T s_eta[TY + 2][TX + 2];
    const int ny = a.ny, nx = a.nx;
    const T dxi = a.dxi;
    const int i0 = by * TY, j0 = bx * TX;
    const int tid = threadIdx.y * TX + threadIdx.x;
    for (int k = tid; k < (TY + 2) * (TX + 2); k += NT) {
        int ci = i0 - 1 + k / (TX + 2), cj = j0 - 1 + k % (TX + 2);
        T e = T(0);
        if (ci >= 0 && ci < ny && cj >= 0 && cj < nx) e = cell_viscosity(a.u, a.v, a.B, ci, cj, ny, nx, dxi, a.P);
        s_eta[k / (TX + 2)][k % (TX + 2)] = e;
    }
    __syncthreads();
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    if (i > ny || j > nx) return;
    const int li = threadIdx.y + 1, lj = threadIdx.x + 1;
    const T* __restrict__ u = a.u;
    const T* __restrict__ v = a.v;
    const T* __restrict__ H = a.H;
    const T* __restrict__ ph = a.phi;

    if (i < ny && j < nx) {
        const long c = (long)i * nx + j;
        T Hc = H[c], pc = ph[c];
        T Hl = H[offC(i, j - 1, ny, nx)], Hr = H[offC(i, j + 1, ny, nx)];
        T Ht = H[offC(i - 1, j, ny, nx)], Hb = H[offC(i + 1, j, ny, nx)];
        T rH = Hc / a.dt;
        if (use_forcing) rH -= a.f_H[c];
        rH -= flux_x(u[offU(i, j, ny, nx)], Hl, Hc, j > 0).res * dxi;
        rH += (T(1) - pc) * (T(1) - ph[offC(i, j - 1, ny, nx)]) * calving_rate * Hc * dxi;
        rH += flux_x(u[offU(i, j + 1, ny, nx)], Hc, Hr, j + 1 < nx).res * dxi;
        rH += (T(1) - pc) * (T(1) - ph[offC(i, j + 1, ny, nx)]) * calving_rate * Hc * dxi;
        rH += flux_y(v[offV(i, j, ny, nx)], Ht, Hc, i > 0).res * dxi;
        rH += (T(1) - pc) * (T(1) - ph[offC(i - 1, j, ny, nx)]) * calving_rate * Hc * dxi;
        rH -= flux_y(v[offV(i + 1, j, ny, nx)], Hc, Hb, i + 1 < ny).res * dxi;
        rH += (T(1) - pc) * (T(1) - ph[offC(i + 1, j, ny, nx)]) * calving_rate * Hc * dxi;
        T mk = use_mask ? a.mask[c] : T(0);
        out_H[c] = (T(1) - mk) * rH + mk * (Hc - a.gamma[c]);
    }
    if (i < ny) {
        const long o = (long)i * (nx + 1) + j;
        T ru;
        if (j == 0 || j == nx) {
            ru = u[o];
        } else {
            ru = use_forcing ? -a.f_u[o] : T(0);
            T ul = u[o];
            T Hc = H[(long)i * nx + j], Hl = H[(long)i * nx + j - 1];
            T vt = v[(long)i * nx + j], vb = v[(long)(i + 1) * nx + j];
            T vlt = v[(long)i * nx + j - 1], vlb = v[(long)(i + 1) * nx + j - 1];
            ru += sxx(ul, u[o + 1], vt, vb, s_eta[li][lj] * Hc, dxi, true).res * dxi;
            ru -= sxx(u[o - 1], ul, vlt, vlb, s_eta[li][lj - 1] * Hl, dxi, true).res * dxi;
            if (i > 0) {
                T eH = T(0.25) * (s_eta[li - 1][lj - 1] * H[(long)(i - 1) * nx + j - 1] +
                                  s_eta[li - 1][lj] * H[(long)(i - 1) * nx + j] + s_eta[li][lj - 1] * Hl +
                                  s_eta[li][lj] * Hc);
                ru += sxy(u[o - (nx + 1)], ul, vlt, vt, eH, dxi, true).res * dxi;
            }
            if (i + 1 < ny) {
                T eH = T(0.25) * (s_eta[li][lj - 1] * Hl + s_eta[li][lj] * Hc +
                                  s_eta[li + 1][lj - 1] * H[(long)(i + 1) * nx + j - 1] +
                                  s_eta[li + 1][lj] * H[(long)(i + 1) * nx + j]);
                ru -= sxy(ul, u[o + (nx + 1)], vlb, vb, eH, dxi, true).res * dxi;
            }
            T pl = ph[(long)i * nx + j - 1], pc = ph[(long)i * nx + j];
            ru += basal_drag(ul, vlt, vt, vlb, vb, pl, pc, a.beta[(long)i * nx + j - 1], a.beta[(long)i * nx + j], a.P)
                      .res;
            ru -= drive_x(Hl, Hc, a.bed[(long)i * nx + j - 1], a.bed[(long)i * nx + j], pl, pc, dxi, true).res;
        }
        out_u[o] = ru;
    }
    if (j < nx) {
        const long o = (long)i * nx + j;
        T rv;
        if (i == 0 || i == ny) {
            rv = v[o];
        } else {
            rv = use_forcing ? -a.f_v[o] : T(0);
            T vt = v[o];
            T Hc = H[o], Ht = H[o - nx];
            T ul = u[(long)i * (nx + 1) + j], ur = u[(long)i * (nx + 1) + j + 1];
            T utl = u[(long)(i - 1) * (nx + 1) + j], utr = u[(long)(i - 1) * (nx + 1) + j + 1];
            rv += syy(utl, utr, v[o - nx], vt, s_eta[li - 1][lj] * Ht, dxi, true).res * dxi;
            rv -= syy(ul, ur, vt, v[o + nx], s_eta[li][lj] * Hc, dxi, true).res * dxi;
            if (j > 0) {
                T eH = T(0.25) * (s_eta[li - 1][lj - 1] * H[o - nx - 1] + s_eta[li - 1][lj] * Ht +
                                  s_eta[li][lj - 1] * H[o - 1] + s_eta[li][lj] * Hc);
                rv -= sxy(utl, ul, v[o - 1], vt, eH, dxi, true).res * dxi;
            }
            if (j + 1 < nx) {
                T eH = T(0.25) * (s_eta[li - 1][lj] * Ht + s_eta[li - 1][lj + 1] * H[o - nx + 1] +
                                  s_eta[li][lj] * Hc + s_eta[li][lj + 1] * H[o + 1]);
                rv += sxy(utr, ur, vt, v[o + 1], eH, dxi, true).res * dxi;
            }
            T pt = ph[o - nx], pc = ph[o];
            rv += basal_drag(vt, utl, utr, ul, ur, pt, pc, a.beta[o - nx], a.beta[o], a.P).res;
            rv -= drive_y(Ht, Hc, a.bed[o - nx], a.bed[o], pt, pc, dxi, true).res;
        }
        out_v[o] = rv;
    }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_residual(SsaArgs<T> a, T* __restrict__ out_u, T* __restrict__ out_v,
                                                 T* __restrict__ out_H, int use_forcing, int use_mask,
                                                 T calving_rate) {
    residual_body<T>(blockIdx.x, blockIdx.y, a, out_u, out_v, out_H, use_forcing, use_mask, calving_rate);
}

// ------------------------------------------------------------------ JVP
// J(x) d of the coupled residual along d = (du, dv, dH) with the grounded
// fraction, mask, bed, B and beta frozen (glide/cuda/residuals.cu:340-660):
// the derivative of k_residual term by term.  Dirichlet facet rows are zero
// and the thickness row is (1 - mask) times the mass-balance derivative, as
// in the reference.  The tile holds each cell's viscosity and its derivative
// along d.
template <typename T>
__global__ __launch_bounds__(NT) void k_jvp(SsaArgs<T> a, const T* __restrict__ du, const T* __restrict__ dv,
                                            const T* __restrict__ dH, T* __restrict__ out_u, T* __restrict__ out_v,
                                            T* __restrict__ out_H, int use_mask, T calving_rate) {
    __shared__ T s_eta[TY + 2][TX + 2];
    __shared__ T s_etd[TY + 2][TX + 2];
    const int ny = a.ny, nx = a.nx;
    const T dxi = a.dxi;
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * TX;
    const int tid = threadIdx.y * TX + threadIdx.x;
    for (int k = tid; k < (TY + 2) * (TX + 2); k += NT) {
        int ci = i0 - 1 + k / (TX + 2), cj = j0 - 1 + k % (TX + 2);
        T e = T(0), ed = T(0);
        if (ci >= 0 && ci < ny && cj >= 0 && cj < nx) {
            CellVel<T> c = load_cell_vel(a.u, a.v, ci, cj, ny, nx);
            CellVel<T> d = load_cell_vel(du, dv, ci, cj, ny, nx);
            T e2 = strain_invariant(c, ci, cj, ny, nx, dxi, a.P.eps_reg);
            T de2 = strain_invariant_dir(c, d, ci, cj, ny, nx, dxi);
            e = T(0.5) * a.B[(long)ci * nx + cj] * glen_power(e2, a.P);
            ed = e * a.P.glen_p * de2 / e2;
        }
        s_eta[k / (TX + 2)][k % (TX + 2)] = e;
        s_etd[k / (TX + 2)][k % (TX + 2)] = ed;
    }
    __syncthreads();
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    if (i > ny || j > nx) return;
    const int li = threadIdx.y + 1, lj = threadIdx.x + 1;
    const T* __restrict__ u = a.u;
    const T* __restrict__ v = a.v;
    const T* __restrict__ H = a.H;
    const T* __restrict__ ph = a.phi;
    // eta H of tile cell (ti, tj) at grid cell c and its derivative along d
    auto eH = [&](int ti, int tj, long c) { return s_eta[ti][tj] * H[c]; };
    auto deH = [&](int ti, int tj, long c) { return s_etd[ti][tj] * H[c] + s_eta[ti][tj] * dH[c]; };

    if (i < ny && j < nx) {
        const long c = (long)i * nx + j;
        T Hc = H[c], pc = ph[c], dHc = dH[c];
        const long cl = offC(i, j - 1, ny, nx), cr = offC(i, j + 1, ny, nx);
        const long ct = offC(i - 1, j, ny, nx), cb = offC(i + 1, j, ny, nx);
        T d = dHc / a.dt;
        Flux<T> f = flux_x(u[offU(i, j, ny, nx)], H[cl], Hc, j > 0);
        d -= (f.d_w * du[offU(i, j, ny, nx)] + f.d_HA * dH[cl] + f.d_HB * dHc) * dxi;
        d += (T(1) - pc) * (T(1) - ph[cl]) * calving_rate * dHc * dxi;
        f = flux_x(u[offU(i, j + 1, ny, nx)], Hc, H[cr], j + 1 < nx);
        d += (f.d_w * du[offU(i, j + 1, ny, nx)] + f.d_HA * dHc + f.d_HB * dH[cr]) * dxi;
        d += (T(1) - pc) * (T(1) - ph[cr]) * calving_rate * dHc * dxi;
        f = flux_y(v[offV(i, j, ny, nx)], H[ct], Hc, i > 0);
        d += (f.d_w * dv[offV(i, j, ny, nx)] + f.d_HA * dH[ct] + f.d_HB * dHc) * dxi;
        d += (T(1) - pc) * (T(1) - ph[ct]) * calving_rate * dHc * dxi;
        f = flux_y(v[offV(i + 1, j, ny, nx)], Hc, H[cb], i + 1 < ny);
        d -= (f.d_w * dv[offV(i + 1, j, ny, nx)] + f.d_HA * dHc + f.d_HB * dH[cb]) * dxi;
        d += (T(1) - pc) * (T(1) - ph[cb]) * calving_rate * dHc * dxi;
        T mk = use_mask ? a.mask[c] : T(0);
        out_H[c] = (T(1) - mk) * d;
    }
    if (i < ny) {
        const long o = (long)i * (nx + 1) + j;
        T d = T(0);
        if (j > 0 && j < nx) {
            const long c = (long)i * nx + j;
            T ul = u[o], dul = du[o];
            T vt = v[c], vb = v[c + nx], vlt = v[c - 1], vlb = v[c + nx - 1];
            T dvt = dv[c], dvb = dv[c + nx], dvlt = dv[c - 1], dvlb = dv[c + nx - 1];
            NormalStress<T> s = sxx(ul, u[o + 1], vt, vb, eH(li, lj, c), dxi, true);
            d += (s.d_ul * dul + s.d_ur * du[o + 1] + s.d_vt * dvt + s.d_vb * dvb + s.d_eH * deH(li, lj, c)) * dxi;
            s = sxx(u[o - 1], ul, vlt, vlb, eH(li, lj - 1, c - 1), dxi, true);
            d -= (s.d_ul * du[o - 1] + s.d_ur * dul + s.d_vt * dvlt + s.d_vb * dvlb +
                  s.d_eH * deH(li, lj - 1, c - 1)) * dxi;
            if (i > 0) {
                T e = T(0.25) * (eH(li - 1, lj - 1, c - nx - 1) + eH(li - 1, lj, c - nx) + eH(li, lj - 1, c - 1) +
                                 eH(li, lj, c));
                T de = T(0.25) * (deH(li - 1, lj - 1, c - nx - 1) + deH(li - 1, lj, c - nx) +
                                  deH(li, lj - 1, c - 1) + deH(li, lj, c));
                ShearStress<T> t = sxy(u[o - (nx + 1)], ul, vlt, vt, e, dxi, true);
                d += (t.d_ut * du[o - (nx + 1)] + t.d_ub * dul + t.d_vl * dvlt + t.d_vr * dvt + t.d_eH * de) * dxi;
            }
            if (i + 1 < ny) {
                T e = T(0.25) * (eH(li, lj - 1, c - 1) + eH(li, lj, c) + eH(li + 1, lj - 1, c + nx - 1) +
                                 eH(li + 1, lj, c + nx));
                T de = T(0.25) * (deH(li, lj - 1, c - 1) + deH(li, lj, c) + deH(li + 1, lj - 1, c + nx - 1) +
                                  deH(li + 1, lj, c + nx));
                ShearStress<T> t = sxy(ul, u[o + (nx + 1)], vlb, vb, e, dxi, true);
                d -= (t.d_ut * dul + t.d_ub * du[o + (nx + 1)] + t.d_vl * dvlb + t.d_vr * dvb + t.d_eH * de) * dxi;
            }
            T pl = ph[c - 1], pc = ph[c];
            Drag<T> g = basal_drag(ul, vlt, vt, vlb, vb, pl, pc, a.beta[c - 1], a.beta[c], a.P);
            d += g.d_w * dul + g.c0 * dvlt + g.c1 * dvt + g.c2 * dvlb + g.c3 * dvb;
            Drive<T> r = drive_x(H[c - 1], H[c], a.bed[c - 1], a.bed[c], pl, pc, dxi, true);
            d -= r.d_HA * dH[c - 1] + r.d_HB * dH[c];

        }
        out_u[o] = d;
    }
    if (j < nx) {
        const long o = (long)i * nx + j;
        T d = T(0);
        if (i > 0 && i < ny) {
            const long ou = (long)i * (nx + 1) + j, ot = ou - (nx + 1);
            T vt = v[o], dvt = dv[o];
            T ul = u[ou], ur = u[ou + 1], utl = u[ot], utr = u[ot + 1];
            T dul = du[ou], dur = du[ou + 1], dutl = du[ot], dutr = du[ot + 1];
            NormalStress<T> s = syy(utl, utr, v[o - nx], vt, eH(li - 1, lj, o - nx), dxi, true);
            d += (s.d_ul * dutl + s.d_ur * dutr + s.d_vt * dv[o - nx] + s.d_vb * dvt +
                  s.d_eH * deH(li - 1, lj, o - nx)) * dxi;
            s = syy(ul, ur, vt, v[o + nx], eH(li, lj, o), dxi, true);
            d -= (s.d_ul * dul + s.d_ur * dur + s.d_vt * dvt + s.d_vb * dv[o + nx] + s.d_eH * deH(li, lj, o)) * dxi;
            if (j > 0) {
                T e = T(0.25) * (eH(li - 1, lj - 1, o - nx - 1) + eH(li - 1, lj, o - nx) + eH(li, lj - 1, o - 1) +
                                 eH(li, lj, o));
                T de = T(0.25) * (deH(li - 1, lj - 1, o - nx - 1) + deH(li - 1, lj, o - nx) +
                                  deH(li, lj - 1, o - 1) + deH(li, lj, o));
                ShearStress<T> t = sxy(utl, ul, v[o - 1], vt, e, dxi, true);
                d -= (t.d_ut * dutl + t.d_ub * dul + t.d_vl * dv[o - 1] + t.d_vr * dvt + t.d_eH * de) * dxi;
            }
            if (j + 1 < nx) {
                T e = T(0.25) * (eH(li - 1, lj, o - nx) + eH(li - 1, lj + 1, o - nx + 1) + eH(li, lj, o) +
                                 eH(li, lj + 1, o + 1));
                T de = T(0.25) * (deH(li - 1, lj, o - nx) + deH(li - 1, lj + 1, o - nx + 1) + deH(li, lj, o) +
                                  deH(li, lj + 1, o + 1));
                ShearStress<T> t = sxy(utr, ur, vt, v[o + 1], e, dxi, true);
                d += (t.d_ut * dutr + t.d_ub * dur + t.d_vl * dvt + t.d_vr * dv[o + 1] + t.d_eH * de) * dxi;
            }
            T pt = ph[o - nx], pc = ph[o];
            Drag<T> g = basal_drag(vt, utl, utr, ul, ur, pt, pc, a.beta[o - nx], a.beta[o], a.P);
            d += g.d_w * dvt + g.c0 * dutl + g.c1 * dutr + g.c2 * dul + g.c3 * dur;
            Drive<T> r = drive_y(H[o - nx], H[o], a.bed[o - nx], a.bed[o], pt, pc, dxi, true);
            d -= r.d_HA * dH[o - nx] + r.d_HB * dH[o];

        }
        out_v[o] = d;
    }
}

// ------------------------------------------------------------------ Vanka
// Gather the patch of cell (i,j); viscosity from the LDS tile (li, lj are the
// cell's tile coordinates including the one-cell ring).
template <typename T, int W = TX + 2, class Ld = LdPlain>
__device__ __forceinline__ void load_patch(Patch<T>& p, const SsaArgs<T>& a, const T (*s_eta)[W], int li, int lj,
                                           int i, int j, const T* __restrict__ fu, const T* __restrict__ fv,
                                           const T* __restrict__ fH, const T* __restrict__ phi) {
    load_patch_with<T, Ld>(p, a, [&](int di, int dj) { return s_eta[li + di][lj + dj]; }, i, j, fu, fv, fH, phi);
}

// Viscosity tile (+ ring) of cells, optionally refreshing the grounded
// fraction of the tile's own cells first (the ring recomputes it from H, bed).
template <typename T, int ROWS = TY>
__device__ __forceinline__ void fill_eta_tile(T (*s_eta)[TX + 2], const SsaArgs<T>& a, int i0, int j0, int tid) {
    for (int k = tid; k < (ROWS + 2) * (TX + 2); k += TX * ROWS) {
        int ci = i0 - 1 + k / (TX + 2), cj = j0 - 1 + k % (TX + 2);
        T e = T(0);
        if (ci >= 0 && ci < a.ny && cj >= 0 && cj < a.nx) e = cell_viscosity(a.u, a.v, a.B, ci, cj, a.ny, a.nx, a.dxi, a.P);
        s_eta[k / (TX + 2)][k % (TX + 2)] = e;
    }
}

// Forward smoothing: every cell runs the local damped Newton iteration on its
// patch against the state at the start of the pass and writes its increments
// (half-weighted for facets) into the workspace planes
// work[0..4] = (du_l, du_r, dv_t, dv_b, dH); the active-set flag goes to mask.
// The damped local Newton iteration of cell (i,j) on its patch against the
// state in a (viscosity from the tile s_eta, whose cell (li, lj) is (i,j)):
// returns the increments w = ((u_l, u_r, v_t, v_b) change / 2, H change) and
// the active-set flag.  edge: the calling wave may hold cells on the domain's
// edge (the loop then tests the Dirichlet pins; interior waves carry no pin
// code).  The wave-uniform edge test keeps one loop copy per wave.
template <typename T, bool LINEAR, typename TJ, int W, class Ld = LdPlain>
__device__ __forceinline__ void cell_newton(const SsaArgs<T>& a, const T (*s_eta)[W], int li, int lj, int i, int j,
                                            const VankaArgs<T>& vk, bool edge, T w[5], T& masked, int& iters) {
    const int ny = a.ny, nx = a.nx;
    CellSys<T> cs;
    T thk;
    {
        Patch<T> p;
        load_patch<T, W, Ld>(p, a, s_eta, li, lj, i, j, a.f_u, a.f_v, a.f_H, a.phi);
        make_cell_sys(cs, p, a.P, vk.calving_rate, a.dxi, a.dt, vk.ssa_damping, vk.mc_damping);
        thk = p.gam;
    }
    CellSys<TJ> cq_own;
    if constexpr (!std::is_same<T, TJ>::value) cast_cell_sys(cq_own, cs);
    const CellSys<TJ>* cq;
    if constexpr (std::is_same<T, TJ>::value) cq = &cs;
    else cq = &cq_own;
    const long c = (long)i * nx + j;
    const T x0[5] = {Ld::get(a.u + (long)i * (nx + 1) + j), Ld::get(a.u + (long)i * (nx + 1) + j + 1), Ld::get(a.v + c),
                     Ld::get(a.v + c + nx), Ld::get(a.H + c)};
    T x[5] = {x0[0], x0[1], x0[2], x0[3], x0[4]};
    T cmp[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) cmp[k] = T(0);
    T rnorm = T(1);
    const T tol = T(1e-6);
    const bool pin0 = j == 0, pin1 = j == nx - 1, pin2 = i == 0, pin3 = i == ny - 1;
    int it = 0;
    auto newton = [&](auto edge_tag) {
        constexpr bool EDGE = decltype(edge_tag)::value;
        while (it < vk.newton_steps && rnorm > tol) {
            TJ J[25];
            T r[5];
            cell_system<T, TJ, LINEAR>(cs, *cq, x, J, r, a.P);
            if constexpr (EDGE) {
                if (pin0) { pin_row(J, 0); r[0] = x[0]; }
                if (pin1) { pin_row(J, 1); r[1] = x[1]; }
                if (pin2) { pin_row(J, 2); r[2] = x[2]; }
                if (pin3) { pin_row(J, 3); r[3] = x[3]; }
            }
            if ((x[4] - a.dt * r[4]) <= thk) {
                masked = T(1);
                pin_row(J, 4);
                r[4] = x[4] - thk;
            } else {
                masked = T(0);
            }
            TJ rj[5] = {TJ(r[0]), TJ(r[1]), TJ(r[2]), TJ(r[3]), TJ(r[4])};
            TJ d[5];
            lu5_solve_rcp(J, rj, d);
            rnorm = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3] + r[4] * r[4];
#pragma unroll
            for (int k = 0; k < 5; ++k) {  // compensated relaxed update, local relaxation 0.5
                T y = T(-0.5) * T(d[k]) - cmp[k];
                T t = x[k] + y;
                cmp[k] = (t - x[k]) - y;
                x[k] = t;
            }
            x[4] = fmax(x[4], thk);
            ++it;
        }
    };
    if (edge) newton(std::true_type{});
    else newton(std::false_type{});
    w[0] = T(0.5) * (x[0] - x0[0]);
    w[1] = T(0.5) * (x[1] - x0[1]);
    w[2] = T(0.5) * (x[2] - x0[2]);
    w[3] = T(0.5) * (x[3] - x0[3]);
    w[4] = x[4] - x0[4];
    iters = it;
}

// Forward smoothing: every cell runs the local damped Newton iteration on its
// patch against the state at the start of the pass and writes its increments
// (half-weighted for facets) into the workspace planes
// work[0..4] = (du_l, du_r, dv_t, dv_b, dH); the active-set flag goes to mask.
template <typename T, bool LINEAR, typename TJ, int ROWS>
__global__ __launch_bounds__(TX * ROWS) void k_vanka_cell(SsaArgs<T> a, T* __restrict__ work, T* __restrict__ mask,
                                                          VankaArgs<T> vk) {
    __shared__ T s_eta[ROWS + 2][TX + 2];
    const int i0 = blockIdx.y * ROWS, j0 = blockIdx.x * TX;
    const int tid = threadIdx.y * TX + threadIdx.x;
    fill_eta_tile<T, ROWS>(s_eta, a, i0, j0, tid);
    __syncthreads();
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    const int ny = a.ny, nx = a.nx;
    if (i >= ny || j >= nx) return;
    const long c = (long)i * nx + j;
    T w[5], masked = mask[c];
    int it = 0;
    const bool edge_wave = i == 0 || i == ny - 1 || j0 == 0 || j0 + TX >= nx;
    cell_newton<T, LINEAR, TJ, TX + 2>(a, s_eta, threadIdx.y + 1, threadIdx.x + 1, i, j, vk, edge_wave, w, masked, it);
    const long n = (long)ny * nx;
#pragma unroll
    for (int q = 0; q < 5; ++q) work[q * n + c] = w[q];
    mask[c] = masked;
    if (vk.iters) vk.iters[c] = it;
}

// A whole forward smoothing pass in one launch, for levels small enough that
// the patch solves are latency-bound: the block solves the cells of one tile
// row (wave 0) together with the cells above them (wave 1) and the cell left
// of them (wave 2), whose bottom / right facet increments the row's facets
// also receive, then writes the updated state of the facets and cells the
// row owns into a second buffer with the arithmetic of k_vanka_apply.  The
// pass reads the state from a (unchanged during the pass, so no block sees
// another block's update) and writes u_dst, v_dst, H_dst; the active-set flag
// is written by the owner in place (no other cell reads it).
template <typename T, bool LINEAR, typename TJ>
__global__ __launch_bounds__(3 * TX) void k_vanka_fused(SsaArgs<T> a, T* __restrict__ u_dst, T* __restrict__ v_dst,
                                                        T* __restrict__ H_dst, T* __restrict__ mask, VankaArgs<T> vk,
                                                        T omega) {
    constexpr int W = TX + 3;
    __shared__ T s_eta[4][W];     // cells (i0 - 2 .. i0 + 1, j0 - 2 .. j0 + TX)
    __shared__ T s_w1[TX + 1];    // right-facet increments of cells (i0, j0 - 1 + q)
    __shared__ T s_w3[TX];        // bottom-facet increments of cells (i0 - 1, j0 + q)
    const int ny = a.ny, nx = a.nx;
    const int i0 = blockIdx.y, j0 = blockIdx.x * TX;
    const int wv = threadIdx.y, lane = threadIdx.x, tid = wv * TX + lane;
    for (int k = tid; k < 4 * W; k += 3 * TX) {
        const int ci = i0 - 2 + k / W, cj = j0 - 2 + k % W;
        T e = T(0);
        if (ci >= 0 && ci < ny && cj >= 0 && cj < nx) e = cell_viscosity(a.u, a.v, a.B, ci, cj, ny, nx, a.dxi, a.P);
        s_eta[k / W][k % W] = e;
    }
    __syncthreads();
    int ci = i0, cj = j0 + lane;
    bool active = cj < nx;
    bool edge = i0 == 0 || i0 == ny - 1 || j0 == 0 || j0 + TX >= nx;
    if (wv == 1) {
        ci = i0 - 1;
        active = ci >= 0 && cj < nx;
        edge = ci <= 0 || ci == ny - 1 || j0 == 0 || j0 + TX >= nx;
    } else if (wv == 2) {
        cj = j0 - 1;
        active = lane == 0 && cj >= 0;
        edge = true;
    }
    T w[5] = {T(0), T(0), T(0), T(0), T(0)};
    T masked = (wv == 0 && active) ? mask[(long)i0 * nx + cj] : T(0);
    int it = 0;
    if (active) {
        cell_newton<T, LINEAR, TJ, W>(a, s_eta, ci - i0 + 2, cj - j0 + 2, ci, cj, vk, edge, w, masked, it);
        if (wv == 0) s_w1[lane + 1] = w[1];
        else if (wv == 1) s_w3[lane] = w[3];
        else s_w1[0] = w[1];
    }
    __syncthreads();
    if (wv != 0 || !active) return;
    const int i = i0, j = cj;
    const long c = (long)i * nx + j, ou = (long)i * (nx + 1) + j;
    {
        T d = w[0];
        if (j > 0) d = d + s_w1[lane];
        u_dst[ou] = a.u[ou] + omega * d;
        if (j == nx - 1) {
            const T e = T(0) + w[1];
            u_dst[ou + 1] = a.u[ou + 1] + omega * e;
        }
    }
    {
        T d = w[2];
        if (i > 0) d = d + s_w3[lane];
        v_dst[c] = a.v[c] + omega * d;
        if (i == ny - 1) {
            const T e = T(0) + w[3];
            v_dst[c + nx] = a.v[c + nx] + omega * e;
        }
    }
    H_dst[c] = a.H[c] + omega * w[4];
    mask[c] = masked;
}

// Barrier of the working workgroups of a persistent launch (a handful, all
// resident at once).  sync[0] counts the arrivals of the launch (zeroed before
// it); barrier k of the launch waits for nb * (k + 1) of them.  Every thread's
// stores have reached the L2 before its workgroup arrives.  When the
// workgroups run on one XCD they share that L2 and read each other's state
// with device-scope loads (LdCoherent), so the barrier needs no cache
// maintenance; otherwise (fenced) the arriving thread's agent-scope fences
// write its L2 back before the arrival and invalidate it after the wait.  The
// wait is bounded: after ctl.spin_limit polls (a negative limit gives up at
// once, which tests the failure path) a workgroup sets sync[1], after which no
// barrier of the launch waits; the kernel then writes NaN into its last pass's
// output, so that the failure shows in the results, and latches bit 0 of
// *ctl.latch, which the host's next synchronising call reports.

// State loads of the persistent kernels: device-scope loads (default), or
// plain loads after an agent-scope acquire at every barrier (a tuning switch:
// the same results; per pass 24.5 against 25.6 us for the 32² forward passes,
// equal at 64², 8.70 against 8.62 us for the 32² adjoint passes).
#ifndef GLIDE_PERSISTENT_PLAIN_LOADS
#define GLIDE_PERSISTENT_PLAIN_LOADS 0
#endif
using LdPersistent = std::conditional_t<GLIDE_PERSISTENT_PLAIN_LOADS != 0, LdPlain, LdCoherent>;

__device__ __forceinline__ void grid_barrier(const PersistentCtl& ctl, unsigned long long target, bool fenced) {
    unsigned long long* sync = ctl.sync;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        if (fenced) __threadfence();
        atomicAdd(sync, 1ull);
        if (ctl.spin_limit < 0) atomicExch(sync + 1, 1ull);
        int spins = 0;
        while (__hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target &&
               !__hip_atomic_load(sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            if (++spins > ctl.spin_limit) {
                atomicExch(sync + 1, 1ull);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (fenced) __threadfence();
        else if (GLIDE_PERSISTENT_PLAIN_LOADS) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// The XCC (XCD) the calling wave runs on.
__device__ __forceinline__ unsigned xcc_id() {
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & 0xf;
}

// Workgroups of a launch are dispatched to the XCDs in turn, so the
// persistent kernels launch kXcds times the workgroups they need and work in
// every kXcds-th one only, which puts the working set on one XCD.
constexpr int kXcds = 8;

// Start of a persistent launch: the working workgroups record their XCDs in
// sync[2] and pass the first barrier (fenced); returns whether they all run
// on one XCD (the same answer in every workgroup).
__device__ __forceinline__ bool persistent_start(const PersistentCtl& ctl, int nb) {
    __shared__ int s_one;
    if (threadIdx.x == 0 && threadIdx.y == 0) atomicOr(ctl.sync + 2, 1ull << xcc_id());
    grid_barrier(ctl, (unsigned long long)nb, true);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        s_one = __popcll(__hip_atomic_load(ctl.sync + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 1;
    __syncthreads();
    return s_one != 0;
}

// End of a persistent launch: whether a barrier timed out (the same answer in
// every workgroup); the first workgroup latches the failure for the host.
__device__ __forceinline__ bool persistent_failed(const PersistentCtl& ctl) {
    __shared__ int s_fail;
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        s_fail = __hip_atomic_load(ctl.sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
        if (s_fail && blockIdx.x == 0 && ctl.latch) atomicOr(ctl.latch, 1u);
    }
    __syncthreads();
    return s_fail != 0;
}

__global__ __launch_bounds__(64) void k_zero_sync(unsigned long long* sync) {
    if (threadIdx.x < 3) sync[threadIdx.x] = 0ull;
}

// n_iter forward smoothing passes of a level at most TX cells wide in one
// launch (bitwise what n_iter k_vanka_fused launches give).  Working
// workgroup b owns the full rows i0 = b * RW .. i0 + RW - 1 (wave q < RW
// solves row i0 + q) and also solves row i0 - 1 (wave RW), whose bottom-facet
// increments its first row's top facets receive.  Pass p reads the state from
// a.u/v/H (p even) or the alternate buffers (p odd) and writes the other set;
// a barrier separates the passes.  If a barrier timed out, the last pass
// writes NaN thicknesses, which the caller's residual norms then show.
template <typename T, bool LINEAR, typename TJ, int RW>
__global__ __launch_bounds__(TX * (RW + 1)) void k_vanka_persistent(SsaArgs<T> a, T* alt_u, T* alt_v, T* alt_H,
                                                                    T* __restrict__ mask, VankaArgs<T> vk, T omega,
                                                                    int n_iter, PersistentCtl ctl) {
    using Ld = LdPersistent;
    constexpr int W = TX + 3;
    __shared__ T s_eta[RW + 3][W];  // cells (i0 - 2 .. i0 + RW, -2 .. TX)
    __shared__ T s_w1[RW][TX + 1];  // right-facet increments of the owned rows' cells (cell j at j + 1)
    __shared__ T s_w3[RW + 1][TX];  // bottom-facet increments of rows i0 + q (q < RW) and i0 - 1 (q = RW)
    if (blockIdx.x % kXcds) return;
    const int nb = gridDim.x / kXcds;
    const int ny = a.ny, nx = a.nx;
    const int i0 = (blockIdx.x / kXcds) * RW;
    const int wv = threadIdx.y, lane = threadIdx.x;
    const bool active = (wv < RW ? i0 + wv : i0 - 1) >= 0 && (wv < RW ? i0 + wv : i0 - 1) < ny && lane < nx;
    const bool own = wv < RW && active;
    const int above = wv == 0 ? RW : wv - 1;  // the wave that solved row ci - 1
    T* const base_u = const_cast<T*>(a.u);
    T* const base_v = const_cast<T*>(a.v);
    T* const base_H = const_cast<T*>(a.H);
    const bool fenced = !persistent_start(ctl, nb) || ctl.force_fenced;
#pragma unroll 1
    for (int p = 0; p < n_iter; ++p) {
        // the lane's indices pass through an empty asm statement every pass, so
        // that the compiler does not hoist the (pass-invariant) addresses of
        // the parameter loads out of the loop and hold them in registers
        int tid = wv * TX + lane, cj = lane, ci = wv < RW ? i0 + wv : i0 - 1;
        asm volatile("" : "+v"(tid), "+v"(cj), "+v"(ci));
        SsaArgs<T> s = a;
        T *u_dst = alt_u, *v_dst = alt_v, *H_dst = alt_H;
        if (p & 1) {
            s.u = alt_u;
            s.v = alt_v;
            s.H = alt_H;
            u_dst = base_u;
            v_dst = base_v;
            H_dst = base_H;
        }
        // the patches read columns -2 .. nx of the tile (all of it when nx = TX)
        const int wn = min(nx + 3, W);
        for (int k = tid; k < (RW + 3) * wn; k += (RW + 1) * TX) {
            const int r = k / wn, q = k % wn;
            const int ei = i0 - 2 + r, ej = q - 2;
            T e = T(0);
            if (ei >= 0 && ei < ny && ej >= 0 && ej < nx)
                e = cell_viscosity<T, Ld>(s.u, s.v, s.B, ei, ej, ny, nx, s.dxi, s.P);
            s_eta[r][q] = e;
        }
        __syncthreads();
        T w[5] = {T(0), T(0), T(0), T(0), T(0)};
        T masked = own ? Ld::get(mask + (long)ci * nx + cj) : T(0);
        int it = 0;
        if (active) {
            cell_newton<T, LINEAR, TJ, W, Ld>(s, s_eta, ci - i0 + 2, cj + 2, ci, cj, vk, true, w, masked, it);
            if (wv < RW) s_w1[wv][lane + 1] = w[1];
            s_w3[wv][lane] = w[3];
        }
        __syncthreads();
        if (own) {
            const int i = ci, j = cj;
            const long c = (long)i * nx + j, ou = (long)i * (nx + 1) + j;
            {
                T d = w[0];
                if (j > 0) d = d + s_w1[wv][lane];
                u_dst[ou] = Ld::get(s.u + ou) + omega * d;
                if (j == nx - 1) {
                    const T e = T(0) + w[1];
                    u_dst[ou + 1] = Ld::get(s.u + ou + 1) + omega * e;
                }
            }
            {
                T d = w[2];
                if (i > 0) d = d + s_w3[above][lane];
                v_dst[c] = Ld::get(s.v + c) + omega * d;
                if (i == ny - 1) {
                    const T e = T(0) + w[3];
                    v_dst[c + nx] = Ld::get(s.v + c + nx) + omega * e;
                }
            }
            H_dst[c] = Ld::get(s.H + c) + omega * w[4];
            mask[c] = masked;
        }
        if (p + 1 < n_iter) grid_barrier(ctl, (unsigned long long)nb * (p + 2), fenced);
    }
    if (persistent_failed(ctl) && own)
        (((n_iter - 1) & 1) ? base_H : alt_H)[(long)(i0 + wv) * nx + lane] = __longlong_as_double(0x7ff8000000000000ll);
}

// x += sign * omega * (sum of the cell increments touching each unknown).
// With phi != nullptr the grounded fraction of the next pass is computed from
// the updated thickness in the same launch (what k_grounded would give).
template <typename T>
__global__ __launch_bounds__(256) void k_vanka_apply(T* __restrict__ u, T* __restrict__ v, T* __restrict__ H,
                                                     const T* __restrict__ work, int ny, int nx, T omega,
                                                     T* __restrict__ phi = nullptr, const T* __restrict__ bed = nullptr,
                                                     T gc = T(0), T relax = T(0)) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j > nx) return;
    const long n = (long)ny * nx;
    if (i < ny) {
        T d = T(0);
        if (j < nx) d = work[(long)i * nx + j];
        if (j > 0) d = d + work[n + (long)i * nx + j - 1];
        u[(long)i * (nx + 1) + j] += omega * d;
    }
    if (j < nx) {
        T d = T(0);
        if (i < ny) d = work[2 * n + (long)i * nx + j];
        if (i > 0) d = d + work[3 * n + (long)(i - 1) * nx + j];
        v[(long)i * nx + j] += omega * d;
        if (i < ny) {
            const long c = (long)i * nx + j;
            const T h = H[c] + omega * work[4 * n + c];
            H[c] = h;
            if (phi) {
                T g = grounded_fraction(h, bed[c], gc);
                phi[c] = relax == T(0) ? g : (T(1) - relax) * g + relax * phi[c];
            }
        }
    }
}

// Local systems at the current state (debug / parity dump).
template <typename T>
__global__ __launch_bounds__(NT) void k_vanka_dump(SsaArgs<T> a, T* __restrict__ Jout, T* __restrict__ rout,
                                                   T calving_rate) {
    __shared__ T s_eta[TY + 2][TX + 2];
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * TX;
    const int tid = threadIdx.y * TX + threadIdx.x;
    fill_eta_tile(s_eta, a, i0, j0, tid);
    __syncthreads();
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    const int ny = a.ny, nx = a.nx;
    if (i >= ny || j >= nx) return;
    Patch<T> p;
    load_patch(p, a, s_eta, threadIdx.y + 1, threadIdx.x + 1, i, j, a.f_u, a.f_v, a.f_H, a.phi);
    const long c = (long)i * nx + j;
    T x[5] = {a.u[(long)i * (nx + 1) + j], a.u[(long)i * (nx + 1) + j + 1], a.v[c], a.v[c + nx], a.H[c]};
    T J[25], r[5];
    local_system(p, x, J, r, a.P, calving_rate, a.dxi, a.dt);
    r[0] -= p.f0; r[1] -= p.f1; r[2] -= p.f2; r[3] -= p.f3; r[4] -= p.f4;
    if (j == 0) { pin_row(J, 0); r[0] = x[0]; }
    if (j == nx - 1) { pin_row(J, 1); r[1] = x[1]; }
    if (i == 0) { pin_row(J, 2); r[2] = x[2]; }
    if (i == ny - 1) { pin_row(J, 3); r[3] = x[3]; }
    if ((x[4] - a.dt * r[4]) <= p.gam) { pin_row(J, 4); r[4] = x[4] - p.gam; }
    for (int k = 0; k < 25; ++k) Jout[25 * c + k] = J[k];
    for (int k = 0; k < 5; ++k) rout[5 * c + k] = r[k];
}

// Adjoint smoothing: J_locᵀ d = r_adj at the forward state, increments into
// the workspace planes (glide/cuda/vanka.cu:662-797).
template <typename T>
__global__ __launch_bounds__(NT) void k_vanka_adjoint_cell(SsaArgs<T> a, const T* __restrict__ ru,
                                                           const T* __restrict__ rv, const T* __restrict__ rH,
                                                           T* __restrict__ work, VankaArgs<T> vk) {
    __shared__ T s_eta[TY + 2][TX + 2];
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * TX;
    const int tid = threadIdx.y * TX + threadIdx.x;
    fill_eta_tile(s_eta, a, i0, j0, tid);
    __syncthreads();
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    const int ny = a.ny, nx = a.nx;
    if (i >= ny || j >= nx) return;
    Patch<T> p;
    load_patch(p, a, s_eta, threadIdx.y + 1, threadIdx.x + 1, i, j, (const T*)nullptr, (const T*)nullptr,
               (const T*)nullptr, a.phi);
    const long c = (long)i * nx + j;
    T x[5] = {a.u[(long)i * (nx + 1) + j], a.u[(long)i * (nx + 1) + j + 1], a.v[c], a.v[c + nx], a.H[c]};
    T J[25], r[5];
    local_system(p, x, J, r, a.P, vk.calving_rate, a.dxi, a.dt);
    J[0] -= vk.ssa_damping; J[6] -= vk.ssa_damping; J[12] -= vk.ssa_damping; J[18] -= vk.ssa_damping;
    J[24] += vk.mc_damping;
    T b[5] = {ru[(long)i * (nx + 1) + j], ru[(long)i * (nx + 1) + j + 1], rv[c], rv[c + nx], rH[c]};
    if (j == 0) { pin_row(J, 0); b[0] = T(0); }
    if (j == nx - 1) { pin_row(J, 1); b[1] = T(0); }
    if (i == 0) { pin_row(J, 2); b[2] = T(0); }
    if (i == ny - 1) { pin_row(J, 3); b[3] = T(0); }
    if (a.mask[c] > T(0.5)) { pin_row(J, 4); b[4] = T(0); }
    T JT[25];
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
        for (int s = 0; s < 5; ++s) JT[q * 5 + s] = J[s * 5 + q];
    T d[5];
    lu5_solve(JT, b, d);
    const long n = (long)ny * nx;
    work[c] = T(0.5) * d[0];
    work[n + c] = T(0.5) * d[1];
    work[2 * n + c] = T(0.5) * d[2];
    work[3 * n + c] = T(0.5) * d[3];
    work[4 * n + c] = d[4];
}

// Adjoint smoother cache.  The forward state is fixed during an adjoint
// solve, so each cell's transposed, damped and pinned patch Jacobian is the
// same in every adjoint pass: k_adjoint_factor factorises it once per level
// and state, and k_vanka_adjoint_cell_cached only substitutes.  Layout
// (planar, n = ny * nx cells): cache[q * n + c], q = 0..24, the in-place
// no-pivot Doolittle factors of J^T (row-major) with the reciprocal pivot on
// the diagonal, 0 for a pinned row (boundary facet or active-set cell), so
// that its increment is exactly zero as when its right-hand side is cleared;
// cache[25 * n + c] the cell viscosity, reused by the transpose operator.
template <typename T>
__device__ __forceinline__ void adjoint_factor_body(int bx, int by, const SsaArgs<T>& a, T* __restrict__ cache,
                                                    const VankaArgs<T>& vk) {
    __shared__ T s_eta[TY + 2][TX + 2];
    const int i0 = by * TY, j0 = bx * TX;
    const int tid = threadIdx.y * TX + threadIdx.x;
    fill_eta_tile(s_eta, a, i0, j0, tid);
    __syncthreads();
    const int i = i0 + threadIdx.y, j = j0 + threadIdx.x;
    const int ny = a.ny, nx = a.nx;
    if (i >= ny || j >= nx) return;
    Patch<T> p;
    load_patch(p, a, s_eta, threadIdx.y + 1, threadIdx.x + 1, i, j, (const T*)nullptr, (const T*)nullptr,
               (const T*)nullptr, a.phi);
    const long c = (long)i * nx + j;
    T x[5] = {a.u[(long)i * (nx + 1) + j], a.u[(long)i * (nx + 1) + j + 1], a.v[c], a.v[c + nx], a.H[c]};
    T J[25], r[5];
    local_system(p, x, J, r, a.P, vk.calving_rate, a.dxi, a.dt);
    J[0] -= vk.ssa_damping; J[6] -= vk.ssa_damping; J[12] -= vk.ssa_damping; J[18] -= vk.ssa_damping;
    J[24] += vk.mc_damping;
    const bool pin[5] = {j == 0, j == nx - 1, i == 0, i == ny - 1, a.mask[c] > T(0.5)};
#pragma unroll
    for (int k = 0; k < 5; ++k)
        if (pin[k]) pin_row(J, k);
    T A[25];
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
        for (int s = 0; s < 5; ++s) A[q * 5 + s] = J[s * 5 + q];
    T inv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        inv[k] = T(1) / A[k * 5 + k];
#pragma unroll
        for (int q = k + 1; q < 5; ++q) {
            A[q * 5 + k] *= inv[k];
#pragma unroll
            for (int s = k + 1; s < 5; ++s) A[q * 5 + s] -= A[q * 5 + k] * A[k * 5 + s];
        }
    }
    const long n = (long)ny * nx;
#pragma unroll
    for (int k = 0; k < 5; ++k) A[k * 6] = pin[k] ? T(0) : inv[k];
#pragma unroll
    for (int q = 0; q < 25; ++q) cache[q * n + c] = A[q];
    cache[25 * n + c] = s_eta[threadIdx.y + 1][threadIdx.x + 1];
}

template <typename T>
__global__ __launch_bounds__(NT) void k_adjoint_factor(SsaArgs<T> a, T* __restrict__ cache, VankaArgs<T> vk) {
    adjoint_factor_body<T>(blockIdx.x, blockIdx.y, a, cache, vk);
}

// Increments of a cell from its 25 cached factors A (in registers) and the
// residual b: w = (d0/2, d1/2, d2/2, d3/2, d4) with J^T d = b (half-weighted
// facets); adjoint_cached_solve loads the factors of cell c first.
template <typename T>
__device__ __forceinline__ void adjoint_solve_factors(const T A[25], const T b[5], T w[5]) {
    T y[5], d[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        T s = b[q];
#pragma unroll
        for (int k = 0; k < q; ++k) s -= A[q * 5 + k] * y[k];
        y[q] = s;
    }
#pragma unroll
    for (int q = 4; q >= 0; --q) {
        T s = y[q];
#pragma unroll
        for (int k = q + 1; k < 5; ++k) s -= A[q * 5 + k] * d[k];
        d[q] = s * A[q * 6];
    }
    w[0] = T(0.5) * d[0];
    w[1] = T(0.5) * d[1];
    w[2] = T(0.5) * d[2];
    w[3] = T(0.5) * d[3];
    w[4] = d[4];
}
template <typename T>
__device__ __forceinline__ void adjoint_cached_solve(const T* __restrict__ cache, long n, long c, const T b[5],
                                                     T w[5]) {
    T A[25];
#pragma unroll
    for (int q = 0; q < 25; ++q) A[q] = cache[q * n + c];
    adjoint_solve_factors(A, b, w);
}
template <typename T>
__device__ __forceinline__ void adjoint_cached_increments(const T* __restrict__ cache, const T* __restrict__ ru,
                                                          const T* __restrict__ rv, const T* __restrict__ rH, int i,
                                                          int j, int ny, int nx, T w[5]) {
    const long n = (long)ny * nx, c = (long)i * nx + j;
    const T b[5] = {ru[(long)i * (nx + 1) + j], ru[(long)i * (nx + 1) + j + 1], rv[c], rv[c + nx], rH[c]};
    adjoint_cached_solve(cache, n, c, b, w);
}

// Adjoint patch solve from the cached factors: the increments of
// k_vanka_adjoint_cell into the workspace planes.
template <typename T>
__global__ __launch_bounds__(NT) void k_vanka_adjoint_cell_cached(const T* __restrict__ cache,
                                                                  const T* __restrict__ ru, const T* __restrict__ rv,
                                                                  const T* __restrict__ rH, T* __restrict__ work,
                                                                  int ny, int nx) {
    const int i = blockIdx.y * TY + threadIdx.y, j = blockIdx.x * TX + threadIdx.x;
    if (i >= ny || j >= nx) return;
    const long n = (long)ny * nx, c = (long)i * nx + j;
    T w[5];
    adjoint_cached_increments(cache, ru, rv, rH, i, j, ny, nx, w);
#pragma unroll
    for (int q = 0; q < 5; ++q) work[q * n + c] = w[q];
}

// A whole adjoint smoothing pass after the transpose operator in one launch:
// the cached patch solves of the tile's cells and of the cells left of and
// above the tile (whose right / bottom facet increments the tile's facets
// also receive), then lambda += omega * (sum of the increments) on the facets
// and cells the tile owns, with the arithmetic of k_vanka_apply.  Each
// multiplier is read and written by its owner only, so the update is in place.
template <typename T>
__global__ __launch_bounds__(NT) void k_vanka_adjoint_fused(const T* __restrict__ cache, const T* __restrict__ ru,
                                                            const T* __restrict__ rv, const T* __restrict__ rH,
                                                            T* __restrict__ lu, T* __restrict__ lv,
                                                            T* __restrict__ lH, int ny, int nx, T omega) {
    __shared__ T s_w1[TY][TX + 1];  // right-facet increment of cells (i0 + a, j0 - 1 + b)
    __shared__ T s_w3[TY + 1][TX];  // bottom-facet increment of cells (i0 - 1 + a, j0 + b)
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * TX;
    const int ty = threadIdx.y, tx = threadIdx.x, tid = ty * TX + tx;
    const int i = i0 + ty, j = j0 + tx;
    const bool own = i < ny && j < nx;
    T w[5] = {T(0), T(0), T(0), T(0), T(0)};
    if (own) {
        adjoint_cached_increments(cache, ru, rv, rH, i, j, ny, nx, w);
        s_w1[ty][tx + 1] = w[1];
        s_w3[ty + 1][tx] = w[3];
    }
    if (tid < TY) {
        const int ci = i0 + tid, cj = j0 - 1;
        if (cj >= 0 && ci < ny) {
            T wr[5];
            adjoint_cached_increments(cache, ru, rv, rH, ci, cj, ny, nx, wr);
            s_w1[tid][0] = wr[1];
        }
    } else if (tid < TY + TX) {
        const int ci = i0 - 1, cj = j0 + (tid - TY);
        if (ci >= 0 && cj < nx) {
            T wr[5];
            adjoint_cached_increments(cache, ru, rv, rH, ci, cj, ny, nx, wr);
            s_w3[0][tid - TY] = wr[3];
        }
    }
    __syncthreads();
    if (!own) return;
    const long c = (long)i * nx + j;
    {
        T d = w[0];
        if (j > 0) d = d + s_w1[ty][tx];
        lu[(long)i * (nx + 1) + j] += omega * d;
        if (j == nx - 1) {
            T e = T(0) + w[1];
            lu[(long)i * (nx + 1) + nx] += omega * e;
        }
    }
    {
        T d = w[2];
        if (i > 0) d = d + s_w3[ty][tx];
        lv[c] += omega * d;
        if (i == ny - 1) {
            T e = T(0) + w[3];
            lv[(long)ny * nx + j] += omega * e;
        }
    }
    lH[c] += omega * w[4];
}

// ------------------------------------------------------------------ transpose operator
// Jᵀλ (minus the adjoint forcing when use_forcing) as a gather.  Phase 1
// computes, for the cells of the tile plus a one-cell ring, the masked
// thickness multiplier, viscosity, its derivative along λ and the normal
// stress rows applied to λ; phase 2 the shear-stress rows of the vertices and
// the drag cross-terms of the facets; phase 3 gives every output its sum in a
// fixed order.  Velocity rows use the symmetry of the velocity block as the
// reference does (glide/cuda/residuals.cu:797-817); Dirichlet facets and
// masked cells return zero (residuals.cu:1122-1160).
template <typename T, int ROWS>
struct VjpTile {
    T lHm[ROWS + 2][TX + 2];
    T eta[ROWS + 2][TX + 2];
    T etad[ROWS + 2][TX + 2];
    T Hs[ROWS + 2][TX + 2];
    T Lxx[ROWS + 2][TX + 2];
    T Lyy[ROWS + 2][TX + 2];
    T Kxx[ROWS + 2][TX + 2];   // d sigma_xx / d(eta H) at the cell
    T Kyy[ROWS + 2][TX + 2];
    T Lxy[ROWS + 1][TX + 1];   // vertices [i0, i0+ROWS] x [j0, j0+TX]
    T Gxy[ROWS + 1][TX + 1];   // d sigma_xy / d(eta H) * 0.25 * (λ combination) * dxi
    T Pu[4][ROWS + 1][TX + 1]; // λu * drag cross-derivatives, u facets rows [i0-1, i0+ROWS-1], cols [j0, j0+TX]
    T Pv[4][ROWS + 1][TX + 1]; // λv * drag cross-derivatives, v facets rows [i0, i0+ROWS], cols [j0-1, j0+TX-1]
};

// The transpose operator for the outputs (facets and cells) i0 <= i < i0 +
// rows, j0 <= j < j0 + cols inside the grid, by nthr threads (tid = 0 ..
// nthr - 1) of one block, through the block storage S (cell arrays over the
// one-cell ring, vertex arrays, and with S::kDrag the drag cross-terms;
// without it the drag must be linear, whose cross-terms vanish), each output
// handed to Out with its tile position and global index.
template <typename T, typename S, typename Out, class Ld = LdPlain>
__device__ __forceinline__ void vjp_body(const SsaArgs<T>& a, const T* __restrict__ lu, const T* __restrict__ lv,
                                         const T* __restrict__ lH, const T* __restrict__ fu, const T* __restrict__ fv,
                                         const T* __restrict__ fH, Out& out, T calving_rate,
                                         const T* __restrict__ eta_cache, int i0, int j0, int rows, int cols, int tid,
                                         int nthr, S& s) {
    const int ny = a.ny, nx = a.nx;
    const T dxi = a.dxi;
    const T* __restrict__ u = a.u;
    const T* __restrict__ v = a.v;
    const T* __restrict__ H = a.H;

    // phase 1: cells
    for (int k = tid; k < (rows + 2) * (cols + 2); k += nthr) {
        const int ri = k / (cols + 2), rj = k % (cols + 2);
        const int ci = i0 - 1 + ri, cj = j0 - 1 + rj;
        T lHm = 0, e = 0, ed = 0, Hh = 0, Lx = 0, Ly = 0, Kx = 0, Ky = 0;
        if (ci >= 0 && ci < ny && cj >= 0 && cj < nx) {
            const long c = (long)ci * nx + cj;
            CellVel<T> cv = load_cell_vel(u, v, ci, cj, ny, nx);
            CellVel<T> cl = load_cell_vel<T, Ld>(lu, lv, ci, cj, ny, nx);
            T e2 = strain_invariant(cv, ci, cj, ny, nx, dxi, a.P.eps_reg);
            T de2 = strain_invariant_dir(cv, cl, ci, cj, ny, nx, dxi);
            e = eta_cache ? eta_cache[c] : T(0.5) * a.B[c] * glen_power(e2, a.P);
            ed = e * a.P.glen_p * de2 / e2;
            Hh = H[c];
            lHm = Ld::get(lH + c) * (T(1) - a.mask[c]);
            NormalStress<T> sx = sxx(cv.ul, cv.ur, cv.vt, cv.vb, e * Hh, dxi, true);
            Lx = sx.d_ul * cl.ul + sx.d_ur * cl.ur + sx.d_vt * cl.vt + sx.d_vb * cl.vb + sx.d_eH * (Hh * ed);
            Kx = sx.d_eH;
            NormalStress<T> sy = syy(cv.ul, cv.ur, cv.vt, cv.vb, e * Hh, dxi, true);
            Ly = sy.d_ul * cl.ul + sy.d_ur * cl.ur + sy.d_vt * cl.vt + sy.d_vb * cl.vb + sy.d_eH * (Hh * ed);
            Ky = sy.d_eH;
        }
        s.lHm(ri, rj) = lHm; s.eta(ri, rj) = e; s.etad(ri, rj) = ed; s.Hs(ri, rj) = Hh;
        s.Lxx(ri, rj) = Lx; s.Lyy(ri, rj) = Ly; s.Kxx(ri, rj) = Kx; s.Kyy(ri, rj) = Ky;
    }
    __syncthreads();
    // phase 2a: vertices (vi, vj) in [i0, i0+rows] x [j0, j0+cols]
    for (int k = tid; k < (rows + 1) * (cols + 1); k += nthr) {
        const int ri = k / (cols + 1), rj = k % (cols + 1);
        const int vi = i0 + ri, vj = j0 + rj;
        T L = 0, G = 0;
        if (vi > 0 && vi < ny && vj > 0 && vj < nx) {
            // cells (vi-1,vj-1), (vi-1,vj), (vi,vj-1), (vi,vj) at tile rows ri, ri+1 and cols rj, rj+1
            T eH = T(0.25) * (s.eta(ri, rj) * s.Hs(ri, rj) + s.eta(ri, rj + 1) * s.Hs(ri, rj + 1) +
                              s.eta(ri + 1, rj) * s.Hs(ri + 1, rj) + s.eta(ri + 1, rj + 1) * s.Hs(ri + 1, rj + 1));
            T eHd = T(0.25) * (s.etad(ri, rj) * s.Hs(ri, rj) + s.etad(ri, rj + 1) * s.Hs(ri, rj + 1) +
                               s.etad(ri + 1, rj) * s.Hs(ri + 1, rj) + s.etad(ri + 1, rj + 1) * s.Hs(ri + 1, rj + 1));
            const long ut = (long)(vi - 1) * (nx + 1) + vj, ub = (long)vi * (nx + 1) + vj;
            const long vl = (long)vi * nx + vj - 1, vr = (long)vi * nx + vj;
            ShearStress<T> q = sxy(u[ut], u[ub], v[vl], v[vr], eH, dxi, true);
            const T lut = Ld::get(lu + ut), lub = Ld::get(lu + ub), lvl = Ld::get(lv + vl), lvr = Ld::get(lv + vr);
            L = q.d_ut * lut + q.d_ub * lub + q.d_vl * lvl + q.d_vr * lvr + q.d_eH * eHd;
            G = q.d_eH * T(0.25) * (lub - lut + lvl - lvr) * dxi;
        }
        s.Lxy(ri, rj) = L; s.Gxy(ri, rj) = G;
    }
    // phase 2b: drag cross-terms of u facets (rows i0-1.., cols j0..) and v facets (rows i0.., cols j0-1..)
    for (int k = tid; S::kDrag && k < (rows + 1) * (cols + 1); k += nthr) {
        const int ri = k / (cols + 1), rj = k % (cols + 1);
        {
            const int fi = i0 - 1 + ri, fj = j0 + rj;
            T q0 = 0, q1 = 0, q2 = 0, q3 = 0;
            if (fi >= 0 && fi < ny && fj >= 0 && fj <= nx && !a.P.linear_drag) {
                Drag<T> t = basal_drag(u[(long)fi * (nx + 1) + fj], v[offV(fi, fj - 1, ny, nx)], v[offV(fi, fj, ny, nx)],
                                       v[offV(fi + 1, fj - 1, ny, nx)], v[offV(fi + 1, fj, ny, nx)],
                                       a.phi[offC(fi, fj - 1, ny, nx)], a.phi[offC(fi, fj, ny, nx)],
                                       a.beta[offC(fi, fj - 1, ny, nx)], a.beta[offC(fi, fj, ny, nx)], a.P);
                T l = Ld::get(lu + (long)fi * (nx + 1) + fj);
                q0 = l * t.c0; q1 = l * t.c1; q2 = l * t.c2; q3 = l * t.c3;
            }
            s.Pu(0, ri, rj) = q0; s.Pu(1, ri, rj) = q1; s.Pu(2, ri, rj) = q2; s.Pu(3, ri, rj) = q3;
        }
        {
            const int fi = i0 + ri, fj = j0 - 1 + rj;
            T q0 = 0, q1 = 0, q2 = 0, q3 = 0;
            if (fi >= 0 && fi <= ny && fj >= 0 && fj < nx && !a.P.linear_drag) {
                Drag<T> t = basal_drag(v[(long)fi * nx + fj], u[offU(fi - 1, fj, ny, nx)], u[offU(fi - 1, fj + 1, ny, nx)],
                                       u[offU(fi, fj, ny, nx)], u[offU(fi, fj + 1, ny, nx)],
                                       a.phi[offC(fi - 1, fj, ny, nx)], a.phi[offC(fi, fj, ny, nx)],
                                       a.beta[offC(fi - 1, fj, ny, nx)], a.beta[offC(fi, fj, ny, nx)], a.P);
                T l = Ld::get(lv + (long)fi * nx + fj);
                q0 = l * t.c0; q1 = l * t.c1; q2 = l * t.c2; q3 = l * t.c3;
            }
            s.Pv(0, ri, rj) = q0; s.Pv(1, ri, rj) = q1; s.Pv(2, ri, rj) = q2; s.Pv(3, ri, rj) = q3;
        }
    }
    __syncthreads();

    // phase 3: outputs
    for (int k = tid; k < rows * cols; k += nthr) {
        const int ti = k / cols, tj = k % cols;
        const int i = i0 + ti, j = j0 + tj;
        if (i < 0 || j < 0 || i > ny || j > nx) continue;
        const int ci = ti + 1, cj = tj + 1;  // cell (i,j) in the ring tile
        const int vi = ti, vj = tj;          // vertex (i,j) in the vertex tile
        if (i < ny) {  // u facet (i,j)
            const long o = (long)i * (nx + 1) + j;
            T acc = T(0);
            if (j > 0 && j < nx) {
                acc = (s.Lxx(ci, cj) - s.Lxx(ci, cj - 1)) * dxi + (s.Lxy(vi, vj) - s.Lxy(vi + 1, vj)) * dxi;
                const long cc = (long)i * nx + j;
                T pl = a.phi[cc - 1], pc = a.phi[cc];
                Drag<T> t = basal_drag(u[o], v[cc - 1], v[cc], v[cc + nx - 1], v[cc + nx], pl, pc, a.beta[cc - 1],
                                       a.beta[cc], a.P);
                acc += Ld::get(lu + o) * t.d_w;
                // cross terms from v facets (i,j) [ubl], (i,j-1) [ubr], (i+1,j) [utl], (i+1,j-1) [utr]
                if (S::kDrag)
                    acc += s.Pv(2, vi, cj) + s.Pv(3, vi, cj - 1) + s.Pv(0, vi + 1, cj) + s.Pv(1, vi + 1, cj - 1);
                Flux<T> f = flux_x(u[o], H[cc - 1], H[cc], true);
                acc += f.d_w * (s.lHm(ci, cj - 1) - s.lHm(ci, cj)) * dxi;
                if (fu) acc -= fu[o];
            }
            out.put_u(ti, tj, o, acc);
        }
        if (j < nx) {  // v facet (i,j)
            const long o = (long)i * nx + j;
            T acc = T(0);
            if (i > 0 && i < ny) {
                acc = (s.Lyy(ci - 1, cj) - s.Lyy(ci, cj)) * dxi + (s.Lxy(vi, vj + 1) - s.Lxy(vi, vj)) * dxi;
                const long ul = (long)i * (nx + 1) + j, utl = ul - (nx + 1);
                T pt = a.phi[o - nx], pc = a.phi[o];
                Drag<T> t = basal_drag(v[o], u[utl], u[utl + 1], u[ul], u[ul + 1], pt, pc, a.beta[o - nx], a.beta[o], a.P);
                acc += Ld::get(lv + o) * t.d_w;
                // cross terms from u facets (i,j) [vtr], (i,j+1) [vtl], (i-1,j) [vbr], (i-1,j+1) [vbl]
                if (S::kDrag)
                    acc += s.Pu(1, vi + 1, vj) + s.Pu(0, vi + 1, vj + 1) + s.Pu(3, vi, vj) + s.Pu(2, vi, vj + 1);
                Flux<T> f = flux_y(v[o], H[o - nx], H[o], true);
                acc += f.d_w * (s.lHm(ci, cj) - s.lHm(ci - 1, cj)) * dxi;
                if (fv) acc -= fv[o];
            }
            out.put_v(ti, tj, o, acc);
        }
        if (i < ny && j < nx) {  // cell (i,j)
            const long c = (long)i * nx + j;
            T acc = T(0);
            if (!(a.mask[c] > T(0))) {
                const T lc = s.lHm(ci, cj);
                const T Hc = s.Hs(ci, cj), ec = s.eta(ci, cj);
                const long ol = (long)i * (nx + 1) + j;
                acc = lc / a.dt;
                // thickness fluxes through the four faces
                {
                    Flux<T> f = flux_x(u[ol], H[offC(i, j - 1, ny, nx)], Hc, j > 0);
                    acc += (s.lHm(ci, cj - 1) - lc) * f.d_HB * dxi;
                }
                {
                    Flux<T> f = flux_x(u[ol + 1], Hc, H[offC(i, j + 1, ny, nx)], j + 1 < nx);
                    acc += (lc - s.lHm(ci, cj + 1)) * f.d_HA * dxi;
                }
                {
                    Flux<T> f = flux_y(v[c], H[offC(i - 1, j, ny, nx)], Hc, i > 0);
                    acc += (lc - s.lHm(ci - 1, cj)) * f.d_HB * dxi;
                }
                {
                    Flux<T> f = flux_y(v[c + nx], Hc, H[offC(i + 1, j, ny, nx)], i + 1 < ny);
                    acc += (s.lHm(ci + 1, cj) - lc) * f.d_HA * dxi;
                }
                if (calving_rate != T(0)) {
                    const T pc = a.phi[c];
                    T k = (T(1) - pc) * ((T(1) - a.phi[offC(i, j - 1, ny, nx)]) + (T(1) - a.phi[offC(i, j + 1, ny, nx)]) +
                                         (T(1) - a.phi[offC(i - 1, j, ny, nx)]) + (T(1) - a.phi[offC(i + 1, j, ny, nx)]));
                    acc += lc * k * calving_rate * dxi;
                }
                // viscosity-thickness products of the cell's normal stresses
                acc += s.Kxx(ci, cj) * ec * (Ld::get(lu + ol) - Ld::get(lu + ol + 1)) * dxi;
                acc += s.Kyy(ci, cj) * ec * (Ld::get(lv + c + nx) - Ld::get(lv + c)) * dxi;
                // four corner vertices
                acc += (s.Gxy(vi, vj) + s.Gxy(vi, vj + 1) + s.Gxy(vi + 1, vj) + s.Gxy(vi + 1, vj + 1)) * ec;
                // driving stress on the four faces
                {
                    const T pc = a.phi[c], zc = a.bed[c];
                    if (j > 0) {
                        Drive<T> d = drive_x(H[c - 1], Hc, a.bed[c - 1], zc, a.phi[c - 1], pc, dxi, true);
                        acc -= Ld::get(lu + ol) * d.d_HB;
                    }
                    if (j + 1 < nx) {
                        Drive<T> d = drive_x(Hc, H[c + 1], zc, a.bed[c + 1], pc, a.phi[c + 1], dxi, true);
                        acc -= Ld::get(lu + ol + 1) * d.d_HA;
                    }
                    if (i > 0) {
                        Drive<T> d = drive_y(H[c - nx], Hc, a.bed[c - nx], zc, a.phi[c - nx], pc, dxi, true);
                        acc -= Ld::get(lv + c) * d.d_HB;
                    }
                    if (i + 1 < ny) {
                        Drive<T> d = drive_y(Hc, H[c + nx], zc, a.bed[c + nx], pc, a.phi[c + nx], dxi, true);
                        acc -= Ld::get(lv + c + nx) * d.d_HA;
                    }
                }
                if (fH) acc -= fH[c];
            }
            out.put_h(ti, tj, c, acc);
        }
    }
}

// Block storage of the transpose operator: a static tile of k_vjp, and a
// whole level carved from dynamic LDS (linear drag only).
template <typename T, int ROWS>
struct VjpTileStore {
    static constexpr bool kDrag = true;
    VjpTile<T, ROWS>& t;
    __device__ T& lHm(int r, int c) { return t.lHm[r][c]; }
    __device__ T& eta(int r, int c) { return t.eta[r][c]; }
    __device__ T& etad(int r, int c) { return t.etad[r][c]; }
    __device__ T& Hs(int r, int c) { return t.Hs[r][c]; }
    __device__ T& Lxx(int r, int c) { return t.Lxx[r][c]; }
    __device__ T& Lyy(int r, int c) { return t.Lyy[r][c]; }
    __device__ T& Kxx(int r, int c) { return t.Kxx[r][c]; }
    __device__ T& Kyy(int r, int c) { return t.Kyy[r][c]; }
    __device__ T& Lxy(int r, int c) { return t.Lxy[r][c]; }
    __device__ T& Gxy(int r, int c) { return t.Gxy[r][c]; }
    __device__ T& Pu(int q, int r, int c) { return t.Pu[q][r][c]; }
    __device__ T& Pv(int q, int r, int c) { return t.Pv[q][r][c]; }
};
template <typename T>
struct VjpLevelStore {
    static constexpr bool kDrag = false;
    T* cell;  // 8 arrays of (rows + 2) x (cols + 2)
    T* vert;  // 2 arrays of (rows + 1) x (cols + 1)
    int cw, nc, vw, nv;
    __device__ T& lHm(int r, int c) { return cell[r * cw + c]; }
    __device__ T& eta(int r, int c) { return cell[nc + r * cw + c]; }
    __device__ T& etad(int r, int c) { return cell[2 * nc + r * cw + c]; }
    __device__ T& Hs(int r, int c) { return cell[3 * nc + r * cw + c]; }
    __device__ T& Lxx(int r, int c) { return cell[4 * nc + r * cw + c]; }
    __device__ T& Lyy(int r, int c) { return cell[5 * nc + r * cw + c]; }
    __device__ T& Kxx(int r, int c) { return cell[6 * nc + r * cw + c]; }
    __device__ T& Kyy(int r, int c) { return cell[7 * nc + r * cw + c]; }
    __device__ T& Lxy(int r, int c) { return vert[r * vw + c]; }
    __device__ T& Gxy(int r, int c) { return vert[nv + r * vw + c]; }
    __device__ T& Pu(int, int, int) { return cell[0]; }  // never used without kDrag
    __device__ T& Pv(int, int, int) { return cell[0]; }
};

// Output sinks of vjp_body: arrays in the grid's own layouts, or tile-local
// LDS arrays indexed by the output's position in the tile.
template <typename T>
struct VjpGridOut {
    T *u, *v, *h;
    __device__ void put_u(int, int, long o, T x) { u[o] = x; }
    __device__ void put_v(int, int, long o, T x) { v[o] = x; }
    __device__ void put_h(int, int, long c, T x) { h[c] = x; }
};
template <typename T, int RC>
struct VjpTileOut {
    T (*u)[RC];
    T (*v)[RC];
    T (*h)[RC];
    __device__ void put_u(int ti, int tj, long, T x) { u[ti][tj] = x; }
    __device__ void put_v(int ti, int tj, long, T x) { v[ti][tj] = x; }
    __device__ void put_h(int ti, int tj, long, T x) { h[ti][tj] = x; }
};

template <typename T, int ROWS>
__global__ __launch_bounds__(TX * ROWS) void k_vjp(SsaArgs<T> a, const T* __restrict__ lu, const T* __restrict__ lv,
                                            const T* __restrict__ lH, const T* __restrict__ fu,
                                            const T* __restrict__ fv, const T* __restrict__ fH, T* __restrict__ out_u,
                                            T* __restrict__ out_v, T* __restrict__ out_H, T calving_rate,
                                            const T* __restrict__ eta_cache) {
    __shared__ VjpTile<T, ROWS> t;
    VjpTileStore<T, ROWS> s{t};
    VjpGridOut<T> o{out_u, out_v, out_H};
    vjp_body(a, lu, lv, lH, fu, fv, fH, o, calving_rate, eta_cache, (int)blockIdx.y * ROWS, (int)blockIdx.x * TX,
             ROWS, TX, (int)(threadIdx.y * TX + threadIdx.x), TX * ROWS, s);
}

// A whole cached adjoint pass in one launch (linear drag): the transpose
// operator minus the forcing for the facets and cells of the tile and of its
// left column and top row (vjp_body over rows i0 - 1 .. i0 + TY and columns
// j0 - 1 .. j0 + TX, into LDS), the cached patch solves of the tile's cells
// and of the cells left of and above it, then the in-place update of
// k_vanka_adjoint_fused, written to the other of two multiplier buffers:
// the transpose operator reads the multipliers of neighbouring tiles, which
// their blocks update during the same pass.  The same arithmetic as k_vjp
// followed by k_vanka_adjoint_fused; it recomputes the ring's residuals (1.4x the
// transpose operator's work) and saves a launch, the residual planes' round
// trip through HBM and the second kernel's ramp: 9.7 against 12.6 us per
// pass at 32², 14.6 against 19.8 us at 256², and still ahead at 512².
// The one-launch adjoint pass uses TY-row tiles (2 TY thread rows, so that the
// ring's patch solves run beside the tile's).  Each tile recomputes the
// transpose operator on its top row and left column, (TY + 2) / TY of the
// work; 8-row tiles (1024 threads, 88 KB of LDS) measured 35.2 against
// 38.6 us per pass at 512² alone, but the coupled step, whose enthalpy
// kernels share the GPU with these passes, ran 0.75 ms slower with them.

// One adjoint pass over the TR x TX tile at (i0, j0) with TX x 2 TR threads:
// k_vjp's body over the tile plus its left column and top row (into LDS), the
// cached patch solves of the tile (thread rows < TR) and of the ring (the
// other threads), then the update of the tile's multipliers into the
// destination buffers.
// MC: one colour phase of the multicolour adjoint pass instead: only cells
// with (i + j) % 2 == colour solve, and each unknown receives its (single)
// solving cell's full increment times omega.
template <typename T, class Ld, int TR, bool MC = false>
__device__ __forceinline__ void adjoint_pass_tile(const SsaArgs<T>& a, const T* __restrict__ cache,
                                                  const T* __restrict__ lu, const T* __restrict__ lv,
                                                  const T* __restrict__ lH, T* __restrict__ lu_dst,
                                                  T* __restrict__ lv_dst, T* __restrict__ lH_dst,
                                                  const T* __restrict__ fu, const T* __restrict__ fv,
                                                  const T* __restrict__ fH, T calving_rate, T omega, int i0, int j0,
                                                  int ty, int tx, int colour = -1) {
    constexpr int RR = TR + 2, RC = TX + 2;
    constexpr int CW = RC + 2, NCELL = (RR + 2) * CW, VW = RC + 1, NVERT = (RR + 1) * VW;
    __shared__ T s_cell[8 * NCELL];
    __shared__ T s_vert[2 * NVERT];
    __shared__ T r_u[RR][RC], r_v[RR][RC], r_h[RR][RC];
    __shared__ T s_w1[TR][TX + 1];  // right-facet increment of cells (i0 + a, j0 - 1 + b)
    __shared__ T s_w3[TR + 1][TX];  // bottom-facet increment of cells (i0 - 1 + a, j0 + b)
    VjpLevelStore<T> store;
    store.cell = s_cell;
    store.vert = s_vert;
    store.cw = CW;
    store.nc = NCELL;
    store.vw = VW;
    store.nv = NVERT;
    VjpTileOut<T, RC> out{r_u, r_v, r_h};
    const int ny = a.ny, nx = a.nx;
    const long n = (long)ny * nx;
    const int tid = ty * TX + tx;
    constexpr int NTF = TX * 2 * TR;
    vjp_body<T, VjpLevelStore<T>, VjpTileOut<T, RC>, Ld>(a, lu, lv, lH, fu, fv, fH, out, calving_rate,
                                                          cache + (long)kAdjointCacheEta * n, i0 - 1, j0 - 1, RR, RC,
                                                          tid, NTF, store);
    __syncthreads();
    auto increments = [&](int ci, int cj, T w[5]) {
        if constexpr (MC) {
            if (((ci + cj) & 1) != colour) {
#pragma unroll
                for (int q = 0; q < 5; ++q) w[q] = T(0);
                return;
            }
        }
        const int ti = ci - i0 + 1, tj = cj - j0 + 1;
        const T b[5] = {r_u[ti][tj], r_u[ti][tj + 1], r_v[ti][tj], r_v[ti + 1][tj], r_h[ti][tj]};
        adjoint_cached_solve(cache, n, (long)ci * nx + cj, b, w);
        if constexpr (MC) {
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = T(2) * w[q];  // full facet weight
        }
    };
    // thread rows 0 .. TR-1 own the tile's cells, the rest solve the ring
    const int i = i0 + ty, j = j0 + tx;
    const bool own = ty < TR && i < ny && j < nx;
    T w[5] = {T(0), T(0), T(0), T(0), T(0)};
    if (own) {
        increments(i, j, w);
        s_w1[ty][tx + 1] = w[1];
        s_w3[ty + 1][tx] = w[3];
    }
    const int rt = tid - TR * TX;
    if (rt >= 0 && rt < TR) {
        const int ci = i0 + rt, cj = j0 - 1;
        if (cj >= 0 && ci < ny) {
            T wr[5];
            increments(ci, cj, wr);
            s_w1[rt][0] = wr[1];
        }
    } else if (rt >= TR && rt < TR + TX) {
        const int ci = i0 - 1, cj = j0 + (rt - TR);
        if (ci >= 0 && cj < nx) {
            T wr[5];
            increments(ci, cj, wr);
            s_w3[0][rt - TR] = wr[3];
        }
    }
    __syncthreads();
    if (!own) return;
    const long c = (long)i * nx + j;
    {
        T d = w[0];
        if (j > 0) d = d + s_w1[ty][tx];
        const long o = (long)i * (nx + 1) + j;
        lu_dst[o] = Ld::get(lu + o) + omega * d;
        if (j == nx - 1) {
            T e = T(0) + w[1];
            lu_dst[o + 1] = Ld::get(lu + o + 1) + omega * e;
        }
    }
    {
        T d = w[2];
        if (i > 0) d = d + s_w3[ty][tx];
        lv_dst[c] = Ld::get(lv + c) + omega * d;
        if (i == ny - 1) {
            T e = T(0) + w[3];
            lv_dst[c + nx] = Ld::get(lv + c + nx) + omega * e;
        }
    }
    lH_dst[c] = Ld::get(lH + c) + omega * w[4];
}

template <typename T, int TR, bool MC = false>
__global__ __launch_bounds__(TX * 2 * TR) void k_adjoint_pass_fused(SsaArgs<T> a, const T* __restrict__ cache,
                                                                           const T* __restrict__ lu,
                                                                           const T* __restrict__ lv,
                                                                           const T* __restrict__ lH,
                                                                           T* __restrict__ lu_dst,
                                                                           T* __restrict__ lv_dst,
                                                                           T* __restrict__ lH_dst,
                                                                           const T* __restrict__ fu,
                                                                           const T* __restrict__ fv,
                                                                           const T* __restrict__ fH, T calving_rate,
                                                                           T omega, int colour) {
    adjoint_pass_tile<T, LdPlain, TR, MC>(a, cache, lu, lv, lH, lu_dst, lv_dst, lH_dst, fu, fv, fH, calving_rate,
                                          omega, blockIdx.y * TR, blockIdx.x * TX, threadIdx.y, threadIdx.x, colour);
}

// n_iter adjoint passes of a level at most TX cells wide in one launch
// (bitwise what n_iter k_adjoint_pass_fused launches give): working workgroup
// b runs the pass of tile row b; pass p reads the multipliers from lu/lv/lH
// (p even) or the alternate buffers (p odd) and writes the other set, with the
// barriers and the one-XCD placement of k_vanka_persistent.  If a barrier
// timed out, the last pass writes NaN multipliers.
template <typename T>
__global__ __launch_bounds__(TX * 2 * TY) void k_adjoint_persistent(SsaArgs<T> a, const T* __restrict__ cache,
                                                                           T* lu, T* lv, T* lH, T* alt_u, T* alt_v,
                                                                           T* alt_H, const T* __restrict__ fu,
                                                                           const T* __restrict__ fv,
                                                                           const T* __restrict__ fH, T calving_rate,
                                                                           T omega, int n_iter,
                                                                           PersistentCtl ctl) {
    if (blockIdx.x % kXcds) return;
    const int nb = gridDim.x / kXcds;
    const int i0 = (blockIdx.x / kXcds) * TY;
    const bool fenced = !persistent_start(ctl, nb) || ctl.force_fenced;
#pragma unroll 1
    for (int p = 0; p < n_iter; ++p) {
        // see k_vanka_persistent: keeps the pass-invariant addresses from being hoisted
        int ty = threadIdx.y, tx = threadIdx.x;
        asm volatile("" : "+v"(ty), "+v"(tx));
        const bool odd = (p & 1) != 0;
        adjoint_pass_tile<T, LdPersistent, TY>(a, cache, odd ? alt_u : lu, odd ? alt_v : lv, odd ? alt_H : lH,
                                         odd ? lu : alt_u, odd ? lv : alt_v, odd ? lH : alt_H, fu, fv, fH,
                                         calving_rate, omega, i0, 0, ty, tx);
        if (p + 1 < n_iter) grid_barrier(ctl, (unsigned long long)nb * (p + 2), fenced);
    }
    const bool failed = persistent_failed(ctl);
    const int i = i0 + threadIdx.y, j = threadIdx.x;
    if (failed && threadIdx.y < TY && i < a.ny && j < a.nx)
        (((n_iter - 1) & 1) ? lH : alt_H)[(long)i * a.nx + j] = __longlong_as_double(0x7ff8000000000000ll);
}

// Doubles of dynamic LDS k_adjoint_coarse needs for an ny x nx level.
inline long adjoint_coarse_lds(int ny, int nx) {
    const long nc = (long)(ny + 3) * (nx + 3), nv = (long)(ny + 2) * (nx + 2);
    return 8 * nc + 2 * nv + (long)ny * (nx + 1) + (long)(ny + 1) * nx + (long)ny * nx;
}

// All n_iter adjoint smoothing passes of a small level (ny * nx <=
// kAdjointCoarseMaxCells) in one workgroup: per pass the transpose operator
// minus the forcing into LDS, the cached patch solves (increments into LDS),
// then the in-place update of k_vanka_adjoint_fused; the multipliers stay in
// global memory, ordered between passes by the workgroup barrier.  The same
// arithmetic as k_vjp followed by k_vanka_adjoint_fused, without a launch per
// pass.
constexpr int kCoarseThreads = 512;
template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void k_adjoint_coarse(SsaArgs<T> a, const T* __restrict__ cache, T* lu,
                                                                   T* lv, T* lH, const T* __restrict__ fu,
                                                                   const T* __restrict__ fv, const T* __restrict__ fH,
                                                                   T calving_rate, T omega, int n_iter) {
    extern __shared__ double lds[];
    const int ny = a.ny, nx = a.nx, rows = ny + 1, cols = nx + 1;
    const int n = ny * nx;
    VjpLevelStore<T> s;
    s.cw = cols + 2;
    s.nc = (rows + 2) * s.cw;
    s.vw = cols + 1;
    s.nv = (rows + 1) * s.vw;
    s.cell = reinterpret_cast<T*>(lds);
    s.vert = s.cell + 8 * s.nc;
    T* ru = s.vert + 2 * s.nv;  // transpose operator outputs
    T* rv = ru + ny * (nx + 1);
    T* rH = rv + (ny + 1) * nx;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const T* eta_cache = cache + (long)kAdjointCacheEta * n;
    VjpGridOut<T> out{ru, rv, rH};
    for (int it = 0; it < n_iter; ++it) {
        vjp_body(a, lu, lv, lH, fu, fv, fH, out, calving_rate, eta_cache, 0, 0, rows, cols, tid, nthr, s);
        __syncthreads();
        // the cell arrays of the transpose operator are free until the next
        // pass: the increments w0..w4 of cell c go to W[q * n + c]
        T* W = s.cell;
#pragma unroll 1
        for (int c = tid; c < n; c += nthr) {
            T w[5];
            adjoint_cached_increments(cache, ru, rv, rH, c / nx, c % nx, ny, nx, w);
#pragma unroll
            for (int q = 0; q < 5; ++q) W[q * n + c] = w[q];
        }
        __syncthreads();
#pragma unroll 1
        for (int c = tid; c < n; c += nthr) {
            const int i = c / nx, j = c % nx;
            {
                T d = W[c];
                if (j > 0) d = d + W[n + c - 1];
                lu[(long)i * (nx + 1) + j] += omega * d;
                if (j == nx - 1) {
                    T e = T(0) + W[n + c];
                    lu[(long)i * (nx + 1) + nx] += omega * e;
                }
            }
            {
                T d = W[2 * n + c];
                if (i > 0) d = d + W[3 * n + c - nx];
                lv[c] += omega * d;
                if (i == ny - 1) {
                    T e = T(0) + W[3 * n + c];
                    lv[(long)ny * nx + j] += omega * e;
                }
            }
            lH[c] += omega * W[4 * n + c];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ gradients
// dJ/dbeta and dJ/dbed (glide/cuda/grad.cu:1-169) as a per-cell gather,
// keeping the reference's facet guards (j < nx-1, i < ny-1) on the
// right/bottom-cell contributions.
template <typename T>
__device__ __forceinline__ void grad_beta_bed_body(int bx, int by, const SsaArgs<T>& a, const T* __restrict__ lu,
                                                   const T* __restrict__ lv, T* __restrict__ gbeta,
                                                   T* __restrict__ gbed, int exact) {
    const int j = bx * 256 + threadIdx.x, i = by;
    const int ny = a.ny, nx = a.nx;
    if (j >= nx) return;
    const T* __restrict__ u = a.u;
    const T* __restrict__ v = a.v;
    const T* __restrict__ H = a.H;
    const T* __restrict__ ph = a.phi;
    const T* __restrict__ bt = a.beta;
    const T* __restrict__ z = a.bed;
    T gb = T(0), gz = T(0);
    const long c = (long)i * nx + j;
    {  // u facet (i, j+1): this cell is its left cell (always counted)
        const int fj = j + 1;
        const long o = (long)i * (nx + 1) + fj;
        if (gbeta) {
            Drag<T> t = basal_drag(u[o], v[offV(i, fj - 1, ny, nx)], v[offV(i, fj, ny, nx)], v[offV(i + 1, fj - 1, ny, nx)],
                                   v[offV(i + 1, fj, ny, nx)], ph[offC(i, fj - 1, ny, nx)], ph[offC(i, fj, ny, nx)],
                                   bt[offC(i, fj - 1, ny, nx)], bt[offC(i, fj, ny, nx)], a.P);
            gb += lu[o] * t.d_bA;
        }
        if (gbed) {
            Drive<T> d = drive_x(H[c], H[offC(i, fj, ny, nx)], z[c], z[offC(i, fj, ny, nx)], ph[c],
                                 ph[offC(i, fj, ny, nx)], a.dxi, fj < nx);
            gz -= lu[o] * d.d_zA;
        }
    }
    if (j < nx - 1 || exact) {  // u facet (i, j): this cell is its right cell
        const long o = (long)i * (nx + 1) + j;
        if (gbeta) {
            Drag<T> t = basal_drag(u[o], v[offV(i, j - 1, ny, nx)], v[offV(i, j, ny, nx)], v[offV(i + 1, j - 1, ny, nx)],
                                   v[offV(i + 1, j, ny, nx)], ph[offC(i, j - 1, ny, nx)], ph[c],
                                   bt[offC(i, j - 1, ny, nx)], bt[c], a.P);
            gb += lu[o] * t.d_bB;
        }
        if (gbed) {
            Drive<T> d = drive_x(H[offC(i, j - 1, ny, nx)], H[c], z[offC(i, j - 1, ny, nx)], z[c],
                                 ph[offC(i, j - 1, ny, nx)], ph[c], a.dxi, j > 0);
            gz -= lu[o] * d.d_zB;
        }
    }
    {  // v facet (i+1, j): this cell is its top cell
        const int fi = i + 1;
        const long o = (long)fi * nx + j;
        if (gbeta) {
            Drag<T> t = basal_drag(v[o], u[offU(fi - 1, j, ny, nx)], u[offU(fi - 1, j + 1, ny, nx)], u[offU(fi, j, ny, nx)],
                                   u[offU(fi, j + 1, ny, nx)], ph[c], ph[offC(fi, j, ny, nx)], bt[c],
                                   bt[offC(fi, j, ny, nx)], a.P);
            gb += lv[o] * t.d_bA;
        }
        if (gbed) {
            Drive<T> d = drive_y(H[c], H[offC(fi, j, ny, nx)], z[c], z[offC(fi, j, ny, nx)], ph[c],
                                 ph[offC(fi, j, ny, nx)], a.dxi, fi < ny);
            gz -= lv[o] * d.d_zA;
        }
    }
    if (i < ny - 1 || exact) {  // v facet (i, j): this cell is its bottom cell
        const long o = (long)i * nx + j;
        if (gbeta) {
            Drag<T> t = basal_drag(v[o], u[offU(i - 1, j, ny, nx)], u[offU(i - 1, j + 1, ny, nx)], u[offU(i, j, ny, nx)],
                                   u[offU(i, j + 1, ny, nx)], ph[offC(i - 1, j, ny, nx)], ph[c],
                                   bt[offC(i - 1, j, ny, nx)], bt[c], a.P);
            gb += lv[o] * t.d_bB;
        }
        if (gbed) {
            Drive<T> d = drive_y(H[offC(i - 1, j, ny, nx)], H[c], z[offC(i - 1, j, ny, nx)], z[c],
                                 ph[offC(i - 1, j, ny, nx)], ph[c], a.dxi, i > 0);
            gz -= lv[o] * d.d_zB;
        }
    }
    if (gbeta) gbeta[c] = gb;
    if (gbed) gbed[c] = gz;
}

template <typename T>
__global__ __launch_bounds__(256) void k_grad_beta_bed(SsaArgs<T> a, const T* __restrict__ lu,
                                                       const T* __restrict__ lv, T* __restrict__ gbeta,
                                                       T* __restrict__ gbed, int exact) {
    grad_beta_bed_body<T>(blockIdx.x, blockIdx.y, a, lu, lv, gbeta, gbed, exact);
}

// ------------------------------------------------------------------ launchers
static inline dim3 tile_grid(int ny, int nx, bool with_facets) {
    int gx = with_facets ? (nx + 1 + TX - 1) / TX : (nx + TX - 1) / TX;
    int gy = with_facets ? (ny + 1 + TY - 1) / TY : (ny + TY - 1) / TY;
    return dim3(gx, gy);
}

template <typename T>
hipError_t launch_grounded(T* phi, const T* H, const T* bed, T c, T relax, long n, hipStream_t st) {
    k_grounded<T><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(phi, H, bed, c, relax, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_residual(const SsaArgs<T>& a, T* ou, T* ov, T* oH, int use_forcing, int use_mask, T calving_rate,
                           hipStream_t st) {
    TimedLaunch tl(kTimeResidual, (long)a.ny * a.nx, st);
    k_residual<T><<<tile_grid(a.ny, a.nx, true), dim3(TX, TY), 0, st>>>(a, ou, ov, oH, use_forcing, use_mask,
                                                                         calving_rate);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_jvp(const SsaArgs<T>& a, const T* du, const T* dv, const T* dH, T* ou, T* ov, T* oH, int use_mask,
                      T calving_rate, hipStream_t st) {
    k_jvp<T><<<tile_grid(a.ny, a.nx, true), dim3(TX, TY), 0, st>>>(a, du, dv, dH, ou, ov, oH, use_mask,
                                                                     calving_rate);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vanka_pass(const SsaArgs<T>& a, T* u, T* v, T* H, T* work, T* mask, const VankaArgs<T>& vk, T omega,
                             T* phi_next, T relax_phi, hipStream_t st) {
    {
        TimedLaunch tl(kTimeVankaCell, (long)a.ny * a.nx, st);
        constexpr int R = kVankaRows;
        const dim3 g((a.nx + TX - 1) / TX, (a.ny + R - 1) / R), b(TX, R);
        if (vk.local_fp32) {
            if (a.P.linear_drag) k_vanka_cell<T, true, float, R><<<g, b, 0, st>>>(a, work, mask, vk);
            else k_vanka_cell<T, false, float, R><<<g, b, 0, st>>>(a, work, mask, vk);
        } else {
            if (a.P.linear_drag) k_vanka_cell<T, true, T, R><<<g, b, 0, st>>>(a, work, mask, vk);
            else k_vanka_cell<T, false, T, R><<<g, b, 0, st>>>(a, work, mask, vk);
        }
    }
    dim3 g((a.nx + 1 + 255) / 256, a.ny + 1);
    TimedLaunch tl(kTimeVankaApply, (long)a.ny * a.nx, st);
    k_vanka_apply<T><<<g, 256, 0, st>>>(u, v, H, work, a.ny, a.nx, omega, phi_next, a.bed, a.P.frd, relax_phi);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vanka_fused(const SsaArgs<T>& a, T* u_dst, T* v_dst, T* H_dst, T* mask, const VankaArgs<T>& vk,
                              T omega, hipStream_t st) {
    TimedLaunch tl(kTimeVankaCell, (long)a.ny * a.nx, st);
    const dim3 g((a.nx + TX - 1) / TX, a.ny), b(TX, 3);
    // the mixed form evaluates and factorises the patch Jacobian in FP32
    if (vk.local_fp32) {
        if (a.P.linear_drag) k_vanka_fused<T, true, float><<<g, b, 0, st>>>(a, u_dst, v_dst, H_dst, mask, vk, omega);
        else k_vanka_fused<T, false, float><<<g, b, 0, st>>>(a, u_dst, v_dst, H_dst, mask, vk, omega);
    } else {
        if (a.P.linear_drag) k_vanka_fused<T, true, T><<<g, b, 0, st>>>(a, u_dst, v_dst, H_dst, mask, vk, omega);
        else k_vanka_fused<T, false, T><<<g, b, 0, st>>>(a, u_dst, v_dst, H_dst, mask, vk, omega);
    }
    return hipGetLastError();
}

bool vanka_persistent_fits(int ny, int nx) {
    // the three sync words sit behind the alternating state in the 5 ny nx workspace
    return nx >= 1 && nx <= TX && ny >= 1 && (ny + kPersistentRows - 1) / kPersistentRows <= kPersistentMaxBlocks &&
           2L * ny * nx >= (long)ny + nx + 3;
}

template <typename T>
hipError_t launch_vanka_persistent(const SsaArgs<T>& a, T* alt_u, T* alt_v, T* alt_H, T* mask,
                                   const VankaArgs<T>& vk, T omega, int n_iter, const PersistentCtl& ctl,
                                   hipStream_t st) {
    if (n_iter <= 0) return hipSuccess;
    if (!vanka_persistent_fits(a.ny, a.nx)) return hipErrorInvalidValue;
    // the arrival counter, the timeout flag and the XCD mask start at zero (a
    // kernel rather than a memset, so that the launch can sit in a captured graph)
    k_zero_sync<<<1, 64, 0, st>>>(ctl.sync);
    TimedLaunch tl(kTimeVankaCell, (long)a.ny * a.nx * n_iter, st);
    const dim3 g(kXcds * ((a.ny + kPersistentRows - 1) / kPersistentRows)), b(TX, kPersistentRows + 1);
    constexpr int R = kPersistentRows;
    if (vk.local_fp32) {
        if (a.P.linear_drag)
            k_vanka_persistent<T, true, float, R><<<g, b, 0, st>>>(a, alt_u, alt_v, alt_H, mask, vk, omega, n_iter, ctl);
        else
            k_vanka_persistent<T, false, float, R><<<g, b, 0, st>>>(a, alt_u, alt_v, alt_H, mask, vk, omega, n_iter, ctl);
    } else {
        if (a.P.linear_drag)
            k_vanka_persistent<T, true, T, R><<<g, b, 0, st>>>(a, alt_u, alt_v, alt_H, mask, vk, omega, n_iter, ctl);
        else
            k_vanka_persistent<T, false, T, R><<<g, b, 0, st>>>(a, alt_u, alt_v, alt_H, mask, vk, omega, n_iter, ctl);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vanka_dump(const SsaArgs<T>& a, T* J, T* r, T calving_rate, hipStream_t st) {
    k_vanka_dump<T><<<tile_grid(a.ny, a.nx, false), dim3(TX, TY), 0, st>>>(a, J, r, calving_rate);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vanka_adjoint_pass(const SsaArgs<T>& a, const T* ru, const T* rv, const T* rH, T* lu, T* lv, T* lH,
                                     T* work, const VankaArgs<T>& vk, T omega, hipStream_t st) {
    {
        TimedLaunch tl(kTimeAdjointCell, (long)a.ny * a.nx, st);
        k_vanka_adjoint_cell<T><<<tile_grid(a.ny, a.nx, false), dim3(TX, TY), 0, st>>>(a, ru, rv, rH, work, vk);
    }
    dim3 g((a.nx + 1 + 255) / 256, a.ny + 1);
    TimedLaunch tl(kTimeAdjointApply, (long)a.ny * a.nx, st);
    k_vanka_apply<T><<<g, 256, 0, st>>>(lu, lv, lH, work, a.ny, a.nx, -omega);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vjp(const SsaArgs<T>& a, const T* lu, const T* lv, const T* lH, const T* fu, const T* fv,
                      const T* fH, T* ou, T* ov, T* oH, T calving_rate, const T* eta_cache, hipStream_t st) {
    TimedLaunch tl(kTimeVjp, (long)a.ny * a.nx, st);
    constexpr int R = kVjpRows;
    const dim3 g((a.nx + 1 + TX - 1) / TX, (a.ny + 1 + R - 1) / R);
    k_vjp<T, R><<<g, dim3(TX, R), 0, st>>>(a, lu, lv, lH, fu, fv, fH, ou, ov, oH, calving_rate, eta_cache);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_adjoint_pass_fused(const SsaArgs<T>& a, const T* cache, const T* lu, const T* lv, const T* lH,
                                     T* lu_dst, T* lv_dst, T* lH_dst, const T* fu, const T* fv, const T* fH,
                                     T calving_rate, T omega, hipStream_t st) {
    if (!a.P.linear_drag) return hipErrorInvalidValue;
    TimedLaunch tl(kTimeAdjointCell, (long)a.ny * a.nx, st);
    k_adjoint_pass_fused<T, TY><<<tile_grid(a.ny, a.nx, false), dim3(TX, 2 * TY), 0, st>>>(
        a, cache, lu, lv, lH, lu_dst, lv_dst, lH_dst, fu, fv, fH, calving_rate, -omega, -1);
    return hipGetLastError();
}

// Tile rows of the multicolour adjoint phase: TY, or GLIDE_MC_ADJ_ROWS32 rows
// in FP32 (the FP64 tile keeps TY rows, whose LDS would double).  8 rows
// halve the ring's share of the transpose operator but measured no faster:
// 25.4 against 26.8 us per phase at 512², 75.3 against 73.5 ms per batched
// 96-domain ensemble step.
#ifndef GLIDE_MC_ADJ_ROWS32
#define GLIDE_MC_ADJ_ROWS32 4
#endif
template <typename T>
constexpr int mc_adj_rows() { return std::is_same<T, float>::value ? GLIDE_MC_ADJ_ROWS32 : TY; }
int tab_adj_pass_rows(bool fp32) { return fp32 ? mc_adj_rows<float>() : mc_adj_rows<double>(); }

template <typename T>
hipError_t launch_adjoint_pass_mc(const SsaArgs<T>& a, const T* cache, const T* lu, const T* lv, const T* lH,
                                  T* lu_dst, T* lv_dst, T* lH_dst, const T* fu, const T* fv, const T* fH,
                                  T calving_rate, T omega, int colour, hipStream_t st) {
    if (!a.P.linear_drag) return hipErrorInvalidValue;
    TimedLaunch tl(std::is_same<T, float>::value ? kTimeAdjointMc32 : kTimeAdjointMc64, (long)a.ny * a.nx, st);
    constexpr int R = mc_adj_rows<T>();
    const dim3 g((a.nx + TX - 1) / TX, (a.ny + R - 1) / R);
    k_adjoint_pass_fused<T, R, true><<<g, dim3(TX, 2 * R), 0, st>>>(a, cache, lu, lv, lH, lu_dst, lv_dst, lH_dst, fu,
                                                                   fv, fH, calving_rate, -omega, colour & 1);
    return hipGetLastError();
}
bool adjoint_persistent_fits(int ny, int nx) {
    // the three sync words sit behind the alternating multipliers in the 5 ny nx workspace
    return nx >= 1 && nx <= TX && ny >= 1 && (ny + TY - 1) / TY <= kAdjointPersistentMaxBlocks &&
           2L * ny * nx >= (long)ny + nx + 3;
}

template <typename T>
hipError_t launch_adjoint_persistent(const SsaArgs<T>& a, const T* cache, T* lu, T* lv, T* lH, T* alt_u, T* alt_v,
                                     T* alt_H, const T* fu, const T* fv, const T* fH, T calving_rate, T omega,
                                     int n_iter, const PersistentCtl& ctl, hipStream_t st) {
    if (n_iter <= 0) return hipSuccess;
    if (!a.P.linear_drag || !adjoint_persistent_fits(a.ny, a.nx)) return hipErrorInvalidValue;
    k_zero_sync<<<1, 64, 0, st>>>(ctl.sync);
    TimedLaunch tl(kTimeAdjointCell, (long)a.ny * a.nx * n_iter, st);
    k_adjoint_persistent<T><<<kXcds * ((a.ny + TY - 1) / TY), dim3(TX, 2 * TY), 0, st>>>(
        a, cache, lu, lv, lH, alt_u, alt_v, alt_H, fu, fv, fH, calving_rate, -omega, n_iter, ctl);
    return hipGetLastError();
}

template <typename T>
bool adjoint_coarse_fits(const SsaArgs<T>& a) {
    if (!a.P.linear_drag || a.ny < 1 || a.nx < 1 || (long)a.ny * a.nx > kAdjointCoarseMaxCells) return false;
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return false;
    return adjoint_coarse_lds(a.ny, a.nx) * (long)sizeof(T) <= (long)lds;
}
template <typename T>
hipError_t launch_adjoint_coarse(const SsaArgs<T>& a, const T* cache, T* lu, T* lv, T* lH, const T* fu, const T* fv,
                                 const T* fH, T calving_rate, T omega, int n_iter, hipStream_t st) {
    if (!adjoint_coarse_fits(a)) return hipErrorInvalidValue;
    const size_t bytes = (size_t)adjoint_coarse_lds(a.ny, a.nx) * sizeof(T);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_adjoint_coarse<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    TimedLaunch tl(kTimeAdjointCell, (long)a.ny * a.nx * n_iter, st);
    k_adjoint_coarse<T><<<1, kCoarseThreads, bytes, st>>>(a, cache, lu, lv, lH, fu, fv, fH, calving_rate, -omega,
                                                          n_iter);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_adjoint_factor(const SsaArgs<T>& a, T* cache, const VankaArgs<T>& vk, hipStream_t st) {
    TimedLaunch tl(kTimeAdjointFactor, (long)a.ny * a.nx, st);
    k_adjoint_factor<T><<<tile_grid(a.ny, a.nx, false), dim3(TX, TY), 0, st>>>(a, cache, vk);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vanka_adjoint_pass_cached(int ny, int nx, const T* cache, const T* ru, const T* rv, const T* rH,
                                            T* lu, T* lv, T* lH, T* work, T omega, bool fused, hipStream_t st) {
    if (fused) {
        TimedLaunch tl(kTimeAdjointCell, (long)ny * nx, st);
        k_vanka_adjoint_fused<T><<<tile_grid(ny, nx, false), dim3(TX, TY), 0, st>>>(cache, ru, rv, rH, lu, lv, lH, ny,
                                                                                   nx, -omega);
        return hipGetLastError();
    }
    {
        TimedLaunch tl(kTimeAdjointCell, (long)ny * nx, st);
        k_vanka_adjoint_cell_cached<T><<<tile_grid(ny, nx, false), dim3(TX, TY), 0, st>>>(cache, ru, rv, rH, work,
                                                                                         ny, nx);
    }
    dim3 g((nx + 1 + 255) / 256, ny + 1);
    TimedLaunch tl(kTimeAdjointApply, (long)ny * nx, st);
    k_vanka_apply<T><<<g, 256, 0, st>>>(lu, lv, lH, work, ny, nx, -omega);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_grad_beta_bed(const SsaArgs<T>& a, const T* lu, const T* lv, T* gbeta, T* gbed, int exact,
                                hipStream_t st) {
    dim3 g((a.nx + 255) / 256, a.ny);
    k_grad_beta_bed<T><<<g, 256, 0, st>>>(a, lu, lv, gbeta, gbed, exact);
    return hipGetLastError();
}

// Gradient of lambda^T R with respect to the rate factor B (no counterpart in
// the reference, whose gradients cover beta, bed, H_prev and smb,
// glide/cuda/grad.cu:1-169).  B enters R only through the cell viscosity
// eta = B/2 (...)^{(1-n)/(2n)}: eta H scales the cell's normal stresses and,
// averaged over four cells, the shear stress of each corner vertex, so
// g_B = S eta H / B with S = d(lambda^T R)/d(eta H) of the cell, the bracket
// the transpose operator's thickness row multiplies by eta (vjp_body), over
// the stress rows of interior facets (Dirichlet rows do not depend on B).
template <typename T>
__global__ __launch_bounds__(256) void k_grad_B(SsaArgs<T> a, const T* __restrict__ lu, const T* __restrict__ lv,
                                                T* __restrict__ gB) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    const int ny = a.ny, nx = a.nx;
    if (j >= nx) return;
    const T dxi = a.dxi;
    const long c = (long)i * nx + j;
    const CellVel<T> cv = load_cell_vel(a.u, a.v, i, j, ny, nx);
    const T e2 = strain_invariant(cv, i, j, ny, nx, dxi, a.P.eps_reg);
    const T eta = T(0.5) * a.B[c] * glen_power(e2, a.P);
    const T Hc = a.H[c];
    const NormalStress<T> sx = sxx(cv.ul, cv.ur, cv.vt, cv.vb, eta * Hc, dxi, true);
    const NormalStress<T> sy = syy(cv.ul, cv.ur, cv.vt, cv.vb, eta * Hc, dxi, true);
    const long ol = (long)i * (nx + 1) + j;
    const T lul = j > 0 ? lu[ol] : T(0), lur = j + 1 < nx ? lu[ol + 1] : T(0);
    const T lvt = i > 0 ? lv[c] : T(0), lvb = i + 1 < ny ? lv[c + nx] : T(0);
    T S = sx.d_eH * (lul - lur) * dxi + sy.d_eH * (lvb - lvt) * dxi;
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // corner vertices (i, j), (i, j+1), (i+1, j), (i+1, j+1)
        const int vi = i + (q >> 1), vj = j + (q & 1);
        if (vi > 0 && vi < ny && vj > 0 && vj < nx) {
            const long ut = (long)(vi - 1) * (nx + 1) + vj, ub = (long)vi * (nx + 1) + vj;
            const long vl = (long)vi * nx + vj - 1, vr = (long)vi * nx + vj;
            const ShearStress<T> sh = sxy(a.u[ut], a.u[ub], a.v[vl], a.v[vr], T(1), dxi, true);
            S += sh.d_eH * T(0.25) * (lu[ub] - lu[ut] + lv[vl] - lv[vr]) * dxi;
        }
    }
    gB[c] = a.B[c] != T(0) ? S * eta * Hc / a.B[c] : T(0);
}

template <typename T>
hipError_t launch_grad_B(const SsaArgs<T>& a, const T* lu, const T* lv, T* gB, hipStream_t st) {
    dim3 g((a.nx + 255) / 256, a.ny);
    k_grad_B<T><<<g, 256, 0, st>>>(a, lu, lv, gB);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero_masked(T* __restrict__ g, const T* __restrict__ mask, long n) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && mask[k] > T(0.5)) g[k] = T(0);
}

template <typename T>
hipError_t launch_zero_masked(T* g, const T* mask, long n, hipStream_t st) {
    k_zero_masked<T><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g, mask, n);
    return hipGetLastError();
}

// ------------------------------------------------------------ domain tables
// One launch over many domains (tab.hpp): each block runs the single-domain
// body of its job with its block index inside the job's own grid.
template <typename T>
__global__ __launch_bounds__(NT) void k_tab_residual(const TabResidual<T>* __restrict__ jobs,
                                                     const int* __restrict__ start, int n) {
    const int e = tab_find(start, n, blockIdx.x);
    const TabResidual<T> J = jobs[e];
    const int lb = blockIdx.x - start[e];
    residual_body<T>(lb % J.gx, lb / J.gx, J.a, J.ou, J.ov, J.oH, J.use_forcing, J.use_mask, J.cr);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_tab_adj_factor(const TabAdjFactor<T>* __restrict__ jobs,
                                                       const int* __restrict__ start, int n) {
    const int e = tab_find(start, n, blockIdx.x);
    const TabAdjFactor<T> J = jobs[e];
    const int lb = blockIdx.x - start[e];
    adjoint_factor_body<T>(lb % J.gx, lb / J.gx, J.a, J.cache, J.vk);
}

template <typename T, int TR>
__global__ __launch_bounds__(TX * 2 * TR) void k_tab_adj_pass(const TabAdjPass<T>* __restrict__ jobs,
                                                              const int* __restrict__ start, int n, int colour) {
    const int e = tab_find(start, n, blockIdx.x);
    const TabAdjPass<T> J = jobs[e];
    const int lb = blockIdx.x - start[e];
    const int bx = lb % J.gx, by = lb / J.gx;
    adjoint_pass_tile<T, LdPlain, TR, true>(J.a, J.cache, J.lu, J.lv, J.lH, J.lu_dst, J.lv_dst, J.lH_dst, J.fu, J.fv,
                                            J.fH, J.cr, J.omega, by * TR, bx * TX, threadIdx.y, threadIdx.x, colour);
}

template <typename T>
__global__ __launch_bounds__(TX * kVjpRows) void k_tab_vjp(const TabVjp<T>* __restrict__ jobs,
                                                           const int* __restrict__ start, int n) {
    constexpr int ROWS = kVjpRows;
    __shared__ VjpTile<T, ROWS> t;
    const int e = tab_find(start, n, blockIdx.x);
    const TabVjp<T> J = jobs[e];
    const int lb = blockIdx.x - start[e];
    const int bx = lb % J.gx, by = lb / J.gx;
    VjpTileStore<T, ROWS> s{t};
    VjpGridOut<T> o{J.ou, J.ov, J.oH};
    vjp_body(J.a, J.lu, J.lv, J.lH, J.fu, J.fv, J.fH, o, J.cr, J.eta_cache, by * ROWS, bx * TX, ROWS, TX,
             (int)(threadIdx.y * TX + threadIdx.x), TX * ROWS, s);
}

template <typename T>
__global__ __launch_bounds__(256) void k_tab_grad(const TabGrad<T>* __restrict__ jobs, const int* __restrict__ start,
                                                  int n) {
    const int e = tab_find(start, n, blockIdx.x);
    const TabGrad<T> J = jobs[e];
    const int lb = blockIdx.x - start[e];
    grad_beta_bed_body<T>(lb % J.gx, lb / J.gx, J.a, J.lu, J.lv, J.gbeta, J.gbed, J.exact);
}

int tab_blocks_residual(int ny, int nx, int* gx) {
    const dim3 g = tile_grid(ny, nx, true);
    *gx = (int)g.x;
    return (int)(g.x * g.y);
}
int tab_blocks_adj_factor(int ny, int nx, int* gx) {
    const dim3 g = tile_grid(ny, nx, false);
    *gx = (int)g.x;
    return (int)(g.x * g.y);
}
int tab_blocks_adj_pass(int ny, int nx, int rows, int* gx) {
    *gx = (nx + TX - 1) / TX;
    return *gx * ((ny + rows - 1) / rows);
}
int tab_blocks_vjp(int ny, int nx, int* gx) {
    *gx = (nx + 1 + TX - 1) / TX;
    return *gx * ((ny + 1 + kVjpRows - 1) / kVjpRows);
}
int tab_blocks_grad(int ny, int nx, int* gx) {
    *gx = (nx + 255) / 256;
    return *gx * ny;
}

template <typename T>
hipError_t launch_tab_residual(const TabResidual<T>* jobs, const int* start, int n, int blocks, hipStream_t st) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    k_tab_residual<T><<<blocks, dim3(TX, TY), 0, st>>>(jobs, start, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_tab_adj_factor(const TabAdjFactor<T>* jobs, const int* start, int n, int blocks, hipStream_t st) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    k_tab_adj_factor<T><<<blocks, dim3(TX, TY), 0, st>>>(jobs, start, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_tab_adj_pass(const TabAdjPass<T>* jobs, const int* start, int n, int blocks, int colour,
                               hipStream_t st) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    constexpr int R = mc_adj_rows<T>();
    k_tab_adj_pass<T, R><<<blocks, dim3(TX, 2 * R), 0, st>>>(jobs, start, n, colour & 1);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_tab_vjp(const TabVjp<T>* jobs, const int* start, int n, int blocks, hipStream_t st) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    k_tab_vjp<T><<<blocks, dim3(TX, kVjpRows), 0, st>>>(jobs, start, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_tab_grad(const TabGrad<T>* jobs, const int* start, int n, int blocks, hipStream_t st) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    k_tab_grad<T><<<blocks, 256, 0, st>>>(jobs, start, n);
    return hipGetLastError();
}

#define GLIDE_TAB_INSTANTIATE(T)                                                                                     \
    template hipError_t launch_tab_residual<T>(const TabResidual<T>*, const int*, int, int, hipStream_t);            \
    template hipError_t launch_tab_adj_factor<T>(const TabAdjFactor<T>*, const int*, int, int, hipStream_t);         \
    template hipError_t launch_tab_adj_pass<T>(const TabAdjPass<T>*, const int*, int, int, int, hipStream_t);        \
    template hipError_t launch_tab_vjp<T>(const TabVjp<T>*, const int*, int, int, hipStream_t);                      \
    template hipError_t launch_tab_grad<T>(const TabGrad<T>*, const int*, int, int, hipStream_t);

GLIDE_TAB_INSTANTIATE(double)
GLIDE_TAB_INSTANTIATE(float)

#define GLIDE_INSTANTIATE(T)                                                                                         \
    template hipError_t launch_grounded<T>(T*, const T*, const T*, T, T, long, hipStream_t);                          \
    template hipError_t launch_residual<T>(const SsaArgs<T>&, T*, T*, T*, int, int, T, hipStream_t);                  \
    template hipError_t launch_jvp<T>(const SsaArgs<T>&, const T*, const T*, const T*, T*, T*, T*, int, T,      \
                                      hipStream_t);                                                           \
    template hipError_t launch_vanka_pass<T>(const SsaArgs<T>&, T*, T*, T*, T*, T*, const VankaArgs<T>&, T, T*, T,   \
                                             hipStream_t);                                                           \
    template hipError_t launch_vanka_dump<T>(const SsaArgs<T>&, T*, T*, T, hipStream_t);                              \
    template hipError_t launch_vanka_fused<T>(const SsaArgs<T>&, T*, T*, T*, T*, const VankaArgs<T>&, T, hipStream_t);\
    template hipError_t launch_vanka_persistent<T>(const SsaArgs<T>&, T*, T*, T*, T*, const VankaArgs<T>&, T, int,    \
                                                  const PersistentCtl&, hipStream_t);                             \
    template hipError_t launch_vanka_adjoint_pass<T>(const SsaArgs<T>&, const T*, const T*, const T*, T*, T*, T*, T*, \
                                                     const VankaArgs<T>&, T, hipStream_t);                           \
    template hipError_t launch_adjoint_pass_fused<T>(const SsaArgs<T>&, const T*, const T*, const T*, const T*, T*,   \
                                                     T*, T*, const T*, const T*, const T*, T, T, hipStream_t);       \
    template hipError_t launch_adjoint_pass_mc<T>(const SsaArgs<T>&, const T*, const T*, const T*, const T*, T*, T*, \
                                                  T*, const T*, const T*, const T*, T, T, int, hipStream_t);         \
    template bool adjoint_coarse_fits<T>(const SsaArgs<T>&);                                                         \
    template hipError_t launch_adjoint_persistent<T>(const SsaArgs<T>&, const T*, T*, T*, T*, T*, T*, T*, const T*,  \
                                                     const T*, const T*, T, T, int, const PersistentCtl&,            \
                                                     hipStream_t);                                                    \
    template hipError_t launch_adjoint_coarse<T>(const SsaArgs<T>&, const T*, T*, T*, T*, const T*, const T*,          \
                                                 const T*, T, T, int, hipStream_t);                                  \
    template hipError_t launch_vjp<T>(const SsaArgs<T>&, const T*, const T*, const T*, const T*, const T*, const T*, \
                                      T*, T*, T*, T, const T*, hipStream_t);                                         \
    template hipError_t launch_adjoint_factor<T>(const SsaArgs<T>&, T*, const VankaArgs<T>&, hipStream_t);          \
    template hipError_t launch_vanka_adjoint_pass_cached<T>(int, int, const T*, const T*, const T*, const T*, T*, T*, \
                                                            T*, T*, T, bool, hipStream_t);                           \
    template hipError_t launch_grad_beta_bed<T>(const SsaArgs<T>&, const T*, const T*, T*, T*, int, hipStream_t); \
    template hipError_t launch_zero_masked<T>(T*, const T*, long, hipStream_t);                                       \
    template hipError_t launch_grad_B<T>(const SsaArgs<T>&, const T*, const T*, T*, hipStream_t);

GLIDE_INSTANTIATE(double)
GLIDE_INSTANTIATE(float)

}  // namespace glide
