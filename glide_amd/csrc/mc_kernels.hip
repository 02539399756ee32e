// gfx950 kernels of the multicolour (red-black) Vanka smoother, the smoother
// of the mixed-precision solver (no counterpart in the reference, whose Vanka
// pass is additive: glide/cuda/vanka.cu:477-660).
//
// A pass is three launches: the cell viscosity of the state the pass starts
// from (frozen for the pass, as in the reference), then the red cells
// ((i + j) even) and the black cells.  A colour phase reads the state from one
// buffer set and writes the complete new state to the other: each facet
// belongs to exactly one cell of a colour, which writes it; the cells of the
// other colour copy their thickness and the domain-edge facets nobody of the
// colour owns.  Cells of one colour read each other's old values, so a phase
// is a block-Jacobi step over non-overlapping patches and its result does not
// depend on scheduling (run-to-run bitwise identical).
//
// Work mapping: a lane owns one cell of the active colour and the inactive
// cell beside it in the same row, so a wavefront covers 128 columns and no
// lane idles on the other colour; the active cells' loads are stride-2 and the
// other colour's phase reads the other half of the same lines.  The local
// Newton iteration runs the folded cell system of the additive smoother
// (cell_system, refined hardware reciprocals), without compensated updates
// and with the caller's relaxation and stopping threshold.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ssa_device.hpp"
#include "ssa_kernels.hpp"
#include "tab.hpp"

namespace glide {

namespace {

constexpr int kMcLanes = 64;  // cell pairs per wavefront row segment
constexpr int kMcRows = 4;    // rows per block

// Kernel bodies take their block index (bx, by) as arguments, so that the
// domain-table launches (one launch over many domains) run them unchanged.
template <typename T>
__device__ __forceinline__ void cell_eta_body(int bx, int by, const SsaArgs<T>& a, T* __restrict__ eta) {
    const int j = bx * 256 + threadIdx.x, i = by;
    if (j >= a.nx) return;
    CellVel<T> c = load_cell_vel<T>(a.u, a.v, i, j, a.ny, a.nx);
    const T e2 = strain_invariant(c, i, j, a.ny, a.nx, a.dxi, a.P.eps_reg);
    eta[(long)i * a.nx + j] = T(0.5) * a.B[(long)i * a.nx + j] * glen_power(e2, a.P);
}

template <typename T>
__global__ __launch_bounds__(256) void k_cell_eta(SsaArgs<T> a, T* __restrict__ eta) {
    cell_eta_body<T>(blockIdx.x, blockIdx.y, a, eta);
}

// Local Newton iteration of cell (i, j) against the state in a; returns the
// new unknowns x and the active-set flag.  EDGE: the cell may lie on the
// domain edge (the Dirichlet pins are tested).
template <typename T, bool LINEAR, bool EDGE>
__device__ __forceinline__ void mc_newton(const SsaArgs<T>& a, const T* __restrict__ eta, int i, int j,
                                          const VankaArgs<T>& vk, T x[5], T& masked) {
    const int ny = a.ny, nx = a.nx;
    CellSys<T> cs;
    T thk;
    {
        Patch<T> p;
        load_patch_with<T>(p, a, [&](int di, int dj) { return eta[offC(i + di, j + dj, ny, nx)]; }, i, j, a.f_u, a.f_v,
                           a.f_H, a.phi);
        make_cell_sys(cs, p, a.P, vk.calving_rate, a.dxi, a.dt, vk.ssa_damping, vk.mc_damping);
        thk = p.gam;
    }
    const long c = (long)i * nx + j;
    x[0] = a.u[(long)i * (nx + 1) + j];
    x[1] = a.u[(long)i * (nx + 1) + j + 1];
    x[2] = a.v[c];
    x[3] = a.v[c + nx];
    x[4] = a.H[c];
    const bool pin0 = j == 0, pin1 = j == nx - 1, pin2 = i == 0, pin3 = i == ny - 1;
    T rnorm = T(1e30);
    for (int it = 0; it < vk.newton_steps && rnorm > vk.tol; ++it) {
        T J[25], r[5];
        cell_system<T, T, LINEAR>(cs, cs, x, J, r, a.P);
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
        T d[5];
        lu5_solve_rcp(J, r, d);
        rnorm = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3] + r[4] * r[4];
#pragma unroll
        for (int k = 0; k < 5; ++k) x[k] -= vk.relax * d[k];
        x[4] = fmax(x[4], thk);
    }
}

// ROWS rows of cell pairs per block (kMcLanes pairs each); the arithmetic of
// a cell does not depend on ROWS.
template <typename T, bool LINEAR, int ROWS = kMcRows>
__device__ __forceinline__ void vanka_mc_body(int bx, int by, const SsaArgs<T>& a, const T* __restrict__ eta,
                                              T* __restrict__ u_dst, T* __restrict__ v_dst, T* __restrict__ H_dst,
                                              T* __restrict__ mask, const VankaArgs<T>& vk, int colour) {
    const int ny = a.ny, nx = a.nx;
    const int i = by * ROWS + threadIdx.y;
    if (i >= ny) return;
    const int pair = bx * kMcLanes + threadIdx.x;
    const int p = (i + colour) & 1;
    const int j = 2 * pair + p;        // the lane's cell of the active colour
    const int jo = 2 * pair + 1 - p;   // and the other colour's cell beside it
    if (j < nx) {
        const long c = (long)i * nx + j;
        T x[5], masked = mask[c];
        // waves holding a row on the domain edge or the first / last column
        // test the Dirichlet pins; the others run a loop copy without them
        const bool edge = i == 0 || i == ny - 1 || bx == 0 || 2 * (bx + 1) * kMcLanes >= nx;
        if (edge) mc_newton<T, LINEAR, true>(a, eta, i, j, vk, x, masked);
        else mc_newton<T, LINEAR, false>(a, eta, i, j, vk, x, masked);
        u_dst[(long)i * (nx + 1) + j] = x[0];
        u_dst[(long)i * (nx + 1) + j + 1] = x[1];
        v_dst[c] = x[2];
        v_dst[c + nx] = x[3];
        H_dst[c] = x[4];
        mask[c] = masked;
    }
    if (jo < nx) {
        const long c = (long)i * nx + jo;
        H_dst[c] = a.H[c];
        if (jo == 0) u_dst[(long)i * (nx + 1)] = a.u[(long)i * (nx + 1)];
        if (jo == nx - 1) u_dst[(long)i * (nx + 1) + nx] = a.u[(long)i * (nx + 1) + nx];
        if (i == 0) v_dst[jo] = a.v[jo];
        if (i == ny - 1) v_dst[(long)ny * nx + jo] = a.v[(long)ny * nx + jo];
    }
}

template <typename T, bool LINEAR>
__global__ __launch_bounds__(kMcLanes* kMcRows) void k_vanka_mc(SsaArgs<T> a, const T* __restrict__ eta,
                                                                T* __restrict__ u_dst, T* __restrict__ v_dst,
                                                                T* __restrict__ H_dst, T* __restrict__ mask,
                                                                VankaArgs<T> vk, int colour) {
    vanka_mc_body<T, LINEAR>(blockIdx.x, blockIdx.y, a, eta, u_dst, v_dst, H_dst, mask, vk, colour);
}

template <typename T>
__global__ __launch_bounds__(256) void k_pin_rows(T* __restrict__ du, T* __restrict__ dv, T* __restrict__ dH,
                                                  const T* __restrict__ mask, int ny, int nx) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > nx) return;
    if (i < ny && (j == 0 || j == nx)) du[(long)i * (nx + 1) + j] = T(0);
    if (j < nx) {
        if (i == 0 || i == ny) dv[(long)i * nx + j] = T(0);
        if (i < ny && mask[(long)i * nx + j] > T(0.5)) dH[(long)i * nx + j] = T(0);
    }
}

}  // namespace

template <typename T>
hipError_t launch_pin_rows(T* du, T* dv, T* dH, const T* mask, int ny, int nx, hipStream_t st) {
    k_pin_rows<T><<<dim3((nx + 1 + 255) / 256, ny + 1), 256, 0, st>>>(du, dv, dH, mask, ny, nx);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cell_eta(const SsaArgs<T>& a, T* eta, hipStream_t st) {
    k_cell_eta<T><<<dim3((a.nx + 255) / 256, a.ny), 256, 0, st>>>(a, eta);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vanka_mc(const SsaArgs<T>& a, const T* eta, T* u_dst, T* v_dst, T* H_dst, T* mask,
                           const VankaArgs<T>& vk, int colour, hipStream_t st) {
    const int pairs = (a.nx + 1) / 2;
    const dim3 g((pairs + kMcLanes - 1) / kMcLanes, (a.ny + kMcRows - 1) / kMcRows), b(kMcLanes, kMcRows);
    TimedLaunch tl(std::is_same<T, float>::value ? kTimeVankaMc32 : kTimeVankaMc64, (long)a.ny * a.nx, st);
    if (a.P.linear_drag)
        k_vanka_mc<T, true><<<g, b, 0, st>>>(a, eta, u_dst, v_dst, H_dst, mask, vk, colour & 1);
    else
        k_vanka_mc<T, false><<<g, b, 0, st>>>(a, eta, u_dst, v_dst, H_dst, mask, vk, colour & 1);
    return hipGetLastError();
}

// ------------------------------------------------------------ domain tables
namespace {
template <typename T>
__global__ __launch_bounds__(256) void k_tab_eta(const TabEta<T>* __restrict__ jobs, const int* __restrict__ start,
                                                 int n) {
    const int e = tab_find(start, n, blockIdx.x);
    const TabEta<T> J = jobs[e];
    const int lb = blockIdx.x - start[e];
    cell_eta_body<T>(lb % J.gx, lb / J.gx, J.a, J.eta);
}

template <typename T, bool LINEAR, int ROWS>
__global__ __launch_bounds__(kMcLanes* ROWS) void k_tab_vanka_mc(const TabVankaMc<T>* __restrict__ jobs,
                                                                 const int* __restrict__ start, int n, int colour) {
    const int e = tab_find(start, n, blockIdx.x);
    const TabVankaMc<T> J = jobs[e];
    const int lb = blockIdx.x - start[e];
    vanka_mc_body<T, LINEAR, ROWS>(lb % J.gx, lb / J.gx, J.a, J.eta, J.u_dst, J.v_dst, J.H_dst, J.mask, J.vk,
                                   colour);
}
}  // namespace

int tab_blocks_eta(int ny, int nx, int* gx) {
    *gx = (nx + 255) / 256;
    return *gx * ny;
}
bool tab_vanka_mc_rows_ok(int rows) { return rows == 1 || rows == 2 || rows == 4 || rows == 8; }
int tab_blocks_vanka_mc(int ny, int nx, int rows, int* gx) {
    const int pairs = (nx + 1) / 2;
    *gx = (pairs + kMcLanes - 1) / kMcLanes;
    return *gx * ((ny + rows - 1) / rows);
}

template <typename T>
hipError_t launch_tab_eta(const TabEta<T>* jobs, const int* start, int n, int blocks, hipStream_t st) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    k_tab_eta<T><<<blocks, 256, 0, st>>>(jobs, start, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_tab_vanka_mc(const TabVankaMc<T>* jobs, const int* start, int n, int blocks, int colour,
                               bool linear, int rows, hipStream_t st) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    const dim3 b(kMcLanes, rows);
    const int c = colour & 1;
#define GLIDE_TAB_MC(R)                                                                        \
    if (rows == R) {                                                                           \
        if (linear) k_tab_vanka_mc<T, true, R><<<blocks, b, 0, st>>>(jobs, start, n, c);      \
        else k_tab_vanka_mc<T, false, R><<<blocks, b, 0, st>>>(jobs, start, n, c);            \
        return hipGetLastError();                                                              \
    }
    GLIDE_TAB_MC(1)
    GLIDE_TAB_MC(2)
    GLIDE_TAB_MC(4)
    GLIDE_TAB_MC(8)
#undef GLIDE_TAB_MC
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------ precision bridge
__global__ __launch_bounds__(256) void k_mixed_batch(MixedBatch b) {
    const MixedJob& q = b.job[blockIdx.y];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < q.n; k += (long)gridDim.x * blockDim.x) {
        switch (q.op) {
            case kMixTo32:
                static_cast<float*>(q.out)[k] = (float)static_cast<const double*>(q.a)[k];
                break;
            case kMixTo64:
                static_cast<double*>(q.out)[k] = (double)static_cast<const float*>(q.a)[k];
                break;
            case kMixDefect:  // out32 = a32 - (float) b64
                static_cast<float*>(q.out)[k] =
                    static_cast<const float*>(q.a)[k] - (float)static_cast<const double*>(q.b)[k];
                break;
            case kMixUpdate:
            case kMixUpdateH: {  // x64 += y32 - fl32(x64); H: x64 = thk where mask32 is set
                double* x = static_cast<double*>(q.out);
                const double x0 = x[k];
                double xn = x0 + ((double)static_cast<const float*>(q.a)[k] - (double)(float)x0);
                if (q.op == kMixUpdateH && static_cast<const float*>(q.b)[k] > 0.5f) xn = q.val;
                x[k] = xn;
                break;
            }
            case kMixAdd64:  // out64 += (double) a32
                static_cast<double*>(q.out)[k] += (double)static_cast<const float*>(q.a)[k];
                break;
            case kMixNeg32:  // out32 = -(float) a64
                static_cast<float*>(q.out)[k] = -(float)static_cast<const double*>(q.a)[k];
                break;
        }
    }
}

hipError_t launch_mixed_batch(const MixedBatch& b, hipStream_t st) {
    if (b.count <= 0) return hipSuccess;
    long nmax = 0;
    for (int q = 0; q < b.count; ++q) nmax = nmax > b.job[q].n ? nmax : b.job[q].n;
    long bx = (nmax + 255) / 256;
    if (bx > 2048) bx = 2048;
    if (bx < 1) bx = 1;
    k_mixed_batch<<<dim3((unsigned)bx, b.count), 256, 0, st>>>(b);
    return hipGetLastError();
}

// ------------------------------------------------------------ Krylov helpers
// The flexible GMRES outer iteration of the mixed adjoint solve works on flat
// FP64 vectors (the u, v and H multipliers of a level one after another);
// these are its inner products and updates.  The sums are formed in a fixed
// order (fixed grid, fixed block reduction), so repeated solves are bitwise
// identical.

// partial[q * gridDim.x + b]: block b's part of <w, V_q> for q < nq (V_q at
// V + q * ld), and of <w, w> at q = nq.
__global__ __launch_bounds__(256) void k_krylov_dots(const double* __restrict__ w, const double* __restrict__ V,
                                                     long ld, long n, int nq, double* __restrict__ partial) {
    __shared__ double red[4][kKrylovMaxDots + 1];
    double acc[kKrylovMaxDots + 1];
#pragma unroll
    for (int q = 0; q <= kKrylovMaxDots; ++q) acc[q] = 0.0;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const double wk = w[k];
#pragma unroll
        for (int q = 0; q < kKrylovMaxDots; ++q)
            if (q < nq) acc[q] = fma(wk, V[q * ld + k], acc[q]);
        acc[kKrylovMaxDots] = fma(wk, wk, acc[kKrylovMaxDots]);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q <= kKrylovMaxDots; ++q) {
        if (q < nq || q == kKrylovMaxDots) {
            double v = acc[q];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0) red[wv][q] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x <= nq) {
        const int q = threadIdx.x == nq ? kKrylovMaxDots : threadIdx.x;
        partial[(long)threadIdx.x * gridDim.x + blockIdx.x] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
    }
}

// out[q] = sum over the nb partials of q, q <= nq (one block per q).
__global__ __launch_bounds__(256) void k_krylov_reduce(const double* __restrict__ partial, int nb,
                                                       double* __restrict__ out) {
    __shared__ double red[4];
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += partial[(long)blockIdx.x * nb + b];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// out[k] += sign * sum_{q < nq} c[q] V[q * ld + k]  (c on the device).
__global__ __launch_bounds__(256) void k_krylov_axpy(double* __restrict__ out, const double* __restrict__ V, long ld,
                                                     long n, int nq, const double* __restrict__ c, double sign) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < nq; ++q) s = fma(c[q], V[q * ld + k], s);
        out[k] = fma(sign, s, out[k]);
    }
}

hipError_t launch_krylov_dots(const double* w, const double* V, long ld, long n, int nq, double* partial,
                              double* out, hipStream_t st) {
    if (nq < 0 || nq > kKrylovMaxDots) return hipErrorInvalidValue;
    k_krylov_dots<<<kKrylovBlocks, 256, 0, st>>>(w, V, ld, n, nq, partial);
    k_krylov_reduce<<<nq + 1, 256, 0, st>>>(partial, kKrylovBlocks, out);
    return hipGetLastError();
}

hipError_t launch_krylov_axpy(double* out, const double* V, long ld, long n, int nq, const double* c, double sign,
                              hipStream_t st) {
    if (nq <= 0 || n <= 0) return hipSuccess;
    long b = (n + 255) / 256;
    if (b > 2048) b = 2048;
    k_krylov_axpy<<<(unsigned)b, 256, 0, st>>>(out, V, ld, n, nq, c, sign);
    return hipGetLastError();
}

#define GLIDE_MC_INSTANTIATE(T)                                                                                     \
    template hipError_t launch_pin_rows<T>(T*, T*, T*, const T*, int, int, hipStream_t);                           \
    template hipError_t launch_cell_eta<T>(const SsaArgs<T>&, T*, hipStream_t);                                    \
    template hipError_t launch_vanka_mc<T>(const SsaArgs<T>&, const T*, T*, T*, T*, T*, const VankaArgs<T>&, int,  \
                                           hipStream_t);

GLIDE_MC_INSTANTIATE(double)
GLIDE_MC_INSTANTIATE(float)

#define GLIDE_MC_TAB_INSTANTIATE(T)                                                                                 \
    template hipError_t launch_tab_eta<T>(const TabEta<T>*, const int*, int, int, hipStream_t);                    \
    template hipError_t launch_tab_vanka_mc<T>(const TabVankaMc<T>*, const int*, int, int, int, bool, int,   \
                                               hipStream_t);

GLIDE_MC_TAB_INSTANTIATE(double)
GLIDE_MC_TAB_INSTANTIATE(float)

}  // namespace glide
