// Multigrid transfers between MAC-grid levels (factor 2 per direction) and the
// element-wise / reduction helpers of the FAS cycles.
//
// Transfer operators follow glide/cuda/transfer.cu: tangential pair means for
// facets, 2x2 mean or max for cells, bilinear facet prolongation with edge
// clamping, cell injection or bilinear prolongation.  Each thread produces one
// output value; rows are contiguous in x so a wavefront reads/writes one row
// segment.
#include <hip/hip_runtime.h>

#include "ssa_kernels.hpp"

namespace glide {

template <typename T>
__global__ __launch_bounds__(256) void k_restrict_vfacet(const T* __restrict__ f, T* __restrict__ c, int nyc, int nxc) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
    if (I > nxc) return;
    const long sf = 2 * nxc + 1;
    c[(long)J * (nxc + 1) + I] = T(0.5) * (f[2L * J * sf + 2 * I] + f[(2L * J + 1) * sf + 2 * I]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_restrict_hfacet(const T* __restrict__ f, T* __restrict__ c, int nyc, int nxc) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
    if (I >= nxc) return;
    const long sf = 2 * nxc;
    c[(long)J * nxc + I] = T(0.5) * (f[2L * J * sf + 2 * I] + f[2L * J * sf + 2 * I + 1]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_restrict_cell(const T* __restrict__ f, T* __restrict__ c, int nyc, int nxc,
                                                       int use_max) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
    if (I >= nxc) return;
    const long sf = 2 * nxc;
    T a = f[2L * J * sf + 2 * I], b = f[2L * J * sf + 2 * I + 1];
    T d = f[(2L * J + 1) * sf + 2 * I], e = f[(2L * J + 1) * sf + 2 * I + 1];
    c[(long)J * nxc + I] = use_max ? fmax(fmax(a, b), fmax(d, e)) : T(0.25) * (a + b + d + e);
}

// Bilinear interpolation weights along one axis for a fine index whose coarse
// coordinate is s, clamped to [0, hi].
template <typename T>
__device__ __forceinline__ void bilin_axis(T s, int hi, int& lo, int& up, T& t) {
    s = fmax(T(0), fmin(s, (T)hi));
    lo = (int)s;
    up = min(lo + 1, hi);
    t = s - (T)lo;
}

// u-facets: fine (nyf, nxf+1) from coarse (nyf/2, nxf/2+1); fine face (j, i)
// sits at coarse coordinates ((j-0.5)/2, i/2).
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_vfacet_bilinear(const T* __restrict__ c, T* __restrict__ f, int nyf,
                                                                 int nxf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i > nxf) return;
    const int nyc = nyf / 2, nxc = nxf / 2;
    int jl, jh, il, ih;
    T ty, tx;
    bilin_axis((T(j) - T(0.5)) * T(0.5), nyc - 1, jl, jh, ty);
    bilin_axis(T(i) * T(0.5), nxc, il, ih, tx);
    const int s = nxc + 1;
    T v00 = c[jl * s + il], v01 = c[jl * s + ih], v10 = c[jh * s + il], v11 = c[jh * s + ih];
    f[(long)j * (nxf + 1) + i] = (T(1) - ty) * ((T(1) - tx) * v00 + tx * v01) + ty * ((T(1) - tx) * v10 + tx * v11);
}

// v-facets: fine (nyf+1, nxf) from coarse (nyf/2+1, nxf/2); fine face (j, i)
// sits at coarse coordinates (j/2, (i-0.5)/2).
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_hfacet_bilinear(const T* __restrict__ c, T* __restrict__ f, int nyf,
                                                                 int nxf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= nxf) return;
    const int nyc = nyf / 2, nxc = nxf / 2;
    int jl, jh, il, ih;
    T ty, tx;
    bilin_axis(T(j) * T(0.5), nyc, jl, jh, ty);
    bilin_axis((T(i) - T(0.5)) * T(0.5), nxc - 1, il, ih, tx);
    T v00 = c[jl * nxc + il], v01 = c[jl * nxc + ih], v10 = c[jh * nxc + il], v11 = c[jh * nxc + ih];
    f[(long)j * nxf + i] = (T(1) - ty) * ((T(1) - tx) * v00 + tx * v01) + ty * ((T(1) - tx) * v10 + tx * v11);
}

template <typename T>
__global__ __launch_bounds__(256) void k_prolong_cell(const T* __restrict__ c, T* __restrict__ f, int nyf, int nxf,
                                                      int bilinear) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= nxf) return;
    const int nyc = nyf / 2, nxc = nxf / 2;
    if (!bilinear) {
        f[(long)j * nxf + i] = c[(long)(j / 2) * nxc + i / 2];
        return;
    }
    int jl, jh, il, ih;
    T ty, tx;
    bilin_axis((T(j) - T(0.5)) * T(0.5), nyc - 1, jl, jh, ty);
    bilin_axis((T(i) - T(0.5)) * T(0.5), nxc - 1, il, ih, tx);
    T v00 = c[jl * nxc + il], v01 = c[jl * nxc + ih], v10 = c[jh * nxc + il], v11 = c[jh * nxc + ih];
    f[(long)j * nxf + i] = (T(1) - ty) * ((T(1) - tx) * v00 + tx * v01) + ty * ((T(1) - tx) * v10 + tx * v11);
}

template <typename T>
__global__ __launch_bounds__(256) void k_axpby(T* __restrict__ out, T a, const T* __restrict__ x, T b,
                                               const T* __restrict__ y, T c, const T* __restrict__ z, long n) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    T r = a * x[k] + b * y[k];
    if (z) r += c * z[k];
    out[k] = r;
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy(T* __restrict__ out, const T* __restrict__ x, long n) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = x[k];
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ out, T val, long n) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = val;
}

// Batched element-wise jobs of one V-cycle stage (independent of each other)
// in one launch: blockIdx.y selects the job, a grid-stride loop covers it.
// Each job keeps the arithmetic of its single-array kernel.
template <typename T>
__global__ __launch_bounds__(256) void k_ew_batch(EwBatch<T> b) {
    const EwJob<T>& q = b.job[blockIdx.y];
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < q.n; k += (long)gridDim.x * blockDim.x) {
        switch (q.op) {
            case kEwAxpby: {
                T r = q.a * q.x[k] + q.b * q.y[k];
                if (q.z) r += q.c * q.z[k];
                q.out[k] = r;
                break;
            }
            case kEwCopy: q.out[k] = q.x[k]; break;
            case kEwFill: q.out[k] = q.a; break;
            default: {  // kEwCorrect: z = z + y, then x = w + z
                const T t = q.x[k] + q.y[k];
                q.out[k] = t;
                q.out2[k] = q.z[k] + t;
                break;
            }
        }
    }
}

// Batched restrictions onto one coarse level (nyc, nxc): blockIdx.z selects
// the job; the arithmetic of k_restrict_vfacet / _hfacet / _cell.
template <typename T>
__global__ __launch_bounds__(256) void k_restrict_batch(RestrictBatch<T> b, int nyc, int nxc) {
    const RestrictJob<T>& q = b.job[blockIdx.z];
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
    const T* __restrict__ f = q.f;
    T* __restrict__ c = q.c;
    if (q.kind == kRestrictU) {
        if (I > nxc || J >= nyc) return;
        const long sf = 2 * nxc + 1;
        c[(long)J * (nxc + 1) + I] = T(0.5) * (f[2L * J * sf + 2 * I] + f[(2L * J + 1) * sf + 2 * I]);
    } else if (q.kind == kRestrictV) {
        if (I >= nxc || J > nyc) return;
        const long sf = 2 * nxc;
        c[(long)J * nxc + I] = T(0.5) * (f[2L * J * sf + 2 * I] + f[2L * J * sf + 2 * I + 1]);
    } else {
        if (I >= nxc || J >= nyc) return;
        const long sf = 2 * nxc;
        T a = f[2L * J * sf + 2 * I], bb = f[2L * J * sf + 2 * I + 1];
        T d = f[(2L * J + 1) * sf + 2 * I], e = f[(2L * J + 1) * sf + 2 * I + 1];
        c[(long)J * nxc + I] = q.kind == kRestrictCellMax ? fmax(fmax(a, bb), fmax(d, e)) : T(0.25) * (a + bb + d + e);
    }
}

__device__ __forceinline__ double block_sum_256(double s, double* sh) {
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    return t;
}

// Sum of squares, first pass: blockIdx.y selects the array.
template <typename T>
__global__ __launch_bounds__(256) void k_sumsq_partial(const T* __restrict__ a, long na, const T* __restrict__ b,
                                                       long nb, const T* __restrict__ c, long nc,
                                                       double* __restrict__ partials) {
    __shared__ double sh[4];
    const T* p = blockIdx.y == 0 ? a : (blockIdx.y == 1 ? b : c);
    const long n = blockIdx.y == 0 ? na : (blockIdx.y == 1 ? nb : nc);
    double s = 0.0;
    if (p)
        for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
            double x = (double)p[k];
            s += x * x;
        }
    double t = block_sum_256(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_sumsq_final(const double* __restrict__ partials, int nblk,
                                                     double* __restrict__ out3) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) s += partials[blockIdx.x * nblk + k];
    double t = block_sum_256(s, sh);
    if (threadIdx.x == 0) out3[blockIdx.x] = t;
}

// ------------------------------------------------------------------ launchers
template <typename T>
hipError_t launch_restrict_vfacet(const T* f, T* c, int nyc, int nxc, hipStream_t st) {
    k_restrict_vfacet<T><<<dim3((nxc + 1 + 255) / 256, nyc), 256, 0, st>>>(f, c, nyc, nxc);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_restrict_hfacet(const T* f, T* c, int nyc, int nxc, hipStream_t st) {
    k_restrict_hfacet<T><<<dim3((nxc + 255) / 256, nyc + 1), 256, 0, st>>>(f, c, nyc, nxc);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_restrict_cell(const T* f, T* c, int nyc, int nxc, int use_max, hipStream_t st) {
    k_restrict_cell<T><<<dim3((nxc + 255) / 256, nyc), 256, 0, st>>>(f, c, nyc, nxc, use_max);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_prolong_vfacet_bilinear(const T* c, T* f, int nyf, int nxf, hipStream_t st) {
    k_prolong_vfacet_bilinear<T><<<dim3((nxf + 1 + 255) / 256, nyf), 256, 0, st>>>(c, f, nyf, nxf);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_prolong_hfacet_bilinear(const T* c, T* f, int nyf, int nxf, hipStream_t st) {
    k_prolong_hfacet_bilinear<T><<<dim3((nxf + 255) / 256, nyf + 1), 256, 0, st>>>(c, f, nyf, nxf);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_prolong_cell(const T* c, T* f, int nyf, int nxf, int bilinear, hipStream_t st) {
    k_prolong_cell<T><<<dim3((nxf + 255) / 256, nyf), 256, 0, st>>>(c, f, nyf, nxf, bilinear);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_axpby(T* out, T a, const T* x, T b, const T* y, T c, const T* z, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_axpby<T><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(out, a, x, b, y, c, z, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_copy(T* out, const T* x, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_copy<T><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(out, x, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_fill(T* out, T val, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_fill<T><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(out, val, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_sumsq3(const T* a, long na, const T* b, long nb, const T* c, long nc, double* partials, double* out3,
                         hipStream_t st) {
    long nmax = na > nb ? na : nb;
    nmax = nmax > nc ? nmax : nc;
    int nblk = (int)((nmax + 255) / 256);
    if (nblk > kSumsqBlocks) nblk = kSumsqBlocks;
    if (nblk < 1) nblk = 1;
    k_sumsq_partial<T><<<dim3(nblk, 3), 256, 0, st>>>(a, na, b, nb, c, nc, partials);
    k_sumsq_final<<<3, 256, 0, st>>>(partials, nblk, out3);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ew_batch(const EwBatch<T>& b, hipStream_t st) {
    if (b.count <= 0) return hipSuccess;
    long nmax = 0;
    for (int q = 0; q < b.count; ++q) nmax = b.job[q].n > nmax ? b.job[q].n : nmax;
    if (nmax <= 0) return hipSuccess;
    long gx = (nmax + 255) / 256;
    if (gx > 2048) gx = 2048;
    k_ew_batch<T><<<dim3((unsigned)gx, b.count), 256, 0, st>>>(b);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_restrict_batch(const RestrictBatch<T>& b, int nyc, int nxc, hipStream_t st) {
    if (b.count <= 0) return hipSuccess;
    k_restrict_batch<T><<<dim3((nxc + 1 + 255) / 256, nyc + 1, b.count), 256, 0, st>>>(b, nyc, nxc);
    return hipGetLastError();
}

#define GLIDE_TRANSFER_INSTANTIATE(T)                                                                     \
    template hipError_t launch_restrict_vfacet<T>(const T*, T*, int, int, hipStream_t);                    \
    template hipError_t launch_restrict_hfacet<T>(const T*, T*, int, int, hipStream_t);                    \
    template hipError_t launch_restrict_cell<T>(const T*, T*, int, int, int, hipStream_t);                 \
    template hipError_t launch_prolong_vfacet_bilinear<T>(const T*, T*, int, int, hipStream_t);            \
    template hipError_t launch_prolong_hfacet_bilinear<T>(const T*, T*, int, int, hipStream_t);            \
    template hipError_t launch_prolong_cell<T>(const T*, T*, int, int, int, hipStream_t);                  \
    template hipError_t launch_axpby<T>(T*, T, const T*, T, const T*, T, const T*, long, hipStream_t);     \
    template hipError_t launch_fill<T>(T*, T, long, hipStream_t);                                          \
    template hipError_t launch_copy<T>(T*, const T*, long, hipStream_t);                                   \
    template hipError_t launch_sumsq3<T>(const T*, long, const T*, long, const T*, long, double*, double*, \
                                         hipStream_t);                                                     \
    template hipError_t launch_ew_batch<T>(const EwBatch<T>&, hipStream_t);                                \
    template hipError_t launch_restrict_batch<T>(const RestrictBatch<T>&, int, int, hipStream_t);

GLIDE_TRANSFER_INSTANTIATE(double)
GLIDE_TRANSFER_INSTANTIATE(float)

}  // namespace glide
