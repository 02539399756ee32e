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
#include "tab.hpp"

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

// ------------------------------------------------------------ domain tables
// Element-wise jobs of many domains in one launch (tab.hpp): each keeps the
// arithmetic of its single-domain kernel (k_ew_batch, k_axpby, k_mixed_batch,
// k_grounded, k_zero_masked).
template <typename T>
__device__ __forceinline__ void tab_ew_one(const TabEw& q, int op, long k) {
    T* out = static_cast<T*>(q.out);
    const T* x = static_cast<const T*>(q.x);
    switch (op) {
        case 0: {  // axpby
            T r = T(q.a) * x[k] + T(q.b) * static_cast<const T*>(q.y)[k];
            if (q.z) r += T(q.c) * static_cast<const T*>(q.z)[k];
            out[k] = r;
            break;
        }
        case 1: out[k] = x[k]; break;
        case 2: out[k] = T(q.a); break;
        default: {  // correct: z = z + y, then x = w + z
            const T t = x[k] + static_cast<const T*>(q.y)[k];
            out[k] = t;
            static_cast<T*>(q.out2)[k] = static_cast<const T*>(q.z)[k] + t;
            break;
        }
    }
}

__global__ __launch_bounds__(256) void k_tab_ew(const TabEw* __restrict__ jobs, const int* __restrict__ start, int n) {
    const int e = tab_find(start, n, blockIdx.x);
    const TabEw q = jobs[e];
    const long base = (long)(blockIdx.x - start[e]) * kTabEwPerBlock;
    const int op = q.op;
#pragma unroll
    for (int r = 0; r < kTabEwPerBlock / 256; ++r) {
        const long k = base + r * 256 + threadIdx.x;
        if (k >= q.n) return;
        if (op <= kTabCorrect32) {
            tab_ew_one<float>(q, op - kTabAxpby32, k);
        } else if (op <= kTabCorrect64) {
            tab_ew_one<double>(q, op - kTabAxpby64, k);
        } else {
            switch (op) {
                case kTabTo32:
                    static_cast<float*>(q.out)[k] = (float)static_cast<const double*>(q.x)[k];
                    break;
                case kTabTo64:
                    static_cast<double*>(q.out)[k] = (double)static_cast<const float*>(q.x)[k];
                    break;
                case kTabDefect:  // out32 = x32 - (float) y64
                    static_cast<float*>(q.out)[k] =
                        static_cast<const float*>(q.x)[k] - (float)static_cast<const double*>(q.y)[k];
                    break;
                case kTabUpdate:
                case kTabUpdateH: {  // out64 += x32 - fl32(out64); H: out64 = a where y32 is set
                    double* xo = static_cast<double*>(q.out);
                    const double x0 = xo[k];
                    double xn = x0 + ((double)static_cast<const float*>(q.x)[k] - (double)(float)x0);
                    if (op == kTabUpdateH && static_cast<const float*>(q.y)[k] > 0.5f) xn = q.a;
                    xo[k] = xn;
                    break;
                }
                case kTabAdd64:
                    static_cast<double*>(q.out)[k] += (double)static_cast<const float*>(q.x)[k];
                    break;
                case kTabNeg32:
                    static_cast<float*>(q.out)[k] = -(float)static_cast<const double*>(q.x)[k];
                    break;
                case kTabGrounded32: {
                    float* phi = static_cast<float*>(q.out);
                    const float c = (float)q.a, relax = (float)q.b;
                    const float g = grounded_fraction(static_cast<const float*>(q.x)[k],
                                                      static_cast<const float*>(q.y)[k], c);
                    phi[k] = relax == 0.0f ? g : (1.0f - relax) * g + relax * phi[k];
                    break;
                }
                case kTabGrounded64: {
                    double* phi = static_cast<double*>(q.out);
                    const double g = grounded_fraction(static_cast<const double*>(q.x)[k],
                                                       static_cast<const double*>(q.y)[k], q.a);
                    phi[k] = q.b == 0.0 ? g : (1.0 - q.b) * g + q.b * phi[k];
                    break;
                }
                default:  // kTabZeroMasked64
                    if (static_cast<const double*>(q.x)[k] > 0.5) static_cast<double*>(q.out)[k] = 0.0;
                    break;
            }
        }
    }
}

// FP32 transfers and pinning of many domains in one launch: the arithmetic of
// k_restrict_batch, k_prolong_* and k_pin_rows (row J = block row, column I).
__global__ __launch_bounds__(256) void k_tab_xfer(const TabXfer* __restrict__ jobs, const int* __restrict__ start,
                                                  int n) {
    typedef float T;
    const int e = tab_find(start, n, blockIdx.x);
    const TabXfer q = jobs[e];
    const int lb = blockIdx.x - start[e];
    const int I = (lb % q.gx) * 256 + threadIdx.x, J = lb / q.gx;
    const int ny = q.ny, nx = q.nx;
    const T* __restrict__ src = q.src;
    T* __restrict__ dst = q.dst;
    switch (q.op) {
        case kTabRestrictU: {
            if (I > nx) return;
            const long sf = 2 * nx + 1;
            dst[(long)J * (nx + 1) + I] = T(0.5) * (src[2L * J * sf + 2 * I] + src[(2L * J + 1) * sf + 2 * I]);
            break;
        }
        case kTabRestrictV: {
            if (I >= nx) return;
            const long sf = 2 * nx;
            dst[(long)J * nx + I] = T(0.5) * (src[2L * J * sf + 2 * I] + src[2L * J * sf + 2 * I + 1]);
            break;
        }
        case kTabRestrictMean:
        case kTabRestrictMax: {
            if (I >= nx) return;
            const long sf = 2 * nx;
            T a = src[2L * J * sf + 2 * I], bb = src[2L * J * sf + 2 * I + 1];
            T d = src[(2L * J + 1) * sf + 2 * I], ee = src[(2L * J + 1) * sf + 2 * I + 1];
            dst[(long)J * nx + I] =
                q.op == kTabRestrictMax ? fmax(fmax(a, bb), fmax(d, ee)) : T(0.25) * (a + bb + d + ee);
            break;
        }
        case kTabProlongVfacet: {
            if (I > nx) return;
            const int nyc = ny / 2, nxc = nx / 2;
            int jl, jh, il, ih;
            T ty, tx;
            bilin_axis((T(J) - T(0.5)) * T(0.5), nyc - 1, jl, jh, ty);
            bilin_axis(T(I) * T(0.5), nxc, il, ih, tx);
            const int s = nxc + 1;
            T v00 = src[jl * s + il], v01 = src[jl * s + ih], v10 = src[jh * s + il], v11 = src[jh * s + ih];
            dst[(long)J * (nx + 1) + I] =
                (T(1) - ty) * ((T(1) - tx) * v00 + tx * v01) + ty * ((T(1) - tx) * v10 + tx * v11);
            break;
        }
        case kTabProlongHfacet: {
            if (I >= nx) return;
            const int nyc = ny / 2, nxc = nx / 2;
            int jl, jh, il, ih;
            T ty, tx;
            bilin_axis(T(J) * T(0.5), nyc, jl, jh, ty);
            bilin_axis((T(I) - T(0.5)) * T(0.5), nxc - 1, il, ih, tx);
            T v00 = src[jl * nxc + il], v01 = src[jl * nxc + ih], v10 = src[jh * nxc + il], v11 = src[jh * nxc + ih];
            dst[(long)J * nx + I] = (T(1) - ty) * ((T(1) - tx) * v00 + tx * v01) + ty * ((T(1) - tx) * v10 + tx * v11);
            break;
        }
        case kTabProlongInject: {
            if (I >= nx) return;
            dst[(long)J * nx + I] = src[(long)(J / 2) * (nx / 2) + I / 2];
            break;
        }
        case kTabProlongBilinear: {
            if (I >= nx) return;
            const int nyc = ny / 2, nxc = nx / 2;
            int jl, jh, il, ih;
            T ty, tx;
            bilin_axis((T(J) - T(0.5)) * T(0.5), nyc - 1, jl, jh, ty);
            bilin_axis((T(I) - T(0.5)) * T(0.5), nxc - 1, il, ih, tx);
            T v00 = src[jl * nxc + il], v01 = src[jl * nxc + ih], v10 = src[jh * nxc + il], v11 = src[jh * nxc + ih];
            dst[(long)J * nx + I] = (T(1) - ty) * ((T(1) - tx) * v00 + tx * v01) + ty * ((T(1) - tx) * v10 + tx * v11);
            break;
        }
        default: {  // kTabPinRows: (i, j) = (J, I) over (ny + 1) x (nx + 1)
            const int i = J, j = I;
            if (j > nx) return;
            if (i < ny && (j == 0 || j == nx)) dst[(long)i * (nx + 1) + j] = T(0);
            if (j < nx) {
                if (i == 0 || i == ny) q.dv[(long)i * nx + j] = T(0);
                if (i < ny && q.mask[(long)i * nx + j] > T(0.5)) q.dH[(long)i * nx + j] = T(0);
            }
            break;
        }
    }
}

int tab_blocks_xfer(int op, int ny, int nx, int* gx) {
    int rows = 0, cols = 0;
    switch (op) {
        case kTabRestrictU: rows = ny; cols = nx + 1; break;
        case kTabRestrictV: rows = ny + 1; cols = nx; break;
        case kTabRestrictMean:
        case kTabRestrictMax: rows = ny; cols = nx; break;
        case kTabProlongVfacet: rows = ny; cols = nx + 1; break;
        case kTabProlongHfacet: rows = ny + 1; cols = nx; break;
        case kTabProlongInject:
        case kTabProlongBilinear: rows = ny; cols = nx; break;
        default: rows = ny + 1; cols = nx + 1; break;
    }
    *gx = (cols + 255) / 256;
    return *gx * rows;
}

hipError_t launch_tab_ew(const TabEw* jobs, const int* start, int n, int blocks, hipStream_t st) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    k_tab_ew<<<blocks, 256, 0, st>>>(jobs, start, n);
    return hipGetLastError();
}
hipError_t launch_tab_xfer(const TabXfer* jobs, const int* start, int n, int blocks, hipStream_t st) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    k_tab_xfer<<<blocks, 256, 0, st>>>(jobs, start, n);
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
