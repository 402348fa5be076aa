// Padding / truncation along one axis of a strided 1-3D array, the spectral
// half of the 3/2-rule dealiasing (SpectralBase._padding_backward /
// _truncation_forward, spectralbase.py:1362-1389; fourier/bases.py:409-435,
// 570-598).  A padded space evaluates its transform on M = floor(pf N) points:
// backward pads the N coefficients into M slots, forward truncates the M
// coefficients of the padded-length transform back to N.  dealias_direct
// (no padding) zeroes the top third of the wavenumbers instead.
//
// One thread per output element; when the axis is contiguous consecutive
// threads walk the axis, otherwise they walk lines, so every access of a
// wave is a contiguous or a line-to-line coalesced segment.
#include <algorithm>

#include "common.h"

namespace sfa {

struct PadMap {
    int64_t sa_in, s1_in, s2_in;
    int64_t sa_out, s1_out, s2_out;
    int64_t d2, nlines;
};

enum { PAD_MODE_PAD = 0, PAD_MODE_TRUNCATE = 1, PAD_MODE_ZERO = 2 };
enum { PAD_KIND_POLY = 0, PAD_KIND_R2C = 1, PAD_KIND_C2C = 2 };

template <typename T>
__device__ inline T pz();
template <>
__device__ inline double pz<double>() { return 0.0; }
template <>
__device__ inline double2 pz<double2>() { return make_double2(0.0, 0.0); }
__device__ inline double pscale(double v, double s) { return v * s; }
__device__ inline double2 pscale(double2 v, double s) { return make_double2(v.x * s, v.y * s); }
__device__ inline double padd(double a, double b) { return a + b; }
__device__ inline double2 padd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
// the real part times s (the Hermitian-symmetric Nyquist coefficient)
__device__ inline double preal(double v, double s) { return v * s; }
__device__ inline double2 preal(double2 v, double s) { return make_double2(v.x * s, 0.0); }

// n_in / n_out: lengths along the axis; N: the unpadded size of the space
// (the real length for R2C); nbc: boundary slots at the end of a composite
// polynomial space; [z0, z1): the zeroed range of PAD_MODE_ZERO.
template <typename T, bool AXIS_FAST>
__global__ void pad_axis_kernel(const T* __restrict__ in, T* __restrict__ out, PadMap g, int n_in,
                                int n_out, int mode, int kind, int N, int nbc, int z0, int z1)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= g.nlines * n_out) return;
    int k;
    int64_t L;
    if (AXIS_FAST) {
        L = gid / n_out;
        k = (int)(gid - L * n_out);
    } else {
        k = (int)(gid / g.nlines);
        L = gid - (int64_t)k * g.nlines;
    }
    const int64_t i1 = L / g.d2, i2 = L - i1 * g.d2;
    const T* src = in + i1 * g.s1_in + i2 * g.s2_in;
    const int64_t sa = g.sa_in;
    const bool even = (N & 1) == 0;
    const int h = N / 2;
    T v = pz<T>();
    if (mode == PAD_MODE_PAD) {
        if (kind == PAD_KIND_POLY) {
            if (k < n_in - nbc) v = src[k * sa];
            else if (k >= n_out - nbc) v = src[(int64_t)(k - n_out + n_in) * sa];
        } else if (kind == PAD_KIND_R2C) {
            if (k < n_in) {
                v = src[k * sa];
                if (even && k == h) v = preal(v, 0.5);
            }
        } else {
            if (k <= h) v = src[k * sa];
            else if (k >= n_out - h) v = src[(int64_t)(k - n_out + N) * sa];
            if (even && (k == h || k == n_out - h)) v = pscale(v, 0.5);
        }
    } else if (mode == PAD_MODE_TRUNCATE) {
        if (kind == PAD_KIND_POLY) {
            if (k < n_out - nbc) v = src[k * sa];
        } else if (kind == PAD_KIND_R2C) {
            v = src[k * sa];
            if (even && k == h) v = preal(v, 2.0);
        } else {
            if (k <= h) v = src[k * sa];
            if (k >= N - h) v = padd(v, src[(int64_t)(k - N + n_in) * sa]);
        }
    } else {
        if (k < z0 || k >= z1) v = src[k * sa];
    }
    out[i1 * g.s1_out + i2 * g.s2_out + (int64_t)k * g.sa_out] = v;
}

}  // namespace sfa

using namespace sfa;

extern "C" int sfa_pad_axis(const void* in, void* out, int dtype, int ndim, const int64_t* in_shape,
                            const int64_t* in_strides, const int64_t* out_shape,
                            const int64_t* out_strides, int axis, int mode, int kind, int64_t N,
                            int64_t nbc, int64_t z0, int64_t z1, void* stream)
{
    Geom gi, go;
    int rc = make_geom(ndim, in_shape, in_strides, axis, &gi);
    if (rc) return rc;
    rc = make_geom(ndim, out_shape, out_strides, axis, &go);
    if (rc) return rc;
    SFA_CHECK(in != nullptr && out != nullptr && in != out, SFA_ERR_ARG,
              "pad_axis: null or aliased arrays");
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "pad_axis: bad dtype %d", dtype);
    SFA_CHECK(mode >= 0 && mode <= 2 && kind >= 0 && kind <= 2, SFA_ERR_ARG,
              "pad_axis: bad mode %d / kind %d", mode, kind);
    SFA_CHECK(kind == PAD_KIND_POLY || dtype == SFA_C128, SFA_ERR_ARG,
              "pad_axis: Fourier coefficients are complex128");
    for (int d = 0; d < ndim; ++d)
        SFA_CHECK(d == axis || in_shape[d] == out_shape[d], SFA_ERR_ARG,
                  "pad_axis: shapes differ off the axis (dim %d)", d);
    const int64_t ni = gi.n, no = go.n;
    SFA_CHECK(N >= 1 && ni <= (1 << 28) && no <= (1 << 28), SFA_ERR_ARG, "pad_axis: bad sizes");
    // lengths each mode and family reads and writes
    const int64_t nspec = (kind == PAD_KIND_R2C) ? N / 2 + 1 : N;
    if (mode == PAD_MODE_PAD) {
        SFA_CHECK(ni == nspec && no >= ni, SFA_ERR_ARG,
                  "pad_axis: pad needs in length %lld (got %lld) <= out length %lld",
                  (long long)nspec, (long long)ni, (long long)no);
        SFA_CHECK(kind != PAD_KIND_C2C || no >= N + (N % 2 == 0 ? 1 : 0), SFA_ERR_ARG,
                  "pad_axis: C2C padding must grow the line");
    } else if (mode == PAD_MODE_TRUNCATE) {
        SFA_CHECK(no == nspec && ni >= no, SFA_ERR_ARG,
                  "pad_axis: truncate needs out length %lld (got %lld) <= in length %lld",
                  (long long)nspec, (long long)no, (long long)ni);
    } else {
        SFA_CHECK(ni == no && z0 >= 0 && z1 >= z0, SFA_ERR_ARG, "pad_axis: bad zero range");
    }
    SFA_CHECK(kind != PAD_KIND_POLY || (nbc >= 0 && nbc <= (mode == PAD_MODE_PAD ? ni : no)),
              SFA_ERR_ARG, "pad_axis: bad boundary slot count %lld", (long long)nbc);
    PadMap g;
    g.sa_in = gi.sa;
    g.s1_in = gi.s1;
    g.s2_in = gi.s2;
    g.sa_out = go.sa;
    g.s1_out = go.s1;
    g.s2_out = go.s2;
    g.d2 = go.d2;
    g.nlines = go.d1 * go.d2;
    const int64_t total = g.nlines * no;
    if (total == 0) return SFA_OK;
    const int bs = 256;
    const dim3 grid((unsigned)((total + bs - 1) / bs));
    hipStream_t s = (hipStream_t)stream;
    const bool fast = go.sa == 1;
    const int zz0 = (int)std::min<int64_t>(z0, no), zz1 = (int)std::min<int64_t>(z1, no);
#define SFA_PAD_LAUNCH(T, F)                                                                     \
    hipLaunchKernelGGL((pad_axis_kernel<T, F>), grid, dim3(bs), 0, s, (const T*)in, (T*)out, g, \
                       (int)ni, (int)no, mode, kind, (int)N, (int)nbc, zz0, zz1)
    if (dtype == SFA_C128) {
        if (fast) SFA_PAD_LAUNCH(double2, true);
        else SFA_PAD_LAUNCH(double2, false);
    } else {
        if (fast) SFA_PAD_LAUNCH(double, true);
        else SFA_PAD_LAUNCH(double, false);
    }
#undef SFA_PAD_LAUNCH
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// ----------------------------------------------------------------- boundary values
// dst(.., k, ..) += sum_m s_m(line) sum_j W[m][k][j] src(.., n_src - nbc + j, ..)
// for k < nrows: the boundary functions' part of a composite space with
// nonhomogeneous boundary values, read from the nbc boundary slots at the end
// of each src line.  Backward adds the lift on the physical points
// (BoundaryValues._add_to_orthogonal, tensorproductspace.py:2070-2089), forward
// subtracts the boundary mass matrix (_add_mass_rhs :2091-2106), the solvers
// subtract the boundary stiffness and mass matrices scaled per mode
// (chebyshev/la.py:307-311, la.py:1212-1229).  One thread per (line, k).
namespace sfa {
template <typename T>
__global__ void boundary_update_kernel(const T* __restrict__ src, T* dst, PadMap g, int n_src,
                                       int nbc, int nrows, int nmat, const double* __restrict__ W,
                                       const double* __restrict__ scales, int64_t scale_stride)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= g.nlines * nrows) return;
    const int k = (int)(gid % nrows);
    const int64_t L = gid / nrows;
    const int64_t i1 = L / g.d2, i2 = L - i1 * g.d2;
    const T* s = src + i1 * g.s1_in + i2 * g.s2_in;
    T acc = pz<T>();
    for (int m = 0; m < nmat; ++m) {
        const double sc = scales ? scales[m * scale_stride + L] : 1.0;
        const double* Wm = W + ((int64_t)m * nrows + k) * nbc;
        for (int j = 0; j < nbc; ++j)
            acc = padd(acc, pscale(s[(int64_t)(n_src - nbc + j) * g.sa_in], sc * Wm[j]));
    }
    T* d = dst + i1 * g.s1_out + i2 * g.s2_out + (int64_t)k * g.sa_out;
    *d = padd(*d, acc);
}
}  // namespace sfa

extern "C" int sfa_boundary_update(const void* src, void* dst, int dtype, int ndim,
                                   const int64_t* src_shape, const int64_t* src_strides,
                                   const int64_t* dst_shape, const int64_t* dst_strides, int axis,
                                   int64_t nbc, int64_t nrows, int64_t nmat, const double* W,
                                   const double* scales, int64_t scale_stride, void* stream)
{
    Geom gi, go;
    int rc = make_geom(ndim, src_shape, src_strides, axis, &gi);
    if (rc) return rc;
    rc = make_geom(ndim, dst_shape, dst_strides, axis, &go);
    if (rc) return rc;
    SFA_CHECK(src != nullptr && dst != nullptr && W != nullptr, SFA_ERR_ARG,
              "boundary_update: null array");
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "boundary_update: bad dtype %d",
              dtype);
    for (int d = 0; d < ndim; ++d)
        SFA_CHECK(d == axis || src_shape[d] == dst_shape[d], SFA_ERR_ARG,
                  "boundary_update: shapes differ off the axis (dim %d)", d);
    SFA_CHECK(nbc >= 1 && nbc <= gi.n && nrows >= 0 && nrows <= go.n && nmat >= 1 && nmat <= 8,
              SFA_ERR_ARG, "boundary_update: bad nbc %lld / nrows %lld / nmat %lld",
              (long long)nbc, (long long)nrows, (long long)nmat);
    // in place: the rows written must not be the boundary slots read
    SFA_CHECK(src != dst || nrows <= gi.n - nbc, SFA_ERR_ARG,
              "boundary_update: in place, rows must stay below the boundary slots");
    const int64_t nlines = go.d1 * go.d2;
    SFA_CHECK(scales == nullptr || scale_stride >= nlines, SFA_ERR_ARG,
              "boundary_update: scale_stride %lld < lines %lld", (long long)scale_stride,
              (long long)nlines);
    PadMap g;
    g.sa_in = gi.sa;
    g.s1_in = gi.s1;
    g.s2_in = gi.s2;
    g.sa_out = go.sa;
    g.s1_out = go.s1;
    g.s2_out = go.s2;
    g.d2 = go.d2;
    g.nlines = nlines;
    const int64_t total = nlines * nrows;
    if (total == 0) return SFA_OK;
    const int bs = 256;
    const dim3 grid((unsigned)((total + bs - 1) / bs));
    if (dtype == SFA_C128)
        hipLaunchKernelGGL(boundary_update_kernel<double2>, grid, dim3(bs), 0, (hipStream_t)stream,
                           (const double2*)src, (double2*)dst, g, (int)gi.n, (int)nbc, (int)nrows,
                           (int)nmat, W, scales, scale_stride);
    else
        hipLaunchKernelGGL(boundary_update_kernel<double>, grid, dim3(bs), 0, (hipStream_t)stream,
                           (const double*)src, (double*)dst, g, (int)gi.n, (int)nbc, (int)nrows,
                           (int)nmat, W, scales, scale_stride);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// ----------------------------------------------------------------- convolution
// Pointwise product of two contiguous arrays (the physical-space step of a
// dealiased convolution, TensorProductSpace.convolve, tensorproductspace.py:637-666).
namespace sfa {
template <typename T>
__device__ inline T pmul(T a, T b);
template <>
__device__ inline double pmul<double>(double a, double b) { return a * b; }
template <>
__device__ inline double2 pmul<double2>(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename T>
__global__ void pointwise_mul_kernel(const T* __restrict__ a, const T* __restrict__ b,
                                     T* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pmul(a[i], b[i]);
}

// Direct convolution sums of two Fourier coefficient vectors (the fast=False
// path).  C2C (convolve_1D, convolve.pyx:9-30): N coefficients in FFT order,
// out 2N entries, out[(m + n) mod 2N] += u_m v_n over the wavenumbers of a
// length Np = N (odd) or N + 1 (even) grid, the Nyquist entries halved.
// R2C (convolve_real_1D, convolve.pyx:34-75): N/2+1 coefficients, negative
// wavenumbers by conjugation, out N + 1 entries, p = m + n >= 0 only.
// One thread per output entry, the wavenumber pairs enumerated in the
// reference's order for that entry.
__device__ inline double2 conv_coef(const double2* u, int m, int N, bool real)
{
    const int am = m < 0 ? -m : m;
    if ((N & 1) == 0 && am == N / 2) {  // Nyquist: halved, never conjugated
        const double2 c = u[am];
        return make_double2(c.x * 0.5, c.y * 0.5);
    }
    if (real) {
        const double2 c = u[am];
        return m < 0 ? make_double2(c.x, -c.y) : c;
    }
    return u[m < 0 ? N + m : m];
}

__global__ void convolve_direct_kernel(const double2* __restrict__ u, const double2* __restrict__ v,
                                       double2* __restrict__ out, int N, int nout, int real)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nout) return;
    const int Np = (N & 1) ? N : N + 1;
    const int kmax = Np / 2;  // wavenumbers -kmax .. kmax
    double2 acc = make_double2(0.0, 0.0);
    // outputs receiving p: C2C q = p mod 2N (p and p - 2N), R2C q = p
    for (int j = 0; j < Np; ++j) {
        const int m = j <= kmax ? j : j - Np;  // fftfreq order, as the reference loops
        for (int r = 0; r < (real ? 1 : 2); ++r) {
            const int p = real ? q : (r == 0 ? q : q - 2 * N);
            const int n = p - m;
            if (n < -kmax || n > kmax) continue;
            const double2 a = conv_coef(u, m, N, real != 0);
            const double2 b = conv_coef(v, n, N, real != 0);
            acc.x += a.x * b.x - a.y * b.y;
            acc.y += a.x * b.y + a.y * b.x;
        }
    }
    out[q] = acc;
}
}  // namespace sfa

extern "C" int sfa_pointwise_mul(const void* a, const void* b, void* out, int dtype, int64_t n,
                                 void* stream)
{
    SFA_CHECK(a != nullptr && b != nullptr && out != nullptr && n >= 0, SFA_ERR_ARG,
              "pointwise_mul: bad arguments");
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "pointwise_mul: bad dtype %d",
              dtype);
    if (n == 0) return SFA_OK;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (dtype == SFA_C128)
        hipLaunchKernelGGL(pointwise_mul_kernel<double2>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double2*)a, (const double2*)b, (double2*)out, n);
    else
        hipLaunchKernelGGL(pointwise_mul_kernel<double>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double*)a, (const double*)b, (double*)out, n);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// y[r n + i] += a x[i] for r < nrep: the lift of nonhomogeneous boundary
// values added to each solution of a stack.  Two doubles per thread (16-byte
// accesses) where x and y are 16-byte aligned and n is even, else one.
template <int V>
__global__ void __launch_bounds__(256) axpy_kernel(const double* __restrict__ x, double* __restrict__ y,
                                                   double a, int64_t nv, int64_t nrep)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv * nrep; i += stride) {
        const int64_t k = i % nv;
        if constexpr (V == 2) {
            const double2 xv = reinterpret_cast<const double2*>(x)[k];
            double2 yv = reinterpret_cast<double2*>(y)[i];
            yv.x += a * xv.x;
            yv.y += a * xv.y;
            reinterpret_cast<double2*>(y)[i] = yv;
        } else {
            y[i] += a * x[k];
        }
    }
}

extern "C" int sfa_axpy(const double* x, double* y, double a, int64_t n, int64_t nrep, void* stream)
{
    SFA_CHECK(x != nullptr && y != nullptr && n >= 0 && nrep >= 0, SFA_ERR_ARG, "axpy: bad arguments");
    if (n == 0 || nrep == 0) return SFA_OK;
    const bool v2 = n % 2 == 0 && ((uintptr_t)x | (uintptr_t)y) % 16 == 0;
    const int64_t nv = v2 ? n / 2 : n, tot = nv * nrep;
    const int64_t blocks = std::min<int64_t>((tot + 255) / 256, 8192);
    if (v2)
        hipLaunchKernelGGL(axpy_kernel<2>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, x, y, a,
                           nv, nrep);
    else
        hipLaunchKernelGGL(axpy_kernel<1>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, x, y, a,
                           nv, nrep);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_convolve_direct(const void* u, const void* v, void* out, int64_t N, int real,
                                   void* stream)
{
    SFA_CHECK(u != nullptr && v != nullptr && out != nullptr, SFA_ERR_ARG,
              "convolve_direct: null array");
    SFA_CHECK(N >= 1 && N <= (1 << 20) && (real == 0 || real == 1), SFA_ERR_ARG,
              "convolve_direct: bad N %lld / real %d", (long long)N, real);
    const int nout = real ? (int)N + 1 : 2 * (int)N;
    hipLaunchKernelGGL(convolve_direct_kernel, dim3((nout + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, (const double2*)u, (const double2*)v, (double2*)out,
                       (int)N, nout, real);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}
