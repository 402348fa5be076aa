// Shared device/host helpers for the shenfun_amd HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/shenfun_amd.h"

namespace sfa {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const char* fmt, ...);

#define SFA_HIP(call)                                                          \
    do {                                                                       \
        hipError_t _e = (call);                                                \
        if (_e != hipSuccess)                                                  \
            return ::sfa::fail(SFA_ERR_HIP, "%s failed: %s (%s:%d)", #call,    \
                               hipGetErrorString(_e), __FILE__, __LINE__);     \
    } while (0)

#define SFA_CHECK(cond, code, ...)                                             \
    do {                                                                       \
        if (!(cond)) return ::sfa::fail((code), __VA_ARGS__);                  \
    } while (0)

#define SFA_LAUNCH_CHECK()                                                     \
    do {                                                                       \
        hipError_t _e = hipGetLastError();                                     \
        if (_e != hipSuccess)                                                  \
            return ::sfa::fail(SFA_ERR_HIP, "kernel launch failed: %s (%s:%d)", \
                               hipGetErrorString(_e), __FILE__, __LINE__);     \
    } while (0)

// ---------------------------------------------------------------- complex
struct cd {  // complex double, 16-byte aligned like hipDoubleComplex
    double x, y;
};

__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cmulc(double2 a, double2 b) {  // a * conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// float2 counterparts (complex64 arithmetic of the mixed-precision kernels)
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// complex type traits: C2 in {double2, float2}
template <typename C2>
struct CT;
template <>
struct CT<double2> {
    using R = double;
    __host__ __device__ static inline double2 mk(double x, double y) { return make_double2(x, y); }
};
template <>
struct CT<float2> {
    using R = float;
    __host__ __device__ static inline float2 mk(float x, float y) { return make_float2(x, y); }
};
template <typename C2>
__host__ __device__ inline C2 cmk(double x, double y)
{
    using R = typename CT<C2>::R;
    return CT<C2>::mk((R)x, (R)y);
}
__host__ __device__ inline double2 to_c64(float2 a) { return make_double2(a.x, a.y); }
__host__ __device__ inline double2 to_c64(double2 a) { return a; }
template <typename C2>
__host__ __device__ inline C2 from_c64(double2 a);
template <>
__host__ __device__ inline double2 from_c64<double2>(double2 a) { return a; }
template <>
__host__ __device__ inline float2 from_c64<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }

// ---------------------------------------------------------------- geometry
// A batch of 1D lines inside an up-to-3D strided array: the line axis has
// length n and stride sa; the remaining two dims (d1, s1) (d2, s2) enumerate
// the lines.  Strides are in elements of the array's own dtype.
struct Geom {
    int64_t n, sa;
    int64_t d1, s1;
    int64_t d2, s2;
};

// Build a Geom from (ndim, shape, strides, axis).  Returns 0 or an error.
int make_geom(int ndim, const int64_t* shape, const int64_t* strides, int axis, Geom* g);

// Device twiddle tables, cached per length for the process lifetime.
// fft_table(n): w[m] = exp(-2 pi i m / n), m = 0..n-1.
// quarter_table(N): w[k] = exp(-i pi k / (2N)), k = 0..N-1.
const double2* fft_table(int n);
const double2* quarter_table(int N);

// While a NoAllocScope lives on this host thread, the lazily built device
// tables (twiddles, Galerkin diagonals, stencils) are looked up only: a table
// that does not exist yet makes its getter return NULL (and record the error)
// instead of calling hipMalloc.  Used by the no-allocation entry points.
bool alloc_forbidden();
struct NoAllocScope {
    NoAllocScope();
    ~NoAllocScope();
    NoAllocScope(const NoAllocScope&) = delete;
    NoAllocScope& operator=(const NoAllocScope&) = delete;
    bool prev;
};

int device_cu_count();

}  // namespace sfa
