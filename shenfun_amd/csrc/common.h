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

int device_cu_count();

}  // namespace sfa
