// Device-side FFT building blocks: small in-register DFTs and LDS Stockham
// stages (mixed radix 2,3,4,5,8,16 + generic prime), FP64.
#pragma once
#include "common.h"

namespace sfa {

// Spectrum statistics of the guarded mixed-precision solve: per band of 16
// kz planes (band 0 = kz 0..16, band b >= 1 = kz 16b+1 .. 16b+16, so the
// band edges are the admissible first complex64 planes 16m+1), the energy of
// the x-pass input lines (f after the z and y transforms, accumulators 0 ..
// kStatBands-1) and that of u's Chebyshev coefficients (kStatBands ..
// 2 kStatBands-1), both summed by the x-pass.  Each accumulator is spread
// over kStatSlots slots of one 128-byte line each, picked by workgroup index:
// thousands of waves adding into a handful of addresses serialise on them
// (a z R2C summing per wave at 256^3 took 117 us instead of 50).  A one-workgroup
// kernel after the x-pass sums the slots into the 2 kStatBands results.
constexpr int kStatBands = 16;
constexpr int kStatSlots = 64;
constexpr int kStatLine = 16;  // doubles per slot
constexpr int kStatAcc = 2 * kStatBands;
constexpr int kStatBytes = kStatAcc * kStatSlots * kStatLine * 8;
__host__ __device__ constexpr int stat_band(int kz) { return kz == 0 ? 0 : (kz - 1) / 16; }
__host__ __device__ constexpr int stat_index(int acc, int slot)
{
    return (acc * kStatSlots + slot) * kStatLine;
}

// Where the summed statistics go: kStatAcc doubles, then a flag set to seq
// once they are written (device addresses of pinned, fine-grained host memory
// that the host spins on).  flag == NULL: no delivery.
struct StatOut {
    double* sums;
    unsigned* flag;
    unsigned seq;
};

// One workgroup (every thread calls it) sums the kStatSlots slots of each
// accumulator of a finished statistics buffer, writes the kStatAcc sums and
// then out.seq to out.flag.
__device__ inline void stat_deliver(const double* __restrict__ stats, const StatOut& out)
{
    for (int acc = threadIdx.x; acc < kStatAcc; acc += blockDim.x) {
        double s = 0.0;
#pragma unroll 16
        for (int sl = 0; sl < kStatSlots; ++sl) s += stats[stat_index(acc, sl)];
        out.sums[acc] = s;
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(out.flag, out.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

struct FFTPlan {
    int nfft;       // complex FFT length
    int nst;        // number of stages
    int radix[24];  // radix per stage (product = nfft)
};

// Host: factor n into stages, preferring radix 16, then 8, 4, 2, then odd primes.
inline FFTPlan make_fft_plan(int n)
{
    FFTPlan p;
    p.nfft = n;
    p.nst = 0;
    int m = n;
    while (m % 16 == 0) { p.radix[p.nst++] = 16; m /= 16; }
    if (m % 8 == 0) { p.radix[p.nst++] = 8; m /= 8; }
    if (m % 4 == 0) { p.radix[p.nst++] = 4; m /= 4; }
    if (m % 2 == 0) { p.radix[p.nst++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0 && p.nst < 24) { p.radix[p.nst++] = f; m /= f; }
        if ((long long)f * f > m && m > 1) {
            if (p.nst < 24) p.radix[p.nst++] = m;
            m = 1;
        }
    }
    return p;
}

// exp(-2 pi i m / 16), m = 0..15
__device__ constexpr double W16C[16] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
__device__ constexpr double W16S[16] = {
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173};

// multiply by exp(s 2 pi i m/16), s = -1 forward, +1 inverse (compile-time m)
template <int M, bool INV, typename C2>
__device__ inline C2 tw16(C2 a)
{
    using R = typename CT<C2>::R;
    constexpr int m = M & 15;
    if constexpr (m == 0) return a;
    else if constexpr (m == 4) {  // -i (fwd) / +i (inv)
        return INV ? CT<C2>::mk(-a.y, a.x) : CT<C2>::mk(a.y, -a.x);
    } else if constexpr (m == 8) {
        return CT<C2>::mk(-a.x, -a.y);
    } else if constexpr (m == 12) {
        return INV ? CT<C2>::mk(a.y, -a.x) : CT<C2>::mk(-a.y, a.x);
    } else {
        const R c = (R)W16C[m];
        const R s = (R)(INV ? -W16S[m] : W16S[m]);
        return CT<C2>::mk(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

template <bool INV, typename C2>
__device__ inline void dft2(C2& a0, C2& a1)
{
    C2 t = a0;
    a0 = cadd(t, a1);
    a1 = csub(t, a1);
}

template <bool INV, typename C2>
__device__ inline void dft4(C2& a0, C2& a1, C2& a2, C2& a3)
{
    C2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    C2 t2 = cadd(a1, a3), t3 = tw16<4, INV>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <bool INV, typename C2>
__device__ inline void dft8(C2* a)
{
    // 8 = 4 (r1) x 2 (r2): r = 2 r1 + r2, q = q1 + 4 q2
    dft4<INV>(a[0], a[2], a[4], a[6]);
    dft4<INV>(a[1], a[3], a[5], a[7]);
    a[3] = tw16<2, INV>(a[3]);
    a[5] = tw16<4, INV>(a[5]);
    a[7] = tw16<6, INV>(a[7]);
    C2 b[8];
#pragma unroll
    for (int q1 = 0; q1 < 4; ++q1) {
        C2 x0 = a[2 * q1], x1 = a[2 * q1 + 1];
        b[q1] = cadd(x0, x1);
        b[q1 + 4] = csub(x0, x1);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = b[i];
}

template <bool INV, typename C2>
__device__ inline void dft16(C2* a)
{
    // 16 = 4 x 4: r = 4 r1 + r2; DFT4 over r1 for each r2 -> q1, twiddle
    // w16^(r2 q1), then DFT4 over r2 for each q1 -> q2; q = q1 + 4 q2
    C2 c[16];
#pragma unroll
    for (int r2 = 0; r2 < 4; ++r2) {
        C2 x0 = a[r2], x1 = a[4 + r2], x2 = a[8 + r2], x3 = a[12 + r2];
        dft4<INV>(x0, x1, x2, x3);
        c[r2 * 4 + 0] = x0;
        c[r2 * 4 + 1] = x1;
        c[r2 * 4 + 2] = x2;
        c[r2 * 4 + 3] = x3;
    }
    c[1 * 4 + 1] = tw16<1, INV>(c[1 * 4 + 1]);
    c[1 * 4 + 2] = tw16<2, INV>(c[1 * 4 + 2]);
    c[1 * 4 + 3] = tw16<3, INV>(c[1 * 4 + 3]);
    c[2 * 4 + 1] = tw16<2, INV>(c[2 * 4 + 1]);
    c[2 * 4 + 2] = tw16<4, INV>(c[2 * 4 + 2]);
    c[2 * 4 + 3] = tw16<6, INV>(c[2 * 4 + 3]);
    c[3 * 4 + 1] = tw16<3, INV>(c[3 * 4 + 1]);
    c[3 * 4 + 2] = tw16<6, INV>(c[3 * 4 + 2]);
    c[3 * 4 + 3] = tw16<9, INV>(c[3 * 4 + 3]);
#pragma unroll
    for (int q1 = 0; q1 < 4; ++q1) {
        C2 x0 = c[0 * 4 + q1], x1 = c[1 * 4 + q1], x2 = c[2 * 4 + q1], x3 = c[3 * 4 + q1];
        dft4<INV>(x0, x1, x2, x3);
        a[q1] = x0;
        a[q1 + 4] = x1;
        a[q1 + 8] = x2;
        a[q1 + 12] = x3;
    }
}

// odd radices 3, 5, 7 by the direct formula with exact constants
template <int R, bool INV, typename C2>
__device__ inline void dft_odd(C2* a)
{
    using T = typename CT<C2>::R;
    constexpr double PI2 = 6.283185307179586476925286766559;
    C2 b[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        C2 acc = a[0];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const int m = (r * q) % R;
            const T c = (T)__builtin_cos(PI2 * m / R);
            const T s = (T)((INV ? 1.0 : -1.0) * __builtin_sin(PI2 * m / R));
            acc.x += a[r].x * c - a[r].y * s;
            acc.y += a[r].x * s + a[r].y * c;
        }
        b[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) a[q] = b[q];
}

template <bool INV, typename C2>
__device__ inline void dft32(C2* a);

// dft<R, INV>(a): in-register DFT of radix R
template <int R, bool INV, typename C2>
__device__ inline void dft(C2* a)
{
    if constexpr (R == 2) dft2<INV>(a[0], a[1]);
    else if constexpr (R == 4) dft4<INV>(a[0], a[1], a[2], a[3]);
    else if constexpr (R == 8) dft8<INV>(a);
    else if constexpr (R == 16) dft16<INV>(a);
    else if constexpr (R == 32) dft32<INV>(a);
    else dft_odd<R, INV>(a);
}

__device__ inline double2 twiddle(const double2* __restrict__ tw, int idx, bool inv)
{
    double2 w = tw[idx];
    if (inv) w.y = -w.y;
    return w;
}

// One out-of-place Stockham stage of radix R on T lines (line stride ldn).
// src/dst in LDS; tw = exp(-2 pi i m / n) table (global, n entries).
template <int R, bool INV>
__device__ inline void stockham_stage(const double2* __restrict__ src, double2* __restrict__ dst,
                                      int n, int ldn, int T, int Ns, const double2* __restrict__ tw,
                                      int tid, int nth)
{
    const int nb = n / R;
    const int tstride = n / (Ns * R);
    const int total = nb * T;
    for (int b = tid; b < total; b += nth) {
        const int t = b / nb;
        const int j = b - t * nb;
        const int k = j % Ns;
        const double2* s = src + t * ldn;
        double2 a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = s[j + r * nb];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) a[r] = cmul(a[r], twiddle(tw, r * k * tstride, INV));
        }
        dft<R, INV>(a);
        double2* d = dst + t * ldn;
        const int base = (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) d[base + q * Ns] = a[q];
    }
}

// Generic radix p stage (p any, O(p^2) per butterfly), reading LDS directly.
template <bool INV>
__device__ inline void stockham_stage_generic(const double2* __restrict__ src,
                                              double2* __restrict__ dst, int n, int ldn, int T,
                                              int Ns, int R, const double2* __restrict__ tw,
                                              int tid, int nth)
{
    const int nb = n / R;
    const int tstride = n / (Ns * R);
    const int pstride = n / R;  // exp(-2 pi i rq / R) = tw[((r q) % R) * pstride]
    const int total = nb * T * R;
    for (int e = tid; e < total; e += nth) {
        const int q = e % R;
        const int b = e / R;
        const int t = b / nb;
        const int j = b - t * nb;
        const int k = j % Ns;
        const double2* s = src + t * ldn;
        double2 acc = make_double2(0.0, 0.0);
        for (int r = 0; r < R; ++r) {
            double2 a = s[j + r * nb];
            if (Ns > 1 && r > 0) a = cmul(a, twiddle(tw, r * k * tstride, INV));
            const double2 w = twiddle(tw, ((r * q) % R) * pstride, INV);
            acc = cadd(acc, cmul(a, w));
        }
        dst[t * ldn + (j - k) * R + k + q * Ns] = acc;
    }
}

// Run all stages; A holds the input.  Returns the buffer holding the result
// (A or B).  Ends with a __syncthreads().
template <bool INV>
__device__ inline double2* run_fft(double2* A, double2* B, const FFTPlan& plan, int ldn, int T,
                                   const double2* __restrict__ tw, int tid, int nth)
{
    double2* src = A;
    double2* dst = B;
    int Ns = 1;
    const int n = plan.nfft;
    for (int s = 0; s < plan.nst; ++s) {
        const int R = plan.radix[s];
        switch (R) {
            case 16: stockham_stage<16, INV>(src, dst, n, ldn, T, Ns, tw, tid, nth); break;
            case 8: stockham_stage<8, INV>(src, dst, n, ldn, T, Ns, tw, tid, nth); break;
            case 4: stockham_stage<4, INV>(src, dst, n, ldn, T, Ns, tw, tid, nth); break;
            case 2: stockham_stage<2, INV>(src, dst, n, ldn, T, Ns, tw, tid, nth); break;
            case 3: stockham_stage<3, INV>(src, dst, n, ldn, T, Ns, tw, tid, nth); break;
            case 5: stockham_stage<5, INV>(src, dst, n, ldn, T, Ns, tw, tid, nth); break;
            case 7: stockham_stage<7, INV>(src, dst, n, ldn, T, Ns, tw, tid, nth); break;
            default: stockham_stage_generic<INV>(src, dst, n, ldn, T, Ns, R, tw, tid, nth); break;
        }
        __syncthreads();
        double2* tmp = src;
        src = dst;
        dst = tmp;
        Ns *= R;
    }
    return src;
}

}  // namespace sfa
