// Register-resident FFT of power-of-two length N in [16, 512], FP64.
//
// N = P * Q with P in {Q, 2Q}.  Each line is handled by Q threads; thread j
// holds P complex values in registers.  Pass 1: x[j + Q r] (r < P) -> DFT-P
// over r, twiddle w_N^(j k1); one exchange through LDS (padded, conflict free
// for the j-contiguous writes); pass 2: DFT-Q over j for k1 = h Q + j
// (h < P/Q).  On exit thread j holds X[k1 + P k2] in a[h Q + k2].
#pragma once
#include "fft_device.h"

namespace sfa {

template <int N>
struct Pow2 {
    static constexpr int Q = (N >= 256) ? 16 : (N >= 64 ? 8 : 4);
    static constexpr int P = N / Q;
    static constexpr int H = P / Q;
    static constexpr int LS = P * (Q + 1) + 1;  // exchange stride per line (complex)
    static_assert(P == Q || P == 2 * Q, "unsupported power-of-two length");
};

// exp(-2 pi i m / 32)
__device__ constexpr double W32C[32] = {
    1.0, 0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
    0.70710678118654752440, 0.55557023301960222474, 0.38268343236508977173, 0.19509032201612826785,
    0.0, -0.19509032201612826785, -0.38268343236508977173, -0.55557023301960222474,
    -0.70710678118654752440, -0.83146961230254523708, -0.92387953251128675613, -0.98078528040323044913,
    -1.0, -0.98078528040323044913, -0.92387953251128675613, -0.83146961230254523708,
    -0.70710678118654752440, -0.55557023301960222474, -0.38268343236508977173, -0.19509032201612826785,
    0.0, 0.19509032201612826785, 0.38268343236508977173, 0.55557023301960222474,
    0.70710678118654752440, 0.83146961230254523708, 0.92387953251128675613, 0.98078528040323044913};

template <int M, bool INV, typename C2>
__device__ inline C2 tw32(C2 a)
{
    using R = typename CT<C2>::R;
    constexpr int m = M & 31;
    if constexpr ((m & 1) == 0) {
        return tw16<(m >> 1), INV>(a);
    } else {
        const R c = (R)W32C[m];
        const R sfwd = (R)(-W32C[(m + 24) & 31]);  // sin(-2 pi m/32) = -cos(2 pi (m - 8)/32)
        const R s = INV ? -sfwd : sfwd;
        return CT<C2>::mk(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

template <bool INV, typename C2>
__device__ inline void dft32(C2* a)
{
    // 32 = 4 (r1) x 8 (r2): r = 8 r1 + r2; DFT-4 over r1 -> q1, twiddle
    // w32^(r2 q1), DFT-8 over r2 -> q2; q = q1 + 4 q2
    C2 c[32];
#pragma unroll
    for (int r2 = 0; r2 < 8; ++r2) {
        C2 x0 = a[r2], x1 = a[8 + r2], x2 = a[16 + r2], x3 = a[24 + r2];
        dft4<INV>(x0, x1, x2, x3);
        c[r2 * 4 + 0] = x0;
        c[r2 * 4 + 1] = x1;
        c[r2 * 4 + 2] = x2;
        c[r2 * 4 + 3] = x3;
    }
#define SFA_T32(R2)                                          \
    c[R2 * 4 + 1] = tw32<R2 * 1, INV>(c[R2 * 4 + 1]);        \
    c[R2 * 4 + 2] = tw32<R2 * 2, INV>(c[R2 * 4 + 2]);        \
    c[R2 * 4 + 3] = tw32<R2 * 3, INV>(c[R2 * 4 + 3]);
    SFA_T32(1) SFA_T32(2) SFA_T32(3) SFA_T32(4) SFA_T32(5) SFA_T32(6) SFA_T32(7)
#undef SFA_T32
#pragma unroll
    for (int q1 = 0; q1 < 4; ++q1) {
        C2 b[8];
#pragma unroll
        for (int r2 = 0; r2 < 8; ++r2) b[r2] = c[r2 * 4 + q1];
        dft8<INV>(b);
#pragma unroll
        for (int q2 = 0; q2 < 8; ++q2) a[q1 + 4 * q2] = b[q2];
    }
}

// First half of a line FFT: DFT-P over r in registers, twiddle w_N^(j k1),
// write to this line's LDS exchange area ex (Pow2<N>::LS complex).  tw: LDS
// table exp(-2 pi i m/N), m < N.
template <int N, bool INV, typename C2>
__device__ inline void pow2_fft_p1(C2 (&a)[Pow2<N>::P], int j, C2* ex, const C2* tw)
{
    constexpr int P = Pow2<N>::P, Q = Pow2<N>::Q;
    dft<P, INV>(a);
#pragma unroll
    for (int k1 = 1; k1 < P; ++k1) {
        C2 w = tw[j * k1];
        if (INV) w.y = -w.y;
        a[k1] = cmul(a[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) ex[k1 * (Q + 1) + j] = a[k1];
}

// First half with the twiddle table laid out k1-major, twT[k1 Q + j] =
// exp(-2 pi i j k1 / N) (same N entries): when the Q threads of a line are
// adjacent lanes they read consecutive entries (tw[j k1] strides by k1, a
// bank conflict for every even k1); lanes that share j still read one entry.
template <int N, bool INV, typename C2>
__device__ inline void pow2_fft_p1t(C2 (&a)[Pow2<N>::P], int j, C2* ex, const C2* twT)
{
    constexpr int P = Pow2<N>::P, Q = Pow2<N>::Q;
    dft<P, INV>(a);
#pragma unroll
    for (int k1 = 1; k1 < P; ++k1) {
        C2 w = twT[k1 * Q + j];
        if (INV) w.y = -w.y;
        a[k1] = cmul(a[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) ex[k1 * (Q + 1) + j] = a[k1];
}

// Second half: after the exchange is visible, thread j reads the DFT-Q inputs
// of k1 = h Q + j from ex and transforms them.  The thread j of this half may
// be a different thread of the block than the one of the first half (the
// exchange doubles as a change of thread mapping).
template <int N, bool INV, typename C2>
__device__ inline void pow2_fft_p2(C2 (&a)[Pow2<N>::P], int j, const C2* ex)
{
    constexpr int P = Pow2<N>::P, Q = Pow2<N>::Q;
#pragma unroll
    for (int h = 0; h < P / Q; ++h) {
        C2 b[Q];
        const int k1 = h * Q + j;
#pragma unroll
        for (int jj = 0; jj < Q; ++jj) b[jj] = ex[k1 * (Q + 1) + jj];
        dft<Q, INV>(b);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) a[h * Q + k2] = b[k2];
    }
}

// One line FFT.  ex: this line's LDS exchange area (Pow2<N>::LS complex);
// tw: LDS table exp(-2 pi i m/N), m < N.  Contains two __syncthreads(): every
// thread of the block must call it the same number of times.
template <int N, bool INV, typename C2>
__device__ inline void pow2_fft(C2 (&a)[Pow2<N>::P], int j, C2* ex, const C2* tw)
{
    pow2_fft_p1<N, INV>(a, j, ex, tw);
    __syncthreads();
    pow2_fft_p2<N, INV>(a, j, ex);
    __syncthreads();
}

// Order this wave's LDS accesses (lanes of one wave exchanging through LDS
// need no workgroup barrier).
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// index of a[h*Q + k2] after pow2_fft for thread j
template <int N>
__device__ inline int pow2_out_index(int j, int h, int k2)
{
    return (h * Pow2<N>::Q + j) + Pow2<N>::P * k2;
}

}  // namespace sfa
