// Banded solves along one axis of strided 1-3D arrays (FP64 factors,
// float64 or complex128 right-hand sides).  All band offsets are even, so the
// even and odd unknowns form two independent chains: one thread per
// (line, parity) chain, consecutive threads on consecutive lines so every
// step of the recurrence is a coalesced access across the wave.
#include "common.h"

namespace sfa {

__device__ inline double sub_mul(double a, double s, double b) { return a - s * b; }
__device__ inline double2 sub_mul(double2 a, double s, double2 b)
{
    return make_double2(a.x - s * b.x, a.y - s * b.y);
}
__device__ inline double mul(double a, double s) { return a * s; }
__device__ inline double2 mul(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline double divd(double a, double s) { return a / s; }
__device__ inline double2 divd(double2 a, double s) { return make_double2(a.x / s, a.y / s); }
__device__ inline double add(double a, double b) { return a + b; }
__device__ inline double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ inline T zero();
template <>
__device__ inline double zero<double>() { return 0.0; }
template <>
__device__ inline double2 zero<double2>() { return make_double2(0.0, 0.0); }

struct LineMap {
    int64_t sa, s1, s2, d2, nlines;
    __device__ inline int64_t off(int64_t L) const
    {
        const int64_t i1 = L / d2;
        return i1 * s1 + (L - i1 * d2) * s2;
    }
};

__device__ inline int last_of_parity(int n, int p) { return (n - 1) - (((n - 1) - p) & 1); }

// ----------------------------------------------------------------- TDMA
template <typename T>
__global__ void tdma_solve_kernel(T* __restrict__ u, LineMap g, const double* __restrict__ data,
                                  int n, int64_t dls, int64_t skip)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = g.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    if (L < skip || p >= n) return;
    T* v = u + g.off(L);
    const int64_t sa = g.sa;
    const double* ld = data + L * dls;
    const double* d = ld + n;
    const double* ud = ld + 2 * n + 2;
    T prev = v[p * sa];
    for (int i = p + 2; i < n; i += 2) {
        T cur = sub_mul(v[i * sa], ld[i - 2], prev);
        v[i * sa] = cur;
        prev = cur;
    }
    const int il = last_of_parity(n, p);
    T next = divd(v[il * sa], d[il]);
    v[il * sa] = next;
    for (int i = il - 2; i >= 0; i -= 2) {
        T cur = divd(sub_mul(v[i * sa], ud[i], next), d[i]);
        v[i * sa] = cur;
        next = cur;
    }
}

__global__ void tdma_lu_kernel(double* __restrict__ data, int n, int64_t nbatch)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 2 * nbatch) return;
    const int p = (int)(gid / nbatch);
    const int64_t b = gid - (int64_t)p * nbatch;
    double* ld = data + b * 3 * n;
    double* d = ld + n;
    double* ud = ld + 2 * n + 2;
    for (int i = p + 2; i < n; i += 2) {
        ld[i - 2] = ld[i - 2] / d[i - 2];
        d[i] = d[i] - ld[i - 2] * ud[i - 2];
    }
}

// ----------------------------------------------------------------- PDMA
template <typename T>
__global__ void pdma_solve_kernel(T* __restrict__ u, LineMap g, const double* __restrict__ data,
                                  int n, int64_t dls, int64_t skip)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = g.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    if (L < skip || p >= n) return;
    T* v = u + g.off(L);
    const int64_t sa = g.sa;
    const double* a = data + L * dls;
    const double* b = a + n;
    const double* d = a + 2 * n;
    const double* e = a + 3 * n + 2;
    const double* f = a + 4 * n + 4;
    // forward
    T pm2 = v[p * sa];  // u[i-2]
    T pm4 = zero<T>();  // u[i-4]
    if (p + 2 < n) {
        T cur = sub_mul(v[(p + 2) * sa], b[p], pm2);
        v[(p + 2) * sa] = cur;
        pm4 = pm2;
        pm2 = cur;
    }
    for (int k = p + 4; k < n; k += 2) {
        T s = add(mul(pm2, b[k - 2]), mul(pm4, a[k - 4]));
        T cur = add(v[k * sa], mul(s, -1.0));
        v[k * sa] = cur;
        pm4 = pm2;
        pm2 = cur;
    }
    // backward
    const int il = last_of_parity(n, p);
    T n2 = divd(v[il * sa], d[il]);  // u[k+2]
    v[il * sa] = n2;
    T n4 = zero<T>();                // u[k+4]
    if (il - 2 >= 0) {
        T cur = divd(sub_mul(v[(il - 2) * sa], e[il - 2], n2), d[il - 2]);
        v[(il - 2) * sa] = cur;
        n4 = n2;
        n2 = cur;
    }
    for (int k = il - 4; k >= 0; k -= 2) {
        T cur = divd(sub_mul(sub_mul(v[k * sa], e[k], n2), f[k], n4), d[k]);
        v[k * sa] = cur;
        n4 = n2;
        n2 = cur;
    }
}

__global__ void pdma_lu_kernel(double* __restrict__ data, int n, int64_t nbatch)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 2 * nbatch) return;
    const int p = (int)(gid / nbatch);
    const int64_t bb = gid - (int64_t)p * nbatch;
    double* a = data + bb * 5 * n;
    double* b = a + n;
    double* d = a + 2 * n;
    double* e = a + 3 * n + 2;
    double* f = a + 4 * n + 4;
    for (int i = p; i < n - 4; i += 2) {
        double lam = b[i] / d[i];
        d[i + 2] -= lam * e[i];
        e[i + 2] -= lam * f[i];
        b[i] = lam;
        lam = a[i] / d[i];
        b[i + 2] -= lam * e[i];
        d[i + 4] -= lam * f[i];
        a[i] = lam;
    }
    for (int i = n - 4; i <= n - 3; ++i) {
        if (i < 0 || ((i & 1) != p)) continue;
        const double lam = b[i] / d[i];
        d[i + 2] -= lam * e[i];
        b[i] = lam;
    }
}

// ----------------------------------------------------------------- DiagMA
template <typename T>
__global__ void diag_solve_kernel(T* __restrict__ u, LineMap g, const double* __restrict__ dd,
                                  int n, int64_t dls, int64_t skip)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = g.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    if (L < skip) return;
    T* v = u + g.off(L);
    const double* d = dd + L * dls;
    for (int i = p; i < n; i += 2) v[i * g.sa] = divd(v[i * g.sa], d[i]);
}

// ----------------------------------------------------------------- other band patterns
// TwoDMA (offsets 0, 2), ThreeDMA (0, 2, 4), FDMA (-2, 0, 2, 4) and HeptaDMA
// (-4, -2, 0, 2, 4, 6, 8) with S = 2 (two parity chains per line), and the
// odd-offset TDMA_O (-1, 0, 1) with S = 1 (one chain).  DIA rows in offset
// order: NL lower bands (row NL - l holds offset -S l by column, after the LU
// the unit-L multipliers), the diagonal (row NL), NU upper bands (row NL + k
// holds offset S k by column).  Forward elimination subtracts the summed lower
// terms, back substitution subtracts the upper terms one by one, as la.pyx's
// *_inner_solve_ptr loops do; terms beyond either end of the line are absent.
template <typename T, int NL, int NU, int S>
__global__ void band_solve_kernel(T* __restrict__ u, LineMap g, const double* __restrict__ data,
                                  int n, int64_t dls, int64_t skip)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = g.nlines;
    if (gid >= S * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    if (L < skip || p >= n) return;
    T* v = u + g.off(L);
    const int64_t sa = g.sa;
    const double* rows = data + L * dls;
    const double* d = rows + (int64_t)NL * n;
    if constexpr (NL > 0) {
        T w[NL];  // w[l] = u[i - S (l + 1)]
#pragma unroll
        for (int l = 0; l < NL; ++l) w[l] = zero<T>();
        w[0] = v[p * sa];
        for (int i = p + S; i < n; i += S) {
            T s = mul(w[0], rows[(int64_t)(NL - 1) * n + i - S]);
#pragma unroll
            for (int l = 1; l < NL; ++l) {
                const int c = i - S * (l + 1);
                if (c >= 0) s = add(s, mul(w[l], rows[(int64_t)(NL - 1 - l) * n + c]));
            }
            const T cur = add(v[i * sa], mul(s, -1.0));
            v[i * sa] = cur;
#pragma unroll
            for (int l = NL - 1; l > 0; --l) w[l] = w[l - 1];
            w[0] = cur;
        }
    }
    const int il = (S == 2) ? last_of_parity(n, p) : n - 1;
    T nx[NU];  // nx[k] = u[i + S (k + 1)]
#pragma unroll
    for (int k = 0; k < NU; ++k) nx[k] = zero<T>();
    for (int i = il; i >= 0; i -= S) {
        T acc = v[i * sa];
#pragma unroll
        for (int k = 1; k <= NU; ++k) {
            const int c = i + S * k;
            if (c < n) acc = sub_mul(acc, rows[(int64_t)(NL + k) * n + c], nx[k - 1]);
        }
        const T cur = divd(acc, d[i]);
        v[i * sa] = cur;
#pragma unroll
        for (int k = NU - 1; k > 0; --k) nx[k] = nx[k - 1];
        nx[0] = cur;
    }
}

// Doolittle LU without pivoting of a general band matrix with offsets
// -NL..NU (unit stride), in place on its DIA rows (A(i, j) at row NL + j - i,
// column j): the multipliers replace the lower bands, U the diagonal and the
// upper bands, the layout band_solve_kernel<T, NL, NU, 1> reads.  One thread
// per system.  Stands in for the reference's general sparse fallback
// (la.py:823-844, SuperLU) on the banded systems the Galerkin spaces give;
// without pivoting it needs nonzero pivots (SPD or diagonally dominant
// matrices, as the mass and Helmholtz matrices are).
template <int NL, int NU>
__global__ void band_lu_kernel(double* __restrict__ data, int n, int64_t nbatch)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbatch) return;
    double* r = data + b * (int64_t)(NL + 1 + NU) * n;
    for (int k = 0; k + 1 < n; ++k) {
        const double piv = r[(int64_t)NL * n + k];
#pragma unroll
        for (int l = 1; l <= NL; ++l) {
            if (k + l >= n) break;
            double* lik = r + (int64_t)(NL - l) * n + k;  // A(k + l, k)
            const double m = *lik / piv;
            *lik = m;
#pragma unroll
            for (int q = 1; q <= NU; ++q) {
                if (k + q >= n) break;
                // A(k + l, k + q) -= m A(k, k + q); offset q - l in [1 - NL, NU - 1]
                r[(int64_t)(NL + q - l) * n + k + q] -= m * r[(int64_t)(NL + q) * n + k + q];
            }
        }
    }
}

// la.pyx:254-263 TDMA_O_LU, one thread per system
__global__ void tdma_o_lu_kernel(double* __restrict__ data, int n, int64_t nbatch)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbatch) return;
    double* ld = data + b * 3 * n;
    double* d = ld + n;
    double* ud = ld + 2 * n + 1;
    for (int i = 1; i < n; ++i) {
        ld[i - 1] = ld[i - 1] / d[i - 1];
        d[i] = d[i] - ld[i - 1] * ud[i - 1];
    }
}

// la.pyx:199-211 FDMA_LU, one thread per (system, parity)
__global__ void fdma_lu_kernel(double* __restrict__ data, int n, int64_t nbatch)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 2 * nbatch) return;
    const int p = (int)(gid / nbatch);
    const int64_t b = gid - (int64_t)p * nbatch;
    double* ld = data + b * 4 * n;
    double* d = ld + n;
    double* u1 = ld + 2 * n + 2;
    double* u2 = ld + 3 * n + 4;
    for (int i = p + 2; i < n; i += 2) {
        ld[i - 2] = ld[i - 2] / d[i - 2];
        d[i] = d[i] - ld[i - 2] * u1[i - 2];
        if (i < n - 2) u1[i] = u1[i] - ld[i - 2] * u2[i - 2];
    }
}

// la.pyx:265-303 HeptaDMA_LU (band spacing 2), one thread per (system, parity)
__global__ void heptadma_lu_kernel(double* __restrict__ data, int n, int64_t nbatch)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 2 * nbatch) return;
    const int p = (int)(gid / nbatch);
    const int64_t bb = gid - (int64_t)p * nbatch;
    double* a = data + bb * 7 * n;
    double* b = a + n;
    double* d = a + 2 * n;
    double* e = a + 3 * n + 2;
    double* f = a + 4 * n + 4;
    double* g = a + 5 * n + 6;
    double* h = a + 6 * n + 8;
    for (int i = p; i < n - 4; i += 2) {
        double lam = b[i] / d[i];
        d[i + 2] -= lam * e[i];
        e[i + 2] -= lam * f[i];
        if (i < n - 6) f[i + 2] -= lam * g[i];
        if (i < n - 8) g[i + 2] -= lam * h[i];
        b[i] = lam;
        lam = a[i] / d[i];
        b[i + 2] -= lam * e[i];
        d[i + 4] -= lam * f[i];
        if (i < n - 6) e[i + 4] -= lam * g[i];
        if (i < n - 8) f[i + 4] -= lam * h[i];
        a[i] = lam;
    }
    for (int i = n - 4; i <= n - 3; ++i) {
        if ((i & 1) != p) continue;
        const double lam = b[i] / d[i];
        d[i + 2] -= lam * e[i];
        b[i] = lam;
    }
}

// ----------------------------------------------------------------- Helmholtz
struct HelmDiags {
    const double *A0, *A2, *A4, *Bm2, *B0, *B2;
    int M;
    int neumann;
};

__global__ void helmholtz_lu_kernel(HelmDiags h, const double* __restrict__ alfa,
                                    const double* __restrict__ beta, int64_t abs_, LineMap g,
                                    double* __restrict__ d0, double* __restrict__ d1,
                                    double* __restrict__ d2, double* __restrict__ Lf)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = g.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    const int M = h.M;
    if (p >= M) return;
    const int64_t o = g.off(L), sa = g.sa;
    const double al = alfa[L * abs_], be = beta[L * abs_];
    const bool nm = h.neumann != 0;
    const bool zb = nm && fabs(be) < 1e-8;
    auto a0 = [&](int i) {
        double v = h.A0[i];
        if (nm) {
            if (i == 0) { if (zb) v = 1.0 / al; }
            else v /= (double)i * i;
        }
        return v;
    };
    auto b0 = [&](int i) {
        double v = h.B0[i];
        if (nm) {
            if (i == 0) { if (zb) v = 0.0; }
            else v /= (double)i * i;
        }
        return v;
    };
    auto a2 = [&](int i) {
        if (i >= M - 2) return 0.0;
        double v = h.A2[i];
        if (nm) v /= (double)(i + 2) * (i + 2);
        return v;
    };
    auto b2 = [&](int i) {
        if (i >= M - 2) return 0.0;
        double v = h.B2[i];
        if (nm) v /= (double)(i + 2) * (i + 2);
        return v;
    };
    auto a4 = [&](int i) {
        if (i >= M - 4) return 0.0;
        double v = h.A4[i];
        if (nm) v /= (double)(i + 4) * (i + 4);
        return v;
    };
    auto bm2 = [&](int i) {
        double v = h.Bm2[i];
        if (nm && i >= 1) v /= (double)i * i;
        return v;
    };
    double d0p = al * a0(p) + be * b0(p);
    double d1p = al * a2(p) + be * b2(p);
    double d2p = al * a4(p);
    d0[o + p * sa] = d0p;
    if (p < M - 2) d1[o + p * sa] = d1p;
    if (p < M - 4) d2[o + p * sa] = d2p;
    for (int i = p + 2; i < M; i += 2) {
        const double l = be * bm2(i - 2) / d0p;
        Lf[o + (i - 2) * sa] = l;
        const double d0v = al * a0(i) + be * b0(i) - l * d1p;
        double d1v = 0.0, d2v = 0.0;
        if (i < M - 2) d1v = al * a2(i) + be * b2(i) - l * d2p;
        if (i < M - 4) d2v = al * a4(i) - l * d2p;
        d0[o + i * sa] = d0v;
        if (i < M - 2) d1[o + i * sa] = d1v;
        if (i < M - 4) d2[o + i * sa] = d2v;
        d0p = d0v;
        d1p = d1v;
        d2p = d2v;
    }
}

template <typename T>
__global__ void helmholtz_solve_kernel(const T* __restrict__ b, T* __restrict__ u, LineMap gb,
                                       LineMap gu, LineMap gf, int M, int neumann,
                                       const double* __restrict__ d0, const double* __restrict__ d1,
                                       const double* __restrict__ d2, const double* __restrict__ Lf)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = gu.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    if (p >= M) return;
    const T* bl = b + gb.off(L);
    T* ul = u + gu.off(L);
    const int64_t of = gf.off(L);
    const int64_t sb = gb.sa, su = gu.sa, sf = gf.sa;
    // forward elimination, y stored in u
    T yp = bl[p * sb];
    ul[p * su] = yp;
    for (int i = p + 2; i < M; i += 2) {
        T yv = sub_mul(bl[i * sb], Lf[of + (i - 2) * sf], yp);
        ul[i * su] = yv;
        yp = yv;
    }
    // back substitution with the running parity sum of the constant tail
    const int il = last_of_parity(M, p);
    T n2 = divd(ul[il * su], d0[of + il * sf]);  // u[i+2]
    ul[il * su] = n2;
    T n4 = zero<T>();                             // u[i+4]
    if (il - 2 >= 0) {
        const int i = il - 2;
        T cur = divd(sub_mul(ul[i * su], d1[of + i * sf], n2), d0[of + i * sf]);
        ul[i * su] = cur;
        n4 = n2;
        n2 = cur;
    }
    T S = zero<T>();
    for (int i = il - 4; i >= 0; i -= 2) {
        T cur = sub_mul(ul[i * su], d1[of + i * sf], n2);
        S = add(S, n4);
        cur = sub_mul(cur, d2[of + i * sf], S);
        cur = divd(cur, d0[of + i * sf]);
        ul[i * su] = cur;
        n4 = n2;
        n2 = cur;
    }
    if (neumann) {
        const double d00 = d0[of];
        if (p == 0 && (d00 - 1.0) * (d00 - 1.0) < 1e-16) ul[0] = zero<T>();
        for (int i = (p == 0 ? 2 : 1); i < M; i += 2) ul[i * su] = divd(ul[i * su], (double)i * i);
    }
}

// Long chains, few lines: one workgroup per (line, parity) chain.  The
// chain's right-hand side and factors are staged in LDS with coalesced loads
// (1/d0 formed in parallel there), one thread runs the two sweeps from LDS,
// and the workgroup stores the result; the thread-per-chain kernel above
// waits on global memory once per row, which costs milliseconds at N = 4096
// when there are too few lines to hide it.  Same recurrences
// (la.pyx:871-912), d0 applied as a multiplication by its reciprocal.
constexpr int HELM_LDS_MIN_M = 512;
constexpr int64_t HELM_LDS_MAX_LINES = 2048;

template <typename T>
__global__ void __launch_bounds__(256)
    helmholtz_solve_lds_kernel(const T* __restrict__ b, T* __restrict__ u, LineMap gb, LineMap gu,
                               LineMap gf, int M, int neumann, const double* __restrict__ d0,
                               const double* __restrict__ d1, const double* __restrict__ d2,
                               const double* __restrict__ Lf)
{
    extern __shared__ __align__(16) unsigned char sm_raw[];
    const int64_t nl = gu.nlines;
    const int p = (int)(blockIdx.x / nl);
    const int64_t L = blockIdx.x - (int64_t)p * nl;
    const int K = (M - p + 1) / 2;  // rows i = p + 2 m, m < K
    T* y = reinterpret_cast<T*>(sm_raw);
    double* Ls = reinterpret_cast<double*>(y + K);
    double* r0 = Ls + K;
    double* e1 = r0 + K;
    double* e2 = e1 + K;
    const T* bl = b + gb.off(L);
    T* ul = u + gu.off(L);
    const int64_t of = gf.off(L);
    const int64_t sb = gb.sa, su = gu.sa, sf = gf.sa;
    for (int m = threadIdx.x; m < K; m += blockDim.x) {
        const int64_t i = p + 2 * m;
        y[m] = bl[i * sb];
        Ls[m] = m > 0 ? Lf[of + (i - 2) * sf] : 0.0;
        r0[m] = 1.0 / d0[of + i * sf];
        e1[m] = d1[of + i * sf];
        e2[m] = d2[of + i * sf];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        T yp = y[0];
#pragma unroll 8
        for (int m = 1; m < K; ++m) {
            yp = sub_mul(y[m], Ls[m], yp);
            y[m] = yp;
        }
        T n2 = mul(y[K - 1], r0[K - 1]);
        y[K - 1] = n2;
        T n4 = zero<T>();
        if (K >= 2) {
            const T cur = mul(sub_mul(y[K - 2], e1[K - 2], n2), r0[K - 2]);
            y[K - 2] = cur;
            n4 = n2;
            n2 = cur;
        }
        T S = zero<T>();
#pragma unroll 8
        for (int m = K - 3; m >= 0; --m) {
            T cur = sub_mul(y[m], e1[m], n2);
            S = add(S, n4);
            cur = sub_mul(cur, e2[m], S);
            cur = mul(cur, r0[m]);
            y[m] = cur;
            n4 = n2;
            n2 = cur;
        }
    }
    __syncthreads();
    for (int m = threadIdx.x; m < K; m += blockDim.x) {
        const int i = p + 2 * m;
        T v = y[m];
        if (neumann) {
            if (i == 0) {
                const double d00 = d0[of];
                if ((d00 - 1.0) * (d00 - 1.0) < 1e-16) v = zero<T>();
            } else {
                v = divd(v, (double)i * i);
            }
        }
        ul[(int64_t)i * su] = v;
    }
}

template <typename T>
__global__ void helmholtz_matvec_kernel(const T* __restrict__ v, T* __restrict__ c, LineMap gv,
                                        LineMap gc, int M, int neumann,
                                        const double* __restrict__ A0, const double* __restrict__ At,
                                        const double* __restrict__ Bm2, const double* __restrict__ B0,
                                        const double* __restrict__ B2, const double* __restrict__ alfa,
                                        const double* __restrict__ beta, int64_t abs_)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = gv.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    const T* vl = v + gv.off(L);
    T* cl = c + gc.off(L);
    const int64_t sv = gv.sa, sc = gc.sa;
    const double al = alfa[L * abs_], be = beta[L * abs_];
    cl[(M + p) * sc] = zero<T>();  // BC slots N-2, N-1
    if (p >= M) return;
    const int il = last_of_parity(M, p);
    T S = zero<T>();  // sum_{j>=1} w_{k+2j} v_{k+2j}
    for (int k = il; k >= 0; k -= 2) {
        if (k + 2 < M) {
            const double w = neumann ? (double)(k + 2) * (k + 2) : 1.0;
            S = add(S, mul(vl[(k + 2) * sv], w));
        }
        T r = mul(vl[k * sv], al * A0[k] + be * B0[k]);
        if (k >= 2) r = add(r, mul(vl[(k - 2) * sv], be * Bm2[k - 2]));
        if (k + 2 < M) {
            r = add(r, mul(vl[(k + 2) * sv], be * B2[k]));
            r = add(r, mul(S, al * At[k]));
        }
        cl[k * sc] = r;
    }
}

template <typename T>
__global__ void poisson_add_kernel(const T* __restrict__ b, T* __restrict__ u, LineMap gb,
                                   LineMap gu, int M, const double* __restrict__ d,
                                   const double* __restrict__ d1, double scale)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = gu.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    if (p >= M) return;
    const T* bl = b + gb.off(L);
    T* ul = u + gu.off(L);
    const int64_t sb = gb.sa, su = gu.sa;
    const int il = last_of_parity(M, p);
    T n2 = divd(bl[il * sb], d[il]);
    ul[il * su] = n2;
    T S = zero<T>();
    for (int k = il - 2; k >= 0; k -= 2) {
        S = add(S, n2);
        T cur = divd(sub_mul(bl[k * sb], d1[k], S), d[k]);
        ul[k * su] = cur;
        n2 = cur;
    }
    if (!(fabs(scale - 1.0) < 1e-8))
        for (int k = p; k < M; k += 2) ul[k * su] = divd(ul[k * su], scale);
}

// ----------------------------------------------------------------- host
static int line_map(int ndim, const int64_t* shape, const int64_t* strides, int axis, LineMap* m,
                    int64_t* n)
{
    Geom g;
    int rc = make_geom(ndim, shape, strides, axis, &g);
    if (rc) return rc;
    m->sa = g.sa;
    m->s1 = g.s1;
    m->s2 = g.s2;
    m->d2 = g.d2;
    m->nlines = g.d1 * g.d2;
    if (n) *n = g.n;
    return SFA_OK;
}

static inline dim3 grid_for(int64_t nthreads, int bs)
{
    return dim3((unsigned)((nthreads + bs - 1) / bs));
}

}  // namespace sfa

using namespace sfa;

template <int KIND>
static int banded_solve(void* u, int dtype, int ndim, const int64_t* shape, const int64_t* strides,
                        int axis, const double* data, int64_t n, int64_t dls, int64_t skip,
                        void* stream)
{
    LineMap g;
    int64_t len;
    int rc = line_map(ndim, shape, strides, axis, &g, &len);
    if (rc) return rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "bad dtype %d", dtype);
    SFA_CHECK(n >= 1 && n <= len, SFA_ERR_ARG, "system size %lld must be in [1, %lld]",
              (long long)n, (long long)len);
    SFA_CHECK(KIND == 0 || n >= (KIND == 1 ? 2 : 4), SFA_ERR_ARG, "system too small (n=%lld)",
              (long long)n);
    SFA_CHECK(dls >= 0 && skip >= 0, SFA_ERR_ARG, "negative stride/skip");
    const int bs = 256;
    dim3 grid = grid_for(2 * g.nlines, bs);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == SFA_C128) {
        double2* uu = (double2*)u;
        if (KIND == 0) hipLaunchKernelGGL(diag_solve_kernel<double2>, grid, dim3(bs), 0, s, uu, g, data, (int)n, dls, skip);
        if (KIND == 1) hipLaunchKernelGGL(tdma_solve_kernel<double2>, grid, dim3(bs), 0, s, uu, g, data, (int)n, dls, skip);
        if (KIND == 2) hipLaunchKernelGGL(pdma_solve_kernel<double2>, grid, dim3(bs), 0, s, uu, g, data, (int)n, dls, skip);
    } else {
        double* uu = (double*)u;
        if (KIND == 0) hipLaunchKernelGGL(diag_solve_kernel<double>, grid, dim3(bs), 0, s, uu, g, data, (int)n, dls, skip);
        if (KIND == 1) hipLaunchKernelGGL(tdma_solve_kernel<double>, grid, dim3(bs), 0, s, uu, g, data, (int)n, dls, skip);
        if (KIND == 2) hipLaunchKernelGGL(pdma_solve_kernel<double>, grid, dim3(bs), 0, s, uu, g, data, (int)n, dls, skip);
    }
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_tdma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                              const int64_t* strides, int axis, const double* data, int64_t n,
                              int64_t data_line_stride, int64_t skip_lines, void* stream)
{
    return banded_solve<1>(u, dtype, ndim, shape, strides, axis, data, n, data_line_stride,
                           skip_lines, stream);
}

extern "C" int sfa_pdma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                              const int64_t* strides, int axis, const double* data, int64_t n,
                              int64_t data_line_stride, int64_t skip_lines, void* stream)
{
    return banded_solve<2>(u, dtype, ndim, shape, strides, axis, data, n, data_line_stride,
                           skip_lines, stream);
}

extern "C" int sfa_diagma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                                const int64_t* strides, int axis, const double* d, int64_t n,
                                int64_t data_line_stride, int64_t skip_lines, void* stream)
{
    return banded_solve<0>(u, dtype, ndim, shape, strides, axis, d, n, data_line_stride,
                           skip_lines, stream);
}

extern "C" int sfa_tdma_lu(double* data, int64_t n, int64_t nbatch, void* stream)
{
    SFA_CHECK(n >= 2 && nbatch >= 1, SFA_ERR_ARG, "bad TDMA size n=%lld nbatch=%lld", (long long)n,
              (long long)nbatch);
    hipLaunchKernelGGL(tdma_lu_kernel, grid_for(2 * nbatch, 256), dim3(256), 0, (hipStream_t)stream,
                       data, (int)n, nbatch);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

template <int NL, int NU, int S>
static int band_solve(const char* name, int64_t nmin, void* u, int dtype, int ndim,
                      const int64_t* shape, const int64_t* strides, int axis, const double* data,
                      int64_t n, int64_t dls, int64_t skip, void* stream)
{
    LineMap g;
    int64_t len;
    int rc = line_map(ndim, shape, strides, axis, &g, &len);
    if (rc) return rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "%s: bad dtype %d", name, dtype);
    SFA_CHECK(n >= nmin && n <= len, SFA_ERR_ARG, "%s: system size %lld must be in [%lld, %lld]",
              name, (long long)n, (long long)nmin, (long long)len);
    SFA_CHECK(n <= (1 << 26), SFA_ERR_ARG, "%s: system size %lld too large", name, (long long)n);
    SFA_CHECK(data != nullptr && dls >= 0 && skip >= 0, SFA_ERR_ARG,
              "%s: null factors or negative stride/skip", name);
    if (g.nlines == 0) return SFA_OK;
    const int bs = 256;
    const dim3 grid = grid_for(S * g.nlines, bs);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == SFA_C128)
        hipLaunchKernelGGL((band_solve_kernel<double2, NL, NU, S>), grid, dim3(bs), 0, s,
                           (double2*)u, g, data, (int)n, dls, skip);
    else
        hipLaunchKernelGGL((band_solve_kernel<double, NL, NU, S>), grid, dim3(bs), 0, s,
                           (double*)u, g, data, (int)n, dls, skip);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

#define SFA_BAND_SOLVE(fn, NL, NU, S, NMIN)                                                      \
    extern "C" int fn(void* u, int dtype, int ndim, const int64_t* shape,                        \
                      const int64_t* strides, int axis, const double* data, int64_t n,           \
                      int64_t data_line_stride, int64_t skip_lines, void* stream)                \
    {                                                                                            \
        return band_solve<NL, NU, S>(#fn, NMIN, u, dtype, ndim, shape, strides, axis, data, n,   \
                                     data_line_stride, skip_lines, stream);                      \
    }
SFA_BAND_SOLVE(sfa_tdma_o_solve, 1, 1, 1, 1)
SFA_BAND_SOLVE(sfa_twodma_solve, 0, 1, 2, 2)
SFA_BAND_SOLVE(sfa_threedma_solve, 0, 2, 2, 4)
SFA_BAND_SOLVE(sfa_fdma_solve, 1, 2, 2, 4)
SFA_BAND_SOLVE(sfa_heptadma_solve, 2, 4, 2, 8)
#undef SFA_BAND_SOLVE

static int band_lu(const char* name, void (*kern)(double*, int, int64_t), int chains, int64_t nmin,
                   double* data, int64_t n, int64_t nbatch, void* stream)
{
    SFA_CHECK(data != nullptr && n >= nmin && n <= (1 << 26) && nbatch >= 1, SFA_ERR_ARG,
              "bad %s size n=%lld nbatch=%lld (n >= %lld)", name, (long long)n,
              (long long)nbatch, (long long)nmin);
    hipLaunchKernelGGL(kern, grid_for(chains * nbatch, 256), dim3(256), 0, (hipStream_t)stream,
                       data, (int)n, nbatch);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// general banded systems, offsets -nl..nu (0 <= nl <= 4, 1 <= nu <= 4), unit stride
using BandSolveFn = int (*)(const char*, int64_t, void*, int, int, const int64_t*,
                            const int64_t*, int, const double*, int64_t, int64_t, int64_t, void*);
#define SFA_ROW(NL) {band_solve<NL, 1, 1>, band_solve<NL, 2, 1>, band_solve<NL, 3, 1>,           \
                     band_solve<NL, 4, 1>}
static const BandSolveFn kBandSolve[5][4] = {SFA_ROW(0), SFA_ROW(1), SFA_ROW(2), SFA_ROW(3),
                                             SFA_ROW(4)};
#undef SFA_ROW
#define SFA_ROW(NL) {band_lu_kernel<NL, 1>, band_lu_kernel<NL, 2>, band_lu_kernel<NL, 3>,        \
                     band_lu_kernel<NL, 4>}
static void (*const kBandLU[5][4])(double*, int, int64_t) = {SFA_ROW(0), SFA_ROW(1), SFA_ROW(2),
                                                               SFA_ROW(3), SFA_ROW(4)};
#undef SFA_ROW

extern "C" int sfa_band_lu(double* data, int64_t n, int nl, int nu, int64_t nbatch, void* stream)
{
    SFA_CHECK(nl >= 0 && nl <= 4 && nu >= 1 && nu <= 4, SFA_ERR_ARG,
              "band_lu: bands -%d..%d outside -4..4 (nu >= 1)", nl, nu);
    return band_lu("band", kBandLU[nl][nu - 1], 1, 1, data, n, nbatch, stream);
}

extern "C" int sfa_band_solve(void* u, int dtype, int ndim, const int64_t* shape,
                              const int64_t* strides, int axis, const double* data, int64_t n,
                              int64_t data_line_stride, int64_t skip_lines, int nl, int nu,
                              void* stream)
{
    SFA_CHECK(nl >= 0 && nl <= 4 && nu >= 1 && nu <= 4, SFA_ERR_ARG,
              "band_solve: bands -%d..%d outside -4..4 (nu >= 1)", nl, nu);
    return kBandSolve[nl][nu - 1]("sfa_band_solve", 1, u, dtype, ndim, shape, strides, axis, data,
                                  n, data_line_stride, skip_lines, stream);
}

extern "C" int sfa_tdma_o_lu(double* data, int64_t n, int64_t nbatch, void* stream)
{
    return band_lu("TDMA_O", tdma_o_lu_kernel, 1, 1, data, n, nbatch, stream);
}

extern "C" int sfa_fdma_lu(double* data, int64_t n, int64_t nbatch, void* stream)
{
    return band_lu("FDMA", fdma_lu_kernel, 2, 4, data, n, nbatch, stream);
}

extern "C" int sfa_heptadma_lu(double* data, int64_t n, int64_t nbatch, void* stream)
{
    return band_lu("HeptaDMA", heptadma_lu_kernel, 2, 8, data, n, nbatch, stream);
}

extern "C" int sfa_pdma_lu(double* data, int64_t n, int64_t nbatch, void* stream)
{
    SFA_CHECK(n >= 5 && nbatch >= 1, SFA_ERR_ARG, "bad PDMA size n=%lld nbatch=%lld", (long long)n,
              (long long)nbatch);
    hipLaunchKernelGGL(pdma_lu_kernel, grid_for(2 * nbatch, 256), dim3(256), 0, (hipStream_t)stream,
                       data, (int)n, nbatch);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_helmholtz_lu(int neumann, int64_t N, const double* A0, const double* A2,
                                const double* A4, const double* Bm2, const double* B0,
                                const double* B2, const double* alfa, const double* beta,
                                int64_t ab_stride, double* d0, double* d1, double* d2, double* L,
                                int ndim, const int64_t* shape, const int64_t* f_strides, int axis,
                                void* stream)
{
    LineMap g;
    int64_t len;
    int rc = line_map(ndim, shape, f_strides, axis, &g, &len);
    if (rc) return rc;
    SFA_CHECK(len == N, SFA_ERR_ARG, "factor axis length %lld != N %lld", (long long)len, (long long)N);
    SFA_CHECK(N >= 8, SFA_ERR_ARG, "Helmholtz solver needs N >= 8 (got %lld)", (long long)N);
    SFA_CHECK(ab_stride == 0 || ab_stride == 1, SFA_ERR_ARG, "ab_stride must be 0 or 1");
    HelmDiags h{A0, A2, A4, Bm2, B0, B2, (int)(N - 2), neumann};
    hipLaunchKernelGGL(helmholtz_lu_kernel, grid_for(2 * g.nlines, 256), dim3(256), 0,
                       (hipStream_t)stream, h, alfa, beta, ab_stride, g, d0, d1, d2, L);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_helmholtz_solve(const void* b, void* u, int dtype, int ndim, const int64_t* shape,
                                   const int64_t* b_strides, const int64_t* u_strides, int axis,
                                   int neumann, const double* d0, const double* d1,
                                   const double* d2, const double* L, const int64_t* f_strides,
                                   void* stream)
{
    LineMap gb, gu, gf;
    int64_t N;
    int rc;
    if ((rc = line_map(ndim, shape, b_strides, axis, &gb, &N))) return rc;
    if ((rc = line_map(ndim, shape, u_strides, axis, &gu, nullptr))) return rc;
    if ((rc = line_map(ndim, shape, f_strides, axis, &gf, nullptr))) return rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "bad dtype %d", dtype);
    SFA_CHECK(N >= 8, SFA_ERR_ARG, "Helmholtz solver needs N >= 8 (got %lld)", (long long)N);
    const int M = (int)N - 2;
    const size_t elem = dtype == SFA_C128 ? sizeof(double2) : sizeof(double);
    const size_t lds = (size_t)((M + 1) / 2) * (elem + 4 * sizeof(double));
    if (M >= HELM_LDS_MIN_M && gu.nlines <= HELM_LDS_MAX_LINES && lds <= 160 * 1024) {
        dim3 grid((unsigned)(2 * gu.nlines));
        if (dtype == SFA_C128) {
            auto k = helmholtz_solve_lds_kernel<double2>;
            if (lds > 64 * 1024)
                hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL(k, grid, dim3(256), lds, (hipStream_t)stream, (const double2*)b,
                               (double2*)u, gb, gu, gf, M, neumann, d0, d1, d2, L);
        } else {
            auto k = helmholtz_solve_lds_kernel<double>;
            if (lds > 64 * 1024)
                hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL(k, grid, dim3(256), lds, (hipStream_t)stream, (const double*)b,
                               (double*)u, gb, gu, gf, M, neumann, d0, d1, d2, L);
        }
        SFA_LAUNCH_CHECK();
        return SFA_OK;
    }
    dim3 grid = grid_for(2 * gu.nlines, 256);
    if (dtype == SFA_C128)
        hipLaunchKernelGGL(helmholtz_solve_kernel<double2>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double2*)b, (double2*)u, gb, gu, gf, M, neumann, d0, d1, d2, L);
    else
        hipLaunchKernelGGL(helmholtz_solve_kernel<double>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double*)b, (double*)u, gb, gu, gf, M, neumann, d0, d1, d2, L);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_helmholtz_matvec(const void* v, void* c, int dtype, int ndim,
                                    const int64_t* shape, const int64_t* v_strides,
                                    const int64_t* c_strides, int axis, int neumann,
                                    const double* A0, const double* At, const double* Bm2,
                                    const double* B0, const double* B2, const double* alfa,
                                    const double* beta, int64_t ab_stride, void* stream)
{
    LineMap gv, gc;
    int64_t N;
    int rc;
    if ((rc = line_map(ndim, shape, v_strides, axis, &gv, &N))) return rc;
    if ((rc = line_map(ndim, shape, c_strides, axis, &gc, nullptr))) return rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "bad dtype %d", dtype);
    SFA_CHECK(N >= 4, SFA_ERR_ARG, "matvec needs N >= 4");
    SFA_CHECK(v != c, SFA_ERR_ARG, "matvec cannot run in place");
    SFA_CHECK(ab_stride == 0 || ab_stride == 1, SFA_ERR_ARG, "ab_stride must be 0 or 1");
    const int M = (int)N - 2;
    dim3 grid = grid_for(2 * gv.nlines, 256);
    if (dtype == SFA_C128)
        hipLaunchKernelGGL(helmholtz_matvec_kernel<double2>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double2*)v, (double2*)c, gv, gc, M, neumann, A0, At, Bm2, B0, B2,
                           alfa, beta, ab_stride);
    else
        hipLaunchKernelGGL(helmholtz_matvec_kernel<double>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double*)v, (double*)c, gv, gc, M, neumann, A0, At, Bm2, B0, B2,
                           alfa, beta, ab_stride);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_poisson_add_solve(const void* b, void* u, int dtype, int ndim,
                                     const int64_t* shape, const int64_t* b_strides,
                                     const int64_t* u_strides, int axis, const double* A0,
                                     const double* A2, double scale, void* stream)
{
    LineMap gb, gu;
    int64_t N;
    int rc;
    if ((rc = line_map(ndim, shape, b_strides, axis, &gb, &N))) return rc;
    if ((rc = line_map(ndim, shape, u_strides, axis, &gu, nullptr))) return rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "bad dtype %d", dtype);
    SFA_CHECK(N >= 4, SFA_ERR_ARG, "Poisson solver needs N >= 4");
    const int M = (int)N - 2;
    dim3 grid = grid_for(2 * gu.nlines, 256);
    if (dtype == SFA_C128)
        hipLaunchKernelGGL(poisson_add_kernel<double2>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double2*)b, (double2*)u, gb, gu, M, A0, A2, scale);
    else
        hipLaunchKernelGGL(poisson_add_kernel<double>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double*)b, (double*)u, gb, gu, M, A0, A2, scale);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// ----------------------------------------------------------------- banded matvec
// c(.., i, ..) (+)= s(line) sum_r D[r][i] v(.., i + off_r, ..) for i < nrows,
// 0 <= i + off_r < ncols (SparseMatrix.matvec, matrixbase.py:87-144, with the
// per-mode scale of TPMatrix.matvec :1437-1449).  D holds the diagonals by row
// (D[r][i] = A[i][i + off_r]); without accumulate, rows >= nrows of c are zeroed.
namespace sfa {
template <typename T>
__global__ void band_matvec_kernel(const T* __restrict__ v, T* __restrict__ c, LineMap gv,
                                   LineMap gc, int nout, int nrows, int ncols, int noffs,
                                   const int64_t* __restrict__ offs, const double* __restrict__ D,
                                   const double* __restrict__ scales, int accumulate)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= gc.nlines * nout) return;
    const int i = (int)(gid % nout);
    const int64_t L = gid / nout;
    const T* vl = v + gv.off(L);
    T* cl = c + gc.off(L) + (int64_t)i * gc.sa;
    T acc = zero<T>();
    if (i < nrows) {
        for (int r = 0; r < noffs; ++r) {
            const int64_t j = i + offs[r];
            if (j >= 0 && j < ncols) acc = add(acc, mul(vl[(int64_t)j * gv.sa], D[(int64_t)r * nrows + i]));
        }
        if (scales) acc = mul(acc, scales[L]);
    } else if (accumulate) {
        return;
    }
    *cl = accumulate ? add(*cl, acc) : acc;
}
}  // namespace sfa

extern "C" int sfa_band_matvec(const void* v, void* c, int dtype, int ndim, const int64_t* v_shape,
                               const int64_t* v_strides, const int64_t* c_shape,
                               const int64_t* c_strides, int axis, int64_t nrows, int64_t ncols,
                               int noffs, const int64_t* offsets, const double* D,
                               const double* scales, int accumulate, void* stream)
{
    LineMap gv, gc;
    int64_t nv, nc;
    int rc = line_map(ndim, v_shape, v_strides, axis, &gv, &nv);
    if (rc) return rc;
    rc = line_map(ndim, c_shape, c_strides, axis, &gc, &nc);
    if (rc) return rc;
    SFA_CHECK(v != nullptr && c != nullptr && D != nullptr && offsets != nullptr && v != c,
              SFA_ERR_ARG, "band_matvec: null or aliased arrays");
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "band_matvec: bad dtype %d", dtype);
    for (int d = 0; d < ndim; ++d)
        SFA_CHECK(d == axis || v_shape[d] == c_shape[d], SFA_ERR_ARG,
                  "band_matvec: shapes differ off the axis (dim %d)", d);
    SFA_CHECK(noffs >= 1, SFA_ERR_ARG, "band_matvec: no diagonals");
    SFA_CHECK(nrows >= 1 && nrows <= nc && ncols >= 1 && ncols <= nv && nc <= (1 << 28), SFA_ERR_ARG,
              "band_matvec: matrix %lld x %lld does not fit the arrays (%lld, %lld)", (long long)nrows,
              (long long)ncols, (long long)nc, (long long)nv);
    const int64_t total = gc.nlines * nc;
    if (total == 0) return SFA_OK;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (dtype == SFA_C128)
        hipLaunchKernelGGL(band_matvec_kernel<double2>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double2*)v, (double2*)c, gv, gc, (int)nc, (int)nrows, (int)ncols, noffs,
                           offsets, D, scales, accumulate);
    else
        hipLaunchKernelGGL(band_matvec_kernel<double>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double*)v, (double*)c, gv, gc, (int)nc, (int)nrows, (int)ncols, noffs,
                           offsets, D, scales, accumulate);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}
