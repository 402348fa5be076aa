// Fused (alpha - lap) u = f solve on Chebyshev(SD/SN) x C2C x R2C, FP64.
//
// Pipeline over the complex workspace W (Nx, Ny, Nz/2+1):
//   z: R2C (scale 1/Nz)            f -> W
//   y: C2C forward (scale 1/Ny)    W -> W
//   x: fused x-pass                W -> W
//        DCT-II * pi/(2Nx) -> composite stencil -> Helmholtz LU computed on
//        the fly (alfa = -1, beta = alpha + ky^2 + kz^2) + solve ->
//        to_ortho -> DCT-III * 0.5 + t_0/2, all in LDS
//   y: C2C inverse                 W -> W
//   z: C2R                         W -> u
#include <math.h>

#include <map>
#include <mutex>
#include <vector>

#include "fft_device.h"

namespace sfa {

struct HelmTables {
    const double *A0, *A2, *A4, *Bm2, *B0, *B2;  // unscaled diagonals, M = N-2
};

// Closed-form Chebyshev SD/SN Galerkin diagonals (chebyshev/matrices.py:221-281,
// 377-427, 1047-1165), uploaded once per (N, neumann) and kept for the process.
static std::mutex g_helm_mu;
static std::map<std::pair<int, int>, double*> g_helm_tabs;  // (N*2+neumann, device) -> 6*M doubles

static const double* helm_tables(int N, int neumann)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_helm_mu);
    auto key = std::make_pair(N * 2 + neumann, dev);
    auto it = g_helm_tabs.find(key);
    if (it != g_helm_tabs.end()) return it->second;
    const int M = N - 2;
    const double pi = 3.14159265358979323846;
    std::vector<double> h(6 * (size_t)M, 0.0);
    double* A0 = h.data();
    double* A2 = A0 + M;
    double* A4 = A2 + M;
    double* Bm2 = A4 + M;
    double* B0 = Bm2 + M;
    double* B2 = B0 + M;
    std::vector<double> hn(N, pi / 2);  // Chebyshev GC norms
    hn[0] = pi;
    for (int k = 0; k < M; ++k) {
        if (!neumann) {
            A0[k] = -2 * pi * (k + 1) * (k + 2);
            if (k < M - 2) A2[k] = -4 * pi * (k + 1);
            if (k < M - 4) A4[k] = -4 * pi * (k + 1);
            B0[k] = hn[k] + hn[k + 2];
            if (k < M - 2) { B2[k] = -hn[1]; Bm2[k] = -hn[1]; }
        } else {
            const double kk = k;
            A0[k] = -2 * pi * kk * kk * (kk + 1) / (kk + 2);
            if (k < M - 2) A2[k] = -4 * pi * (kk + 2) * (kk + 2) * (kk + 1) / ((kk + 2) * (kk + 2));
            if (k < M - 4) A4[k] = -4 * pi * (kk + 4) * (kk + 4) * (kk + 1) / ((kk + 2) * (kk + 2));
            const double al = (kk / (kk + 2)) * (kk / (kk + 2));
            const double al2 = ((kk + 2) / (kk + 4)) * ((kk + 2) / (kk + 4));
            (void)al2;
            B0[k] = hn[k] + al * al * hn[k + 2];
            if (k < M - 2) { B2[k] = -al * hn[k + 2]; Bm2[k] = B2[k]; }
        }
    }
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    g_helm_tabs[key] = d;
    return d;
}

struct XPassArgs {
    double2* W;
    int64_t sa, s1;  // x stride, y stride (z stride 1)
    int64_t d1, d2;  // Ny, Nz/2+1
    int N, ldn, T;
    int neumann;
    double alpha;
    int Ny;
    const double* K2;      // N-2 stencil entries
    const double* tabs;    // 6*(N-2) diagonals
    const double2* tw;     // FFT table N
    const double2* tq;     // quarter table N
    FFTPlan plan;
};

__device__ inline int perm_dct(int i, int N) { return (i & 1) ? (N - 1 - (i >> 1)) : (i >> 1); }

__global__ void __launch_bounds__(256) xpass_kernel(XPassArgs a)
{
    extern __shared__ double2 smem[];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int64_t ntile = (a.d2 + a.T - 1) / a.T;
    const int64_t jy = blockIdx.x / ntile;
    const int64_t jz0 = (blockIdx.x - jy * ntile) * a.T;
    const int Tv = (int)((a.d2 - jz0) < a.T ? (a.d2 - jz0) : a.T);
    const int N = a.N, ldn = a.ldn, M = N - 2;
    double2* A = smem;
    double2* B = smem + (size_t)a.T * ldn;
    double* t0s = reinterpret_cast<double*>(smem + (size_t)2 * a.T * ldn);
    double2* W = a.W + jy * a.s1 + jz0;

    // 1. load with the DCT-II even/odd permutation
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e % Tv, i = e / Tv;
        A[t * ldn + perm_dct(i, N)] = W[(int64_t)i * a.sa + t];
    }
    __syncthreads();
    double2* F = run_fft<false>(A, B, a.plan, ldn, Tv, a.tw, tid, nth);
    double2* G = (F == A) ? B : A;
    // 2. DCT-II post-processing and scale pi/(2N)
    const double sc = 3.14159265358979323846 / (2.0 * N);
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e / N, k = e - t * N;
        const double2 P = F[t * ldn + k];
        const double2 Q = cconj(F[t * ldn + (k == 0 ? 0 : N - k)]);
        const double2 w = a.tq[k];
        const double2 s1 = cmul(w, cadd(P, Q));
        const double2 s2 = cmul(w, csub(P, Q));
        G[t * ldn + k] = make_double2(s1.x * sc, s2.y * sc);
    }
    __syncthreads();
    // 3. composite stencil -> F
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e / N, k = e - t * N;
        double2 v = G[t * ldn + k];
        if (k < N - 2) {
            const double kk = a.K2[k];
            const double2 u = G[t * ldn + k + 2];
            v.x += kk * u.x;
            v.y += kk * u.y;
        }
        F[t * ldn + k] = v;
    }
    __syncthreads();
    // 4. Helmholtz LU on the fly + solve, one thread per (line, parity)
    if (tid < 2 * Tv) {
        const int t = tid >> 1, p = tid & 1;
        const int64_t jz = jz0 + t;
        const double ky = (jy < a.Ny / 2 || (a.Ny % 2 == 1 && jy == a.Ny / 2)) ? (double)jy
                                                                             : (double)(jy - a.Ny);
        const double kz = (double)jz;
        const double be = a.alpha + ky * ky + kz * kz;
        const double al = -1.0;
        const bool nm = a.neumann != 0;
        const bool zb = nm && fabs(be) < 1e-8;
        const double* tA0 = a.tabs;
        const double* tA2 = tA0 + M;
        const double* tA4 = tA2 + M;
        const double* tBm2 = tA4 + M;
        const double* tB0 = tBm2 + M;
        const double* tB2 = tB0 + M;
        double2* X = F + t * ldn;  // rhs in, y/d0 then u out
        double2* E = G + t * ldn;  // (d1/d0, d2/d0) per row
        double d0p = 0, d1p = 0, d2p = 0;
        double2 yp = make_double2(0.0, 0.0);
        for (int i = p; i < M; i += 2) {
            double a0 = tA0[i], b0 = tB0[i];
            double a2 = (i < M - 2) ? tA2[i] : 0.0, b2 = (i < M - 2) ? tB2[i] : 0.0;
            double a4 = (i < M - 4) ? tA4[i] : 0.0;
            if (nm) {
                if (i == 0) {
                    if (zb) { a0 = 1.0 / al; b0 = 0.0; }
                } else {
                    a0 /= (double)i * i;
                    b0 /= (double)i * i;
                }
                a2 /= (double)(i + 2) * (i + 2);
                b2 /= (double)(i + 2) * (i + 2);
                a4 /= (double)(i + 4) * (i + 4);
            }
            double d0v, d1v, d2v;
            double2 yv = X[i];
            if (i < 2) {
                d0v = al * a0 + be * b0;
                d1v = al * a2 + be * b2;
                d2v = al * a4;
            } else {
                double bm = tBm2[i - 2];
                if (nm && i - 2 >= 1) bm /= (double)(i - 2) * (i - 2);
                const double l = be * bm / d0p;
                d0v = al * a0 + be * b0 - l * d1p;
                d1v = (i < M - 2) ? al * a2 + be * b2 - l * d2p : 0.0;
                d2v = (i < M - 4) ? al * a4 - l * d2p : 0.0;
                yv.x -= l * yp.x;
                yv.y -= l * yp.y;
            }
            const double inv = 1.0 / d0v;
            X[i] = make_double2(yv.x * inv, yv.y * inv);
            E[i] = make_double2(d1v * inv, d2v * inv);
            yp = yv;
            d0p = d0v;
            d1p = d1v;
            d2p = d2v;
        }
        const int il = (M - 1) - (((M - 1) - p) & 1);
        double2 n2 = X[il];
        double2 n4 = make_double2(0.0, 0.0);
        double2 S = make_double2(0.0, 0.0);
        if (il - 2 >= 0) {
            const int i = il - 2;
            const double2 e = E[i];
            double2 cur = X[i];
            cur.x -= e.x * n2.x;
            cur.y -= e.x * n2.y;
            X[i] = cur;
            n4 = n2;
            n2 = cur;
        }
        for (int i = il - 4; i >= 0; i -= 2) {
            const double2 e = E[i];
            S.x += n4.x;
            S.y += n4.y;
            double2 cur = X[i];
            cur.x -= e.x * n2.x + e.y * S.x;
            cur.y -= e.x * n2.y + e.y * S.y;
            X[i] = cur;
            n4 = n2;
            n2 = cur;
        }
        if (nm) {
            // gauge of la.pyx:907-909 uses the first LU diagonal d0[0]
            double a00 = tA0[0], b00 = tB0[0];
            if (zb) { a00 = 1.0 / al; b00 = 0.0; }
            const double d00 = al * a00 + be * b00;
            if (p == 0 && (d00 - 1.0) * (d00 - 1.0) < 1e-16) X[0] = make_double2(0.0, 0.0);
            for (int i = (p == 0 ? 2 : 1); i < M; i += 2) {
                const double s = 1.0 / ((double)i * i);
                X[i] = make_double2(X[i].x * s, X[i].y * s);
            }
        }
        if (p == 0) {  // BC slots of a fresh solution are zero
            X[M] = make_double2(0.0, 0.0);
            X[M + 1] = make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    // 5. to_ortho -> G, keep t_0
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e / N, k = e - t * N;
        double2 v = (k < N - 2) ? F[t * ldn + k] : make_double2(0.0, 0.0);
        if (k >= 2) {
            const double kk = a.K2[k - 2];
            const double2 u = F[t * ldn + k - 2];
            v.x += kk * u.x;
            v.y += kk * u.y;
        }
        G[t * ldn + k] = v;
        if (k == 0) {
            t0s[2 * t] = v.x;
            t0s[2 * t + 1] = v.y;
        }
    }
    __syncthreads();
    // 6. DCT-III pre-processing -> F
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e / N, k = e - t * N;
        const double2 ck = G[t * ldn + k];
        const double2 cn = (k == 0) ? make_double2(0.0, 0.0) : G[t * ldn + N - k];
        const double2 z = make_double2(ck.x + cn.y, ck.y - cn.x);
        F[t * ldn + k] = cmul(cconj(a.tq[k]), z);
    }
    __syncthreads();
    double2* H = run_fft<true>(F, G, a.plan, ldn, Tv, a.tw, tid, nth);
    // 7. store 0.5 y + t_0/2 with the inverse permutation
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e % Tv, i = e / Tv;
        const double2 y = H[t * ldn + perm_dct(i, N)];
        W[(int64_t)i * a.sa + t] = make_double2(0.5 * y.x + 0.5 * t0s[2 * t], 0.5 * y.y + 0.5 * t0s[2 * t + 1]);
    }
}

// K2 stencil of the Chebyshev SD/SN bases (chebyshev/bases.py:546, 937)
static std::map<std::pair<int, int>, double*> g_k2;
static const double* cheb_k2(int N, int neumann)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_helm_mu);
    auto key = std::make_pair(N * 2 + neumann, dev);
    auto it = g_k2.find(key);
    if (it != g_k2.end()) return it->second;
    std::vector<double> h(N - 2);
    for (int k = 0; k < N - 2; ++k) {
        const double kk = k;
        h[k] = neumann ? -(kk / (kk + 2)) * (kk / (kk + 2)) : -1.0;
    }
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    g_k2[key] = d;
    return d;
}

int xpass_launch(double2* W, int Nx, int Ny, int Nzh, double alpha, int neumann, hipStream_t s)
{
    XPassArgs a;
    a.W = W;
    a.sa = (int64_t)Ny * Nzh;
    a.s1 = Nzh;
    a.d1 = Ny;
    a.d2 = Nzh;
    a.N = Nx;
    a.ldn = Nx + 2;
    int T = (int)(65536 / (2 * 16 * (int64_t)a.ldn));
    if (T < 1) T = 1;
    if (T > 16) T = 16;
    if (T > Nzh) T = Nzh;
    a.T = T;
    a.neumann = neumann;
    a.alpha = alpha;
    a.Ny = Ny;
    a.K2 = cheb_k2(Nx, neumann);
    a.tabs = helm_tables(Nx, neumann);
    a.tw = fft_table(Nx);
    a.tq = quarter_table(Nx);
    SFA_CHECK(a.K2 && a.tabs && a.tw && a.tq, SFA_ERR_HIP, "table allocation failed");
    a.plan = make_fft_plan(Nx);
    const size_t lds = (size_t)2 * T * a.ldn * sizeof(double2) + (size_t)2 * T * sizeof(double);
    SFA_CHECK(lds <= 160 * 1024, SFA_ERR_UNSUPPORTED, "Nx=%d too large for the fused x-pass", Nx);
    if (lds > 64 * 1024)
        hipFuncSetAttribute((const void*)xpass_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
    const int64_t nblk = (int64_t)Ny * ((Nzh + T - 1) / T);
    hipLaunchKernelGGL(xpass_kernel, dim3((unsigned)nblk), dim3(256), lds, s, a);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

}  // namespace sfa

using namespace sfa;

extern "C" int64_t sfa_helmholtz3d_workspace(int64_t Nx, int64_t Ny, int64_t Nz, int precision)
{
    (void)precision;
    return Nx * Ny * (Nz / 2 + 1) * (int64_t)sizeof(double2);
}

extern "C" int sfa_helmholtz3d(const double* f, double* u, int64_t Nx, int64_t Ny, int64_t Nz,
                               double alpha, int neumann, int precision, int refine, void* work,
                               int64_t work_bytes, void* stream)
{
    (void)refine;
    SFA_CHECK(Nx >= 8 && Ny >= 1 && Nz >= 2 && Nz % 2 == 0, SFA_ERR_ARG,
              "need Nx >= 8, Ny >= 1, even Nz >= 2");
    SFA_CHECK(precision == 0, SFA_ERR_UNSUPPORTED, "precision %d not built yet", precision);
    SFA_CHECK(work_bytes >= sfa_helmholtz3d_workspace(Nx, Ny, Nz, precision), SFA_ERR_ARG,
              "workspace too small");
    const int64_t Nzh = Nz / 2 + 1;
    double2* W = (double2*)work;
    const int64_t rshape[3] = {Nx, Ny, Nz};
    const int64_t rstr[3] = {Ny * Nz, Nz, 1};
    const int64_t cstr[3] = {Ny * Nzh, Nzh, 1};
    const int64_t cshape[3] = {Nx, Ny, Nzh};
    int rc;
    if ((rc = sfa_fft_r2c(f, W, 3, rshape, rstr, cstr, 2, 1.0 / (double)Nz, stream))) return rc;
    if ((rc = sfa_fft_c2c(W, W, 3, cshape, cstr, cstr, 1, -1, 1.0 / (double)Ny, stream))) return rc;
    if ((rc = xpass_launch(W, (int)Nx, (int)Ny, (int)Nzh, alpha, neumann, (hipStream_t)stream)))
        return rc;
    if ((rc = sfa_fft_c2c(W, W, 3, cshape, cstr, cstr, 1, +1, 1.0, stream))) return rc;
    if ((rc = sfa_fft_c2r(W, u, 3, rshape, cstr, rstr, 2, 1.0, stream))) return rc;
    return SFA_OK;
}
