// Batched 2D Helmholtz solves (alpha - lap) u = f on Legendre(SD or SN) x
// Fourier(R2C) in one persistent launch.
//
// Per problem (Nx x Ny physical, M = Nx - 2 composite dofs, Nyh = Ny/2 + 1):
//   T1 = P f                P = Phi^T diag(w): composite Legendre scalar product
//                           (spectralbase.py:550-558 with the stencil of
//                           legendre/bases.py:374 / :502, spectralbase.py:1462-1474)
//   T2 = T1 Df              R2C along y with 1/Ny (fourier/bases.py:157-162),
//                           stored as complex (re, im) pairs
//   per ky: (A + (alpha + ky^2) B) u_hat = T2[:, ky]   (la.py:1346-1364: the
//                           per-mode TDMA of SolverGeneric1ND; A, B closed forms of
//                           legendre/matrices.py:110-162, 271-315)
//   U1 = Phi u_hat          composite synthesis (to_ortho + recursive backward,
//                           spectralbase.py:1616-1634, :484-491)
//   u = U1 Di               C2R along y (imag of k = 0 and Nyquist ignored)
// Phi[j, m] = L_m(x_j) + K2_m L_{m+2}(x_j) on the Legendre-Gauss nodes; Df, Di
// are the real DFT matrices.  All four contractions are FP64 GEMMs on MFMA
// (v_mfma_f64_16x16x4f64).  One workgroup owns a problem from start to end; the
// host assigns problems to workgroups largest first (static LPT schedule), so
// the launch needs no atomics and every workgroup exits after its list.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>
#include <queue>
#include <vector>

#include "common.h"
#include "mfma_gemm.h"

namespace sfa {

// ----------------------------------------------------------------- problem table
struct B2Prob {
    int64_t f_off, u_off;     // element offsets of f and u (row-major Nx x Ny float64)
    int64_t phi_off, p_off;   // Phi (Nx x M) and P (M x Nx) in the matrix arena (doubles)
    int64_t df_off, di_off;   // Df (Ny x (Ny+2)) and Di ((Ny+2) x Ny)
    int32_t Nx, Ny, bc, pad;  // bc: 0 Dirichlet (SD), 1 Neumann (SN)
    double alpha;
};

// ----------------------------------------------------------------- MFMA GEMM
constexpr int B2_THREADS = GEMM_THREADS;

// C (m x n, ldc) = A (m x k, lda) B (k x n, ldb), all row-major FP64, by the
// whole workgroup (4 waves in a 2 x 2 grid over BM x BN block tiles, K in
// steps of BK).  sA: BK x (BM + 1) doubles, sB: BK x BN doubles of LDS.
// Out-of-range operand elements load as zero; stores are bounds-checked.
// Each output is one MFMA accumulation chain over K in ascending steps of 4
// (up to the last step holding k), so the result does not depend on BM, BN,
// BK, on which workgroup computes the tile, or on the order of the tiles.
template <int BM, int BN, int BK>
__device__ void block_gemm(const double* __restrict__ A, int64_t lda, const double* __restrict__ B,
                           int64_t ldb, double* __restrict__ C, int64_t ldc, int m, int n, int k,
                           double* sA, double* sB)
{
    for (int m0 = 0; m0 < m; m0 += BM)
        for (int n0 = 0; n0 < n; n0 += BN) gemm_tile<BM, BN, BK>(A, lda, B, ldb, C, ldc, m, n, k, m0, n0, sA, sB);
    __syncthreads();  // C complete and LDS free before the caller's next stage
}

// ----------------------------------------------------------------- per-ky solve
// Legendre composite Galerkin matrices (legendre/matrices.py:110-162, 271-315),
// h_k = 2/(2k+1):  SD  A_kk = 4k+6, B_kk = h_k + h_{k+2}, B_{k,k+2} = -h_{k+2};
// SN  a_k = k(k+1)/((k+2)(k+3)), A_kk = h_k a_k (k+1/2)((k+2)(k+3) - k(k+1)),
//     B_kk = h_k + a_k^2 h_{k+2}, B_{k,k+2} = -a_k h_{k+2}.  B is symmetric.
__device__ inline void leg_rows(int i, int bc, double& Akk, double& Bkk, double& Bup)
{
    const double k = (double)i;
    const double h0 = 2.0 / (2.0 * k + 1.0), h2 = 2.0 / (2.0 * k + 5.0);
    if (bc == 0) {
        Akk = 4.0 * k + 6.0;
        Bkk = h0 + h2;
        Bup = -h2;
    } else {
        const double a = k * (k + 1.0) / ((k + 2.0) * (k + 3.0));
        Akk = h0 * a * (k + 0.5) * ((k + 2.0) * (k + 3.0) - k * (k + 1.0));
        Bkk = h0 + a * a * h2;
        Bup = -a * h2;
    }
}

// Thomas algorithm per (ky, parity chain) on the complex columns of W0 (M rows,
// row pitch ldw doubles, column ky at 2 ky).  cp: scratch, M * Nyh doubles,
// row-major (row i, column ky), so a wave's stores of one step are contiguous
// (ky-major, a chain's own column, they were 32 separate lines per wave: the
// solve launch of the tile-parallel form took 1.21 ms instead of 0.83 ms for
// the 4096-problem batch).  Staging row chunks in LDS between barriers measured
// slower still (1.43 ms).  Neumann with beta = 0 (ky = 0, alpha = 0): the singular row 0
// is replaced by u_0 = 0 (the constraint (0, 0) of the reference's mode-0
// solver).
// One row of the forward elimination (the reference's TDMA order), x = (re, im)
// of row i in, its reduced values out.
__device__ __forceinline__ void thomas_row(int i, int M, int bc, double beta, double2& x, double& cc,
                                           double& cprev, double& rprev_re, double& rprev_im,
                                           double& lprev_up)
{
    double Akk, Bkk, Bup;
    leg_rows(i, bc, Akk, Bkk, Bup);
    const double d = Akk + beta * Bkk;
    const double up = (i + 2 < M) ? beta * Bup : 0.0;
    const double lo = (i >= 2) ? lprev_up : 0.0;  // B symmetric: B_{i,i-2} = B_{i-2,i}
    double rr_re, rr_im;
    if (bc == 1 && i == 0 && fabs(beta) < 1e-12) {
        cc = 0.0;
        rr_re = rr_im = 0.0;
    } else {
        const double dd = 1.0 / (d - lo * cprev);
        cc = up * dd;
        rr_re = (x.x - lo * rprev_re) * dd;
        rr_im = (x.y - lo * rprev_im) * dd;
    }
    x = make_double2(rr_re, rr_im);
    cprev = cc;
    rprev_re = rr_re;
    rprev_im = rr_im;
    lprev_up = up;
}

__device__ __forceinline__ void batch_solve(double* __restrict__ W0, int64_t ldw, int M, int Nyh, int bc,
                                            double alpha, double* __restrict__ cp)
{
    // A chain's row loads do not depend on its recurrence: the sweeps take the
    // rows D at a time from one of two register buffers while the next D rows
    // load into the other, so a step does not wait out a load of its own
    // (clamped row indices keep the loads unconditional).
    constexpr int D = 8;
    for (int c = threadIdx.x; c < 2 * Nyh; c += blockDim.x) {
        const int ky = c >> 1, p = c & 1;
        const double beta = alpha + (double)ky * (double)ky;
        double* cpc = cp + ky;  // cpc[i * Nyh]
        double2* xc = reinterpret_cast<double2*>(W0 + 2 * ky);  // row i at xc[i * ldw / 2]
        const int64_t lx = ldw / 2;
        double cprev = 0.0, rprev_re = 0.0, rprev_im = 0.0, lprev_up = 0.0;
        int last = -1;
        double2 bA[D], bB[D];
        auto fwd = [&](int i0, double2 (&cur)[D], double2 (&nxt)[D]) {
#pragma unroll
            for (int d = 0; d < D; ++d) nxt[d] = xc[min(i0 + 2 * (D + d), M - 1) * lx];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int i = i0 + 2 * d;
                if (i < M) {
                    double cc;
                    thomas_row(i, M, bc, beta, cur[d], cc, cprev, rprev_re, rprev_im, lprev_up);
                    cpc[(int64_t)i * Nyh] = cc;
                    xc[i * lx] = cur[d];
                    last = i;
                }
            }
        };
#pragma unroll
        for (int d = 0; d < D; ++d) bA[d] = xc[min(p + 2 * d, M - 1) * lx];
        for (int i0 = p; i0 < M; i0 += 4 * D) {
            fwd(i0, bA, bB);
            if (i0 + 2 * D < M) fwd(i0 + 2 * D, bB, bA);
        }
        // back substitution over the same rows, last to first
        double xr = 0.0, xi = 0.0;
        double2 cA[D], cB[D];
        double kA[D], kB[D];
        auto bwd = [&](int i0, double2 (&cx)[D], double (&ck)[D], double2 (&nx)[D], double (&nk)[D]) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int i = max(i0 - 2 * (D + d), 0);
                nx[d] = xc[i * lx];
                nk[d] = cpc[(int64_t)i * Nyh];
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int i = i0 - 2 * d;
                if (i >= 0) {
                    xr = cx[d].x - ck[d] * xr;
                    xi = cx[d].y - ck[d] * xi;
                    xc[i * lx] = make_double2(xr, xi);
                }
            }
        };
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = max(last - 2 * d, 0);
            cA[d] = xc[i * lx];
            kA[d] = cpc[(int64_t)i * Nyh];
        }
        for (int i0 = last; i0 >= 0; i0 -= 4 * D) {
            bwd(i0, cA, kA, cB, kB);
            if (i0 - 2 * D >= 0) bwd(i0 - 2 * D, cB, kB, cA, kA);
        }
    }
    __syncthreads();
}

// ----------------------------------------------------------------- persistent kernel
struct B2Args {
    const double* F;
    double* U;
    const B2Prob* probs;
    const int* blk_ptr;  // nblocks + 1
    const int* order;    // problem ids, grouped by workgroup
    const double* mats;  // matrix arena
    double* work;        // per workgroup: wstride doubles
    int64_t wstride;
};

template <int BM, int BN, int BK>
__global__ void __launch_bounds__(B2_THREADS) batch2d_kernel(B2Args a)
{
    __shared__ double sA[BK * (BM + 1)];
    __shared__ double sB[BK * BN];
    const int b = blockIdx.x;
    double* W0 = a.work + (int64_t)b * a.wstride;
    for (int q = a.blk_ptr[b]; q < a.blk_ptr[b + 1]; ++q) {
        const B2Prob pr = a.probs[a.order[q]];
        const int Nx = pr.Nx, Ny = pr.Ny, M = Nx - 2, Nyh = Ny / 2 + 1, L = Ny + 2;
        const double* f = a.F + pr.f_off;
        double* u = a.U + pr.u_off;
        double* W1 = W0 + (int64_t)M * L;  // Nx x L
        const double* Phi = a.mats + pr.phi_off;
        const double* P = a.mats + pr.p_off;
        const double* Df = a.mats + pr.df_off;
        const double* Di = a.mats + pr.di_off;
        // The four contractions through one GEMM call site (a loop over the
        // stages, not unrolled): with four inlined copies the compiler held the
        // address setup of all of them, 163 VGPRs against 67 for one tile.
#pragma unroll 1
        for (int st = 0; st < 4; ++st) {
            const double *A, *B;
            double* C;
            int64_t lda, ldb, ldc;
            int m, n, k;
            if (st == 0) {  // T1 = P f (M x Ny), held in u
                A = P, lda = Nx, B = f, ldb = Ny, C = u, ldc = Ny, m = M, n = Ny, k = Nx;
            } else if (st == 1) {  // T2 = T1 Df (M x L) in W0
                A = u, lda = Ny, B = Df, ldb = L, C = W0, ldc = L, m = M, n = L, k = Ny;
            } else if (st == 2) {  // U1 = Phi u_hat (Nx x L) in W1
                // per-ky solve in place first, W1 as the Thomas scratch
                batch_solve(W0, L, M, Nyh, pr.bc, pr.alpha, W1);
                A = Phi, lda = M, B = W0, ldb = L, C = W1, ldc = L, m = Nx, n = L, k = M;
            } else {  // u = U1 Di (Nx x Ny)
                A = W1, lda = L, B = Di, ldb = Ny, C = u, ldc = Ny, m = Nx, n = Ny, k = L;
            }
            block_gemm<BM, BN, BK>(A, lda, B, ldb, C, ldc, m, n, k, sA, sB);
        }
    }
}

// ----------------------------------------------------------------- tile-parallel mode
// For batches with fewer problems than CUs: each GEMM stage runs as one
// launch over all (problem, output tile) tasks, the solve as one launch with a
// workgroup per problem, so a single large problem spreads over the device.
// Per-problem intermediates live at work + woff[p] (W0: (Nx-2) x (Ny+2),
// then W1: Nx x (Ny+2)); T1 is held in u as in the persistent kernel.
struct B2Task {
    int32_t prob, m0, n0, tile;  // tile: 0 = 32 x 32, 1 = 64 x 64 block tile
};

struct B2WideArgs {
    const double* F;
    double* U;
    const B2Prob* probs;
    const B2Task* tasks;   // this stage's tasks
    const int64_t* woff;   // per problem, in doubles
    const int* order;      // the solve launch's problem per workgroup
    const double* mats;
    double* work;
};

constexpr int B2W_BK = 16;
// block tiles of the tile classes (a 128 x 64 class spills in this kernel)
constexpr int kB2WClasses = 2;
constexpr int kB2WTile[kB2WClasses][2] = {{32, 32}, {64, 64}};

template <int BM, int BN>
__device__ void wide_gemm(const B2WideArgs& a, const B2Task& tk, int stage, double* sA, double* sB)
{
    const B2Prob pr = a.probs[tk.prob];
    const int Nx = pr.Nx, Ny = pr.Ny, M = Nx - 2, L = Ny + 2;
    double* u = a.U + pr.u_off;
    double* W0 = a.work + a.woff[tk.prob];
    double* W1 = W0 + (int64_t)M * L;
    if (stage == 0)  // T1 = P f (M x Ny) in u
        gemm_tile<BM, BN, B2W_BK>(a.mats + pr.p_off, Nx, a.F + pr.f_off, Ny, u, Ny, M, Ny, Nx, tk.m0, tk.n0,
                                  sA, sB);
    else if (stage == 1)  // T2 = T1 Df (M x L) in W0
        gemm_tile<BM, BN, B2W_BK>(u, Ny, a.mats + pr.df_off, L, W0, L, M, L, Ny, tk.m0, tk.n0, sA, sB);
    else if (stage == 2)  // U1 = Phi u_hat (Nx x L) in W1
        gemm_tile<BM, BN, B2W_BK>(a.mats + pr.phi_off, M, W0, L, W1, L, Nx, L, M, tk.m0, tk.n0, sA, sB);
    else  // u = U1 Di (Nx x Ny)
        gemm_tile<BM, BN, B2W_BK>(W1, L, a.mats + pr.di_off, Ny, u, Ny, Nx, Ny, L, tk.m0, tk.n0, sA, sB);
}

// One task per workgroup, all of a launch's tasks of tile class CLS; every
// output keeps its one MFMA chain over K, so the results do not depend on the
// class.
template <int STAGE, int CLS>
__global__ void __launch_bounds__(B2_THREADS) batch2d_tile_kernel(B2WideArgs a)
{
    constexpr int BM = kB2WTile[CLS][0], BN = kB2WTile[CLS][1];
    __shared__ double sA[B2W_BK * (BM + 1)];
    __shared__ double sB[B2W_BK * BN];
    wide_gemm<BM, BN>(a, a.tasks[blockIdx.x], STAGE, sA, sB);
}

template <int STAGE>
static int launch_wide_stage(const B2WideArgs& a, int cls, int64_t ntask, hipStream_t s)
{
    const dim3 blk(B2_THREADS), grid((unsigned)ntask);
    if (cls == 1)
        hipLaunchKernelGGL((batch2d_tile_kernel<STAGE, 1>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((batch2d_tile_kernel<STAGE, 0>), grid, blk, 0, s, a);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

__global__ void __launch_bounds__(B2_THREADS) batch2d_wide_solve_kernel(B2WideArgs a)
{
    const int p = a.order[blockIdx.x];
    const B2Prob pr = a.probs[p];
    const int Nx = pr.Nx, Ny = pr.Ny, M = Nx - 2, L = Ny + 2;
    double* W0 = a.work + a.woff[p];
    batch_solve(W0, L, M, Ny / 2 + 1, pr.bc, pr.alpha, W0 + (int64_t)M * L);
}

// ----------------------------------------------------------------- matrix builders
// Phi (Nx x M) and P (M x Nx) for Legendre-Gauss nodes x, weights w: one thread
// per node, the recurrence L_{k+1} = ((2k+1) x L_k - k L_{k-1})/(k+1) in the
// anm/anp form of transforms.pyx:219-323.
__global__ void b2_phi_kernel(const double* __restrict__ x, const double* __restrict__ w, int Nx,
                              int bc, double* __restrict__ Phi, double* __restrict__ P)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= Nx) return;
    const int M = Nx - 2;
    const double xj = x[j], wj = w[j];
    double Lm = 1.0, L0 = xj;
    double L1 = 1.5 * xj * L0 - 0.5 * Lm;  // L_2
    for (int m = 0; m < M; ++m) {
        // Lm = L_m, L0 = L_{m+1}, L1 = L_{m+2}
        const double km = (double)m;
        const double K2 = (bc == 0) ? -1.0 : -km * (km + 1.0) / ((km + 2.0) * (km + 3.0));
        const double v = Lm + K2 * L1;
        Phi[(int64_t)j * M + m] = v;
        P[(int64_t)m * Nx + j] = wj * v;
        const int n = m + 2;  // L_{n+1} = (1/anm_n) x L_n - (anp_n/anm_n) L_{n-1}
        const double anm = (n + 1.0) / (2.0 * n + 1.0), anp = (double)n / (2.0 * n + 1.0);
        const double Ln = (1.0 / anm) * xj * L1 - (anp / anm) * L0;
        Lm = L0;
        L0 = L1;
        L1 = Ln;
    }
}

// Df (Ny x (Ny+2)): [y][2k] = cos(2 pi k y/Ny)/Ny, [y][2k+1] = -sin(..)/Ny;
// Di ((Ny+2) x Ny): [2k][y] = c_k cos, [2k+1][y] = -c_k sin, c_k = 1 at k = 0 and
// Ny/2, else 2.  Angles reduced exactly ((k y) mod Ny) before sincospi.
__global__ void b2_dft_kernel(int Ny, double* __restrict__ Df, double* __restrict__ Di)
{
    const int Nyh = Ny / 2 + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Ny * Nyh) return;
    const int y = (int)(t / Nyh), k = (int)(t % Nyh);
    const int r = (int)(((int64_t)k * y) % Ny);
    double s, c;
    sincospi(2.0 * (double)r / (double)Ny, &s, &c);
    const int L = Ny + 2;
    Df[(int64_t)y * L + 2 * k] = c / Ny;
    Df[(int64_t)y * L + 2 * k + 1] = -s / Ny;
    const double ck = (k == 0 || 2 * k == Ny) ? 1.0 : 2.0;
    Di[(int64_t)(2 * k) * Ny + y] = ck * c;
    Di[(int64_t)(2 * k + 1) * Ny + y] = (k == 0 || 2 * k == Ny) ? 0.0 : -ck * s;
}

// MFMA lane-layout probe: A (16 x 4) = i + 100 k, B (4 x 16) = j + 1000 k;
// out[l * 4 + r] = D register r of lane l.
__global__ void mfma_probe_kernel(double* out)
{
    const int l = threadIdx.x;
    const double a = (double)(l % 16) + 100.0 * (l / 16);
    const double b = (double)(l % 16) + 1000.0 * (l / 16);
    v4d d = mfma16(a, b, (v4d){0.0, 0.0, 0.0, 0.0});
    for (int r = 0; r < 4; ++r) out[l * 4 + r] = d[r];
}

// ----------------------------------------------------------------- host side
// Legendre-Gauss nodes (ascending) and weights: Newton on the three-term
// recurrence in long double from the cosine guess, weights 2/((1-x^2) P'^2),
// symmetrised; cached per N for the process.
static std::mutex g_lg_mu;
static std::map<int, std::pair<std::vector<double>, std::vector<double>>> g_lg;

static void lg_nodes(int N, std::vector<double>& xo, std::vector<double>& wo)
{
    {
        std::lock_guard<std::mutex> lk(g_lg_mu);
        auto it = g_lg.find(N);
        if (it != g_lg.end()) {
            xo = it->second.first;
            wo = it->second.second;
            return;
        }
    }
    typedef long double ld;
    const ld pi = 3.141592653589793238462643383279502884L;
    std::vector<ld> x(N), w(N);
    auto pdp = [&](ld xv, ld& p, ld& dp) {
        ld p0 = 1.0L, p1 = xv;
        for (int k = 1; k < N; ++k) {
            const ld p2 = ((2 * k + 1) * xv * p1 - k * p0) / (k + 1);
            p0 = p1;
            p1 = p2;
        }
        p = p1;
        dp = N * (xv * p1 - p0) / (xv * xv - 1.0L);
    };
    for (int i = 0; i < N; ++i) {
        ld xv = std::cos(pi * ((ld)(i + 1) - 0.25L) / ((ld)N + 0.5L));
        for (int it = 0; it < 100; ++it) {
            ld p, dp;
            pdp(xv, p, dp);
            const ld dx = p / dp;
            xv -= dx;
            if (std::fabs(dx) < 1e-19L) break;
        }
        ld p, dp;
        pdp(xv, p, dp);
        x[i] = xv;
        w[i] = 2.0L / ((1.0L - xv * xv) * dp * dp);
    }
    // ascending, rounded, symmetrised
    std::vector<int> o(N);
    std::iota(o.begin(), o.end(), 0);
    std::sort(o.begin(), o.end(), [&](int a, int b) { return x[a] < x[b]; });
    std::vector<double> xs(N), ws(N);
    for (int i = 0; i < N; ++i) {
        xs[i] = (double)x[o[i]];
        ws[i] = (double)w[o[i]];
    }
    xo.assign(N, 0.0);
    wo.assign(N, 0.0);
    for (int i = 0; i < N; ++i) {
        xo[i] = 0.5 * (xs[i] - xs[N - 1 - i]);
        wo[i] = 0.5 * (ws[i] + ws[N - 1 - i]);
    }
    std::lock_guard<std::mutex> lk(g_lg_mu);
    g_lg[N] = std::make_pair(xo, wo);
}

// Device copies of the nodes and weights per N (x then w), for the process.
static std::map<std::pair<int, int>, double*> g_lg_dev;  // (N, device)

static const double* lg_nodes_dev(int N)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    {
        std::lock_guard<std::mutex> lk(g_lg_mu);
        auto it = g_lg_dev.find(std::make_pair(N, dev));
        if (it != g_lg_dev.end()) return it->second;
    }
    std::vector<double> x, w;
    lg_nodes(N, x, w);
    x.insert(x.end(), w.begin(), w.end());
    double* d = nullptr;
    if (hipMalloc(&d, x.size() * sizeof(double)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_lg_mu);
    g_lg_dev[std::make_pair(N, dev)] = d;
    return d;
}

static int64_t leg_elems(int Nx) { return 2 * (int64_t)Nx * (Nx - 2); }
static int64_t dft_elems(int Ny) { return 2 * (int64_t)Ny * (Ny + 2); }

static int build_leg(int Nx, int bc, double* out, hipStream_t s)
{
    const double* xw = lg_nodes_dev(Nx);
    SFA_CHECK(xw, SFA_ERR_HIP, "node table allocation failed");
    hipLaunchKernelGGL(b2_phi_kernel, dim3((unsigned)((Nx + 127) / 128)), dim3(128), 0, s, xw, xw + Nx, Nx,
                       bc, out, out + (int64_t)Nx * (Nx - 2));
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

static int build_dft(int Ny, double* out, hipStream_t s)
{
    const int64_t nt = (int64_t)Ny * (Ny / 2 + 1);
    hipLaunchKernelGGL(b2_dft_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, Ny, out,
                       out + (int64_t)Ny * (Ny + 2));
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

static int64_t r256(int64_t b) { return (b + 255) / 256 * 256; }

// plan layout: problem table | blk_ptr | order [| matrix arena (self-contained plans)]
struct B2Layout {
    int64_t off_probs, off_blk, off_order, off_mats, bytes;
};

static void b2_layout(int nprob, int nblocks, int64_t mats_elems, B2Layout* L)
{
    L->off_probs = 0;
    L->off_blk = r256((int64_t)nprob * sizeof(B2Prob));
    L->off_order = L->off_blk + r256((int64_t)(nblocks + 1) * 4);
    L->off_mats = L->off_order + r256((int64_t)nprob * 4);
    L->bytes = L->off_mats + r256(mats_elems * 8);
}

// distinct matrices of a batch: (Nx, bc) -> Phi offset, Ny -> Df offset
static int64_t b2_arena(int nprob, const int* Nx, const int* Ny, const int* bc,
                        std::map<std::pair<int, int>, int64_t>* phi, std::map<int, int64_t>* dft)
{
    int64_t mats = 0;
    for (int i = 0; i < nprob; ++i) {
        const auto key = std::make_pair(Nx[i], bc[i] ? 1 : 0);
        if (!phi->count(key)) {
            (*phi)[key] = mats;
            mats += leg_elems(Nx[i]);
        }
        if (!dft->count(Ny[i])) {
            (*dft)[Ny[i]] = mats;
            mats += dft_elems(Ny[i]);
        }
    }
    return mats;
}

static double b2_cost(int Nx, int Ny)
{
    // flops of the four contractions plus a memory and fixed term (units ~ flops)
    const double x = Nx, y = Ny;
    return 4.0 * x * y * (x + y) + 64.0 * x * y + 2.0e5;
}

struct B2Info {
    int nprob, nblocks;
    int64_t off_probs, off_blk, off_order;
    const double* mats;
    int64_t wstride;  // doubles per workgroup
};
static std::mutex g_b2_mu;
static std::map<const void*, B2Info> g_b2;
// tile-parallel plan: problem table | tasks of the four GEMM stages | woff
struct B2WideInfo {
    int nprob;
    int cls[4];  // tile class of each stage
    int64_t ntask[4], off_task[4];
    int64_t off_probs, off_woff, work_doubles;
    const double* mats;
};
static std::map<const void*, B2WideInfo> g_b2w;

static int64_t b2_wstride(int nprob, const int* Nx, const int* Ny)
{
    int64_t mx = 0;
    for (int i = 0; i < nprob; ++i) {
        const int64_t L = Ny[i] + 2;
        mx = std::max(mx, (int64_t)(Nx[i] - 2) * L + (int64_t)Nx[i] * L);
    }
    return (mx + 31) / 32 * 32;
}

// Problem table and static LPT schedule into the plan; registers the plan.
// Validated problem table.
static int b2_table(int nprob, const int* Nx, const int* Ny, const double* alpha, const int* bc,
                    const int64_t* f_off, const int64_t* u_off, int64_t f_elems, int64_t u_elems,
                    const int64_t* phi_off, const int64_t* df_off, int64_t mats_elems,
                    std::vector<B2Prob>& probs)
{
    probs.resize(nprob);
    for (int i = 0; i < nprob; ++i) {
        const int64_t n = (int64_t)Nx[i] * Ny[i];
        SFA_CHECK(f_off[i] >= 0 && f_off[i] + n <= f_elems, SFA_ERR_ARG,
                  "problem %d: f range [%lld, %lld) outside the %lld-element input", i,
                  (long long)f_off[i], (long long)(f_off[i] + n), (long long)f_elems);
        SFA_CHECK(u_off[i] >= 0 && u_off[i] + n <= u_elems, SFA_ERR_ARG,
                  "problem %d: u range outside the %lld-element output", i, (long long)u_elems);
        SFA_CHECK(phi_off[i] >= 0 && phi_off[i] + leg_elems(Nx[i]) <= mats_elems && df_off[i] >= 0 &&
                      df_off[i] + dft_elems(Ny[i]) <= mats_elems,
                  SFA_ERR_ARG, "problem %d: matrix offsets outside the %lld-element arena", i,
                  (long long)mats_elems);
        B2Prob& p = probs[i];
        p.f_off = f_off[i];
        p.u_off = u_off[i];
        p.phi_off = phi_off[i];
        p.p_off = phi_off[i] + (int64_t)Nx[i] * (Nx[i] - 2);
        p.df_off = df_off[i];
        p.di_off = df_off[i] + (int64_t)Ny[i] * (Ny[i] + 2);
        p.Nx = Nx[i];
        p.Ny = Ny[i];
        p.bc = bc[i] ? 1 : 0;
        p.pad = 0;
        p.alpha = alpha[i];
    }
    return SFA_OK;
}

static int b2_plan_core(int nprob, const int* Nx, const int* Ny, const double* alpha, const int* bc,
                        const int64_t* f_off, const int64_t* u_off, int64_t f_elems, int64_t u_elems,
                        const double* mats, const int64_t* phi_off, const int64_t* df_off,
                        int64_t mats_elems, int nblocks, void* plan, int64_t plan_bytes,
                        hipStream_t s)
{
    B2Layout L;
    b2_layout(nprob, nblocks, 0, &L);
    SFA_CHECK(plan_bytes >= L.bytes, SFA_ERR_ARG, "plan buffer too small (%lld < %lld)",
              (long long)plan_bytes, (long long)L.bytes);
    std::vector<B2Prob> probs;
    int rc = b2_table(nprob, Nx, Ny, alpha, bc, f_off, u_off, f_elems, u_elems, phi_off, df_off,
                      mats_elems, probs);
    if (rc) return rc;
    // static LPT schedule: largest first onto the least loaded workgroup
    std::vector<int> ids(nprob);
    std::iota(ids.begin(), ids.end(), 0);
    std::stable_sort(ids.begin(), ids.end(),
                     [&](int a, int b) { return b2_cost(Nx[a], Ny[a]) > b2_cost(Nx[b], Ny[b]); });
    typedef std::pair<double, int> LB;
    std::priority_queue<LB, std::vector<LB>, std::greater<LB>> heap;
    for (int b = 0; b < nblocks; ++b) heap.push(LB(0.0, b));
    std::vector<std::vector<int>> lists(nblocks);
    for (int id : ids) {
        LB t = heap.top();
        heap.pop();
        lists[t.second].push_back(id);
        t.first += b2_cost(Nx[id], Ny[id]);
        heap.push(t);
    }
    std::vector<int> blk(nblocks + 1, 0), order;
    order.reserve(nprob);
    for (int b = 0; b < nblocks; ++b) {
        blk[b] = (int)order.size();
        for (int id : lists[b]) order.push_back(id);
    }
    blk[nblocks] = (int)order.size();
    char* pb = (char*)plan;
    SFA_HIP(hipMemcpyAsync(pb + L.off_probs, probs.data(), probs.size() * sizeof(B2Prob),
                           hipMemcpyHostToDevice, s));
    SFA_HIP(hipMemcpyAsync(pb + L.off_blk, blk.data(), blk.size() * 4, hipMemcpyHostToDevice, s));
    SFA_HIP(hipMemcpyAsync(pb + L.off_order, order.data(), order.size() * 4, hipMemcpyHostToDevice, s));
    SFA_HIP(hipStreamSynchronize(s));  // host vectors are copied before they go out of scope
    std::lock_guard<std::mutex> lk(g_b2_mu);
    g_b2w.erase(plan);  // the address may have held a tile-parallel plan
    g_b2[plan] = B2Info{nprob, nblocks, L.off_probs, L.off_blk, L.off_order, mats,
                        b2_wstride(nprob, Nx, Ny)};
    return SFA_OK;
}


// Tile class of each stage: the largest whose task count still gives every
// one of the kWideCus compute units a few workgroups (more operand reuse per
// tile where the batch is large, more tiles where a few problems must fill the
// device); ntask[st] is that stage's task count.
constexpr int kWideCus = 256;  // compute units of an MI355X

static void b2_wide_shape(int nprob, const int* Nx, const int* Ny, int64_t ntask[4], int64_t* work,
                          int cls[4] = nullptr)
{
    auto cdiv = [](int64_t a, int64_t b) { return (a + b - 1) / b; };
    *work = 0;
    int64_t cnt[4][kB2WClasses] = {};
    for (int i = 0; i < nprob; ++i) {
        const int64_t M = Nx[i] - 2, L = Ny[i] + 2, X = Nx[i], Y = Ny[i];
        const int64_t mm[4] = {M, M, X, X}, nn[4] = {Y, L, L, Y};
        for (int st = 0; st < 4; ++st)
            for (int c = 0; c < kB2WClasses; ++c)
                cnt[st][c] += cdiv(mm[st], kB2WTile[c][0]) * cdiv(nn[st], kB2WTile[c][1]);
        *work += (M * L + X * L + 31) / 32 * 32;
    }
    for (int st = 0; st < 4; ++st) {
        int c = kB2WClasses - 1;
        while (c > 0 && cnt[st][c] < 4 * kWideCus) --c;
        ntask[st] = cnt[st][c];
        if (cls) cls[st] = c;
    }
}

// plan layout: problem table | tasks of each stage | per-problem workspace
// offsets | solve order (problems by decreasing rows, the longest chains first)
static int64_t b2_wide_bytes(int nprob, const int64_t ntask[4], int64_t off_task[4], int64_t* off_woff)
{
    int64_t o = r256((int64_t)nprob * sizeof(B2Prob));
    for (int st = 0; st < 4; ++st) {
        off_task[st] = o;
        o += r256(ntask[st] * (int64_t)sizeof(B2Task));
    }
    *off_woff = o;
    o += r256((int64_t)nprob * 8);
    return o + r256((int64_t)nprob * 4);  // the solve order after the offsets
}

template <int BM, int BN, int BK>
static int launch_b2(const B2Args& a, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL((batch2d_kernel<BM, BN, BK>), dim3((unsigned)nblocks), dim3(B2_THREADS), 0, s, a);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

}  // namespace sfa

using namespace sfa;

// block tiles (BM, BN, BK) of the launch variants
// (64, 128, 32) and (128, 64, 32) need more than 256 VGPRs + AGPRs and spill: left out
static const int kB2Variants[][3] = {{64, 64, 16}, {32, 64, 16}, {64, 128, 16}, {128, 64, 16},
                                     {64, 64, 32}};

extern "C" int sfa_batch2d_variants(int* bm, int* bn, int cap)
{
    const int n = (int)(sizeof(kB2Variants) / sizeof(kB2Variants[0]));
    for (int i = 0; i < n && i < cap; ++i) {
        if (bm) bm[i] = kB2Variants[i][0];
        if (bn) bn[i] = kB2Variants[i][1] * 1000 + kB2Variants[i][2];
    }
    return n;
}

static int b2_check(int nprob, const int* Nx, const int* Ny, const int* bc, int nblocks)
{
    SFA_CHECK(nprob >= 1 && Nx && Ny && bc, SFA_ERR_ARG, "need nprob >= 1 and size arrays");
    SFA_CHECK(nblocks >= 1 && nblocks <= 65536, SFA_ERR_ARG, "nblocks must be in [1, 65536]");
    for (int i = 0; i < nprob; ++i) {
        SFA_CHECK(Nx[i] >= 8 && Nx[i] <= 1024 && Nx[i] % 2 == 0, SFA_ERR_ARG,
                  "problem %d: Nx=%d must be even in [8, 1024]", i, Nx[i]);
        SFA_CHECK(Ny[i] >= 4 && Ny[i] <= 1024 && Ny[i] % 2 == 0, SFA_ERR_ARG,
                  "problem %d: Ny=%d must be even in [4, 1024]", i, Ny[i]);
        SFA_CHECK(bc[i] == 0 || bc[i] == 1, SFA_ERR_ARG, "problem %d: bc must be 0 or 1", i);
    }
    return SFA_OK;
}

extern "C" int64_t sfa_batch2d_plan_bytes(int nprob, const int* Nx, const int* Ny, const int* bc,
                                          int nblocks)
{
    if (b2_check(nprob, Nx, Ny, bc, nblocks)) return -1;
    std::map<std::pair<int, int>, int64_t> phi;
    std::map<int, int64_t> dft;
    B2Layout L;
    b2_layout(nprob, nblocks, b2_arena(nprob, Nx, Ny, bc, &phi, &dft), &L);
    return L.bytes;
}

extern "C" int64_t sfa_batch2d_plan_ext_bytes(int nprob, int nblocks)
{
    if (nprob < 1 || nblocks < 1) return -1;
    B2Layout L;
    b2_layout(nprob, nblocks, 0, &L);
    return L.bytes;
}

extern "C" int64_t sfa_batch2d_workspace(int nprob, const int* Nx, const int* Ny, int nblocks)
{
    if (nprob < 1 || nblocks < 1 || !Nx || !Ny) return -1;
    return b2_wstride(nprob, Nx, Ny) * 8 * (int64_t)nblocks;
}

extern "C" int64_t sfa_batch2d_leg_elems(int Nx) { return (Nx >= 8 && Nx % 2 == 0) ? leg_elems(Nx) : -1; }
extern "C" int64_t sfa_batch2d_dft_elems(int Ny) { return (Ny >= 4 && Ny % 2 == 0) ? dft_elems(Ny) : -1; }

extern "C" int sfa_batch2d_leg_matrices(int Nx, int bc, double* out, void* stream)
{
    SFA_CHECK(Nx >= 8 && Nx <= 1024 && Nx % 2 == 0 && (bc == 0 || bc == 1) && out, SFA_ERR_ARG,
              "need even Nx in [8, 1024], bc 0/1, out");
    return build_leg(Nx, bc, out, (hipStream_t)stream);
}

extern "C" int sfa_batch2d_dft_matrices(int Ny, double* out, void* stream)
{
    SFA_CHECK(Ny >= 4 && Ny <= 1024 && Ny % 2 == 0 && out, SFA_ERR_ARG, "need even Ny in [4, 1024], out");
    return build_dft(Ny, out, (hipStream_t)stream);
}

extern "C" int sfa_batch2d_plan(int nprob, const int* Nx, const int* Ny, const double* alpha,
                                const int* bc, const int64_t* f_off, const int64_t* u_off,
                                int64_t f_elems, int64_t u_elems, int nblocks, void* plan,
                                int64_t plan_bytes, void* stream)
{
    int rc;
    if ((rc = b2_check(nprob, Nx, Ny, bc, nblocks))) return rc;
    SFA_CHECK(alpha && f_off && u_off && plan, SFA_ERR_ARG, "NULL argument");
    std::map<std::pair<int, int>, int64_t> phi;
    std::map<int, int64_t> dft;
    const int64_t mats = b2_arena(nprob, Nx, Ny, bc, &phi, &dft);
    B2Layout L;
    b2_layout(nprob, nblocks, mats, &L);
    SFA_CHECK(plan_bytes >= L.bytes, SFA_ERR_ARG, "plan buffer too small (%lld < %lld)",
              (long long)plan_bytes, (long long)L.bytes);
    hipStream_t s = (hipStream_t)stream;
    double* dm = (double*)((char*)plan + L.off_mats);
    for (auto& kv : phi)
        if ((rc = build_leg(kv.first.first, kv.first.second, dm + kv.second, s))) return rc;
    for (auto& kv : dft)
        if ((rc = build_dft(kv.first, dm + kv.second, s))) return rc;
    std::vector<int64_t> po(nprob), dfo(nprob);
    for (int i = 0; i < nprob; ++i) {
        po[i] = phi[std::make_pair(Nx[i], bc[i] ? 1 : 0)];
        dfo[i] = dft[Ny[i]];
    }
    return b2_plan_core(nprob, Nx, Ny, alpha, bc, f_off, u_off, f_elems, u_elems, dm, po.data(),
                        dfo.data(), mats, nblocks, plan, plan_bytes, s);
}

extern "C" int sfa_batch2d_plan_ext(int nprob, const int* Nx, const int* Ny, const double* alpha,
                                    const int* bc, const int64_t* f_off, const int64_t* u_off,
                                    int64_t f_elems, int64_t u_elems, const double* mats,
                                    const int64_t* phi_off, const int64_t* df_off,
                                    int64_t mats_elems, int nblocks, void* plan,
                                    int64_t plan_bytes, void* stream)
{
    int rc;
    if ((rc = b2_check(nprob, Nx, Ny, bc, nblocks))) return rc;
    SFA_CHECK(alpha && f_off && u_off && plan && mats && phi_off && df_off, SFA_ERR_ARG,
              "NULL argument");
    return b2_plan_core(nprob, Nx, Ny, alpha, bc, f_off, u_off, f_elems, u_elems, mats, phi_off,
                        df_off, mats_elems, nblocks, plan, plan_bytes, (hipStream_t)stream);
}

extern "C" int sfa_batch2d_solve(const double* F, double* U, const void* plan, void* work,
                                 int64_t work_bytes, int variant, void* stream)
{
    B2Info pi;
    {
        std::lock_guard<std::mutex> lk(g_b2_mu);
        auto it = g_b2.find(plan);
        SFA_CHECK(it != g_b2.end(), SFA_ERR_ARG, "plan was not built by sfa_batch2d_plan(_ext)");
        pi = it->second;
    }
    SFA_CHECK(F && U && work, SFA_ERR_ARG, "NULL argument");
    SFA_CHECK(work_bytes >= pi.wstride * 8 * (int64_t)pi.nblocks, SFA_ERR_ARG,
              "workspace too small (%lld < %lld)", (long long)work_bytes,
              (long long)(pi.wstride * 8 * (int64_t)pi.nblocks));
    const int nv = sfa_batch2d_variants(nullptr, nullptr, 0);
    SFA_CHECK(variant >= 0 && variant < nv, SFA_ERR_ARG, "variant must be in [0, %d)", nv);
    const char* pb = (const char*)plan;
    B2Args a;
    a.F = F;
    a.U = U;
    a.probs = (const B2Prob*)(pb + pi.off_probs);
    a.blk_ptr = (const int*)(pb + pi.off_blk);
    a.order = (const int*)(pb + pi.off_order);
    a.mats = pi.mats;
    a.work = (double*)work;
    a.wstride = pi.wstride;
    hipStream_t s = (hipStream_t)stream;
    switch (variant) {
        case 0: return launch_b2<64, 64, 16>(a, pi.nblocks, s);
        case 1: return launch_b2<32, 64, 16>(a, pi.nblocks, s);
        case 2: return launch_b2<64, 128, 16>(a, pi.nblocks, s);
        case 3: return launch_b2<128, 64, 16>(a, pi.nblocks, s);
        default: return launch_b2<64, 64, 32>(a, pi.nblocks, s);
    }
}

extern "C" int sfa_batch2d_release(const void* plan)
{
    std::lock_guard<std::mutex> lk(g_b2_mu);
    g_b2.erase(plan);
    g_b2w.erase(plan);
    return SFA_OK;
}

extern "C" int64_t sfa_batch2d_wide_plan_bytes(int nprob, const int* Nx, const int* Ny)
{
    if (nprob < 1 || !Nx || !Ny) return -1;
    int64_t nt[4], ot[4], ow, work;
    b2_wide_shape(nprob, Nx, Ny, nt, &work);
    return b2_wide_bytes(nprob, nt, ot, &ow);
}

extern "C" int64_t sfa_batch2d_wide_workspace(int nprob, const int* Nx, const int* Ny)
{
    if (nprob < 1 || !Nx || !Ny) return -1;
    int64_t nt[4], work;
    b2_wide_shape(nprob, Nx, Ny, nt, &work);
    return work * 8;
}

extern "C" int sfa_batch2d_wide_plan(int nprob, const int* Nx, const int* Ny, const double* alpha,
                                     const int* bc, const int64_t* f_off, const int64_t* u_off,
                                     int64_t f_elems, int64_t u_elems, const double* mats,
                                     const int64_t* phi_off, const int64_t* df_off,
                                     int64_t mats_elems, void* plan, int64_t plan_bytes, void* stream)
{
    int rc;
    if ((rc = b2_check(nprob, Nx, Ny, bc, 1))) return rc;
    SFA_CHECK(alpha && f_off && u_off && plan && mats && phi_off && df_off, SFA_ERR_ARG,
              "NULL argument");
    std::vector<B2Prob> probs;
    if ((rc = b2_table(nprob, Nx, Ny, alpha, bc, f_off, u_off, f_elems, u_elems, phi_off, df_off,
                       mats_elems, probs)))
        return rc;
    B2WideInfo wi;
    wi.nprob = nprob;
    int* cls = wi.cls;
    b2_wide_shape(nprob, Nx, Ny, wi.ntask, &wi.work_doubles, cls);
    const int64_t bytes = b2_wide_bytes(nprob, wi.ntask, wi.off_task, &wi.off_woff);
    SFA_CHECK(plan_bytes >= bytes, SFA_ERR_ARG, "plan buffer too small (%lld < %lld)",
              (long long)plan_bytes, (long long)bytes);
    wi.off_probs = 0;
    wi.mats = mats;
    std::vector<B2Task> tasks[4];
    std::vector<int64_t> woff(nprob);
    int64_t w = 0;
    for (int i = 0; i < nprob; ++i) {
        const int M = Nx[i] - 2, L = Ny[i] + 2, X = Nx[i], Y = Ny[i];
        const int mm[4] = {M, M, X, X}, nn[4] = {Y, L, L, Y};
        for (int st = 0; st < 4; ++st) {
            const int bm = kB2WTile[cls[st]][0], bn = kB2WTile[cls[st]][1];
            for (int m0 = 0; m0 < mm[st]; m0 += bm)
                for (int n0 = 0; n0 < nn[st]; n0 += bn) tasks[st].push_back(B2Task{i, m0, n0, cls[st]});
        }
        woff[i] = w;
        w += ((int64_t)M * L + (int64_t)X * L + 31) / 32 * 32;
    }
    hipStream_t s = (hipStream_t)stream;
    char* pb = (char*)plan;
    SFA_HIP(hipMemcpyAsync(pb, probs.data(), probs.size() * sizeof(B2Prob), hipMemcpyHostToDevice, s));
    for (int st = 0; st < 4; ++st)
        SFA_HIP(hipMemcpyAsync(pb + wi.off_task[st], tasks[st].data(), tasks[st].size() * sizeof(B2Task),
                               hipMemcpyHostToDevice, s));
    SFA_HIP(hipMemcpyAsync(pb + wi.off_woff, woff.data(), woff.size() * 8, hipMemcpyHostToDevice, s));
    // the solve's chains run sequentially over a problem's rows: the problems
    // with the most rows start first so they do not finish last
    std::vector<int> order(nprob);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return Nx[x] > Nx[y]; });
    SFA_HIP(hipMemcpyAsync(pb + wi.off_woff + r256((int64_t)nprob * 8), order.data(), order.size() * 4,
                           hipMemcpyHostToDevice, s));
    SFA_HIP(hipStreamSynchronize(s));  // host vectors are copied before they go out of scope
    std::lock_guard<std::mutex> lk(g_b2_mu);
    g_b2.erase(plan);  // the address may have held a persistent-mode plan
    g_b2w[plan] = wi;
    return SFA_OK;
}

extern "C" int sfa_batch2d_wide_solve(const double* F, double* U, const void* plan, void* work,
                                      int64_t work_bytes, void* stream)
{
    B2WideInfo wi;
    {
        std::lock_guard<std::mutex> lk(g_b2_mu);
        auto it = g_b2w.find(plan);
        SFA_CHECK(it != g_b2w.end(), SFA_ERR_ARG, "plan was not built by sfa_batch2d_wide_plan");
        wi = it->second;
    }
    SFA_CHECK(F && U && work, SFA_ERR_ARG, "NULL argument");
    SFA_CHECK(work_bytes >= wi.work_doubles * 8, SFA_ERR_ARG, "workspace too small (%lld < %lld)",
              (long long)work_bytes, (long long)(wi.work_doubles * 8));
    const char* pb = (const char*)plan;
    B2WideArgs a;
    a.F = F;
    a.U = U;
    a.probs = (const B2Prob*)pb;
    a.woff = (const int64_t*)(pb + wi.off_woff);
    a.order = (const int*)(pb + wi.off_woff + r256((int64_t)wi.nprob * 8));
    a.mats = wi.mats;
    a.work = (double*)work;
    hipStream_t s = (hipStream_t)stream;
    const dim3 blk(B2_THREADS);
    int rc;
    a.tasks = (const B2Task*)(pb + wi.off_task[0]);
    if ((rc = launch_wide_stage<0>(a, wi.cls[0], wi.ntask[0], s))) return rc;
    a.tasks = (const B2Task*)(pb + wi.off_task[1]);
    if ((rc = launch_wide_stage<1>(a, wi.cls[1], wi.ntask[1], s))) return rc;
    hipLaunchKernelGGL(batch2d_wide_solve_kernel, dim3((unsigned)wi.nprob), blk, 0, s, a);
    SFA_LAUNCH_CHECK();
    a.tasks = (const B2Task*)(pb + wi.off_task[2]);
    if ((rc = launch_wide_stage<2>(a, wi.cls[2], wi.ntask[2], s))) return rc;
    a.tasks = (const B2Task*)(pb + wi.off_task[3]);
    return launch_wide_stage<3>(a, wi.cls[3], wi.ntask[3], s);
}

extern "C" int sfa_mfma_f64_layout_probe(double* out, void* stream)
{
    SFA_CHECK(out != nullptr, SFA_ERR_ARG, "out is NULL");
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, out);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}
