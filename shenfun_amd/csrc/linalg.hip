// Dense per-line transforms: Legendre transform matrices and a strided
// batched matrix-vector product along one axis (FP64, LDS tiled).
#include "common.h"
#include "mfma_gemm.h"

namespace sfa {

// One thread per node x_j: three-term recurrence of transforms.pyx:219-323
// (L_{m+1} = (1/anm_m) x L_m - (anp_m/anm_m) L_{m-1}, anm_m = (m+1)/(2m+1),
// anp_m = m/(2m+1)) and the (optionally composite) matrix entries.
__global__ void legendre_matrix_kernel(const double* __restrict__ x, const double* __restrict__ w,
                                       int Nx, int N, const double* __restrict__ K2, int analysis,
                                       double* __restrict__ M)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= Nx) return;
    const double xj = x[j];
    const double wj = analysis ? w[j] : 1.0;
    // window of L_{k}, L_{k+1}, L_{k+2}
    double Lm = 1.0;  // L_0
    double L0 = xj;   // L_1 = x / anm_0, anm_0 = 1
    double L1 = (1.0 / (2.0 / 3.0)) * xj * L0 - ((1.0 / 3.0) / (2.0 / 3.0)) * Lm;  // L_2
    for (int k = 0; k < N; ++k) {
        // Lm = L_k, L0 = L_{k+1}, L1 = L_{k+2}
        double v = Lm;
        const bool comp = (K2 != nullptr);
        if (comp) v = (k < N - 2) ? (Lm + K2[k] * L1) : (analysis ? Lm : 0.0);
        if (analysis)
            M[(int64_t)k * Nx + j] = wj * v;
        else
            M[(int64_t)j * N + k] = v;
        // advance: L_{k+3} = (1/anm_{k+2}) x L_{k+2} - (anp_{k+2}/anm_{k+2}) L_{k+1}
        const int m = k + 2;
        const double anm = (m + 1.0) / (2.0 * m + 1.0);
        const double anp = (double)m / (2.0 * m + 1.0);
        const double s1 = 1.0 / anm, s2 = anp / anm;
        const double Ln = s1 * xj * L1 - s2 * L0;
        Lm = L0;
        L0 = L1;
        L1 = Ln;
    }
}

// One thread per node x_j: the generic three-term recurrence of
// transforms.pyx:219-323 on the caller's DIA rows of the recursion matrix,
// x P_n = anm_n P_{n+1} + ann_n P_n + anp_n P_{n-1} (ann = 0 without a middle
// row), P_0 = 1, in the reference's operation order.
__global__ void recursion_matrix_kernel(const double* __restrict__ x, const double* __restrict__ w,
                                        int Nx, int N, const double* __restrict__ anm,
                                        const double* __restrict__ ann,
                                        const double* __restrict__ anp, int analysis,
                                        double* __restrict__ M)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= Nx) return;
    const double xj = x[j];
    const double wj = analysis ? w[j] : 1.0;
    const double a0 = ann ? ann[0] : 0.0, a1 = ann ? ann[1] : 0.0;
    double Lnm = 1.0;
    double Ln = (xj - a0) / anm[0];
    double Lnp = (xj - a1) / anm[1] * Ln - anp[1] / anm[1] * Lnm;
    for (int i = 0; i < N; ++i) {
        if (analysis)
            M[(int64_t)i * Nx + j] = Lnm * wj;
        else
            M[(int64_t)j * N + i] = Lnm;
        const double s1 = 1.0 / anm[i + 2];
        const double s2 = anp[i + 2] / anm[i + 2];
        const double a00 = ann ? ann[i + 2] : 0.0;
        Lnm = Ln;
        Ln = Lnp;
        Lnp = s1 * (xj - a00) * Ln - s2 * Lnm;
    }
}

// out(line, k) = sum_j M[k][j] in(line, j); tile 64 k x 16 lines, j chunks of 32.
template <bool CPLX>
__global__ void __launch_bounds__(256) line_matmul_kernel(
    const char* __restrict__ in, char* __restrict__ out, int nin, int nout,
    const double* __restrict__ M, int64_t in_sa, int64_t in_s1, int64_t in_s2, int64_t out_sa,
    int64_t out_s1, int64_t out_s2, int64_t d2, int64_t nlines, int t_fast_in, int t_fast_out)
{
    constexpr int BK = 64, BL = 16, BJ = 32;
    __shared__ double Ms[BK][BJ + 1];
    __shared__ double2 Xs[BJ][BL + 1];
    const int tid = threadIdx.x;
    const int64_t L0 = (int64_t)blockIdx.x * BL;
    const int k0 = blockIdx.y * BK;
    const int tl = tid % BL;       // line within tile
    const int tk = (tid / BL) * 4;  // first of 4 rows
    double2 acc[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = make_double2(0.0, 0.0);

    for (int j0 = 0; j0 < nin; j0 += BJ) {
        // M tile
        for (int e = tid; e < BK * BJ; e += 256) {
            const int kk = e / BJ, jj = e % BJ;
            const int k = k0 + kk, j = j0 + jj;
            Ms[kk][jj] = (k < nout && j < nin) ? M[(int64_t)k * nin + j] : 0.0;
        }
        // X tile
        for (int e = tid; e < BJ * BL; e += 256) {
            int ll, jj;
            if (t_fast_in) { ll = e % BL; jj = e / BL; }
            else { jj = e % BJ; ll = e / BJ; }
            const int64_t L = L0 + ll;
            const int j = j0 + jj;
            double2 v = make_double2(0.0, 0.0);
            if (L < nlines && j < nin) {
                const int64_t i1 = L / d2, i2 = L - i1 * d2;
                const int64_t off = i1 * in_s1 + i2 * in_s2 + (int64_t)j * in_sa;
                if (CPLX) v = reinterpret_cast<const double2*>(in)[off];
                else v.x = reinterpret_cast<const double*>(in)[off];
            }
            Xs[jj][ll] = v;
        }
        __syncthreads();
#pragma unroll 8
        for (int jj = 0; jj < BJ; ++jj) {
            const double2 xv = Xs[jj][tl];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double m = Ms[tk + r][jj];
                acc[r].x += m * xv.x;
                if (CPLX) acc[r].y += m * xv.y;
            }
        }
        __syncthreads();
    }
    // store via LDS transpose for coalescing when lines are contiguous
    const int64_t L = L0 + tl;
    if (L < nlines) {
        const int64_t i1 = L / d2, i2 = L - i1 * d2;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = k0 + tk + r;
            if (k < nout) {
                const int64_t off = i1 * out_s1 + i2 * out_s2 + (int64_t)k * out_sa;
                if (CPLX) reinterpret_cast<double2*>(out)[off] = acc[r];
                else reinterpret_cast<double*>(out)[off] = acc[r].x;
            }
        }
    }
    (void)t_fast_out;
}

// Few-lines form of the same product: one wave per MV_R consecutive rows k,
// lanes striding j (each M row read once, coalesced), MV_NB lines per
// blockIdx.y group.  Each output is lane-strided partial sums followed by a
// fixed butterfly, so it depends only on its own line's data.
constexpr int MV_R = 2, MV_NB = 4;

template <bool CPLX>
__global__ void __launch_bounds__(256) line_matvec_kernel(
    const char* __restrict__ in, char* __restrict__ out, int nin, int nout,
    const double* __restrict__ M, int64_t in_sa, int64_t in_s1, int64_t in_s2, int64_t out_sa,
    int64_t out_s1, int64_t out_s2, int64_t d2, int64_t nlines)
{
    const int lane = threadIdx.x & 63;
    const int k0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * MV_R;
    if (k0 >= nout) return;  // whole wave; no workgroup barrier below
    const int64_t L0 = (int64_t)blockIdx.y * MV_NB;
    const int nb = (int)(nlines - L0 < MV_NB ? nlines - L0 : MV_NB);
    int64_t base[MV_NB];
#pragma unroll
    for (int b = 0; b < MV_NB; ++b) {
        const int64_t L = L0 + (b < nb ? b : 0);
        const int64_t i1 = L / d2, i2 = L - i1 * d2;
        base[b] = i1 * in_s1 + i2 * in_s2;
    }
    double ax[MV_R][MV_NB], ay[MV_R][MV_NB];
#pragma unroll
    for (int r = 0; r < MV_R; ++r)
#pragma unroll
        for (int b = 0; b < MV_NB; ++b) ax[r][b] = ay[r][b] = 0.0;
    const double* Mk[MV_R];
#pragma unroll
    for (int r = 0; r < MV_R; ++r) Mk[r] = M + (int64_t)(k0 + r < nout ? k0 + r : k0) * nin;
#pragma unroll 2
    for (int j = lane; j < nin; j += 64) {
        double m[MV_R];
#pragma unroll
        for (int r = 0; r < MV_R; ++r) m[r] = Mk[r][j];
#pragma unroll
        for (int b = 0; b < MV_NB; ++b) {
            if (b < nb) {
                const int64_t off = base[b] + (int64_t)j * in_sa;
                double2 v;
                if (CPLX) v = reinterpret_cast<const double2*>(in)[off];
                else v = make_double2(reinterpret_cast<const double*>(in)[off], 0.0);
#pragma unroll
                for (int r = 0; r < MV_R; ++r) {
                    ax[r][b] += m[r] * v.x;
                    if (CPLX) ay[r][b] += m[r] * v.y;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < MV_R; ++r)
#pragma unroll
        for (int b = 0; b < MV_NB; ++b)
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                ax[r][b] += __shfl_xor(ax[r][b], o);
                if (CPLX) ay[r][b] += __shfl_xor(ay[r][b], o);
            }
    if (lane == 0) {
#pragma unroll
        for (int b = 0; b < MV_NB; ++b) {
            if (b >= nb) break;
            const int64_t L = L0 + b, i1 = L / d2, i2 = L - i1 * d2;
#pragma unroll
            for (int r = 0; r < MV_R; ++r) {
                const int k = k0 + r;
                if (k < nout) {
                    const int64_t off = i1 * out_s1 + i2 * out_s2 + (int64_t)k * out_sa;
                    if (CPLX) reinterpret_cast<double2*>(out)[off] = make_double2(ax[r][b], ay[r][b]);
                    else reinterpret_cast<double*>(out)[off] = ax[r][b];
                }
            }
        }
    }
}

// Lines along a strided axis whose other dims flatten to unit-stride line
// indices: the lines are the columns of X (nin x ncol, row pitch ldx doubles;
// complex lines are two real columns), and the contraction is the FP64 MFMA
// GEMM Y = M X, one workgroup per 64 x 64 tile of Y.
__global__ void __launch_bounds__(GEMM_THREADS)
    line_gemm_kernel(const double* __restrict__ M, const double* __restrict__ X, double* __restrict__ Y,
                     int nout, int nin, int64_t ncol, int64_t ldx, int64_t ldy)
{
    __shared__ double sA[16 * (64 + 1)];
    __shared__ double sB[16 * 64];
    gemm_tile<64, 64, 16>(M, nin, X, ldx, Y, ldy, nout, (int)ncol, nin,
                          (int)blockIdx.y * 64, (int)blockIdx.x * 64, sA, sB);
}

}  // namespace sfa

using namespace sfa;

extern "C" int sfa_legendre_matrix(const double* x, const double* w, int64_t Nx, int64_t N,
                                   const double* K2, int analysis, double* M, void* stream)
{
    SFA_CHECK(Nx >= 1 && N >= 1, SFA_ERR_ARG, "Nx, N must be >= 1");
    SFA_CHECK(!analysis || w != nullptr, SFA_ERR_ARG, "analysis matrix needs weights");
    SFA_CHECK(K2 == nullptr || N >= 3, SFA_ERR_ARG, "composite basis needs N >= 3");
    dim3 grid((unsigned)((Nx + 127) / 128)), block(128);
    hipLaunchKernelGGL(legendre_matrix_kernel, grid, block, 0, (hipStream_t)stream, x, w, (int)Nx,
                       (int)N, K2, analysis, M);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_recursion_matrix(const double* x, const double* w, int64_t Nx, const double* a,
                                    int64_t a_rows, int64_t a_len, int64_t N, int analysis,
                                    double* M, void* stream)
{
    SFA_CHECK(Nx >= 1 && N >= 1, SFA_ERR_ARG, "Nx, N must be >= 1");
    SFA_CHECK(a != nullptr && (a_rows == 2 || a_rows == 3), SFA_ERR_ARG,
              "a must hold 2 (anm, anp) or 3 (anm, ann, anp) DIA rows");
    SFA_CHECK(a_len >= N + 2 && a_len >= 2, SFA_ERR_ARG, "recursion rows must have >= N + 2 entries");
    SFA_CHECK(!analysis || w != nullptr, SFA_ERR_ARG, "analysis matrix needs weights");
    const double* anm = a;
    const double* ann = a_rows == 3 ? a + a_len : nullptr;
    const double* anp = a + (a_rows - 1) * a_len;
    dim3 grid((unsigned)((Nx + 127) / 128)), block(128);
    hipLaunchKernelGGL(recursion_matrix_kernel, grid, block, 0, (hipStream_t)stream, x, w, (int)Nx,
                       (int)N, anm, ann, anp, analysis, M);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_line_matmul(const void* in, void* out, int dtype, int ndim,
                               const int64_t* in_shape, const int64_t* in_strides,
                               const int64_t* out_strides, int axis, int64_t nout, const double* M,
                               void* stream)
{
    Geom gi, go;
    int rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "bad dtype %d", dtype);
    SFA_CHECK(in != out, SFA_ERR_ARG, "line_matmul cannot run in place");
    if ((rc = make_geom(ndim, in_shape, in_strides, axis, &gi))) return rc;
    if ((rc = make_geom(ndim, in_shape, out_strides, axis, &go))) return rc;
    SFA_CHECK(nout >= 1, SFA_ERR_ARG, "nout must be >= 1");
    const int64_t nlines = gi.d1 * gi.d2;
    auto absl = [](int64_t v) { return v < 0 ? -v : v; };
    const int tfi = absl(gi.s2) < absl(gi.sa) ? 1 : 0;
    const int tfo = absl(go.s2) < absl(go.sa) ? 1 : 0;
    // lines enumerated in C order with d1 slow: both geometries share d1, d2
    auto flat = [&](const Geom& g) {  // line L at offset L: a column of a row-major matrix
        return g.s2 == 1 && (g.d1 == 1 || g.s1 == g.d2) && g.sa >= nlines;
    };
    if (nlines > 16 && flat(gi) && flat(go) && nlines * (dtype == SFA_C128 ? 2 : 1) <= INT32_MAX) {
        const int w = dtype == SFA_C128 ? 2 : 1;  // doubles per element
        const int64_t ncol = nlines * w;
        dim3 g((unsigned)((ncol + 63) / 64), (unsigned)((nout + 63) / 64));
        hipLaunchKernelGGL(line_gemm_kernel, g, dim3(GEMM_THREADS), 0, (hipStream_t)stream, M,
                           (const double*)in, (double*)out, (int)nout, (int)gi.n, ncol, gi.sa * w,
                           go.sa * w);
        SFA_LAUNCH_CHECK();
        return SFA_OK;
    }
    if (nlines <= 16) {  // few lines: row-parallel matvec form
        dim3 g((unsigned)((nout + 4 * MV_R - 1) / (4 * MV_R)), (unsigned)((nlines + MV_NB - 1) / MV_NB));
        if (dtype == SFA_C128)
            hipLaunchKernelGGL(line_matvec_kernel<true>, g, dim3(256), 0, (hipStream_t)stream,
                               (const char*)in, (char*)out, (int)gi.n, (int)nout, M, gi.sa, gi.s1,
                               gi.s2, go.sa, go.s1, go.s2, gi.d2, nlines);
        else
            hipLaunchKernelGGL(line_matvec_kernel<false>, g, dim3(256), 0, (hipStream_t)stream,
                               (const char*)in, (char*)out, (int)gi.n, (int)nout, M, gi.sa, gi.s1,
                               gi.s2, go.sa, go.s1, go.s2, gi.d2, nlines);
        SFA_LAUNCH_CHECK();
        return SFA_OK;
    }
    dim3 grid((unsigned)((nlines + 15) / 16), (unsigned)((nout + 63) / 64)), block(256);
    if (dtype == SFA_C128)
        hipLaunchKernelGGL(line_matmul_kernel<true>, grid, block, 0, (hipStream_t)stream,
                           (const char*)in, (char*)out, (int)gi.n, (int)nout, M, gi.sa, gi.s1,
                           gi.s2, go.sa, go.s1, go.s2, gi.d2, nlines, tfi, tfo);
    else
        hipLaunchKernelGGL(line_matmul_kernel<false>, grid, block, 0, (hipStream_t)stream,
                           (const char*)in, (char*)out, (int)gi.n, (int)nout, M, gi.sa, gi.s1,
                           gi.s2, go.sa, go.s1, go.s2, gi.d2, nlines, tfi, tfo);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// ----------------------------------------------------------------- point evaluation
// out[p] = sum_{k0,k1,k2} c[k0][k1][k2] E0[p][k0] E1[p][k1] E2[p][k2] for
// scattered points p (TensorProductSpace.eval, tensorproductspace.py:668-760):
// E_d (npts, n_d) complex128 are each axis' basis functions at the points'
// coordinate d (stencils, weights and boundary functions folded in by the
// caller), c contiguous complex128 (n0, n1, n2) (n2 = 1, E2 = 1 for 2D).  One
// workgroup per point: threads stride the (k0, k1) pairs, each sums over k2,
// then a workgroup reduction.
namespace sfa {
__global__ void __launch_bounds__(256)
    tensor_eval_kernel(const double2* __restrict__ c, int n0, int n1, int n2,
                       const double2* __restrict__ E0, const double2* __restrict__ E1,
                       const double2* __restrict__ E2, double2* __restrict__ out)
{
    __shared__ double2 red[256];
    const int p = blockIdx.x, tid = threadIdx.x;
    const double2* e0 = E0 + (int64_t)p * n0;
    const double2* e1 = E1 + (int64_t)p * n1;
    const double2* e2 = E2 + (int64_t)p * n2;
    double2 acc = make_double2(0.0, 0.0);
    for (int q = tid; q < n0 * n1; q += blockDim.x) {
        const int k0 = q / n1, k1 = q - k0 * n1;
        const double2* row = c + ((int64_t)k0 * n1 + k1) * n2;
        double2 s = make_double2(0.0, 0.0);
        for (int k2 = 0; k2 < n2; ++k2) {
            const double2 a = row[k2], b = e2[k2];
            s.x += a.x * b.x - a.y * b.y;
            s.y += a.x * b.y + a.y * b.x;
        }
        const double2 a = e0[k0], b = e1[k1];
        const double2 w = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
        acc.x += w.x * s.x - w.y * s.y;
        acc.y += w.x * s.y + w.y * s.x;
    }
    red[tid] = acc;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (tid < h) {
            red[tid].x += red[tid + h].x;
            red[tid].y += red[tid + h].y;
        }
        __syncthreads();
    }
    if (tid == 0) out[p] = red[0];
}
}  // namespace sfa

extern "C" int sfa_tensor_eval(const void* c, int64_t n0, int64_t n1, int64_t n2, const void* E0,
                               const void* E1, const void* E2, void* out, int64_t npts, void* stream)
{
    SFA_CHECK(c && E0 && E1 && E2 && out, SFA_ERR_ARG, "tensor_eval: null array");
    SFA_CHECK(n0 >= 1 && n1 >= 1 && n2 >= 1 && n0 * n1 * n2 < (1ll << 30) && npts >= 0 &&
                  npts < (1ll << 31),
              SFA_ERR_ARG, "tensor_eval: bad sizes");
    if (npts == 0) return SFA_OK;
    hipLaunchKernelGGL(tensor_eval_kernel, dim3((unsigned)npts), dim3(256), 0, (hipStream_t)stream,
                       (const double2*)c, (int)n0, (int)n1, (int)n2, (const double2*)E0,
                       (const double2*)E1, (const double2*)E2, (double2*)out);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}
