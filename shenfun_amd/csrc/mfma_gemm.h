// FP64 MFMA GEMM tile shared by the batched 2D solver and the dense line
// contraction: C (m x n) = A (m x k) B (k x n), row-major, one workgroup of
// GEMM_THREADS threads per BM x BN output tile.
#pragma once
#include "common.h"

namespace sfa {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int GEMM_THREADS = 256;

// MFMA f64 16x16x4 operand layout (lane l of the wave): A[i][k] at i = l % 16,
// k = l / 16; B[k][j] at j = l % 16, k = l / 16; D[i][j] in register r of lane
// l with i = 4 r + l / 16, j = l % 16 (checked by sfa_mfma_f64_layout_probe).
__device__ inline v4d mfma16(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}


// C (m x n, ldc) = A (m x k, lda) B (k x n, ldb), all row-major FP64, by the
// whole workgroup (4 waves in a 2 x 2 grid over a BM x BN block tile, K in
// steps of BK).  sA: BK x (BM + 1) doubles, sB: BK x BN doubles of LDS.
// Out-of-range operand elements load as zero; stores are bounds-checked.
// Each output is one MFMA accumulation chain over K in ascending order, so
// the result does not depend on BM, BN or BK.
template <int BM, int BN, int BK>
__device__ void gemm_tile(const double* __restrict__ A, int64_t lda, const double* __restrict__ B,
                          int64_t ldb, double* __restrict__ C, int64_t ldc, int m, int n, int k,
                          int m0, int n0, double* sA, double* sB)
{
    constexpr int SA = BM + 1;
    constexpr int TI = BM / 32, TJ = BN / 32;
    constexpr int LA = BM * BK / GEMM_THREADS, LB = BK * BN / GEMM_THREADS;
    static_assert(LA >= 1 && LB >= 1 && TI >= 1 && TJ >= 1, "tile too small");
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));  // per-tile indices: nothing hoisted across the caller's tile loop
    const int lane = tid & 63, wv = tid >> 6;
    const int wi = wv >> 1, wj = wv & 1;
    const int lr = lane & 15, lk = lane >> 4;
    v4d acc[TI][TJ];
#pragma unroll
    for (int a = 0; a < TI; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
    bool rin[TI], cin[TJ];  // this wave's 16 x 16 sub-tiles that intersect C
#pragma unroll
    for (int a = 0; a < TI; ++a) rin[a] = m0 + wi * (BM / 2) + a * 16 < m;
#pragma unroll
    for (int b = 0; b < TJ; ++b) cin[b] = n0 + wj * (BN / 2) + b * 16 < n;
    double ra[LA], rb[LB];
    auto load = [&](int k0) {
#pragma unroll
        for (int r = 0; r < LA; ++r) {
            const int e = tid + GEMM_THREADS * r;
            const int i = e / BK, kk = e % BK;
            const int gi = m0 + i, gk = k0 + kk;
            ra[r] = (gi < m && gk < k) ? A[(int64_t)gi * lda + gk] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < LB; ++r) {
            const int e = tid + GEMM_THREADS * r;
            const int kk = e / BN, j = e % BN;
            const int gk = k0 + kk, gj = n0 + j;
            rb[r] = (gk < k && gj < n) ? B[(int64_t)gk * ldb + gj] : 0.0;
        }
    };
    load(0);
    for (int k0 = 0; k0 < k; k0 += BK) {
        __syncthreads();  // the previous step's fragment reads are done
#pragma unroll
        for (int r = 0; r < LA; ++r) {
            const int e = tid + GEMM_THREADS * r;
            sA[(e % BK) * SA + e / BK] = ra[r];
        }
#pragma unroll
        for (int r = 0; r < LB; ++r) {
            const int e = tid + GEMM_THREADS * r;
            sB[(e / BN) * BN + e % BN] = rb[r];
        }
        __syncthreads();
        if (k0 + BK < k) load(k0 + BK);  // prefetch overlaps the MFMAs
        // MFMAs whose operands are all zero padding are skipped: K steps of 4
        // past k, and 16 x 16 sub-tiles wholly outside C (never stored).  All
        // conditions are wave-uniform; a skipped step would add exact zeros.
#pragma unroll
        for (int s = 0; s < BK / 4; ++s) {
            if (k0 + 4 * s < k) {
                const int kk = 4 * s + lk;
                double af[TI], bf[TJ];
#pragma unroll
                for (int a = 0; a < TI; ++a) af[a] = sA[kk * SA + wi * (BM / 2) + a * 16 + lr];
#pragma unroll
                for (int b = 0; b < TJ; ++b) bf[b] = sB[kk * BN + wj * (BN / 2) + b * 16 + lr];
#pragma unroll
                for (int a = 0; a < TI; ++a)
#pragma unroll
                    for (int b = 0; b < TJ; ++b)
                        if (rin[a] && cin[b]) acc[a][b] = mfma16(af[a], bf[b], acc[a][b]);
            }
        }
    }
#pragma unroll
    for (int a = 0; a < TI; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b) {
            const int gj = n0 + wj * (BN / 2) + b * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wi * (BM / 2) + a * 16 + 4 * r + lk;
                if (gi < m && gj < n) C[(int64_t)gi * ldc + gj] = acc[a][b][r];
            }
        }
}

}  // namespace sfa
