// Register-resident power-of-two FFT kernels (FP64) for the hot shapes:
//   * C2C along a strided axis whose neighbour dimension is contiguous
//     (the y axis of (Nx, Ny, Nz/2+1)): 16 lines per block, lanes run over
//     the lines so every global access is a 256-byte row segment;
//   * R2C / C2R along the contiguous axis (z): rows packed as N/2 complex.
// Also the v2 fused x-pass (DCT-II -> stencil -> Helmholtz scan solve ->
// to_ortho -> DCT-III) on 16 flattened (ky, kz) lines per block.
#include <stdlib.h>

#include "fft_pow2.h"
#include "helm_scan.h"

namespace sfa {

constexpr int TL = 16;  // lines per block

// ----------------------------------------------------------------- C2C strided
struct StridedLines {
    const void* in;
    void* out;
    int64_t in_sa, in_s1, in_s2, out_sa, out_s1, out_s2, d2, nlines;
    double scale;
};

// one block of NL * Q threads: NL lines starting at line blk * NL
template <int N, bool INV, typename C2, int NL = TL>
__device__ __forceinline__ void c2c_strided_body(const StridedLines& g, const double2* __restrict__ twg,
                                                 int64_t blk)
{
    using Pl = Pow2<N>;
    constexpr int P = Pl::P, Q = Pl::Q, LS = Pl::LS;
    extern __shared__ __align__(16) unsigned char sm_raw[];
    C2* tw = reinterpret_cast<C2*>(sm_raw);
    C2* ex = tw + N;
    const C2* in = reinterpret_cast<const C2*>(g.in);
    C2* out = reinterpret_cast<C2*>(g.out);
    const int tid = threadIdx.x, nth = blockDim.x;
    const int t = tid % NL, j = tid / NL;
    const int64_t L = blk * NL + t;
    const bool valid = L < g.nlines;
    const int64_t i1 = valid ? L / g.d2 : 0, i2 = valid ? L - i1 * g.d2 : 0;
    for (int i = tid; i < N; i += nth) tw[i] = from_c64<C2>(twg[i]);
    C2 a[P];
    const C2* src = in + i1 * g.in_s1 + i2 * g.in_s2;
#pragma unroll
    for (int r = 0; r < P; ++r) a[r] = valid ? src[(int64_t)(j + Q * r) * g.in_sa] : cmk<C2>(0.0, 0.0);
    __syncthreads();
    pow2_fft<N, INV>(a, j, ex + t * LS, tw);
    if (valid) {
        C2* dst = out + i1 * g.out_s1 + i2 * g.out_s2;
        const typename CT<C2>::R sc = (typename CT<C2>::R)g.scale;
#pragma unroll
        for (int h = 0; h < P / Q; ++h)
#pragma unroll
            for (int k2 = 0; k2 < Q; ++k2)
                dst[(int64_t)pow2_out_index<N>(j, h, k2) * g.out_sa] = cscale(a[h * Q + k2], sc);
    }
}

template <int N, bool INV, typename C2 = double2>
__global__ void __launch_bounds__(TL * Pow2<N>::Q)
    c2c_strided_kernel(StridedLines g, const double2* __restrict__ twg)
{
    c2c_strided_body<N, INV, C2>(g, twg, blockIdx.x);
}

// the mixed form's two sub-arrays in one grid: the complex128 planes' blocks
// first, then the complex64 planes' blocks (two launches left the GPU
// under-filled during the few complex128 blocks).  out.flag != NULL: block 0
// (dispatched first) delivers the statistics in stats, which the previous
// kernel finished, while the other blocks transform.
template <int N, bool INV>
__global__ void __launch_bounds__(TL * Pow2<N>::Q)
    c2c_strided_dual_kernel(StridedLines g64, StridedLines g32, int64_t nblk64,
                            const double2* __restrict__ twg, const double* stats, StatOut out)
{
    int64_t b = blockIdx.x;
    if (out.flag != nullptr) {
        if (b == 0) {
            stat_deliver(stats, out);
            return;
        }
        --b;
    }
    if (b < nblk64)
        c2c_strided_body<N, INV, double2>(g64, twg, b);
    else
        c2c_strided_body<N, INV, float2>(g32, twg, b - nblk64);
}

// ----------------------------------------------------------------- R2C contiguous
// Row exchange areas of the z kernels: with Q < 16 lanes per row, a stride of
// Q (mod 16) complex elements puts consecutive rows of a wave in
// complementary LDS banks (the plain P(Q+1)+1 stride made them collide).
template <int H>
constexpr int zrow_stride()
{
    return Pow2<H>::Q >= 16 ? Pow2<H>::LS
                            : Pow2<H>::LS + ((Pow2<H>::Q - Pow2<H>::LS % 16) + 16) % 16;
}

// rows of 2H reals (stride in_rs), output rows of H+1 complex (stride out_rs).
// NR rows R0 .. R0 + NR - 1 (of nrows) per workgroup of NR * Q threads.
template <int H, int NR = TL>
__device__ __forceinline__ void r2c_contig_body(const double* __restrict__ in, double2* __restrict__ out,
                                                int64_t in_rs, int64_t out_rs, int64_t d2, int64_t in_s1,
                                                int64_t out_s1, int64_t nrows, double scale,
                                                const double2* __restrict__ twg /* 2H entries */,
                                                int64_t R0)
{
    using Pl = Pow2<H>;
    constexpr int P = Pl::P, Q = Pl::Q, LS = zrow_stride<H>();
    constexpr int N = 2 * H;
    extern __shared__ double2 sm[];
    // LDS: NR lines x LS exchange, then exp(-2 pi i m / H), m < H.  The
    // post-twiddles exp(-2 pi i k / N) are read from the global table twg
    // (L2-resident): keeping them out of LDS lets four blocks share a CU.
    double2* ex = sm;
    double2* twH = ex + NR * LS;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int j = tid % Q, t = tid / Q;  // j fastest: contiguous row segments
    const int64_t R = R0 + t;
    const bool valid = R < nrows;
    const int64_t i1 = valid ? R / d2 : 0, i2 = valid ? R - i1 * d2 : 0;
    for (int i = tid; i < H; i += nth) twH[i] = twg[2 * i];
    double2 a[P];
    const double* src = in + i1 * in_s1 + i2 * in_rs;
#pragma unroll
    for (int r = 0; r < P; ++r) {
        const int m = j + Q * r;
        a[r] = valid ? *reinterpret_cast<const double2*>(src + 2 * m) : make_double2(0.0, 0.0);
    }
    // post-twiddles of this thread's outputs k = j + Q i, issued early
    double2 twr[P + 1];
#pragma unroll
    for (int i = 0; i <= P; ++i) twr[i] = (j + Q * i <= H) ? twg[j + Q * i] : make_double2(0.0, 0.0);
    __syncthreads();
    double2* ln = ex + t * LS;
    pow2_fft<H, false>(a, j, ln, twH);
#pragma unroll
    for (int h = 0; h < P / Q; ++h)
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) ln[pow2_out_index<H>(j, h, k2)] = a[h * Q + k2];
    __syncthreads();
    if (valid) {
        double2* dst = out + i1 * out_s1 + i2 * out_rs;
#pragma unroll
        for (int i = 0; i <= P; ++i) {
            const int k = j + Q * i;
            if (k > H) break;
            const double2 zk = ln[k == H ? 0 : k];
            const double2 zc = cconj(ln[k == 0 ? 0 : H - k]);
            const double2 d = cmul(twr[i], csub(zk, zc));
            dst[k] = make_double2(0.5 * (zk.x + zc.x + d.y) * scale, 0.5 * (zk.y + zc.y - d.x) * scale);
        }
    }
}

template <int H>
__global__ void __launch_bounds__(TL * Pow2<H>::Q)
    r2c_contig_kernel(const double* __restrict__ in, double2* __restrict__ out, int64_t in_rs,
                      int64_t out_rs, int64_t d2, int64_t in_s1, int64_t out_s1, int64_t nrows,
                      double scale, const double2* __restrict__ twg /* 2H entries */)
{
    r2c_contig_body<H>(in, out, in_rs, out_rs, d2, in_s1, out_s1, nrows, scale, twg,
                       (int64_t)blockIdx.x * TL);
}

// Row store of the C2R kernels: out = scale z, plus, with ADD, the same row
// of add (the rows of add are contiguous, 2H doubles each); all of a
// thread's add loads are issued before its first store.
template <int H, bool ADD>
__device__ __forceinline__ void c2r_store(const double2 (&a)[Pow2<H>::P], int j, double* __restrict__ dst,
                                          const double* __restrict__ arow, double scale)
{
    using Pl = Pow2<H>;
    constexpr int P = Pl::P, Q = Pl::Q;
    double2 ad[P];
    if constexpr (ADD) {
#pragma unroll
        for (int h = 0; h < P / Q; ++h)
#pragma unroll
            for (int k2 = 0; k2 < Q; ++k2)
                ad[h * Q + k2] = *reinterpret_cast<const double2*>(arow + 2 * pow2_out_index<H>(j, h, k2));
    }
#pragma unroll
    for (int h = 0; h < P / Q; ++h)
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) {
            const int m = pow2_out_index<H>(j, h, k2);
            const double2 z = a[h * Q + k2];
            double2 o = make_double2(z.x * scale, z.y * scale);
            if constexpr (ADD) {
                o.x += ad[h * Q + k2].x;
                o.y += ad[h * Q + k2].y;
            }
            *reinterpret_cast<double2*>(dst + 2 * m) = o;
        }
}

// ADD: output rows are contiguous (row R at out + 2H R) and row R also gets
// row R % add_rows of add.
template <int H, bool ADD = false, int NR = TL>
__device__ __forceinline__ void c2r_contig_body(const double2* __restrict__ in, double* __restrict__ out,
                                                int64_t in_rs, int64_t out_rs, int64_t d2, int64_t in_s1,
                                                int64_t out_s1, int64_t nrows, double scale,
                                                const double2* __restrict__ twg /* 2H entries */,
                                                int64_t R0, const double* __restrict__ add = nullptr,
                                                int64_t add_rows = 1)
{
    using Pl = Pow2<H>;
    constexpr int P = Pl::P, Q = Pl::Q, LS = zrow_stride<H>();
    constexpr int N = 2 * H;
    extern __shared__ double2 sm[];
    double2* ex = sm;
    double2* twH = ex + NR * LS;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int j = tid % Q, t = tid / Q;
    const int64_t R = R0 + t;
    const bool valid = R < nrows;
    const int64_t i1 = valid ? R / d2 : 0, i2 = valid ? R - i1 * d2 : 0;
    for (int i = tid; i < H; i += nth) twH[i] = twg[2 * i];
    double2* ln = ex + t * LS;
    const double2* src = in + i1 * in_s1 + i2 * in_rs;
    // all of this thread's loads in flight before the LDS stores, the
    // pre-twiddles exp(-2 pi i k / N) of its k = j + Q r with them (LDS, not
    // registers, limits this kernel's occupancy)
    double2 vin[P + 1], twr[P];
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        const int k = j + Q * i;
        vin[i] = (valid && k <= H) ? src[k] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int r = 0; r < P; ++r) twr[r] = twg[j + Q * r];
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        const int k = j + Q * i;
        if (k <= H) {
            double2 v = vin[i];
            if (k == 0 || k == H) v.y = 0.0;
            ln[k] = v;
        }
    }
    __syncthreads();
    double2 a[P];
#pragma unroll
    for (int r = 0; r < P; ++r) {
        const int k = j + Q * r;
        const double2 xk = ln[k];
        const double2 xc = cconj(ln[H - k]);
        const double2 d = cmul(cconj(twr[r]), csub(xk, xc));
        a[r] = make_double2(xk.x + xc.x - d.y, xk.y + xc.y + d.x);
    }
    __syncthreads();
    pow2_fft<H, true>(a, j, ln, twH);
    if (valid)
        c2r_store<H, ADD>(a, j, out + i1 * out_s1 + i2 * out_rs, ADD ? add + (R % add_rows) * N : nullptr,
                          scale);
}

template <int H, bool ADD = false>
__global__ void __launch_bounds__(TL * Pow2<H>::Q)
    c2r_contig_kernel(const double2* __restrict__ in, double* __restrict__ out, int64_t in_rs,
                      int64_t out_rs, int64_t d2, int64_t in_s1, int64_t out_s1, int64_t nrows,
                      double scale, const double2* __restrict__ twg /* 2H entries */,
                      const double* __restrict__ add, int64_t add_rows)
{
    c2r_contig_body<H, ADD>(in, out, in_rs, out_rs, d2, in_s1, out_s1, nrows, scale, twg,
                            (int64_t)blockIdx.x * TL, add, add_rows);
}

// ----------------------------------------------------------------- host dispatch
static StridedLines strided_lines(const Geom& gi, const Geom& go, const void* in, void* out,
                                  double scale)
{
    StridedLines g;
    g.in = in;
    g.out = out;
    g.in_sa = gi.sa;
    g.in_s1 = gi.s1;
    g.in_s2 = gi.s2;
    g.out_sa = go.sa;
    g.out_s1 = go.s1;
    g.out_s2 = go.s2;
    g.d2 = gi.d2;
    g.nlines = gi.d1 * gi.d2;
    g.scale = scale;
    return g;
}

template <int N, bool INV, typename C2 = double2>
static int launch_c2c_strided(const Geom& gi, const Geom& go, const void* in, void* out,
                              double scale, hipStream_t s)
{
    using Pl = Pow2<N>;
    const double2* tw = fft_table(N);
    SFA_CHECK(tw, SFA_ERR_HIP, "twiddle table allocation failed");
    const int64_t nl = gi.d1 * gi.d2;
    const size_t lds = sizeof(C2) * ((size_t)N + (size_t)TL * Pl::LS);
    auto k = c2c_strided_kernel<N, INV, C2>;
    if (lds > 64 * 1024)
        hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k, dim3((unsigned)((nl + TL - 1) / TL)), dim3(TL * Pl::Q), lds, s,
                       strided_lines(gi, go, in, out, scale), tw);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

template <int N, bool INV>
static int launch_c2c_strided_dual(const Geom& g64, void* W64, const Geom& g32, void* W32,
                                   double scale, hipStream_t s, const double* stats, StatOut out)
{
    using Pl = Pow2<N>;
    const double2* tw = fft_table(N);
    SFA_CHECK(tw, SFA_ERR_HIP, "twiddle table allocation failed");
    const int64_t nb64 = (g64.d1 * g64.d2 + TL - 1) / TL, nb32 = (g32.d1 * g32.d2 + TL - 1) / TL;
    SFA_CHECK(nb64 + nb32 < (1ll << 31) - 1, SFA_ERR_UNSUPPORTED, "too many lines");
    const size_t lds = sizeof(double2) * ((size_t)N + (size_t)TL * Pl::LS);  // the larger body
    auto k = c2c_strided_dual_kernel<N, INV>;
    if (lds > 64 * 1024)
        hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const int64_t nblk = nb64 + nb32 + (out.flag != nullptr ? 1 : 0);
    hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(TL * Pl::Q), lds, s,
                       strided_lines(g64, g64, W64, W64, scale), strided_lines(g32, g32, W32, W32, scale),
                       nb64, tw, stats, out);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// in-place y C2C of the mixed workspace's complex128 and complex64 planes in
// one launch; returns 1 if handled (power-of-two Ny up to 256).  At 256^3 it
// takes 38.5 -> 34.2 us per pass: the 272 complex128 blocks alone leave a
// tail on 256 CUs.  Not at Ny = 512, where the complex128 body's 256 VGPRs
// would hold the complex64 blocks to one wave per SIMD (217 -> 261 us).
// out.flag != NULL: the launch also delivers the statistics in stats.
int fast_c2c_mixed(const Geom& g64, void* W64, const Geom& g32, void* W32, int sign, double scale,
                   hipStream_t s, int* rc, const double* stats, StatOut out)
{
    if (!(g64.s2 == 1 && g32.s2 == 1 && g64.sa != 1 && g32.sa != 1 && g64.n == g32.n)) return 0;
    const bool inv = sign > 0;
#define SFA_C2CM(NN)                                                                      \
    case NN:                                                                              \
        *rc = inv ? launch_c2c_strided_dual<NN, true>(g64, W64, g32, W32, scale, s, stats, out)   \
                  : launch_c2c_strided_dual<NN, false>(g64, W64, g32, W32, scale, s, stats, out); \
        return 1;
    switch (g64.n) {
        SFA_C2CM(16) SFA_C2CM(32) SFA_C2CM(64) SFA_C2CM(128) SFA_C2CM(256)
        default: return 0;
    }
#undef SFA_C2CM
}

// returns 1 if handled (rc in *rc), 0 if the shape needs the generic kernel
int fast_c2c(const Geom& gi, const Geom& go, const void* in, void* out, int sign, double scale,
             hipStream_t s, int* rc)
{
    const int64_t n = gi.n;
    // strided axis with a contiguous neighbour dim (lines run along s2 == 1)
    if (!(gi.s2 == 1 && go.s2 == 1 && gi.sa != 1 && go.sa != 1)) return 0;
    const bool inv = sign > 0;
#define SFA_C2C(NN)                                                                       \
    case NN:                                                                              \
        *rc = inv ? launch_c2c_strided<NN, true>(gi, go, in, out, scale, s)               \
                  : launch_c2c_strided<NN, false>(gi, go, in, out, scale, s);             \
        return 1;
    switch (n) {
        SFA_C2C(16) SFA_C2C(32) SFA_C2C(64) SFA_C2C(128) SFA_C2C(256) SFA_C2C(512)
        default: return 0;
    }
#undef SFA_C2C
}

// complex64 strided C2C (geometry as fast_c2c); returns 1 if handled
int fast_c2c_f32(const Geom& gi, const Geom& go, const void* in, void* out, int sign, double scale,
                 hipStream_t s, int* rc)
{
    if (!(gi.s2 == 1 && go.s2 == 1 && gi.sa != 1 && go.sa != 1)) return 0;
    const bool inv = sign > 0;
#define SFA_C2CF(NN)                                                                      \
    case NN:                                                                              \
        *rc = inv ? launch_c2c_strided<NN, true, float2>(gi, go, in, out, scale, s)       \
                  : launch_c2c_strided<NN, false, float2>(gi, go, in, out, scale, s);     \
        return 1;
    switch (gi.n) {
        SFA_C2CF(16) SFA_C2CF(32) SFA_C2CF(64) SFA_C2CF(128) SFA_C2CF(256) SFA_C2CF(512)
        default: return 0;
    }
#undef SFA_C2CF
}

template <int H, bool R2C>
static int launch_rc(const Geom& gi, const Geom& go, const void* in, void* out, double scale,
                     hipStream_t s, const double* add, int64_t add_rows)
{
    using Pl = Pow2<H>;
    const double2* tw = fft_table(2 * H);
    SFA_CHECK(tw, SFA_ERR_HIP, "twiddle table allocation failed");
    const int64_t nr = gi.d1 * gi.d2;
    const size_t lds = sizeof(double2) * ((size_t)TL * zrow_stride<H>() + H);
    dim3 grid((unsigned)((nr + TL - 1) / TL)), block(TL * Pl::Q);
    if (R2C) {
        auto k = r2c_contig_kernel<H>;
        if (lds > 64 * 1024)
            hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k, grid, block, lds, s, (const double*)in, (double2*)out, gi.s2, go.s2,
                           gi.d2, gi.s1, go.s1, nr, scale, tw);
    } else {
        auto k = add ? c2r_contig_kernel<H, true> : c2r_contig_kernel<H, false>;
        if (lds > 64 * 1024)
            hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k, grid, block, lds, s, (const double2*)in, (double*)out, gi.s2, go.s2,
                           gi.d2, gi.s1, go.s1, nr, scale, tw, add, add_rows);
    }
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// real length n along a contiguous axis (sa == 1 on both sides).  add (C2R
// only, may be NULL): add_rows contiguous rows of n doubles, row R % add_rows
// added to output row R; needs contiguous output rows.
int fast_rc(bool r2c, const Geom& gi, const Geom& go, const void* in, void* out, double scale,
            hipStream_t s, int* rc, const double* add, int64_t add_rows)
{
    if (!(gi.sa == 1 && go.sa == 1)) return 0;
    if (add && (r2c || add_rows < 1 || go.s2 != gi.n || go.s1 != go.d2 * gi.n ||
                (reinterpret_cast<uintptr_t>(add) & 15)))
        return 0;
    // rows must not overlap the packed complex loads: real rows 16-byte aligned
    const int64_t H = gi.n / 2;
    const int64_t real_rs = r2c ? gi.s2 : go.s2;
    const int64_t real_s1 = r2c ? gi.s1 : go.s1;
    if ((real_rs & 1) || (real_s1 & 1)) return 0;
#define SFA_RC(HH)                                                                            \
    case HH:                                                                                  \
        *rc = r2c ? launch_rc<HH, true>(gi, go, in, out, scale, s, nullptr, 1)                \
                  : launch_rc<HH, false>(gi, go, in, out, scale, s, add, add_rows);           \
        return 1;
    switch (H) {
        SFA_RC(16) SFA_RC(32) SFA_RC(64) SFA_RC(128) SFA_RC(256)
        default: return 0;
    }
#undef SFA_RC
}

}  // namespace sfa

namespace sfa {

// x-pass LDS line layout: a line's natural-index element k lives at
// k + k / PADW.  The solve's 16 lanes read rows 2E apart; without the pad
// element every 2E rows they would all hit the same LDS banks.
template <int N>
struct XLay {
    static constexpr int E = ((N - 2 + 1) / 2 + 15) / 16;  // rows per lane of a 16-lane chain
    static constexpr int PADW = 2 * E;
    static constexpr int NP = N + N / PADW;                // padded line length
    static constexpr int LA = (Pow2<N>::LS > NP + 1) ? Pow2<N>::LS : NP + 1;  // line area stride
    __host__ __device__ static constexpr int pd(int k) { return k + k / PADW; }
    // dynamic LDS bytes for storage type C2
    __host__ __device__ static constexpr size_t lds(size_t c2) { return c2 * (2 * N + TL * LA + TL) + 8 * NP; }
};

struct XPass2Args {
    void* W;         // lines of C2 values; element (x, L) of array b at W[b * bstride + x * sa + L]
    int64_t sa;      // x stride (= Ny * Kp)
    int64_t nl;      // lines of all arrays: nb * nlb
    int64_t nlb;     // lines per array (= Ny * Kp)
    int64_t bstride; // elements between consecutive arrays (= Nx * sa)
    int Kp;          // kz planes held by this array (line pitch per ky)
    int kz0;         // first kz of this array
    int Nzh, Ny;
    int neumann;
    double alpha;
    const double* K2;
    XPlan pl;
    const double2* tw;   // exp(-2 pi i m/N)
    const double2* tq;   // exp(-i pi k/(2N))
    double* stats;       // STATS kernels: band energies of the solution's coefficients
};

// Statistics of the guarded mixed form: lane jb == 0 of mapping-B line tb of
// block bid adds w_kz e to accumulator first + band(kz) (slot bid % kStatSlots).
template <int N>
__device__ __forceinline__ void xpass_stat_add(const XPass2Args& a, int64_t bid, int tb, int jb, int first,
                                               double e)
{
    const uint32_t Ls = (uint32_t)bid * TL + tb;  // lines < 2^31 (checked at launch)
    if (jb != 0 || Ls >= (uint32_t)a.nl) return;
    const uint32_t Lr = Ls % (uint32_t)a.nlb;
    const uint32_t ky = Lr / (uint32_t)a.Kp, kz = a.kz0 + (Lr - ky * (uint32_t)a.Kp);
    if (kz >= a.Nzh) return;  // a pitch-padding line
    const double wz = (kz == 0 || kz == a.Nzh - 1) ? 1.0 : 2.0;
    atomicAdd(a.stats + stat_index(first + stat_band((int)kz), (int)(bid % kStatSlots)), wz * e);
}

// C2 = double2: FP64 everywhere.  C2 = float2: complex64 storage and DCT/FFT
// arithmetic, the Helmholtz solve still FP64 (values widened on entry).
// STATS: also accumulate, per band of 16 kz planes (w_kz = 2 for 0 < kz < Nz/2,
// else 1; slot bid % kStatSlots), the energy of the line's input,
// w_kz sum_x |F_x|^2, into accumulator band of a.stats (from the forward FFT
// by Parseval, sum |V_k|^2 = N sum |F_x|^2), and that of the solution's
// Chebyshev coefficients t (after to_ortho), w_kz (|t_0|^2 + 1/2 sum_{k>=1}
// |t_k|^2), into accumulator kStatBands + band.
template <int N, typename C2, bool STATS, bool NM>
__device__ __forceinline__ void xpass2_body(const XPass2Args& a, int64_t bid)
{
    using Pl = Pow2<N>;
    using R = typename CT<C2>::R;
    constexpr int P = Pl::P, Q = Pl::Q, LS = Pl::LS;
    constexpr int M = N - 2;
    constexpr int K = (M + 1) / 2;  // rows of the even chain
    constexpr int E = XLay<N>::E;
    constexpr int PADW = XLay<N>::PADW, LA = XLay<N>::LA;
    static_assert(E == (K + 15) / 16, "layout rows per lane");
    auto pd = [](int k) { return XLay<N>::pd(k); };
    constexpr int NTH = TL * Q;
    constexpr int ROUNDS = (2 * TL * 16) / NTH;
    static_assert(64 % Q == 0, "a line's threads must lie in one wave");
    extern __shared__ __align__(16) unsigned char sm_raw[];
    // LDS: twiddles, DCT quarter-wave table, stencil, the line exchange areas,
    // t_0 per line.  Tables live in LDS so that no global load is hoisted
    // across the solve (it would hold registers through it).
    C2* tw = reinterpret_cast<C2*>(sm_raw);
    C2* tqs = tw + N;
    double* k2s = reinterpret_cast<double*>(tqs + N);
    C2* ex = reinterpret_cast<C2*>(k2s + XLay<N>::NP);
    C2* t0s = ex + TL * LA;
    C2* Wd = reinterpret_cast<C2*>(a.W);
    const int tid = threadIdx.x;
    // mapping A (global loads/stores): lanes run over the TL lines, so each
    // x row is one contiguous TL-line segment.  Mapping B (everything between
    // the two FFT exchanges): a line's Q threads are adjacent lanes of one
    // wave, so those phases synchronise per wave, not per workgroup.
    // line t of group g in mapping A: element x = 0 at Lo
    // line t of block g in mapping A, element x = 0 at Lo (line indices in 32
    // bits: fewer than 2^31 lines, checked at launch)
    auto line_of = [&](int64_t g, int t, int64_t& Lo) {
        const uint32_t L = (uint32_t)g * TL + t;
        const uint32_t Lb = L / (uint32_t)a.nlb;
        Lo = (int64_t)Lb * a.bstride + (L - Lb * (uint32_t)a.nlb);
        return L < (uint32_t)a.nl;
    };
    // 1. load with the DCT-II even/odd permutation: v[m] = x[2m] (m < N/2), x[2(N-1-m)+1]
    // A line past the end loads line 0 (in bounds) and its results are never
    // stored or counted: the loads stay outside any branch, so the waits after
    // them can count them instead of draining every load.
    auto load = [&](int64_t g, int t, int j, C2 (&w)[P]) {
        int64_t Lo;
        const int64_t Ls = line_of(g, t, Lo) ? Lo : 0;
#pragma unroll
        for (int r = 0; r < P; ++r) {
            const int m = j + Q * r;
            const int x = (m < N / 2) ? 2 * m : 2 * (N - 1 - m) + 1;
            w[r] = Wd[(int64_t)x * a.sa + Ls];
        }
    };
    // The block's loads go out together before any wait: this thread's table
    // entries (clamped indices, no branch), its line's values, the plan lookups.
    constexpr int TPI = (N + NTH - 1) / NTH;
    double2 twr[TPI], tqr[TPI];
    double k2r[TPI];
#pragma unroll
    for (int e = 0; e < TPI; ++e) {
        const int i = min(tid + e * NTH, N - 1);
        twr[e] = a.tw[(i % Q) * (i / Q)];  // k1-major: tw[k1 Q + j] = w^(j k1)
        tqr[e] = a.tq[i];
        k2r[e] = a.K2[min(i, N - 3)];
    }
    const int t = tid % TL, j = tid / TL;
    const int tb = tid / Q, jb = tid % Q;
    C2 v[P];
    load(bid, t, j, v);
#pragma unroll
    for (int e = 0; e < TPI; ++e) {  // (threads past N rewrite entry N - 1 with its own value)
        const int i = min(tid + e * NTH, N - 1);
        tw[i] = from_c64<C2>(twr[e]);
        tqs[i] = from_c64<C2>(tqr[e]);
        k2s[pd(i)] = (i < N - 2) ? k2r[e] : 0.0;
    }
    // plan entries of this block's solve lines (the 16-lane group's chains are
    // lines ct0 .. ct0 + ROUNDS/2 - 1, each with both parities), looked up here
    // so the line-map latency overlaps the loads and FFT; before the solve only
    // the independent svals / restart-state loads remain
    static_assert(ROUNDS % 2 == 0, "chains per group must pair both parities");
    constexpr int NLR = ROUNDS / 2;
    int ug[NLR];
#pragma unroll
    for (int q = 0; q < NLR; ++q) {
        // branch-free, so that no wait for this load is placed here
        const uint32_t Lc = (uint32_t)bid * TL + (tid / 16) * NLR + q;
        const bool inl = Lc < (uint32_t)a.nl;
        const uint32_t Lr = (inl ? Lc : 0u) % (uint32_t)a.nlb;  // line within its array
        const uint32_t ky = Lr / (uint32_t)a.Kp, kz = a.kz0 + (Lr - ky * (uint32_t)a.Kp);
        const bool ok = inl && kz < (uint32_t)a.Nzh;  // else past the end or a pitch-padding line
        const int uq = a.pl.linemap[ok ? ky * a.Nzh + kz : 0];
        ug[q] = ok ? uq : 0;
    }
    __syncthreads();  // tables
    pow2_fft_p1t<N, false>(v, j, ex + t * LA, tw);
    __syncthreads();  // exchange, and the switch to mapping B
    // the solve rounds' shifts and restart states, loaded now (the line map has
    // arrived with the line values) so they arrive during the FFT's second half
    // and the DCT post-processing
    double beR[ROUNDS];
    double4 b4R[ROUNDS];
#pragma unroll
    for (int rho = 0; rho < ROUNDS; ++rho) {
        const int u = ug[rho >> 1];
        beR[rho] = a.alpha + a.pl.svals[u];
        b4R[rho] = a.pl.bs[((int64_t)2 * u + (rho & 1)) * 16 + (tid & 15)];
    }
    C2* ln = ex + tb * LA;
    pow2_fft_p2<N, false>(v, jb, ln);
    if (STATS) {  // the line's input energy, added at once (nothing held through the solve)
        double ef = 0.0;
#pragma unroll
        for (int r = 0; r < P; ++r) {
            const double2 c = to_c64(v[r]);
            ef += c.x * c.x + c.y * c.y;
        }
#pragma unroll
        for (int off = Q / 2; off >= 1; off >>= 1) ef += __shfl_xor(ef, off);
        xpass_stat_add<N>(a, bid, tb, jb, 0, ef * (1.0 / N));
    }
    wave_sync();
#pragma unroll
    for (int h = 0; h < P / Q; ++h)
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) ln[pd(pow2_out_index<N>(jb, h, k2))] = v[h * Q + k2];
    wave_sync();
    // 2a. DCT-II post-processing X_k = sc (Re(w (V_k + conj V_{N-k})) + i Im(w (V_k - conj V_{N-k})))
    const R sc = (R)(3.14159265358979323846 / (2.0 * N));
#pragma unroll
    for (int h = 0; h < P / Q; ++h)
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) {
            const int k = pow2_out_index<N>(jb, h, k2);
            const C2 Pv = v[h * Q + k2];
            const C2 Qv = cconj(ln[pd(k == 0 ? 0 : N - k)]);
            const C2 w = tqs[k];
            const C2 s1 = cmul(w, cadd(Pv, Qv));
            const C2 s2 = cmul(w, csub(Pv, Qv));
            v[h * Q + k2] = CT<C2>::mk(s1.x * sc, s2.y * sc);
        }
    wave_sync();
#pragma unroll
    for (int h = 0; h < P / Q; ++h)
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) ln[pd(pow2_out_index<N>(jb, h, k2))] = v[h * Q + k2];
    wave_sync();
    // 2b. stencil (FP64) + Helmholtz solve, 16 lanes per (line, parity) chain.
    // Group g = tid / 16 takes chains g * ROUNDS + rho: the parity alternates
    // with rho, so the row operators (functions of the row index) cannot be
    // hoisted out of the loop and held in registers across it.  Its lines are
    // mapping-B lines of its own wave.
#pragma unroll 1
    for (int rho = 0; rho < ROUNDS; ++rho) {
        const int c = (tid / 16) * ROUNDS + rho;
        const int ct = c >> 1, p = c & 1, lane = tid & 15;
        C2* X = ex + ct * LA;
        double be = beR[0];
        double4 b4 = b4R[0];
#pragma unroll
        for (int q = 1; q < ROUNDS; ++q)
            if (rho == q) {
                be = beR[q];
                b4 = b4R[q];
            }
        double2 f[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = 2 * (lane * E + e) + p;
            if (i < M) {
                const double2 x0 = to_c64(X[pd(i)]), x2 = to_c64(X[pd(i + 2)]);
                const double kk = k2s[pd(i)];
                f[e] = make_double2(x0.x + kk * x2.x, x0.y + kk * x2.y);
            } else {
                f[e] = make_double2(0.0, 0.0);
            }
        }
        helm_chain_solve_bs<E, 16, C2, PADW>(f, X, p, M, be, NM, b4, lane);
        // boundary slots M, M+1 (u carries no value there).  The odd chain
        // runs after the even one and its stencil reads X[M+1], so it clears
        // them once its own reads (same group, earlier in program order) are done.
        if (p == 1 && lane == 0) {
            X[pd(M)] = cmk<C2>(0.0, 0.0);
            X[pd(M + 1)] = cmk<C2>(0.0, 0.0);
        }
        wave_sync();
    }
    // 3a. to_ortho t_k = u_k (k < N-2) + K2_{k-2} u_{k-2}, for this thread's k = jb + Q r
    const C2 t0 = ln[pd(0)];
    if (jb == 0) t0s[tb] = t0;
#pragma unroll
    for (int r = 0; r < P; ++r) {
        const int k = jb + Q * r;
        C2 cc = (k < N - 2) ? ln[pd(k)] : cmk<C2>(0.0, 0.0);
        if (k >= 2) {
            const R kk = (R)k2s[pd(k - 2)];
            const C2 uu = ln[pd(k - 2)];
            cc.x += kk * uu.x;
            cc.y += kk * uu.y;
        }
        v[r] = cc;
    }
    if (STATS) {
        double el = 0.0;
#pragma unroll
        for (int r = 0; r < P; ++r) {
            const double2 c = to_c64(v[r]);
            el += ((jb + Q * r) == 0 ? 1.0 : 0.5) * (c.x * c.x + c.y * c.y);
        }
#pragma unroll
        for (int off = Q / 2; off >= 1; off >>= 1) el += __shfl_xor(el, off);
        xpass_stat_add<N>(a, bid, tb, jb, kStatBands, el);
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < P; ++r) ln[pd(jb + Q * r)] = v[r];
    wave_sync();
    // 3b. DCT-III pre-processing V_k = conj(w_k) ((a_k + b_{N-k}) + i (b_k - a_{N-k}))
#pragma unroll
    for (int r = 0; r < P; ++r) {
        const int k = jb + Q * r;
        const C2 cn = (k == 0) ? cmk<C2>(0.0, 0.0) : ln[pd(N - k)];
        const C2 z = CT<C2>::mk(v[r].x + cn.y, v[r].y - cn.x);
        v[r] = cmul(cconj(tqs[k]), z);
    }
    wave_sync();
    pow2_fft_p1t<N, true>(v, jb, ln, tw);
    __syncthreads();  // exchange, and the switch back to mapping A
    pow2_fft_p2<N, true>(v, j, ex + t * LA);
    const C2 t0a = t0s[t];
    // 4. store 0.5 y + t_0/2 with the inverse permutation
    int64_t Lo;
    if (line_of(bid, t, Lo)) {
        const R half = (R)0.5;
#pragma unroll
        for (int h = 0; h < P / Q; ++h)
#pragma unroll
            for (int k2 = 0; k2 < Q; ++k2) {
                const int m = pow2_out_index<N>(j, h, k2);
                const int x = (m < N / 2) ? 2 * m : 2 * (N - 1 - m) + 1;
                const C2 y = v[h * Q + k2];
                Wd[(int64_t)x * a.sa + Lo] = CT<C2>::mk(half * y.x + half * t0a.x, half * y.y + half * t0a.y);
            }
    }
}

// workgroups per CU the x-pass is compiled for: at N = 512 one block holds
// 16 lines of 512 complex values in LDS (~156 KB in FP64)
template <int N>
constexpr int xpass_minb() { return N >= 512 ? 1 : 2; }

// FP64 x-pass.  NM: Neumann (Chebyshev SN) rows, else Dirichlet (SD); one
// kernel each so that the Dirichlet kernels carry no Neumann code
template <int N, bool NM>
__global__ void __launch_bounds__(TL * Pow2<N>::Q, xpass_minb<N>()) xpass2_kernel(XPass2Args a)
{
    xpass2_body<N, double2, false, NM>(a, blockIdx.x);
}

// Mixed precision: the complex128 planes' blocks [0, nb64) and the complex64
// planes' blocks after them in one grid, so the few FP64 blocks run beside the
// FP32 ones instead of in a launch of their own.
// Block order of the complex64 part of the dual x-pass: where a ky row is a
// whole number of blocks (bpk) and the launch holds one array, dispatch slot
// d runs the block of ky and slot d + 8 the block of Ny - ky with the same kz
// range, 8 such pairs per 16 slots.  Workgroups go to the 8 XCDs round-robin,
// so the two land on one XCD at about the same time, and their lines, which
// share beta = alpha + ky^2 + kz^2, read one set of plan restart states
// through that XCD's L2 (mixed x-pass at 256^3 76.3 -> 75.5 us, the same
// bits).  ky = 0 and Ny/2, their own partners, come last.  A permutation of
// [0, Ny bpk) (checked exhaustively on the host for Ny <= 512, bpk <= 16).
__device__ __forceinline__ int64_t xpass_pair_order(const XPass2Args& a, int64_t d)
{
    if (a.Kp % TL != 0 || a.nl != a.nlb) return d;
    const int64_t bpk = a.Kp / TL, Ny = a.Ny;
    const int64_t E = (Ny / 2 - 1) * bpk;  // blocks of ky in [1, Ny/2)
    if (d >= 2 * E) {
        const int64_t r = d - 2 * E;
        return r < bpk ? r : (Ny / 2) * bpk + (r - bpk);
    }
    const int64_t G = E / 8, rem = E % 8;
    int64_t e;
    bool upper;
    if (d < 16 * G) {
        e = 8 * (d / 16) + d % 8;
        upper = (d % 16) >= 8;
    } else {  // the last, partial group: rem blocks, then their partners
        const int64_t r2 = d - 16 * G;
        upper = r2 >= rem;
        e = 8 * G + (upper ? r2 - rem : r2);
    }
    const int64_t ky = 1 + e / bpk, c = e % bpk;
    return (upper ? Ny - ky : ky) * bpk + c;
}

template <int N, bool STATS, bool NM>
__global__ void __launch_bounds__(TL * Pow2<N>::Q, xpass_minb<N>()) xpass2_dual_kernel(XPass2Args a64, XPass2Args a32,
                                                                        int64_t nb64)
{
    const int64_t b = blockIdx.x;
    if (b < nb64)
        xpass2_body<N, double2, STATS, NM>(a64, b);
    else
        xpass2_body<N, float2, STATS, NM>(a32, xpass_pair_order(a32, b - nb64));
}

template <int N>
static int launch_xpass2(XPass2Args& a, hipStream_t s)
{
    using Pl = Pow2<N>;
    const size_t lds = XLay<N>::lds(sizeof(double2));
    auto k = a.neumann ? xpass2_kernel<N, true> : xpass2_kernel<N, false>;
    if (lds > 64 * 1024)
        hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k, dim3((unsigned)((a.nl + TL - 1) / TL)), dim3(TL * Pl::Q), lds, s, a);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

template <int N>
static int launch_xpass2_dual(XPass2Args& a64, XPass2Args& a32, hipStream_t s)
{
    using Pl = Pow2<N>;
    const size_t lds = XLay<N>::lds(sizeof(double2));
    auto k = a64.neumann ? (a64.stats ? xpass2_dual_kernel<N, true, true> : xpass2_dual_kernel<N, false, true>)
                         : (a64.stats ? xpass2_dual_kernel<N, true, false> : xpass2_dual_kernel<N, false, false>);
    if (lds > 64 * 1024)
        hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const int64_t nb64 = (a64.nl + TL - 1) / TL, nb32 = (a32.nl + TL - 1) / TL;
    hipLaunchKernelGGL(k, dim3((unsigned)(nb64 + nb32)), dim3(TL * Pl::Q), lds, s, a64, a32, nb64);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// Fused FP64 x-pass on the kz planes [kz0, kz0 + Kp) held (line pitch Kp)
// in W.  Returns 1 if handled.
static bool xpass_args(void* W, int Nx, int Ny, int Nzh, int Kp, int kz0, double alpha, int neumann,
                       const double* K2, const XPlan& pl, int64_t nb, XPass2Args* a)
{
    a->W = W;
    a->sa = (int64_t)Ny * Kp;
    a->nlb = (int64_t)Ny * Kp;
    a->nl = nb * a->nlb;
    a->bstride = (int64_t)Nx * a->sa;
    a->Kp = Kp;
    a->kz0 = kz0;
    a->Nzh = Nzh;
    a->Ny = Ny;
    a->neumann = neumann;
    a->alpha = alpha;
    a->K2 = K2;
    a->pl = pl;
    a->tw = fft_table(Nx);
    a->tq = quarter_table(Nx);
    a->stats = nullptr;
    return a->tw && a->tq;
}

static int fast_xpass_sub(void* W, int Nx, int Ny, int Nzh, int Kp, int kz0, double alpha, int neumann,
                          const double* K2, const XPlan& pl, int64_t nb, hipStream_t s, int* rc)
{
    if (!(Nx == 64 || Nx == 128 || Nx == 256 || Nx == 512) || pl.linemap == nullptr || pl.bs == nullptr)
        return 0;
    XPass2Args a;
    if (!xpass_args(W, Nx, Ny, Nzh, Kp, kz0, alpha, neumann, K2, pl, nb, &a)) {
        *rc = fail(SFA_ERR_HIP, "table allocation failed");
        return 1;
    }
    if (a.nl >= (int64_t(1) << 31)) {
        *rc = fail(SFA_ERR_UNSUPPORTED, "x-pass: 2^31 or more lines");
        return 1;
    }
    switch (Nx) {
        case 64: *rc = launch_xpass2<64>(a, s); break;
        case 128: *rc = launch_xpass2<128>(a, s); break;
        case 256: *rc = launch_xpass2<256>(a, s); break;
        default: *rc = launch_xpass2<512>(a, s); break;
    }
    return 1;
}

// Mixed precision x-pass: planes [0, Kh) complex128 in W64 (pitch Kh) and
// [Kh, Nzh) complex64 in W32 (pitch Nzh - Kh), one launch; nb arrays back to
// back in each of W64 and W32.
int fast_xpass_mixed(void* W64, void* W32, int Nx, int Ny, int Nzh, int Kh, double alpha,
                     int neumann, const double* K2, const XPlan& pl, int64_t nb, hipStream_t s, int* rc,
                     double* stats)
{
    if (!(Nx == 64 || Nx == 128 || Nx == 256 || Nx == 512) || pl.linemap == nullptr || pl.bs == nullptr)
        return 0;
    XPass2Args a64, a32;
    if (!xpass_args(W64, Nx, Ny, Nzh, Kh, 0, alpha, neumann, K2, pl, nb, &a64) ||
        !xpass_args(W32, Nx, Ny, Nzh, Nzh - Kh, Kh, alpha, neumann, K2, pl, nb, &a32)) {
        *rc = fail(SFA_ERR_HIP, "table allocation failed");
        return 1;
    }
    if (a64.nl + a32.nl >= (int64_t(1) << 31)) {
        *rc = fail(SFA_ERR_UNSUPPORTED, "x-pass: 2^31 or more lines");
        return 1;
    }
    a64.stats = a32.stats = stats;
    switch (Nx) {
        case 64: *rc = launch_xpass2_dual<64>(a64, a32, s); break;
        case 128: *rc = launch_xpass2_dual<128>(a64, a32, s); break;
        case 256: *rc = launch_xpass2_dual<256>(a64, a32, s); break;
        default: *rc = launch_xpass2_dual<512>(a64, a32, s); break;
    }
    return 1;
}

// W: nb arrays (Nx, Ny, Kp) back to back, lines (ky, kz) at pitch Kp >= Nzh
// per ky; the Kp - Nzh padding lines are transformed too (their values are
// never read back).
int fast_xpass(double2* W, int Nx, int Ny, int Nzh, int Kp, double alpha, int neumann,
               const double* K2, const XPlan& pl, int64_t nb, hipStream_t s, int* rc)
{
    return fast_xpass_sub(W, Nx, Ny, Nzh, Kp, 0, alpha, neumann, K2, pl, nb, s, rc);
}

}  // namespace sfa

namespace sfa {

// ----------------------------------------------------------------- kz-split z passes
// R2C of contiguous rows of 2H reals (FP64 arithmetic), output split by kz:
// k < Kh -> o64 (row pitch Kh, complex128), k >= Kh -> o32 (row pitch H+1-Kh, complex64).
// KSMALL: Kh <= 3Q, so only a thread's first three outputs (k = j + Q i,
// i < 3) can go to o64 and the others are stored without a per-element branch.
template <int H, bool KSMALL, int NR = TL>
__device__ __forceinline__ void r2c_split_body(const double* __restrict__ in, double2* __restrict__ o64,
                                               float2* __restrict__ o32, int64_t nrows, int Kh,
                                               double scale, const double2* __restrict__ twg, int64_t R0)
{
    using Pl = Pow2<H>;
    constexpr int P = Pl::P, Q = Pl::Q, LS = zrow_stride<H>();
    constexpr int N = 2 * H;
    extern __shared__ double2 sm[];
    double2* ex = sm;
    double2* twH = ex + NR * LS;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int j = tid % Q, t = tid / Q;
    const int64_t R = R0 + t;
    const bool valid = R < nrows;
    const int Kl = H + 1 - Kh;
    for (int i = tid; i < H; i += nth) twH[i] = twg[2 * i];
    double2 a[P];
    const double* src = in + (valid ? R : 0) * N;
#pragma unroll
    for (int r = 0; r < P; ++r) {
        const int m = j + Q * r;
        a[r] = valid ? *reinterpret_cast<const double2*>(src + 2 * m) : make_double2(0.0, 0.0);
    }
    // post-twiddles of this thread's outputs k = j + Q i, issued early
    double2 twr[P + 1];
#pragma unroll
    for (int i = 0; i <= P; ++i) twr[i] = (j + Q * i <= H) ? twg[j + Q * i] : make_double2(0.0, 0.0);
    __syncthreads();
    double2* ln = ex + t * LS;
    pow2_fft<H, false>(a, j, ln, twH);
#pragma unroll
    for (int h = 0; h < P / Q; ++h)
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) ln[pow2_out_index<H>(j, h, k2)] = a[h * Q + k2];
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int i = 0; i <= P; ++i) {
            const int k = j + Q * i;
            if (k > H) break;
            const double2 zk = ln[k == H ? 0 : k];
            const double2 zc = cconj(ln[k == 0 ? 0 : H - k]);
            const double2 d = cmul(twr[i], csub(zk, zc));
            const double2 X = make_double2(0.5 * (zk.x + zc.x + d.y) * scale, 0.5 * (zk.y + zc.y - d.x) * scale);
            if (KSMALL && i >= 3) o32[R * Kl + (k - Kh)] = from_c64<float2>(X);
            else if (k < Kh) o64[R * Kh + k] = X;
            else o32[R * Kl + (k - Kh)] = from_c64<float2>(X);
        }
    }
}

// zero[0, nzero): cleared on the side (grid-stride), for the accumulators that
// later kernels of the same pass add into, so no launch of its own is needed
template <int H, bool KSMALL>
__global__ void __launch_bounds__(TL * Pow2<H>::Q)
    r2c_split_kernel(const double* __restrict__ in, double2* __restrict__ o64,
                     float2* __restrict__ o32, int64_t nrows, int Kh, double scale,
                     const double2* __restrict__ twg, double* __restrict__ zero, int nzero)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nzero; i += gridDim.x * blockDim.x)
        zero[i] = 0.0;
    r2c_split_body<H, KSMALL>(in, o64, o32, nrows, Kh, scale, twg, (int64_t)blockIdx.x * TL);
}

// C2R (FP64 arithmetic) of rows merged from o64 (k < Kh) and o32 (k >= Kh).
// KSMALL: Kh <= 3Q, so only this thread's first three elements (k = j + Q i,
// i < 3) can lie in the complex128 planes; every other load is a plain
// complex64 load, and all loads are issued before any conversion (a
// per-element select of a converted complex64 load waited on each load in turn).
// ADD: row R also gets row R % add_rows of add.
template <int H, bool KSMALL, bool ADD = false, int NR = TL>
__device__ __forceinline__ void c2r_merge_body(const double2* __restrict__ i64, const float2* __restrict__ i32,
                                               double* __restrict__ out, int64_t nrows, int Kh,
                                               double scale, const double2* __restrict__ twg, int64_t R0,
                                               const double* __restrict__ add = nullptr, int64_t add_rows = 1)
{
    using Pl = Pow2<H>;
    constexpr int P = Pl::P, Q = Pl::Q, LS = zrow_stride<H>();
    constexpr int N = 2 * H;
    extern __shared__ double2 sm[];
    double2* ex = sm;
    double2* twH = ex + NR * LS;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int j = tid % Q, t = tid / Q;
    const int64_t R = R0 + t;
    const bool valid = R < nrows;
    const int Kl = H + 1 - Kh;
    for (int i = tid; i < H; i += nth) twH[i] = twg[2 * i];
    double2* ln = ex + t * LS;
    double2 vin[P + 1], twr[P];
    if (KSMALL) {
        // every load unconditional from a clamped in-range address (row 0 for
        // rows past the end, k clamped into its sub-array), validity applied
        // afterwards by selects: no branch separates the loads
        constexpr int NLO = 3;
        float2 v32[P + 1];
        double2 lo[NLO];
        const int64_t Rc = valid ? R : 0;
        const float2* r32 = i32 + Rc * Kl;  // element k >= Kh at r32[k - Kh]
#pragma unroll
        for (int i = 0; i <= P; ++i) {
            const int k = j + Q * i;
            const int kc = k < Kh ? Kh : (k > H ? H : k);
            v32[i] = r32[kc - Kh];
        }
#pragma unroll
        for (int i = 0; i < NLO; ++i) {
            const int k = j + Q * i;
            lo[i] = i64[Rc * Kh + (k < Kh ? k : 0)];
        }
#pragma unroll
        for (int i = 0; i <= P; ++i) {
            const int k = j + Q * i;
            const double2 v = (i < NLO && k < Kh) ? lo[i < NLO ? i : 0] : to_c64(v32[i]);
            vin[i] = (valid && k <= H) ? v : make_double2(0.0, 0.0);
        }
    } else {
#pragma unroll
        for (int i = 0; i <= P; ++i) {
            const int k = j + Q * i;
            double2 v = make_double2(0.0, 0.0);
            if (valid && k <= H) v = (k < Kh) ? i64[R * Kh + k] : to_c64(i32[R * Kl + (k - Kh)]);
            vin[i] = v;
        }
    }
#pragma unroll
    for (int r = 0; r < P; ++r) twr[r] = twg[j + Q * r];  // pre-twiddles, issued with the loads
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        const int k = j + Q * i;
        if (k <= H) {
            double2 v = vin[i];
            if (k == 0 || k == H) v.y = 0.0;
            ln[k] = v;
        }
    }
    __syncthreads();
    double2 a[P];
#pragma unroll
    for (int r = 0; r < P; ++r) {
        const int k = j + Q * r;
        const double2 xk = ln[k];
        const double2 xc = cconj(ln[H - k]);
        const double2 d = cmul(cconj(twr[r]), csub(xk, xc));
        a[r] = make_double2(xk.x + xc.x - d.y, xk.y + xc.y + d.x);
    }
    __syncthreads();
    pow2_fft<H, true>(a, j, ln, twH);
    if (valid)
        c2r_store<H, ADD>(a, j, out + R * N, ADD ? add + (R % add_rows) * N : nullptr, scale);
}

template <int H, bool KSMALL, bool ADD = false>
__global__ void __launch_bounds__(TL * Pow2<H>::Q)
    c2r_merge_kernel(const double2* __restrict__ i64, const float2* __restrict__ i32,
                     double* __restrict__ out, int64_t nrows, int Kh, double scale,
                     const double2* __restrict__ twg, const double* __restrict__ add, int64_t add_rows)
{
    c2r_merge_body<H, KSMALL, ADD>(i64, i32, out, nrows, Kh, scale, twg, (int64_t)blockIdx.x * TL, add,
                                   add_rows);
}

template <int H, bool SPLIT>
static int launch_split(const void* in, void* o64, void* o32, int64_t nrows, int Kh, double scale,
                        hipStream_t s, double* zero, int nzero, const double* add, int64_t add_rows)
{
    using Pl = Pow2<H>;
    const double2* tw = fft_table(2 * H);
    SFA_CHECK(tw, SFA_ERR_HIP, "twiddle table allocation failed");
    const size_t lds = sizeof(double2) * ((size_t)TL * zrow_stride<H>() + H);
    dim3 grid((unsigned)((nrows + TL - 1) / TL)), block(TL * Pl::Q);
    if (SPLIT) {
        auto k = Kh <= 3 * Pl::Q ? r2c_split_kernel<H, true> : r2c_split_kernel<H, false>;
        if (lds > 64 * 1024)
            hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k, grid, block, lds, s, (const double*)in, (double2*)o64, (float2*)o32,
                           nrows, Kh, scale, tw, zero, nzero);
    } else {
        auto k = Kh <= 3 * Pl::Q ? (add ? c2r_merge_kernel<H, true, true> : c2r_merge_kernel<H, true, false>)
                                 : (add ? c2r_merge_kernel<H, false, true> : c2r_merge_kernel<H, false, false>);
        if (lds > 64 * 1024)
            hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k, grid, block, lds, s, (const double2*)o64, (const float2*)o32,
                           (double*)in, nrows, Kh, scale, tw, add, add_rows);
    }
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// split (r2c=true: in -> o64/o32) or merge (r2c=false: o64/o32 -> in) of
// contiguous rows; the split also clears zero[0, nzero)
int split_rc(bool r2c, int H, const void* in, void* o64, void* o32, int64_t nrows, int Kh,
             double scale, hipStream_t s, int* rc, double* zero, int nzero, const double* add,
             int64_t add_rows)
{
    if (add && (r2c || add_rows < 1 || (reinterpret_cast<uintptr_t>(add) & 15))) return 0;
#define SFA_SP(HH)                                                                                       \
    case HH:                                                                                             \
        *rc = r2c ? launch_split<HH, true>(in, o64, o32, nrows, Kh, scale, s, zero, nzero, nullptr, 1)  \
                  : launch_split<HH, false>(in, o64, o32, nrows, Kh, scale, s, nullptr, 0, add, add_rows); \
        return 1;
    switch (H) {
        SFA_SP(16) SFA_SP(32) SFA_SP(64) SFA_SP(128) SFA_SP(256)
        default: return 0;
    }
#undef SFA_SP
}

}  // namespace sfa
