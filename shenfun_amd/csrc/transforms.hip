// Per-axis transforms on strided 1-3D arrays: C2C/R2C/C2R FFT, Chebyshev
// DCT-II/III (with the composite stencil / to_ortho fused), dense line
// matmul (Legendre transforms), Legendre matrix builder.
//
// Each block stages T lines of the transformed axis in LDS (two ping-pong
// buffers of T x ldn complex), runs the mixed-radix Stockham FFT there and
// writes back.  Loads/stores walk the tile with the stride-1 index fastest
// (lines when the axis is strided, line elements when the axis is contiguous)
// so global accesses coalesce.
#include <vector>

#include "fft_device.h"

namespace sfa {

// OP_DCT1F / OP_DCT1B: Chebyshev-Gauss-Lobatto scalar product / backward via
// DCT-I (REDFT00) of length n, computed as the complex FFT of the line's even
// extension (length 2(n-1)): the sine terms cancel, so the first n outputs
// are the DCT-I of the real and imaginary parts at once.
// OP_R2CF / OP_C2RF: odd-length R2C / C2R as a full-length complex FFT (real
// input widened; Hermitian spectrum completed on load), the even lengths pack
// pairs of reals instead (OP_R2C / OP_C2R)
enum LineOp {
    OP_C2C = 0, OP_R2C = 1, OP_C2R = 2, OP_DCT2 = 3, OP_DCT3 = 4, OP_DCT1F = 5, OP_DCT1B = 6,
    OP_R2CF = 7, OP_C2RF = 8
};

struct LineArgs {
    const char* in;
    char* out;
    int64_t in_sa, in_s1, in_s2;
    int64_t out_sa, out_s1, out_s2;
    int64_t d1, d2;
    int n;      // logical length along the axis (real length for R2C/C2R)
    int ldn;    // LDS line stride (complex elements)
    int T;      // lines per block
    int in_cplx, out_cplx;
    int t_fast_in, t_fast_out;
    int inv;    // C2C direction
    double scale;
    const double* K2;
    const double2* tw;   // FFT table of length plan.nfft
    const double2* tw2;  // R2C/C2R: table of length n; DCT: quarter table of length n
    FFTPlan plan;
};

__device__ inline double2 ld_elem(const char* base, int64_t off, int cplx)
{
    if (cplx) return reinterpret_cast<const double2*>(base)[off];
    return make_double2(reinterpret_cast<const double*>(base)[off], 0.0);
}

__device__ inline void st_elem(char* base, int64_t off, int cplx, double2 v)
{
    if (cplx)
        reinterpret_cast<double2*>(base)[off] = v;
    else
        reinterpret_cast<double*>(base)[off] = v.x;
}

// decode a flat tile index into (t, i) with either t or i fastest
__device__ inline void tile_index(int e, int T, int len, int t_fast, int& t, int& i)
{
    if (t_fast) {
        t = e % T;
        i = e / T;
    } else {
        i = e % len;
        t = e / len;
    }
}

__device__ inline int dct_perm(int i, int N) { return (i & 1) ? (N - 1 - (i >> 1)) : (i >> 1); }

template <int OP, bool INV>
__global__ void __launch_bounds__(256) line_kernel(LineArgs a)
{
    extern __shared__ double2 smem[];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int64_t ntile = (a.d2 + a.T - 1) / a.T;
    const int64_t i1 = blockIdx.x / ntile;
    const int64_t t0 = (blockIdx.x - i1 * ntile) * a.T;
    const int T = a.T;
    const int Tv = (int)((a.d2 - t0) < T ? (a.d2 - t0) : T);  // valid lines in this tile
    const int ldn = a.ldn;
    double2* A = smem;
    double2* B = smem + (size_t)T * ldn;
    double* t0s = reinterpret_cast<double*>(smem + (size_t)2 * T * ldn);  // 4*T doubles
    const char* inb = a.in;
    char* outb = a.out;
    const int64_t ibase = i1 * a.in_s1 + t0 * a.in_s2;
    const int64_t obase = i1 * a.out_s1 + t0 * a.out_s2;
    const int n = a.n;
    const int nf = a.plan.nfft;

    // ------------------------------------------------------------ load
    if (OP == OP_C2C || OP == OP_R2CF) {
        for (int e = tid; e < Tv * n; e += nth) {
            int t, i;
            tile_index(e, Tv, n, a.t_fast_in, t, i);
            A[t * ldn + i] = ld_elem(inb, ibase + t * a.in_s2 + i * a.in_sa, a.in_cplx);
        }
    } else if (OP == OP_R2C) {
        const int h = n / 2;
        for (int e = tid; e < Tv * h; e += nth) {
            int t, m;
            tile_index(e, Tv, h, a.t_fast_in, t, m);
            const double* p = reinterpret_cast<const double*>(inb);
            const int64_t o = ibase + t * a.in_s2 + (int64_t)(2 * m) * a.in_sa;
            A[t * ldn + m] = make_double2(p[o], p[o + a.in_sa]);
        }
    } else if (OP == OP_C2RF) {
        // X_k for k <= n/2 from the input, X_{n-k} = conj(X_k) above; Im X_0 dropped
        const int h = n / 2;
        for (int e = tid; e < Tv * n; e += nth) {
            int t, i;
            tile_index(e, Tv, n, a.t_fast_in, t, i);
            const int k = i <= h ? i : n - i;
            double2 v = ld_elem(inb, ibase + t * a.in_s2 + k * a.in_sa, 1);
            if (i == 0) v.y = 0.0;
            A[t * ldn + i] = i <= h ? v : cconj(v);
        }
    } else if (OP == OP_C2R) {
        const int h = n / 2;
        for (int e = tid; e < Tv * (h + 1); e += nth) {
            int t, k;
            tile_index(e, Tv, h + 1, a.t_fast_in, t, k);
            double2 v = ld_elem(inb, ibase + t * a.in_s2 + k * a.in_sa, 1);
            if (k == 0 || k == h) v.y = 0.0;
            B[t * ldn + k] = v;
        }
        __syncthreads();
        for (int e = tid; e < Tv * h; e += nth) {
            const int t = e / h, k = e - t * h;
            const double2 xk = B[t * ldn + k];
            const double2 xc = cconj(B[t * ldn + h - k]);
            const double2 w = cconj(a.tw2[k]);  // exp(+2 pi i k / n)
            const double2 d = cmul(w, csub(xk, xc));
            // Z = (xk + xc) + i * w (xk - xc)
            A[t * ldn + k] = make_double2(xk.x + xc.x - d.y, xk.y + xc.y + d.x);
        }
    } else if (OP == OP_DCT2) {
        for (int e = tid; e < Tv * n; e += nth) {
            int t, i;
            tile_index(e, Tv, n, a.t_fast_in, t, i);
            A[t * ldn + dct_perm(i, n)] = ld_elem(inb, ibase + t * a.in_s2 + i * a.in_sa, a.in_cplx);
        }
    } else if (OP == OP_DCT3) {
        for (int e = tid; e < Tv * n; e += nth) {
            int t, i;
            tile_index(e, Tv, n, a.t_fast_in, t, i);
            B[t * ldn + i] = ld_elem(inb, ibase + t * a.in_s2 + i * a.in_sa, a.in_cplx);
        }
        __syncthreads();
        if (a.K2 != nullptr) {  // to_ortho into A, then copy back to B
            for (int e = tid; e < Tv * n; e += nth) {
                const int t = e / n, k = e - t * n;
                double2 v = (k < n - 2) ? B[t * ldn + k] : make_double2(0.0, 0.0);
                if (k >= 2) {
                    const double kk = a.K2[k - 2];
                    const double2 u = B[t * ldn + k - 2];
                    v.x += kk * u.x;
                    v.y += kk * u.y;
                }
                A[t * ldn + k] = v;
            }
            __syncthreads();
            for (int e = tid; e < Tv * n; e += nth) {
                const int t = e / n, k = e - t * n;
                B[t * ldn + k] = A[t * ldn + k];
            }
            __syncthreads();
        }
        for (int t = tid; t < Tv; t += nth) {
            t0s[2 * t] = B[t * ldn].x;
            t0s[2 * t + 1] = B[t * ldn].y;
        }
        // V[k] = exp(+i pi k/(2N)) ((a_k + b_{N-k}) + i (b_k - a_{N-k})), a_N = b_N = 0
        for (int e = tid; e < Tv * n; e += nth) {
            const int t = e / n, k = e - t * n;
            const double2 ck = B[t * ldn + k];
            const double2 cn = (k == 0) ? make_double2(0.0, 0.0) : B[t * ldn + n - k];
            const double2 z = make_double2(ck.x + cn.y, ck.y - cn.x);
            A[t * ldn + k] = cmul(cconj(a.tw2[k]), z);
        }
    } else if (OP == OP_DCT1F) {
        const int m2 = 2 * (n - 1);
        for (int e = tid; e < Tv * n; e += nth) {
            int t, i;
            tile_index(e, Tv, n, a.t_fast_in, t, i);
            const double2 v = ld_elem(inb, ibase + t * a.in_s2 + i * a.in_sa, a.in_cplx);
            A[t * ldn + i] = v;
            if (i > 0 && i < n - 1) A[t * ldn + m2 - i] = v;
        }
    } else if (OP == OP_DCT1B) {
        const int m2 = 2 * (n - 1);
        for (int e = tid; e < Tv * n; e += nth) {
            int t, i;
            tile_index(e, Tv, n, a.t_fast_in, t, i);
            B[t * ldn + i] = ld_elem(inb, ibase + t * a.in_s2 + i * a.in_sa, a.in_cplx);
        }
        __syncthreads();
        // to_ortho (composite) t_k = u_k (k < n-2) + K2_{k-2} u_{k-2}, into the even extension
        for (int e = tid; e < Tv * n; e += nth) {
            const int t = e / n, k = e - t * n;
            double2 v = B[t * ldn + k];
            if (a.K2 != nullptr) {
                if (k >= n - 2) v = make_double2(0.0, 0.0);
                if (k >= 2) {
                    const double kk = a.K2[k - 2];
                    const double2 u = B[t * ldn + k - 2];
                    v.x += kk * u.x;
                    v.y += kk * u.y;
                }
            }
            A[t * ldn + k] = v;
            if (k > 0 && k < n - 1) A[t * ldn + m2 - k] = v;
            if (k == 0) {
                t0s[4 * t] = v.x;
                t0s[4 * t + 1] = v.y;
            }
            if (k == n - 1) {
                t0s[4 * t + 2] = v.x;
                t0s[4 * t + 3] = v.y;
            }
        }
    }
    __syncthreads();

    // ------------------------------------------------------------ FFT
    double2* F;
    if (OP == OP_C2C)
        F = run_fft<INV>(A, B, a.plan, ldn, Tv, a.tw, tid, nth);
    else if (OP == OP_R2C || OP == OP_R2CF || OP == OP_DCT2 || OP == OP_DCT1F || OP == OP_DCT1B)
        F = run_fft<false>(A, B, a.plan, ldn, Tv, a.tw, tid, nth);
    else
        F = run_fft<true>(A, B, a.plan, ldn, Tv, a.tw, tid, nth);
    double2* G = (F == A) ? B : A;

    // ------------------------------------------------------------ store
    const double sc = a.scale;
    if (OP == OP_R2CF) {
        const int h = n / 2;
        for (int e = tid; e < Tv * (h + 1); e += nth) {
            int t, k;
            tile_index(e, Tv, h + 1, a.t_fast_out, t, k);
            st_elem(outb, obase + t * a.out_s2 + k * a.out_sa, 1, cscale(F[t * ldn + k], sc));
        }
    } else if (OP == OP_C2C || OP == OP_C2RF) {
        for (int e = tid; e < Tv * n; e += nth) {
            int t, i;
            tile_index(e, Tv, n, a.t_fast_out, t, i);
            st_elem(outb, obase + t * a.out_s2 + i * a.out_sa, a.out_cplx, cscale(F[t * ldn + i], sc));
        }
    } else if (OP == OP_R2C) {
        const int h = n / 2;
        for (int e = tid; e < Tv * (h + 1); e += nth) {
            int t, k;
            tile_index(e, Tv, h + 1, a.t_fast_out, t, k);
            const double2 zk = F[t * ldn + (k == h ? 0 : k)];
            const double2 zc = cconj(F[t * ldn + (k == 0 ? 0 : h - k)]);
            const double2 d = cmul(a.tw2[k], csub(zk, zc));  // w (Zk - Zc), w = exp(-2 pi i k/n)
            // X = 0.5 (Zk + Zc) - 0.5 i w (Zk - Zc)
            const double2 X = make_double2(0.5 * (zk.x + zc.x + d.y), 0.5 * (zk.y + zc.y - d.x));
            st_elem(outb, obase + t * a.out_s2 + k * a.out_sa, 1, cscale(X, sc));
        }
    } else if (OP == OP_C2R) {
        const int h = n / 2;
        for (int e = tid; e < Tv * h; e += nth) {
            int t, m;
            tile_index(e, Tv, h, a.t_fast_out, t, m);
            const double2 z = F[t * ldn + m];
            double* p = reinterpret_cast<double*>(outb);
            const int64_t o = obase + t * a.out_s2 + (int64_t)(2 * m) * a.out_sa;
            p[o] = z.x * sc;
            p[o + a.out_sa] = z.y * sc;
        }
    } else if (OP == OP_DCT2) {
        // X_k = Re(w (P + Q)) + i Im(w (P - Q)), P = V_k, Q = conj(V_{N-k})
        for (int e = tid; e < Tv * n; e += nth) {
            const int t = e / n, k = e - t * n;
            const double2 P = F[t * ldn + k];
            const double2 Q = cconj(F[t * ldn + (k == 0 ? 0 : n - k)]);
            const double2 w = a.tw2[k];
            const double2 s1 = cmul(w, cadd(P, Q));
            const double2 s2 = cmul(w, csub(P, Q));
            G[t * ldn + k] = make_double2(s1.x * sc, s2.y * sc);
        }
        __syncthreads();
        for (int e = tid; e < Tv * n; e += nth) {
            int t, k;
            tile_index(e, Tv, n, a.t_fast_out, t, k);
            double2 v = G[t * ldn + k];
            if (a.K2 != nullptr && k < n - 2) {
                const double kk = a.K2[k];
                const double2 u = G[t * ldn + k + 2];
                v.x += kk * u.x;
                v.y += kk * u.y;
            }
            st_elem(outb, obase + t * a.out_s2 + k * a.out_sa, a.out_cplx, v);
        }
    } else if (OP == OP_DCT1F) {
        // s_k = sc Y_k, composite: s_k += K2_k s_{k+2} (k < n-2)
        for (int e = tid; e < Tv * n; e += nth) {
            int t, k;
            tile_index(e, Tv, n, a.t_fast_out, t, k);
            double2 v = cscale(F[t * ldn + k], sc);
            if (a.K2 != nullptr && k < n - 2) {
                const double kk = a.K2[k] * sc;
                const double2 u = F[t * ldn + k + 2];
                v.x += kk * u.x;
                v.y += kk * u.y;
            }
            st_elem(outb, obase + t * a.out_s2 + k * a.out_sa, a.out_cplx, v);
        }
    } else if (OP == OP_DCT1B) {
        // u_j = 0.5 Y_j + 0.5 t_0 + 0.5 (-1)^j t_{n-1}
        for (int e = tid; e < Tv * n; e += nth) {
            int t, j;
            tile_index(e, Tv, n, a.t_fast_out, t, j);
            const double2 y = F[t * ldn + j];
            const double sg = (j & 1) ? -0.5 : 0.5;
            const double2 v = make_double2(
                (0.5 * y.x + 0.5 * t0s[4 * t] + sg * t0s[4 * t + 2]) * sc,
                (0.5 * y.y + 0.5 * t0s[4 * t + 1] + sg * t0s[4 * t + 3]) * sc);
            st_elem(outb, obase + t * a.out_s2 + j * a.out_sa, a.out_cplx, v);
        }
    } else if (OP == OP_DCT3) {
        for (int e = tid; e < Tv * n; e += nth) {
            int t, i;
            tile_index(e, Tv, n, a.t_fast_out, t, i);
            const double2 y = F[t * ldn + dct_perm(i, n)];
            const double2 v = make_double2((0.5 * y.x + 0.5 * t0s[2 * t]) * sc,
                                           (0.5 * y.y + 0.5 * t0s[2 * t + 1]) * sc);
            st_elem(outb, obase + t * a.out_s2 + i * a.out_sa, a.out_cplx, v);
        }
    }
}

// --------------------------------------------------------------- host side

static int choose_T(int ldn, int64_t d2)
{
    // two LDS buffers of T*ldn complex, keep <= 64 KiB (2 blocks / CU)
    int T = (int)(65536 / (2 * 16 * (int64_t)ldn));
    if (T < 1) T = 1;
    if (T > 64) T = 64;
    if (T > d2) T = (int)d2;
    return T;
}

static int launch_line(int op, bool inv, LineArgs& a, hipStream_t stream)
{
    const int ldn = a.ldn;
    const size_t lds = (size_t)2 * a.T * ldn * sizeof(double2) + (size_t)4 * a.T * sizeof(double);
    SFA_CHECK(lds <= 160 * 1024, SFA_ERR_UNSUPPORTED, "transform length %d too large for LDS", a.n);
    const int64_t ntile = (a.d2 + a.T - 1) / a.T;
    const int64_t nblk = ntile * a.d1;
    SFA_CHECK(nblk < (1ll << 31), SFA_ERR_UNSUPPORTED, "too many lines");
    dim3 grid((unsigned)nblk), block(256);
#define SFA_LK(OPV, INVV)                                                                     \
    do {                                                                                     \
        auto kfn = line_kernel<OPV, INVV>;                                                   \
        if (lds > 64 * 1024) hipFuncSetAttribute((const void*)kfn,                          \
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                 (int)lds);                                  \
        hipLaunchKernelGGL(kfn, grid, block, lds, stream, a);                                \
    } while (0)
    switch (op) {
        case OP_C2C:
            if (inv) SFA_LK(OP_C2C, true);
            else SFA_LK(OP_C2C, false);
            break;
        case OP_R2C: SFA_LK(OP_R2C, false); break;
        case OP_C2R: SFA_LK(OP_C2R, true); break;
        case OP_R2CF: SFA_LK(OP_R2CF, false); break;
        case OP_C2RF: SFA_LK(OP_C2RF, true); break;
        case OP_DCT2: SFA_LK(OP_DCT2, false); break;
        case OP_DCT3: SFA_LK(OP_DCT3, true); break;
        case OP_DCT1F: SFA_LK(OP_DCT1F, false); break;
        case OP_DCT1B: SFA_LK(OP_DCT1B, false); break;
    }
#undef SFA_LK
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

int fast_c2c(const Geom& gi, const Geom& go, const void* in, void* out, int sign, double scale,
             hipStream_t s, int* rc);
int fast_rc(bool r2c, const Geom& gi, const Geom& go, const void* in, void* out, double scale,
            hipStream_t s, int* rc, const double* add = nullptr, int64_t add_rows = 1);

static inline bool aligned16(const void* p) { return (((uintptr_t)p) & 15) == 0; }

// Fill LineArgs from geometries (in_geom for the input, out_geom for the output).
static int setup_line(LineArgs& a, const void* in, void* out, const Geom& gi, const Geom& go,
                      int n, int nfft, int in_cplx, int out_cplx)
{
    SFA_CHECK(gi.d1 == go.d1 && gi.d2 == go.d2, SFA_ERR_ARG, "input/output line counts differ");
    a.in = (const char*)in;
    a.out = (char*)out;
    a.in_sa = gi.sa;
    a.in_s1 = gi.s1;
    a.in_s2 = gi.s2;
    a.out_sa = go.sa;
    a.out_s1 = go.s1;
    a.out_s2 = go.s2;
    a.d1 = gi.d1;
    a.d2 = gi.d2;
    a.n = n;
    a.in_cplx = in_cplx;
    a.out_cplx = out_cplx;
    a.ldn = nfft + 2;
    a.T = choose_T(a.ldn, a.d2);
    auto absl = [](int64_t v) { return v < 0 ? -v : v; };
    a.t_fast_in = absl(gi.s2) < absl(gi.sa) ? 1 : 0;
    a.t_fast_out = absl(go.s2) < absl(go.sa) ? 1 : 0;
    a.plan = make_fft_plan(nfft);
    a.tw = fft_table(nfft);
    SFA_CHECK(a.tw != nullptr, SFA_ERR_HIP, "twiddle table allocation failed");
    a.tw2 = nullptr;
    a.K2 = nullptr;
    a.inv = 0;
    a.scale = 1.0;
    return SFA_OK;
}

}  // namespace sfa

using namespace sfa;

extern "C" int sfa_fft_c2c(const void* in, void* out, int ndim, const int64_t* shape,
                           const int64_t* in_strides, const int64_t* out_strides, int axis,
                           int sign, double scale, void* stream)
{
    Geom gi, go;
    int rc;
    if ((rc = make_geom(ndim, shape, in_strides, axis, &gi))) return rc;
    if ((rc = make_geom(ndim, shape, out_strides, axis, &go))) return rc;
    SFA_CHECK(sign == 1 || sign == -1, SFA_ERR_ARG, "sign must be +-1");
    if (aligned16(in) && aligned16(out) && fast_c2c(gi, go, in, out, sign, scale, (hipStream_t)stream, &rc))
        return rc;
    LineArgs a;
    if ((rc = setup_line(a, in, out, gi, go, (int)gi.n, (int)gi.n, 1, 1))) return rc;
    a.scale = scale;
    return launch_line(OP_C2C, sign > 0, a, (hipStream_t)stream);
}

extern "C" int sfa_fft_r2c(const double* in, void* out, int ndim, const int64_t* real_shape,
                           const int64_t* in_strides, const int64_t* out_strides, int axis,
                           double scale, void* stream)
{
    Geom gi, go;
    int rc;
    if ((rc = make_geom(ndim, real_shape, in_strides, axis, &gi))) return rc;
    if ((rc = make_geom(ndim, real_shape, out_strides, axis, &go))) return rc;
    const int n = (int)gi.n;
    SFA_CHECK(n >= 1, SFA_ERR_ARG, "R2C length must be >= 1");
    SFA_CHECK(in != nullptr && out != nullptr, SFA_ERR_ARG, "R2C: null array");
    if (n % 2) {
        LineArgs a;
        if ((rc = setup_line(a, in, out, gi, go, n, n, 0, 1))) return rc;
        a.scale = scale;
        return launch_line(OP_R2CF, false, a, (hipStream_t)stream);
    }
    if (aligned16(in) && aligned16(out) && fast_rc(true, gi, go, in, out, scale, (hipStream_t)stream, &rc))
        return rc;
    LineArgs a;
    if ((rc = setup_line(a, in, out, gi, go, n, n / 2, 0, 1))) return rc;
    a.tw2 = fft_table(n);
    SFA_CHECK(a.tw2 != nullptr, SFA_ERR_HIP, "twiddle table allocation failed");
    a.scale = scale;
    return launch_line(OP_R2C, false, a, (hipStream_t)stream);
}

int fft_c2r_add(const void* in, double* out, int ndim, const int64_t* real_shape, const int64_t* in_strides,
                const int64_t* out_strides, int axis, double scale, void* stream, const double* add,
                int64_t add_rows);

extern "C" int sfa_fft_c2r(const void* in, double* out, int ndim, const int64_t* real_shape,
                           const int64_t* in_strides, const int64_t* out_strides, int axis,
                           double scale, void* stream)
{
    return fft_c2r_add(in, out, ndim, real_shape, in_strides, out_strides, axis, scale, stream, nullptr, 1);
}

// sfa_fft_c2r that also adds add (add_rows contiguous rows of the transform
// length, row R % add_rows to output row R, the output rows contiguous) to
// the result: inside the transform's stores where its fast kernel runs, else
// by a pass over the output afterwards.
int fft_c2r_add(const void* in, double* out, int ndim, const int64_t* real_shape, const int64_t* in_strides,
                const int64_t* out_strides, int axis, double scale, void* stream, const double* add,
                int64_t add_rows)
{
    Geom gi, go;
    int rc;
    if ((rc = make_geom(ndim, real_shape, in_strides, axis, &gi))) return rc;
    if ((rc = make_geom(ndim, real_shape, out_strides, axis, &go))) return rc;
    const int n = (int)gi.n;
    SFA_CHECK(n >= 1, SFA_ERR_ARG, "C2R length must be >= 1");
    SFA_CHECK(in != nullptr && out != nullptr, SFA_ERR_ARG, "C2R: null array");
    if (n % 2) {
        SFA_CHECK(add == nullptr, SFA_ERR_UNSUPPORTED, "C2R with add: odd length");
        LineArgs a;
        if ((rc = setup_line(a, in, out, gi, go, n, n, 1, 0))) return rc;
        a.scale = scale;
        return launch_line(OP_C2RF, true, a, (hipStream_t)stream);
    }
    if (aligned16(in) && aligned16(out) &&
        fast_rc(false, gi, go, in, out, scale, (hipStream_t)stream, &rc, add, add_rows))
        return rc;
    LineArgs a;
    if ((rc = setup_line(a, in, out, gi, go, n, n / 2, 1, 0))) return rc;
    a.tw2 = fft_table(n);
    SFA_CHECK(a.tw2 != nullptr, SFA_ERR_HIP, "twiddle table allocation failed");
    a.scale = scale;
    if ((rc = launch_line(OP_C2R, true, a, (hipStream_t)stream))) return rc;
    if (add) {
        SFA_CHECK(go.s2 == n && go.s1 == go.d2 * n && (go.d1 * go.d2) % add_rows == 0, SFA_ERR_ARG,
                  "C2R with add: output rows must be contiguous");
        return sfa_axpy(add, out, 1.0, add_rows * n, go.d1 * go.d2 / add_rows, stream);
    }
    return SFA_OK;
}

static int cheb_common(int op, const void* in, void* out, int dtype, int ndim, const int64_t* shape,
                       const int64_t* in_strides, const int64_t* out_strides, int axis,
                       const double* K2, double scale, void* stream)
{
    Geom gi, go;
    int rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "bad dtype %d", dtype);
    if ((rc = make_geom(ndim, shape, in_strides, axis, &gi))) return rc;
    if ((rc = make_geom(ndim, shape, out_strides, axis, &go))) return rc;
    const int n = (int)gi.n;
    SFA_CHECK(n >= 3 || K2 == nullptr, SFA_ERR_ARG, "composite basis needs N >= 3");
    LineArgs a;
    const int c = dtype == SFA_C128;
    if ((rc = setup_line(a, in, out, gi, go, n, n, c, c))) return rc;
    a.tw2 = quarter_table(n);
    SFA_CHECK(a.tw2 != nullptr, SFA_ERR_HIP, "twiddle table allocation failed");
    a.K2 = K2;
    a.scale = scale;
    return launch_line(op, op == OP_DCT3, a, (hipStream_t)stream);
}

extern "C" int sfa_cheb_scalar_product(const void* in, void* out, int dtype, int ndim,
                                       const int64_t* shape, const int64_t* in_strides,
                                       const int64_t* out_strides, int axis, const double* K2,
                                       void* stream)
{
    const double N = (double)shape[axis >= 0 && axis < ndim ? axis : 0];
    return cheb_common(OP_DCT2, in, out, dtype, ndim, shape, in_strides, out_strides, axis, K2,
                       3.14159265358979323846 / (2.0 * N), stream);
}

extern "C" int sfa_cheb_backward(const void* in, void* out, int dtype, int ndim,
                                 const int64_t* shape, const int64_t* in_strides,
                                 const int64_t* out_strides, int axis, const double* K2,
                                 void* stream)
{
    return cheb_common(OP_DCT3, in, out, dtype, ndim, shape, in_strides, out_strides, axis, K2, 1.0,
                       stream);
}

// Gauss-Lobatto: DCT-I of length n through the FFT of the even extension.
static int cheb_gl_common(int op, const void* in, void* out, int dtype, int ndim,
                          const int64_t* shape, const int64_t* in_strides,
                          const int64_t* out_strides, int axis, const double* K2, double scale,
                          void* stream)
{
    Geom gi, go;
    int rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "bad dtype %d", dtype);
    if ((rc = make_geom(ndim, shape, in_strides, axis, &gi))) return rc;
    if ((rc = make_geom(ndim, shape, out_strides, axis, &go))) return rc;
    const int n = (int)gi.n;
    SFA_CHECK(n >= 2, SFA_ERR_ARG, "Gauss-Lobatto needs N >= 2");
    SFA_CHECK(n >= 3 || K2 == nullptr, SFA_ERR_ARG, "composite basis needs N >= 3");
    LineArgs a;
    const int c = dtype == SFA_C128;
    if ((rc = setup_line(a, in, out, gi, go, n, 2 * (n - 1), c, c))) return rc;
    a.K2 = K2;
    a.scale = scale;
    return launch_line(op, false, a, (hipStream_t)stream);
}

extern "C" int sfa_cheb_scalar_product_gl(const void* in, void* out, int dtype, int ndim,
                                          const int64_t* shape, const int64_t* in_strides,
                                          const int64_t* out_strides, int axis, const double* K2,
                                          void* stream)
{
    const double N = (double)shape[axis >= 0 && axis < ndim ? axis : 0];
    return cheb_gl_common(OP_DCT1F, in, out, dtype, ndim, shape, in_strides, out_strides, axis, K2,
                          3.14159265358979323846 / (2.0 * (N - 1.0)), stream);
}

extern "C" int sfa_cheb_backward_gl(const void* in, void* out, int dtype, int ndim,
                                    const int64_t* shape, const int64_t* in_strides,
                                    const int64_t* out_strides, int axis, const double* K2,
                                    void* stream)
{
    return cheb_gl_common(OP_DCT1B, in, out, dtype, ndim, shape, in_strides, out_strides, axis, K2,
                          1.0, stream);
}
