// Chebyshev Galerkin biharmonic solver a0 S + alfa A + beta B on the
// ShenBiharmonic basis (chebyshev/la.py:354-558; la.pyx:975-1310; the matvec
// of Matvec.pyx:820-923), along any axis of a 1-3 dimensional array.
//
// All offsets are even, so even and odd unknowns i = 2r + p form two
// independent chains of length M_p = (N - 4 - p + 1) / 2 (r = chain row).  In
// chain indices the matrix has two lower bands (alfa A and beta B), three
// upper band entries per row and, from S, a dense upper tail of rank two:
//   S(r, c) = p_r / (j + 3) + q_r (j + 2)^2 / (j + 3),  c >= r + 3, j = 2c + p,
//   p_r = 8 pi k (k+1)(k+2)(k+4), q_r = 24 pi (k+1)(k+2), k = 2r + p.
// Gaussian elimination with the two lower bands keeps U's rows of that form:
// U(r, c >= r+3) = a0 (ak_r / (j+3) + bk_r (j+2)^2 / (j+3)) with ak, bk
// eliminated like the rows (Biharmonic_factor_oe_pr, la.pyx:1211-1233).  The
// LU here follows that structure row by row, O(M) per chain (the reference
// updates whole rows, O(M^2)); the solve is the reference's forward sweep and
// back substitution with the two running tail sums (la.pyx:1243-1310).
//
// One thread per (line, parity) chain, consecutive threads on consecutive
// lines, so every row access is coalesced across lines.  Factor arrays have
// the reference's layout: C order, shape (2,) + u's shape with the axis
// length replaced by the even chain's length.
#include "common.h"

namespace sfa {
namespace {

struct Lines {  // lines of a strided array: C order over the non-axis dims
    int64_t sa, s1, s2, d2, nlines;
    __device__ inline int64_t off(int64_t L) const
    {
        const int64_t i1 = L / d2;
        return i1 * s1 + (L - i1 * d2) * s2;
    }
};

int lines_of(int ndim, const int64_t* shape, const int64_t* strides, int axis, Lines* m, int64_t* n)
{
    Geom g;
    const int rc = make_geom(ndim, shape, strides, axis, &g);
    if (rc) return rc;
    m->sa = g.sa;
    m->s1 = g.s1;
    m->s2 = g.s2;
    m->d2 = g.d2;
    m->nlines = g.d1 * g.d2;
    if (n) *n = g.n;
    return SFA_OK;
}

// Lines of the factor arrays ((2,) + shape with the axis length Mh, C order);
// *P = elements per parity.
int factor_lines(int ndim, const int64_t* shape, int axis, int64_t Mh, Lines* m, int64_t* P)
{
    int64_t fs[3], st[3];
    for (int i = 0; i < ndim; ++i) fs[i] = (i == axis) ? Mh : shape[i];
    int64_t acc = 1;
    for (int i = ndim - 1; i >= 0; --i) {
        st[i] = acc;
        acc *= fs[i];
    }
    *P = acc;
    return lines_of(ndim, fs, st, axis, m, nullptr);
}

struct BihDiags {
    // full diagonals as the reference passes them: S[0], S[2], S[4] (upper),
    // A[-2], A[0], A[2], B[-4], B[-2], B[0], B[2], B[4]; lower diagonal k is
    // stored from its first entry, D[k][i] = M(i + |k|, i)
    const double *sii, *siu, *siuu, *ail, *aii, *aiu, *bill, *bil, *bii, *biu, *biuu;
};

__device__ inline double mulsub(double a, double s, double b) { return a - s * b; }
__device__ inline double2 mulsub(double2 a, double s, double2 b)
{
    return make_double2(a.x - s * b.x, a.y - s * b.y);
}
__device__ inline double divr(double a, double s) { return a / s; }
__device__ inline double2 divr(double2 a, double s) { return make_double2(a.x / s, a.y / s); }
__device__ inline double addt(double a, double b) { return a + b; }
__device__ inline double2 addt(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double scl(double a, double s) { return a * s; }
__device__ inline double2 scl(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
template <typename T>
__device__ inline T zero();
template <>
__device__ inline double zero<double>() { return 0.0; }
template <>
__device__ inline double2 zero<double2>() { return make_double2(0.0, 0.0); }

constexpr double kPi = 3.14159265358979323846;

// Row r's tail coefficients of S: p_r, q_r (k = 2r + parity).
__device__ inline void s_tail(int k, double* p, double* q)
{
    const double kk = k;
    *p = 8.0 * kk * (kk + 1) * (kk + 2) * (kk + 4) * kPi;
    *q = 24.0 * (kk + 1) * (kk + 2) * kPi;
}

__global__ void biharmonic_lu_kernel(Lines fl, int64_t P, int Nu, double a0,
                                     const double* __restrict__ alfa, const double* __restrict__ beta,
                                     int64_t ab_stride, BihDiags dg, double* __restrict__ u0,
                                     double* __restrict__ u1, double* __restrict__ u2,
                                     double* __restrict__ l0, double* __restrict__ l1,
                                     double* __restrict__ ak, double* __restrict__ bk)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = fl.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    const int M = (Nu - p + 1) / 2;
    if (M <= 0) return;
    const int64_t o = p * P + fl.off(L), rs = fl.sa;
    const double a = a0, b = alfa[L * ab_stride], c = beta[L * ab_stride];
    // state of rows r-1 and r-2
    double u0m1 = 0, u0m2 = 0, u1m1 = 0, u1m2 = 0, u2m1 = 0, u2m2 = 0;
    double tam1 = 0, tam2 = 0, tbm1 = 0, tbm2 = 0;
    for (int r = 0; r < M; ++r) {
        const int i = 2 * r + p;
        double lb2 = 0.0, lb1 = 0.0;  // L(r, r-2), L(r, r-1)
        if (r >= 2) {
            lb2 = (c * dg.bill[i - 4]) / u0m2;
            l1[o + (r - 2) * rs] = lb2;
        }
        if (r >= 1) {
            double arm1 = b * dg.ail[i - 2] + c * dg.bil[i - 2];
            if (r >= 2) arm1 -= lb2 * u1m2;
            lb1 = arm1 / u0m1;
            l0[o + (r - 1) * rs] = lb1;
        }
        // U(r-1, c), U(r-2, c) in the tails: a0 (ta / (j+3) + tb (j+2)^2 / (j+3))
        auto tail = [&](double ta, double tb, int col) {
            const double j = 2.0 * col + p;
            return a * (ta / (j + 3.0) + tb * ((j + 2.0) * (j + 2.0)) / (j + 3.0));
        };
        const double v0 = a * dg.sii[i] + b * dg.aii[i] + c * dg.bii[i] - lb1 * u1m1 - lb2 * u2m2;
        double v1 = 0.0, v2 = 0.0;
        if (r + 1 < M) {
            v1 = a * dg.siu[i] + b * dg.aiu[i] + c * dg.biu[i] - lb1 * u2m1;
            if (r >= 2) v1 -= lb2 * tail(tam2, tbm2, r + 1);
        }
        if (r + 2 < M) {
            v2 = a * dg.siuu[i] + c * dg.biuu[i];
            if (r >= 1) v2 -= lb1 * tail(tam1, tbm1, r + 2);
            if (r >= 2) v2 -= lb2 * tail(tam2, tbm2, r + 2);
        }
        double pr, qr;
        s_tail(i, &pr, &qr);
        const double ta = pr - lb1 * tam1 - lb2 * tam2;
        const double tb = qr - lb1 * tbm1 - lb2 * tbm2;
        u0[o + r * rs] = v0;
        if (r + 1 < M) u1[o + r * rs] = v1;
        if (r + 2 < M) u2[o + r * rs] = v2;
        if (ak && r + 3 <= M) {  // the rows la.pyx:1211-1233 defines (r <= M - 3)
            ak[o + r * rs] = ta;
            bk[o + r * rs] = tb;
        }
        u0m2 = u0m1;
        u0m1 = v0;
        u1m2 = u1m1;
        u1m1 = v1;
        u2m2 = u2m1;
        u2m1 = v2;
        tam2 = tam1;
        tam1 = ta;
        tbm2 = tbm1;
        tbm1 = tb;
    }
}

// Biharmonic_factor_oe_pr (la.pyx:1211-1233): ak, bk from l0, l1.
__global__ void biharmonic_factor_kernel(Lines fl, int64_t P, int Nu, const double* __restrict__ l0,
                                         const double* __restrict__ l1, double* __restrict__ ak,
                                         double* __restrict__ bk)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = fl.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    const int M = (Nu - p + 1) / 2;
    const int64_t o = p * P + fl.off(L), rs = fl.sa;
    double am1 = 0, am2 = 0, bm1 = 0, bm2 = 0;
    for (int r = 0; r + 3 <= M; ++r) {
        double pr, qr;
        s_tail(2 * r + p, &pr, &qr);
        double ta = pr, tb = qr;
        if (r >= 1) {
            ta -= l0[o + (r - 1) * rs] * am1;
            tb -= l0[o + (r - 1) * rs] * bm1;
        }
        if (r >= 2) {
            ta -= l1[o + (r - 2) * rs] * am2;
            tb -= l1[o + (r - 2) * rs] * bm2;
        }
        ak[o + r * rs] = ta;
        bk[o + r * rs] = tb;
        am2 = am1;
        am1 = ta;
        bm2 = bm1;
        bm1 = tb;
    }
}

template <typename T>
__global__ void biharmonic_solve_kernel(const T* __restrict__ b, T* __restrict__ u, Lines gb, Lines gu,
                                        Lines fl, int64_t P, int Nu, double a0,
                                        const double* __restrict__ u0, const double* __restrict__ u1,
                                        const double* __restrict__ u2, const double* __restrict__ l0,
                                        const double* __restrict__ l1, const double* __restrict__ ak,
                                        const double* __restrict__ bk)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = gu.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    const int M = (Nu - p + 1) / 2;
    if (M <= 0) return;
    const T* bl = b + gb.off(L);
    T* ul = u + gu.off(L);
    const int64_t o = p * P + fl.off(L), rs = fl.sa;
    const int64_t sb = 2 * gb.sa, su = 2 * gu.sa;  // chain steps
    bl += p * gb.sa;
    ul += p * gu.sa;
    // forward: y_r = f_r - l0_{r-1} y_{r-1} - l1_{r-2} y_{r-2}, stored in u
    T y1 = zero<T>(), y2 = zero<T>();
    for (int r = 0; r < M; ++r) {
        T y = bl[r * sb];
        if (r >= 1) y = mulsub(y, l0[o + (r - 1) * rs], y1);
        if (r >= 2) y = mulsub(y, l1[o + (r - 2) * rs], y2);
        ul[r * su] = y;
        y2 = y1;
        y1 = y;
    }
    // back substitution with the tail sums s1 = sum x/(j+3), s2 = sum x (j+2)^2/(j+3)
    T x1 = zero<T>(), x2 = zero<T>(), x3 = zero<T>();
    T s1 = zero<T>(), s2 = zero<T>();
    for (int r = M - 1; r >= 0; --r) {
        T v = ul[r * su];
        if (r + 1 < M) v = mulsub(v, u1[o + r * rs], x1);
        if (r + 2 < M) v = mulsub(v, u2[o + r * rs], x2);
        if (r + 3 < M) {
            const double j = 2.0 * (r + 3) + p;
            const T w = divr(x3, j + 3.0);
            s1 = addt(s1, w);
            s2 = addt(s2, scl(w, (j + 2.0) * (j + 2.0)));
            v = mulsub(v, ak[o + r * rs] * a0, s1);
            v = mulsub(v, bk[o + r * rs] * a0, s2);
        }
        v = divr(v, u0[o + r * rs]);
        ul[r * su] = v;
        x3 = x2;
        x2 = x1;
        x1 = v;
    }
}

// c = (a0 S + alfa A + beta B) v on the unknowns (entries N-4.. of c zero):
// the band part directly, S's tail through reverse running sums per chain.
template <typename T>
__global__ void biharmonic_matvec_kernel(const T* __restrict__ v, T* __restrict__ cc, Lines gv, Lines gc,
                                         int Nu, int N, double a0, const double* __restrict__ alfa,
                                         const double* __restrict__ beta, int64_t ab_stride, BihDiags dg)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = gv.nlines;
    if (gid >= 2 * nl) return;
    const int p = (int)(gid / nl);
    const int64_t L = gid - (int64_t)p * nl;
    const int M = (Nu - p + 1) / 2;
    const T* vl = v + gv.off(L);
    T* cl = cc + gc.off(L);
    const int64_t sv = gv.sa, sc = gc.sa;
    const double a = a0, b = alfa[L * ab_stride], c = beta[L * ab_stride];
    T s1 = zero<T>(), s2 = zero<T>();
    for (int r = M - 1; r >= 0; --r) {
        const int i = 2 * r + p;
        if (r + 3 < M) {
            const int j = i + 6;
            const T w = divr(vl[j * sv], j + 3.0);
            s1 = addt(s1, w);
            s2 = addt(s2, scl(w, (j + 2.0) * (j + 2.0)));
        }
        double pr, qr;
        s_tail(i, &pr, &qr);
        T acc = scl(s1, a * pr);
        acc = addt(acc, scl(s2, a * qr));
        acc = addt(acc, scl(vl[i * sv], a * dg.sii[i] + b * dg.aii[i] + c * dg.bii[i]));
        if (r + 1 < M) acc = addt(acc, scl(vl[(i + 2) * sv], a * dg.siu[i] + b * dg.aiu[i] + c * dg.biu[i]));
        if (r + 2 < M) acc = addt(acc, scl(vl[(i + 4) * sv], a * dg.siuu[i] + c * dg.biuu[i]));
        if (r >= 1) acc = addt(acc, scl(vl[(i - 2) * sv], b * dg.ail[i - 2] + c * dg.bil[i - 2]));
        if (r >= 2) acc = addt(acc, scl(vl[(i - 4) * sv], c * dg.bill[i - 4]));
        cl[i * sc] = acc;
    }
    if (p == 0)
        for (int i = Nu; i < N; ++i) cl[i * sc] = zero<T>();
}

inline dim3 grid_of(int64_t nthreads) { return dim3((unsigned)((nthreads + 255) / 256)); }

}  // namespace
}  // namespace sfa

using namespace sfa;

extern "C" int sfa_biharmonic_lu(int ndim, const int64_t* shape, int axis, double a0,
                                 const double* alfa, const double* beta, int64_t ab_stride,
                                 const double* const* diags, double* u0, double* u1, double* u2,
                                 double* l0, double* l1, double* ak, double* bk, void* stream)
{
    SFA_CHECK(ndim >= 1 && ndim <= 3 && axis >= 0 && axis < ndim, SFA_ERR_ARG, "bad ndim/axis");
    const int64_t N = shape[axis];
    SFA_CHECK(N >= 8, SFA_ERR_ARG, "biharmonic needs N >= 8 (N - 4 unknowns)");
    SFA_CHECK(ab_stride == 0 || ab_stride == 1, SFA_ERR_ARG, "ab_stride must be 0 or 1");
    SFA_CHECK(diags && u0 && u1 && u2 && l0 && l1 && alfa && beta, SFA_ERR_ARG, "NULL argument");
    SFA_CHECK((ak == nullptr) == (bk == nullptr), SFA_ERR_ARG, "ak and bk go together");
    for (int k = 0; k < 11; ++k) SFA_CHECK(diags[k] != nullptr, SFA_ERR_ARG, "diagonal %d is NULL", k);
    const int Nu = (int)N - 4;
    Lines fl;
    int64_t P;
    int rc = factor_lines(ndim, shape, axis, (Nu + 1) / 2, &fl, &P);
    if (rc) return rc;
    BihDiags dg{diags[0], diags[1], diags[2], diags[3], diags[4], diags[5],
                diags[6], diags[7], diags[8], diags[9], diags[10]};
    hipLaunchKernelGGL(biharmonic_lu_kernel, grid_of(2 * fl.nlines), dim3(256), 0, (hipStream_t)stream,
                       fl, P, Nu, a0, alfa, beta, ab_stride, dg, u0, u1, u2, l0, l1, ak, bk);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_biharmonic_factor(int ndim, const int64_t* shape, int axis, const double* l0,
                                     const double* l1, double* ak, double* bk, void* stream)
{
    SFA_CHECK(ndim >= 1 && ndim <= 3 && axis >= 0 && axis < ndim, SFA_ERR_ARG, "bad ndim/axis");
    const int64_t N = shape[axis];
    SFA_CHECK(N >= 8, SFA_ERR_ARG, "biharmonic needs N >= 8");
    SFA_CHECK(l0 && l1 && ak && bk, SFA_ERR_ARG, "NULL argument");
    const int Nu = (int)N - 4;
    Lines fl;
    int64_t P;
    int rc = factor_lines(ndim, shape, axis, (Nu + 1) / 2, &fl, &P);
    if (rc) return rc;
    hipLaunchKernelGGL(biharmonic_factor_kernel, grid_of(2 * fl.nlines), dim3(256), 0,
                       (hipStream_t)stream, fl, P, Nu, l0, l1, ak, bk);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_biharmonic_solve(const void* b, void* u, int dtype, int ndim, const int64_t* shape,
                                    const int64_t* b_strides, const int64_t* u_strides, int axis,
                                    double a0, const double* u0, const double* u1, const double* u2,
                                    const double* l0, const double* l1, const double* ak,
                                    const double* bk, void* stream)
{
    Lines gb, gu, fl;
    int64_t N, P;
    int rc;
    if ((rc = lines_of(ndim, shape, b_strides, axis, &gb, &N))) return rc;
    if ((rc = lines_of(ndim, shape, u_strides, axis, &gu, nullptr))) return rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "bad dtype %d", dtype);
    SFA_CHECK(N >= 8, SFA_ERR_ARG, "biharmonic needs N >= 8");
    SFA_CHECK(b && u && u0 && u1 && u2 && l0 && l1 && ak && bk, SFA_ERR_ARG, "NULL argument");
    const int Nu = (int)N - 4;
    if ((rc = factor_lines(ndim, shape, axis, (Nu + 1) / 2, &fl, &P))) return rc;
    dim3 grid = grid_of(2 * gu.nlines);
    if (dtype == SFA_C128)
        hipLaunchKernelGGL(biharmonic_solve_kernel<double2>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double2*)b, (double2*)u, gb, gu, fl, P, Nu, a0, u0, u1, u2, l0, l1,
                           ak, bk);
    else
        hipLaunchKernelGGL(biharmonic_solve_kernel<double>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double*)b, (double*)u, gb, gu, fl, P, Nu, a0, u0, u1, u2, l0, l1,
                           ak, bk);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

extern "C" int sfa_biharmonic_matvec(const void* v, void* c, int dtype, int ndim, const int64_t* shape,
                                     const int64_t* v_strides, const int64_t* c_strides, int axis,
                                     double a0, const double* alfa, const double* beta,
                                     int64_t ab_stride, const double* const* diags, void* stream)
{
    Lines gv, gc;
    int64_t N;
    int rc;
    if ((rc = lines_of(ndim, shape, v_strides, axis, &gv, &N))) return rc;
    if ((rc = lines_of(ndim, shape, c_strides, axis, &gc, nullptr))) return rc;
    SFA_CHECK(dtype == SFA_F64 || dtype == SFA_C128, SFA_ERR_ARG, "bad dtype %d", dtype);
    SFA_CHECK(N >= 8, SFA_ERR_ARG, "biharmonic needs N >= 8");
    SFA_CHECK(v != c, SFA_ERR_ARG, "matvec cannot run in place");
    SFA_CHECK(ab_stride == 0 || ab_stride == 1, SFA_ERR_ARG, "ab_stride must be 0 or 1");
    SFA_CHECK(diags && alfa && beta, SFA_ERR_ARG, "NULL argument");
    for (int k = 0; k < 11; ++k) SFA_CHECK(diags[k] != nullptr, SFA_ERR_ARG, "diagonal %d is NULL", k);
    BihDiags dg{diags[0], diags[1], diags[2], diags[3], diags[4], diags[5],
                diags[6], diags[7], diags[8], diags[9], diags[10]};
    dim3 grid = grid_of(2 * gv.nlines);
    if (dtype == SFA_C128)
        hipLaunchKernelGGL(biharmonic_matvec_kernel<double2>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double2*)v, (double2*)c, gv, gc, (int)N - 4, (int)N, a0, alfa, beta,
                           ab_stride, dg);
    else
        hipLaunchKernelGGL(biharmonic_matvec_kernel<double>, grid, dim3(256), 0, (hipStream_t)stream,
                           (const double*)v, (double*)c, gv, gc, (int)N - 4, (int)N, a0, alfa, beta,
                           ab_stride, dg);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}
