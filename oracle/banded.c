/* Plain-C restatement of shenfun's serial Cython loops (test oracle and timed
 * CPU baseline only; never linked into the product library).
 *
 * Arrays are C-order (N, nlines): the solved/transformed axis is axis 0 with
 * stride nlines, exactly the access pattern of the reference's
 * Solve_axis_3D and Solve_Helmholtz_3D_ptr loops for axis 0 of a
 * C-contiguous array.  Complex data are interleaved (re, im) doubles.
 *
 * nthreads > 1 splits the lines over OpenMP threads (stand-in for running
 * the reference under mpirun -np n); nthreads == 1 is the serial loop.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

/* la.pyx:871-912 Solve_Helmholtz_1D_ptr (complex T), one line. */
static void helm_solve_line(int N, long st, const cplx* fk, cplx* u, int neumann,
                            const double* d0, const double* d1, const double* d2,
                            const double* L, cplx* y)
{
    int M = N - 2;
    y[0] = fk[0];
    y[1] = fk[st];
    for (int i = 2; i < M; ++i) {
        double l = L[(long)(i - 2) * st];
        y[i].re = fk[(long)i * st].re - l * y[i - 2].re;
        y[i].im = fk[(long)i * st].im - l * y[i - 2].im;
    }
#define U(i) u[(long)(i) * st]
#define D(a, i) a[(long)(i) * st]
    U(M - 1).re = y[M - 1].re / D(d0, M - 1); U(M - 1).im = y[M - 1].im / D(d0, M - 1);
    U(M - 2).re = y[M - 2].re / D(d0, M - 2); U(M - 2).im = y[M - 2].im / D(d0, M - 2);
    U(M - 3).re = (y[M - 3].re - D(d1, M - 3) * U(M - 1).re) / D(d0, M - 3);
    U(M - 3).im = (y[M - 3].im - D(d1, M - 3) * U(M - 1).im) / D(d0, M - 3);
    U(M - 4).re = (y[M - 4].re - D(d1, M - 4) * U(M - 2).re) / D(d0, M - 4);
    U(M - 4).im = (y[M - 4].im - D(d1, M - 4) * U(M - 2).im) / D(d0, M - 4);
    cplx se = {0, 0}, so = {0, 0};
    for (int i = M - 5; i >= 0; --i) {
        cplx v;
        v.re = y[i].re - D(d1, i) * U(i + 2).re;
        v.im = y[i].im - D(d1, i) * U(i + 2).im;
        if (i % 2 == 0) {
            se.re += U(i + 4).re; se.im += U(i + 4).im;
            v.re -= D(d2, i) * se.re; v.im -= D(d2, i) * se.im;
        } else {
            so.re += U(i + 4).re; so.im += U(i + 4).im;
            v.re -= D(d2, i) * so.re; v.im -= D(d2, i) * so.im;
        }
        U(i).re = v.re / D(d0, i);
        U(i).im = v.im / D(d0, i);
    }
    if (neumann) {
        if ((d0[0] - 1.0) * (d0[0] - 1.0) < 1e-16) { U(0).re = 0; U(0).im = 0; }
        for (int i = 1; i < M; ++i) { U(i).re /= (double)i * i; U(i).im /= (double)i * i; }
    }
#undef U
#undef D
}

/* Solve_Helmholtz_3D_ptr for axis 0 (la.pyx:914-947): factors share u's layout. */
void orc_helmholtz_solve_axis0(int N, long nlines, const double* fk, double* u,
                               int neumann, const double* d0, const double* d1,
                               const double* d2, const double* L, int nthreads)
{
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        cplx* y = (cplx*)malloc(sizeof(cplx) * (size_t)N);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (long j = 0; j < nlines; ++j)
            helm_solve_line(N, nlines, (const cplx*)fk + j, (cplx*)u + j, neumann,
                            d0 + j, d1 + j, d2 + j, L + j, y);
        free(y);
    }
}

/* la.pyx:520-532 TDMA_inner_solve_ptr with per-line factors
 * (SolverGeneric1ND_solve_2D, la.pyx:411-426, naxes = 0): data is
 * (nlines, 3, n) and line j < skip is left alone (is_zero_index). */
void orc_tdma_solve_modes_axis0(int n, long nlines, double* u, const double* data,
                                long skip, int nthreads)
{
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (long j = skip; j < nlines; ++j) {
        cplx* v = (cplx*)u + j;
        const double* ld = data + (size_t)j * 3 * n;
        const double* d = ld + n;
        const double* ud = ld + 2 * n + 2;
        long st = nlines;
        for (int i = 2; i < n; ++i) {
            v[i * st].re -= ld[i - 2] * v[(i - 2) * st].re;
            v[i * st].im -= ld[i - 2] * v[(i - 2) * st].im;
        }
        v[(n - 1) * st].re /= d[n - 1]; v[(n - 1) * st].im /= d[n - 1];
        v[(n - 2) * st].re /= d[n - 2]; v[(n - 2) * st].im /= d[n - 2];
        for (int i = n - 3; i >= 0; --i) {
            v[i * st].re = (v[i * st].re - ud[i] * v[(i + 2) * st].re) / d[i];
            v[i * st].im = (v[i * st].im - ud[i] * v[(i + 2) * st].im) / d[i];
        }
    }
}

/* transforms.pyx:270-323 _scalar_product_ptr (complex), per line, axis 0.
 * anm/anp: recursion DIA rows (length >= N+3). */
void orc_leg_scalar_product_axis0(int N, long nlines, const double* in, double* out,
                                  const double* x, const double* w,
                                  const double* anm, const double* anp, int nthreads)
{
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double* Lnm = (double*)malloc(sizeof(double) * N);
        double* Ln = (double*)malloc(sizeof(double) * N);
        double* Lnp = (double*)malloc(sizeof(double) * N);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (long l = 0; l < nlines; ++l) {
            const cplx* ui = (const cplx*)in + l;
            cplx* uo = (cplx*)out + l;
            for (int i = 0; i < N; ++i) {
                Lnm[i] = 1.0;
                Ln[i] = x[i] / anm[0];
                Lnp[i] = x[i] / anm[1] * Ln[i] - anp[1] / anm[1] * Lnm[i];
            }
            for (int i = 0; i < N; ++i) {
                double s1 = 1.0 / anm[i + 2], s2 = anp[i + 2] / anm[i + 2];
                cplx s = {0, 0};
                for (int j = 0; j < N; ++j) {
                    double lw = Lnm[j] * w[j];
                    s.re += lw * ui[(long)j * nlines].re;
                    s.im += lw * ui[(long)j * nlines].im;
                    Lnm[j] = Ln[j];
                    Ln[j] = Lnp[j];
                    Lnp[j] = s1 * x[j] * Ln[j] - s2 * Lnm[j];
                }
                uo[(long)i * nlines] = s;
            }
        }
        free(Lnm); free(Ln); free(Lnp);
    }
}

/* transforms.pyx:219-267 _evaluate_expansion_all_ptr (complex), axis 0. */
void orc_leg_backward_axis0(int N, long nlines, const double* in, double* out,
                            const double* x, const double* anm, const double* anp,
                            int nthreads)
{
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double* Lnm = (double*)malloc(sizeof(double) * N);
        double* Ln = (double*)malloc(sizeof(double) * N);
        double* Lnp = (double*)malloc(sizeof(double) * N);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (long l = 0; l < nlines; ++l) {
            const cplx* ui = (const cplx*)in + l;
            cplx* uo = (cplx*)out + l;
            for (int i = 0; i < N; ++i) {
                Lnm[i] = 1.0;
                Ln[i] = x[i] / anm[0];
                uo[(long)i * nlines].re = 0; uo[(long)i * nlines].im = 0;
                Lnp[i] = x[i] / anm[1] * Ln[i] - anp[1] / anm[1] * Lnm[i];
            }
            for (int i = 0; i < N; ++i) {
                double s1 = 1.0 / anm[i + 2], s2 = anp[i + 2] / anm[i + 2];
                cplx c = ui[(long)i * nlines];
                for (int j = 0; j < N; ++j) {
                    uo[(long)j * nlines].re += Lnm[j] * c.re;
                    uo[(long)j * nlines].im += Lnm[j] * c.im;
                    Lnm[j] = Ln[j];
                    Ln[j] = Lnp[j];
                    Lnp[j] = s1 * x[j] * Ln[j] - s2 * Lnm[j];
                }
            }
        }
        free(Lnm); free(Ln); free(Lnp);
    }
}
