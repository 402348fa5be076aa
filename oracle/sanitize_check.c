/* Memory and undefined-behaviour check of the oracle's C restatement
 * (banded.c): built with -fsanitize=address,undefined by
 * tests/test_oracle_sanitize.py and run on small random inputs of every
 * function, each array allocated at exactly the size the function's
 * contract names, serial and with two OpenMP threads.  Test infrastructure
 * only.  Exit status 0 and finite outputs: clean. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

void orc_helmholtz_solve_axis0(int N, long nlines, const double* fk, double* u, int neumann,
                               const double* d0, const double* d1, const double* d2,
                               const double* L, int nthreads);
void orc_tdma_solve_modes_axis0(int n, long nlines, double* u, const double* data, long skip,
                                int nthreads);
void orc_leg_scalar_product_axis0(int N, long nlines, const double* in, double* out,
                                  const double* x, const double* w, const double* anm,
                                  const double* anp, int nthreads);
void orc_leg_backward_axis0(int N, long nlines, const double* in, double* out, const double* x,
                            const double* anm, const double* anp, int nthreads);

static unsigned long long rng = 88172645463325252ull;
static double urand(void)
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (double)(rng >> 11) / 9007199254740992.0;
}

static double* alloc(size_t n, double lo, double hi)
{
    double* p = (double*)malloc(n * sizeof(double));
    if (!p) exit(3);
    for (size_t i = 0; i < n; ++i) p[i] = lo + (hi - lo) * urand();
    return p;
}

static int finite_all(const double* p, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(p[i])) return 0;
    return 1;
}

int main(void)
{
    int bad = 0;
    for (int nt = 1; nt <= 2; ++nt) {
        for (int N = 6; N <= 17; N += 11) {
            const long nl = 5;
            /* Helmholtz: fk, u complex (N, nl); factors (N, nl) doubles */
            for (int nm = 0; nm <= 1; ++nm) {
                double* fk = alloc((size_t)2 * N * nl, -1, 1);
                double* u = alloc((size_t)2 * N * nl, 0, 0);
                double* d0 = alloc((size_t)N * nl, 2, 3);
                double* d1 = alloc((size_t)N * nl, -0.2, 0.2);
                double* d2 = alloc((size_t)N * nl, -0.1, 0.1);
                double* L = alloc((size_t)N * nl, -0.3, 0.3);
                orc_helmholtz_solve_axis0(N, nl, fk, u, nm, d0, d1, d2, L, nt);
                bad |= !finite_all(u, (size_t)2 * (N - 2) * nl);
                free(fk); free(u); free(d0); free(d1); free(d2); free(L);
            }
            /* TDMA per mode: u complex (n, nl), data (nl, 3, n) */
            {
                const int n = N;
                double* u = alloc((size_t)2 * n * nl, -1, 1);
                double* data = alloc((size_t)nl * 3 * n, 0.1, 0.3);
                for (long j = 0; j < nl; ++j)
                    for (int i = 0; i < n; ++i) data[(size_t)j * 3 * n + n + i] = 2.0 + urand();
                orc_tdma_solve_modes_axis0(n, nl, u, data, 1, nt);
                bad |= !finite_all(u, (size_t)2 * n * nl);
                free(u); free(data);
            }
            /* Legendre recursion transforms: in/out complex (N, nl), x, w (N),
               anm, anp (N + 3) */
            {
                double* in = alloc((size_t)2 * N * nl, -1, 1);
                double* out = alloc((size_t)2 * N * nl, 0, 0);
                double* back = alloc((size_t)2 * N * nl, 0, 0);
                double* x = alloc((size_t)N, -1, 1);
                double* w = alloc((size_t)N, 0, 0.5);
                double* anm = alloc((size_t)N + 3, 1, 2);
                double* anp = alloc((size_t)N + 3, -0.5, 0.5);
                orc_leg_scalar_product_axis0(N, nl, in, out, x, w, anm, anp, nt);
                orc_leg_backward_axis0(N, nl, out, back, x, anm, anp, nt);
                bad |= !finite_all(out, (size_t)2 * N * nl) || !finite_all(back, (size_t)2 * N * nl);
                free(in); free(out); free(back); free(x); free(w); free(anm); free(anp);
            }
        }
    }
    if (bad) {
        fprintf(stderr, "non-finite output\n");
        return 2;
    }
    printf("oracle sanitize check: clean\n");
    return 0;
}
