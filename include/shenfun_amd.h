/*
 * shenfun_amd: C ABI of the MI355X (gfx950) spectral-Galerkin Helmholtz hot path.
 *
 * Plain pointers and sizes only.  Every array argument is a DEVICE pointer
 * (e.g. a PyTorch-ROCm tensor's data_ptr()), shapes and strides are int64 in
 * ELEMENTS of the array's own dtype, `stream` is a hipStream_t passed as
 * void* (NULL = default stream).  All calls are asynchronous on `stream`,
 * return SFA_OK (0) or a negative status, and leave a message for
 * sfa_last_error().  Arrays have ndim 1..3; `axis` is the transformed or
 * solved axis.  dtype is SFA_F64 (float64) or SFA_C128 (complex128,
 * interleaved re/im).
 *
 * Each entry point names the reference (shenfun 4.3.0) interface it replaces,
 * as path:line inside the shenfun source tree.
 */
#ifndef SHENFUN_AMD_H
#define SHENFUN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFA_F64 0
#define SFA_C128 1

#define SFA_OK 0
#define SFA_ERR_ARG (-1)
#define SFA_ERR_HIP (-2)
#define SFA_ERR_UNSUPPORTED (-3)
#define SFA_ERR_NOMEM (-4)

/* Library version (major*10000 + minor*100 + patch). */
int sfa_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char* sfa_last_error(void);

/* ------------------------------------------------------------------ */
/* Transforms along one axis (the FFTW planners' execute step)         */
/* ------------------------------------------------------------------ */

/* Complex-to-complex DFT along `axis`, out = scale * sum_j in_j e^{sign 2 pi i jk/n}.
 * Replaces the fftw.fftn / ifftn planners of fourier/bases.py:515-533 (C2C);
 * shenfun uses sign=-1, scale=1/N for the scalar product (:157-162) and
 * sign=+1, scale=1 for backward (normalise_idft=False, :533).  in == out is allowed. */
int sfa_fft_c2c(const void* in, void* out, int ndim, const int64_t* shape,
                const int64_t* in_strides, const int64_t* out_strides, int axis,
                int sign, double scale, void* stream);

/* Real-to-complex DFT along `axis` (real length shape[axis] = N; odd N runs
 * as a full-length complex FFT), out has N/2+1 entries along axis.  Replaces fftw.rfftn of
 * fourier/bases.py:249-265 (R2C scalar product, scale = 1/N). */
int sfa_fft_r2c(const double* in, void* out, int ndim, const int64_t* real_shape,
                const int64_t* in_strides, const int64_t* out_strides, int axis,
                double scale, void* stream);

/* Complex-to-real inverse DFT along `axis` (real length real_shape[axis] = N;
 * in has N/2+1 entries, the imaginary parts of entry 0 and, for even N, of
 * entry N/2 ignored; odd N runs as a full-length complex FFT of the Hermitian
 * extension).
 * Replaces fftw.irfftn of fourier/bases.py:380-407 (normalise_idft=False, scale 1). */
int sfa_fft_c2r(const void* in, double* out, int ndim, const int64_t* real_shape,
                const int64_t* in_strides, const int64_t* out_strides, int axis,
                double scale, void* stream);

/* Chebyshev-Gauss scalar product along `axis`: out_k = pi/(2N) REDFT10(in)_k,
 * then, if K2 != NULL (device, N-2 entries), the composite stencil
 * out_k += K2[k] out_{k+2} for k < N-2.  Replaces
 * chebyshev/bases.py:340-353 (Orthogonal._evaluate_scalar_product, DCTWrap
 * :100-115) and spectralbase.py:1462-1474 (CompositeBase stencil matvec). */
int sfa_cheb_scalar_product(const void* in, void* out, int dtype, int ndim,
                            const int64_t* shape, const int64_t* in_strides,
                            const int64_t* out_strides, int axis, const double* K2,
                            void* stream);

/* Chebyshev-Gauss synthesis along `axis`: if K2 != NULL first t = to_ortho(in)
 * (t_k = in_k + K2[k-2] in_{k-2}, BC slots N-2, N-1 ignored), then
 * out_j = sum_k t_k T_k(x_j) = 0.5 REDFT01(t)_j + t_0/2.  Replaces
 * chebyshev/bases.py:315-338 and spectralbase.py:1450-1460, 1616-1634. */
int sfa_cheb_backward(const void* in, void* out, int dtype, int ndim,
                      const int64_t* shape, const int64_t* in_strides,
                      const int64_t* out_strides, int axis, const double* K2,
                      void* stream);

/* Chebyshev-Gauss-Lobatto variants (x_j = cos(pi j/(N-1))): scalar product
 * (T_k, f)_w = REDFT00(f)_k pi/(2(N-1)) (chebyshev/bases.py:349-353) with the
 * composite stencil fused as above, and backward u_j = REDFT00(c)_j / 2 + c_0/2
 * + (-1)^j c_{N-1}/2 (chebyshev/bases.py:331-338) with to_ortho fused.  DCT-I
 * of length N runs as the complex FFT of the even extension (length 2(N-1)).
 * Same arguments as sfa_cheb_scalar_product / sfa_cheb_backward; N >= 2. */
int sfa_cheb_scalar_product_gl(const void* in, void* out, int dtype, int ndim, const int64_t* shape,
                               const int64_t* in_strides, const int64_t* out_strides, int axis,
                               const double* K2, void* stream);
int sfa_cheb_backward_gl(const void* in, void* out, int dtype, int ndim, const int64_t* shape,
                         const int64_t* in_strides, const int64_t* out_strides, int axis,
                         const double* K2, void* stream);

/* Padding / truncation along `axis` for dealiasing (padded spaces evaluate
 * their transform on M = floor(padding_factor N) points).  mode 0 pads (in:
 * the N coefficients, out: the M slots), replaces _padding_backward; mode 1
 * truncates (in: M, out: N), replaces _truncation_forward; mode 2 copies with
 * entries [z0, z1) zeroed (dealias_direct, same length).  kind 0: Chebyshev /
 * Legendre, the first N - nbc entries kept and the nbc boundary slots moved to
 * the end (spectralbase.py:1362-1389); kind 1: R2C, in/out complex, N the real
 * length, Nyquist halved on padding and doubled on truncation, imaginary part
 * dropped (fourier/bases.py:409-435); kind 2: C2C, wavenumbers 0..N/2 and the
 * last N/2 kept, Nyquist split between +-N/2 (fourier/bases.py:570-598).
 * Shapes must agree off the axis; in and out must not alias. */
int sfa_pad_axis(const void* in, void* out, int dtype, int ndim, const int64_t* in_shape,
                 const int64_t* in_strides, const int64_t* out_shape, const int64_t* out_strides,
                 int axis, int mode, int kind, int64_t N, int64_t nbc, int64_t z0, int64_t z1,
                 void* stream);

/* Nonhomogeneous boundary values of a composite space: for k < nrows,
 * dst(.., k, ..) += sum_m scales[m*scale_stride + line] * sum_j W[m][k][j] *
 * src(.., n_src - nbc + j, ..), the nbc boundary slots at the end of each src
 * line; W device (nmat, nrows, nbc) row-major, scales device (nmat,
 * scale_stride) or NULL (all 1), lines in C order over the non-axis dims.
 * Replaces BoundaryValues._add_to_orthogonal (tensorproductspace.py:2070-2089,
 * the lift added by backward: W = the boundary basis on the physical points),
 * _add_mass_rhs (:2091-2106, W = -boundary mass matrix) and the bc-matrix
 * right-hand-side corrections of Helmholtz.__call__ (chebyshev/la.py:307-311)
 * and SolverGeneric1ND (la.py:1212-1229).  src == dst is allowed when
 * nrows <= n_src - nbc (rows below the slots read). */
int sfa_boundary_update(const void* src, void* dst, int dtype, int ndim, const int64_t* src_shape,
                        const int64_t* src_strides, const int64_t* dst_shape,
                        const int64_t* dst_strides, int axis, int64_t nbc, int64_t nrows,
                        int64_t nmat, const double* W, const double* scales, int64_t scale_stride,
                        void* stream);

/* Pointwise product out = a * b of two contiguous arrays of n elements
 * (float64 or complex128): the physical-space step of the dealiased
 * convolution (TensorProductSpace.convolve tensorproductspace.py:637-666,
 * R2C/C2C.convolve(fast=True) fourier/bases.py:437-475, 600-640). */
int sfa_pointwise_mul(const void* a, const void* b, void* out, int dtype, int64_t n, void* stream);

/* y[r n + i] += a x[i] for i < n, r < nrep (float64, contiguous).  No single
 * reference counterpart: the fused 3D solve's nonhomogeneous boundary values
 * enter as u = u_0 + L, u_0 the solution with homogeneous values and L the
 * lift ((alpha - lap) L = 0 with the given values), which the reference forms
 * through its bc matrices and the boundary basis in backward
 * (chebyshev/la.py:303-307, tensorproductspace.py:2070-2089). */
int sfa_axpy(const double* x, double* y, double a, int64_t n, int64_t nrep, void* stream);

/* Direct convolution sums of two Fourier coefficient vectors (complex128,
 * contiguous): real = 0, C2C, N coefficients in FFT order, out 2N entries
 * (convolve_1D, optimization/cython/convolve.pyx:9-30); real = 1, R2C, N/2+1
 * coefficients of a real length-N signal, out N+1 entries (convolve_real_1D
 * convolve.pyx:34-75).  Nyquist entries halved for even N.  The fast=False
 * path of R2C/C2C.convolve. */
int sfa_convolve_direct(const void* u, const void* v, void* out, int64_t N, int real, void* stream);

/* Build the dense Legendre transform matrix on the device (the plan step of
 * the recursive transforms, spectralbase.py:484-491, 550-558).
 * analysis=1: M[k][j] = w_j (L_k(x_j) + K2[k] L_{k+2}(x_j))  (k < N-2 when
 *             K2 != NULL; rows N-2, N-1 keep w_j L_k(x_j)), shape (N, Nx).
 * analysis=0: M[j][k] = L_k(x_j) + K2[k] L_{k+2}(x_j) (k < N-2) for the
 *             composite synthesis (columns N-2, N-1 zero when K2 != NULL;
 *             plain L_k(x_j) when K2 == NULL), shape (Nx, N).
 * x, w (w unused for analysis=0), K2: device arrays.  L_k by the three-term
 * recurrence of transforms.pyx:219-323. */
int sfa_legendre_matrix(const double* x, const double* w, int64_t Nx, int64_t N,
                        const double* K2, int analysis, double* M, void* stream);

/* Dense transform matrix from the caller's recursion rows: the plan step of
 * the recursive seam lib.scalar_product(input, output, x, w, axis, a) /
 * lib.evaluate_expansion_all(input, output, x, axis, a) (spectralbase.py:
 * 484-491, 552-558; transforms.pyx:175-323).  a: device (a_rows, a_len)
 * row-major DIA rows of the recursion matrix, a_rows = 2 (anm, anp) or 3
 * (anm, ann, anp), x P_n = anm_n P_{n+1} + ann_n P_n + anp_n P_{n-1}, P_0 = 1,
 * a_len >= N + 2.  analysis=1: M[i][j] = P_i(x_j) w_j, shape (N, Nx);
 * analysis=0: M[j][i] = P_i(x_j), shape (Nx, N).  x (Nx), w (Nx): device. */
int sfa_recursion_matrix(const double* x, const double* w, int64_t Nx, const double* a,
                         int64_t a_rows, int64_t a_len, int64_t N, int analysis, double* M,
                         void* stream);

/* out(.., k, ..) = sum_j M[k][j] in(.., j, ..) along `axis` (M row-major
 * nout x shape[axis], FP64).  With the matrices above this replaces the
 * recursive Legendre kernels transforms.pyx:175-190 (scalar_product) and
 * :199-216 (evaluate_expansion_all).  in != out. */
int sfa_line_matmul(const void* in, void* out, int dtype, int ndim,
                    const int64_t* in_shape, const int64_t* in_strides,
                    const int64_t* out_strides, int axis, int64_t nout,
                    const double* M, void* stream);

/* Evaluation of a 1-3D expansion at scattered points: out[p] = sum over
 * (k0, k1, k2) of c[k0][k1][k2] E0[p][k0] E1[p][k1] E2[p][k2]; E_d device
 * complex128 (npts, n_d), each axis' basis functions at the points'
 * coordinate d (stencil, R2C weights and boundary functions folded in);
 * c contiguous complex128 (n0, n1, n2), n2 = 1 and E2 = ones for 2D, n1 = n2 = 1
 * for 1D.  Replaces TensorProductSpace.eval (tensorproductspace.py:668-760)
 * and SpectralBase.eval (spectralbase.py eval -> evaluate_basis_all). */
int sfa_tensor_eval(const void* c, int64_t n0, int64_t n1, int64_t n2, const void* E0,
                    const void* E1, const void* E2, void* out, int64_t npts, void* stream);

/* ------------------------------------------------------------------ */
/* Banded solves (la.py + optimization/cython/la.pyx)                  */
/* ------------------------------------------------------------------ */

/* In-place LU of nbatch TDMA systems, data (nbatch, 3, n) DIA rows
 * (offsets -2, 0, 2 as la.py:256-258).  Replaces TDMA_LU la.pyx:243-252. */
int sfa_tdma_lu(double* data, int64_t n, int64_t nbatch, void* stream);

/* In-place LU of nbatch PDMA systems, data (nbatch, 5, n) (offsets
 * -4,-2,0,2,4).  Replaces PDMA_LU la.pyx:213-241. */
int sfa_pdma_lu(double* data, int64_t n, int64_t nbatch, void* stream);

/* Solve in place along `axis` with LU'd DIA data.  Line l (C order over the
 * non-axis dims) uses data + l*data_line_stride (0 = one system for all
 * lines, TDMA_Solve la.pyx:77-91 / Solve_axis_2D/3D :429-460); lines
 * l < skip_lines are left untouched (SolverGeneric1ND_solve_data
 * la.pyx:371-426 with is_zero_index).  Only entries 0..n-1 along axis are
 * touched (n <= shape[axis]).  Kernels: TDMA_inner_solve_ptr :520-532,
 * PDMA_inner_solve_ptr :496-514, DiagMA_inner_solve_ptr :586-590. */
int sfa_tdma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                   const int64_t* strides, int axis, const double* data, int64_t n,
                   int64_t data_line_stride, int64_t skip_lines, void* stream);
int sfa_pdma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                   const int64_t* strides, int axis, const double* data, int64_t n,
                   int64_t data_line_stride, int64_t skip_lines, void* stream);
int sfa_diagma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                     const int64_t* strides, int axis, const double* d, int64_t n,
                     int64_t data_line_stride, int64_t skip_lines, void* stream);

/* General banded systems with offsets -nl..nu (0 <= nl <= 4, 1 <= nu <= 4):
 * Doolittle LU without pivoting in place on nbatch stacked DIA blocks of
 * (nl + 1 + nu) rows x n (A(i, j) at row nl + j - i, column j), then the
 * forward/backward substitution along axis like the other band solves.
 * Replaces the reference's general sparse fallback Solve (la.py:823-844,
 * chosen by matrixbase.py:439-474 for any other pattern) on the banded
 * systems of the Galerkin spaces; nonzero pivots required (SPD or
 * diagonally dominant matrices). */
int sfa_band_lu(double* data, int64_t n, int nl, int nu, int64_t nbatch, void* stream);
int sfa_band_solve(void* u, int dtype, int ndim, const int64_t* shape, const int64_t* strides,
                   int axis, const double* data, int64_t n, int64_t data_line_stride,
                   int64_t skip_lines, int nl, int nu, void* stream);

/* The other band patterns of SparseMatrix.get_solver (matrixbase.py:439-474),
 * DIA rows in offset order, same layout and line conventions as above.
 * LU in place, nbatch systems of (rows, n):
 *   TDMA_O   (-1, 0, 1), 3 rows, n >= 1: replaces TDMA_O_LU la.pyx:254-263
 *   FDMA     (-2, 0, 2, 4), 4 rows, n >= 4: replaces FDMA_LU la.pyx:199-211
 *   HeptaDMA (-4, -2, 0, 2, 4, 6, 8), 7 rows, n >= 8: replaces HeptaDMA_LU
 *            la.pyx:265-303
 * TwoDMA (0, 2) and ThreeDMA (0, 2, 4) are upper triangular: no LU. */
int sfa_tdma_o_lu(double* data, int64_t n, int64_t nbatch, void* stream);
int sfa_fdma_lu(double* data, int64_t n, int64_t nbatch, void* stream);
int sfa_heptadma_lu(double* data, int64_t n, int64_t nbatch, void* stream);

/* Solves in place along `axis` (arguments as sfa_tdma_solve).  Replace
 * TDMA_O_inner_solve_ptr la.pyx:538-549 / TDMA_O_Solve :93-107,
 * TwoDMA_inner_solve_ptr :555-563 / TwoDMA_Solve :45, ThreeDMA_inner_solve_ptr
 * :569-580 / ThreeDMA_Solve :29 (its +4 band is read from DIA row 2, where the
 * reference reads row 1), FDMA_inner_solve_ptr :596-611 / FDMA_Solve :165,
 * HeptaDMA_inner_solve_ptr :466-490 / HeptaDMA_Solve :181.  Minimum n: TDMA_O 1,
 * TwoDMA 2, ThreeDMA 4, FDMA 4, HeptaDMA 8. */
int sfa_tdma_o_solve(void* u, int dtype, int ndim, const int64_t* shape,
                     const int64_t* strides, int axis, const double* data, int64_t n,
                     int64_t data_line_stride, int64_t skip_lines, void* stream);
int sfa_twodma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                     const int64_t* strides, int axis, const double* data, int64_t n,
                     int64_t data_line_stride, int64_t skip_lines, void* stream);
int sfa_threedma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                       const int64_t* strides, int axis, const double* data, int64_t n,
                       int64_t data_line_stride, int64_t skip_lines, void* stream);
int sfa_fdma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                   const int64_t* strides, int axis, const double* data, int64_t n,
                   int64_t data_line_stride, int64_t skip_lines, void* stream);
int sfa_heptadma_solve(void* u, int dtype, int ndim, const int64_t* shape,
                       const int64_t* strides, int axis, const double* data, int64_t n,
                       int64_t data_line_stride, int64_t skip_lines, void* stream);

/* Banded matrix-vector product along `axis`: c_i (+)= scales[line] *
 * sum_r D[r][i] v_{i + offsets[r]} for i < nrows and 0 <= i + offsets[r] < ncols;
 * D device (noffs, nrows), the diagonals by row (D[r][i] = A[i][i + offsets[r]]),
 * offsets device (noffs int64), scales device per line (C order over the
 * non-axis dims) or NULL.  accumulate = 0 writes c (rows >= nrows zeroed),
 * 1 adds to rows < nrows.  Replaces SparseMatrix.matvec (matrixbase.py:87-144)
 * and TPMatrix.matvec (:1437-1449, one non-periodic axis).  v != c. */
int sfa_band_matvec(const void* v, void* c, int dtype, int ndim, const int64_t* v_shape,
                    const int64_t* v_strides, const int64_t* c_shape, const int64_t* c_strides,
                    int axis, int64_t nrows, int64_t ncols, int noffs, const int64_t* offsets,
                    const double* D, const double* scales, int accumulate, void* stream);

/* Chebyshev Helmholtz LU (alfa*A + beta*B), N basis size (M = N-2 unknowns).
 * A0 (M), A2 (M-2), A4 (M-4): stiffness diagonals 0, +2, +4; Bm2 (M-2), B0 (M),
 * B2 (M-2): mass diagonals (unscaled; the Neumann 1/j^2 column scaling and
 * k=0 gauge of la.pyx:735-749 are applied inside).  alfa/beta: per line
 * (ab_stride 1, C order over the non-axis dims) or one value (ab_stride 0).
 * d0, d1, d2, L: real arrays of `shape` with `f_strides` (shape[axis] = N),
 * entries beyond the reference's loop ranges are not written.
 * Replaces LU_Helmholtz la.pyx:708-843 (LU_Helmholtz_1D :717-763). */
int sfa_helmholtz_lu(int neumann, int64_t N, const double* A0, const double* A2,
                     const double* A4, const double* Bm2, const double* B0,
                     const double* B2, const double* alfa, const double* beta,
                     int64_t ab_stride, double* d0, double* d1, double* d2, double* L,
                     int ndim, const int64_t* shape, const int64_t* f_strides, int axis,
                     void* stream);

/* Chebyshev Helmholtz solve along `axis`: u = H^{-1} b using the factors above.
 * b may alias u.  Entries N-2, N-1 of u along axis are not written.
 * Replaces Solve_Helmholtz la.pyx:845-973 (Solve_Helmholtz_1D_ptr :871-912). */
int sfa_helmholtz_solve(const void* b, void* u, int dtype, int ndim, const int64_t* shape,
                        const int64_t* b_strides, const int64_t* u_strides, int axis,
                        int neumann, const double* d0, const double* d1,
                        const double* d2, const double* L, const int64_t* f_strides,
                        void* stream);

/* c = (alfa*A + beta*B) v along `axis` for the Chebyshev SD/SN Galerkin
 * matrices: A given by A0 (M) and its row-constant upper tail At (M-2,
 * At[k] = A[k][k+2j] for all j >= 1 after the SN column scaling), B by Bm2,
 * B0, B2.  Entries N-2, N-1 of c are set to 0.  Replaces Helmholtz_matvec
 * Matvec.pyx:656-720 and Helmholtz_Neumann_matvec :722-818.  v != c. */
int sfa_helmholtz_matvec(const void* v, void* c, int dtype, int ndim, const int64_t* shape,
                         const int64_t* v_strides, const int64_t* c_strides, int axis,
                         int neumann, const double* A0, const double* At,
                         const double* Bm2, const double* B0, const double* B2,
                         const double* alfa, const double* beta, int64_t ab_stride,
                         void* stream);

/* Chebyshev Galerkin biharmonic a0 S + alfa A + beta B on the ShenBiharmonic
 * basis along `axis` of an array of `shape` (N = shape[axis] >= 8, N - 4
 * unknowns in two parity chains).  diags: 11 device pointers, the diagonals
 * as the reference passes them: S[0], S[2], S[4], A[-2], A[0], A[2], B[-4],
 * B[-2], B[0], B[2], B[4] (a lower diagonal -k stored from its first entry,
 * D[i] = M(i + k, i)).  alfa/beta per line (ab_stride 1, C order over the
 * non-axis dims) or one value (0).  Factor arrays u0, u1, u2, l0, l1, ak, bk:
 * C-contiguous (2,) + shape with the axis length (N - 3) / 2 (the reference's
 * layout, chebyshev/la.py:480-492).  ak, bk may be NULL (else written too).
 * S's dense upper tail is rank two per row and kept so through the
 * elimination, so the LU is O(N) per chain.
 * Replaces LU_Biharmonic la.pyx:975-1138 (+ Biharmonic_factor_pr :1141-1233). */
int sfa_biharmonic_lu(int ndim, const int64_t* shape, int axis, double a0, const double* alfa,
                      const double* beta, int64_t ab_stride, const double* const* diags,
                      double* u0, double* u1, double* u2, double* l0, double* l1, double* ak,
                      double* bk, void* stream);
/* ak, bk from l0, l1 alone.  Replaces Biharmonic_factor_pr la.pyx:1141-1233. */
int sfa_biharmonic_factor(int ndim, const int64_t* shape, int axis, const double* l0,
                          const double* l1, double* ak, double* bk, void* stream);
/* u = (a0 S + alfa A + beta B)^-1 b along axis; b may alias u; entries
 * N-4 .. N-1 of u are not written.  Replaces Biharmonic_Solve la.pyx:1235-1310. */
int sfa_biharmonic_solve(const void* b, void* u, int dtype, int ndim, const int64_t* shape,
                         const int64_t* b_strides, const int64_t* u_strides, int axis, double a0,
                         const double* u0, const double* u1, const double* u2, const double* l0,
                         const double* l1, const double* ak, const double* bk, void* stream);
/* c = (a0 S + alfa A + beta B) v along axis (entries N-4 .. N-1 of c zero);
 * v != c.  Replaces Biharmonic_matvec Matvec.pyx:820-923. */
int sfa_biharmonic_matvec(const void* v, void* c, int dtype, int ndim, const int64_t* shape,
                          const int64_t* v_strides, const int64_t* c_strides, int axis, double a0,
                          const double* alfa, const double* beta, int64_t ab_stride,
                          const double* const* diags, void* stream);

/* Chebyshev Dirichlet Poisson A u = b (alpha = 0), A0 (M) diagonal, A2 (M-2)
 * row-constant tail; u /= scale unless |scale-1| < 1e-8.  b may alias u.
 * Replaces Poisson_Solve_ADD la.pyx:614-706 (chebyshev/la.py:11-63). */
int sfa_poisson_add_solve(const void* b, void* u, int dtype, int ndim, const int64_t* shape,
                          const int64_t* b_strides, const int64_t* u_strides, int axis,
                          const double* A0, const double* A2, double scale, void* stream);

/* ------------------------------------------------------------------ */
/* Fused pipelines                                                     */
/* ------------------------------------------------------------------ */

/* (alpha - lap) u = f on Chebyshev(SD or SN) x C2C x R2C, f and u physical
 * (Nx, Ny, Nz) C-contiguous float64: scalar product -> Helmholtz solve
 * (alfa = -1, beta = alpha + ky^2 + kz^2) -> backward.  Equivalent to the
 * reference sequence TensorProductSpace.scalar_product
 * (tensorproductspace.py:146-190) -> chebyshev.la.Helmholtz.__call__
 * (chebyshev/la.py:288-317) -> TensorProductSpace.backward
 * (tensorproductspace.py:46-96).
 *
 * Planning (the counterpart of chebyshev.la.Helmholtz.__init__, which
 * factorises once per solver, chebyshev/la.py:223-276): sfa_helmholtz3d_plan
 * fills a device buffer of sfa_helmholtz3d_plan_bytes() bytes with, for every
 * distinct beta, the LU recurrence state at 16 segment starts of each parity
 * chain (the solve regenerates the factors of a segment from its start
 * state), and the line -> entry map, asynchronously on `stream` (the line
 * map and the distinct ky^2 + kz^2 of a (Ny, Nz) are built on the host and
 * uploaded once per process and device, a blocking step of the first plan
 * of that shape).  plan may be NULL in sfa_helmholtz3d, in which case the
 * factors are computed on the fly (wavefront scan of the LU recurrence).
 * precision: 0 = FP64 throughout.  1 = mixed: after the FP64 R2C of f, the
 * kz planes [0, fp64_planes) stay complex128 with FP64 transforms and the
 * planes [fp64_planes, Nz/2] are stored and transformed as complex64; the
 * Helmholtz solve and the final C2R to u are FP64.  fp64_planes <= 0 picks
 * 17 (kz < 17 in FP64); the value is clamped to [2, Nz/2 - 1].  The error of
 * the complex64 part is about 6e-8 times the largest coefficient above the
 * cut relative to the largest of f, so the cut is a wavenumber, not a
 * fraction of Nz.  precision 1 applies no check of that error: the guarded
 * form (sfa_helmholtz3d_guarded) is the safe one.  Mixed precision needs a
 * plan and power-of-two Nx in {64, 128, 256, 512}, Ny in [16, 512], Nz/2 in
 * [16, 256].
 * work: device workspace of at least sfa_helmholtz3d_workspace() bytes (the
 * complex workspace plus a small statistics area used by the guarded form). */
int64_t sfa_helmholtz3d_plan_bytes(int64_t Nx, int64_t Ny, int64_t Nz);
int sfa_helmholtz3d_plan(int64_t Nx, int64_t Ny, int64_t Nz, double alpha, int neumann,
                         void* plan, int64_t plan_bytes, void* stream);
/* Forget a plan buffer registered by sfa_helmholtz3d_plan (call before its
 * memory is freed or reused: a plan is recognised by its address).  No
 * reference counterpart (chebyshev.la.Helmholtz's factors are Python-owned). */
int sfa_helmholtz3d_plan_release(const void* plan);
/* Register with a plan a lift L (Nx * Ny * Nz doubles on the device, the
 * physical grid, 16-byte aligned; NULL clears it) that every later solve with
 * that plan (sfa_helmholtz3d, _many, _guarded, _noalloc, _stage_times) adds
 * to u, inside the final z C2R's stores where its fast kernel runs, else by
 * a pass over u: nonhomogeneous boundary values as u = u_0 + L, u_0 the
 * solve with homogeneous values and L the solution of (alpha - lap) L = 0
 * with the given values, the same discrete solution as the reference's bc
 * matrices (chebyshev/la.py:303-307) with the boundary basis added in
 * backward (tensorproductspace.py:2070-2089).  The caller keeps L alive while
 * it is registered. */
int sfa_helmholtz3d_set_lift(const void* plan, const double* lift);
int64_t sfa_helmholtz3d_workspace(int64_t Nx, int64_t Ny, int64_t Nz, int precision);
int sfa_helmholtz3d(const double* f, double* u, int64_t Nx, int64_t Ny, int64_t Nz,
                    double alpha, int neumann, int precision, int fp64_planes, const void* plan,
                    void* work, int64_t work_bytes, void* stream);
/* nbatch solves of one shape, alpha and bc in one pass of launches: f and u
 * hold nbatch (Nx, Ny, Nz) arrays back to back, work at least nbatch x
 * sfa_helmholtz3d_workspace() bytes; each array's result is bitwise equal to
 * its sfa_helmholtz3d solve.  Same reference sequence as sfa_helmholtz3d per
 * array; no reference counterpart for the batch itself (shenfun solves one
 * TensorProductSpace array per call). */
int sfa_helmholtz3d_many(const double* f, double* u, int64_t nbatch, int64_t Nx, int64_t Ny,
                         int64_t Nz, double alpha, int neumann, int precision, int fp64_planes,
                         const void* plan, void* work, int64_t work_bytes, void* stream);

/* Guarded mixed precision: the mixed pipeline of sfa_helmholtz3d_many with
 * the first complex64 plane kh_hint (<= 0: 17), whose x-pass also sums, per
 * band of 16 kz planes, the energy of its input lines (f after the z and y
 * transforms) and of u's Chebyshev coefficients.  After the x-pass the host reads these sums (the
 * inverse passes are already queued) and predicts the relative max error of
 * the complex64 part: 64 eps32 sqrt((w^2 E_f,hi + E_u,hi) / E_u), E_hi the
 * energies of the planes at and above the cut, w = 1 / (alpha + Kh^2) the
 * solve's gain there.  If it exceeds tol the solve is redone from f with the
 * smallest cut 16m + 1 whose prediction passes, or in FP64 when none does
 * (f is left unchanged).  Out (each may be NULL): kh_used = first complex64
 * plane of the returned u (Nz/2 + 1: all FP64), kh_next = the smallest
 * passing cut for this f (0: none), err_pred = the prediction for kh_used (0
 * for FP64).  Synchronises with the host once per call, after the x-pass:
 * the first workgroup of the inverse y pass writes the sums and a sequence
 * number to pinned host memory, and the calling thread spins on it (querying
 * the stream once per millisecond for errors) while the inverse passes run.
 * Calls from one host thread to one device share that buffer and run one at
 * a time (each returns after its wait).  work: at least nbatch x the complex workspace + the statistics area
 * (sfa_helmholtz3d_workspace() bytes for nbatch = 1).  Same reference sequence
 * as sfa_helmholtz3d; the guard has no reference counterpart (shenfun is FP64
 * only). */
int sfa_helmholtz3d_guarded(const double* f, double* u, int64_t nbatch, int64_t Nx, int64_t Ny,
                            int64_t Nz, double alpha, int neumann, int kh_hint, double tol,
                            const void* plan, void* work, int64_t work_bytes, int* kh_used,
                            int* kh_next, double* err_pred, void* stream);

/* No-allocation variant of sfa_helmholtz3d (precision 0 or 1, no guard).
 * It has no counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a file-scope device
 * array of sfa_helmholtz3d_noalloc_max_bytes() bytes (32 MiB) per GPU, held
 * until the process ends; the call makes no hipMalloc/hipFree and builds no
 * table: it needs a plan from sfa_helmholtz3d_plan (which also builds the
 * device tables) on the same device.  Shapes: Nx in {64, 128, 256, 512},
 * power-of-two Ny in [16, 512] and Nz in [32, 512], whose workspace
 * (sfa_helmholtz3d_workspace() minus the statistics area) fits the array
 * (e.g. 128^3, 256 x 128 x 64); larger shapes are refused with
 * SFA_ERR_ARG.  Every call on one GPU uses that GPU's one array: calls that
 * overlap on one GPU (other streams or host threads) race.  The result is
 * bitwise equal to sfa_helmholtz3d's. */
int64_t sfa_helmholtz3d_noalloc_max_bytes(void);
int sfa_helmholtz3d_noalloc(const double* f, double* u, int64_t Nx, int64_t Ny, int64_t Nz,
                            double alpha, int neumann, int precision, int fp64_planes,
                            const void* plan, void* stream);

/* Profiling variant of sfa_helmholtz3d: runs the pipeline (precision and
 * fp64_planes as there) nrep times with hipEvents recorded on `stream` between
 * its stages and adds each stage's milliseconds into stage_ms[5] (z R2C, y
 * C2C, x-pass, y C2C inverse, z C2R; in mixed precision a stage covers its
 * complex128 and complex64 launches; precision 2 times the guarded form's
 * kernels, with their statistics, at the cut fp64_planes and without the
 * host check).  Synchronises the stream.  No reference
 * counterpart: it exists to attribute the pipeline's time to its kernels. */
int sfa_helmholtz3d_stage_times(const double* f, double* u, int64_t Nx, int64_t Ny, int64_t Nz,
                                double alpha, int neumann, int precision, int fp64_planes,
                                const void* plan, void* work, int64_t work_bytes, int nrep,
                                float* stage_ms, void* stream);

/* ------------------------------------------------------------------ */
/* Batched 2D solves                                                   */
/* ------------------------------------------------------------------ */

/* nprob independent (alpha_i - lap) u_i = f_i on Legendre(SD: bc 0, SN: bc 1)
 * x Fourier R2C, sizes Nx_i x Ny_i (both even, Nx in [8, 1024], Ny in
 * [4, 1024]), f_i and u_i row-major float64 at element offsets f_off[i],
 * u_off[i] of F (f_elems) and U (u_elems).  One call of sfa_batch2d_solve
 * replaces, per problem, the reference sequence
 * TensorProductSpace.scalar_product (tensorproductspace.py:146-190; Legendre
 * recursive scalar product spectralbase.py:550-558, R2C fourier/bases.py:157-162)
 * -> SolverGeneric1ND.__call__ (la.py:1390; per-mode TDMA of
 * A + (alpha + ky^2) B, legendre/matrices.py:110-162, 271-315)
 * -> TensorProductSpace.backward (tensorproductspace.py:46-96).  For the
 * Neumann ky = 0 mode with alpha = 0 the solve pins u_0 = 0.
 *
 * Planning: sfa_batch2d_plan fills a device buffer of sfa_batch2d_plan_bytes()
 * with the problem table, a static largest-first schedule of the problems
 * over nblocks workgroups, and the composite Legendre and real DFT matrices
 * per distinct Nx / Ny; it synchronises `stream`.  sfa_batch2d_solve runs all
 * problems in one persistent launch of nblocks workgroups with tile variant
 * `variant` (0 <= variant < sfa_batch2d_variants(); its block tile BM x BN and
 * K step BK are reported as bm[] = BM, bn[] = 1000 BN + BK).  work: sfa_batch2d_workspace() bytes.  F and U must
 * not overlap.  No reference counterpart exists for the batching itself: the
 * reference solves one problem per call. */
int64_t sfa_batch2d_plan_bytes(int nprob, const int* Nx, const int* Ny, const int* bc, int nblocks);
int sfa_batch2d_plan(int nprob, const int* Nx, const int* Ny, const double* alpha, const int* bc,
                     const int64_t* f_off, const int64_t* u_off, int64_t f_elems, int64_t u_elems,
                     int nblocks, void* plan, int64_t plan_bytes, void* stream);
int64_t sfa_batch2d_workspace(int nprob, const int* Nx, const int* Ny, int nblocks);
int sfa_batch2d_variants(int* bm, int* bn, int cap);
int sfa_batch2d_solve(const double* F, double* U, const void* plan, void* work, int64_t work_bytes,
                      int variant, void* stream);
/* Forget a plan's host-side registration (before its buffer is freed). */
int sfa_batch2d_release(const void* plan);

/* The same batches with the matrices held by the caller (a resident cache
 * shared by many batches, as a serving loop keeps it): sfa_batch2d_leg_matrices
 * writes Phi (Nx x (Nx-2)) then P ((Nx-2) x Nx) of the composite Legendre
 * basis (bc 0 SD, 1 SN) into sfa_batch2d_leg_elems(Nx) doubles at out;
 * sfa_batch2d_dft_matrices writes Df (Ny x (Ny+2)) then Di ((Ny+2) x Ny) into
 * sfa_batch2d_dft_elems(Ny) doubles.  sfa_batch2d_plan_ext builds a plan of
 * sfa_batch2d_plan_ext_bytes() bytes (problem table and schedule only) whose
 * problem i uses the matrices at mats + phi_off[i] and mats + df_off[i]
 * (mats_elems doubles in all); mats must stay allocated while the plan is
 * used.  Synchronises `stream`. */
int64_t sfa_batch2d_leg_elems(int Nx);
int64_t sfa_batch2d_dft_elems(int Ny);
int sfa_batch2d_leg_matrices(int Nx, int bc, double* out, void* stream);
int sfa_batch2d_dft_matrices(int Ny, double* out, void* stream);
int64_t sfa_batch2d_plan_ext_bytes(int nprob, int nblocks);
int sfa_batch2d_plan_ext(int nprob, const int* Nx, const int* Ny, const double* alpha, const int* bc,
                         const int64_t* f_off, const int64_t* u_off, int64_t f_elems,
                         int64_t u_elems, const double* mats, const int64_t* phi_off,
                         const int64_t* df_off, int64_t mats_elems, int nblocks, void* plan,
                         int64_t plan_bytes, void* stream);

/* Tile-parallel mode of the same batched solve, for batches with fewer
 * problems than CUs (one large problem would otherwise run on one workgroup):
 * each of the four contractions runs as one launch over all (problem, 32 x 32
 * output tile) tasks and the per-ky solve as one launch with a workgroup per
 * problem.  Same problem table and external matrices as sfa_batch2d_plan_ext
 * (tensorproductspace.py:146-190, la.py:1390, tensorproductspace.py:46-96 per
 * problem); every output is bitwise equal to sfa_batch2d_solve's.  The plan
 * takes sfa_batch2d_wide_plan_bytes() bytes, the workspace
 * sfa_batch2d_wide_workspace() bytes.  sfa_batch2d_wide_plan synchronises
 * `stream`; release the plan with sfa_batch2d_release. */
int64_t sfa_batch2d_wide_plan_bytes(int nprob, const int* Nx, const int* Ny);
int64_t sfa_batch2d_wide_workspace(int nprob, const int* Nx, const int* Ny);
int sfa_batch2d_wide_plan(int nprob, const int* Nx, const int* Ny, const double* alpha, const int* bc,
                          const int64_t* f_off, const int64_t* u_off, int64_t f_elems,
                          int64_t u_elems, const double* mats, const int64_t* phi_off,
                          const int64_t* df_off, int64_t mats_elems, void* plan, int64_t plan_bytes,
                          void* stream);
int sfa_batch2d_wide_solve(const double* F, double* U, const void* plan, void* work,
                           int64_t work_bytes, void* stream);

/* Self-check of the FP64 MFMA 16x16x4 lane layout the batched GEMMs assume:
 * writes the 64 x 4 accumulator registers of one MFMA of A = i + 100 k and
 * B = j + 1000 k into out (256 doubles).  No reference counterpart. */
int sfa_mfma_f64_layout_probe(double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SHENFUN_AMD_H */
