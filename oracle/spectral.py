"""numpy/scipy FP64 restatement of shenfun 4.3.0's hot path (test oracle only).

Every function cites the reference file:line it restates (paths relative to
the shenfun source tree).  All arithmetic is float64 / complex128.  Arrays are
numpy arrays; the axis being transformed or solved is given explicitly, like
the reference's ``axis`` arguments.

Conventions restated (Appendix of the reference's own docs and code):

* Chebyshev Gauss (GC) nodes x_j = cos(pi(2j+1)/(2N)), descending, w_j = pi/N
  (chebyshev/bases.py:213-216 via numpy.polynomial.chebyshev.chebgauss);
  Gauss-Lobatto (GL) nodes x_j = cos(pi j/(N-1)), descending, w_j = pi/(N-1)
  with both ends halved (chebyshev/bases.py:202-207).
* Legendre Gauss (LG) nodes ascending, Gauss weights
  (legendre/bases.py:121-126, legendre/fastgl/fastgl_wrap.pyx:22-29).
* Fourier nodes 2 pi j / N, scalar product = FFT / N, backward = unnormalised
  inverse FFT (fourier/bases.py:73-82, 157-162, 380-407, 526-533).
"""
from __future__ import annotations

import numpy as np
import scipy.fft

PI = np.pi

# --------------------------------------------------------------------------
# Quadrature
# --------------------------------------------------------------------------


def cheb_points_weights(N, quad='GC'):
    """Chebyshev-Gauss (GC) or Gauss-Lobatto (GL) points (descending) and weights.

    chebyshev/bases.py:213-216 (``n_cheb.chebgauss``): x_j = cos(pi(2j+1)/(2N)),
    w_j = pi/N; chebyshev/bases.py:202-207 (GL, ``-chebpts2``): x_j = cos(pi j/(N-1)),
    w_j = pi/(N-1), halved at both ends.
    """
    j = np.arange(N, dtype=float)
    if quad == 'GL':
        w = np.full(N, PI / (N - 1))
        w[0] /= 2
        w[-1] /= 2
        return np.cos(PI * j / (N - 1)), w
    x = np.cos(PI * (2.0 * j + 1.0) / (2.0 * N))
    w = np.full(N, PI / N)
    return x, w


def _legendre_and_derivative(n, x):
    """P_n(x) and P_n'(x) by the three-term recurrence (in x's precision)."""
    p0 = np.ones_like(x)
    if n == 0:
        return p0, np.zeros_like(x)
    p1 = x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1)
    return p1, dp


def leg_points_weights(N):
    """Legendre-Gauss points (ascending) and weights.

    Restates the *result* of legendre/bases.py:121-126 ->
    legendre/fastgl/fastgl_wrap.pyx:22-29 (nodes x_i = GLPair(N, N-i).x(),
    i.e. ascending, with Gauss weights).  Own method: cosine initial guess,
    Newton on the three-term recurrence carried out in x87 extended precision
    (np.longdouble), weights 2/((1-x^2) P_N'(x)^2) in the same precision, then
    rounded to float64 and symmetrised.  Pinned against fastgl itself
    (compiled from the reference sources into oracle/_ref) to a few ulp.
    """
    if N == 1:
        return np.array([0.0]), np.array([2.0])
    ld = np.longdouble
    k = np.arange(1, N + 1, dtype=ld)
    pi = ld(np.pi)
    x = np.cos(pi * (k - ld(0.25)) / (N + ld(0.5))) * (1 - (N - 1) / (ld(8.0) * N**3))
    for _ in range(100):
        p, dp = _legendre_and_derivative(N, x)
        dx = p / dp
        x = x - dx
        if np.max(np.abs(dx)) < ld(1e-19):
            break
    p, dp = _legendre_and_derivative(N, x)
    x = x - p / dp
    _, dp = _legendre_and_derivative(N, x)
    w = 2 / ((1 - x * x) * dp * dp)
    order = np.argsort(x)
    x = x[order].astype(float)
    w = w[order].astype(float)
    x = 0.5 * (x - x[::-1])
    w = 0.5 * (w + w[::-1])
    return x, w


def fourier_points(N):
    """fourier/bases.py:76: x_j = 2 pi j / N."""
    return np.arange(N, dtype=float) * 2 * PI / N


def r2c_wavenumbers(N):
    """fourier/bases.py:334-342: np.fft.rfftfreq(N, 1/N) (Nyquist +N/2)."""
    return np.fft.rfftfreq(N, 1.0 / N).astype(int)


def c2c_wavenumbers(N):
    """fourier/bases.py:535-543: np.fft.fftfreq(N, 1/N) (index N/2 -> -N/2)."""
    return np.fft.fftfreq(N, 1.0 / N).astype(int)


# --------------------------------------------------------------------------
# Helpers
# --------------------------------------------------------------------------


def _bcast(v, ndim, axis):
    s = [np.newaxis] * ndim
    s[axis] = slice(None)
    return v[tuple(s)]


def _take(a, sl, axis):
    s = [slice(None)] * a.ndim
    s[axis] = sl
    return tuple(s)


# --------------------------------------------------------------------------
# Orthogonal transforms along one axis
# --------------------------------------------------------------------------


def _dct(a, typ, axis):
    # chebyshev/bases.py:100-115 (DCTWrap): complex data -> DCT of real and
    # imaginary parts separately.  scipy.fft.dct(norm=None) has the same
    # unnormalised definition as FFTW's REDFT10 (type 2) / REDFT01 (type 3) /
    # REDFT00 (type 1).
    if np.iscomplexobj(a):
        return (scipy.fft.dct(a.real, type=typ, axis=axis)
                + 1j * scipy.fft.dct(a.imag, type=typ, axis=axis))
    return scipy.fft.dct(a, type=typ, axis=axis)


def cheb_scalar_product(f, axis=0, quad='GC'):
    """(T_k, f)_w with GC or GL quadrature.

    chebyshev/bases.py:344-347: GC out = REDFT10(f) * pi/(2N);
    chebyshev/bases.py:349-353: GL out = REDFT00(f) * pi/(2(N-1)).
    """
    N = f.shape[axis]
    if quad == 'GL':
        return _dct(f, 1, axis) * (PI / (2 * (N - 1)))
    return _dct(f, 2, axis) * (PI / (2 * N))


def cheb_backward(c, axis=0, quad='GC'):
    """u_j = sum_k c_k T_k(x_j) at GC or GL nodes.

    chebyshev/bases.py:326-329: GC out = 0.5*REDFT01(c) + c_0/2;
    chebyshev/bases.py:331-338: GL out = 0.5*REDFT00(c) + c_0/2 + (-1)^j c_{N-1}/2.
    """
    if quad == 'GL':
        N = c.shape[axis]
        out = 0.5 * _dct(c, 1, axis)
        out += c[_take(c, slice(0, 1), axis)] / 2
        sgn = _bcast(np.where(np.arange(N) % 2 == 0, 0.5, -0.5), c.ndim, axis)
        return out + sgn * c[_take(c, slice(N - 1, N), axis)]
    out = 0.5 * _dct(c, 3, axis)
    out += c[_take(c, slice(0, 1), axis)] / 2
    return out


def _leg_recursion_rows(N):
    """DIA rows (anm, anp) of legendre/bases.py:239-242's recursion matrix.

    get_recursion_matrix(M, M).diags('dia').data has offsets (-1, +1):
    anm[j] = (j+1)/(2j+1)   (entry [j+1, j]),
    anp[j] = j/(2j+1)       (entry [j-1, j]).
    """
    j = np.arange(N, dtype=float)
    anm = (j + 1) / (2 * j + 1)
    anp = j / (2 * j + 1)
    return anm, anp


def leg_scalar_product(f, axis=0, x=None, w=None):
    """s_k = sum_j f_j w_j L_k(x_j), the recursive kernel.

    optimization/cython/transforms.pyx:270-323 (_scalar_product_ptr), called
    from spectralbase.py:550-558 with a = recursion matrix (N+3, N+3) DIA data
    (2 rows => ann = 0).  Same recurrence and update order.
    """
    N = f.shape[axis]
    if x is None:
        x, w = leg_points_weights(N)
    anm, anp = _leg_recursion_rows(N + 3)
    fm = np.moveaxis(f, axis, 0)
    out = np.zeros_like(fm)
    xb = _bcast(x, fm.ndim, 0)
    wf = fm * _bcast(w, fm.ndim, 0)
    Lnm = np.ones_like(xb)
    Ln = (xb - 0.0) / anm[0]
    Lnp = (xb - 0.0) / anm[1] * Ln - anp[1] / anm[1] * Lnm
    for i in range(N):
        s1 = 1.0 / anm[i + 2]
        s2 = anp[i + 2] / anm[i + 2]
        out[i] = np.sum(Lnm * wf, axis=0)
        Lnm = Ln
        Ln = Lnp
        Lnp = s1 * xb * Ln - s2 * Lnm
    return np.moveaxis(out, 0, axis)


def leg_backward(c, axis=0, x=None):
    """u_j = sum_k c_k L_k(x_j), the recursive kernel.

    optimization/cython/transforms.pyx:219-267 (_evaluate_expansion_all_ptr),
    called from spectralbase.py:484-491.
    """
    N = c.shape[axis]
    if x is None:
        x, _ = leg_points_weights(N)
    Nx = x.shape[0]
    anm, anp = _leg_recursion_rows(N + 3)
    cm = np.moveaxis(c, axis, 0)
    out = np.zeros((Nx,) + cm.shape[1:], dtype=cm.dtype)
    xb = _bcast(x, cm.ndim, 0)
    Lnm = np.ones_like(xb)
    Ln = xb / anm[0]
    Lnp = xb / anm[1] * Ln - anp[1] / anm[1] * Lnm
    for i in range(N):
        s1 = 1.0 / anm[i + 2]
        s2 = anp[i + 2] / anm[i + 2]
        out += Lnm * cm[i:i + 1]
        Lnm = Ln
        Ln = Lnp
        Lnp = s1 * xb * Ln - s2 * Lnm
    return np.moveaxis(out, 0, axis)


# --------------------------------------------------------------------------
# Composite (Shen) bases: stencils
# --------------------------------------------------------------------------


def stencil_k2(family, bc, N):
    """Second stencil diagonal K2(k), k = 0..N-3 (phi_k = P_k + K2(k) P_{k+2}).

    Chebyshev SD chebyshev/bases.py:546 {0: 1, 2: -1};
    Chebyshev SN chebyshev/bases.py:937 {0: 1, 2: -(n/(n+2))^2};
    Legendre SD legendre/bases.py:374 {0: 1, 2: -1};
    Legendre SN legendre/bases.py:502 {0: 1, 2: -n(n+1)/(n^2+5n+6)}.
    """
    k = np.arange(N - 2, dtype=float)
    if bc == 'D':
        return -np.ones(N - 2)
    if family == 'chebyshev':
        return -(k / (k + 2)) ** 2
    return -k * (k + 1) / (k * k + 5 * k + 6)


def composite_scalar_product(t, K2, axis=0):
    """Scalar product in the composite basis from the orthogonal one.

    spectralbase.py:1462-1474 + 1504-1526: out = K @ t with K the N x N
    stencil matrix (diagonals 0 and +2).  Rows N-2, N-1 keep t (the scalar
    main diagonal 1 is not truncated, spectralbase.py:1519-1523).
    """
    tm = np.moveaxis(t, axis, 0)
    N = tm.shape[0]
    out = tm.copy()
    out[:N - 2] += _bcast(K2, tm.ndim, 0) * tm[2:]
    return np.moveaxis(out, 0, axis)


def to_ortho(u, K2, axis=0):
    """Composite -> orthogonal coefficients, homogeneous BCs.

    spectralbase.py:1616-1634: t[k] += u[k], t[k+2] += K2(k) u[k] for
    k < N-2; BC slots N-2, N-1 contribute nothing.
    """
    um = np.moveaxis(u, axis, 0)
    N = um.shape[0]
    t = np.zeros_like(um)
    t[:N - 2] += um[:N - 2]
    t[2:] += _bcast(K2, um.ndim, 0) * um[:N - 2]
    return np.moveaxis(t, 0, axis)


# --------------------------------------------------------------------------
# Per-basis forward/backward (scalar product + synthesis)
# --------------------------------------------------------------------------


def basis_scalar_product(family, bc, f, axis=0, quad='GC'):
    """SpectralBase.scalar_product (spectralbase.py:200-233) for one axis.

    family in {'chebyshev', 'legendre', 'fourier-r2c', 'fourier-c2c'};
    bc in {None, 'D', 'N'} (None = orthogonal basis); quad 'GC' or 'GL'
    (Chebyshev only).
    """
    N = f.shape[axis]
    if family == 'fourier-r2c':
        return np.fft.rfft(f, axis=axis) / N
    if family == 'fourier-c2c':
        return np.fft.fft(f, axis=axis) / N
    if family == 'chebyshev':
        t = cheb_scalar_product(f, axis, quad)
    else:
        t = leg_scalar_product(f, axis)
    if bc is None:
        return t
    return composite_scalar_product(t, stencil_k2(family, bc, N), axis)


def basis_backward(family, bc, c, axis=0, N=None, quad='GC'):
    """SpectralBase.backward (spectralbase.py:265-313) for one axis."""
    if family == 'fourier-r2c':
        n = N if N is not None else 2 * (c.shape[axis] - 1)
        # fourier/bases.py:407 irfftn(normalise_idft=False)
        return np.fft.irfft(c, n=n, axis=axis) * n
    if family == 'fourier-c2c':
        # fourier/bases.py:533 ifftn(normalise_idft=False)
        return np.fft.ifft(c, axis=axis) * c.shape[axis]
    Nc = c.shape[axis]
    if bc is not None:
        c = to_ortho(c, stencil_k2(family, bc, Nc), axis)
    if family == 'chebyshev':
        return cheb_backward(c, axis, quad)
    return leg_backward(c, axis)


def tps_scalar_product(f, bases):
    """TensorProductSpace.scalar_product: last axis first.

    tensorproductspace.py:146-190 with the axis order of :423-447.
    ``bases`` is a list of (family, bc) per axis.
    """
    out = f
    for axis in range(f.ndim - 1, -1, -1):
        fam, bc = bases[axis]
        out = basis_scalar_product(fam, bc, out, axis)
    return out


def tps_backward(c, bases, shape):
    """TensorProductSpace.backward: first axis first (tensorproductspace.py:46-96)."""
    out = c
    for axis in range(c.ndim):
        fam, bc = bases[axis]
        out = basis_backward(fam, bc, out, axis, N=shape[axis])
    return out


# --------------------------------------------------------------------------
# Dealiasing: padded transforms (3/2 rule) and direct 2/3-rule truncation
# --------------------------------------------------------------------------


def _nbc(bc):
    return 0 if bc is None else 2


def pad_backward(c, family, bc, N, M, axis=0):
    """_padding_backward: N coefficients -> M slots along ``axis``.

    Polynomial (spectralbase.py:1368-1389): dofs 0..N-nbc-1 kept, the nbc
    boundary slots moved to the end.  R2C (fourier/bases.py:419-429): the
    N/2+1 entries kept, even N: the Nyquist entry real and halved.  C2C
    (fourier/bases.py:580-594): wavenumbers 0..N/2 and the last N/2 kept,
    even N: both copies of the Nyquist entry halved.  ``N`` is the space's
    size (the real length for R2C), ``M`` the padded physical length.
    """
    cm = np.moveaxis(np.asarray(c), axis, 0)
    if family == 'fourier-r2c':
        out = np.zeros((M // 2 + 1,) + cm.shape[1:], dtype=cm.dtype)
        out[:N // 2 + 1] = cm
        if N % 2 == 0:
            out[N // 2] = out[N // 2].real * 0.5
    elif family == 'fourier-c2c':
        out = np.zeros((M,) + cm.shape[1:], dtype=cm.dtype)
        h = N // 2
        out[:h + 1] = cm[:h + 1]
        if h > 0:
            out[M - h:] = cm[N - h:]
        if N % 2 == 0:
            out[h] *= 0.5
            out[M - h] *= 0.5
    else:
        nb = _nbc(bc)
        out = np.zeros((M,) + cm.shape[1:], dtype=cm.dtype)
        out[:N - nb] = cm[:N - nb]
        if nb:
            out[M - nb:] = cm[N - nb:]
    return np.moveaxis(out, 0, axis)


def truncate_forward(c, family, bc, N, axis=0):
    """_truncation_forward: M coefficients -> the N of the space.

    Polynomial (spectralbase.py:1362-1366): dofs 0..N-nbc-1, boundary slots 0.
    R2C (fourier/bases.py:409-418): first N/2+1 entries, even N: Nyquist
    real and doubled.  C2C (fourier/bases.py:570-578): wavenumbers 0..N/2 plus
    the last N/2 added (the two Nyquist halves meet at N/2 for even N).
    """
    cm = np.moveaxis(np.asarray(c), axis, 0)
    M = cm.shape[0]
    if family == 'fourier-r2c':
        out = cm[:N // 2 + 1].copy()
        if N % 2 == 0:
            out[N // 2] = out[N // 2].real * 2
    elif family == 'fourier-c2c':
        h = N // 2
        out = np.zeros((N,) + cm.shape[1:], dtype=cm.dtype)
        out[:h + 1] = cm[:h + 1]
        if h > 0:
            out[N - h:] += cm[M - h:]
    else:
        nb = _nbc(bc)
        out = np.zeros((N,) + cm.shape[1:], dtype=cm.dtype)
        out[:N - nb] = cm[:N - nb]
    return np.moveaxis(out, 0, axis)


def dealias_direct_zero(c, family, N, axis=0):
    """dealias_direct: the top third of the wavenumbers zeroed before backward.

    Polynomial: k >= 2N//3 (spectralbase.py:1374-1376); R2C: k >= N//3
    (fourier/bases.py:430-433); C2C: slice(N//3, -(N//3)+1)
    (fourier/bases.py:595-598).
    """
    cm = np.moveaxis(np.array(c, copy=True), axis, 0)
    if family == 'fourier-r2c':
        cm[N // 3:] = 0
    elif family == 'fourier-c2c':
        cm[slice(N // 3, -(N // 3) + 1)] = 0
    else:
        cm[2 * N // 3:] = 0
    return np.moveaxis(cm, 0, axis)


def padded_scalar_product(family, bc, f, N, axis=0):
    """Scalar product of a padded space: on the M = f.shape[axis] points,
    then truncated (spectralbase.py:225-229)."""
    return truncate_forward(basis_scalar_product(family, bc, f, axis), family, bc, N, axis)


def padded_backward(family, bc, c, N, M, axis=0):
    """Backward of a padded space: padded coefficients synthesised on M
    points (spectralbase.py:291-310)."""
    return basis_backward(family, bc, pad_backward(c, family, bc, N, M, axis), axis, N=M)


def convolve_direct(u, v, N, real):
    """Direct convolution sums (optimization/cython/convolve.pyx): C2C
    convolve_1D :9-30 (out 2N, python-style negative indices) or R2C
    convolve_real_1D :34-75 (out N+1, p >= 0, negative wavenumbers by
    conjugation); wavenumbers of a length Np = N (odd) / N + 1 (even) grid in
    fftfreq order, Nyquist entries halved for even N.  Pure-Python loops:
    small N only."""
    Np = N if N % 2 else N + 1
    k = np.fft.fftfreq(Np, 1.0 / Np).astype(int)
    uv = np.zeros(N + 1 if real else 2 * N, dtype=complex)

    def coef(a, m):
        if N % 2 == 0 and abs(m) == N // 2:
            return a[abs(m)] * 0.5
        if real:
            return a[m] if m >= 0 else np.conj(a[-m])
        return a[m]

    for m in k:
        for n in k:
            p = m + n
            if real and p < 0:
                continue
            uv[p] += coef(u, m) * coef(v, n)
    return uv


# --------------------------------------------------------------------------
# Petrov-Galerkin test space Phi2
# --------------------------------------------------------------------------


def phi2_stencil(family, N):
    """{offset: S(k)}, k < N-4: Chebyshev chebyshev/bases.py:1441-1445,
    Legendre legendre/bases.py:631 (the ``_stencil`` dicts)."""
    k = np.arange(N - 4, dtype=float)
    if family == 'chebyshev':
        return {0: 1 / (2 * PI * (k + 1) * (k + 2)), 2: -1 / (PI * (k + 1) * (k + 3)),
                4: 1 / (2 * PI * (k + 2) * (k + 3))}
    return {0: 1 / (2 * (2 * k + 3)), 2: -(2 * k + 5) / (2 * k + 7) / (2 * k + 3),
            4: 1 / (2 * (2 * k + 7))}


def biharmonic_stencil(family, kind, N):
    """{offset: S(k)}, k < N-4 (SB) or N-8 (P4): ShenBiharmonic
    chebyshev/bases.py ShenBiharmonic ``_stencil`` {0: 1, 2: -(2n+4)/(n+3),
    4: (n+1)/(n+3)}, legendre/bases.py {0: 1, 2: -(4n+10)/(2n+7),
    4: (2n+3)/(2n+7)}; Phi4 chebyshev/bases.py / legendre/bases.py Phi4."""
    if kind == 'SB':
        k = np.arange(N - 4, dtype=float)
        if family == 'chebyshev':
            return {0: np.ones_like(k), 2: -(2 * k + 4) / (k + 3), 4: (k + 1) / (k + 3)}
        return {0: np.ones_like(k), 2: -(4 * k + 10) / (2 * k + 7), 4: (2 * k + 3) / (2 * k + 7)}
    n = np.arange(N - 8, dtype=float)
    if family == 'chebyshev':
        return {0: 1 / (8 * PI * (n + 1) * (n + 2) * (n + 3) * (n + 4)),
                2: -1 / (2 * PI * (n + 1) * (n + 3) * (n + 4) * (n + 5)),
                4: 3 / (4 * PI * (n + 2) * (n + 3) * (n + 5) * (n + 6)),
                6: -1 / (2 * PI * (n + 3) * (n + 4) * (n + 5) * (n + 7)),
                8: 1 / (8 * PI * (n + 4) * (n + 5) * (n + 6) * (n + 7))}
    return {0: 1 / (2 * (8 * n ** 3 + 60 * n ** 2 + 142 * n + 105)),
            2: -2 / (8 * n ** 3 + 84 * n ** 2 + 262 * n + 231),
            4: 3 * (2 * n + 9) / ((2 * n + 5) * (2 * n + 7) * (2 * n + 11) * (2 * n + 13)),
            6: -2 / (8 * n ** 3 + 132 * n ** 2 + 694 * n + 1155),
            8: 1 / (2 * (8 * n ** 3 + 156 * n ** 2 + 1006 * n + 2145))}


def stencil_polys(family, st):
    """The basis functions of a stencil dict as numpy polynomial objects."""
    P = _poly_cls(family)
    n = len(next(iter(st.values())))
    return [sum(st[o][k] * P.basis(k + o) for o in st) for k in range(n)]


def sb_backward(family, c, x):
    """ShenBiharmonic synthesis sum_k c_k phi_k(x) (dofs k < N-4)."""
    N = len(c)
    return sum(c[k] * p(x) for k, p in enumerate(stencil_polys(family, biharmonic_stencil(family, 'SB', N))))


def phi2_scalar_product(family, f, axis=0):
    """(f, phi_k)_w on the N points of f: the orthogonal scalar product, then
    the stencil (spectralbase.py:1462-1474 with Phi2's stencil matrix,
    :1504-1526); slots N-4 .. N-1 zero."""
    N = f.shape[axis]
    t = np.moveaxis(cheb_scalar_product(f, axis) if family == 'chebyshev'
                    else leg_scalar_product(f, axis), axis, 0)
    out = np.zeros_like(t)
    for off, v in phi2_stencil(family, N).items():
        out[:N - 4] += _bcast(v, t.ndim, 0) * t[off:off + N - 4]
    return np.moveaxis(out, 0, axis)


def pg_matrices(family, bc, N):
    """Dense Petrov-Galerkin matrices (n x n, n = N-2) of an N-point trial space
    and its Phi2 test space on N + 2 points: A[k, j] = (phi_k, psi_j'')_w,
    B[k, j] = (phi_k, psi_j)_w, by Gauss quadrature on N + 8 points."""
    n = N - 2
    Q = N + 8
    x, w = cheb_points_weights(Q) if family == 'chebyshev' else leg_points_weights(Q)
    P = _poly_cls(family)
    S = phi2_stencil(family, N + 2)
    phi = np.stack([sum(S[o][k] * P.basis(k + o)(x) for o in S) for k in range(n)], axis=1)
    K2 = stencil_k2(family, bc, N)
    psi = [P.basis(j) + K2[j] * P.basis(j + 2) for j in range(n)]
    A = (phi * w[:, None]).T @ np.stack([p.deriv(2)(x) for p in psi], axis=1)
    B = (phi * w[:, None]).T @ np.stack([p(x) for p in psi], axis=1)
    return A, B


# --------------------------------------------------------------------------
# Nonhomogeneous boundary values (Dirichlet or Neumann at both ends)
# --------------------------------------------------------------------------


def _bnd_value(family, j, deriv, side):
    """P_j^(deriv)(+-1) in closed form: T_j(+-1) = (+-1)^j, T_j'(+-1) =
    (+-1)^(j+1) j^2; L_j(+-1) = (+-1)^j, L_j'(+-1) = (+-1)^(j+1) j(j+1)/2
    (chebyshev/bases.py, legendre/bases.py ``bnd_values``)."""
    sgn = -1.0 if side == 'L' else 1.0
    if deriv == 0:
        return sgn ** j
    d = j * j if family == 'chebyshev' else j * (j + 1) / 2
    return sgn ** (j + 1) * d


def bc_basis(family, kind):
    """get_bc_basis (utilities/findbasis.py:68-140) for kind 'D' or 'N' at both
    ends: the 2 x 2 matrix of the boundary functionals on P_first,
    P_first+1, first the lowest index that makes it invertible; rows of the
    result are the boundary functions' orthogonal coefficients (left, right)."""
    deriv = 0 if kind == 'D' else 1
    for first in range(deriv + 1):
        A = np.array([[_bnd_value(family, first + m, deriv, side) for m in range(2)]
                      for side in ('L', 'R')])
        if abs(np.linalg.det(A)) > 1e-12:
            break
    X = np.linalg.solve(A, np.eye(2)).T
    out = np.zeros((2, first + 2))
    out[:, first:] = X
    return out


def _poly_cls(family):
    return np.polynomial.Chebyshev if family == 'chebyshev' else np.polynomial.Legendre


def bc_functions(family, kind, x, deriv=0):
    """phi_bc_j^(deriv)(x), shape (len(x), 2)."""
    P = _poly_cls(family)
    C = bc_basis(family, kind)
    return np.stack([P(C[j]).deriv(deriv)(x) if deriv else P(C[j])(x) for j in range(2)], axis=1)


def bc_matrices(family, kind, N):
    """(A, B), each (N-2, 2): B[k, j] = (phi_bc_j, phi_k)_w, A[k, j] =
    (phi_bc_j'', phi_k)_w for Chebyshev, -(phi_bc_j'', phi_k) for Legendre
    (the units of its stiffness matrix (phi', phi')).  Quadrature on N + 6
    points: exact, and independent of the space's own rule."""
    Q = N + 6
    if family == 'chebyshev':
        x, w = cheb_points_weights(Q)
    else:
        x, w = leg_points_weights(Q)
    P = _poly_cls(family)
    K2 = stencil_k2(family, kind, N)
    Phi = np.stack([P.basis(k)(x) + K2[k] * P.basis(k + 2)(x) for k in range(N - 2)], axis=1)
    B = (Phi * w[:, None]).T @ bc_functions(family, kind, x)
    A = (Phi * w[:, None]).T @ bc_functions(family, kind, x, 2)
    return (A if family == 'chebyshev' else -A), B


def backward_bc(family, kind, c, x, axis=0):
    """Backward of a composite space with nonhomogeneous boundary values:
    the homogeneous synthesis plus the boundary functions weighted by the
    coefficients' boundary slots N-2, N-1 (to_ortho + _add_to_orthogonal,
    spectralbase.py:1616-1634, tensorproductspace.py:2070-2089), on points x."""
    cm = np.moveaxis(np.asarray(c), axis, 0)
    N = cm.shape[0]
    u = np.moveaxis(basis_backward(family, kind, c, axis), axis, 0)
    lift = bc_functions(family, kind, x)  # (len(x), 2)
    u = u + np.tensordot(lift, cm[N - 2:], axes=(1, 0))
    return np.moveaxis(u, 0, axis)


def forward_bc(family, kind, vals, f):
    """1D forward with boundary values ``vals``: scalar product, slots :=
    vals, minus the boundary mass matrix times vals (_add_mass_rhs,
    tensorproductspace.py:2091-2106), then the dense mass solve."""
    N = f.shape[0]
    s = basis_scalar_product(family, kind, f, 0)
    s[N - 2:] = vals
    _, Bbc = bc_matrices(family, kind, N)
    rhs = s[:N - 2] - Bbc @ np.asarray(vals, dtype=float)
    Bf = {('chebyshev', 'D'): cheb_B_SD, ('chebyshev', 'N'): cheb_B_SN,
          ('legendre', 'D'): leg_B_SD, ('legendre', 'N'): leg_B_SN}[(family, kind)]
    s[:N - 2] = np.linalg.solve(dense(Bf(N), N - 2), rhs)
    return s


# --------------------------------------------------------------------------
# Galerkin matrices (closed forms).  Returned as dict offset -> diagonal.
# --------------------------------------------------------------------------


def cheb_norm_sq(N, quad='GC'):
    """chebyshev/bases.py:271-277 (l2_norm_sq): pi/2, first entry pi, and with
    GL the last entry pi (matrices assemble by quadrature, matrixbase.py:635-636)."""
    h = np.full(N, PI / 2)
    h[0] = PI
    if quad == 'GL':
        h[-1] = PI
    return h


def leg_norm_sq(N):
    """Legendre ||L_k||^2 = 2/(2k+1) (jacobi/bases.py JacobiBase, alpha=beta=0)."""
    k = np.arange(N, dtype=float)
    return 2.0 / (2 * k + 1)


def cheb_B_SD(N, quad='GC'):
    """chebyshev/matrices.py:221-249 BSDSDmat: {-2: -h1, 0: h[:-2]+h[2:], 2: -h1}."""
    h = cheb_norm_sq(N, quad)
    return {-2: np.full(N - 4, -h[1]), 0: h[:-2] + h[2:], 2: np.full(N - 4, -h[1])}


def cheb_A_SD(N, nbands=None):
    """chebyshev/matrices.py:1047-1079 ASDSDmat (phi''_j, phi_k)_w.

    a_kk = -2 pi (k+1)(k+2); a_{k,k+2j} = -4 pi (k+1) for j >= 1.
    """
    k = np.arange(N, dtype=float)
    d = {0: -2 * PI * (k[:N - 2] + 1) * (k[:N - 2] + 2)}
    for i in range(2, N - 2, 2):
        if nbands is not None and i > 2 * nbands:
            break
        d[i] = -4 * PI * (k[:N - 2 - i] + 1)
    return d


def cheb_B_SN(N, quad='GC'):
    """chebyshev/matrices.py:377-397 BSNSNmat."""
    h = cheb_norm_sq(N, quad)
    k = np.arange(N, dtype=float)
    alpha = (k / (k + 2)) ** 2
    M = N - 2
    d = {0: h[:-2] + alpha[:-2] ** 2 * h[2:],
         2: -alpha[:M - 2] * h[2:M]}
    d[-2] = d[2].copy()
    return d


def cheb_A_SN(N):
    """chebyshev/matrices.py:1096-1121 ASNSNmat."""
    k = np.arange(N - 2, dtype=float)
    d = {0: -2 * PI * k ** 2 * (k + 1) / (k + 2)}
    for i in range(2, N - 2, 2):
        d[i] = -4 * PI * (k[:-i] + i) ** 2 * (k[:-i] + 1) / (k[:-i] + 2) ** 2
    return d


def leg_B_SD(N):
    """legendre/matrices.py:110-134 BSDSDmat (unscaled)."""
    h = leg_norm_sq(N)
    d = {0: h[:-2] + h[2:], -2: -h[2:-2]}
    d[2] = d[-2].copy()
    return d


def leg_A_SD(N):
    """legendre/matrices.py:271-290 ASDSDmat: (phi'_j, phi'_k) = 4k+6 (diagonal)."""
    k = np.arange(N - 2, dtype=float)
    return {0: 4 * k + 6}


def leg_B_SN(N):
    """legendre/matrices.py:139-160 BSNSNmat."""
    k = np.arange(N - 2, dtype=float)
    alpha = k * (k + 1) / (k + 2) / (k + 3)
    h = leg_norm_sq(N)
    d = {0: h[:-2] + alpha ** 2 * h[2:], 2: -alpha[:-2] * h[2:-2]}
    d[-2] = d[2].copy()
    return d


def leg_A_SN(N):
    """legendre/matrices.py:295-312 ASNSNmat (diagonal)."""
    k = np.arange(N - 2, dtype=float)
    alpha = k * (k + 1) / (k + 2) / (k + 3)
    d0 = 2.0 / (2 * k + 1)
    return {0: d0 * alpha * (k + 0.5) * ((k + 2) * (k + 3) - k * (k + 1))}


def dense(d, M):
    """Dense M x M matrix from a diagonal dict (scalars broadcast)."""
    A = np.zeros((M, M))
    for off, v in d.items():
        n = M - abs(off)
        if n <= 0:
            continue
        vv = np.broadcast_to(np.atleast_1d(v), (n,)) if np.ndim(v) == 0 or len(np.atleast_1d(v)) != n else v
        if off >= 0:
            A[np.arange(n), np.arange(n) + off] = vv[:n]
        else:
            A[np.arange(n) - off, np.arange(n)] = vv[:n]
    return A


# --------------------------------------------------------------------------
# Banded LU / solves (la.py + optimization/cython/la.pyx), vectorised over
# all lines: the solved axis is moved to 0 and the recurrences run along it.
# --------------------------------------------------------------------------


def tdma_dia(d, n):
    """SparseMatrix {-2,0,2} -> scipy-style DIA rows (3, n) (la.py:256-258)."""
    data = np.zeros((3, n))
    data[0, :n - 2] = np.broadcast_to(d[-2], (n - 2,))
    data[1, :] = np.broadcast_to(d[0], (n,))
    data[2, 2:] = np.broadcast_to(d[2], (n - 2,))
    return data


def tdma_lu(data):
    """la.pyx:243-252 TDMA_LU (in place on a copy; returns it)."""
    data = np.array(data, dtype=float, copy=True)
    n = data.shape[-1]
    ld = data[..., 0, :]
    d = data[..., 1, :]
    ud = data[..., 2, 2:]
    for i in range(2, n):
        ld[..., i - 2] = ld[..., i - 2] / d[..., i - 2]
        d[..., i] = d[..., i] - ld[..., i - 2] * ud[..., i - 2]
    return data


def tdma_solve(u, data, axis=0):
    """la.pyx:520-532 TDMA_inner_solve_ptr along ``axis`` (returns new array).

    ``data`` is (3, n) shared by every line, or (lines..., 3, n) broadcastable
    against the non-axis dims of u (per-mode factors, la.pyx:371-426).
    """
    um = np.moveaxis(np.array(u, copy=True), axis, 0)
    data = np.asarray(data)
    n = data.shape[-1]

    def row(r, i):
        v = data[..., r, i]
        return v if np.ndim(v) == 0 else v

    for i in range(2, n):
        um[i] -= row(0, i - 2) * um[i - 2]
    um[n - 1] = um[n - 1] / row(1, n - 1)
    um[n - 2] = um[n - 2] / row(1, n - 2)
    for i in range(n - 3, -1, -1):
        um[i] = (um[i] - row(2, i + 2) * um[i + 2]) / row(1, i)
    return np.moveaxis(um, 0, axis)


def pdma_dia(d, n):
    """SparseMatrix {-4,-2,0,2,4} -> DIA rows (5, n)."""
    data = np.zeros((5, n))
    data[0, :n - 4] = np.broadcast_to(d[-4], (n - 4,))
    data[1, :n - 2] = np.broadcast_to(d[-2], (n - 2,))
    data[2, :] = np.broadcast_to(d[0], (n,))
    data[3, 2:] = np.broadcast_to(d[2], (n - 2,))
    data[4, 4:] = np.broadcast_to(d[4], (n - 4,))
    return data


def pdma_lu(data):
    """la.pyx:213-241 PDMA_LU (k = 2 band spacing)."""
    data = np.array(data, dtype=float, copy=True)
    n = data.shape[-1]
    a = data[0, :n - 4]
    b = data[1, :n - 2]
    d = data[2, :]
    e = data[3, 2:]
    f = data[4, 4:]
    m = e.shape[0]
    k = n - m
    for i in range(n - 2 * k):
        lam = b[i] / d[i]
        d[i + k] -= lam * e[i]
        e[i + k] -= lam * f[i]
        b[i] = lam
        lam = a[i] / d[i]
        b[i + k] -= lam * e[i]
        d[i + 2 * k] -= lam * f[i]
        a[i] = lam
    i = n - 4
    lam = b[i] / d[i]
    d[i + k] -= lam * e[i]
    b[i] = lam
    i = n - 3
    lam = b[i] / d[i]
    d[i + k] -= lam * e[i]
    b[i] = lam
    return data


def pdma_solve(u, data, axis=0):
    """la.pyx:496-514 PDMA_inner_solve_ptr along ``axis``."""
    um = np.moveaxis(np.array(u, copy=True), axis, 0)
    n = data.shape[-1]
    a = data[0]
    b = data[1]
    d = data[2]
    e = data[3, 2:]
    f = data[4, 4:]
    um[2] -= b[0] * um[0]
    um[3] -= b[1] * um[1]
    for k in range(4, n):
        um[k] -= b[k - 2] * um[k - 2] + a[k - 4] * um[k - 4]
    um[n - 1] /= d[n - 1]
    um[n - 2] /= d[n - 2]
    um[n - 3] = (um[n - 3] - e[n - 3] * um[n - 1]) / d[n - 3]
    um[n - 4] = (um[n - 4] - e[n - 4] * um[n - 2]) / d[n - 4]
    for k in range(n - 5, -1, -1):
        um[k] = (um[k] - e[k] * um[k + 2] - f[k] * um[k + 4]) / d[k]
    return np.moveaxis(um, 0, axis)


def diagma_solve(u, d, axis=0):
    """la.pyx:586-590 DiagMA_inner_solve_ptr: u[i] /= d[i], i < len(d)."""
    um = np.moveaxis(np.array(u, copy=True), axis, 0)
    n = len(d)
    um[:n] = um[:n] / _bcast(np.asarray(d), um.ndim, 0)
    return np.moveaxis(um, 0, axis)


def band_dia(d, n, offsets):
    """SparseMatrix dict -> scipy-style DIA rows (len(offsets), n), rows in
    ``offsets`` order; offset o >= 0 fills columns o.., o < 0 columns ..n+o
    (la.py:256-258, the ``mat.diags('dia').data`` layout)."""
    data = np.zeros((len(offsets), n))
    for r, o in enumerate(offsets):
        if o not in d:
            continue
        if o >= 0:
            data[r, o:] = np.broadcast_to(d[o], (n - o,))
        else:
            data[r, :n + o] = np.broadcast_to(d[o], (n + o,))
    return data


def tdma_o_lu(data):
    """la.pyx:254-263 TDMA_O_LU (offsets -1, 0, 1)."""
    data = np.array(data, dtype=float, copy=True)
    n = data.shape[-1]
    ld = data[0, :n - 1]
    d = data[1, :]
    ud = data[2, 1:]
    for i in range(1, n):
        ld[i - 1] = ld[i - 1] / d[i - 1]
        d[i] = d[i] - ld[i - 1] * ud[i - 1]
    return data


def tdma_o_solve(u, data, axis=0):
    """la.pyx:538-549 TDMA_O_inner_solve_ptr along ``axis``."""
    um = np.moveaxis(np.array(u, copy=True), axis, 0)
    n = data.shape[-1]
    ld = data[0]
    d = data[1]
    ud = data[2, 1:]
    for i in range(1, n):
        um[i] -= ld[i - 1] * um[i - 1]
    um[n - 1] = um[n - 1] / d[n - 1]
    for i in range(n - 2, -1, -1):
        um[i] = (um[i] - ud[i] * um[i + 1]) / d[i]
    return np.moveaxis(um, 0, axis)


def twodma_solve(u, data, axis=0):
    """la.pyx:555-563 TwoDMA_inner_solve_ptr (offsets 0, 2; no LU)."""
    um = np.moveaxis(np.array(u, copy=True), axis, 0)
    n = data.shape[-1]
    d = data[0]
    u1 = data[1, 2:]
    um[n - 1] = um[n - 1] / d[n - 1]
    um[n - 2] = um[n - 2] / d[n - 2]
    for i in range(n - 3, -1, -1):
        um[i] = (um[i] - u1[i] * um[i + 2]) / d[i]
    return np.moveaxis(um, 0, axis)


def threedma_solve(u, data, axis=0, compat=False):
    """la.pyx:569-580 ThreeDMA_inner_solve_ptr (offsets 0, 2, 4; no LU).

    The reference takes the +4 band from DIA row 1 (``u2 = &data[m1+4]``,
    la.pyx:574; ``data[1, 4:]``, la.py:688), i.e. the +2 band shifted by two
    rows, so its result is the solution of the matrix only when those two
    bands coincide.  compat=True restates exactly that; the default reads row
    2, the +4 band, and so solves the matrix the class documents
    (la.py:661-668).
    """
    um = np.moveaxis(np.array(u, copy=True), axis, 0)
    n = data.shape[-1]
    d = data[0]
    u1 = data[1, 2:]
    u2 = data[1, 4:] if compat else data[2, 4:]
    um[n - 1] = um[n - 1] / d[n - 1]
    um[n - 2] = um[n - 2] / d[n - 2]
    um[n - 3] = (um[n - 3] - u1[n - 3] * um[n - 1]) / d[n - 3]
    um[n - 4] = (um[n - 4] - u1[n - 4] * um[n - 2]) / d[n - 4]
    for i in range(n - 5, -1, -1):
        um[i] = (um[i] - u1[i] * um[i + 2] - u2[i] * um[i + 4]) / d[i]
    return np.moveaxis(um, 0, axis)


def fdma_lu(data):
    """la.pyx:199-211 FDMA_LU (offsets -2, 0, 2, 4)."""
    data = np.array(data, dtype=float, copy=True)
    n = data.shape[-1]
    ld = data[0, :n - 2]
    d = data[1, :]
    u1 = data[2, 2:]
    u2 = data[3, 4:]
    for i in range(2, n):
        ld[i - 2] = ld[i - 2] / d[i - 2]
        d[i] = d[i] - ld[i - 2] * u1[i - 2]
        if i < n - 2:
            u1[i] = u1[i] - ld[i - 2] * u2[i - 2]
    return data


def fdma_solve(u, data, axis=0):
    """la.pyx:596-611 FDMA_inner_solve_ptr along ``axis``."""
    um = np.moveaxis(np.array(u, copy=True), axis, 0)
    n = data.shape[-1]
    ld = data[0]
    d = data[1]
    u1 = data[2, 2:]
    u2 = data[3, 4:]
    for i in range(2, n):
        um[i] -= ld[i - 2] * um[i - 2]
    um[n - 1] = um[n - 1] / d[n - 1]
    um[n - 2] = um[n - 2] / d[n - 2]
    um[n - 3] = (um[n - 3] - u1[n - 3] * um[n - 1]) / d[n - 3]
    um[n - 4] = (um[n - 4] - u1[n - 4] * um[n - 2]) / d[n - 4]
    for i in range(n - 5, -1, -1):
        um[i] = (um[i] - u1[i] * um[i + 2] - u2[i] * um[i + 4]) / d[i]
    return np.moveaxis(um, 0, axis)


def heptadma_lu(data):
    """la.pyx:265-303 HeptaDMA_LU (offsets -4, -2, 0, 2, 4, 6, 8; k = 2)."""
    data = np.array(data, dtype=float, copy=True)
    n = data.shape[-1]
    a = data[0, :n - 4]
    b = data[1, :n - 2]
    d = data[2, :]
    e = data[3, 2:]
    f = data[4, 4:]
    g = data[5, 6:]
    h = data[6, 8:]
    k = 2
    for i in range(n - 2 * k):
        lam = b[i] / d[i]
        d[i + k] -= lam * e[i]
        e[i + k] -= lam * f[i]
        if i < n - 6:
            f[i + k] -= lam * g[i]
        if i < n - 8:
            g[i + k] -= lam * h[i]
        b[i] = lam
        lam = a[i] / d[i]
        b[i + k] -= lam * e[i]
        d[i + 2 * k] -= lam * f[i]
        if i < n - 6:
            e[i + 2 * k] -= lam * g[i]
        if i < n - 8:
            f[i + 2 * k] -= lam * h[i]
        a[i] = lam
    for i in (n - 4, n - 3):
        lam = b[i] / d[i]
        d[i + k] -= lam * e[i]
        b[i] = lam
    return data


def heptadma_solve(u, data, axis=0):
    """la.pyx:466-490 HeptaDMA_inner_solve_ptr along ``axis``; the reference's
    eight end rows are the general row with the terms beyond n - 1 dropped."""
    um = np.moveaxis(np.array(u, copy=True), axis, 0)
    n = data.shape[-1]
    a = data[0]
    b = data[1]
    d = data[2]
    up = [data[3, 2:], data[4, 4:], data[5, 6:], data[6, 8:]]
    um[2] -= b[0] * um[0]
    um[3] -= b[1] * um[1]
    for k in range(4, n):
        um[k] -= b[k - 2] * um[k - 2] + a[k - 4] * um[k - 4]
    for k in range(n - 1, -1, -1):
        acc = um[k]
        for j, c in enumerate(up):
            if k + 2 * (j + 1) < n:
                acc = acc - c[k] * um[k + 2 * (j + 1)]
        um[k] = acc / d[k]
    return np.moveaxis(um, 0, axis)


# --------------------------------------------------------------------------
# Chebyshev Helmholtz (chebyshev/la.py:131-351, la.pyx:708-973)
# --------------------------------------------------------------------------


def helmholtz_lu(family_bc, N, alfa, beta, quad='GC'):
    """LU_Helmholtz_1D (la.pyx:717-763) vectorised over modes.

    family_bc: 'D' (ASDSD/BSDSD) or 'N' (ASNSN/BSNSN) Chebyshev bases of size
    N (M = N-2 unknowns).  alfa/beta: arrays (any shape, broadcast together)
    of per-mode scales.  Returns d0, d1, d2, L with shape (N,) + mode shape;
    entries beyond the reference's loop ranges are 0, like the zero-initialised
    arrays of chebyshev/la.py:270-276.
    """
    neumann = family_bc == 'N'
    if neumann:
        A, B = cheb_A_SN(N), cheb_B_SN(N, quad)
    else:
        A, B = cheb_A_SD(N, nbands=2), cheb_B_SD(N, quad)
    M = N - 2
    alfa = np.asarray(alfa, dtype=float)
    beta = np.asarray(beta, dtype=float)
    mshape = np.broadcast(alfa, beta).shape
    alfa = np.broadcast_to(alfa, mshape)
    beta = np.broadcast_to(beta, mshape)
    ex = (slice(None),) + (np.newaxis,) * len(mshape)

    def full(v, n):
        return np.broadcast_to(np.atleast_1d(np.asarray(v, float)), (n,)).copy()

    A0 = full(A[0], M)
    A2 = full(A[2], M - 2)
    A4 = full(A[4], M - 4)
    Bm2 = full(B[-2], M - 2)
    B0 = full(B[0], M)
    B2 = full(B[2], M - 2)
    A0 = np.broadcast_to(A0[ex], (M,) + mshape).copy()
    B0 = np.broadcast_to(B0[ex], (M,) + mshape).copy()
    if neumann:
        # la.pyx:735-749
        zero_beta = np.abs(beta) < 1e-8
        A0[0] = np.where(zero_beta, 1.0 / np.where(zero_beta, alfa, 1.0), A0[0])
        B0[0] = np.where(zero_beta, 0.0, B0[0])
        i = np.arange(1, M, dtype=float)
        A0[1:] /= (i ** 2)[ex]
        B0[1:] /= (i ** 2)[ex]
        i = np.arange(2, M, dtype=float)
        A2 = A2 / i ** 2
        B2 = B2 / i ** 2
        i = np.arange(4, M, dtype=float)
        A4 = A4 / i ** 2
        i = np.arange(1, M - 2, dtype=float)
        Bm2 = Bm2.copy()
        Bm2[1:] /= i ** 2
    d0 = np.zeros((N,) + mshape)
    d1 = np.zeros((N,) + mshape)
    d2 = np.zeros((N,) + mshape)
    L = np.zeros((N,) + mshape)
    d0[0] = alfa * A0[0] + beta * B0[0]
    d0[1] = alfa * A0[1] + beta * B0[1]
    d1[0] = alfa * A2[0] + beta * B2[0]
    d1[1] = alfa * A2[1] + beta * B2[1]
    d2[0] = alfa * A4[0]
    d2[1] = alfa * A4[1]
    for i in range(2, M):
        L[i - 2] = beta * Bm2[i - 2] / d0[i - 2]
        d0[i] = alfa * A0[i] + beta * B0[i] - L[i - 2] * d1[i - 2]
        if i < M - 2:
            d1[i] = alfa * A2[i] + beta * B2[i] - L[i - 2] * d2[i - 2]
        if i < M - 4:
            d2[i] = alfa * A4[i] - L[i - 2] * d2[i - 2]
    return d0, d1, d2, L


def helmholtz_solve(b, d0, d1, d2, L, neumann=False, axis=0, u=None):
    """Solve_Helmholtz_1D_ptr (la.pyx:871-912) along ``axis``.

    Factor arrays have shape (N,) + mode shape where the mode shape broadcasts
    against b with the axis removed.  Only entries 0..N-3 of u are written;
    the two BC slots of ``u`` keep their input values (u defaults to zeros,
    like a fresh shenfun Function).
    """
    bm = np.moveaxis(np.asarray(b), axis, 0)
    N = bm.shape[0]
    M = N - 2
    if u is None:
        um = np.zeros_like(bm)
    else:
        um = np.moveaxis(np.array(u, copy=True), axis, 0)

    def f(a):
        return np.moveaxis(np.asarray(a), 0, 0)

    d0, d1, d2, L = f(d0), f(d1), f(d2), f(L)
    y = np.zeros((M,) + bm.shape[1:], dtype=bm.dtype)
    y[0] = bm[0]
    y[1] = bm[1]
    for i in range(2, M):
        y[i] = bm[i] - L[i - 2] * y[i - 2]
    um[M - 1] = y[M - 1] / d0[M - 1]
    um[M - 2] = y[M - 2] / d0[M - 2]
    um[M - 3] = (y[M - 3] - d1[M - 3] * um[M - 1]) / d0[M - 3]
    um[M - 4] = (y[M - 4] - d1[M - 4] * um[M - 2]) / d0[M - 4]
    s_even = np.zeros(bm.shape[1:], dtype=bm.dtype)
    s_odd = np.zeros(bm.shape[1:], dtype=bm.dtype)
    for i in range(M - 5, -1, -1):
        v = y[i] - d1[i] * um[i + 2]
        if i % 2 == 0:
            s_even = s_even + um[i + 4]
            v = v - d2[i] * s_even
        else:
            s_odd = s_odd + um[i + 4]
            v = v - d2[i] * s_odd
        um[i] = v / d0[i]
    if neumann:
        # la.pyx:907-912
        gauge = (d0[0] - 1.0) * (d0[0] - 1.0) < 1e-16
        um[0] = np.where(gauge, 0.0, um[0])
        i = np.arange(1, M, dtype=float)
        um[1:M] = um[1:M] / _bcast(i * i, um.ndim, 0)
    return np.moveaxis(um, 0, axis)


def helmholtz_matvec(v, alfa, beta, family_bc='D', axis=0):
    """c = (alfa*A + beta*B) v for the Chebyshev SD/SN Galerkin matrices.

    Matvec.pyx:656-720 (Dirichlet) and :722-818 (Neumann) compute this with
    running even/odd tail sums; here the product is formed from the closed
    form diagonals directly.  Only entries 0..N-3 of c are set; BC slots 0.
    """
    vm = np.moveaxis(np.asarray(v), axis, 0)
    N = vm.shape[0]
    M = N - 2
    if family_bc == 'D':
        A, B = cheb_A_SD(N), cheb_B_SD(N)
    else:
        A, B = cheb_A_SN(N), cheb_B_SN(N)
    alfa = np.asarray(alfa, float)
    beta = np.asarray(beta, float)
    c = np.zeros_like(vm)
    vv = vm[:M]
    for off, dvals in A.items():
        n = M - off
        dv = np.broadcast_to(np.atleast_1d(dvals), (n,))
        c[:n] += alfa * _bcast(dv, vm.ndim, 0) * vv[off:off + n]
    for off, dvals in B.items():
        n = M - abs(off)
        dv = np.broadcast_to(np.atleast_1d(dvals), (n,))
        if off >= 0:
            c[:n] += beta * _bcast(dv, vm.ndim, 0) * vv[off:off + n]
        else:
            c[-off:-off + n] += beta * _bcast(dv, vm.ndim, 0) * vv[:n]
    return np.moveaxis(c, 0, axis)


def poisson_add_solve(b, N, scale=1.0, axis=0):
    """Chebyshev Dirichlet Poisson A u = b: Poisson_Solve_ADD_1D_ptr (la.pyx:679-706).

    d = A[0], d1 = A[2] (row-constant tail -4 pi (k+1)); running parity sums.
    """
    A = cheb_A_SD(N, nbands=1)
    d = A[0]
    d1 = A[2]
    bm = np.moveaxis(np.asarray(b), axis, 0)
    M = N - 2
    u = np.zeros_like(bm)
    se = np.zeros(bm.shape[1:], dtype=bm.dtype)
    so = np.zeros(bm.shape[1:], dtype=bm.dtype)
    u[M - 1] = bm[M - 1] / d[M - 1]
    u[M - 2] = bm[M - 2] / d[M - 2]
    for k in range(M - 3, -1, -1):
        if k % 2 == 0:
            se = se + u[k + 2]
            u[k] = bm[k] - d1[k] * se
        else:
            so = so + u[k + 2]
            u[k] = bm[k] - d1[k] * so
        u[k] = u[k] / d[k]
    if not abs(scale - 1) < 1e-8:
        u[:M] = u[:M] / scale
    return np.moveaxis(u, 0, axis)


def ann_solve(b, N, scale=1.0, axis=0, constraints=((0, 0),)):
    """Chebyshev Neumann Poisson A u = b: ANNSolver.__call__ (chebyshev/la.py:
    74-128) with u = b.  The SN stiffness is upper triangular with row-constant
    entries after the column scaling m^2, so the unknowns v_m = m^2 u_m follow
    from a back substitution with running parity sums for k = M-1 .. 1; u_0
    keeps b_0 until the constraints set their rows; u /= scale; the boundary
    slots are set to the (homogeneous) boundary values."""
    A = cheb_A_SN(N)
    M = N - 2
    um = np.moveaxis(np.array(b, copy=True), axis, 0)
    bs = um[:M].copy()
    us = um[:M]
    j2 = np.arange(M, dtype=float) ** 2
    j2[0] = 1
    j2 = 1.0 / j2
    d = A[0] * j2
    d1 = A[2] * j2[2:]
    se = np.zeros(um.shape[1:], dtype=um.dtype)
    so = np.zeros(um.shape[1:], dtype=um.dtype)
    us[M - 1] = bs[M - 1] / d[M - 1]
    us[M - 2] = bs[M - 2] / d[M - 2]
    for k in range(M - 3, 0, -1):
        if k % 2 == 0:
            se = se + us[k + 2]
            us[k] = bs[k] - d1[k] * se
        else:
            so = so + us[k + 2]
            us[k] = bs[k] - d1[k] * so
        us[k] = us[k] / d[k]
    us *= j2.reshape((M,) + (1,) * (um.ndim - 1))
    um /= scale
    for row, val in constraints:
        um[row] = val
    um[M:] = 0
    return np.moveaxis(um, 0, axis)


def cheb_SB_matrices(N, quad='GC'):
    """Galerkin matrices of the Chebyshev ShenBiharmonic basis (N-4 unknowns),
    dicts offset -> diagonal: SSBSBmat (phi_j^(4), phi_k)_w, upper with every
    even offset (chebyshev/matrices.py:1444-1459), ASBSBmat (phi_j'', phi_k)_w,
    offsets -2, 0, 2 (:1008-1017), BSBSBmat (phi_j, phi_k)_w, offsets -4..4
    from the norms and the stencil (:515-528)."""
    ki = np.arange(N - 4)
    k = ki.astype(float)
    S = {0: PI * 8 * (k + 1) ** 2 * (k + 2) * (k + 4)}
    for j in range(2, N - 4, 2):
        m = 8 * (ki[:-j] + 1) * (ki[:-j] + 2) * (ki[:-j] * (ki[:-j] + 4) + 3 * (ki[j:] + 2) ** 2)
        S[j] = m * PI / (k[j:] + 3)
    A = {-2: 2 * (k[2:] - 1) * (k[2:] + 2) * PI,
         0: -4 * ((k + 1) * (k + 2) ** 2) / (k + 3) * PI,
         2: 2 * (k[:-2] + 1) * (k[:-2] + 2) * PI}
    h = cheb_norm_sq(N, quad)
    st = biharmonic_stencil('chebyshev', 'SB', N)
    S2 = np.concatenate([st[2], np.zeros(2)])  # the reference's S[2] has N-2 entries
    S4 = st[4]
    B = {0: h[:-4] + h[2:-2] * S2[:-2] ** 2 + h[4:] * S4 ** 2,
         2: h[2:-4] * S2[:-4] + h[4:-2] * S4[:-2] * S2[2:-2],
         4: h[4:-4] * S4[:-4]}
    B[-2] = B[2].copy()
    B[-4] = B[4].copy()
    return S, A, B


def _lu_oe_biharmonic(odd, a, b, c, sii, siu, siuu, ail, aii, aiu, bill, bil, bii, biu, biuu):
    """LU_oe_Biharmonic_1D (la.pyx:1018-1138) of one parity chain: three rolling
    rows c0, c1, c2 of a S + b A + c B, the dense upper tail of S from its
    closed form, eliminated with the two lower bands."""
    M = sii.shape[0]
    c0, c1, c2 = np.zeros(M), np.zeros(M), np.zeros(M)
    u0, u1, u2, l0, l1 = (np.zeros(M) for _ in range(5))

    def tail(kr, j):
        return 8 * (kr + 1) * (kr + 2) * (kr * (kr + 4) + 3 * (j + 2) * (j + 2)) * a * PI / (j + 3.)

    c0[0] = a * sii[0] + b * aii[0] + c * bii[0]
    c0[1] = a * siu[0] + b * aiu[0] + c * biu[0]
    c0[2] = a * siuu[0] + c * biuu[0]
    c0[3] = tail(odd, 6 + odd)
    c0[4] = tail(odd, 8 + odd)
    c1[0] = b * ail[0] + c * bil[0]
    c1[1] = a * sii[1] + b * aii[1] + c * bii[1]
    c1[2] = a * siu[1] + b * aiu[1] + c * biu[1]
    c1[3] = a * siuu[1] + c * biuu[1]
    c1[4] = tail(odd + 2, 8 + odd)
    c2[0] = c * bill[0]
    c2[1] = b * ail[1] + c * bil[1]
    c2[2] = a * sii[2] + b * aii[2] + c * bii[2]
    c2[3] = a * siu[2] + b * aiu[2] + c * biu[2]
    c2[4] = a * siuu[2] + c * biuu[2]
    for i in range(5, M):
        j = 2 * i + odd
        c0[i] = tail(odd, j)
        c1[i] = tail(odd + 2, j)
        c2[i] = tail(odd + 4, j)
    u0[0], u1[0], u2[0] = c0[0], c0[1], c0[2]
    for kk in range(1, M):
        l0[kk - 1] = c1[kk - 1] / u0[kk - 1]
        if kk < M - 1:
            l1[kk - 1] = c2[kk - 1] / u0[kk - 1]
        c1[kk:] = c1[kk:] - l0[kk - 1] * c0[kk:]
        if kk < M - 1:
            c2[kk:] = c2[kk:] - l1[kk - 1] * c0[kk:]
        c0[kk:] = c1[kk:]
        c1[kk:] = c2[kk:]
        if kk < M - 2:
            c2[kk] = c * bill[kk]
            c2[kk + 1] = b * ail[kk + 1] + c * bil[kk + 1]
            c2[kk + 2] = a * sii[kk + 2] + b * aii[kk + 2] + c * bii[kk + 2]
            if kk < M - 3:
                c2[kk + 3] = a * siu[kk + 2] + b * aiu[kk + 2] + c * biu[kk + 2]
            if kk < M - 4:
                c2[kk + 4] = a * siuu[kk + 2] + c * biuu[kk + 2]
            if kk < M - 5:
                kr = 2 * (kk + 2) + odd
                for i in range(kk + 5, M):
                    c2[i] = tail(kr, 2 * i + odd)
        u0[kk] = c0[kk]
        if kk < M - 1:
            u1[kk] = c0[kk + 1]
        if kk < M - 2:
            u2[kk] = c0[kk + 2]
    return u0, u1, u2, l0, l1


def _factor_oe_pr(odd, l0, l1):
    """Biharmonic_factor_oe_pr (la.pyx:1211-1233): the two coefficients of U's
    tail rows, (a_k / (j+3) + b_k (j+2)^2 / (j+3)) a0, eliminated like the rows."""
    M = l0.shape[0] + 1
    a = np.zeros(M - 1)
    b = np.zeros(M - 1)
    k = odd
    a[0] = 8 * k * (k + 1) * (k + 2) * (k + 4) * PI
    b[0] = 24 * (k + 1) * (k + 2) * PI
    k = 2 + odd
    a[1] = 8 * k * (k + 1) * (k + 2) * (k + 4) * PI - l0[0] * a[0]
    b[1] = 24 * (k + 1) * (k + 2) * PI - l0[0] * b[0]
    for k in range(2, M - 3):
        kk = 2 * k + odd
        pp = 8 * kk * (kk + 1) * (kk + 2) * (kk + 4)
        rr = 24 * (kk + 1) * (kk + 2)
        a[k] = pp * PI - l0[k - 1] * a[k - 1] - l1[k - 2] * a[k - 2]
        b[k] = rr * PI - l0[k - 1] * b[k - 1] - l1[k - 2] * b[k - 2]
    return a, b


def biharmonic_lu(N, a0, alfa, beta):
    """LU_Biharmonic_1D + Biharmonic_factor_pr_1D (la.pyx:991-1016, 1203-1209)
    for a0 S + alfa A + beta B of the Chebyshev ShenBiharmonic basis: factors
    (u0, u1, u2, l0, l1, ak, bk), each (2, M) with M the even chain's length."""
    S, A, B = cheb_SB_matrices(N)
    dg = [S[0], S[2], S[4], A[-2], A[0], A[2], B[-4], B[-2], B[0], B[2], B[4]]
    M = dg[0][::2].shape[0]
    out = [np.zeros((2, M)) for _ in range(7)]
    for odd in (0, 1):
        f = _lu_oe_biharmonic(odd, a0, alfa, beta, *[d[odd::2] for d in dg])
        for t, v in zip(out, f):
            t[odd, :len(v)] = v
    for odd in (0, 1):
        # on the (2, M) rows, as Biharmonic_factor_pr_1D passes them
        ak, bk = _factor_oe_pr(odd, out[3][odd], out[4][odd])
        out[5][odd, :len(ak)] = ak
        out[6][odd, :len(bk)] = bk
    return tuple(out)


def biharmonic_solve(b, factors, a0, axis=0):
    """Solve_oe_Biharmonic_1D (la.pyx:1286-1310): forward with the two lower
    bands, back substitution with the three upper bands and the tail through
    running sums s1 = sum u/(j+3), s2 = sum u (j+2)^2/(j+3), per parity chain,
    along ``axis``; factors as biharmonic_lu returns them (a tuple: one mode
    for every line) or a list of such per line (C order over the other dims).  Entries
    N-4.. of u are zero."""
    bm = np.moveaxis(np.asarray(b), axis, 0)
    shp = bm.shape
    lines = bm.reshape(shp[0], -1)
    out = np.zeros_like(lines)
    for li in range(lines.shape[1]):
        u0, u1, u2, l0, l1, ak, bk = factors[li] if isinstance(factors, list) else factors
        for odd in (0, 1):
            fk = lines[odd:shp[0] - 4:2, li]
            M = fk.shape[0]
            y = np.zeros(M, dtype=lines.dtype)
            y[0] = fk[0]
            y[1] = fk[1] - l0[odd, 0] * y[0]
            for i in range(2, M):
                y[i] = fk[i] - l0[odd, i - 1] * y[i - 1] - l1[odd, i - 2] * y[i - 2]
            uk = np.zeros(M, dtype=lines.dtype)
            U0, U1, U2 = u0[odd], u1[odd], u2[odd]
            uk[M - 1] = y[M - 1] / U0[M - 1]
            uk[M - 2] = (y[M - 2] - U1[M - 2] * uk[M - 1]) / U0[M - 2]
            uk[M - 3] = (y[M - 3] - U1[M - 3] * uk[M - 2] - U2[M - 3] * uk[M - 1]) / U0[M - 3]
            s1 = s2 = 0.0
            for kk in range(M - 4, -1, -1):
                j = 2 * kk + odd + 6
                s1 = s1 + uk[kk + 3] / (j + 3.)
                s2 = s2 + (uk[kk + 3] / (j + 3.)) * ((j + 2) * (j + 2))
                uk[kk] = (y[kk] - U1[kk] * uk[kk + 1] - U2[kk] * uk[kk + 2]
                          - ak[odd, kk] * a0 * s1 - bk[odd, kk] * a0 * s2) / U0[kk]
            out[odd:shp[0] - 4:2, li] = uk
    return np.moveaxis(out.reshape(shp), 0, axis)


# --------------------------------------------------------------------------
# End-to-end pipelines (the configurations' work)
# --------------------------------------------------------------------------


def helmholtz3d_cheb_fourier(f, alpha, bc='D'):
    """(alpha - lap) u = f on Chebyshev(bc) x C2C x R2C, FP64.

    Scalar product (tensorproductspace.py:146-190) -> Helmholtz with
    alfa = -1, beta = alpha + ky^2 + kz^2 (chebyshev/la.py:223-317, scales from
    matrixbase.py:1328-1365) -> backward (tensorproductspace.py:46-96).
    Returns (u physical, f_hat, u_hat).
    """
    Nx, Ny, Nz = f.shape
    bases = [('chebyshev', bc), ('fourier-c2c', None), ('fourier-r2c', None)]
    f_hat = tps_scalar_product(f, bases)
    ky = c2c_wavenumbers(Ny).astype(float)
    kz = r2c_wavenumbers(Nz).astype(float)
    beta = alpha + ky[:, None] ** 2 + kz[None, :] ** 2
    alfa = -np.ones_like(beta)
    d0, d1, d2, L = helmholtz_lu(bc, Nx, alfa, beta)
    u_hat = helmholtz_solve(f_hat, d0, d1, d2, L, neumann=(bc == 'N'), axis=0)
    u = tps_backward(u_hat, bases, f.shape)
    return u, f_hat, u_hat


def legendre_fourier_2d(f, alpha, bc='D'):
    """(alpha - lap) u = f on Legendre(bc) x R2C (alpha = 0: Poisson), FP64.

    SolverGeneric1ND (la.py:1183-1434): per ky the TDMA of A + (alpha+ky^2) B
    (legendre/matrices.py:110-137, 271-292); ky = 0 solved with its own
    solver (DiagMA when alpha = 0), la.py:1346-1364.
    Returns (u physical, f_hat, u_hat).
    """
    Nx, Ny = f.shape
    bases = [('legendre', bc), ('fourier-r2c', None)]
    f_hat = tps_scalar_product(f, bases)
    ky = r2c_wavenumbers(Ny).astype(float)
    if bc == 'D':
        A, B = leg_A_SD(Nx), leg_B_SD(Nx)
    else:
        A, B = leg_A_SN(Nx), leg_B_SN(Nx)
    M = Nx - 2
    u_hat = np.zeros_like(f_hat)
    for j, k in enumerate(ky):
        sc = alpha + k * k
        dmat = {0: A[0] + sc * B[0], -2: sc * B[-2], 2: sc * B[2]}
        data = tdma_lu(tdma_dia(dmat, M))
        col = tdma_solve(f_hat[:M, j], data)
        u_hat[:M, j] = col
    u = tps_backward(u_hat, bases, f.shape)
    return u, f_hat, u_hat


# --------------------------------------------------------------------------
# Synthetic inputs of the configurations (seeded)
# --------------------------------------------------------------------------


def poisson2d_manufactured(Nx=64, Ny=64, seed=0):
    """Config: 2D Poisson -lap u = f, Legendre SD(Nx) x R2C(Ny).

    u = (1-x^2) sum_{k=1..8} a_k cos(k y + phi_k), a_k ~ N(0,1)/k^2,
    phi_k ~ U(0, 2 pi).  Draw order: a (8 normals) then phi (8 uniforms).
    Returns (x, y, f, u_exact, u_hat_exact) with u_hat in the R2C layout:
    u = (2/3) phi_0(x) sum a_k cos(k y + phi_k)  =>
    u_hat[0, k] = (2/3) a_k e^{i phi_k} / 2, k = 1..8.
    """
    rng = np.random.default_rng(seed)
    k = np.arange(1, 9, dtype=float)
    a = rng.standard_normal(8) / k ** 2
    phi = rng.uniform(0.0, 2 * PI, 8)
    x, _ = leg_points_weights(Nx)
    y = fourier_points(Ny)
    X = x[:, None]
    Y = y[None, :]
    u = np.zeros((Nx, Ny))
    f = np.zeros((Nx, Ny))
    for kk, ak, pk in zip(k, a, phi):
        c = np.cos(kk * Y + pk)
        u += (1 - X ** 2) * ak * c
        # -u_xx - u_yy: u_xx = -2 a c, u_yy = -k^2 (1-x^2) a c
        f += 2 * ak * c + kk ** 2 * (1 - X ** 2) * ak * c
    uh = np.zeros((Nx, Ny // 2 + 1), dtype=complex)
    for kk, ak, pk in zip(k.astype(int), a, phi):
        uh[0, kk] = (2.0 / 3.0) * ak * np.exp(1j * pk) / 2
    return x, y, f, u, uh


def helmholtz3d_rhs(Nx, Ny, Nz, seed=0, dtype=np.float64):
    """Smooth random f for the 3D Chebyshev x Fourier^2 Helmholtz config.

    Spectral coefficients complex N(0,1) (1 + kx^2 + ky^2 + kz^2)^-2 on
    T_kx x exp(i ky y) x exp(i kz z) (kx = 0..Nx-1, C2C ky, R2C kz); Hermitian
    symmetry is enforced by synthesising f with an FP64 irfft along z.
    Draw order: real parts then imaginary parts, each of shape (Nx, Ny, Nz//2+1).
    """
    rng = np.random.default_rng(seed)
    shp = (Nx, Ny, Nz // 2 + 1)
    c = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    kx = np.arange(Nx, dtype=float)[:, None, None]
    ky = c2c_wavenumbers(Ny).astype(float)[None, :, None]
    kz = r2c_wavenumbers(Nz).astype(float)[None, None, :]
    c *= (1 + kx ** 2 + ky ** 2 + kz ** 2) ** -2.0
    # c2c inverse along y then c2r along z: the c2r takes the Hermitian part
    g = np.fft.irfftn(c, s=(Ny, Nz), axes=(1, 2)) * (Ny * Nz)
    f = cheb_backward(g, axis=0)
    return np.ascontiguousarray(f.astype(dtype))
