"""Galerkin matrices of the hot path in shenfun's diagonal-dict form (host, FP64).

Mirrors matrixbase.SparseMatrix (matrixbase.py:26-531), SpectralMatrix
(:534-801) and TPMatrix (:1275-1646) for the bases of this package, with the
closed-form diagonals of

* Chebyshev: BSDSDmat chebyshev/matrices.py:221-249, ASDSDmat :1047-1079,
  BSNSNmat :377-397, ASNSNmat :1096-1121;
* Legendre: BSDSDmat legendre/matrices.py:110-134, ASDSDmat :271-290,
  BSNSNmat :139-160, ASNSNmat :295-312.

These are tiny O(N) host arrays; the device kernels receive them once.
"""
from numbers import Number

import numpy as np
import scipy.sparse
import torch

from . import _lib

PI = np.pi


class SparseMatrix(dict):
    """Dict offset -> diagonal (array or scalar) of a square banded matrix."""

    def __init__(self, d, shape, scale=1.0):
        super().__init__(sorted(d.items()))
        self.shape = tuple(shape)
        self.scale = scale

    def get_key(self):
        return self.__class__.__name__

    def sorted_keys(self):
        return np.array(sorted(self.keys()))

    def diagonal_array(self, key, scaled=False):
        """Full-length diagonal ``key`` as an array (scalars broadcast)."""
        n = self.shape[0] - abs(key)
        v = np.broadcast_to(np.atleast_1d(np.asarray(self[key], dtype=float)), (n,)).copy()
        if scaled:
            v *= float(np.asarray(self.scale).ravel()[0]) if np.ndim(self.scale) else self.scale
        return v

    def diags(self, format="dia", scaled=True):
        keys = sorted(self.keys())
        data = [self.diagonal_array(k, scaled=scaled) for k in keys]
        m = scipy.sparse.diags(data, keys, shape=self.shape, format=format)
        return m

    def todense(self):
        return self.diags("csr").toarray()

    def __mul__(self, x):
        assert isinstance(x, Number) or np.ndim(x) == 0
        return SparseMatrix({k: self.diagonal_array(k) for k in self}, self.shape,
                            scale=self.scale * float(x))

    __rmul__ = __mul__

    def __add__(self, other):
        assert self.shape == other.shape
        out = {}
        for m in (self, other):
            for k in m:
                v = m.diagonal_array(k) * float(m.scale)
                out[k] = out[k] + v if k in out else v
        return SparseMatrix(out, self.shape, scale=1.0)

    def band_rows(self):
        """(offsets, D) with D[r][i] = A[i][i + offsets[r]] (scale applied), the
        layout of sfa_band_matvec."""
        N = self.shape[0]
        offs = sorted(self.keys())
        D = np.zeros((len(offs), N))
        for r, o in enumerate(offs):
            v = self.diagonal_array(o) * float(np.asarray(self.scale).reshape(-1)[0])
            if o >= 0:
                D[r, :len(v)] = v
            else:
                D[r, -o:-o + len(v)] = v
        return offs, D

    def _device_rows(self, device):
        key = ("_band_dev", str(device))
        cache = self.__dict__.setdefault("_dev_cache", {})
        if key not in cache:
            offs, D = self.band_rows()
            cache[key] = (torch.tensor(offs, dtype=torch.int64, device=device),
                          torch.from_numpy(D).to(device))
        return cache[key]

    def matvec(self, v, c, format=None, axis=0, scales=None, accumulate=False):
        """c = self . v along ``axis`` on the device (matrixbase.py:87-144):
        rows >= the matrix size are zeroed; ``scales`` (device, one per line
        of the non-axis dims in C order) multiplies each line; ``accumulate``
        adds into c instead."""
        for t, name in ((v, "v"), (c, "c")):
            if not isinstance(t, torch.Tensor) or not t.is_cuda:
                raise TypeError(f"{name} must be a CUDA (HIP) torch tensor")
        if v.dtype != c.dtype or v.dtype not in (torch.float64, torch.complex128):
            raise TypeError("v and c must both be float64 or both complex128")
        offs, D = self._device_rows(v.device)
        code = _lib.C128 if v.dtype == torch.complex128 else _lib.F64
        _lib.call("sfa_band_matvec", _lib.ptr(v), _lib.ptr(c), code, v.dim(), _lib.arr(v.shape),
                  _lib.arr(v.stride()), _lib.arr(c.shape), _lib.arr(c.stride()), axis,
                  self.shape[0], self.shape[1], offs.numel(), _lib.ptr(offs), _lib.ptr(D),
                  _lib.ptr(scales) if scales is not None else None, int(bool(accumulate)),
                  _lib.stream_ptr())
        return c

    def get_solver(self):
        """Band pattern -> solver class (matrixbase.py:439-474)."""
        from . import la
        if float(self.scale) == 0:
            return la.Solve
        keys = tuple(self.sorted_keys())
        if keys == (-1, 0, 1):
            return la.TDMA_O if self.issymmetric else la.Solve
        for cls in (la.DiagMA, la.TwoDMA, la.TDMA, la.ThreeDMA, la.FDMA, la.PDMA, la.HeptaDMA):
            if keys == cls.offsets:
                return cls
        return la.Solve

    @property
    def issymmetric(self):
        """Symmetric band pattern with equal mirrored diagonals (matrixbase.py:493-503)."""
        if np.sum(np.array(list(self.keys()))) != 0:
            return False
        for key in self:
            if key <= 0:
                continue
            if -key not in self or not np.all(
                    abs(self.diagonal_array(key) - self.diagonal_array(-key)) < 1e-16):
                return False
        return True


class SpectralMatrix(SparseMatrix):
    """SparseMatrix with test/trial spaces (matrixbase.py:534-801)."""

    def __init__(self, test, trial, scale=1.0):
        self.testfunction = test
        self.trialfunction = trial
        d = self.assemble()
        N = test[0].N - 2
        super().__init__(d, (N, N), scale=scale)

    @property
    def axis(self):
        return self.testfunction[0].axis

    @property
    def tensorproductspace(self):
        return self.testfunction[0].tensorproductspace

    def assemble(self):  # pragma: no cover - abstract
        raise NotImplementedError


def _cheb_h(N, quad="GC"):
    """Quadrature norms of T_k (chebyshev/bases.py:271-277; matrices are
    assembled by quadrature, matrixbase.py:635-636): GL doubles the last."""
    h = np.full(N, PI / 2)
    h[0] = PI
    if quad == "GL":
        h[-1] = PI
    return h


class ChebBSDSDmat(SpectralMatrix):
    """chebyshev/matrices.py:221-249: {-2: -pi/2, 0: h_k + h_{k+2}, 2: -pi/2},
    h the quadrature norms of the space (GL: h_{N-1} = pi)."""

    def assemble(self):
        N = self.testfunction[0].N
        h = _cheb_h(N, self.testfunction[0].quad)
        return {-2: -h[1], 0: h[:-2] + h[2:], 2: -h[1]}

    def get_key(self):
        return "BSDSDmat"


class ChebASDSDmat(SpectralMatrix):
    """chebyshev/matrices.py:1047-1079: a_kk = -2 pi (k+1)(k+2), a_{k,k+2j} = -4 pi (k+1)."""

    def assemble(self):
        N = self.testfunction[0].N
        k = np.arange(N, dtype=float)
        d = {0: -2 * PI * (k[:N - 2] + 1) * (k[:N - 2] + 2)}
        for i in range(2, N - 2, 2):
            d[i] = -4 * PI * (k[:N - 2 - i] + 1)
        return d

    def get_key(self):
        return "ASDSDmat"


class ChebBSNSNmat(SpectralMatrix):
    """chebyshev/matrices.py:377-397."""

    def assemble(self):
        N = self.testfunction[0].N
        h = _cheb_h(N, self.testfunction[0].quad)
        k = np.arange(N, dtype=float)
        al = (k / (k + 2)) ** 2
        M = N - 2
        d = {0: h[:-2] + al[:-2] ** 2 * h[2:], 2: -al[:M - 2] * h[2:M]}
        d[-2] = d[2].copy()
        return d

    def get_key(self):
        return "BSNSNmat"


class ChebASNSNmat(SpectralMatrix):
    """chebyshev/matrices.py:1096-1121."""

    def assemble(self):
        N = self.testfunction[0].N
        k = np.arange(N - 2, dtype=float)
        d = {0: -2 * PI * k ** 2 * (k + 1) / (k + 2)}
        for i in range(2, N - 2, 2):
            d[i] = -4 * PI * (k[:-i] + i) ** 2 * (k[:-i] + 1) / (k[:-i] + 2) ** 2
        return d

    def get_key(self):
        return "ASNSNmat"


def _leg_h(N):
    k = np.arange(N, dtype=float)
    return 2.0 / (2 * k + 1)


class LegBSDSDmat(SpectralMatrix):
    """legendre/matrices.py:110-134 (unscaled basis)."""

    def assemble(self):
        h = _leg_h(self.testfunction[0].N)
        d = {0: h[:-2] + h[2:], -2: -h[2:-2]}
        d[2] = d[-2].copy()
        return d

    def get_key(self):
        return "BSDSDmat"


class LegASDSDmat(SpectralMatrix):
    """legendre/matrices.py:271-290: (phi'_j, phi'_k) = 4k + 6."""

    def assemble(self):
        k = np.arange(self.testfunction[0].N - 2, dtype=float)
        return {0: 4 * k + 6}

    def get_key(self):
        return "ASDSDmat"


class LegBSNSNmat(SpectralMatrix):
    """legendre/matrices.py:139-160."""

    def assemble(self):
        N = self.testfunction[0].N
        k = np.arange(N - 2, dtype=float)
        al = k * (k + 1) / (k + 2) / (k + 3)
        h = _leg_h(N)
        d = {0: h[:-2] + al ** 2 * h[2:], 2: -al[:-2] * h[2:-2]}
        d[-2] = d[2].copy()
        return d

    def get_key(self):
        return "BSNSNmat"


class LegASNSNmat(SpectralMatrix):
    """legendre/matrices.py:295-312."""

    def assemble(self):
        k = np.arange(self.testfunction[0].N - 2, dtype=float)
        al = k * (k + 1) / (k + 2) / (k + 3)
        d0 = 2.0 / (2 * k + 1)
        return {0: d0 * al * (k + 0.5) * ((k + 2) * (k + 3) - k * (k + 1))}

    def get_key(self):
        return "ASNSNmat"


class _ChebSBmat(SpectralMatrix):
    """Galerkin matrix of the Chebyshev ShenBiharmonic basis (N-4 unknowns)."""

    def __init__(self, test, trial, scale=1.0):
        self.testfunction = test
        self.trialfunction = trial
        d = self.assemble()
        N = test[0].N - 4
        SparseMatrix.__init__(self, d, (N, N), scale=scale)


class ChebSSBSBmat(_ChebSBmat):
    """(phi_j^(4), phi_k)_w, chebyshev/matrices.py:1444-1459: upper triangular
    with every even offset, s_kj = 8 pi (k+1)(k+2)(k(k+4) + 3(j+2)^2)/(j+3) for
    j > k and 8 pi (k+1)^2 (k+2)(k+4) on the diagonal."""

    def assemble(self):
        N = self.testfunction[0].N
        ki = np.arange(N - 4)
        k = ki.astype(float)
        d = {0: PI * 8 * (k + 1) ** 2 * (k + 2) * (k + 4)}
        for j in range(2, N - 4, 2):
            m = 8 * (ki[:-j] + 1) * (ki[:-j] + 2) * (ki[:-j] * (ki[:-j] + 4) + 3 * (ki[j:] + 2) ** 2)
            d[j] = m * PI / (k[j:] + 3)
        return d

    def get_key(self):
        return "SSBSBmat"


class ChebASBSBmat(_ChebSBmat):
    """(phi_j'', phi_k)_w, chebyshev/matrices.py:1008-1017 (offsets -2, 0, 2)."""

    def assemble(self):
        k = np.arange(self.testfunction[0].N - 4, dtype=float)
        return {-2: 2 * (k[2:] - 1) * (k[2:] + 2) * PI,
                0: -4 * ((k + 1) * (k + 2) ** 2) / (k + 3) * PI,
                2: 2 * (k[:-2] + 1) * (k[:-2] + 2) * PI}

    def get_key(self):
        return "ASBSBmat"


class ChebBSBSBmat(_ChebSBmat):
    """(phi_j, phi_k)_w, chebyshev/matrices.py:515-528: the norms against the
    stencil phi_k = T_k + s2_k T_{k+2} + s4_k T_{k+4}, s2 = -(2k+4)/(k+3),
    s4 = (k+1)/(k+3) (offsets -4..4)."""

    def assemble(self):
        N = self.testfunction[0].N
        h = _cheb_h(N, self.testfunction[0].quad)
        k = np.arange(N, dtype=float)
        s2 = -(2 * k + 4) / (k + 3)
        s4 = (k + 1) / (k + 3)
        d = {0: h[:-4] + h[2:-2] * s2[:N - 4] ** 2 + h[4:] * s4[:N - 4] ** 2,
             2: h[2:-4] * s2[:N - 6] + h[4:-2] * s4[:N - 6] * s2[2:N - 4],
             4: h[4:-4] * s4[:N - 8]}
        d[-2] = d[2].copy()
        d[-4] = d[4].copy()
        return d

    def get_key(self):
        return "BSBSBmat"


def chebyshev_biharmonic_matrices(space):
    """(S, A, B): the Galerkin biharmonic, stiffness and mass matrices of a
    Chebyshev ShenBiharmonic space, as chebyshev.la.Biharmonic takes them."""
    return (ChebSSBSBmat((space, 0), (space, 4)), ChebASBSBmat((space, 0), (space, 2)),
            ChebBSBSBmat((space, 0), (space, 0)))


def mass_matrix(space):
    """(phi_j, phi_k)_w for a composite space."""
    fam, bc = space.family(), space.bc_kind
    cls = {("chebyshev", "D"): ChebBSDSDmat, ("chebyshev", "N"): ChebBSNSNmat,
           ("legendre", "D"): LegBSDSDmat, ("legendre", "N"): LegBSNSNmat}[(fam, bc)]
    return cls((space, 0), (space, 0))


def stiffness_matrix(space):
    """Chebyshev: (phi''_j, phi_k)_w (ASDSDmat/ASNSNmat, negative definite);
    Legendre: (phi'_j, phi'_k) (positive), as in the reference's mat dicts."""
    fam, bc = space.family(), space.bc_kind
    cls = {("chebyshev", "D"): ChebASDSDmat, ("chebyshev", "N"): ChebASNSNmat,
           ("legendre", "D"): LegASDSDmat, ("legendre", "N"): LegASNSNmat}[(fam, bc)]
    if fam == "chebyshev":
        return cls((space, 0), (space, 2))
    return cls((space, 1), (space, 1))


class TPMatrix:
    """Tensor-product matrix: one SpectralMatrix along ``naxes`` and Fourier
    diagonals folded into ``scale`` (TPMatrix.get_simplified,
    matrixbase.py:1328-1365)."""

    def __init__(self, pmat, space, scale, naxis):
        self.pmat = pmat
        self.space = space
        self.trialspace = space
        self.scale = np.asarray(scale, dtype=float)
        self.naxes = [naxis]
        self.dimensions = space.dimensions

    @property
    def mats(self):
        out = [None] * self.dimensions
        out[self.naxes[0]] = self.pmat
        return out

    def get_key(self):
        return self.pmat.get_key()

    @property
    def testfunction(self):
        return self.pmat.testfunction

    @property
    def axis(self):
        return self.naxes[0]

    @property
    def tensorproductspace(self):
        return self.space

    def line_scales(self, shape, device):
        """The per-mode scale broadcast over the non-axis dims of ``shape``,
        one value per line in C order (device)."""
        ax = self.naxes[0]
        full = np.broadcast_to(np.asarray(self.scale, dtype=float),
                               [1 if i == ax else n for i, n in enumerate(shape)])
        return torch.from_numpy(np.array(np.moveaxis(full, ax, -1)).reshape(-1)).to(device)

    def matvec(self, v, c, format=None, accumulate=False):
        """c = self . v: the matrix along the non-periodic axis times the
        per-mode scale (TPMatrix.matvec, matrixbase.py:1437-1449)."""
        if getattr(self.pmat, "is_bc_matrix", False):
            raise TypeError("boundary matrices act on the boundary slots: use the solvers' matvec")
        return self.pmat.matvec(v, c, axis=self.naxes[0], scales=self.line_scales(v.shape, v.device),
                                accumulate=accumulate)


class BCMatrix:
    """Boundary matrix (N-2, 2) of a composite space with nonhomogeneous
    boundary values: rows the test functions, columns the boundary functions
    (the trial-side boundary matrices that extract_bc_matrices separates,
    matrixbase.py:1769-1793).  ``which`` is 'A' (stiffness units of the space)
    or 'B' (mass)."""

    is_bc_matrix = True

    def __init__(self, space, which, test=None):
        self.space = space
        self.which = which
        if test is None:
            A, B = space.bc_matrices()
            self.mat = A if which == "A" else B
        else:
            # Petrov-Galerkin: (test_k, phi_bc_j'') or (test_k, phi_bc_j) by exact
            # algebra in the orthogonal basis, as GalerkinMatrix
            P = np.polynomial.chebyshev if space.family() == "chebyshev" else np.polynomial.legendre
            der = P.chebder if space.family() == "chebyshev" else P.legder
            n, L = test.dim(), test.N
            Dc = np.zeros((2, L))
            for j, c in enumerate(space.bc_basis()):
                c = der(c, 2) if which == "A" else c
                Dc[j, :len(c)] = c
            self.mat = (test.stencil_matrix()[:n] * space.ortho_norms(L)[None, :]) @ Dc.T
        self.shape = self.mat.shape
        self.scale = 1.0
        self.testfunction = (space if test is None else test, 0)

    def get_key(self):
        return f"BC{self.which}mat"

    @property
    def axis(self):
        return self.space.axis


def _clean_diagonals(D, reltol=1e-8):
    """Nonzero diagonals of the dense matrix D as {offset: diagonal}: a
    diagonal is dropped when its 2-norm over max|row sum| / ncols is below
    reltol (SparseMatrix.clean_diagonals, matrixbase.py:510-528)."""
    relmax = np.abs(D.sum(axis=1)).max() / D.shape[1]
    if relmax == 0:
        relmax = 1.0
    d = {}
    for off in range(-D.shape[0] + 1, D.shape[1]):
        v = np.diagonal(D, off)
        if np.linalg.norm(v) / relmax >= reltol:
            d[off] = v.copy()
    return d


class GalerkinMatrix(SparseMatrix):
    """(test_k, d^deriv/dx^deriv trial_j)_w for composite test/trial spaces
    given by their stencils, assembled once on the host by exact algebra in
    the orthogonal basis: the trial functions' derivative coefficients (numpy
    chebder / legder), weighted by the exact norms (P_m, P_m)_w, against the
    test stencil; then the nonzero diagonals (the reference assembles the
    same matrices by quadrature or closed forms and cleans them the same
    way).  Used for the Petrov-Galerkin and biharmonic systems."""

    def __init__(self, test, trial, deriv, key):
        n = trial.dim()
        if test.dim() != n:
            raise ValueError("test and trial spaces must have the same number of unknowns")
        # band window.  test_k spans P_k .. P_{k+s_te}, trial_j P_j .. P_{j+s_tr}:
        # the mass matrix lies in [-s_tr, s_te].  A Petrov-Galerkin test
        # function is (1-x^2)^q times a polynomial, so deriv integrations by
        # parts move the derivative onto it and the window is
        # [deriv - s_tr, s_te]; Legendre Galerkin (weight 1, trial functions
        # vanishing with their slopes) is symmetric, [-s_tr, s_te].  Outside
        # the window the algebra below would only leave the cancellation
        # error of very large derivative coefficients.
        s_te, s_tr = max(test.stencil()), max(trial.stencil())
        if deriv == 0:
            lo = -s_tr
        elif getattr(test, "kind", None) in ("P2", "P4"):
            lo = deriv - s_tr
        elif trial.family() == "legendre" and test.bc_kind == trial.bc_kind:
            lo = -s_tr
        else:
            raise NotImplementedError("Chebyshev Galerkin matrices of derivatives are not banded; "
                                      "use the Petrov-Galerkin test space get_testspace('PG')")
        L = max(test.N, trial.N)
        P = np.polynomial.chebyshev if trial.family() == "chebyshev" else np.polynomial.legendre
        der = P.chebder if trial.family() == "chebyshev" else P.legder
        St = trial.stencil_matrix()[:n]
        Dc = np.zeros((n, L))
        for j in range(n):
            c = der(St[j], deriv) if deriv else St[j]
            Dc[j, :len(c)] = c
        Sv = np.zeros((n, L))
        Sv[:, :test.N] = test.stencil_matrix()[:n]
        D = (Sv * trial.ortho_norms(L)[None, :]) @ Dc.T
        off = np.arange(n)[None, :] - np.arange(n)[:, None]
        D[(off < lo) | (off > s_te)] = 0.0
        super().__init__(_clean_diagonals(D), (n, n))
        self.testfunction = (test, 0)
        self.trialfunction = (trial, deriv)
        self._key = key

    def get_key(self):
        return self._key


def _k2(T, ax):
    shape = list(T.shape(True))
    shape[ax] = 1
    k2 = np.zeros(shape)
    for i, b in enumerate(T.bases):
        if i != ax:
            k = b.wavenumbers().astype(float)
            s = [1] * T.dimensions
            s[i] = len(k)
            k2 = k2 + (k ** 2).reshape(s)
    return k2


def biharmonic_tpmatrices(T, a=1.0, b=0.0, c=0.0, test=None):
    """TPMatrix list of inner(v, a*biharmonic(u) + b*div(grad(u)) + c*u) for T
    with one ShenBiharmonic axis and Fourier elsewhere, v in ``test``
    (T.get_testspace('PG'): Phi4, the HeptaDMA systems; or T itself,
    Galerkin, banded for Legendre): (v, 4th x-derivative of u) with scale a,
    (v, 2nd x-derivative) with b - 2a|k|^2, (v, u) with a|k|^4 - b|k|^2 + c
    (forms/inner.py -> get_simplified_tpmatrices, matrixbase.py:1648-1677)."""
    ax = T.nonperiodic_axis()
    S = T.bases[ax]
    if getattr(S, "kind", None) != "SB":
        raise ValueError("biharmonic_tpmatrices needs a ShenBiharmonic axis")
    V = (test if test is not None else T).bases[ax]
    k2 = _k2(T, ax)
    if S.family() == "chebyshev" and V is S:
        # Galerkin: S is dense above the diagonal; solved by la.Biharmonic
        Sm, Am, Bm = chebyshev_biharmonic_matrices(S)
        return [TPMatrix(Sm, T, np.full(k2.shape, float(a)), ax), TPMatrix(Am, T, b - 2 * a * k2, ax),
                TPMatrix(Bm, T, a * k2 ** 2 - b * k2 + c, ax)]
    return [TPMatrix(GalerkinMatrix(V, S, 4, "SBBmat"), T, np.full(k2.shape, float(a)), ax),
            TPMatrix(GalerkinMatrix(V, S, 2, "ABBmat"), T, b - 2 * a * k2, ax),
            TPMatrix(GalerkinMatrix(V, S, 0, "BBBmat"), T, a * k2 ** 2 - b * k2 + c, ax)]


def helmholtz_tpmatrices(T, alpha=0.0, test=None):
    """The simplified TPMatrix list of inner(v, alpha*u - div(grad(u))) for a
    tensor-product space T with one Chebyshev/Legendre axis and Fourier
    elsewhere: [A (stiffness) with its scale, B (mass) with scale
    alpha + |k|^2] (forms/inner.py:374-493 -> get_simplified_tpmatrices,
    matrixbase.py:1648-1677).  With ``test = T.get_testspace('PG')`` the
    Petrov-Galerkin pair (Phi2, u'') with scale -1 and (Phi2, u) with
    scale alpha + |k|^2 (bands 0, 2 and -2..4: solved as FDMA)."""
    ax = T.nonperiodic_axis()
    S = T.bases[ax]
    if test is not None and test is not T and getattr(test.bases[ax], "kind", None) == "P2":
        P2 = test.bases[ax]
        k2 = _k2(T, ax)
        mats = [TPMatrix(GalerkinMatrix(P2, S, 2, "PGAmat"), T, np.full(k2.shape, -1.0), ax),
                TPMatrix(GalerkinMatrix(P2, S, 0, "PGBmat"), T, alpha + k2, ax)]
        if getattr(S, "has_nonhomogeneous_bcs", False):
            mats += [TPMatrix(BCMatrix(S, "A", test=P2), T, np.full(k2.shape, -1.0), ax),
                     TPMatrix(BCMatrix(S, "B", test=P2), T, alpha + k2, ax)]
        return mats
    shape = list(T.shape(True))
    shape[ax] = 1
    k2 = np.zeros(shape)
    for i, b in enumerate(T.bases):
        if i == ax:
            continue
        k = b.wavenumbers().astype(float)
        s = [1] * T.dimensions
        s[i] = len(k)
        k2 = k2 + (k ** 2).reshape(s)
    if getattr(S, "kind", None) in ("DN", "ND"):
        # mixed bases: (phi_k, phi_j'') and the mass matrix by exact algebra
        # (Legendre; Chebyshev raises: its Galerkin stiffness is not banded)
        A = GalerkinMatrix(S, S, 2, f"A{S.kind}{S.kind}mat")
        B = GalerkinMatrix(S, S, 0, f"B{S.kind}{S.kind}mat")
        mats = [TPMatrix(A, T, np.full(shape, -1.0), ax), TPMatrix(B, T, alpha + k2, ax)]
        if S.has_nonhomogeneous_bcs:
            # (phi_k, psi_j'') and (phi_k, psi_j) of the boundary functions psi_j
            mats += [TPMatrix(BCMatrix(S, "A", test=S), T, np.full(shape, -1.0), ax),
                     TPMatrix(BCMatrix(S, "B", test=S), T, alpha + k2, ax)]
        return mats
    A = stiffness_matrix(S)
    B = mass_matrix(S)
    # Chebyshev: -lap -> -(phi'', phi)_w = -A ; Legendre: (phi', phi') = +A
    a_scale = -1.0 if S.family() == "chebyshev" else 1.0
    mats = [TPMatrix(A, T, np.full(shape, a_scale), ax), TPMatrix(B, T, alpha + k2, ax)]
    if getattr(S, "has_nonhomogeneous_bcs", False):
        # the boundary functions' part, moved to the right-hand side by the solvers
        mats += [TPMatrix(BCMatrix(S, "A"), T, np.full(shape, a_scale), ax),
                 TPMatrix(BCMatrix(S, "B"), T, alpha + k2, ax)]
    return mats


def _split_bc(mats):
    """(regular matrices, boundary TPMatrices) (extract_bc_matrices, matrixbase.py:1769-1793)."""
    reg = [m for m in mats if not getattr(getattr(m, "pmat", m), "is_bc_matrix", False)]
    bc = [m for m in mats if getattr(getattr(m, "pmat", m), "is_bc_matrix", False)]
    return reg, bc


def apply_bc_matrices(bc_mats, b, u, sign=-1.0):
    """b += sign * (sum of the boundary matrices times u's boundary slots, each
    scaled per mode); sign -1 moves them to the right-hand side
    (Helmholtz.__call__, chebyshev/la.py:307-311; SolverGeneric1ND,
    la.py:1390-1410), +1 adds them to a matvec (la.py:1212-1229).  u's
    boundary slots are set first (final values)."""
    if not bc_mats:
        return b
    S = bc_mats[0].pmat.space
    S.set_boundary_dofs(u, final=True)
    ax = S.axis
    W = np.stack([m.pmat.mat * float(m.pmat.scale) for m in bc_mats])
    rows = []
    for m in bc_mats:
        sc = np.asarray(m.scale, dtype=float)
        if b.dim() == 1:
            rows.append(np.full(1, float(sc.reshape(-1)[0])))
        else:
            full = np.broadcast_to(sc, [1 if i == ax else n for i, n in enumerate(b.shape)])
            rows.append(np.moveaxis(full, ax, -1).reshape(-1))  # C order over the other dims
    scales = np.stack(rows)
    return S._bc_update(u, b, sign * W, S.N - 2, scales)
