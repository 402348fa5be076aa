"""Host side of the mixed Dirichlet/Neumann bases and of the general banded
fallback (no GPU): the stencils against the boundary conditions they encode,
the Galerkin matrices' band structure, solver selection and the C-ABI's
argument checks."""
import numpy as np
import pytest

import shenfun_amd as sfa
from shenfun_amd import _lib

DN = {"left": {"D": 0}, "right": {"N": 0}}
ND = {"left": {"N": 0}, "right": {"D": 0}}


@pytest.mark.parametrize("fam", ["C", "L"])
@pytest.mark.parametrize("bc,kind", [(DN, "DN"), (ND, "ND")])
def test_mixed_stencils_satisfy_their_conditions(fam, bc, kind):
    """DN: phi_k(-1) = phi_k'(1) = 0; ND: phi_k'(-1) = phi_k(1) = 0, for every
    k < N - 2, each phi_k of exact degree k + 2 (numpy polynomial classes)."""
    V = sfa.FunctionSpace(16, fam, bc=bc)
    assert V.dim() == 14 and V.short_name() == kind
    P = np.polynomial.Chebyshev if fam == "C" else np.polynomial.Legendre
    for row in V.stencil_matrix()[:V.dim()]:
        p = P(row)
        left, right = (p(-1.0), p.deriv()(1.0)) if kind == "DN" else (p.deriv()(-1.0), p(1.0))
        assert abs(left) < 1e-12 and abs(right) < 1e-11
        assert row[np.nonzero(row)[0][-1]] != 0


@pytest.mark.parametrize("bc", [DN, ND])
def test_mixed_legendre_galerkin_matrices(bc):
    """Legendre: (phi_k, phi_j'') is diagonal (the bases are built for it) and
    equals -(phi_k', phi_j') by quadrature; the mass matrix has bands -2..2,
    symmetric, and goes to the general banded solver."""
    V = sfa.FunctionSpace(14, "L", bc=bc)
    A = sfa.matrices.GalerkinMatrix(V, V, 2, "A")
    B = sfa.matrices.GalerkinMatrix(V, V, 0, "B")
    assert sorted(A.keys()) == [0]
    assert sorted(B.keys()) == [-2, -1, 0, 1, 2]
    x, w = np.polynomial.legendre.leggauss(40)
    S = V.stencil_matrix()[:V.dim()]
    Vx = np.polynomial.legendre.legvander(x, V.N - 1)
    D1 = np.stack([np.polynomial.legendre.legval(x, np.polynomial.legendre.legder(r)) for r in S])
    phi = Vx @ S.T
    assert np.allclose(A.diags().toarray(), -(D1 * w) @ D1.T, rtol=0, atol=1e-10)
    assert np.allclose(B.diags().toarray(), (phi.T * w) @ phi, rtol=0, atol=1e-12)
    assert B.get_solver() is sfa.la.Solve


def test_mixed_chebyshev_stiffness_refused():
    V = sfa.FunctionSpace(12, "C", bc=DN)
    T = sfa.TensorProductSpace(None, (V, sfa.FunctionSpace(8, "F", dtype="d")))
    with pytest.raises(NotImplementedError):
        sfa.helmholtz_tpmatrices(T, 1.0)


def test_general_band_extent_and_abi_checks():
    assert sfa.la._band_extent([-2, -1, 0, 1, 2]) == (2, 2)
    assert sfa.la._band_extent([0, 3]) == (0, 3)
    assert sfa.la._band_extent([-1, 0]) == (1, 1)
    with pytest.raises(NotImplementedError):
        sfa.la._band_extent([-5, 0, 1])
    L = _lib.lib()
    assert L.sfa_band_lu(None, 8, 5, 1, 1, None) == -1  # nl > 4
    assert L.sfa_band_lu(None, 8, 1, 1, 1, None) == -1  # null data
    assert L.sfa_band_solve(None, 0, 1, _lib.arr([8]), _lib.arr([1]), 0, None, 8, 0, 0, 2, 0,
                            None) == -1  # nu < 1


@pytest.mark.parametrize("fam", ["C", "L"])
@pytest.mark.parametrize("kind", ["DN", "ND"])
def test_mixed_boundary_functions(fam, kind):
    """Boundary functions: slot 0 carries the left condition, slot 1 the right
    one, each with the other end's condition zero (DN: P0, P0 + P1; ND:
    -P0 + P1, P0, the reference's phi_{N-2}, phi_{N-1})."""
    V = sfa.FunctionSpace(10, fam, bc={"left": {kind[0]: 0.7}, "right": {kind[1]: -0.4}})
    assert V.bc_values == (0.7, -0.4) and V.has_nonhomogeneous_bcs
    P = np.polynomial.Chebyshev if fam == "C" else np.polynomial.Legendre
    cond = {"D": lambda p, e: p(e), "N": lambda p, e: p.deriv()(e)}
    for j, row in enumerate(V.bc_basis()):
        p = P(row)
        assert np.isclose(cond[kind[0]](p, -1.0), 1.0 if j == 0 else 0.0)
        assert np.isclose(cond[kind[1]](p, 1.0), 1.0 if j == 1 else 0.0)


def test_non_reference_domains_refused():
    """A domain other than the reference one would scale wavenumbers and
    derivatives; it is refused instead of silently ignored."""
    assert isinstance(sfa.FunctionSpace(8, "F", dtype="d", domain=(0, 2 * np.pi)), sfa.R2C)
    assert sfa.FunctionSpace(8, "C", domain=(-1, 1)).N == 8
    with pytest.raises(NotImplementedError):
        sfa.FunctionSpace(8, "F", dtype="D", domain=(0, 1))
    with pytest.raises(NotImplementedError):
        sfa.FunctionSpace(8, "L", bc=(0, 0), domain=(0, 1))


def test_time_dependent_boundary_values_update():
    """bc.update(t) substitutes t in the values as given (BoundaryValues.update);
    a value vanishing at the current time still makes the space nonhomogeneous."""
    import sympy as sp
    t, y = sp.symbols("t,y", real=True)
    V = sfa.FunctionSpace(10, "C", bc=(sp.sin(t), 2 * sp.cos(t)))
    assert V.bc_values == (0.0, 2.0) and V.has_nonhomogeneous_bcs
    assert np.allclose(V.bc.update(0.5), (np.sin(0.5), 2 * np.cos(0.5)))
    assert np.allclose(V.bc.update(), (np.sin(0.5), 2 * np.cos(0.5)))  # time kept
    W = sfa.FunctionSpace(10, "L", bc=(sp.sin(t) * sp.cos(y), 0))
    assert W.bc_values == (0.0, 0.0) and W.has_nonhomogeneous_bcs
    a, b = W.bc.update(1.0)
    assert b == 0.0 and sp.simplify(a - np.sin(1.0) * sp.cos(y)) == 0
    # a clone (dealiased / test spaces) starts from the values as given
    assert V._clone(10).bc_values == (0.0, 2.0)
