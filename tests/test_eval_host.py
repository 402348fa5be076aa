"""Host-side basis-function matrices behind eval (no GPU): closed forms of the
orthogonal polynomials, the composite stencils' boundary values, and the R2C
weights against numpy's inverse real FFT."""
import numpy as np
import pytest

import shenfun_amd as sfa


def test_chebyshev_orthogonal_is_cos():
    x = np.linspace(-1, 1, 11)
    P = sfa.FunctionSpace(9, "C").evaluate_basis_all(x)
    assert np.allclose(P, np.cos(np.outer(np.arccos(x), np.arange(9))), atol=1e-13)


def test_legendre_orthogonal_recurrence():
    x = np.linspace(-1, 1, 7)
    P = sfa.FunctionSpace(6, "L").evaluate_basis_all(x)
    L3 = 0.5 * (5 * x ** 3 - 3 * x)
    assert np.allclose(P[:, 3], L3) and np.allclose(P[:, 0], 1)


@pytest.mark.parametrize("fam", ["C", "L"])
@pytest.mark.parametrize("bc,deriv,vals", [((0.5, -2.0), 0, (0.5, -2.0)),
                                           ({"left": {"N": 1.5}, "right": {"N": -0.5}}, 1, (1.5, -0.5))])
def test_composite_boundary_values(fam, bc, deriv, vals):
    """Basis functions k < dim vanish (value or slope) at +-1; the boundary
    columns carry unit value/slope at their own end, so c[dim:] = the
    boundary values reproduces them."""
    V = sfa.FunctionSpace(12, fam, bc=bc)
    e = np.array([-1.0, 1.0])
    if deriv == 0:
        P = V.evaluate_basis_all(e)
    else:
        h = 1e-6
        P = (V.evaluate_basis_all(e + h) - V.evaluate_basis_all(e - h)) / (2 * h)
        # one-sided points outside [-1, 1] are fine for polynomials
    d = V.dim()
    assert np.allclose(P[:, :d], 0, atol=1e-6)
    assert np.allclose(P[:, d:], np.eye(2), atol=1e-6)
    assert np.allclose(P[:, d:] @ np.array(vals), vals, atol=1e-6)


@pytest.mark.parametrize("N", [8, 9])
def test_r2c_weights_reproduce_irfft(N):
    R = sfa.FunctionSpace(N, "F", dtype="d")
    rng = np.random.default_rng(N)
    c = rng.standard_normal(N // 2 + 1) + 1j * rng.standard_normal(N // 2 + 1)
    c[0] = c[0].real
    if N % 2 == 0:
        c[-1] = c[-1].real
    x = np.arange(N) * 2 * np.pi / N
    u = np.real(R._eval_matrix(x) @ c)
    assert np.allclose(u, np.fft.irfft(c, N) * N, atol=1e-12)


def test_biharmonic_stencil_vanishes_at_ends():
    V = sfa.FunctionSpace(14, "C", bc=(0, 0, 0, 0))
    e = np.array([-1.0, 1.0])
    h = 1e-5
    P = V.evaluate_basis_all(e)
    dP = (V.evaluate_basis_all(e + h) - V.evaluate_basis_all(e - h)) / (2 * h)
    assert np.allclose(P[:, :V.dim()], 0, atol=1e-10)
    assert np.allclose(dP[:, :V.dim()], 0, atol=1e-5)


def test_tensor_eval_checks_shapes():
    T = sfa.TensorProductSpace(None, (sfa.FunctionSpace(8, "C"), sfa.FunctionSpace(8, "F", dtype="d")))
    with pytest.raises(ValueError):
        T.eval(np.zeros((3, 4)), np.zeros(T.shape(True), complex))
    with pytest.raises(ValueError):
        T.eval(np.zeros((2, 4)), np.zeros((8, 4), complex))
