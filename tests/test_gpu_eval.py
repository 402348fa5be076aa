"""GPU parity: point evaluation (SpectralBase.eval, TensorProductSpace.eval)
against closed-form functions that lie in the discrete spaces, and against
backward at the quadrature points (the reference's test_eval,
tests/test_transforms.py:135-155)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else a
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def sfa():
    import shenfun_amd
    return shenfun_amd


SPACES_1D = [
    ("C", {}), ("C", {"quad": "GL"}), ("L", {}),
    ("C", {"bc": (0, 0)}), ("L", {"bc": (0, 0)}),
    ("C", {"bc": (0.5, -1.5)}), ("L", {"bc": (0.25, 2.0)}),
    ("C", {"bc": {"left": {"N": 0.3}, "right": {"N": -0.7}}}),
    ("L", {"bc": {"left": {"N": 0}, "right": {"N": 0}}}),
    ("C", {"bc": (0, 0, 0, 0)}), ("L", {"bc": (0, 0, 0, 0)}),
    ("F", {"dtype": "d"}), ("F", {"dtype": "D"}),
]


@pytest.mark.parametrize("N", [16, 15])
@pytest.mark.parametrize("fam,kw", SPACES_1D)
def test_eval_1d_matches_backward_and_function(sfa, fam, kw, N):
    """eval at the quadrature points equals backward; at random points it
    equals the function whose samples were forward-transformed (the function
    lies in the space, so forward recovers it exactly)."""
    print(f"[eval 1d {fam} {kw} N={N}]", flush=True)
    rng = np.random.default_rng(N)
    V = sfa.FunctionSpace(N, fam, **kw)
    x = np.asarray(V.mesh())
    if fam == "F":
        K = np.fft.fftfreq(N, 1.0 / N)
        K = K[np.abs(K) < N / 2]  # drop the Nyquist mode
        a = rng.standard_normal(K.size) + 1j * rng.standard_normal(K.size)
        if kw["dtype"] == "d":
            f = lambda t: np.real(np.exp(1j * np.outer(t, K)) @ a)  # noqa: E731
        else:
            f = lambda t: np.exp(1j * np.outer(t, K)) @ a  # noqa: E731
        xs = rng.uniform(0, 2 * np.pi, 40)
    else:
        # a polynomial of the space: backward of random coefficients
        c = sfa.Function(V)
        d = V.dim()
        c[:d] = dev(rng.standard_normal(d))
        f_on_mesh = V.backward(c).cpu().numpy()
        P = np.polynomial.Chebyshev if fam == "C" else np.polynomial.Legendre
        p = P.fit(x, f_on_mesh, N - 1, domain=[-1, 1])
        f = p
        xs = rng.uniform(-1, 1, 40)
    u_hat = V.forward(dev(f(x)))
    assert rel(V.eval(x, u_hat), V.backward(u_hat).cpu().numpy()) < 1e-12
    out = V.eval(xs, u_hat)
    assert rel(out, f(xs)) < 1e-11
    # output_array is filled in place
    buf = torch.zeros(xs.size, dtype=out.dtype, device="cuda")
    assert V.eval(xs, u_hat, output_array=buf) is buf and rel(buf, f(xs)) < 1e-11


@pytest.mark.parametrize("axes", [None, (1, 0, 2)])
def test_eval_tps3d_manufactured(sfa, axes):
    """u = g(x) + q(x) cos(y + 0.3) cos(2z) + sin(3z) x on Chebyshev D(a, b) x
    C2C x R2C, evaluated at random points."""
    print(f"[eval tps3d axes={axes}]", flush=True)
    a, b = 0.6, -0.4
    SD = sfa.FunctionSpace(20, "C", bc=(a, b))
    T = sfa.TensorProductSpace(None, (SD, sfa.FunctionSpace(16, "F", dtype="D"),
                                      sfa.FunctionSpace(16, "F", dtype="d")), axes=axes)
    P = np.polynomial.Polynomial
    g = P([(a + b) / 2, (b - a) / 2 - 0.5, 0, 0.5])
    q = P([0, 0, 1, 0, -1])

    def ue(x, y, z):
        return g(x) + q(x) * np.cos(y + 0.3) * np.cos(2 * z) + (1 - x * x) * np.sin(3 * z)

    X = T.mesh()
    u_hat = T.forward(dev(np.broadcast_to(ue(*X), T.shape(False)).copy()))
    rng = np.random.default_rng(1)
    pts = np.stack([rng.uniform(-1, 1, 300), rng.uniform(0, 2 * np.pi, 300),
                    rng.uniform(0, 2 * np.pi, 300)])
    out = T.eval(pts, u_hat)
    assert out.dtype == torch.float64
    assert rel(out, ue(*pts)) < 1e-11
    # numpy coefficients are accepted
    assert rel(T.eval(pts, u_hat.cpu().numpy()), ue(*pts)) < 1e-11


def test_eval_tps2d_legendre_c2c_complex(sfa):
    """Complex physical data on Legendre x C2C."""
    print("[eval tps2d complex]", flush=True)
    T = sfa.TensorProductSpace(None, (sfa.FunctionSpace(18, "L"),
                                      sfa.FunctionSpace(12, "F", dtype="D")))

    def ue(x, y):
        return (x ** 3 - 0.5 * x) * np.exp(2j * y) + (1 + x) * np.exp(-3j * y) + 0.25

    X = T.mesh()
    u_hat = T.forward(dev(ue(*X)))
    rng = np.random.default_rng(2)
    pts = np.stack([rng.uniform(-1, 1, 100), rng.uniform(0, 2 * np.pi, 100)])
    out = T.eval(pts, u_hat)
    assert out.dtype == torch.complex128 and rel(out, ue(*pts)) < 1e-11


def test_eval_tps_at_mesh_equals_backward(sfa):
    """eval on every quadrature point equals backward, for Legendre N x Chebyshev
    SB x R2C of odd length."""
    print("[eval tps mesh]", flush=True)
    T = sfa.TensorProductSpace(None, (sfa.FunctionSpace(10, "L", bc={"left": {"N": 0}, "right": {"N": 0}}),
                                      sfa.FunctionSpace(12, "C", bc=(0, 0, 0, 0)),
                                      sfa.FunctionSpace(9, "F", dtype="d")))
    rng = np.random.default_rng(3)
    u = T.backward(T.forward(dev(rng.standard_normal(T.shape(False)))))
    u_hat = T.forward(u)
    X = [np.broadcast_to(m, T.shape(False)).ravel() for m in T.mesh()]
    assert rel(T.eval(np.stack(X), u_hat), u.cpu().numpy().ravel()) < 1e-12


def test_eval_complex_coefficients_on_polynomial_space(sfa):
    """Complex coefficients of a Chebyshev space (a line of a tensor-product
    array) keep their imaginary part."""
    print("[eval complex poly]", flush=True)
    V = sfa.FunctionSpace(12, "C", bc=(0, 0))
    rng = np.random.default_rng(4)
    c = np.zeros(12, complex)
    c[:10] = rng.standard_normal(10) + 1j * rng.standard_normal(10)
    xs = rng.uniform(-1, 1, 30)
    ref = V.evaluate_basis_all(xs) @ c
    out = V.eval(xs, dev(c))
    assert out.dtype == torch.complex128 and rel(out, ref) < 1e-12
