"""GPU parity: nonhomogeneous Dirichlet / Neumann boundary values (the boundary
basis lift on backward, _add_mass_rhs on forward, the solvers' bc matrices)
against the oracle and against manufactured solutions of the discrete space."""
import numpy as np
import pytest
import torch

from oracle import spectral as S

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


def _bc(kind, a, b):
    return (a, b) if kind == "D" else {"left": {"N": a}, "right": {"N": b}}


def _poly_with_bcs(family, kind, vals, N, rng):
    P = S._poly_cls(family)
    p = P(rng.standard_normal(N))
    d = 0 if kind == "D" else 1
    bnd = [(p.deriv(d) if d else p)(x) for x in (-1.0, 1.0)]
    C = S.bc_basis(family, kind)
    for j in range(2):
        p = p + (vals[j] - bnd[j]) * P(C[j])
    return p


@pytest.mark.parametrize("family,fam", [("chebyshev", "C"), ("legendre", "L")])
@pytest.mark.parametrize("kind", ["D", "N"])
def test_bc_1d_forward_backward(sfa, family, fam, kind):
    print(f"[bc 1d {family} {kind}]", flush=True)
    N = 24
    vals = (0.75, -1.25)
    V = sfa.FunctionSpace(N, fam, bc=_bc(kind, *vals))
    p = _poly_with_bcs(family, kind, vals, N, np.random.default_rng(3))
    x = np.asarray(V.mesh())
    c = V.forward(dev(p(x)))
    ref = S.forward_bc(family, kind, vals, p(x))
    assert rel(c, ref) < 1e-12
    assert np.allclose(c[N - 2:].cpu().numpy(), vals)
    assert rel(V.backward(c), p(x)) < 1e-12
    # padded backward evaluates the same polynomial on the 3/2 grid
    Pd = V.get_dealiased(1.5)
    xp = np.asarray(Pd.mesh())
    assert len(xp) == 3 * N // 2 and rel(Pd.backward(c), p(xp)) < 1e-12
    # Function of the space carries the boundary values in its slots
    assert np.allclose(sfa.Function(V)[N - 2:].cpu().numpy(), vals)


@pytest.mark.parametrize("kind", ["D", "N"])
def test_bc_helmholtz3d_manufactured(sfa, kind):
    """alpha u - lap u = f on Chebyshev x C2C x R2C with u(+-1) = (a, b) (D) or
    u_x(+-1) = (a, b) (N): u = g(x) + q(x) cos(y + 0.3) cos(2 z) lies in the
    discrete space, so drop-in chain scalar_product -> la.Helmholtz (with its
    bc matrices) -> backward recovers it to roundoff."""
    print(f"[bc helmholtz3d {kind}]", flush=True)
    a, b = 0.6, -0.4
    alpha = 2.0
    N = (20, 16, 16)
    SD = sfa.FunctionSpace(N[0], "C", bc=_bc(kind, a, b))
    K1 = sfa.FunctionSpace(N[1], "F", dtype="D")
    K2 = sfa.FunctionSpace(N[2], "F", dtype="d")
    T = sfa.TensorProductSpace(None, (SD, K1, K2))
    X = T.mesh()
    x = X[0]
    P = np.polynomial.Polynomial
    if kind == "D":
        g = P([(a + b) / 2, (b - a) / 2 - 0.5, 0, 0.5])
        q = P([0, 0, 1, 0, -1])  # x^2 - x^4
    else:
        g = P([0.3, (a + b) / 2, (b - a) / 4])
        q = P([1, 0, -2, 0, 1])  # (1 - x^2)^2
    cyz = np.cos(X[1] + 0.3) * np.cos(2 * X[2])
    u = g(x) + q(x) * cyz
    f = alpha * u - (g.deriv(2)(x) + q.deriv(2)(x) * cyz - 5 * q(x) * cyz)
    f_hat = T.scalar_product(dev(f))
    H = sfa.la.Helmholtz(*sfa.helmholtz_tpmatrices(T, alpha))
    assert len(H.bc_mats) == 2
    u_hat = H(f_hat, sfa.Function(T))
    assert rel(T.backward(u_hat), u) < 1e-11
    # the forward transform of u gives the same coefficients (slots included)
    assert rel(T.forward(dev(u)), u_hat.cpu().numpy()) < 1e-11


@pytest.mark.parametrize("kind", ["D", "N"])
def test_bc_legendre2d_solver_generic(sfa, kind):
    """Legendre x R2C with boundary values through SolverGeneric1ND."""
    print(f"[bc legendre2d {kind}]", flush=True)
    a, b = -0.5, 1.5
    alpha = 1.5
    SD = sfa.FunctionSpace(24, "L", bc=_bc(kind, a, b))
    K = sfa.FunctionSpace(16, "F", dtype="d")
    T = sfa.TensorProductSpace(None, (SD, K))
    X = T.mesh()
    x = X[0]
    P = np.polynomial.Polynomial
    if kind == "D":
        g = P([(a + b) / 2, (b - a) / 2])
        q = P([1, 0, -1])
    else:
        g = P([0.1, (a + b) / 2, (b - a) / 4])
        q = P([1, 0, -2, 0, 1])
    cy = np.cos(2 * X[1])
    u = g(x) + q(x) * cy
    f = alpha * u - (g.deriv(2)(x) + q.deriv(2)(x) * cy - 4 * q(x) * cy)
    f_hat = T.scalar_product(dev(f))
    sol = sfa.la.SolverGeneric1ND(sfa.helmholtz_tpmatrices(T, alpha))
    u_hat = sol(f_hat, sfa.Function(T))
    assert rel(T.backward(u_hat), u) < 1e-11


@pytest.mark.parametrize("form", ["callable", "sympy"])
def test_bc_varying_along_fourier_axes(sfa, form):
    """Dirichlet values u(-1, y, z) = gL(y, z), u(1, y, z) = gR(y, z) given as
    functions (demo/poisson2D.py's ue.subs(x, a) pattern): the slots hold their
    Fourier coefficients, and the drop-in chain recovers the manufactured u."""
    print(f"[bc varying {form}]", flush=True)
    import sympy as sp
    alpha = 1.5
    y, z = sp.symbols("y,z", real=True)
    gLs = 0.5 + 0.3 * sp.cos(y) * sp.sin(2 * z)
    gRs = -0.2 + 0.4 * sp.sin(2 * y + 0.1) + 0.1 * sp.cos(3 * z)
    if form == "sympy":
        bc = (gLs, gRs)
    else:
        fL = sp.lambdify((y, z), gLs, "numpy")
        fR = sp.lambdify((y, z), gRs, "numpy")
        bc = (lambda yy, zz: fL(yy, zz), lambda yy, zz: fR(yy, zz))
    SD = sfa.FunctionSpace(20, "C", bc=bc)
    T = sfa.TensorProductSpace(None, (SD, sfa.FunctionSpace(16, "F", dtype="D"),
                                      sfa.FunctionSpace(16, "F", dtype="d")))
    X = T.mesh()
    x, Y, Z = X
    gL = sp.lambdify((y, z), gLs, "numpy")(Y, Z)
    gR = sp.lambdify((y, z), gRs, "numpy")(Y, Z)
    lapL = sp.lambdify((y, z), sp.diff(gLs, y, 2) + sp.diff(gLs, z, 2), "numpy")(Y, Z)
    lapR = sp.lambdify((y, z), sp.diff(gRs, y, 2) + sp.diff(gRs, z, 2), "numpy")(Y, Z)
    q = np.polynomial.Polynomial([0, 0, 1, 0, -1])  # x^2 - x^4
    c = np.cos(Y) * np.cos(2 * Z)
    u = gL * (1 - x) / 2 + gR * (1 + x) / 2 + q(x) * c
    lap = lapL * (1 - x) / 2 + lapR * (1 + x) / 2 + q.deriv(2)(x) * c - 5 * q(x) * c
    f = alpha * u - lap
    f_hat = T.scalar_product(dev(np.broadcast_to(f, T.shape(False)).copy()))
    H = sfa.la.Helmholtz(*sfa.helmholtz_tpmatrices(T, alpha))
    u_hat = H(f_hat, sfa.Function(T))
    ue = np.broadcast_to(u, T.shape(False))
    assert rel(T.backward(u_hat), ue) < 1e-11
    assert rel(T.forward(dev(ue.copy())), u_hat.cpu().numpy()) < 1e-11


def test_bc_time_dependent_helmholtz2d(sfa):
    """u(x, y, t) = a(t, y) (1 - x)/2 + b(t) (1 + x)/2 + q(x) cos(2y) on
    Chebyshev-D x R2C with sympy boundary values in t (and y): one solver,
    built once, recovers u at several times after bc.update(t) (the
    RayleighBenard demo pattern, demo/RayleighBenard.py:80)."""
    print("[bc time-dependent]", flush=True)
    import sympy as sp
    t, y = sp.symbols("t,y", real=True)
    aS = sp.sin(t) * sp.cos(y) + 0.5
    bS = sp.cos(2 * t)
    alpha = 2.0
    SD = sfa.FunctionSpace(20, "C", bc=(aS, bS))
    T = sfa.TensorProductSpace(None, (SD, sfa.FunctionSpace(16, "F", dtype="d")))
    X = T.mesh()
    x, Y = X
    q = np.polynomial.Polynomial([0, 0, 1, 0, -1])
    H = sfa.la.Helmholtz(*sfa.helmholtz_tpmatrices(T, alpha))
    for tt in (0.0, 0.7, 1.9):
        SD.bc.update(tt)
        a = np.sin(tt) * np.cos(Y) + 0.5
        b = np.cos(2 * tt)
        u = a * (1 - x) / 2 + b * (1 + x) / 2 + q(x) * np.cos(2 * Y)
        lap = -np.sin(tt) * np.cos(Y) * (1 - x) / 2 + q.deriv(2)(x) * np.cos(2 * Y) \
            - 4 * q(x) * np.cos(2 * Y)
        f = alpha * u - lap
        f_hat = T.scalar_product(dev(np.broadcast_to(f, T.shape(False)).copy()))
        u_hat = H(f_hat, sfa.Function(T))
        ue = np.broadcast_to(u, T.shape(False))
        assert rel(T.backward(u_hat), ue) < 1e-11, tt
        assert rel(T.forward(dev(ue.copy())), u_hat.cpu().numpy()) < 1e-11, tt


@pytest.mark.parametrize("N", [(64, 16, 32), (64, 16, 256)])
@pytest.mark.parametrize("kind", ["D", "N"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bc_fused_solver_matches_chain(sfa, kind, precision, N):
    """The fused HelmholtzSolver3D with nonhomogeneous boundary values
    (u = u_0 + lift, the lift added by the final z C2R's stores) returns the
    drop-in chain's solution (scalar_product -> la.Helmholtz with its bc
    matrices -> backward) for a random smooth f, recovers a manufactured
    solution of the discrete space, and adds the lift to every member of a
    stack; clearing the registered lift gives the homogeneous solve back."""
    print(f"[bc fused {kind} {precision} {N}]", flush=True)
    a, b = 0.6, -0.4
    alpha = 2.0
    SD = sfa.FunctionSpace(N[0], "C", bc=_bc(kind, a, b))
    T = sfa.TensorProductSpace(None, (SD, sfa.FunctionSpace(N[1], "F", dtype="D"),
                                      sfa.FunctionSpace(N[2], "F", dtype="d")))
    X = T.mesh()
    x = X[0]
    P = np.polynomial.Polynomial
    if kind == "D":
        g = P([(a + b) / 2, (b - a) / 2 - 0.5, 0, 0.5])
        q = P([0, 0, 1, 0, -1])
    else:
        g = P([0.3, (a + b) / 2, (b - a) / 4])
        q = P([1, 0, -2, 0, 1])
    cyz = np.cos(X[1] + 0.3) * np.cos(2 * X[2])
    u_ex = g(x) + q(x) * cyz
    f_ex = alpha * u_ex - (g.deriv(2)(x) + q.deriv(2)(x) * cyz - 5 * q(x) * cyz)
    Hf = sfa.HelmholtzSolver3D(T, alpha, precision=precision)
    u = Hf(dev(np.broadcast_to(f_ex, T.shape(False))))
    assert rel(u, np.broadcast_to(u_ex, T.shape(False))) < 1e-11
    # random smooth f against the chain on the same space
    rng = np.random.default_rng(7)
    c = rng.standard_normal((N[0], N[1], N[2] // 2 + 1)) * \
        (1.0 + np.arange(N[0])[:, None, None] ** 2) ** -1.5 * \
        (1.0 + np.abs(np.fft.fftfreq(N[1], 1 / N[1]))[None, :, None] ** 2) ** -1.5 * \
        (1.0 + np.arange(N[2] // 2 + 1)[None, None, :] ** 2) ** -1.5
    f = np.fft.irfftn(c, s=(N[1], N[2]), axes=(1, 2)) * (N[1] * N[2])
    f = np.ascontiguousarray(S.cheb_backward(f, axis=0))
    H = sfa.la.Helmholtz(*sfa.helmholtz_tpmatrices(T, alpha))
    u_chain = T.backward(H(T.scalar_product(dev(f)), sfa.Function(T))).cpu().numpy()
    u = Hf(dev(f))
    assert rel(u, u_chain) < (1e-12 if precision == "fp64" else 1e-10)
    U = Hf.solve_many(dev(np.stack([f, f_ex * np.ones(T.shape(False))])))
    assert rel(U[0], u_chain) < (1e-12 if precision == "fp64" else 1e-10)
    assert rel(U[1], np.broadcast_to(u_ex, T.shape(False))) < 1e-11
    # without the lift: u_0 = u - L
    L = Hf.lift()
    sfa._lib.call("sfa_helmholtz3d_set_lift", sfa._lib.ptr(Hf.plan), None)
    u0 = torch.empty_like(L)
    Hf._solve(dev(f), u0, 1, Hf.work)
    assert rel(u0 + L, u_chain) < (1e-12 if precision == "fp64" else 1e-10)
    sfa._lib.call("sfa_helmholtz3d_set_lift", sfa._lib.ptr(Hf.plan), sfa._lib.ptr(L))


def test_bc_fused_solver_time_dependent(sfa):
    """Time-dependent Dirichlet values that vary along y: after bc.update(t)
    the fused solver recomputes its lift and matches the manufactured u."""
    print("[bc fused time-dependent]", flush=True)
    import sympy as sp
    t, y = sp.symbols("t,y", real=True)
    aS = sp.sin(t) * sp.cos(y) + 0.5
    bS = sp.cos(2 * t)
    alpha = 2.0
    SD = sfa.FunctionSpace(64, "C", bc=(aS, bS))
    T = sfa.TensorProductSpace(None, (SD, sfa.FunctionSpace(16, "F", dtype="D"),
                                      sfa.FunctionSpace(16, "F", dtype="d")))
    x, Y, Z = T.mesh()
    q = np.polynomial.Polynomial([0, 0, 1, 0, -1])
    Hf = sfa.HelmholtzSolver3D(T, alpha)
    for tt in (0.0, 0.7, 1.9):
        SD.bc.update(tt)
        av = np.sin(tt) * np.cos(Y) + 0.5
        bv = np.cos(2 * tt)
        cz = np.cos(2 * Y) * np.cos(Z)
        u = av * (1 - x) / 2 + bv * (1 + x) / 2 + q(x) * cz
        lap = -np.sin(tt) * np.cos(Y) * (1 - x) / 2 + q.deriv(2)(x) * cz - 5 * q(x) * cz
        f = np.broadcast_to(alpha * u - lap, T.shape(False)).copy()
        assert rel(Hf(dev(f)), np.broadcast_to(u, T.shape(False))) < 1e-11, tt


@pytest.mark.parametrize("n,nrep,offset", [(1000, 3, 0), (777, 2, 0), (64, 5, 1)])
def test_axpy_stack(sfa, n, nrep, offset):
    """sfa_axpy: y[r n + i] += a x[i] over a stack, the 16-byte form (even n,
    aligned) and the scalar form (odd n, or a misaligned view)."""
    print(f"[axpy {n} {nrep} {offset}]", flush=True)
    g = torch.Generator().manual_seed(n + nrep)
    xb = torch.randn(n + offset, dtype=torch.float64, generator=g).cuda()
    yb = torch.randn(n * nrep + offset, dtype=torch.float64, generator=g).cuda()
    x, y = xb[offset:], yb[offset:]
    ref = (y.cpu().view(nrep, n) + 0.75 * x.cpu()[None, :]).reshape(-1)
    sfa._lib.call("sfa_axpy", sfa._lib.ptr(x), sfa._lib.ptr(y), 0.75, n, nrep, sfa._lib.stream_ptr())
    torch.cuda.synchronize()
    # the device may fuse the multiply-add (one rounding): within an ulp
    assert torch.allclose(y.cpu(), ref, rtol=4e-16, atol=4e-16)
