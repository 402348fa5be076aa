"""GPU parity: batched uneven 2D Legendre x Fourier Helmholtz solves in one
persistent launch vs the FP64 oracle (per problem, legendre_fourier_2d)."""
import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sfa():
    import shenfun_amd
    return shenfun_amd


def _rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_mfma_f64_layout(sfa):
    """The FP64 MFMA 16x16x4 accumulator layout the batched GEMMs assume:
    D[i][j] in register r of lane l with i = 4 r + l // 16, j = l % 16."""
    from shenfun_amd import _lib
    print("[mfma layout probe]", flush=True)
    out = torch.zeros(256, dtype=torch.float64, device="cuda")
    _lib.call("sfa_mfma_f64_layout_probe", _lib.ptr(out), _lib.stream_ptr())
    got = out.cpu().numpy().reshape(64, 4)
    i = np.arange(16)[:, None]
    k = np.arange(4)[None, :]
    A = (i + 100 * k).astype(float)                       # 16 x 4
    B = (np.arange(16)[None, :] + 1000 * np.arange(4)[:, None]).astype(float)  # 4 x 16
    D = A @ B
    exp = np.zeros((64, 4))
    for lane in range(64):
        for r in range(4):
            exp[lane, r] = D[4 * r + lane // 16, lane % 16]
    assert np.array_equal(got, exp)


PROBLEMS = [(32, 32, 1.0, "D"), (34, 46, 0.0, "D"), (64, 64, 5.3, "N"), (100, 38, 17.0, "D"),
            (130, 100, 0.5, "N"), (256, 128, 50.0, "D"), (512, 256, 99.0, "D"),
            (48, 250, 3.0, "N"), (8, 4, 1.0, "D"), (10, 6, 2.0, "N"), (510, 34, 0.25, "N")]


@pytest.fixture(scope="module")
def batch(sfa):
    from shenfun_amd.workloads import batch2d_rhs
    rng = np.random.default_rng(7)
    fs = [batch2d_rhs(nx, ny, rng) for nx, ny, _, _ in PROBLEMS]
    H = sfa.BatchedHelmholtz2D([p[0] for p in PROBLEMS], [p[1] for p in PROBLEMS],
                               [p[2] for p in PROBLEMS], [p[3] for p in PROBLEMS])
    F = H.pack(fs)
    return H, F, fs


def test_batch2d_matches_oracle(sfa, batch):
    H, F, fs = batch
    print("[batch2d vs oracle]", flush=True)
    U = H(F)
    for (nx, ny, al, bc), f, u in zip(PROBLEMS, fs, H.unpack(U)):
        u_ref = S.legendre_fourier_2d(f, al, bc)[0]
        err = _rel(u, u_ref)
        print(f"  Nx={nx} Ny={ny} alpha={al} bc={bc}: max rel err {err:.2e}", flush=True)
        assert err < 1e-11


def test_batch2d_variants_and_schedules_bitwise(sfa, batch):
    """Every tile variant and workgroup count sums each output in the same
    order (K steps ascending, one MFMA chain per element): identical bits."""
    from shenfun_amd.batched import tile_variants
    H, F, _ = batch
    print("[batch2d variants]", flush=True)
    H.build(H.cus)
    H.variant = 0
    ref = H(F).clone()
    for nb in (7, H.cus, 2 * H.cus):
        H.build(nb)
        for v in range(len(tile_variants())):
            H.variant = v
            print(f"  nblocks={nb} variant={v}", flush=True)
            assert torch.equal(H(F), ref)
    H.build(H.cus)
    H.variant = 0


def test_batch2d_alone_equals_in_batch(sfa):
    """A problem solved alone gives the same bits as inside a batch of others."""
    from shenfun_amd.workloads import batch2d_problems, batch2d_rhs
    nx, ny, al, rng = batch2d_problems(40, seed=3)
    fs = [batch2d_rhs(int(a), int(b), rng) for a, b in zip(nx, ny)]
    H = sfa.BatchedHelmholtz2D(nx, ny, al, "D")
    U = H(H.pack(fs))
    parts = H.unpack(U)
    for i in (0, 7, 19, 39):
        print(f"[alone {i}: {nx[i]}x{ny[i]}]", flush=True)
        H1 = sfa.BatchedHelmholtz2D(nx[i:i + 1], ny[i:i + 1], al[i:i + 1], "D", nblocks=1)
        u1 = H1(H1.pack([fs[i]]))
        assert torch.equal(u1.view(int(nx[i]), int(ny[i])), parts[i])


def test_batch2d_graph_replay(sfa, batch):
    """The launch captured in a HIP graph: every replay equals the eager result."""
    H, F, _ = batch
    print("[batch2d graph]", flush=True)
    U_eager = H(F).clone()
    U = torch.empty_like(F)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        H(F, U)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        H(F, U)
    for rep in range(3):
        U.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(U, U_eager), f"replay {rep}"
        (F * 2.0).sum().item()  # eager work between replays


def test_batch2d_c4_workload_sample(sfa):
    """First problems of the 4096-problem seeded workload vs the oracle."""
    from shenfun_amd.workloads import batch2d_problems, batch2d_rhs
    nx, ny, al, rng = batch2d_problems(4096, seed=0)
    n = 12
    fs = [batch2d_rhs(int(a), int(b), rng) for a, b in zip(nx[:n], ny[:n])]
    H = sfa.BatchedHelmholtz2D(nx[:n], ny[:n], al[:n], "D")
    U = H(H.pack(fs))
    for i, u in enumerate(H.unpack(U)):
        u_ref = S.legendre_fourier_2d(fs[i], al[i], "D")[0]
        err = _rel(u, u_ref)
        print(f"[c4 sample {i}: {nx[i]}x{ny[i]} alpha={al[i]:.2f}] err {err:.2e}", flush=True)
        assert err < 1e-11


def test_batch2d_wide_equals_persistent(sfa, batch):
    """The tile-parallel mode gives the persistent kernel's bits, on the mixed
    batch and on a lone large problem, and on repeated calls."""
    from shenfun_amd.workloads import batch2d_rhs
    H, F, _ = batch
    print("[batch2d wide: mixed batch]", flush=True)
    H.build(H.cus)
    H.variant = 0
    ref = H(F).clone()
    W = sfa.BatchedHelmholtz2D(H.Nx, H.Ny, H.alpha, ["D" if b == 0 else "N" for b in H.bc], mode="wide")
    for rep in range(2):
        assert torch.equal(W(F), ref), f"call {rep}"
    # (96, 400), (70, 600), (512, 512): more solve chains (2 Ny/2 + 2) than
    # the workgroup's 256 threads
    for nx, ny, al, bc in [(512, 512, 3.0, "D"), (390, 72, 0.75, "N"), (96, 400, 2.0, "N"),
                           (70, 600, 1.5, "D")]:
        print(f"[batch2d wide: lone {nx}x{ny} {bc}]", flush=True)
        f = batch2d_rhs(nx, ny, np.random.default_rng(nx))
        P1 = sfa.BatchedHelmholtz2D([nx], [ny], [al], bc)
        W1 = sfa.BatchedHelmholtz2D([nx], [ny], [al], bc, mode="wide")
        F1 = P1.pack([f])
        u = W1(F1)
        assert torch.equal(u, P1(F1))
        err = _rel(u.view(nx, ny), S.legendre_fourier_2d(f, al, bc)[0])
        print(f"  max rel err {err:.2e}", flush=True)
        assert err < 1e-11


def test_batch2d_wide_large_tiles_and_autotune(sfa):
    """A batch large enough for the tile-parallel form's 64 x 64 tiles (at
    least 4 per CU of them per stage): the same bits as the persistent kernel,
    a sample of problems against the oracle, and the autotuner, which times
    both forms, leaves a configuration that still gives those bits."""
    from shenfun_amd.workloads import batch2d_problems, batch2d_rhs
    print("[batch2d wide 64x64 tiles: 600 problems]", flush=True)
    nx, ny, al, rng = batch2d_problems(600, 3)
    fs = [batch2d_rhs(int(a), int(b), rng) for a, b in zip(nx, ny)]
    P = sfa.BatchedHelmholtz2D(nx, ny, al, "D")
    F = P.pack(fs)
    ref = P(F).clone()
    W = sfa.BatchedHelmholtz2D(nx, ny, al, "D", mode="wide")
    assert torch.equal(W(F), ref)
    for i in range(0, 600, 97):
        err = _rel(W.unpack(ref)[i], S.legendre_fourier_2d(fs[i], float(al[i]), "D")[0])
        assert err < 1e-11, (i, err)
    tune = P.autotune(F, blocks_per_cu=(2,), variants=[0, 3], reps=1)
    modes = {r["mode"] for r in tune["table"]}
    print(f"  autotune: {tune['best']['label']} ({sorted(modes)})", flush=True)
    assert modes == {"persistent", "wide"}
    assert torch.equal(P(F), ref)
