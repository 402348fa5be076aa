"""GPU: the continuous-batching request loop.  Every served result matches the
FP64 oracle, and the same request served alone gives the same result as in
the mix (stated tolerance: max relative difference <= 1e-13; the batched
kernels treat each problem independently, so it is normally 0).  Requests
upload their right-hand side from pinned host memory and carry their own
alpha (1D batches factorise per request, 3D requests plan anew); malformed
and oversized requests end with their status while the rest are served."""
import numpy as np
import pytest
import torch

from oracle import spectral as S

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _oracle(r):
    f = r.f.cpu().numpy()
    if r.family == "cheb1d":
        N, bc, al = r.key
        fh = S.basis_scalar_product("chebyshev", bc, f)
        d0, d1, d2, L = S.helmholtz_lu(bc, N, np.array(-1.0), np.array(al))
        uh = S.helmholtz_solve(fh, d0, d1, d2, L, neumann=(bc == "N"))
        return S.basis_backward("chebyshev", bc, uh)
    if r.family == "leg2d":
        return S.legendre_fourier_2d(f, r.key[3], r.key[2])[0]
    return S.helmholtz3d_cheb_fourier(f, r.key[2], r.key[1])[0]


@pytest.fixture(scope="module")
def served():
    from shenfun_amd.serving import Server, request_stream
    reqs = request_stream(160, rate=1e5, seed=11, classes=(6, 10, 4))
    srv = Server(slots=4)
    print("[serve mix]", flush=True)
    srv.run(reqs, realtime=False)
    return srv, reqs


def test_served_results_match_oracle(served):
    srv, reqs = served
    assert all(r.status == "ok" for r in reqs)
    per = {"cheb1d": 0, "leg2d": 0, "cheb3d": 0}
    for r in reqs:
        # every request has its own alpha: check the first few of each family
        # (3D up to 128^3 points, for the oracle's time)
        if per[r.family] >= 8 or (r.family == "cheb3d" and np.prod(r.key[0]) > 128 ** 3):
            continue
        per[r.family] += 1
        err = _rel(r.u, _oracle(r))
        print(f"[{r.family} {r.key}] err {err:.2e}", flush=True)
        assert err < 1e-11
    assert srv.batches["leg2d"] < sum(r.family == "leg2d" for r in reqs)  # 2D requests were batched


def test_alone_equals_in_mix(served):
    from shenfun_amd.serving import Server
    _, reqs = served
    picks = {}
    for r in reqs:
        picks.setdefault((r.family, r.key), r)
    worst = 0.0
    for r in [v for v in picks.values() if v.family != "cheb3d"][:8] + \
            [v for v in picks.values() if v.family == "cheb3d"][:3]:
        alone = type(r)(r.rid, r.family, r.key, r.f, 0.0)
        Server(slots=1).run([alone], realtime=False)
        d = _rel(alone.u, r.u.cpu().numpy())
        worst = max(worst, d)
        print(f"[alone {r.family} {r.key}] max rel diff {d:.1e}", flush=True)
        assert d <= 1e-13
    print(f"worst alone-vs-mix difference {worst:.1e}")


def test_memory_admission_serialises(served):
    """A memory budget that fits the largest single request but no two of
    them admits one request at a time (nothing is refused); results are
    unchanged."""
    from shenfun_amd.serving import Request, Server
    _, reqs = served
    sub = [Request(r.rid, r.family, r.key, r.f, 0.0) for r in reqs[:40]]
    one = max(Server._bytes([r]) for r in sub)
    srv = Server(slots=4, mem_budget=one)
    srv.run(sub, realtime=False)
    for a, b in zip(sub, reqs[:40]):
        assert a.status == "ok", a.status
        assert _rel(a.u, b.u.cpu().numpy()) <= 1e-13


@pytest.mark.parametrize("upload,per_request", [(True, True), (False, False)])
def test_realtime_stream_latency(served, upload, per_request):
    """Open-loop Poisson arrivals: every request completes, latencies are
    positive and p99 >= p50 (host uploads with alpha per request, and
    resident inputs with alpha per class)."""
    from shenfun_amd.serving import Server, request_stream
    reqs = request_stream(300, rate=5000, seed=12, classes=(6, 10, 4), upload=upload,
                          alpha_per_request=per_request)
    srv = Server(slots=4)
    srv.run(reqs, realtime=True)
    st = Server.stats(reqs)
    print(st, flush=True)
    assert st["served"] == 300 and 0 < st["p50_ms"] <= st["p99_ms"]


def test_statuses_in_a_served_mix(served):
    """Malformed and oversized requests mixed into a stream end "invalid" /
    "refused" at arrival; every other request is served and matches its
    result in the reference mix bitwise."""
    from shenfun_amd.serving import Request, Server
    _, reqs = served
    sub = [Request(r.rid, r.family, r.key, r.f, 0.0) for r in reqs[:30]]
    bad = [Request(1000, "cheb1d", (64, "Q", 1.0), reqs[0].f, 0.0),
           Request(1001, "leg2d", (31, 16, "D", 1.0), torch.zeros(31, 16, dtype=torch.float64), 0.0),
           Request(1002, "cheb3d", ((256, 256, 256), "D", 1.0),
                   torch.zeros(256, 256, 256, dtype=torch.float64), 0.0)]
    srv = Server(slots=4, mem_budget=256 << 20)
    srv.run(sub + bad, realtime=False)
    assert bad[0].status.startswith("invalid") and bad[1].status.startswith("invalid")
    assert bad[2].status.startswith("refused")
    for a, b in zip(sub, reqs[:30]):
        assert a.status == "ok" and _rel(a.u, b.u.cpu().numpy()) <= 1e-13
