"""Host logic of the serving loop (no GPU): the LRU factor cache with
refcounts, the pending queues, and the per-request outcomes of malformed,
oversized, over-budget, failing, timed-out and cancelled requests, on a
server whose launches and events are replaced by host stand-ins."""
import numpy as np
import torch

from shenfun_amd.serving import FactorCache, Request, Server


def test_factor_cache_lru_and_refcounts():
    c = FactorCache(budget_bytes=100)
    built = []

    def mk(name, nb):
        def build():
            built.append(name)
            return name, nb
        return build

    assert c.get("a", mk("a", 40)) == "a"
    assert c.get("b", mk("b", 40)) == "b"
    c.acquire(["a"])
    assert c.get("c", mk("c", 40)) == "c"  # over budget: evicts b (a is in use)
    assert "b" not in c.items and "a" in c.items
    assert c.get("a", mk("a", 40)) == "a" and built.count("a") == 1  # hit
    c.release(["a"])
    c.get("d", mk("d", 40))
    assert c.nbytes <= 100
    assert c.hits == 1 and c.misses == 4


class _Event:
    """Completes after a fixed number of polls."""

    def __init__(self, polls=2):
        self.left = polls

    def query(self):
        self.left -= 1
        return self.left < 0


class HostServer(Server):
    """Server with host stand-ins for the launches: records the batches and
    fails the launch of any request whose key carries alpha == 13.0."""

    def __init__(self, **kw):
        super().__init__(device="cpu", streams=[object() for _ in range(kw.pop("slots", 2))], **kw)
        self.launched = []

    def _launch(self, batch, stream, slot):
        if any(float(r.key[-1]) == 13.0 for r in batch):
            raise RuntimeError("sfa_x failed (3): injected")
        self.launched.append([r.rid for r in batch])
        for r in batch:
            r.u = r.f
        return (), []

    def _record(self, stream):
        return _Event()

    def _poll(self, now):
        # no leg2d plan registry on the host
        still = []
        for ev, slot, reqs, nbytes, keys, keep in self.inflight:
            if ev.query():
                for r in reqs:
                    self._finish(r, "ok", now())
                self.free_slots.append(slot)
                self.mem_used -= nbytes
            else:
                still.append((ev, slot, reqs, nbytes, keys, keep))
        self.inflight = still


def _f(*shape):
    return torch.zeros(shape, dtype=torch.float64)


def test_pending_queues_fifo_batching():
    """The oldest pending request heads each batch; 1D requests join their
    class (N, bc) whatever their alpha, 2D requests of any size join one
    launch, 3D requests join only their class (shape, bc, alpha); caps hold;
    every request is served exactly once."""
    k1a, k1b = (64, "D", 1.0), (64, "N", 2.0)
    k3a, k3b = ((32, 32, 32), "D", 1.0), ((32, 32, 32), "N", 1.0)
    seq = [("cheb1d", k1a), ("leg2d", (32, 32, "D", 1.0)), ("cheb1d", k1b), ("cheb1d", (64, "D", 5.0)),
           ("cheb3d", k3a), ("leg2d", (64, 40, "N", 2.0)), ("cheb1d", k1a),
           ("cheb1d", k1a), ("cheb1d", k1a), ("leg2d", (48, 36, "D", 3.0)), ("cheb3d", k3b),
           ("cheb3d", k3a), ("cheb3d", k3a), ("cheb3d", k3a)]
    shapes = {"cheb1d": lambda k: (k[0],), "leg2d": lambda k: (k[0], k[1]), "cheb3d": lambda k: k[0]}
    reqs = [Request(i, fam, key, _f(*shapes[fam](key)), arrival=0.0) for i, (fam, key) in enumerate(seq)]
    s = HostServer(slots=1, max_batch_1d=3, max_batch_2d=4, max_batch_3d=2)
    s.run(reqs, realtime=False)
    assert s.launched == [[0, 3, 6], [1, 5, 9], [2], [4, 11], [7, 8], [10], [12, 13]]
    assert all(r.status == "ok" for r in reqs)


def test_malformed_oversized_and_failing_requests():
    """Malformed requests end "invalid", a request larger than the memory
    budget "refused", a failing launch "error" for its own batch only; the
    rest are served."""
    good = [Request(0, "cheb1d", (64, "D", 1.0), _f(64)),
            Request(1, "leg2d", (32, 16, "N", 2.0), _f(32, 16)),
            Request(2, "cheb3d", ((16, 8, 8), "D", 0.5), _f(16, 8, 8))]
    bad = [Request(10, "cheb4d", (64, "D", 1.0), _f(64)),                 # unknown family
           Request(11, "cheb1d", (64, "X", 1.0), _f(64)),                 # bad bc
           Request(12, "cheb1d", (64, "D", float("nan")), _f(64)),        # bad alpha
           Request(13, "cheb1d", (64, "D", -1.0), _f(64)),                # negative alpha
           Request(14, "leg2d", (33, 16, "D", 1.0), _f(33, 16)),          # odd Nx
           Request(15, "leg2d", (32, 16, "D", 1.0), _f(16, 32)),          # wrong shape
           Request(16, "cheb1d", (64, "D", 1.0), _f(64).float()),         # wrong dtype
           Request(17, "cheb1d", (64, "D", 1.0), np.zeros(64)),           # not a tensor
           Request(18, "cheb3d", ((16, 8, 7), "D", 1.0), _f(16, 8, 7)),   # odd Nz
           Request(19, "leg2d", (32, 16, "N", 0.0), _f(32, 16))]          # singular
    huge = Request(20, "cheb3d", ((512, 512, 512), "D", 1.0), _f(1, 1, 1).expand(512, 512, 512))
    failing = [Request(30, "cheb1d", (128, "D", 13.0), _f(128)),
               Request(31, "cheb1d", (128, "D", 13.0), _f(128))]
    s = HostServer(slots=2, mem_budget=64 << 20)
    reqs = good + bad + [huge] + failing
    s.run(reqs, realtime=False)
    assert all(r.status == "ok" for r in good)
    assert all(r.status.startswith("invalid: ") for r in bad), [r.status for r in bad]
    assert huge.status.startswith("refused: ")
    assert all(r.status.startswith("error: RuntimeError") and r.u is None for r in failing)
    st = Server.stats(reqs)
    assert st["served"] == 3 and st["status"] == {"ok": 3, "invalid": 10, "refused": 1, "error": 2}


def test_over_budget_batches_are_shortened_not_refused():
    """A budget that holds one request but not the whole class admits the
    class one request at a time; nothing is refused."""
    reqs = [Request(i, "cheb1d", (1024, "D", 1.0), _f(1024)) for i in range(5)]
    one = Server._bytes(reqs[:1])
    s = HostServer(slots=4, mem_budget=one + one // 2)
    s.run(reqs, realtime=False)
    assert s.launched == [[0], [1], [2], [3], [4]]
    assert all(r.status == "ok" for r in reqs)


def test_timeout_and_cancel():
    """A pending request whose deadline passes ends "timeout"; a cancelled
    pending request ends "cancelled"; admitted ones cannot be cancelled."""
    s = HostServer(slots=1)
    a = Request(0, "cheb1d", (64, "D", 1.0), _f(64))
    b = Request(1, "cheb1d", (128, "D", 1.0), _f(128), deadline=0.5)
    c = Request(2, "cheb1d", (256, "D", 1.0), _f(256))
    for r in (a, b, c):
        s._push(r)
    assert s.cancel(2) and not s.cancel(2) and not s.cancel(99)
    s._admit(lambda: 0.0)      # the one slot takes a
    assert a.t_admit is not None and not s.cancel(0)
    s._expire(1.0)             # b's deadline passes while it waits
    assert b.status == "timeout" and c.status == "cancelled"
    assert s._npending == 0
    for _ in range(5):
        s._poll(lambda: 2.0)
        s._admit(lambda: 2.0)
    assert a.status == "ok" and s.launched == [[0]]


def test_cache_builds_on_the_missing_batch_stream(monkeypatch):
    """A miss builds inside the stream context of the batch that missed and
    records a ready event on that stream; a hit on another stream waits for
    that event before using the entry."""
    import contextlib
    from shenfun_amd import serving

    current = []
    recorded, waited = [], []

    @contextlib.contextmanager
    def fake_stream(s):
        current.append(s)
        yield
        current.pop()

    class FakeEvent:
        def record(self, s):
            recorded.append(s)

    class FakeStream:
        def __init__(self, name):
            self.name = name

        def wait_event(self, ev):
            waited.append((self.name, ev))

    monkeypatch.setattr(serving.torch.cuda, "stream", fake_stream)
    monkeypatch.setattr(serving.torch.cuda, "Event", FakeEvent)
    c = FactorCache(budget_bytes=1 << 20)
    s1, s2 = FakeStream("s1"), FakeStream("s2")
    seen = []

    def build():
        seen.append(current[-1] if current else None)
        return "obj", 10

    assert serving._use(s1, c.lookup("k", build, s1)) == "obj"
    assert seen == [s1] and recorded == [s1]
    obj = serving._use(s2, c.lookup("k", build, s2))
    assert obj == "obj" and seen == [s1]  # hit: no rebuild
    assert [w[0] for w in waited] == ["s1", "s2"] and isinstance(waited[1][1], FakeEvent)
