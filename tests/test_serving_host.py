"""Host logic of the serving loop (no GPU): LRU factor cache with refcounts."""
from shenfun_amd.serving import FactorCache


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


def _bare_server(max1d=3, max2d=4):  # max_batch_3d = 2
    """A Server without device state: only the pending-queue logic."""
    import collections
    from shenfun_amd.serving import Server
    s = Server.__new__(Server)
    s.pending = collections.deque()
    s._q1d = {}
    s._q2d = collections.deque()
    s._q3d = {}
    s._npending = 0
    s.max_batch_1d, s.max_batch_2d, s.max_batch_3d = max1d, max2d, 2
    return s


def test_pending_queues_fifo_batching():
    """The oldest pending request heads each batch; 1D requests join only
    their class (N, bc, alpha), 2D requests of any size join one launch, 3D
    requests join only their class (shape, bc, alpha); caps hold; every
    request is admitted exactly once."""
    from shenfun_amd.serving import Request
    k1a, k1b = (64, "D", 1.0), (64, "D", 2.0)
    k3a, k3b = ((32, 32, 32), "D", 1.0), ((32, 32, 32), "N", 1.0)
    seq = [("cheb1d", k1a), ("leg2d", (32, 32, "D", 1.0)), ("cheb1d", k1b), ("cheb1d", k1a),
           ("cheb3d", k3a), ("leg2d", (64, 40, "N", 2.0)), ("cheb1d", k1a),
           ("cheb1d", k1a), ("cheb1d", k1a), ("leg2d", (48, 36, "D", 3.0)), ("cheb3d", k3b),
           ("cheb3d", k3a), ("cheb3d", k3a), ("cheb3d", k3a)]
    reqs = [Request(i, fam, key, None, arrival=float(i)) for i, (fam, key) in enumerate(seq)]
    s = _bare_server()
    for r in reqs:
        s._push(r)
    batches = []
    while s._npending:
        b = s._next_batch()
        batches.append([r.rid for r in b])
        s._take(b, len(batches))
    assert batches == [[0, 3, 6], [1, 5, 9], [2], [4, 11], [7, 8], [10], [12, 13]]
    assert all(r.t_admit is not None for r in reqs)
    assert s._next_batch() is None and not s._q1d and not s._q2d and not s._q3d
