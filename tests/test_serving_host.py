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
