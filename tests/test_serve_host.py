"""Host logic of the serving loop's resident-context pool (no GPU): keys,
least-recently-used eviction and the HBM budget accounting."""
import numpy as np

from powerfit_amd.serve import FitRequest, FitServer


class _Ctx:
    pass


def _req(key, shape=(8, 8, 8), laplace=False):
    z = np.zeros(shape)
    return FitRequest(0, z, z, z, np.eye(3)[None], laplace, 0.0, key)


def test_pool_key_separates_shape_and_laplace():
    srv = FitServer()
    assert srv._key(_req(None)) is None
    a, b, c = _req("m"), _req("m", (8, 8, 10)), _req("m", laplace=True)
    assert len({srv._key(a), srv._key(b), srv._key(c)}) == 3
    assert FitServer(keep_targets=False)._key(a) is None


def test_idle_contexts_evicted_least_recently_used_first():
    srv = FitServer(hbm_budget=100)
    ctx = {n: _Ctx() for n in "abcd"}
    srv.idle[("m1",)] = [(ctx["a"], 30), (ctx["b"], 20)]
    srv.idle[("m2",)] = [(ctx["c"], 25)]
    srv.idle[("m3",)] = [(ctx["d"], 15)]
    srv.used = 90
    srv._evict_idle(35)               # needs 25 freed: a (30) goes, oldest first
    assert srv.used == 60 and [c for c, _ in srv.idle[("m1",)]] == [ctx["b"]]
    srv._evict_idle(70)               # needs 30 more: b (20), then c (25)
    assert srv.used == 15 and list(srv.idle) == [("m3",)]
    srv.release_idle()
    assert srv.used == 0 and not srv.idle


def test_take_idle_returns_and_removes_one_context():
    srv = FitServer()
    c1, c2 = _Ctx(), _Ctx()
    key = srv._key(_req("m"))
    srv.idle[key] = [(c1, 10), (c2, 10)]
    assert srv._take_idle(key)[0] is c2 and srv._take_idle(key)[0] is c1
    assert key not in srv.idle and srv._take_idle(key) is None and srv._take_idle(None) is None
