"""Host logic of the serving loop (suntans_amd.serve) without a GPU:
admission by HBM budget and eviction, with stand-ins for the device grids
and the batch.

* Only simulations that can be rebuilt (kept host state) are evicted: a
  resident simulation without host state keeps its place, and a request
  that cannot fit beside it is refused with an error instead of waiting
  forever.
* Among rebuildable idle simulations the least recently used leaves first,
  and its q, s and h come back on readmission.
"""
import numpy as np
import pytest

from suntans_amd import serve


class FakeGrid:
    def __init__(self, nbytes, fields=None):
        self.nbytes = nbytes
        self.fields = dict(fields or {})
        self.closed = False

    def bytes(self):
        return self.nbytes

    def sizes(self):
        return {"Nflat": 4}

    def download(self, name, n):
        return self.fields.get(name, np.zeros(n)).copy()

    def upload(self, name, arr):
        self.fields[name] = np.asarray(arr, float).copy()

    def close(self):
        self.closed = True


class FakeBatch:
    def advance(self, iters):
        pass

    def poll(self):
        return []

    def close(self):
        pass


def server(sims, budget):
    srv = serve.Server.__new__(serve.Server)
    srv.mixed, srv.inner_tol, srv.eps, srv.tvd = False, 1e-4, 1e-10, 3
    srv.sims = {s.sid: s for s in sims}
    srv.budget, srv.max_slots = budget, 8
    srv.batch = FakeBatch()
    srv.evictions = srv.admissions = 0
    return srv


def sim(sid, nbytes, rebuildable, resident, last_used=0.0):
    s = serve.Sim(sid=sid, N=4, Nc=2, phys={}, est_bytes=nbytes, last_used=last_used)
    if rebuildable:
        s.make_grid = lambda nb=nbytes: FakeGrid(nb)
    if resident:
        s.dg = FakeGrid(nbytes, {"q": np.full(4, float(sid)), "s": np.full(4, 2.0 * sid), "h": np.full(2, 3.0)})
    return s


def test_not_rebuildable_is_never_evicted():
    a = sim(0, 60, rebuildable=False, resident=True)
    b = sim(1, 60, rebuildable=True, resident=False)
    srv = server([a, b], budget=100)
    assert srv._make_resident(b) is False
    assert a.dg is not None and not a.dg.closed and srv.evictions == 0


def test_stuck_queue_raises():
    a = sim(0, 60, rebuildable=False, resident=True)
    b = sim(1, 60, rebuildable=True, resident=False)
    srv = server([a, b], budget=100)
    with pytest.raises(RuntimeError, match="cannot be placed"):
        srv.run([(0.0, 1)])


def test_nonresident_without_host_state_raises():
    a = sim(0, 10, rebuildable=False, resident=False)
    srv = server([a], budget=100)
    with pytest.raises(RuntimeError, match="no host state"):
        srv._make_resident(a)


def test_lru_rebuildable_evicted_and_restored():
    a = sim(0, 40, rebuildable=True, resident=True, last_used=2.0)
    b = sim(1, 40, rebuildable=True, resident=True, last_used=1.0)
    c = sim(2, 40, rebuildable=False, resident=True, last_used=0.0)
    d = sim(3, 40, rebuildable=True, resident=False)
    srv = server([a, b, c, d], budget=130)
    assert srv._make_resident(d)
    # c is the least recently used but cannot be rebuilt; b goes
    assert b.dg is None and a.dg is not None and c.dg is not None and srv.evictions == 1
    assert np.array_equal(b.saved["q"], np.full(4, 1.0))
    d.dg = None  # make room again for b's return
    assert srv._make_resident(b)
    assert np.array_equal(b.dg.fields["q"], np.full(4, 1.0)) and np.array_equal(b.dg.fields["s"], np.full(4, 2.0))
    assert b.saved == {}
