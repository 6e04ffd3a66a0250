"""Host logic of the serving loop on stand-in work (no GPU): admission by
HBM footprint never exceeds the budget, requests are admitted in arrival
order without overtaking, the worker cap holds, and latencies are measured
from arrival."""
import threading
import time

import pytest

from specfem3d_geotech_amd.serving import Request, Server, footprint_bytes, make_requests


class NullCtx:
    def close(self):
        pass


class StandInServer(Server):
    def __init__(self, **kw):
        super().__init__(context_factory=NullCtx, **kw)
        self.log = []
        self.lock = threading.Lock()

    def run_request(self, req, ctx):
        with self.lock:
            self.log.append(("start", req.tag, self._reserved, self._running))
        time.sleep(0.02 + 0.0001 * (req.nelmt % 100))
        return req.tag


def _reqs(n, sizes, rate=200.0):
    out = []
    for i in range(n):
        out.append(Request("slope" if i % 3 else "excavation", lambda: None, {},
                           sizes[i % len(sizes)], arrival=i / rate, tag=i))
    return out


def test_footprint_grows_with_mesh_and_columns():
    assert footprint_bytes("slope", 2000) < footprint_bytes("slope", 4000)
    assert footprint_bytes("slope", 2000, 8) < footprint_bytes("slope", 2000, 16)
    assert footprint_bytes("excavation", 2000) < footprint_bytes("slope", 2000, 8)


@pytest.mark.parametrize("workers", [1, 3, 8])
def test_admission_respects_budget_and_order(workers):
    sizes = [1000, 5000, 2000, 8000]
    budget = 2 * footprint_bytes("slope", 8000)
    srv = StandInServer(hbm_budget=budget, max_workers=workers)
    reqs = _reqs(24, sizes)
    st = srv.serve(reqs)
    assert st["requests"] == 24
    assert srv.peak_reserved <= budget
    assert srv.peak_running <= workers
    starts = [e[1] for e in srv.log]
    assert sorted(starts) == list(range(24))
    admits = sorted(reqs, key=lambda r: r.t_admit)
    assert [r.tag for r in admits] == list(range(24))  # no overtaking
    for r in reqs:
        assert r.result == r.tag
        assert r.t_admit >= r.t_arrival - 1e-6
        assert r.latency >= r.t_done - r.t_admit - 1e-9
    assert st["p50_s"] <= st["p99_s"] <= st["max_s"]


def test_request_over_budget_is_refused():
    srv = StandInServer(hbm_budget=1000, max_workers=2)
    with pytest.raises(ValueError):
        srv.serve(_reqs(1, [5000]))


def test_request_stream_mix_and_sizes():
    reqs = make_requests(200, rate=4.0, nelmt_range=(5000, 200000), seed=4)
    kinds = [r.kind for r in reqs]
    share = kinds.count("slope") / len(kinds)
    assert 0.6 < share < 0.8
    ne = [r.nelmt for r in reqs]
    assert min(ne) >= 2500 and max(ne) <= 400000
    gaps = [b.arrival - a.arrival for a, b in zip(reqs, reqs[1:])]
    assert 0.15 < sum(gaps) / len(gaps) < 0.35   # mean 1/rate
    m = reqs[0].make_mesh()
    assert abs(m.nelmt - reqs[0].nelmt) == 0
