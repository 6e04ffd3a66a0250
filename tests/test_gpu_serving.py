"""Serving loop on the device: a stream of slope and excavation requests on
meshes of different sizes, admitted by HBM footprint and running
concurrently on their own streams.  Each request's result is the same, bit
for bit, when it runs alone and when it runs in the mix."""
import copy

import numpy as np
import pytest

from specfem3d_geotech_amd.serving import Server, make_requests

pytestmark = pytest.mark.gpu


def _summary(req):
    r = req.result
    if req.kind == "slope":
        return (r.fos_bracket, [tuple(sorted(x.items())) for x in r.rows], r.x.tobytes())
    return ([tuple(sorted(x.items())) for x in r.rows], r.nodalu.tobytes())


def test_results_identical_alone_and_in_mix():
    reqs = make_requests(6, rate=20.0, nelmt_range=(300, 900), seed=11, nl_maxiter=40)
    kinds = [r.kind for r in reqs]
    assert "slope" in kinds and "excavation" in kinds
    alone_reqs = [copy.copy(reqs[kinds.index("slope")]),
                  copy.copy(reqs[kinds.index("excavation")])]
    mix = Server(max_workers=6, slope_columns=4)
    st = mix.serve(reqs)
    print("mix:", st)
    assert st["peak_concurrent"] >= 2
    for a in alone_reqs:
        a.arrival = 0.0
        a.result = None
        srv = Server(max_workers=1, slope_columns=4)
        srv.serve([a])
        m = [r for r in reqs if r.tag == a.tag][0]
        print(a.kind, a.tag, "alone", a.result.rows[:2])
        assert _summary(a) == _summary(m)
