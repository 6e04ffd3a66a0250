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


def test_served_results_match_reference():
    """A served slope sweep (the mixed multi-trial engine) and a served
    staged excavation (the packed FP64 operator), in one mix with other
    requests, against the compiled reference's own results on the same
    meshes: the slope's FOS bracket and the plastic iteration count of every
    converged SRF; every excavation stage's plastic iteration count, CG
    iteration count within 1% and displacement maximum within 1e-8."""
    from conftest import control_from_golden, golden, mesh_from_golden
    from specfem3d_geotech_amd.serving import Request
    zs = golden("slope_8x4x6_layered")
    zp = golden("pit_8x6x6")
    cp = zp["control"]
    nexcav, ninc, *ids = [int(v) for v in zp["excav"]]
    pit_ctl = dict(cg_tol=float(cp[0]), cg_maxiter=int(cp[1]), nl_tol=float(cp[2]),
                   nl_maxiter=int(cp[3]), srf=[float(s) for s in zp["srf"]],
                   nexcav=nexcav, nexcavid=[1] * nexcav, excavid=ids, ninc=ninc)
    slope = Request("slope", lambda: mesh_from_golden(zs), control_from_golden(zs),
                    int(zs["conn"].shape[0]), 0.0, tag="slope_8x4x6")
    pit = Request("excavation", lambda: mesh_from_golden(zp), pit_ctl,
                  int(zp["conn"].shape[0]), 0.05, tag="pit_8x6x6")
    others = make_requests(3, rate=20.0, nelmt_range=(300, 900), seed=5, nl_maxiter=40)
    for r in others:
        r.arrival += 0.02
    st = Server(max_workers=5, slope_columns=4).serve([slope, pit] + others)
    print("mix:", st)
    ref = zs["summary"]
    nl_max = int(zs["control"][3])
    assert slope.result.fos_bracket == (float(ref[-2, 0]), float(ref[-1, 0]))
    for r, rr in zip(slope.result.rows, ref):
        print("served slope", r, rr.tolist())
        if rr[2] < nl_max:
            assert r["nl_tot"] == rr[2]
    for r, rr in zip(pit.result.rows, zp["summary"]):
        print("served pit", r, rr.tolist())
        assert r["nl_tot"] == rr[2]
        assert abs(r["cg_tot"] - rr[1]) <= 0.01 * rr[1]
        assert abs(r["uxmax"] - rr[3]) <= 1e-8 * rr[3]
    assert len(pit.result.rows) == zp["summary"].shape[0]
