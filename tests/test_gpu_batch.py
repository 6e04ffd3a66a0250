"""Mixed-precision multi-trial engine (sgeo_batch, through the C ABI).

* A trial's bits do not depend on the batch: the same analysis run alone
  (one column) and among other analyses that start at other times gives
  identical rows, displacements and plastic-point flags.
* Against the FP64 reference (golden summaries from the compiled reference):
  a PCG solve within 1e-8 of the reference's; over a sweep the same FOS
  bracket, the same plastic iteration counts on converged SRFs, and
  displacements that differ from the reference's by less than the
  reference's own CG truncation error.
* One solve of K x = f against the FP64 oracle's PCG run to a tight
  tolerance: the FP32 inner CG with FP64 residual replacement, stopped by
  the reference's own test at cg_tol, is as accurate as the FP64 PCG
  stopped by that test (the FP64 reference's own error at cg_tol = 1e-8 is
  2.2e-8 relative L2 on this mesh).
"""
import numpy as np
import pytest

import oracle as O
from conftest import control_from_golden, golden, mesh_from_golden
from specfem3d_geotech_amd.batch import (DEFAULT_INNER_TOL, Analysis, BatchEngine,
                                         device_mesh, run_analyses, slope3d_mixed)
from specfem3d_geotech_amd.spec import hex2spec

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _analyses(mesh, n, srf, seed):
    """`n` strength realisations around the mesh's own (the first is the
    mesh's own strengths)."""
    _, _, _, phi, coh, psi = mesh.material_arrays()
    rng = np.random.default_rng(seed)
    out = [Analysis(coh=coh.copy(), phi=phi.copy(), psi=psi.copy(), srf=list(srf),
                    tag=0)]
    for i in range(1, n):
        out.append(Analysis(coh=coh * rng.uniform(0.8, 1.25, coh.shape),
                            phi=phi * rng.uniform(0.85, 1.15, phi.shape),
                            psi=psi.copy(), srf=list(srf), tag=i))
    return out


@pytest.mark.parametrize("period", [1, 4])
def test_batch_trial_identical_alone_and_in_mix(ctx, period):
    """A trial gives the same bits alone (1 column) and behind other
    analyses in an 8-column engine, also when FP64 residuals are formed on
    every 4th step only (the replacement iterate then depends on the period,
    not on the other columns)."""
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    ctl["nl_maxiter"] = 40
    srf = [1.0, 1.3, 1.6, 1.9]
    _, ym, nu, *_ = mesh.material_arrays()
    kw = dict(ym=ym, nu=nu, cg_tol=ctl["cg_tol"], cg_maxiter=ctl["cg_maxiter"],
              nl_tol=ctl["nl_tol"], nl_maxiter=ctl["nl_maxiter"], keep_state=True)
    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        alone = _analyses(mesh, 6, srf, seed=7)
        target = alone[3]
        eng = BatchEngine(dm, ext, 1)
        try:
            eng.set_residual_period(period)
            run_analyses(eng, [target], **kw)
        finally:
            eng.close()
        mix = _analyses(mesh, 6, srf, seed=7)
        # the target enters third, behind analyses already in flight
        order = [0, 1, 3, 2, 4, 5]
        eng = BatchEngine(dm, ext, 8)
        try:
            eng.set_residual_period(period)
            run_analyses(eng, [mix[i] for i in order], srf_parallel=2, **kw)
        finally:
            eng.close()
        got = mix[3]
        print("alone", target.rows, target.bracket)
        print("mix  ", got.rows, got.bracket)
        assert got.bracket == target.bracket
        assert got.rows == target.rows
        assert np.array_equal(got.x_last, target.x_last)
        assert np.array_equal(got.flags_last, target.flags_last)
    finally:
        dm.close()


def test_residual_period_arguments(ctx):
    """The residual period is 1, 2, 4 or 8, and neither it nor the x
    deferral can change while trials run (their residual steps and x writes
    fall at fixed counts of their own iterations); setting the value in
    force again is accepted."""
    from specfem3d_geotech_amd._lib import SgeoError
    from specfem3d_geotech_amd.plastic import dt_viscoplas
    z = golden("slope_4x2x3")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    _, ym, nu, phi, coh, psi = mesh.material_arrays()
    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        eng = BatchEngine(dm, ext, 2, resid_period=8)
        try:
            for bad in (0, 3, 16):
                with pytest.raises(SgeoError):
                    eng.set_residual_period(bad)
            eng.begin(0, ym, nu, phi, coh, psi, dt_viscoplas(nu, phi, ym))
            eng.advance(ctl["cg_tol"], ctl["cg_maxiter"], DEFAULT_INNER_TOL, ctl["nl_tol"],
                        ctl["nl_maxiter"], max_steps=16)
            with pytest.raises(SgeoError):
                eng.set_residual_period(4)
            eng.set_residual_period(8)
            with pytest.raises(SgeoError):
                eng.set_deferred_x(False)
            eng.set_deferred_x(True)
        finally:
            eng.close()
    finally:
        dm.close()


@pytest.mark.parametrize("period", [2, 8])
def test_deferred_x_update_keeps_the_bits(ctx, period):
    """With FP64 residuals every `period`-th step, x is written on every
    period-th iteration only, which applies the updates of the iterations
    before it in their order; solves that end between two writes are
    completed before the body pass.  The rows and x of a 4-column sweep are
    the same bits with the deferral on and off (the stop test on a deferred
    iteration could only differ inside a window of relative width about
    period cg_tol^2 around its threshold)."""
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    _, ym, nu, *_ = mesh.material_arrays()
    kw = dict(ym=ym, nu=nu, cg_tol=ctl["cg_tol"], cg_maxiter=ctl["cg_maxiter"],
              nl_tol=ctl["nl_tol"], nl_maxiter=40, keep_state=True, srf_parallel=2)
    out = []
    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        for defer in (True, False):
            an = _analyses(mesh, 4, [1.0, 1.3, 1.6], seed=5)
            eng = BatchEngine(dm, ext, 4, resid_period=period)
            try:
                eng.set_deferred_x(defer)
                run_analyses(eng, an, **kw)
                st = eng.stats()
            finally:
                eng.close()
            print("deferred" if defer else "every iteration", st,
                  [a.bracket for a in an], flush=True)
            out.append([(a.rows, a.bracket, a.x_last.tobytes(), a.flags_last.tobytes())
                        for a in an])
    finally:
        dm.close()
    assert out[0] == out[1]


@pytest.mark.parametrize("name", ["slope_4x2x3", "slope_8x4x6_layered"])
def test_mixed_solve_within_1e8_of_reference_solve(ctx, name):
    """One PCG solve, the reference's own first call (K x = extload from
    x = 0, recorded by the pcg_solver tap): the FP32 inner CG with FP64
    residual replacement, stopped by the reference's test at cg_tol, gives
    the reference's FP64 solution to 1e-8 relative L2 with the engine's
    defaults (FP64 residuals every 8th step, due once r.z has fallen by
    1e-4), within 5% of its iteration count."""
    from specfem3d_geotech_amd.plastic import dt_viscoplas
    z = golden(name)
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    _, ym, nu, phi, coh, psi = mesh.material_arrays()
    assert int(z["call_idx"][0]) == 0 and not z["call_uin"][0].any()
    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        eng = BatchEngine(dm, ext, 1)
        try:
            eng.begin(0, ym, nu, phi, coh, psi, dt_viscoplas(nu, phi, ym))
            r = eng.plastic_iter(np.array([0], np.int32), ctl["cg_tol"], ctl["cg_maxiter"],
                                 DEFAULT_INNER_TOL, ctl["nl_tol"])
            x, _ = eng.download(0)
        finally:
            eng.close()
    finally:
        dm.close()
    e = rel_l2(x, z["call_uout"][0])
    print(f"{name}: mixed solve vs the reference's {e:.3e}; cg_iter "
          f"{int(r['cg_iter'][0])} vs {int(z['cg_log'][0])}")
    assert e <= 1e-8
    # FP32 inner arithmetic: the reference's stop test fires within a few
    # iterations of the reference's own count
    assert abs(int(r["cg_iter"][0]) - int(z["cg_log"][0])) <= 0.05 * int(z["cg_log"][0])


@pytest.mark.parametrize("name,ncol,period", [("slope_4x2x3", 1, 1),
                                              ("slope_8x4x6_layered", 1, 1),
                                              ("slope_8x4x6_layered", 4, 1),
                                              ("slope_8x4x6_layered", 4, 8)])
def test_mixed_sweep_matches_reference(ctx, name, ncol, period):
    """Same FOS bracket as the FP64 reference, the same plastic iteration
    count at every converged SRF, and at every converged SRF a displacement
    that differs from the reference's by less than the reference's own
    distance from the exactly solved problem (its PCG stops at a relative
    update of cg_tol = 1e-8; the viscoplastic loop carries each solve's
    truncation into x): ||x_mixed - x_ref|| <= ||x_ref - x_tight||, x_ref and
    x_tight the oracle's sweep of that SRF at cg_tol and at 1e-13."""
    z = golden(name)
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    xs = {}
    res = slope3d_mixed(ctx, mesh, ncol=ncol, resid_period=period,
                        on_srf=lambda row, x, fl: xs.__setitem__(row["srf"], x.copy()),
                        **ctl)
    ref = z["summary"]
    print(name, "stats", res.extra["stats"])
    assert len(res.rows) == ref.shape[0]
    for r, rr in zip(res.rows, ref):
        print(name, r, rr.tolist())
        assert r["srf"] == rr[0]
        if rr[2] < ctl["nl_maxiter"]:
            assert r["nl_tot"] == rr[2]
            one = dict(ctl, srf=[float(rr[0])])
            x_ref = O.slope3d(spec, mesh, **one)["x"]
            x_tight = O.slope3d(spec, mesh, **dict(one, cg_tol=1e-13))["x"]
            e, e_ref = rel_l2(xs[r["srf"]], x_ref), rel_l2(x_ref, x_tight)
            print(f"  x vs reference {e:.3e}; reference vs exact solves {e_ref:.3e}")
            assert e <= max(1e-8, e_ref)
        else:
            assert r["nl_tot"] == ctl["nl_maxiter"]
    bracket = (float(ref[-2, 0]) if ref.shape[0] > 1 else None, float(ref[-1, 0])) \
        if ref[-1, 2] == ctl["nl_maxiter"] else None
    assert res.fos_bracket == bracket


def test_refined_solve_matches_fp64(ctx):
    """First plastic iteration = the elastic solve K x = extload from x = 0."""
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    km, dp, ext_o = O.stiffness_bodyload(spec, gam, nu, ym)
    f = z["extload"]
    x_tight, it_t, ec_t = O.pcg_solver(spec.neq, km, np.zeros(spec.neq + 1), f,
                                       z["dinv"], spec.gdof_elmt, 1e-14, 20000)
    x_ref, _, _ = O.pcg_solver(spec.neq, km, np.zeros(spec.neq + 1), f, z["dinv"],
                               spec.gdof_elmt, 1e-8, 5000)
    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        eng = BatchEngine(dm, ext, 1)
        try:
            from specfem3d_geotech_amd.plastic import dt_viscoplas
            eng.begin(0, ym, nu, phi, coh, psi, dt_viscoplas(nu, phi, ym))
            r = eng.plastic_iter(np.array([0], np.int32), 1e-8, 5000, DEFAULT_INNER_TOL,
                                 5e-4)
            x, _ = eng.download(0)
        finally:
            eng.close()
    finally:
        dm.close()
    e_mixed = rel_l2(x, x_tight)
    e_fp64 = rel_l2(x_ref, x_tight)
    print(f"mixed {e_mixed:.3e} (inner its {int(r['cg_iter'][0])}), "
          f"fp64 cg_tol 1e-8: {e_fp64:.3e}; tight solve its {it_t}")
    assert ec_t == 0
    assert e_mixed <= 2.0 * max(1e-8, e_fp64)


def test_batched_realisations_match_oracle_brackets(ctx):
    """Form-A strength realisations through the continuous-batching scheduler
    (more analyses than columns, so columns are reused) against the oracle's
    FP64 sweep of each realisation: identical FOS brackets; plastic iteration
    counts of converged SRFs equal or one apart (uerr <= nl_tol decided on
    solutions that agree to the PCG tolerance)."""
    from specfem3d_geotech_amd import workloads
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    ctl["nl_maxiter"] = 60
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    real = workloads.form_a(6, len(gam))
    an = [Analysis(coh=c, phi=p, psi=psi.copy(), srf=list(ctl["srf"]), tag=i)
          for i, (c, p) in enumerate(real)]
    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        eng = BatchEngine(dm, ext, 4)
        try:
            run_analyses(eng, an, ym=ym, nu=nu, cg_tol=ctl["cg_tol"],
                         cg_maxiter=ctl["cg_maxiter"], nl_tol=ctl["nl_tol"],
                         nl_maxiter=ctl["nl_maxiter"])
        finally:
            eng.close()
    finally:
        dm.close()
    for a, (c, p) in zip(an, real):
        o = O.slope3d(spec, mesh.with_strengths(c, p), **ctl)
        rows = o["rows"]
        br = None
        if rows[-1]["nl_tot"] == ctl["nl_maxiter"]:
            br = (rows[-2]["srf"] if len(rows) > 1 else None, rows[-1]["srf"])
        print(f"c={c[0]:.3f} phi={p[0]:.3f}: batched {a.bracket}, oracle {br}")
        assert a.bracket == br
        for k, r in enumerate(rows):
            if r["nl_tot"] < ctl["nl_maxiter"]:
                assert abs(a.rows[k]["nl_tot"] - r["nl_tot"]) <= 1


def test_graph_steps_match_eager_launches(ctx):
    """CG steps replayed from captured HIP graphs give the same bits as the
    same steps launched one by one (several replays per graph, with host
    synchronisations between them)."""
    z = golden("slope_4x2x3")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    _, ym, nu, phi, coh, psi = mesh.material_arrays()
    out = []
    for graphs in (True, False):
        dm, ext = device_mesh(ctx, mesh, spec)
        try:
            eng = BatchEngine(dm, ext, 2)
            eng.set_graphs(graphs)
            a = Analysis(coh=coh, phi=phi, psi=psi, srf=list(ctl["srf"]))
            try:
                run_analyses(eng, [a], ym=ym, nu=nu, cg_tol=ctl["cg_tol"],
                             cg_maxiter=ctl["cg_maxiter"], nl_tol=ctl["nl_tol"],
                             nl_maxiter=ctl["nl_maxiter"], srf_parallel=2,
                             keep_state=True)
                steps = eng.stats()["batched_steps"]
            finally:
                eng.close()
        finally:
            dm.close()
        out.append((a.rows, a.bracket, a.x_last.tobytes(), steps))
        print("graphs" if graphs else "eager", a.rows, steps)
    assert out[0][:3] == out[1][:3]
    assert out[0][3] > 100  # many graph replays


def test_launch_geometry_does_not_change_bits(ctx):
    """The MFMA product's launch geometry (blocks per CU, the XCD-aware
    split of its element walk) changes which block computes an element,
    never how: a 64-column engine gives the same rows and displacement bits
    under each."""
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    _, ym, nu, *_ = mesh.material_arrays()
    kw = dict(ym=ym, nu=nu, cg_tol=ctl["cg_tol"], cg_maxiter=ctl["cg_maxiter"],
              nl_tol=ctl["nl_tol"], nl_maxiter=40, keep_state=True, srf_parallel=2)
    out = []
    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        for geom in [(3, True), (1, False), (2, True), (4, False)]:
            an = _analyses(mesh, 20, [1.0, 1.5], seed=11)
            eng = BatchEngine(dm, ext, 64)
            try:
                eng.set_geometry(*geom)
                run_analyses(eng, an, **kw)
            finally:
                eng.close()
            print(geom, [a.bracket for a in an[:4]], flush=True)
            out.append([(a.rows, a.bracket, a.x_last.tobytes()) for a in an])
    finally:
        dm.close()
    for o in out[1:]:
        assert o == out[0]


def test_column_released_mid_solve_restarts_clean(ctx):
    """A column released in the middle of a PCG call (an abandoned
    speculative SRF) stops on the device at once; a trial begun on it
    afterwards gives the same bits as on a fresh engine, and the engine's
    iteration count grows only by the surviving column's iterations."""
    from specfem3d_geotech_amd.plastic import dt_viscoplas, strength_reduction
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    _, ym, nu, phi, coh, psi = mesh.material_arrays()
    kw = dict(cg_tol=ctl["cg_tol"], cg_maxiter=ctl["cg_maxiter"],
              inner_tol=DEFAULT_INNER_TOL, nl_tol=ctl["nl_tol"], nl_maxiter=20)

    def begin(eng, col, s):
        cohf, nuf, phif, psif, _ = strength_reduction(s, False, coh, nu, phi, psi)
        eng.begin(col, ym, nuf, phif, cohf, psif, dt_viscoplas(nuf, phif, ym))

    def finish(eng, col):
        while True:
            r = eng.advance(**kw)
            if r["ended"][col]:
                return eng.download(col), int(r["nl_tot"][col]), int(r["cg_tot"][col])

    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        fresh = BatchEngine(dm, ext, 2)
        try:
            begin(fresh, 0, 1.3)
            want = finish(fresh, 0)
        finally:
            fresh.close()
        eng = BatchEngine(dm, ext, 2)
        try:
            begin(eng, 0, 1.0)
            begin(eng, 1, 1.6)
            eng.advance(max_steps=24, **kw)  # both columns inside their PCG calls
            eng.release(1)
            s0 = eng.stats()
            eng.advance(max_steps=16, **kw)
            s1 = eng.stats()
            # only column 0 iterates now: at most one CG iteration per step
            assert s1["cg_iterations"] - s0["cg_iterations"] <= \
                s1["batched_steps"] - s0["batched_steps"]
            begin(eng, 1, 1.3)
            got = finish(eng, 1)
        finally:
            eng.close()
    finally:
        dm.close()
    assert got[1:] == want[1:]
    assert np.array_equal(got[0][0], want[0][0]) and np.array_equal(got[0][1], want[0][1])
