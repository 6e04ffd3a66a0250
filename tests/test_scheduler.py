"""Host logic of the multi-trial scheduler (run_analyses) on a stand-in engine.

The stand-in gives every trial a known plastic-iteration count from its
reduced cohesion (a trial fails, i.e. reaches nl_maxiter, when c/SRF < 5),
so the FOS bracket of each analysis is known in closed form and the
scheduler's bookkeeping -- first-failure rule, speculative SRFs, column
reuse, elastic-solve memo, arrivals -- is checked without a GPU.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from specfem3d_geotech_amd.batch import Analysis, run_analyses

NL_MAX = 12


class StandIn:
    """Engine stand-in: every running column does one plastic iteration per
    round; a round ends the call when a trial ends or the elastic memo
    appears (as sgeo_batch_advance does)."""

    def __init__(self, ncol):
        self.ncol = ncol
        self.dmesh = SimpleNamespace(spec=SimpleNamespace(gdof=np.array([[1, 2, 3]])))
        self.col = [None] * ncol
        self.begins = []
        self.active_per_call = []
        self.memo = False
        self.steps = 0

    def begin(self, col, ym, nuf, phif, cohf, psif, dt, x0=None, x0_cg=0):
        assert self.col[col] is None, "column reused before release"
        c = float(cohf[0])
        need = NL_MAX if c < 5.0 else 1 + int(round(20.0 / c))
        self.col[col] = dict(c=c, need=need, it=0, cg=0, seeded=self.memo, ended=False)
        self.begins.append((col, c, self.memo))

    def release(self, col):
        self.col[col] = None

    def nl_iters(self):
        return np.array([0 if s is None or s["ended"] else s["it"] for s in self.col],
                        np.int32)

    def advance(self, cg_tol, cg_maxiter, inner_tol, nl_tol, nl_maxiter,
                max_steps=1 << 40):
        n = self.ncol
        out = dict(ended=np.zeros(n, np.int32), nl_tot=np.zeros(n, np.int32),
                   cg_tot=np.zeros(n, np.int32), fmax=np.zeros(n),
                   errcode=np.zeros(n, np.int32), memo_ready=self.memo)
        self.active_per_call.append(sum(1 for s in self.col
                                        if s is not None and not s["ended"]))
        memo0 = self.memo
        while True:
            any_end = False
            for j, s in enumerate(self.col):
                if s is None or s["ended"]:
                    continue
                s["it"] += 1
                s["cg"] += 11 if (s["it"] == 1) else 10 + s["it"]
                if s["it"] == 1 and not s["seeded"]:
                    self.memo = True
                if (s["it"] >= s["need"] and s["need"] < NL_MAX) or s["it"] >= nl_maxiter:
                    s["ended"] = True
                    any_end = True
                    out["ended"][j] = 1
                    out["nl_tot"][j] = s["it"]
                    out["cg_tot"][j] = s["cg"]
                    out["fmax"][j] = s["c"]
            self.steps += 1
            if any_end or (self.memo and not memo0) or self.steps % max(1, max_steps) == 0:
                break
        out["memo_ready"] = self.memo
        return out

    def download(self, col):
        s = self.col[col]
        x = np.array([0.0, s["c"], -2 * s["c"], 0.5])
        return x, np.zeros((1, 27), np.uint8)

    def stats(self):
        return dict(steps=self.steps)


def expected_bracket(coh, srf):
    prev = None
    for s in srf:
        c = coh / s if s != 1.0 else coh
        if c < 5.0:
            return (prev, s)
        prev = s
    return None


def _run(ncol, cohs, srf, **kw):
    eng = StandIn(ncol)
    an = [Analysis(coh=np.array([c]), phi=np.array([20.0]), psi=np.array([0.0]),
                   srf=list(srf), tag=i) for i, c in enumerate(cohs)]
    run_analyses(eng, an, ym=np.array([1e5]), nu=np.array([0.3]), cg_tol=1e-8,
                 cg_maxiter=100, nl_tol=5e-4, nl_maxiter=NL_MAX, **kw)
    return eng, an


SRF = [1.0, 1.25, 1.5, 1.75, 2.0, 2.25]
COHS = [6.0, 8.0, 10.0, 12.0, 20.0, 3.0, 9.5]


@pytest.mark.parametrize("ncol,srf_parallel", [(1, 1), (4, 1), (4, 3), (8, 6)])
def test_brackets_and_rows(ncol, srf_parallel):
    eng, an = _run(ncol, COHS, SRF, srf_parallel=srf_parallel)
    _, ref = _run(1, COHS, SRF, srf_parallel=1)
    for a, r in zip(an, ref):
        assert a.done
        assert a.bracket == expected_bracket(float(a.coh[0]), SRF)
        # rows are the sequential sweep's: every SRF up to the first failure
        n = len(a.rows)
        assert sorted(a.rows) == list(range(n))
        assert a.rows == r.rows
        for k, row in a.rows.items():
            c = a.coh[0] / SRF[k] if SRF[k] != 1.0 else a.coh[0]
            assert row["fmax"] == pytest.approx(c)
            assert row["uxmax"] == pytest.approx(2 * c)
    assert all(s is None for s in eng.col)


def test_elastic_memo_used_after_first_trial():
    eng, an = _run(4, COHS, SRF)
    # the first trial solves the elastic system alone; later ones reuse it
    assert eng.begins[0][2] is False
    assert all(m for (_, _, m) in eng.begins[1:])
    assert eng.active_per_call[0] == 1 and max(eng.active_per_call) == 4
    r0 = an[0].rows[0]
    assert r0["cg_tot"] == 11 + sum(10 + i for i in range(2, r0["nl_tot"] + 1))


def test_arrivals_delay_admission():
    t = [0.0]

    def clock():
        t[0] += 0.01
        return t[0]

    eng, an = _run(2, COHS[:3], SRF, arrivals=[0.0, 0.0, 0.5], clock=clock)
    assert an[2].t_start >= an[2].t_arrival
    assert all(a.done for a in an)


@pytest.mark.parametrize("ncol,srf_parallel,defer", [(4, 3, 2), (8, 6, 3), (8, 6, 1)])
def test_failing_trial_holds_speculation(ncol, srf_parallel, defer):
    """Trials above a trial heading for nl_maxiter are released and rerun
    only if it converges: rows and brackets equal the sequential sweep's,
    with fewer trials started above the first failure."""
    eng, an = _run(ncol, COHS, SRF, srf_parallel=srf_parallel, defer_nl=defer,
                   poll_steps=1)
    eng0, ref = _run(ncol, COHS, SRF, srf_parallel=srf_parallel)
    for a, r in zip(an, ref):
        assert a.done and a.bracket == r.bracket == expected_bracket(float(a.coh[0]), SRF)
        assert a.rows == r.rows
    assert all(s is None for s in eng.col)
