"""Multi-trial FOS sweeps on one resident mesh (the batched form).

The viscoplastic method keeps the elastic stiffness fixed
(src/slope.F90:288), and every strength-reduction trial restarts from zero
state (src/slope.F90:248-249, 281-282), so trials of one mesh are
independent: here up to 64 of them share each CG iteration's pass over K
(sgeo_batch, mixed precision with FP64 refinement).

A *FOS analysis* is the reference's sweep for one set of strengths: SRFs in
list order until the first SRF whose plastic loop reaches nl_maxiter
(src/slope.F90:441); its bracket is (last stable SRF, first failing SRF).
The scheduler keeps every column busy: each analysis walks its SRF list
(optionally several SRFs of one analysis speculatively at once), and a
column freed by a finished analysis takes the next waiting analysis.  The
first PCG call of every trial solves K x = extload from x = 0 -- the same
system for every trial of the mesh -- so it is solved once and reused.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib as L
from ._lib import check, f64, i32, ptr
from .mesh import HexMesh
from .plastic import dt_viscoplas, strength_reduction
from .slope import SlopeResult, slope3d, uxmax_umax
from .solver import K_F64, Context, DeviceMesh, nodal_fields_call
from .spec import SpecMesh, hex2spec_device


# FP64 residuals on every 8th engine step: one pass over the packed FP64
# operator then serves the starts and replacements of 8 steps
DEFAULT_RESID_PERIOD = 8
# A column is due for an FP64 residual once r.z has fallen by this factor
# since its last one.  Replacing early keeps the FP32 recursion's drift
# small: with residuals every 8th step, the reference's first solve of the
# layered 8x4x6 slope is matched to 4.4e-9 with 1e-4 and 2.7e-9 with 1e-3,
# but to 1.1e-8 with 1e-5, at the same CG iteration count; on c1 with 64
# trials a step costs 4.69 ms with 1e-4 and 4.78 ms with 1e-3.
DEFAULT_INNER_TOL = 1e-4


class BatchEngine:
    """sgeo_batch: `ncol` trial columns on one DeviceMesh."""

    def __init__(self, dmesh: DeviceMesh, extload: np.ndarray, ncol: int,
                 resid_period: int = DEFAULT_RESID_PERIOD):
        self.dmesh = dmesh
        self.ncol = ncol
        self._ext = f64(extload)
        h = C.c_void_p()
        check(L.lib().sgeo_batch_create(dmesh.h, ncol, ptr(self._ext), C.byref(h)),
              "sgeo_batch_create")
        self.h = h
        self.set_residual_period(resid_period)

    def begin(self, col, ym, nuf, phif, cohf, psif, dt, x0=None, x0_cg=0):
        x0a = f64(x0) if x0 is not None else None
        check(L.lib().sgeo_batch_begin(self.h, col, ptr(f64(ym)), ptr(f64(nuf)),
                                       ptr(f64(phif)), ptr(f64(cohf)), ptr(f64(psif)),
                                       dt, ptr(x0a), int(x0_cg)),
              "sgeo_batch_begin")

    def begin_increment(self, col, ym, nuf, phif, cohf, psif, dt, stress0=None):
        """Excavation load increment on column `col`: state reset, the
        stress state set from stress0 (nelmt, 27, 6) or kept."""
        st = f64(stress0) if stress0 is not None else None
        check(L.lib().sgeo_batch_begin_increment(
            self.h, int(col), ptr(f64(ym)), ptr(f64(nuf)), ptr(f64(phif)),
            ptr(f64(cohf)), ptr(f64(psif)), float(dt), ptr(st)),
            "sgeo_batch_begin_increment")

    def download_stress(self, col) -> np.ndarray:
        st = np.zeros((self.dmesh.spec.nelmt, 27, 6))
        check(L.lib().sgeo_batch_download_stress(self.h, int(col), ptr(st)),
              "sgeo_batch_download_stress")
        return st

    def accumulate_nodal(self, col, first: bool):
        """Excavation: this increment's plastic strain and stress
        concentration factor gathered at the nodes of column `col`."""
        self.dmesh.ensure_nodes()
        check(L.lib().sgeo_batch_accumulate_nodal(self.h, int(col), int(first)),
              "sgeo_batch_accumulate_nodal")

    def release(self, col):
        check(L.lib().sgeo_batch_release(self.h, col), "sgeo_batch_release")

    def set_graphs(self, enable: bool):
        check(L.lib().sgeo_batch_set_graphs(self.h, int(enable)), "sgeo_batch_set_graphs")

    def autotune(self, reps: int = 5) -> Dict:
        """Launch geometry of the MFMA product, timed on this mesh with the
        columns iterating now (at least one must be); keeps the fastest."""
        bpc, xs = C.c_int32(), C.c_int32()
        best = C.c_double()
        table = np.zeros(8)
        check(L.lib().sgeo_batch_autotune(self.h, int(reps), C.byref(bpc), C.byref(xs),
                                          C.byref(best), ptr(table)), "sgeo_batch_autotune")
        cands = [dict(blocks_per_cu=b, xcd_split=x, ms=float(t))
                 for (x, b), t in zip([(x, b) for x in (1, 0) for b in (4, 3, 2, 1)], table)]
        return dict(blocks_per_cu=bpc.value, xcd_split=xs.value, ms=best.value,
                    candidates=cands)

    def plastic_iter(self, is_last, cg_tol, cg_maxiter, inner_tol, nl_tol):
        n = self.ncol
        il = i32(is_last)
        it = np.zeros(n, np.int32); ec = np.zeros(n, np.int32)
        ue = np.zeros(n); cv = np.zeros(n, np.int32); fm = np.zeros(n)
        check(L.lib().sgeo_batch_plastic_iter(
            self.h, ptr(il, C.c_int32), cg_tol, cg_maxiter, inner_tol, nl_tol,
            ptr(it, C.c_int32), ptr(ec, C.c_int32), ptr(ue), ptr(cv, C.c_int32),
            ptr(fm)), "sgeo_batch_plastic_iter")
        return dict(cg_iter=it, errcode=ec, uerr=ue, nl_isconv=cv, fmax=fm)

    def advance(self, cg_tol, cg_maxiter, inner_tol, nl_tol, nl_maxiter,
                max_steps=1 << 40):
        n = self.ncol
        ended = np.zeros(n, np.int32); nl = np.zeros(n, np.int32)
        cg = np.zeros(n, np.int32); fm = np.zeros(n); ec = np.zeros(n, np.int32)
        memo = C.c_int32()
        check(L.lib().sgeo_batch_advance(
            self.h, cg_tol, cg_maxiter, inner_tol, nl_tol, nl_maxiter, int(max_steps),
            ptr(ended, C.c_int32), ptr(nl, C.c_int32), ptr(cg, C.c_int32), ptr(fm),
            ptr(ec, C.c_int32), C.byref(memo)), "sgeo_batch_advance")
        return dict(ended=ended, nl_tot=nl, cg_tot=cg, fmax=fm, errcode=ec,
                    memo_ready=bool(memo.value))

    def nodal_fields(self, col: int) -> Dict[str, np.ndarray]:
        """The nodal result fields of the trial held by column `col`,
        computed on the device (sgeo_batch_nodal_fields)."""
        self.dmesh.ensure_nodes()
        return nodal_fields_call(L.lib().sgeo_batch_nodal_fields,
                                 "sgeo_batch_nodal_fields", self.dmesh.spec.nnode,
                                 self.h, int(col))

    def set_geometry(self, blocks_per_cu: int, xcd_split: bool):
        check(L.lib().sgeo_batch_set_geometry(self.h, int(blocks_per_cu), int(xcd_split)),
              "sgeo_batch_set_geometry")

    def set_residual_period(self, period: int):
        """FP64 residuals on every `period`-th CG step of the engine only
        (1, 2, 4 or 8).  A column due for a replacement iterates on until
        then, so results depend on the period, but never on the batch."""
        check(L.lib().sgeo_batch_set_residual_period(self.h, int(period)),
              "sgeo_batch_set_residual_period")

    def set_deferred_x(self, enable: bool):
        """x is written on every period-th iteration only (up to the 8th),
        which applies the deferred updates in order; x has the same bits
        either way.  Cannot change while trials run."""
        check(L.lib().sgeo_batch_set_deferred_x(self.h, int(bool(enable))),
              "sgeo_batch_set_deferred_x")

    def nl_iters(self) -> np.ndarray:
        out = np.zeros(self.ncol, np.int32)
        check(L.lib().sgeo_batch_nl_iters(self.h, ptr(out, C.c_int32)),
              "sgeo_batch_nl_iters")
        return out

    def download(self, col):
        sp = self.dmesh.spec
        x = np.zeros(sp.neq + 1)
        flags = np.zeros((sp.nelmt, 27), np.uint8)
        check(L.lib().sgeo_batch_download(self.h, col, ptr(x), ptr(flags, C.c_uint8)),
              "sgeo_batch_download")
        return x, flags

    def stats(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(L.lib().sgeo_batch_stats(self.h, C.byref(a), C.byref(b), C.byref(c)),
              "sgeo_batch_stats")
        return dict(cg_iterations=a.value, fp64_residuals=b.value, batched_steps=c.value)

    def nbytes(self):
        b = C.c_size_t()
        check(L.lib().sgeo_batch_bytes(self.h, C.byref(b)), "sgeo_batch_bytes")
        return b.value

    def close(self):
        if self.h:
            L.lib().sgeo_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Analysis:
    """One FOS analysis: a strength set and its SRF list."""
    coh: np.ndarray
    phi: np.ndarray
    psi: np.ndarray
    srf: Sequence[float]
    tag: object = None
    rows: Dict[int, dict] = field(default_factory=dict)   # srf index -> row
    bracket: Optional[tuple] = None
    done: bool = False
    next_srf: int = 0
    fail_idx: Optional[int] = None
    t_arrival: Optional[float] = None
    t_start: Optional[float] = None
    t_end: Optional[float] = None
    flags_last: Optional[np.ndarray] = None
    x_last: Optional[np.ndarray] = None
    hold: Optional[int] = None  # SRF index of a trial heading for failure

    @property
    def latency(self):
        return None if self.t_end is None else self.t_end - self.t_arrival


class PCGNotConverged(RuntimeError):
    """errcode -1 of a trial's PCG call; the reference stops (error_stop)."""


def run_analyses(engine, analyses: List[Analysis], *, ym, nu, cg_tol: float,
                 cg_maxiter: int, nl_tol: float, nl_maxiter: int,
                 inner_tol: float = DEFAULT_INNER_TOL, srf_parallel: int = 1,
                 keep_state: bool = False, on_trial=None,
                 arrivals: Optional[List[float]] = None,
                 clock=time.perf_counter, on_done=None, poll_steps: int = 64,
                 progress=None, progress_steps: int = 2000,
                 defer_nl: Optional[int] = None):
    """Runs every analysis to its FOS bracket, keeping the engine's columns
    busy (continuous batching).

    Each analysis walks its SRF list; srf_parallel > 1 lets it run that many
    consecutive SRFs at once (the bracket is still the first failing SRF in
    list order, src/slope.F90:441, and trials above a known failure are
    abandoned).  A column freed by an ended trial takes the next SRF of an
    admitted analysis, else the next waiting analysis.  `arrivals` (seconds
    from the start, one per analysis) admits a request stream: while
    requests are still to come the engine returns every `poll_steps` CG
    steps so they are admitted as soon as a column is free.  Until the
    engine holds the elastic solution (the first PCG call of any trial) only
    one trial runs.  `progress`, if given, is called with the engine's
    statistics at least every `progress_steps` CG steps.

    `defer_nl` limits speculation: once a trial of an analysis has run
    defer_nl plastic iterations without converging (a failing SRF runs to
    nl_maxiter), the analysis's trials at higher SRFs are released and none
    is started above it until it ends; if it converges after all, they run
    again from the start.  A trial's results do not depend on when it runs,
    so rows and brackets are unchanged.

    `on_trial(analysis, srf_index, x, flags, fields)` is called as each
    trial ends, while its column still holds the trial's state; `fields()`
    computes that state's nodal result fields.  Returns the engine's
    statistics."""
    spec = engine.dmesh.spec
    ncol = engine.ncol
    cols: List[Optional[tuple]] = [None] * ncol  # (analysis, srf index)
    queue = list(range(len(analyses)))
    admitted: List[Analysis] = []
    t0 = clock()
    for i, a in enumerate(analyses):
        a.t_arrival = t0 + (arrivals[i] if arrivals is not None else 0.0)
    memo_ready = False

    def running_of(a):
        return sum(1 for c in cols if c is not None and c[0] is a)

    def next_k(a):
        """Lowest SRF index of `a` neither done nor running, below a known
        failure and not above a held trial."""
        busy = {c[1] for c in cols if c is not None and c[0] is a}
        for k in range(len(a.srf)):
            if k in a.rows or k in busy:
                continue
            if a.fail_idx is not None and k >= a.fail_idx:
                return None
            if a.hold is not None and k > a.hold:
                return None
            return k
        return None

    def apply_holds():
        nl = engine.nl_iters()
        for a in admitted:
            if a.done:
                continue
            mine = [(j, c[1]) for j, c in enumerate(cols) if c is not None and c[0] is a]
            slow = [k for j, k in mine if nl[j] >= defer_nl]
            a.hold = min(slow) if slow else None
            if a.hold is None:
                continue
            for j, k in mine:
                if k > a.hold:
                    engine.release(j)
                    cols[j] = None

    def start_trial(j, a, k):
        s = a.srf[k]
        cohf, nuf, phif, psif, _ = strength_reduction(s, False, a.coh, nu, a.phi,
                                                      a.psi)
        engine.begin(j, ym, nuf, phif, cohf, psif, dt_viscoplas(nuf, phif, ym))
        cols[j] = (a, k)
        if a.t_start is None:
            a.t_start = clock()

    def fill():
        now = clock()
        for j in range(ncol):
            if cols[j] is not None:
                continue
            if not memo_ready and any(c is not None for c in cols):
                return  # one trial computes the elastic solution first
            placed = False
            for a in admitted:  # next SRF of an admitted analysis
                if not a.done and running_of(a) < srf_parallel:
                    k = next_k(a)
                    if k is not None:
                        start_trial(j, a, k)
                        placed = True
                        break
            if placed:
                continue
            ready = [i for i in queue if arrivals is None or arrivals[i] <= now - t0]
            if not ready:
                continue
            queue.remove(ready[0])
            a = analyses[ready[0]]
            admitted.append(a)
            start_trial(j, a, next_k(a))

    def finish_if_possible(a):
        if a.done:
            return
        # the bracket is known once every SRF before the first failure has
        # converged (the reference's sweep stops at the first failure)
        if a.fail_idx is not None and all(k in a.rows for k in range(a.fail_idx)):
            a.bracket = (a.srf[a.fail_idx - 1] if a.fail_idx > 0 else None,
                         a.srf[a.fail_idx])
            a.done = True
        elif a.fail_idx is None and len(a.rows) == len(a.srf):
            a.done = True  # no failure in the list
        if not a.done:
            return
        a.t_end = clock()
        if a.fail_idx is not None:  # rows the reference would not print
            a.rows = {k: v for k, v in a.rows.items() if k <= a.fail_idx}
        for j, c in enumerate(cols):  # abandon speculative trials
            if c is not None and c[0] is a:
                engine.release(j)
                cols[j] = None
        if on_done:
            on_done(a)

    fill()
    while True:
        if all(c is None for c in cols):
            if not queue:
                break
            nxt = min(arrivals[i] for i in queue) if arrivals is not None else 0.0
            wait = t0 + nxt - clock()
            if wait > 0:
                time.sleep(min(wait, 0.05))
            fill()
            continue
        pending = bool(queue) and arrivals is not None
        cap = poll_steps if pending else (progress_steps if progress else 1 << 40)
        if defer_nl is not None and srf_parallel > 1:
            cap = min(cap, poll_steps)  # look at the trials' progress now and then
        r = engine.advance(cg_tol, cg_maxiter, inner_tol, nl_tol, nl_maxiter,
                           max_steps=cap)
        if progress:
            progress(engine.stats())
        memo_ready = r["memo_ready"]
        for j in np.nonzero(r["ended"])[0]:
            if cols[j] is None:  # abandoned above: its analysis already ended
                continue
            a, k = cols[j]
            if r["errcode"][j] != 0:
                raise PCGNotConverged("ERROR: PCG solver doesn't converge!")
            x, flags = engine.download(j)
            ux, um = uxmax_umax(spec, x)
            a.rows[k] = dict(srf=float(a.srf[k]), cg_tot=int(r["cg_tot"][j]),
                             nl_tot=int(r["nl_tot"][j]), uxmax=ux, umax=um,
                             fmax=float(r["fmax"][j]))
            if keep_state:
                a.flags_last, a.x_last = flags, x
            if on_trial:  # the column still holds the trial's state
                on_trial(a, k, x, flags, lambda j=j: engine.nodal_fields(j))
            if r["nl_tot"][j] >= nl_maxiter and (a.fail_idx is None or k < a.fail_idx):
                a.fail_idx = k
            engine.release(j)
            cols[j] = None
            finish_if_possible(a)
        if defer_nl is not None and srf_parallel > 1:
            apply_holds()
        fill()
    return engine.stats()


def device_mesh(ctx: Context, mesh: HexMesh, spec: Optional[SpecMesh] = None):
    """Resident mesh with its elastic stiffness; returns (DeviceMesh, extload)."""
    spec = spec or hex2spec_device(ctx, mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    dm = DeviceMesh(ctx, spec, len(gam))
    ext, _ = dm.build_stiffness(gam, ym, nu, kprec=K_F64)
    return dm, ext


def _trial_sink(on_srf, output):
    if on_srf is None and output is None:
        return None

    def sink(an, k, x, flags, fields):
        if on_srf:
            on_srf(an.rows[k], x, flags)
        if output is not None:
            output.srf_done(k + 1, an.rows[k], x, fields=fields)
    return sink


def slope3d_mixed(ctx: Context, mesh: HexMesh, *, cg_tol: float, cg_maxiter: int,
                  nl_tol: float, nl_maxiter: int, srf: Sequence[float],
                  inner_tol: float = DEFAULT_INNER_TOL, ncol: int = 1, spec=None,
                  on_srf=None,
                  progress=None, output=None, phinu: bool = False,
                  defer_nl: Optional[int] = None,
                  resid_period: int = DEFAULT_RESID_PERIOD) -> SlopeResult:
    """The reference's SRF sweep (src/slope.F90:244-443) in the mixed-precision
    form: FP32 stiffness and inner CG, FP64 refinement to cg_tol.  With
    ncol > 1 up to ncol SRFs of the sweep run at once; the rows and the
    bracket are those of the sequential sweep.  `on_srf` is called with
    (row, x, plastic flags) as each SRF trial ends (in completion order),
    `progress` with the engine statistics now and then; `output`
    (output.SlopeOutput) receives the rows, displacements and nodal fields
    for the reference's result files.  With `phinu`, SRFs whose Poisson's
    ratio correction changes the stiffness run the whole sweep on the FP64
    drop-in instead (the engine's trials share one operator).
    `resid_period`: FP64 residuals on every resid_period-th engine step
    (BatchEngine.set_residual_period)."""
    spec = spec or hex2spec_device(ctx, mesh)
    _, ym, nu, phi, coh, psi = mesh.material_arrays()
    if phinu and any(strength_reduction(s, True, coh, nu, phi, psi)[4] == 1 for s in srf):
        # src/plastic_library.f90:72-86 changes nu with the SRF, and with it
        # the stiffness (src/slope.F90:259-265): such trials no longer share
        # one operator, which the multi-trial engine relies on, so the sweep
        # runs on the FP64 drop-in, which rebuilds K per SRF on the device
        res = slope3d(ctx, mesh, cg_tol=cg_tol, cg_maxiter=cg_maxiter, nl_tol=nl_tol,
                      nl_maxiter=nl_maxiter, srf=srf, phinu=True, spec=spec,
                      on_srf=on_srf, output=output)
        res.extra["engine"] = "fp64 drop-in (phinu changes the stiffness per SRF)"
        return res
    dm, ext = device_mesh(ctx, mesh, spec)
    eng = BatchEngine(dm, ext, ncol)
    a = Analysis(coh=coh, phi=phi, psi=psi, srf=list(srf))
    try:
        eng.set_residual_period(resid_period)
        run_analyses(eng, [a], ym=ym, nu=nu, cg_tol=cg_tol, cg_maxiter=cg_maxiter,
                     nl_tol=nl_tol, nl_maxiter=nl_maxiter, inner_tol=inner_tol,
                     srf_parallel=ncol, keep_state=True, progress=progress,
                     defer_nl=defer_nl,
                     on_trial=_trial_sink(on_srf, output))
        stats = eng.stats()
    finally:
        eng.close()
        dm.close()
    rows = [a.rows[k] for k in sorted(a.rows)]
    return SlopeResult(rows=rows, cg_log=[], fos_bracket=a.bracket, x=a.x_last,
                       flags=a.flags_last, extra=dict(stats=stats))
