"""Request loop for FOS analyses of uneven meshes on one GPU (serving form).

Requests (slope FOS sweeps and staged excavations) arrive over time.  A
request is admitted as soon as HBM and a worker slot are free.  Its HBM
footprint is estimated from its mesh size, and the admitted footprints never
exceed the budget.

Each admitted request runs in its own worker thread with its own library
context, and so on its own HIP stream.  The kernels of concurrent requests
therefore overlap on the device, and the ctypes calls release the GIL.  A
request keeps its mesh, stiffness and state resident until it finishes.

- Slope requests run the mixed-precision multi-trial engine with several
  SRFs of the sweep in flight (batch.slope3d_mixed).
- Excavation requests run the staged driver (excavation.excavation3d).

Every request's arithmetic is independent of the others: its own buffers,
stream and deterministic reductions.  A result is therefore the same alone
and in the mix.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from .mesh import HexMesh, Material, pit_mesh, slope_mesh
from .solver import Context


@dataclass
class Request:
    """One analysis: kind 'slope' (FOS sweep) or 'excavation' (staged)."""
    kind: str
    make_mesh: Callable[[], HexMesh]
    control: Dict
    nelmt: int
    arrival: float = 0.0
    tag: object = None
    result: object = None
    error: Optional[BaseException] = None
    t_arrival: Optional[float] = None
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    footprint: int = 0

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


def footprint_bytes(kind: str, nelmt: int, ncol: int = 8) -> int:
    """HBM a request holds at its peak.

    - Stiffness: FP64 full + packed FP64 + packed FP32, and the element
      product buffers.
    - Slope, mixed multi-trial form with `ncol` columns: interleaved vectors
      (about 3.7 equations per element, 4 FP64 + 9 FP32 per column: x, oldx,
      load and body; r and a ring of 8 search directions), element
      products (FP32 and FP64), per-point state (evpt and stress, FP64),
      body loads.
    - Excavation: one FP64 trial (x, oldx, load, body, ext, PCG vectors) and
      its per-point state.

    The estimate is conservative (about 10% above what sgeo_mesh_bytes and
    sgeo_batch_bytes report for these meshes)."""
    neq = int(3.7 * nelmt) + 1
    k = nelmt * (6561 * 8 + 3328 * 12 + 81 * (8 + 4 + 4)) + neq * 8 * 6
    if kind == "slope":
        st = ncol * (neq * (4 * 8 + 9 * 4) + nelmt * 81 * (4 + 8 + 8)
                     + nelmt * 27 * (6 * 8 * 2 + 1))
    else:
        st = neq * 8 * 10 + nelmt * (27 * (6 * 8 * 2 + 1 + 8) + 81 * 8)
    return int(1.1 * (k + st))


class Server:
    """Admission by HBM footprint, one worker thread per admitted request."""

    def __init__(self, device: int = 0, hbm_budget: Optional[int] = None,
                 max_workers: int = 8, slope_columns: int = 8,
                 context_factory=None, defer_frac: Optional[float] = None,
                 on_event: Optional[Callable[[Request, str], None]] = None):
        """on_event(request, text), if given, is called from the workers as a
        request's work proceeds (SRF trials and plastic iterations ending,
        batched-step counts)."""
        self.device = device
        self.on_event = on_event
        self.max_workers = max_workers
        self.slope_columns = slope_columns
        self.defer_frac = defer_frac
        self.context_factory = context_factory or (lambda: Context(device))
        if hbm_budget is None:
            c = self.context_factory()
            try:
                hbm_budget = int(0.85 * c.info()["hbm_bytes"])
            finally:
                c.close()
        self.hbm_budget = hbm_budget
        self._lock = threading.Lock()
        self._reserved = 0
        self._running = 0
        self.peak_reserved = 0
        self.peak_running = 0

    def run_request(self, req: Request, ctx: Context):
        mesh = req.make_mesh()
        ev = self.on_event
        if req.kind == "slope":
            from .batch import slope3d_mixed
            # SRFs run slope_columns at a time; with defer_frac, trials above
            # one that has run that fraction of nl_maxiter wait for it
            defer = (max(10, int(self.defer_frac * int(req.control["nl_maxiter"])))
                     if self.defer_frac else None)
            hooks = {}
            if ev:
                hooks = dict(
                    on_srf=lambda row, x, fl: ev(req, f"SRF {row['srf']} nl {row['nl_tot']}"),
                    progress=lambda st: ev(req, f"{st['batched_steps']} steps"))
            return slope3d_mixed(ctx, mesh, ncol=self.slope_columns, defer_nl=defer,
                                 **hooks, **req.control)
        from .excavation import excavation3d
        from .solver import KFMT_PACKED
        hooks = {}
        if ev:
            hooks = dict(
                progress=lambda row: ev(req, f"stage {row['stage']} nl {row['nl_tot']}"),
                iter_progress=lambda st, inc, nl, cg: nl % 10 == 0 and ev(
                    req, f"stage {st} increment {inc} plastic iteration {nl}"))
        # the packed FP64 operator: half the stiffness bytes of the full
        # (reference-layout) one, the same arithmetic to rounding
        return excavation3d(ctx, mesh, kfmt=KFMT_PACKED, **hooks, **req.control)

    def _worker(self, req: Request, done: threading.Event):
        ctx = None
        try:
            ctx = self.context_factory()
            req.result = self.run_request(req, ctx)
        except BaseException as e:  # reported with the request
            req.error = e
        finally:
            if ctx is not None:
                ctx.close()
            req.t_done = time.perf_counter()
            with self._lock:
                self._reserved -= req.footprint
                self._running -= 1
            done.set()

    def serve(self, requests: List[Request], on_done=None) -> Dict:
        """Runs the requests (arrival = seconds after the start); returns the
        latency and throughput statistics."""
        t0 = time.perf_counter()
        for r in requests:
            r.t_arrival = t0 + r.arrival
            r.footprint = footprint_bytes(r.kind, r.nelmt, self.slope_columns)
            if r.footprint > self.hbm_budget:
                raise ValueError(f"request {r.tag}: {r.footprint} B exceeds the HBM budget")
        queue = sorted(requests, key=lambda r: r.arrival)
        threads = []
        wake = threading.Event()
        pending = list(queue)
        while pending or any(t.is_alive() for t in threads):
            now = time.perf_counter()
            admitted = False
            # admit in arrival order while the head fits (no overtaking)
            while pending and pending[0].t_arrival <= now:
                head = pending[0]
                with self._lock:
                    fits = (self._reserved + head.footprint <= self.hbm_budget and
                            self._running < self.max_workers)
                    if fits:
                        self._reserved += head.footprint
                        self._running += 1
                        self.peak_reserved = max(self.peak_reserved, self._reserved)
                        self.peak_running = max(self.peak_running, self._running)
                if not fits:
                    break
                pending.pop(0)
                head.t_admit = time.perf_counter()
                th = threading.Thread(target=self._worker, args=(head, wake), daemon=True)
                th.start()
                threads.append(th)
                admitted = True
            if admitted:
                continue
            nxt = pending[0].t_arrival - now if pending else 0.5
            wake.wait(timeout=max(0.001, min(nxt, 0.05)))
            wake.clear()
            for r in queue:
                if r.t_done is not None and on_done and not getattr(r, "_reported", False):
                    r._reported = True
                    on_done(r)
        for th in threads:
            th.join()
        wall = time.perf_counter() - t0
        errs = [r for r in requests if r.error is not None]
        if errs:
            raise RuntimeError(f"{len(errs)} request(s) failed: {errs[0].error!r}") \
                from errs[0].error
        lat = np.array([r.latency for r in requests])
        return dict(requests=len(requests), wall_s=wall,
                    analyses_per_hour=len(requests) * 3600.0 / wall,
                    p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                    max_s=float(lat.max()), peak_hbm_reserved=self.peak_reserved,
                    peak_concurrent=self.peak_running)


# ----------------------------------------------------------------------------
# the serving workload: Poisson arrivals, log-uniform mesh sizes, 70% slopes
# (angle 20-45 deg, 1-3 layers), 30% staged box-pit excavations (3-6 stages)
# ----------------------------------------------------------------------------

SLOPE_CONTROL = dict(cg_tol=1e-8, cg_maxiter=5000, nl_tol=5e-4, nl_maxiter=300,
                     srf=[round(1.0 + 0.05 * i, 2) for i in range(21)])


def _dims(nelmt: int, ratio=(2.0, 1.0, 1.0)):
    a = (nelmt / (ratio[0] * ratio[1] * ratio[2])) ** (1.0 / 3.0)
    return tuple(max(2, int(round(a * r))) for r in ratio)


def make_requests(n: int, *, rate: float = 4.0, nelmt_range=(5000, 200000),
                  slope_share: float = 0.7, seed: int = 4,
                  nl_maxiter: Optional[int] = None) -> List[Request]:
    """A seeded request stream.  Slopes: 2H:1V-box geometry with the face
    angle drawn from U[20, 45] deg and 1-3 layers (c, phi per layer drawn
    around 10 kPa / 20 deg).  Excavations: a box with a centred pit removed
    in 3-6 stages (two load increments each), Mohr-Coulomb soil."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    lo, hi = nelmt_range
    for i in range(n):
        ne = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        if rng.uniform() < slope_share:
            nx, ny, nz = _dims(ne, (2.0, 1.0, 1.0))
            angle = float(rng.uniform(20.0, 45.0))
            nl = int(rng.integers(1, 4))
            mats = [Material(gam=20.0, ym=1e5, nu=0.3,
                             phi=float(rng.uniform(18.0, 30.0)),
                             coh=float(rng.uniform(8.0, 20.0))) for _ in range(nl)]
            splits = sorted(rng.uniform(6.0, 18.0, nl - 1).tolist(), reverse=True)

            def mk(nx=nx, ny=ny, nz=nz, angle=angle, mats=mats, splits=splits):
                return slope_mesh(nx, ny, nz, angle_deg=angle, materials=mats,
                                  layer_z=splits)
            ctl = dict(SLOPE_CONTROL)
            if nl_maxiter:
                ctl["nl_maxiter"] = nl_maxiter
            out.append(Request("slope", mk, ctl, nx * ny * nz, float(t[i]), tag=i))
        else:
            nx, ny, nz = _dims(ne, (2.0, 1.5, 1.0))
            nz = max(nz, 6)
            nst = int(rng.integers(3, 7))
            soil = Material(gam=20.0, ym=1e5, nu=0.3, phi=float(rng.uniform(25.0, 35.0)),
                            coh=float(rng.uniform(15.0, 30.0)))
            lx, ly, lz = 2.0 * nx, 2.0 * ny, 2.0 * nz
            depth = lz / nz * max(1, nz // (2 * nst))

            def mk(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, nst=nst, soil=soil,
                   depth=depth):
                return pit_mesh(nx, ny, nz, lx=lx, ly=ly, lz=lz,
                                pit=(lx / 4, 3 * lx / 4, ly / 4, 3 * ly / 4),
                                nstage=nst, stage_depth=depth, soil=soil)
            ctl = dict(cg_tol=1e-8, cg_maxiter=5000, nl_tol=1e-5,
                       nl_maxiter=nl_maxiter or 100, srf=[1.0], nexcav=nst,
                       nexcavid=[1] * nst, excavid=list(range(2, nst + 2)), ninc=2)
            out.append(Request("excavation", mk, ctl, nx * ny * nz, float(t[i]), tag=i))
    return out
