#!/usr/bin/env python3
"""Benchmark of the EBE-PCG + viscoplastic hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--mode sweep|pcg|analysis|serve|mixed] [--form ...]
                    [--workload c1|c0|c2]

Workloads (synthetic, seeded; specfem3d_geotech_amd.workloads):
  c1  144,000-element two-layer 2H:1V slope with +-2% interior jitter
      (3,489,480 equations).
  c0  1,152-element homogeneous slope (the reference's CPU-runnable case).
  c2  400,000-element two-layer slope, face angle 31.4 deg (9,687,800
      equations), the mixed-precision configuration.

Modes:
  sweep (default)  the batched mixed-precision form on the c1 mesh: 64
            strength realisations (form A: lognormal c, mean 10 kPa, CoV 0.3;
            lognormal tan phi, mean tan 20 deg, CoV 0.2), each an FOS sweep
            over SRF 1.00:0.05:2.00, 64 trials resident at once.  A step is
            one batched CG iteration (every running trial advances one CG
            iteration; trials that finish a PCG call take their plastic body
            pass and go on).  value = DOF-iterations/s over all trials
            (equations x CG iterations / s).  The same run times the plain
            FP64 drop-in PCG on the same mesh for the per-trial speedup.
  pcg       the plain FP64 drop-in: one step = one CG iteration of the
            elastic solve (fixed count); value = DOF-iterations/s.
  serve     a Poisson stream (--rate, default 4/s) of --requests analyses on
            seeded meshes of --min/--max-elements (log-uniform): 70% slope FOS
            sweeps (angle 20-45 deg, 1-3 layers; mixed multi-trial engine,
            --columns SRFs in flight, default 8), 30% staged box-pit
            excavations (3-6 stages);
            admission by HBM footprint, concurrent requests on their own
            streams.  value = analyses/hour, with p50/p99 latency.
  mixed     the mixed-precision form (FP32 stiffness and inner CG, FP64
            refinement) against the FP64 drop-in on the same mesh (default
            --workload c2 settings: the 400,000-element slope), SRF by SRF
            (--srf list): displacement error and plastic-flag differences.
  analysis  complete FOS analyses: --form fp64 runs the reference's sweep
            with the FP64 drop-in, one analysis after another; --form batched
            runs --analyses N form-A analyses through the continuous-batching
            scheduler.  value = FOS analyses/hour, with p50/p99 latency.
The element-product kernel of each mode is timed with HIP events on the
library's stream and reported against the HBM roofline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = ("FOS analyses/hour + p50/p99 latency; EBE-PCG HBM GB/s vs peak, "
          "1xMI355X")
DATA = "synthetic (seeded slope mesh in the reference's input format)"

# single-trial forms: (operator precision, storage, reduction, bytes/entry, entries)
FORMS = {
    "fp64": ("f64", "packed", "tree", 8, 3328),
    "fp64full": ("f64", "full", "tree", 8, 6561),
    "fp64ref": ("f64", "full", "sequential", 8, 6561),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--mode", default="sweep",
                    choices=["sweep", "pcg", "analysis", "serve", "mixed"])
    ap.add_argument("--form", default=None,
                    choices=list(FORMS) + ["batched"])
    ap.add_argument("--workload", default="c1", choices=["c1", "c0", "c2"])
    ap.add_argument("--srf", default=None,
                    help="mixed mode: comma-separated SRF list")
    ap.add_argument("--trials", type=int, default=64)
    ap.add_argument("--analyses", type=int, default=None)
    ap.add_argument("--ref-seconds", type=int, default=20,
                    help="analysis mode: seconds of the compiled reference's CPU sample")
    ap.add_argument("--ref-procs", type=int, default=16,
                    help="analysis mode: concurrent reference processes (one core each)")
    ap.add_argument("--rate", type=float, default=None,
                    help="analysis mode: Poisson arrival rate (1/s) of requests")
    ap.add_argument("--srf-parallel", type=int, default=1,
                    help="analysis mode, batched: SRFs of one analysis in flight")
    ap.add_argument("--defer-nl", type=int, default=None,
                    help="analysis mode, batched: release an analysis's trials above a "
                    "trial that has run this many plastic iterations without converging "
                    "(latency over throughput; results unchanged)")
    ap.add_argument("--own-strengths", action="store_true",
                    help="analysis mode: the mesh's own strengths for every analysis")
    ap.add_argument("--requests", type=int, default=32)
    ap.add_argument("--min-elements", type=int, default=500)
    ap.add_argument("--max-elements", type=int, default=5000)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--columns", type=int, default=8,
                    help="serve mode: trial columns per slope request (SRFs in flight)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="batched engine: launch every step's kernels one by one "
                         "(kernel-trace profilers record graph replays poorly)")
    ap.add_argument("--geometry", default=None,
                    help="sweep mode: MFMA product launch geometry 'blocks_per_cu,xcd_split' "
                         "instead of the autotuner's pick (e.g. 3,1)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--residual-period", type=int, default=8,
                    help="batched engine: FP64 residuals on every n-th CG step (1, 2, 4, 8)")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="batched engine: FP64 residual once r.z fell by this factor "
                    "(default: the engine's)")
    ap.add_argument("--roofline-steps", type=int, default=64,
                    help="sweep mode: steps of the HIP-event window after the timed one")
    ap.add_argument("--cpu-sample-iters", type=int, default=8,
                    help="CG iterations of each reference process in the CPU baseline")
    ap.add_argument("--cpu-port-iters", type=int, default=20,
                    help="CG iterations of the oracle restatement in the CPU baseline")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def single_trial_pcg(ctx, dm, ext, ym, nu, phi, coh, psi, steps, warm):
    """FP64 drop-in: `steps` unconverged CG iterations of the elastic solve
    (warm-up first); returns (seconds, launches, kernel_ms)."""
    from specfem3d_geotech_amd.plastic import dt_viscoplas, strength_reduction
    from specfem3d_geotech_amd.slope import DeviceSlope
    st = DeviceSlope(dm, ext)
    try:
        cohf, nuf, phif, psif, _ = strength_reduction(1.0, False, coh, nu, phi, psi)
        dt = dt_viscoplas(nuf, phif, ym)
        st.begin_srf(ym, nuf, phif, cohf, psif, dt)
        st.plastic_iter(1e-300, warm, 1.0, False)
        st.begin_srf(ym, nuf, phif, cohf, psif, dt)
        ctx.synchronize()
        ctx.profile(1)
        t0 = time.perf_counter()
        r = st.plastic_iter(1e-300, steps, 1.0, False)
        ctx.synchronize()
        dt_s = time.perf_counter() - t0
        n, ms = ctx.profile(0)
        assert r["errcode"] == -1 and r["cg_iter"] == steps + 1
    finally:
        st.close()
    return dt_s, n, ms


def measured_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the committed PMC summary
    (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE runs scaled by what
    the counters read for a copy of known size), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
            return json.load(fh)[kernel_key]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def roofline(kernel, bytes_per_launch, n, ms, extra=None, traffic=None):
    avg = ms / max(n, 1)
    ach = bytes_per_launch / (avg * 1e-3) / 1e9 if n else 0.0
    out = dict(kernel=kernel, bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS,
               unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4), avg_launch_ms=round(avg, 4),
               launches=n, bytes_per_launch=int(bytes_per_launch), traffic=traffic)
    if extra:
        out.update(extra)
    return out


def mode_sweep(a, ctx, wl, mesh, spec, prep_s, copy_gbs):
    import numpy as np
    from specfem3d_geotech_amd import workloads
    from specfem3d_geotech_amd.batch import DEFAULT_INNER_TOL, BatchEngine, device_mesh
    from specfem3d_geotech_amd.plastic import dt_viscoplas, strength_reduction
    from specfem3d_geotech_amd.solver import K_F64, KFMT_PACKED
    steps = a.steps or 1000
    warm = a.warmup if a.warmup is not None else 100
    prof_steps = a.roofline_steps
    B = a.trials
    ctl = wl.control
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    t0 = time.time()
    dm, ext = device_mesh(ctx, mesh, spec)
    kbuild_s = time.time() - t0
    E, N = spec.nelmt, spec.neq
    # plain FP64 drop-in on the same mesh (packed FP64 operator)
    dm.set_operator(K_F64, KFMT_PACKED)
    t64, n64, ms64 = single_trial_pcg(ctx, dm, ext, ym, nu, phi, coh, psi, 200, 10)
    fp64_ms_iter = t64 / 200 * 1e3
    eng = BatchEngine(dm, ext, B)
    if a.no_graphs:
        eng.set_graphs(False)
    eng.set_residual_period(a.residual_period)
    real = workloads.form_a(B, len(gam))
    srf = list(ctl["srf"])
    nxt = [0] * B  # next SRF index per column (each column runs one analysis)
    itol = a.inner_tol if a.inner_tol is not None else DEFAULT_INNER_TOL
    kw = dict(cg_tol=ctl["cg_tol"], cg_maxiter=ctl["cg_maxiter"], inner_tol=itol,
              nl_tol=ctl["nl_tol"], nl_maxiter=ctl["nl_maxiter"])
    done_rows = []

    def start(j):
        c, p = real[j]
        cohf, nuf, phif, psif, _ = strength_reduction(srf[nxt[j]], False, c, nu, p, psi)
        eng.begin(j, ym, nuf, phif, cohf, psif, dt_viscoplas(nuf, phif, ym))
        nxt[j] += 1

    def handle(r):
        for j in np.nonzero(r["ended"])[0]:
            if r["errcode"][j] != 0:
                raise RuntimeError("ERROR: PCG solver doesn't converge!")
            done_rows.append((int(j), srf[nxt[j] - 1], int(r["nl_tot"][j]),
                              int(r["cg_tot"][j])))
            failed = r["nl_tot"][j] >= ctl["nl_maxiter"]
            eng.release(j)
            if failed or nxt[j] >= len(srf):  # analysis ended: start it over
                nxt[j] = 0
            start(j)

    try:
        t0 = time.time()
        start(0)  # the elastic solution first, then every column seeded with it
        while True:
            r = eng.advance(max_steps=1 << 30, **kw)
            handle(r)
            if r["memo_ready"]:
                break
        for j in range(1, B):
            start(j)
        log(f"elastic solve and start: {time.time() - t0:.1f} s")
        s0 = eng.stats()
        while eng.stats()["batched_steps"] - s0["batched_steps"] < warm:
            handle(eng.advance(max_steps=warm, **kw))
        ctx.synchronize()
        tune = eng.autotune(5)
        if a.geometry:  # a fixed launch geometry instead of the tuner's pick
            bpc, xs = (int(v) for v in a.geometry.split(","))
            eng.set_geometry(bpc, bool(xs))
            tune.update(blocks_per_cu=bpc, xcd_split=xs, fixed=True)
        log("autotune (MFMA product launch geometry): " + ", ".join(
            f"{c['blocks_per_cu']}/CU {'xcd' if c['xcd_split'] else 'flat'} {c['ms']:.3f} ms"
            for c in tune["candidates"]))
        def window(nsteps):
            s_a = eng.stats()
            t_a = time.perf_counter()
            while True:
                left = nsteps - (eng.stats()["batched_steps"] - s_a["batched_steps"])
                if left <= 0:
                    break
                handle(eng.advance(max_steps=left, **kw))
            ctx.synchronize()
            return time.perf_counter() - t_a, s_a, eng.stats()

        # the timed window: the production path (steps replayed from HIP
        # graphs, no per-launch events), exactly `steps` batched steps
        ctx.profile(0)
        ended0 = len(done_rows)
        dt_s, s0, s1 = window(steps)
        ended_in_window = len(done_rows) - ended0
        log(f"  timed: {steps} steps in {dt_s:.3f} s")
        # the roofline window right after: the same steps launched one by one
        # with HIP events around every MFMA product on the library's stream
        ctx.profile(1)
        dtp_s, sp0, sp1 = window(prof_steps)
        nl, kms = ctx.profile(0)
        nbytes = eng.nbytes()
    finally:
        eng.close()
        dm.close()
    nsteps = s1["batched_steps"] - s0["batched_steps"]
    ncg = s1["cg_iterations"] - s0["cg_iterations"]
    b_eff = ncg / max(nsteps, 1)  # trials iterating per step
    per_trial_ms = dt_s / max(ncg, 1) * 1e3
    # algorithmic bytes of one element-product launch at the roofline
    # window's columns: packed FP32 K, gdof, the gathered P, the written Y
    # and the written FP64 p.Kp share of the iterating columns
    b_prof = (sp1["cg_iterations"] - sp0["cg_iterations"]) / max(
        sp1["batched_steps"] - sp0["batched_steps"], 1)
    kbytes = E * (3328 * 4 + 81 * 4 + (2 * 81 * 4 + 8) * b_prof)
    res = dict(
        metric=METRIC, value=N * ncg / dt_s, unit="DOF-iterations/s", n_gpus=1,
        steps=nsteps, warmup=warm, ms_per_step=dt_s / nsteps * 1e3,
        higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32 inner CG (MFMA) + f64 dots/x/residual/yield", data=DATA,
        config=dict(workload=wl.description + f"; {B} form-A strength realisations "
                    "resident, each column walking its own SRF sweep 1.00:0.05:2.00 in "
                    f"steady state; the window is {nsteps} batched CG steps (each advances "
                    "every iterating trial one CG iteration), with the plastic body passes, "
                    "PCG starts and FP64 residuals that fall inside it; "
                    f"{ended_in_window} SRF trials ended in the window",
                    nelmt=E, neq=N, trials=B,
                    form="batched mixed precision", k_storage="packed f32 (S=3328) + "
                    "packed f64 for the residual", cg_tol=ctl["cg_tol"], inner_tol=itol,
                    residual_period=a.residual_period,
                    nl_tol=ctl["nl_tol"], nl_maxiter=ctl["nl_maxiter"],
                    batch_hbm_gb=round(nbytes / 1e9, 2), copy_kernel_gbs=round(copy_gbs, 1),
                    ingest_s=round(prep_s, 3), k_build_s=round(kbuild_s, 3)),
        trial_iterations=ncg, mean_trials_iterating=round(b_eff, 2),
        fp64_residuals=s1["fp64_residuals"] - s0["fp64_residuals"],
        per_trial_iteration_ms=round(per_trial_ms, 5),
        fp64_dropin=dict(ms_per_iteration=round(fp64_ms_iter, 4),
                         dof_iterations_per_s=N / (fp64_ms_iter * 1e-3),
                         operator="packed f64",
                         ebe_kernel_ms=round(ms64 / max(n64, 1), 4)),
        speedup_vs_fp64_per_trial_iteration=round(fp64_ms_iter / per_trial_ms, 2),
        trials_ended_in_window=ended_in_window,
        autotune=dict(blocks_per_cu=tune["blocks_per_cu"], xcd_split=tune["xcd_split"],
                      candidates_ms={f"{c['blocks_per_cu']}/CU,"
                                     f"{'xcd' if c['xcd_split'] else 'flat'}":
                                     round(c["ms"], 4) for c in tune["candidates"]}),
        timing=dict(headline="HIP-graph replays of the batched steps (production path), "
                    "wall clock around exactly `steps` steps",
                    roofline=f"{prof_steps} further steps launched one by one with HIP "
                             "events around every k_bmv_mfma launch on the library's stream",
                    roofline_window_ms_per_step=round(dtp_s / max(prof_steps, 1) * 1e3, 4)),
        roofline=roofline("k_bmv_mfma<4> (v_mfma_f32_16x16x4f32)", kbytes, nl, kms,
                          dict(bytes_model="E*(3328*4 + 81*4 + (2*81*4 + 8)*b), "
                               "b = mean trials iterating per step of the roofline window",
                               b=round(b_prof, 2),
                               traffic_source="profiles/batched_step_c1_b64_traffic.md "
                               "(c1, 64 trials iterating)"),
                          traffic=measured_traffic("k_bmv_mfma<4>")
                          if (a.workload == "c1" and B == 64) else None))
    return res


def mode_pcg(a, ctx, wl, mesh, spec, prep_s, copy_gbs):
    from specfem3d_geotech_amd.solver import (K_F32, K_F64, KFMT_FULL, KFMT_PACKED,
                                              REDUCE_SEQUENTIAL, DeviceMesh)
    form = a.form or "fp64"
    kp, kf, red, bk, S = FORMS[form]
    if red == "sequential":
        ctx.set_reduction(REDUCE_SEQUENTIAL)
    steps = a.steps or 300
    warm = a.warmup if a.warmup is not None else 20
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    dm = DeviceMesh(ctx, spec, len(gam))
    t0 = time.time()
    ext, _ = dm.build_stiffness(gam, ym, nu, kprec=K_F64)
    dm.set_operator(K_F64 if kp == "f64" else K_F32,
                    KFMT_PACKED if kf == "packed" else KFMT_FULL)
    kbuild_s = time.time() - t0
    try:
        dt_s, n, ms = single_trial_pcg(ctx, dm, ext, ym, nu, phi, coh, psi, steps, warm)
    finally:
        dm.close()
    E, N = spec.nelmt, spec.neq
    kbytes = E * (S * bk + 81 * 4 + 2 * 81 * 8)
    return dict(
        metric=METRIC, value=N * steps / dt_s, unit="DOF-iterations/s", n_gpus=1,
        steps=steps, warmup=warm, ms_per_step=dt_s / steps * 1e3, higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="f64", data=DATA,
        config=dict(workload=wl.description, nelmt=E, neq=N, form=form, mode="pcg",
                    k_storage=f"{kf} {kp} (S={S})", reduction=red,
                    copy_kernel_gbs=round(copy_gbs, 1), ingest_s=round(prep_s, 3),
                    k_build_s=round(kbuild_s, 4)),
        roofline=roofline({("packed", "f64"): "k_ebe_packed_pipe",
                           ("packed", "f32"): "k_ebe_packed<float>",
                           ("full", "f64"): "k_ebe_elem<double>",
                           ("full", "f32"): "k_ebe_elem<float>"}.get((kf, kp), kf),
                          kbytes, n, ms,
                          dict(bytes_model=f"E*({S}*{bk} + 81*4 + 2*81*8)")))


def percentile(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def mode_analysis(a, ctx, wl, mesh, spec, prep_s, copy_gbs):
    import numpy as np
    from specfem3d_geotech_amd import workloads
    from specfem3d_geotech_amd.solver import K_F64, KFMT_PACKED
    form = a.form or "fp64"
    ctl = dict(wl.control)
    E, N = spec.nelmt, spec.neq
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    nan = a.analyses or (1 if form != "batched" else a.trials)
    real = workloads.form_a(nan, len(gam))
    if a.own_strengths:
        real = [(coh.copy(), phi.copy())] * nan
    lat, rows_out = [], []
    # HIP-event timing makes the batched engine launch its steps one by one
    # instead of replaying graphs, so it is only taken for the FP64 form
    ctx.profile(1 if form != "batched" else 0)
    if form == "batched":
        from specfem3d_geotech_amd.batch import Analysis, BatchEngine, device_mesh, \
            run_analyses
        dm, ext = device_mesh(ctx, mesh, spec)
        eng = BatchEngine(dm, ext, a.trials)
        if a.no_graphs:
            eng.set_graphs(False)
        an = [Analysis(coh=c, phi=p, psi=psi.copy(), srf=list(ctl["srf"]), tag=i)
              for i, (c, p) in enumerate(real)]
        arrivals = None
        if a.rate:
            rng = np.random.default_rng(4)
            arrivals = list(np.cumsum(rng.exponential(1.0 / a.rate, nan)))

        def done(x):
            log(f"  analysis {x.tag} bracket {x.bracket} latency {x.latency:.1f} s, "
                f"{len(x.rows)} SRFs")

        def prog(st):
            log(f"    {st['batched_steps']} steps, {st['cg_iterations']} CG iterations, "
                f"{time.perf_counter() - t0:.0f} s")
        t0 = time.perf_counter()
        try:
            stats = run_analyses(eng, an, ym=ym, nu=nu, cg_tol=ctl["cg_tol"],
                                 cg_maxiter=ctl["cg_maxiter"], nl_tol=ctl["nl_tol"],
                                 nl_maxiter=ctl["nl_maxiter"], arrivals=arrivals,
                                 on_done=done, srf_parallel=a.srf_parallel,
                                 defer_nl=a.defer_nl, progress=prog, progress_steps=4000)
        finally:
            eng.close()
            dm.close()
        dt_s = time.perf_counter() - t0
        lat = [x.latency for x in an]
        rows_out = [dict(tag=x.tag, c=float(x.coh[0]), phi=float(x.phi[0]),
                         bracket=x.bracket, cg_tot=sum(r["cg_tot"] for r in x.rows.values()),
                         nl_tot=sum(r["nl_tot"] for r in x.rows.values()))
                    for x in an]
        extra = dict(engine=stats)
        kname = "k_bmv_mfma"
    else:
        from specfem3d_geotech_amd.slope import slope3d
        kp, kf, red, bk, S = FORMS[form]
        t0 = time.perf_counter()
        for i, (c, p) in enumerate(real):
            m2 = mesh.with_strengths(c, p) if (a.analyses and not a.own_strengths) \
                else mesh
            t1 = time.perf_counter()
            last = [time.perf_counter()]

            def prog(row):
                log("  row", json.dumps(row), f"{time.perf_counter() - last[0]:.1f} s")
                last[0] = time.perf_counter()

            def it_prog(nl, cg):
                if nl % 20 == 0:
                    log(f"    plastic iteration {nl}, cg {cg}")
            res = slope3d(ctx, m2, spec=spec, kprec=K_F64, kfmt=KFMT_PACKED, progress=prog,
                          iter_progress=it_prog, **ctl)
            lat.append(time.perf_counter() - t1)
            rows_out.append(dict(bracket=res.fos_bracket, rows=res.rows,
                                 cg_total=int(sum(res.cg_log))))
        dt_s = time.perf_counter() - t0
        extra = {}
        kname = "k_ebe_packed_pipe"
    nl, kms = ctx.profile(0)
    out = dict(
        metric=METRIC, value=nan * 3600.0 / dt_s, unit="FOS analyses/hour", n_gpus=1,
        steps=nan, warmup=0, ms_per_step=dt_s / nan * 1e3, higher_is_better=True,
        scaling="weak", vs_baseline=None,
        dtype="f64" if form != "batched" else "f32 inner CG (MFMA) + f64", data=DATA,
        config=dict(workload=wl.description, nelmt=E, neq=N, form=form, mode="analysis",
                    analyses=nan, trials=a.trials if form == "batched" else 1,
                    arrival_rate=a.rate, srf_parallel=a.srf_parallel, defer_nl=a.defer_nl,
                    strengths="own" if (a.own_strengths or not a.analyses and
                                        form != "batched") else "form A"),
        latency_s=dict(p50=percentile(lat, 50), p99=percentile(lat, 99), max=max(lat)),
        analyses=rows_out, kernel=dict(name=kname, launches=nl,
                                       avg_ms=kms / max(nl, 1)), **extra)
    cgs = [r.get("cg_tot", r.get("cg_total", 0)) for r in rows_out]
    out["cg_iterations_per_analysis"] = sum(cgs) / max(len(cgs), 1)
    return out


def mode_mixed(a, ctx, wl, mesh, spec, prep_s, copy_gbs):
    """The mixed-precision form against the plain FP64 drop-in, SRF by SRF on
    one mesh: each form runs the same SRF list (FP64: packed FP64 operator,
    one SRF after another; mixed: FP32 stiffness and inner CG with FP64
    refinement, up to 8 SRFs at once).  Per SRF: both summary rows, the
    relative L2 difference of the displacement, and the number of GLL
    points whose plastic flag differs."""
    import numpy as np
    from specfem3d_geotech_amd.batch import DEFAULT_INNER_TOL, slope3d_mixed
    from specfem3d_geotech_amd.slope import slope3d
    from specfem3d_geotech_amd.solver import K_F64, KFMT_PACKED
    srf = [float(v) for v in a.srf.split(",")] if a.srf else [1.0, 1.05, 1.1, 1.15]
    ctl = dict(wl.control, srf=srf)
    ncol = 1
    while ncol < min(len(srf), 8):
        ncol *= 2
    keep64, keepm = {}, {}

    def keep(store, tag):
        def f(row, x, flags):
            store[row["srf"]] = (row, x, flags)
            log(f"  {tag} row", json.dumps(row))
        return f
    t0 = time.perf_counter()
    r64 = slope3d(ctx, mesh, spec=spec, kprec=K_F64, kfmt=KFMT_PACKED,
                  on_srf=keep(keep64, "fp64"),
                  iter_progress=lambda nl, cg: nl % 10 == 0 and log(
                      f"    fp64 plastic iteration {nl}, cg {cg}"), **ctl)
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    rm = slope3d_mixed(ctx, mesh, spec=spec, ncol=ncol, on_srf=keep(keepm, "mixed"),
                       progress=lambda st: log(f"    mixed {st['batched_steps']} steps"),
                       **ctl)
    tm = time.perf_counter() - t0
    per = []
    for s in srf:
        if s not in keep64 or s not in keepm:
            continue
        (ra, xa, fa), (rb, xb, fb) = keep64[s], keepm[s]
        per.append(dict(
            srf=s, fp64=ra, mixed=rb,
            x_rel_l2=float(np.linalg.norm(xb - xa) / max(np.linalg.norm(xa), 1e-300)),
            uxmax_rel=abs(rb["uxmax"] - ra["uxmax"]) / ra["uxmax"],
            nl_tot_equal=ra["nl_tot"] == rb["nl_tot"],
            plastic_flags_differing=int(np.count_nonzero(fa != fb))))
    cg64 = sum(r["cg_tot"] for r in r64.rows)
    cgm = sum(r["cg_tot"] for r in rm.rows)
    E, N = spec.nelmt, spec.neq
    return dict(
        metric=METRIC, value=N * cgm / tm, unit="DOF-iterations/s", n_gpus=1,
        steps=cgm, warmup=0, ms_per_step=tm / max(cgm, 1) * 1e3, higher_is_better=True,
        scaling="weak", vs_baseline=None,
        dtype="f32 stiffness + inner CG, f64 refinement/dots/x/yield", data=DATA,
        config=dict(workload=wl.description, nelmt=E, neq=N, mode="mixed", srf=srf,
                    srf_in_flight=ncol, cg_tol=ctl["cg_tol"], inner_tol=DEFAULT_INNER_TOL,
                    nl_tol=ctl["nl_tol"], nl_maxiter=ctl["nl_maxiter"],
                    ingest_s=round(prep_s, 3)),
        fp64=dict(seconds=round(t64, 2), cg_iterations=cg64,
                  dof_iterations_per_s=N * cg64 / t64, bracket=r64.fos_bracket),
        mixed=dict(seconds=round(tm, 2), cg_iterations=cgm, bracket=rm.fos_bracket,
                   engine=rm.extra.get("stats")),
        per_srf=per,
        max_x_rel_l2=max((p["x_rel_l2"] for p in per), default=None))


def mode_serve(a):
    from specfem3d_geotech_amd.serving import Server, make_requests
    reqs = make_requests(a.requests, rate=a.rate or 4.0,
                         nelmt_range=(a.min_elements, a.max_elements), seed=4)
    t_start = time.perf_counter()

    def event(r, text):
        log(f"    {time.perf_counter() - t_start:7.1f} s  request {r.tag} ({r.kind}, "
            f"{r.nelmt} el): {text}")
    srv = Server(max_workers=a.workers, slope_columns=a.columns, on_event=event)

    def done(r):
        log(f"  request {r.tag} {r.kind} {r.nelmt} el: latency {r.latency:.2f} s")
    st = srv.serve(reqs, on_done=done)
    kinds = [r.kind for r in reqs]
    return dict(
        metric=METRIC, value=st["analyses_per_hour"], unit="FOS analyses/hour",
        n_gpus=1, steps=len(reqs), warmup=0, ms_per_step=st["wall_s"] / len(reqs) * 1e3,
        higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32 inner CG (MFMA) + f64 (slopes); f64 (excavations)", data=DATA,
        config=dict(workload=f"serving: Poisson {a.rate or 4.0}/s, {len(reqs)} requests, "
                    f"meshes {a.min_elements}-{a.max_elements} elements log-uniform, "
                    f"{kinds.count('slope')} slope sweeps + {kinds.count('excavation')} "
                    "staged excavations", workers=a.workers, columns=a.columns,
                    mode="serve"),
        latency_s=dict(p50=st["p50_s"], p99=st["p99_s"], max=st["max_s"]),
        serving=st)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; runs on more GPUs are not part "
              "of it (trials would shard across GPUs with the mesh replicated)",
              file=sys.stderr)
        sys.exit(2)
    from specfem3d_geotech_amd import workloads

    # the compiled reference on the CPU runs first, as child processes of a
    # process that has not touched the GPU yet
    ref_cpu = cpu = None
    if a.mode == "analysis" and not a.no_cpu_baseline:
        wl0 = workloads.get(a.workload)
        ref_cpu = reference_cpu_sample(wl0, wl0.make(), a.ref_seconds, a.ref_procs)
    if a.mode in ("sweep", "pcg") and not a.no_cpu_baseline:
        cpu = cpu_baseline(workloads.get(a.workload).make(), a.cpu_sample_iters,
                           a.ref_procs, a.cpu_port_iters)
    from specfem3d_geotech_amd.solver import Context

    if a.mode == "serve":
        print(json.dumps(mode_serve(a)), flush=True)
        return
    from specfem3d_geotech_amd.spec import hex2spec_device
    wl = workloads.get(a.workload)
    ctx = Context(0)
    hex2spec_device(ctx, workloads.get("c0").make())  # loads the numbering kernels
    # ingest: the seeded mesh, then its GLL numbering on the device
    t0 = time.time()
    mesh = wl.make()
    spec = hex2spec_device(ctx, mesh)
    prep_s = time.time() - t0
    copy_gbs = ctx.calibrate_copy(4 << 30, 10)
    try:
        if a.mode == "sweep":
            res = mode_sweep(a, ctx, wl, mesh, spec, prep_s, copy_gbs)
        elif a.mode == "pcg":
            res = mode_pcg(a, ctx, wl, mesh, spec, prep_s, copy_gbs)
        elif a.mode == "mixed":
            res = mode_mixed(a, ctx, wl, mesh, spec, prep_s, copy_gbs)
        else:
            res = mode_analysis(a, ctx, wl, mesh, spec, prep_s, copy_gbs)
    finally:
        ctx.close()
    if cpu is not None:
        res["cpu_baseline"] = cpu
    if ref_cpu is not None:
        # analyses/hour of the reference on these cores, from its measured
        # PCG rate and this run's CG iterations per analysis; both the
        # plastic body's share of the reference's time and (in the batched
        # form) the elastic solves the engine reuses are left out, which
        # favours the reference
        cg_per = res.get("cg_iterations_per_analysis")
        if cg_per:
            ref_cpu["value"] = ref_cpu["procs"] * ref_cpu["cg_per_s_per_process"] * 3600.0 \
                / cg_per
            ref_cpu["unit"] = "FOS analyses/hour"
            ref_cpu["cg_iterations_per_analysis"] = cg_per
        res["cpu_baseline"] = ref_cpu
    print(json.dumps(res), flush=True)


def reference_cpu_sample(wl, mesh, seconds, procs):
    """The compiled reference (oracle/_ref/semgeotech_tap: the reference's own
    Fortran sources built with amdflang -O2, plus a link-time tap that logs
    each pcg_solver call's cg_iter and wall time) on this workload's input:
    `procs` independent serial processes, one core each, stopped after
    `seconds`.  Returns each process's PCG rate (CG iterations per second of
    solver time, completed calls only), or None when the binary is absent."""
    import shutil
    import subprocess
    import tempfile
    from specfem3d_geotech_amd.mesh import write_reference_input
    binp = os.path.join(ROOT, "oracle", "_ref", "semgeotech_tap")
    if not os.path.exists(binp):
        log("reference CPU sample skipped: oracle/_ref/semgeotech_tap is not built")
        return None
    ctl = dict(wl.control)
    # the reference's parser keeps 80 characters of a value list
    # (src/string_library.f90:547): the input lists the sweep's first 8 SRFs,
    # more than a sample of this length reaches
    srf = list(ctl["srf"])[:8]
    work = tempfile.mkdtemp(prefix="refcpu_")
    jobs = []
    for i in range(procs):
        d = os.path.join(work, f"p{i}")
        sem = write_reference_input(mesh, d, "ref", cg_tol=ctl["cg_tol"],
                                    cg_maxiter=ctl["cg_maxiter"], nl_tol=ctl["nl_tol"],
                                    nl_maxiter=ctl["nl_maxiter"], srf=srf)
        env = dict(os.environ, SGEO_TAP_DIR=d, SGEO_TAP_MODE="time", OMP_NUM_THREADS="1")
        cmd = f"ulimit -s unlimited; exec timeout -s KILL {int(seconds)} {binp} {sem}"
        jobs.append((subprocess.Popen(["bash", "-c", cmd], cwd=d, env=env,
                                      stdout=subprocess.DEVNULL,
                                      stderr=subprocess.DEVNULL), d))
    log(f"reference CPU sample: {procs} processes for {seconds} s")
    rates, neq = [], None
    for p, d in jobs:
        p.wait()
        cg = t = 0.0
        try:
            with open(os.path.join(d, "pcg_times.txt")) as fh:
                for ln in fh:
                    _, it, ec, sec = ln.split()
                    if int(ec) == 0:
                        cg += int(it)
                        t += float(sec)
        except OSError:
            pass
        if t > 0:
            rates.append(cg / t)
    if not rates:
        log("reference CPU sample: no PCG call completed")
        return None
    shutil.rmtree(work, ignore_errors=True)
    from specfem3d_geotech_amd.spec import hex2spec
    neq = hex2spec(mesh).neq
    r = sum(rates) / len(rates)
    return dict(kind="reference", cores=procs, procs=procs,
                cg_per_s_per_process=r, dof_iterations_per_s=procs * r * neq,
                sample=f"{procs} serial processes of the compiled reference "
                       f"(amdflang -O2, its own sources) on the {mesh.nelmt}-element "
                       f"input, {seconds} s each; completed pcg_solver calls timed by "
                       "a link-time tap")


def cpu_baseline(mesh, iters, procs, port_iters):
    """CPU baselines of the EBE-PCG on the benchmark's own mesh, on this
    host's cores, before the GPU is touched.

    kind "reference": the reference's own compiled pcg_solver (the
    unmodified src/solver.f90 built by oracle/Makefile, driven by
    oracle/_ref/ref_pcg_bench) on the elastic system K u = extload of this
    mesh, `procs` concurrent serial processes (one core each, one shared
    memory-mapped copy of K), `iters` unconverged iterations each.  K,
    dprecon and extload come from the oracle's C restatement (OpenMP), which
    equals the reference's stiffness_bodyload bit for bit; the reference's
    own serial K build of a 144,000-element mesh takes minutes and is not
    part of the timed solve.  Secondary (kind "port"): the oracle's C
    restatement of pcg_solver (OpenMP element products) on the same K."""
    import shutil
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from specfem3d_geotech_amd.spec import hex2spec
    cores = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    t0 = time.perf_counter()
    km, dp, ext = O.stiffness_bodyload(spec, gam, nu, ym)
    dinv = np.zeros_like(dp)
    dinv[1:] = 1.0 / dp[1:]
    log(f"cpu baseline: K of {spec.nelmt} elements built on the host in "
        f"{time.perf_counter() - t0:.1f} s")
    out = None
    binp = os.path.join(ROOT, "oracle", "_ref", "ref_pcg_bench")
    if os.path.exists(binp):
        base = "/dev/shm" if os.path.isdir("/dev/shm") else None
        work = tempfile.mkdtemp(prefix="refpcg_", dir=base)
        try:
            km.tofile(os.path.join(work, "k.bin"))
            ext.tofile(os.path.join(work, "f.bin"))
            dinv.tofile(os.path.join(work, "dinv.bin"))
            np.ascontiguousarray(spec.gdof_elmt, dtype=np.int32).tofile(
                os.path.join(work, "gdof.bin"))
            cmd = (f"ulimit -s unlimited; exec {binp} {work} {spec.neq} {spec.nelmt} "
                   f"{iters}")
            env = dict(os.environ, OMP_NUM_THREADS="1")
            jobs = [subprocess.Popen(["bash", "-c", cmd], env=env, stdout=subprocess.PIPE,
                                     stderr=subprocess.DEVNULL, text=True)
                    for _ in range(procs)]
            log(f"cpu baseline: the reference's pcg_solver, {procs} processes x {iters} "
                "iterations")
            secs = []
            for j in jobs:
                o, _ = j.communicate()
                f = o.split()
                if j.returncode == 0 and len(f) == 3 and int(f[0]) == iters + 1:
                    secs.append(float(f[2]))
        finally:
            shutil.rmtree(work, ignore_errors=True)
        if secs:
            per = [spec.neq * iters / t for t in secs]
            out = dict(value=sum(per), unit="DOF-iterations/s", cores=len(secs),
                       kind="reference",
                       sample=f"the reference's compiled pcg_solver (src/solver.f90, "
                              f"amdflang -O2) on this {spec.nelmt}-element mesh's elastic "
                              f"system, {len(secs)} concurrent serial processes x {iters} "
                              "unconverged FP64 CG iterations",
                       seconds_per_iteration=round(sum(secs) / len(secs) / iters, 4))
    if out is None:
        log("cpu baseline: oracle/_ref/ref_pcg_bench not built; the oracle's restatement only")
    t0 = time.perf_counter()
    _, it, ec = O.pcg_solver(spec.neq, km, np.zeros(spec.neq + 1), ext, dinv,
                             spec.gdof_elmt, 1e-300, port_iters)
    dt = time.perf_counter() - t0
    del km
    port = dict(value=spec.neq * port_iters / dt, unit="DOF-iterations/s",
                cores=int(os.environ.get("OMP_NUM_THREADS", cores)), kind="port",
                sample=f"{port_iters} unconverged FP64 CG iterations of the elastic solve "
                       f"on the same {spec.nelmt}-element mesh (oracle C restatement, "
                       "OpenMP element products, serial element-order assembly)",
                ms_per_iter=round(dt / port_iters * 1e3, 2))
    if out is None:
        return port
    out["port"] = port
    return out


if __name__ == "__main__":
    main()
