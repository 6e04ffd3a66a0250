#!/usr/bin/env python3
"""Benchmark of the EBE-PCG + viscoplastic hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode analysis|pcg]
                    [--form fp64|mixed|batched] [--workload c1|c0]

Workloads (synthetic, seeded; see specfem3d_geotech_amd.workloads):
  c1  144,000-element two-layer 2H:1V slope with +-2% interior jitter
      (3,489,480 equations), the plain FP64 GPU drop-in configuration.
  c0  1,152-element homogeneous slope (the reference's CPU-runnable case).

Modes:
  analysis  one step = one complete FOS analysis (strength-reduction sweep
            1.00:0.05:2.00 up to the first SRF whose plastic loop reaches
            nl_maxiter); value = FOS analyses/hour.
  pcg       one step = one CG iteration of the elastic solve (fixed count,
            unconverged); value = DOF-iterations/s.
In both modes the element-product kernel is timed with HIP events on its own
stream over the timed region and reported against the HBM roofline.
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

# form -> (operator precision, storage, reduction order, bytes/entry, entries)
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
    ap.add_argument("--mode", default="pcg", choices=["analysis", "pcg"])
    ap.add_argument("--form", default="fp64", choices=list(FORMS))
    ap.add_argument("--workload", default="c1", choices=["c1", "c0"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-iters", type=int, default=30)
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part "
              "of it (trials or mesh partitions would shard across GPUs, see "
              "DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from specfem3d_geotech_amd import workloads
    from specfem3d_geotech_amd.solver import (K_F32, K_F64, KFMT_FULL, KFMT_PACKED,
                                              REDUCE_SEQUENTIAL, Context,
                                              DeviceMesh)
    from specfem3d_geotech_amd.slope import DeviceSlope
    from specfem3d_geotech_amd.plastic import dt_viscoplas, strength_reduction

    wl = workloads.get(a.workload)
    t0 = time.time()
    mesh, spec = wl.build()
    prep_s = time.time() - t0
    ctx = Context(0)
    kp, kf, red, bk, S = FORMS[a.form]
    if red == "sequential":
        ctx.set_reduction(REDUCE_SEQUENTIAL)
    copy_gbs = ctx.calibrate_copy(4 << 30, 10)
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    dm = DeviceMesh(ctx, spec, len(gam))
    t0 = time.time()
    ext, _ = dm.build_stiffness(gam, ym, nu, kprec=K_F64)
    dm.set_operator(K_F64 if kp == "f64" else K_F32,
                    KFMT_PACKED if kf == "packed" else KFMT_FULL)
    kbuild_s = time.time() - t0
    E, N = spec.nelmt, spec.neq
    # algorithmic bytes of one element-product launch: K (E*S*bk) plus the
    # element dof indices (E*81*4), the gathered p (E*81*8) and Y (E*81*8)
    k_bytes = E * S * bk + E * 81 * 4 + 2 * E * 81 * 8
    result = dict(metric="FOS analyses/hour + p50/p99 latency; EBE-PCG HBM GB/s "
                         "vs peak, 1xMI355X")

    if a.mode == "pcg":
        steps = a.steps or 300
        warm = a.warmup if a.warmup is not None else 20
        st = DeviceSlope(dm, ext)
        cohf, nuf, phif, psif, _ = strength_reduction(1.0, False, coh, nu, phi, psi)
        st.begin_srf(ym, nuf, phif, cohf, psif, dt_viscoplas(nuf, phif, ym))
        st.plastic_iter(1e-300, warm, 1.0, False)  # warm-up CG iterations
        st.begin_srf(ym, nuf, phif, cohf, psif, dt_viscoplas(nuf, phif, ym))
        ctx.synchronize()
        ctx.profile(1)
        t0 = time.perf_counter()
        r = st.plastic_iter(1e-300, steps, 1.0, False)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        nlaunch, kms = ctx.profile(0)
        assert r["errcode"] == -1 and r["cg_iter"] == steps + 1
        ms_step = dt / steps * 1e3
        result.update(value=N * steps / dt, unit="DOF-iterations/s",
                      ms_per_step=ms_step, steps=steps, warmup=warm)
        st.close()
    else:
        from specfem3d_geotech_amd.slope import slope3d
        steps = a.steps or 1
        ctl = dict(wl.control)
        dm.close()
        ctx.profile(1)
        t0 = time.perf_counter()
        analyses = []
        for k in range(steps):
            res = slope3d(ctx, mesh, spec=spec, kprec=K_F64 if kp == "f64" else K_F32,
                          kfmt=KFMT_PACKED if kf == "packed" else KFMT_FULL,
                          progress=lambda r: print("  row", json.dumps(r), flush=True),
                          **ctl)
            analyses.append(dict(bracket=res.fos_bracket, rows=res.rows,
                                 n_pcg=len(res.cg_log), cg_total=int(sum(res.cg_log))))
        dt = time.perf_counter() - t0
        nlaunch, kms = ctx.profile(0)
        cg_total = sum(x["cg_total"] for x in analyses)
        result.update(value=steps * 3600.0 / dt, unit="FOS analyses/hour",
                      ms_per_step=dt / steps * 1e3, steps=steps, warmup=0,
                      analyses=analyses, dof_iterations_per_s=N * cg_total / dt,
                      seconds_per_analysis=dt / steps)
        dm = DeviceMesh(ctx, spec, len(gam))  # closed below

    avg_ms = kms / max(nlaunch, 1)
    achieved = k_bytes / (avg_ms * 1e-3) / 1e9
    result.update(
        n_gpus=1, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64", data="synthetic (seeded slope mesh, reference input format)",
        config=dict(workload=wl.description, nelmt=E, neq=N, form=a.form,
                    mode=a.mode, k_storage=f"{kf} {kp} (S={S})", reduction=red,
                    copy_kernel_gbs=round(copy_gbs, 1),
                    host_prep_s=round(prep_s, 2), k_build_s=round(kbuild_s, 4)),
        roofline=dict(kernel=("k_ebe_packed" if kf == "packed" else "k_ebe_elem")
                      + ("<double>" if kp == "f64" else "<float>"), bound="hbm",
                      achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(achieved / HBM_PEAK_GBS, 4),
                      avg_launch_ms=round(avg_ms, 4), launches=nlaunch,
                      bytes_per_launch=k_bytes, traffic=None))
    dm.close()
    ctx.close()

    if not a.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(spec, mesh, a.cpu_sample_iters,
                                              ms_gpu_iter=result.get("ms_per_step"))
    print(json.dumps(result), flush=True)


def cpu_baseline(spec, mesh, iters, ms_gpu_iter=None):
    """The oracle's C restatement (OpenMP) on the same operator: K build,
    then `iters` unconverged CG iterations of the elastic solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    cores = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    gam, ym, nu, *_ = mesh.material_arrays()
    km, dp, ext = O.stiffness_bodyload(spec, gam, nu, ym)
    dinv = np.zeros_like(dp)
    dinv[1:] = 1.0 / dp[1:]
    t0 = time.perf_counter()
    _, it, ec = O.pcg_solver(spec.neq, km, np.zeros(spec.neq + 1), ext, dinv,
                             spec.gdof_elmt, 1e-300, iters)
    dt = time.perf_counter() - t0
    del km
    return dict(value=spec.neq * iters / dt, unit="DOF-iterations/s",
                cores=int(os.environ.get("OMP_NUM_THREADS", cores)), kind="port",
                sample=f"{iters} unconverged FP64 CG iterations of the elastic "
                       f"solve on the same {spec.nelmt}-element mesh (oracle C "
                       f"restatement, OpenMP element products, serial "
                       f"element-order assembly)", ms_per_iter=dt / iters * 1e3)


if __name__ == "__main__":
    main()
