"""Writes the golden fixtures of tests/golden/ from the reference itself.

Runs the unmodified serial reference built by oracle/Makefile
(oracle/_ref/semgeotech_tap, with the pcg_solver / mohcouf taps of
oracle/ref_tap.c) on small seeded meshes written in the reference's own
input format, and on the reference's example `input/validation1`, then
stores inputs and outputs as .npz (no pickles).  Needs /root/reference, so
it runs in the build container only; the fixtures it writes are committed.

    make -C oracle && python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from specfem3d_geotech_amd.mesh import Material, pit_mesh, read_reference_input, \
    slope_mesh, write_reference_input  # noqa: E402

REF_TAP = os.path.join(ROOT, "oracle", "_ref", "semgeotech_tap")
GLL_PROBE = os.path.join(ROOT, "oracle", "_ref", "gll_probe")
FN_PROBE = os.path.join(ROOT, "oracle", "_ref", "fn_probe")
REF_INPUT = "/root/reference/input"
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "semgeotech")
LISTFMT_PROBE = os.path.join(ROOT, "oracle", "_ref", "listfmt_probe")


def run_tap(sem: str, workdir: str, mode: str = "full", extra_env=None) -> None:
    for f in ("pcg_static.bin", "pcg_calls.bin", "pcg_times.txt", "mohcouf.bin"):
        p = os.path.join(workdir, f)
        if os.path.exists(p):
            os.remove(p)
    env = dict(os.environ, SGEO_TAP_DIR=workdir, SGEO_TAP_MODE=mode, **(extra_env or {}))
    # flang places the reference's automatic arrays on the stack
    r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {REF_TAP} {sem}"],
                       cwd=workdir, env=env, capture_output=True, text=True)
    if "Result summary" not in open(find_summary(workdir)).read():
        raise RuntimeError(f"reference run failed: {r.stdout[-2000:]} {r.stderr}")


def find_summary(workdir: str) -> str:
    for root, _, files in os.walk(workdir):
        for f in files:
            if f.endswith("_summary"):
                return os.path.join(root, f)
    raise FileNotFoundError("no summary file")


def parse_summary(path: str) -> np.ndarray:
    """Rows (srf, cg_tot, nl_tot, uxmax, umax, fmax) as float64; the
    reference writes list-directed full precision (src/slope.F90:415)."""
    rows = []
    lines = open(path).read().splitlines()
    start = [i for i, ln in enumerate(lines) if "SRF, CGITER" in ln][0] + 1
    for ln in lines[start:]:
        if "ELAPSED" in ln or not ln.strip():
            break
        rows.append([float(t) for t in ln.split()])
    return np.array(rows, dtype=np.float64)


def read_static(workdir: str):
    raw = np.fromfile(os.path.join(workdir, "pcg_static.bin"), dtype=np.uint8)
    hdr = raw[:16].view(np.int32)
    neq, nel, nedof = int(hdr[0]), int(hdr[1]), int(hdr[2])
    o = 16
    k = raw[o:o + 8 * nedof * nedof * nel].view(np.float64).reshape(nel, nedof, nedof)
    o += 8 * nedof * nedof * nel
    dinv = raw[o:o + 8 * (neq + 1)].view(np.float64).copy()
    o += 8 * (neq + 1)
    g = raw[o:o + 4 * nedof * nel].view(np.int32).reshape(nel, nedof).copy()
    return neq, k, dinv, g


def read_calls(workdir: str, neq: int):
    raw = np.fromfile(os.path.join(workdir, "pcg_calls.bin"), dtype=np.uint8)
    rec = 16 + 3 * 8 * (neq + 1)
    n = raw.size // rec
    cg = np.zeros(n, np.int32)
    uin = np.zeros((n, neq + 1)); f = np.zeros((n, neq + 1)); uout = np.zeros((n, neq + 1))
    for c in range(n):
        r = raw[c * rec:(c + 1) * rec]
        cg[c] = r[:16].view(np.int32)[2]
        uin[c] = r[16:16 + 8 * (neq + 1)].view(np.float64)
        f[c] = r[16 + 8 * (neq + 1):16 + 16 * (neq + 1)].view(np.float64)
        uout[c] = r[16 + 16 * (neq + 1):].view(np.float64)
    return cg, uin, f, uout


def read_calls_var(workdir: str):
    """PCG call records whose neq changes from call to call (excavation
    stages renumber the equations): list of (neq, cg_iter, u_in, f, u_out)."""
    raw = np.fromfile(os.path.join(workdir, "pcg_calls.bin"), dtype=np.uint8)
    out, o = [], 0
    while o < raw.size:
        hdr = raw[o:o + 16].view(np.int32)
        neq, cg = int(hdr[0]), int(hdr[2])
        n1 = neq + 1
        v = raw[o + 16:o + 16 + 24 * n1].view(np.float64).reshape(3, n1).copy()
        out.append((neq, cg, v[0], v[1], v[2]))
        o += 16 + 24 * n1
    return out


def parse_stage_summary(path: str) -> np.ndarray:
    """Excavation rows (stage, cg_tot, nl_tot, uxmax, umax, fmax)
    (src/excavation.F90:634)."""
    lines = open(path).read().splitlines()
    start = [i for i, ln in enumerate(lines) if "STEP, CGITER" in ln][0] + 1
    toks = []  # list-directed output may wrap a row over two lines
    for ln in lines[start:]:
        if "ELAPSED" in ln or not ln.strip():
            break
        toks += [float(t) for t in ln.split()]
    return np.array(toks, dtype=np.float64).reshape(-1, 6)


def excavation_case(name, mesh, control, excav):
    """A seeded pit excavated in stages, run through the reference's
    excavation3d with the pcg_solver tap."""
    work = tempfile.mkdtemp(prefix=f"golden_{name}_")
    ex = (f", nexcav={excav['nexcav']}, excavid="
          + " ".join(str(v) for v in excav["excavid"]) + f", ninc={excav['ninc']}")
    sem = write_reference_input(mesh, work, name, extra_control=ex, **control)
    run_tap(sem, work)
    calls = read_calls_var(work)
    summ = parse_stage_summary(find_summary(work))
    neqs = np.array([c[0] for c in calls], np.int32)
    # first call of each stage: neq changes (the initial elastic solve is call 0)
    firsts = [0] + [i for i in range(1, len(calls)) if neqs[i] != neqs[i - 1]]
    if len(firsts) < 2 or firsts[1] != 1:
        firsts.insert(1, 1)
    out = dict(mesh_arrays(mesh))
    out.update(summary=summ, cg_log=np.array([c[1] for c in calls], np.int32),
               neq_log=neqs, x_initial=calls[0][4], f_initial=calls[0][3],
               stage_first_call=np.array(firsts[1:], np.int32),
               control=np.array([control["cg_tol"], control["cg_maxiter"],
                                 control["nl_tol"], control["nl_maxiter"]]),
               srf=np.array(control["srf"]),
               excav=np.array([excav["nexcav"], excav["ninc"]] + list(excav["excavid"]),
                              np.int32))
    for k, i in enumerate(firsts[1:]):
        out[f"stage{k + 1}_f"] = calls[i][3]
        out[f"stage{k + 1}_x"] = calls[i][4]
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    shutil.rmtree(work)
    print(name, "calls", len(calls), "summary", summ.tolist())


def excavation_example_case(workdir=None):
    """The reference's own example input/excavation_3d (3 stages removing
    material ids 2, 3, 4; 10 load increments), timing tap only."""
    own = workdir is None
    if own:
        workdir = tempfile.mkdtemp(prefix="golden_ex3d_")
        os.makedirs(os.path.join(workdir, "input"))
        os.makedirs(os.path.join(workdir, "output"))
        for f in os.listdir(REF_INPUT):
            if f.startswith("excavation_3d") and "single" not in f and \
                    "elastic" not in f and not f.endswith("psem"):
                shutil.copy(os.path.join(REF_INPUT, f), os.path.join(workdir, "input"))
        run_tap(os.path.join(workdir, "input", "excavation_3d.sem"), workdir, mode="time")
    sem = os.path.join(workdir, "input", "excavation_3d.sem")
    summ = parse_stage_summary(find_summary(workdir))
    times = np.loadtxt(os.path.join(workdir, "pcg_times.txt"), ndmin=2)
    mesh, ctl = read_reference_input(sem)
    out = dict(mesh_arrays(mesh))
    out.update(cg_log=times[:, 1].astype(np.int32), summary=summ,
               control=np.array([ctl["cg_tol"], ctl["cg_maxiter"], ctl["nl_tol"],
                                 ctl["nl_maxiter"]]), srf=np.array(ctl["srf"]),
               excav=np.array([ctl["nexcav"], ctl["ninc"]] + list(ctl["excavid"]),
                              np.int32))
    np.savez_compressed(os.path.join(HERE, "excavation_3d.npz"), **out)
    if own:
        shutil.rmtree(workdir)
    print("excavation_3d", summ.tolist())


def excavation_example_spread():
    """The same example run twice more with libm results moved by one ulp on
    about 1/64 of the arguments (tests/golden/ulp_shim.c preloaded): asin
    only, then asin, sin, cos, tan and sincos.  The stage rows of these runs
    measure how far the reference's own results move under last-bit changes
    of its transcendental functions; they are stored beside the unperturbed
    summary as spread_rows[variant, stage, (stage, cg_tot, nl_tot, uxmax,
    umax, fmax)]."""
    work = tempfile.mkdtemp(prefix="golden_ex3d_spread_")
    shim = os.path.join(work, "ulp_shim.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-o", shim,
                    os.path.join(HERE, "ulp_shim.c"), "-ldl", "-lm"], check=True)
    funcs = ["asin", "all"]
    jobs = []
    for v in funcs:  # both variants at once, each in its own directory
        wd = os.path.join(work, v)
        os.makedirs(os.path.join(wd, "input"))
        os.makedirs(os.path.join(wd, "output"))
        for f in os.listdir(REF_INPUT):
            if f.startswith("excavation_3d") and "single" not in f and \
                    "elastic" not in f and not f.endswith("psem"):
                shutil.copy(os.path.join(REF_INPUT, f), os.path.join(wd, "input"))
        env = dict(os.environ, SGEO_TAP_DIR=wd, SGEO_TAP_MODE="time",
                   LD_PRELOAD=shim, SGEO_ULP_FUNCS=v)
        sem = os.path.join(wd, "input", "excavation_3d.sem")
        jobs.append(subprocess.Popen(
            ["bash", "-c", f"ulimit -s unlimited; exec {REF_TAP} {sem}"], cwd=wd,
            env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
    for j in jobs:
        j.wait()
    rows = np.stack([parse_stage_summary(find_summary(os.path.join(work, v)))
                     for v in funcs])
    path = os.path.join(HERE, "excavation_3d.npz")
    with np.load(path) as z:
        out = {k: z[k] for k in z.files}
    out.update(spread_rows=rows, spread_funcs=np.array(funcs))
    np.savez_compressed(path, **out)
    shutil.rmtree(work)
    print("excavation_3d spread", rows.tolist())


def list_format_case():
    """The reference compiler's list-directed output (the summary file's
    number format, src/slope.F90:217-219,415): doubles of every magnitude
    and sign, values near the F/E switch, SRF-like and O(1) values, and
    summary rows (srf, cg_tot, nl_tot, uxmax, umax, fmax) whose length makes
    the runtime wrap its record."""
    rng = np.random.default_rng(11)
    special = [0.0, 1.0, -1.0, 0.5, 0.1, 0.099999, 0.0951, 0.0949, 1e-8, 5e-4, 1.2, 1.05,
               2.0, 1e14, 9.99e14, 1e15, 1e16, 123456789012345678.0, 0.125, 659.0,
               100000.0, 1e-300, 1.7976931348623157e308, 5e-324, 3.5001576573408467]
    vals = np.concatenate([
        np.array(special),
        10.0 ** rng.uniform(-20, 20, 1500) * rng.choice([-1.0, 1.0], 1500),
        np.round(rng.uniform(0.0, 3.0, 200), 2),
        rng.uniform(0.0, 1.0, 300),
        rng.uniform(0.09, 0.11, 200),
        10.0 ** rng.uniform(13.5, 16.5, 100)])
    m = 300
    srf = np.array([round(v, int(d)) for v, d in zip(rng.uniform(0.5, 3.0, m),
                                                      rng.integers(0, 3, m))])
    cg = rng.integers(1, 400000, m).astype(np.int32)
    nl = rng.integers(1, 5001, m).astype(np.int32)
    ux = 10.0 ** rng.uniform(-6, 1, m)
    um = ux * rng.uniform(1.0, 1.5, m)
    fm = np.where(rng.uniform(size=m) < 0.2, 0.0, 10.0 ** rng.uniform(-4, 2, m))
    work = tempfile.mkdtemp(prefix="golden_listfmt_")
    path = os.path.join(work, "in.bin")
    rec = np.zeros(m, dtype=[("srf", "<f8"), ("cg", "<i4"), ("nl", "<i4"), ("ux", "<f8"),
                             ("um", "<f8"), ("fm", "<f8")])
    rec["srf"], rec["cg"], rec["nl"], rec["ux"], rec["um"], rec["fm"] = srf, cg, nl, ux, um, fm
    with open(path, "wb") as fh:
        fh.write(np.int32(vals.size).tobytes())
        fh.write(vals.astype("<f8").tobytes())
        fh.write(np.int32(m).tobytes())
        fh.write(rec.tobytes())
    out = subprocess.run([LISTFMT_PROBE, path], capture_output=True, text=True,
                         check=True).stdout
    lines = out.split("\n")
    strings = np.array(lines[:vals.size])
    rows_text = "\n".join(lines[vals.size:])
    np.savez_compressed(os.path.join(HERE, "list_format.npz"), values=vals, strings=strings,
                        row_srf=srf, row_cg=cg, row_nl=nl, row_ux=ux, row_um=um, row_fm=fm,
                        rows_text=np.array(rows_text))
    shutil.rmtree(work)
    print("list_format", vals.size, "values,", m, "rows")


SAVE_ALL = ("disp", "stress", "psigma", "nsigma", "scf", "vmeps")


def output_case():
    """The reference's own output files for a two-SRF sweep of the 4x2x3
    slope with every result field it can be asked for (save: disp, stress,
    psigma, nsigma, scf, vmeps): the EnSight Gold case and geometry files
    of the hex8 mesh and of the GLL mesh, the field files per SRF, and the
    summary file (src/semgeotech.F90:138-322, src/postprocess.f90:128-248,
    src/visual.f90:17-271, src/slope.F90:214-220, 414-439)."""
    mesh = slope_mesh(4, 2, 3)
    work = tempfile.mkdtemp(prefix="golden_out_")
    sem = write_reference_input(mesh, work, "slope_out", cg_tol=1e-8, cg_maxiter=5000,
                                nl_tol=5e-4, nl_maxiter=300, srf=[1.0, 1.2],
                                save=SAVE_ALL)
    r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {REF_BIN} {sem}"],
                       cwd=work, capture_output=True, text=True)
    outdir = os.path.join(work, "output")
    names = sorted(os.listdir(outdir))
    if "slope_out_summary" not in names:
        raise RuntimeError(f"reference run failed: {r.stdout[-2000:]} {r.stderr}")
    out = dict(mesh_arrays(mesh))
    out["file_names"] = np.array(names)
    for f in names:
        out["file:" + f] = np.frombuffer(open(os.path.join(outdir, f), "rb").read(),
                                         dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "output_slope_4x2x3.npz"), **out)
    shutil.rmtree(work)
    print("output", names)


def excavation_overburden_case():
    """The small pit with the overburden initial stress (stress0: type=1,
    src/postprocess.f90:10-33; src/excavation.F90:238-241) instead of the
    gravity solve: the reference's stage rows and every pcg_solver call's
    cg_iter (tap, time mode)."""
    soil = Material(gam=20.0, ym=1e5, nu=0.3, phi=25.0, coh=15.0)
    mesh = pit_mesh(6, 4, 4, lx=12.0, ly=8.0, lz=8.0, pit=(3.0, 9.0, 2.0, 6.0),
                    nstage=2, stage_depth=2.0, soil=soil)
    work = tempfile.mkdtemp(prefix="golden_exob_")
    s0 = dict(z0=8.0, s0=-5.0, k0=0.5)
    sem = write_reference_input(
        mesh, work, "pit_ob", cg_tol=1e-8, cg_maxiter=5000, nl_tol=1e-5, nl_maxiter=100,
        srf=[1.0], extra_control=", nexcav=2, excavid=2 3, ninc=2",
        extra_lines=f"stress0: type=1, z0={s0['z0']!r}, s0={s0['s0']!r}, k0={s0['k0']!r}\n")
    run_tap(sem, work, mode="time")
    summ = parse_stage_summary(find_summary(work))
    times = np.loadtxt(os.path.join(work, "pcg_times.txt"), ndmin=2)
    out = dict(mesh_arrays(mesh))
    out.update(summary=summ, cg_log=times[:, 1].astype(np.int32),
               control=np.array([1e-8, 5000, 1e-5, 100]), srf=np.array([1.0]),
               excav=np.array([2, 2, 2, 3], np.int32),
               stress0=np.array([s0["z0"], s0["s0"], s0["k0"]]))
    np.savez_compressed(os.path.join(HERE, "pit_overburden_6x4x4.npz"), **out)
    shutil.rmtree(work)
    print("excavation_overburden", summ.tolist(), len(out["cg_log"]), "calls")


def output_excavation_case():
    """The reference's own output files for a small pit excavated in two
    stages with two load increments each, every result field saved: the
    in-situ step 0 and each stage's geometry of the intact mesh, its fields
    at the intact nodes, and the summary (src/excavation.F90:284-700,
    src/semgeotech.F90:125-266)."""
    soil = Material(gam=20.0, ym=1e5, nu=0.3, phi=25.0, coh=15.0)
    mesh = pit_mesh(6, 4, 4, lx=12.0, ly=8.0, lz=8.0, pit=(3.0, 9.0, 2.0, 6.0),
                    nstage=2, stage_depth=2.0, soil=soil)
    work = tempfile.mkdtemp(prefix="golden_exout_")
    sem = write_reference_input(mesh, work, "pit_out", cg_tol=1e-8, cg_maxiter=5000,
                                nl_tol=1e-5, nl_maxiter=100, srf=[1.0],
                                extra_control=", nexcav=2, excavid=2 3, ninc=2",
                                save=SAVE_ALL)
    r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {REF_BIN} {sem}"],
                       cwd=work, capture_output=True, text=True)
    outdir = os.path.join(work, "output")
    names = sorted(os.listdir(outdir))
    if "pit_out_summary" not in names:
        raise RuntimeError(f"reference run failed: {r.stdout[-2000:]} {r.stderr}")
    out = dict(mesh_arrays(mesh))
    out["file_names"] = np.array(names)
    out["excav"] = np.array([2, 2, 2, 3], np.int32)  # nexcav, ninc, excavid
    for f in names:
        out["file:" + f] = np.frombuffer(open(os.path.join(outdir, f), "rb").read(),
                                         dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "output_pit_6x4x4.npz"), **out)
    shutil.rmtree(work)
    print("output_excavation", names)


def mesh_arrays(m):
    gam, ym, nu, phi, coh, psi = m.material_arrays()
    return dict(coord=m.coord, conn=m.conn, mat_id=m.mat_id,
                materials=np.stack([gam, ym, nu, phi, coh, psi], 1),
                bc_ux=m.bc["ux"], bc_uy=m.bc["uy"], bc_uz=m.bc["uz"])


def slope_case(name, mesh, control, keep_calls):
    work = tempfile.mkdtemp(prefix=f"golden_{name}_")
    sem = write_reference_input(mesh, work, name, **control)
    run_tap(sem, work)
    neq, k, dinv, g = read_static(work)
    cg, uin, f, uout = read_calls(work, neq)
    moh = np.fromfile(os.path.join(work, "mohcouf.bin"), dtype=np.float64)
    summ = parse_summary(find_summary(work))
    out = dict(mesh_arrays(mesh))
    out.update(
        neq=np.int32(neq), gdof_elmt=g, dinv=dinv,
        k_sha256=np.frombuffer(hashlib.sha256(np.ascontiguousarray(k).tobytes())
                               .digest(), dtype=np.uint8),
        k_elem0=k[0].copy(), extload=f[0].copy(), cg_log=cg, summary=summ,
        call_idx=np.array(keep_calls, np.int32),
        call_uin=uin[keep_calls], call_f=f[keep_calls], call_uout=uout[keep_calls],
        # f of every GLL point at the final body pass of the last SRF
        last_pass_f=moh[-27 * mesh.nelmt:].copy(),
        # ... and at the final body pass of every SRF (row s)
        final_pass_f=np.stack([moh[(p - 1) * 27 * mesh.nelmt:p * 27 * mesh.nelmt]
                               for p in np.cumsum(summ[:, 2].astype(int))]),
        control=np.array([control["cg_tol"], control["cg_maxiter"], control["nl_tol"],
                          control["nl_maxiter"]]),
        srf=np.array(control["srf"]))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    shutil.rmtree(work)
    print(name, "neq", neq, "calls", len(cg), "summary", summ.tolist())


def _chunk_run(mesh, control, srfs, work, name):
    """One reference run over the SRF list `srfs` (tap, full mode) in its own
    directory; returns the process."""
    os.makedirs(work, exist_ok=True)
    sem = write_reference_input(mesh, work, name, **dict(control, srf=srfs))
    env = dict(os.environ, SGEO_TAP_DIR=work, SGEO_TAP_MODE="full")
    return subprocess.Popen(["bash", "-c", f"ulimit -s unlimited; exec {REF_TAP} {sem}"],
                            cwd=work, env=env, stdout=subprocess.DEVNULL,
                            stderr=subprocess.DEVNULL)


def sweep_case(name, mesh, control, chunk=7):
    """A full SRF sweep of `mesh` through the reference, run as independent
    chunks of `chunk` SRFs in parallel: the reference's parser keeps 80
    characters of a value list (src/string_library.f90:547), and every SRF
    trial restarts from zero state (src/slope.F90:248-249, 281-282), so the
    chunks' rows concatenate to the one-run sweep, truncated at the first
    failing SRF (src/slope.F90:441).  Stored: the summary rows, every PCG
    call's cg_iter, the final displacement x of every SRF (the last PCG
    call's result), the plastic-point flags (f >= 0) of every SRF's final
    body pass, packed 8 to a byte, and K's sha256, dinv and extload."""
    srf = list(control["srf"])
    chunks = [srf[i:i + chunk] for i in range(0, len(srf), chunk)]
    base = tempfile.mkdtemp(prefix=f"golden_{name}_")
    procs = [_chunk_run(mesh, control, c, os.path.join(base, f"c{i}"), name)
             for i, c in enumerate(chunks)]
    for p in procs:
        p.wait()
    rows, cg_log, xs, flags = [], [], [], []
    neq = k_sha = dinv = ext = None
    npt = 27 * mesh.nelmt
    for i, c in enumerate(chunks):
        work = os.path.join(base, f"c{i}")
        summ = parse_summary(find_summary(work))
        if i == 0:
            neq, k, dinv, _ = read_static(work)
            k_sha = hashlib.sha256(np.ascontiguousarray(k).tobytes()).digest()
            del k
        n1 = neq + 1
        rec = 16 + 24 * n1
        calls = np.memmap(os.path.join(work, "pcg_calls.bin"), dtype=np.uint8, mode="r")
        ncall = calls.size // rec
        cgs = np.array([calls[q * rec:q * rec + 16].view(np.int32)[2]
                        for q in range(ncall)], np.int32)
        if i == 0:
            ext = calls[16 + 8 * n1:16 + 16 * n1].view(np.float64).copy()
        moh = np.memmap(os.path.join(work, "mohcouf.bin"), dtype=np.float64, mode="r")
        ends = np.cumsum(summ[:, 2].astype(int))
        assert ends[-1] == ncall and moh.size == ncall * npt, (ends, ncall, moh.size)
        for s, e in enumerate(ends):
            q = e - 1
            xs.append(calls[q * rec + 16 + 16 * n1:(q + 1) * rec].view(np.float64).copy())
            flags.append(np.packbits(np.asarray(moh[q * npt:(q + 1) * npt]) >= 0.0))
        rows.append(summ)
        cg_log.append(cgs)
        del calls, moh
        if summ[-1, 2] == control["nl_maxiter"]:
            break  # the sweep ends at its first failing SRF
    shutil.rmtree(base)
    summ = np.concatenate(rows)
    out = dict(mesh_arrays(mesh))
    out.update(
        neq=np.int32(neq), dinv=dinv, extload=ext,
        k_sha256=np.frombuffer(k_sha, dtype=np.uint8),
        summary=summ, cg_log=np.concatenate(cg_log), x_final=np.stack(xs),
        flags_final=np.stack(flags),
        control=np.array([control["cg_tol"], control["cg_maxiter"], control["nl_tol"],
                          control["nl_maxiter"]]),
        srf=np.array(srf))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(name, "neq", neq, "calls", out["cg_log"].size, "summary", summ.tolist())


def c0_case():
    """The 1,152-element homogeneous slope of the CPU-runnable configuration
    (24x4x12, c=10 kPa, phi=20 deg), its full SRF sweep 1.00:0.05:2.00."""
    from specfem3d_geotech_amd import workloads
    w = workloads.get("c0")
    sweep_case("c0_sweep", w.make(), dict(w.control))


def c0_spread_case():
    """How far the reference's own c0 displacements move when its libm moves
    by one ulp on about 1/64 of the arguments (tests/golden/ulp_shim.c
    preloaded: asin only, then asin, sin, cos, tan and sincos): the converged
    SRFs of the c0 sweep rerun under each variant, and the relative L2
    distance of each SRF's final x from the unperturbed one stored beside the
    c0 golden as x_spread[variant, srf] (variants: asin, or all five, moved on
    1/64 of the arguments; all five moved on half of them).  A device math library differs from
    the host libm in the last bit on some arguments, and the viscoplastic
    iteration carries such differences into x: this is the reference's own
    sensitivity to them."""
    from specfem3d_geotech_amd import workloads
    path = os.path.join(HERE, "c0_sweep.npz")
    with np.load(path) as z:
        out = {k: z[k] for k in z.files}
    ctl = dict(workloads.get("c0").control)
    summ = out["summary"]
    conv = [float(s) for s, nl in zip(summ[:, 0], summ[:, 2]) if nl < ctl["nl_maxiter"]]
    ctl["srf"] = conv
    mesh = workloads.get("c0").make()
    base = tempfile.mkdtemp(prefix="golden_c0spread_")
    shim = os.path.join(base, "ulp_shim.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-o", shim,
                    os.path.join(HERE, "ulp_shim.c"), "-ldl", "-lm"], check=True)
    # asin or all five functions moved on 1/64 of the arguments, and all
    # five on half of them (a libm whose last bits are independent of the
    # host's)
    variants = [("asin", 64), ("all", 64), ("all", 2)]
    funcs = [f"{v}/{n}" for v, n in variants]
    procs = []
    for (v, every), name in zip(variants, funcs):
        work = os.path.join(base, name.replace("/", "_"))
        os.makedirs(work)
        sem = write_reference_input(mesh, work, "c0s", **ctl)
        env = dict(os.environ, SGEO_TAP_DIR=work, SGEO_TAP_MODE="full", LD_PRELOAD=shim,
                   SGEO_ULP_FUNCS=v, SGEO_ULP_EVERY=str(every))
        procs.append(subprocess.Popen(
            ["bash", "-c", f"ulimit -s unlimited; exec {REF_TAP} {sem}"], cwd=work, env=env,
            stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
    for p in procs:
        p.wait()
    n1 = int(out["neq"]) + 1
    rec = 16 + 24 * n1
    spread = np.zeros((len(funcs), len(conv)))
    rows = []
    for i, v in enumerate(funcs):
        work = os.path.join(base, v.replace("/", "_"))
        s = parse_summary(find_summary(work))
        rows.append(s)
        calls = np.memmap(os.path.join(work, "pcg_calls.bin"), dtype=np.uint8, mode="r")
        for k, e in enumerate(np.cumsum(s[:, 2].astype(int))):
            q = e - 1
            x = calls[q * rec + 16 + 16 * n1:(q + 1) * rec].view(np.float64)
            xr = out["x_final"][k]
            spread[i, k] = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        del calls
    shutil.rmtree(base)
    out.update(x_spread=spread, spread_rows=np.stack(rows), spread_funcs=np.array(funcs))
    np.savez_compressed(path, **out)
    print("c0 spread", spread.tolist())


def c1_like_case():
    """A reduced mesh with the 144,000-element configuration's features: the
    2H:1V slope at 30x5x15 elements, two layers split at z = 10 m (upper
    c=8 kPa, phi=25 deg; lower c=20 kPa, phi=30 deg), interior nodes
    jittered by U[-0.02,0.02] x local spacing (seed 1); full SRF sweep."""
    from specfem3d_geotech_amd import workloads
    w = workloads.get("c1")
    sweep_case("c1_like_sweep", workloads.c1_like(30, 5, 15), dict(w.control))


def form_a_case(n=64, jobs=6):
    """Form-A strength realisations (lognormal c and tan phi, workloads.form_a,
    seed 3) of the layered 8x4x6 slope, each swept over SRF 1.0:0.2:2.0 with
    nl_maxiter 60 by the reference: the summary rows of every realisation
    (rows of realisation i at offsets[i]:offsets[i+1])."""
    from specfem3d_geotech_amd import workloads
    layered = slope_mesh(8, 4, 6, materials=[
        Material(gam=20.0, ym=1e5, nu=0.3, phi=25.0, coh=8.0),
        Material(gam=20.0, ym=1e5, nu=0.3, phi=30.0, coh=20.0)],
        layer_z=[10.0], jitter=0.02, seed=1)
    ctl = dict(cg_tol=1e-8, cg_maxiter=5000, nl_tol=5e-4, nl_maxiter=60,
               srf=[1.0, 1.2, 1.4, 1.6, 1.8, 2.0])
    real = workloads.form_a(n, len(layered.materials))
    base = tempfile.mkdtemp(prefix="golden_forma_")
    pending, rows = list(range(n)), {}
    running = []
    while pending or running:
        while pending and len(running) < jobs:
            i = pending.pop(0)
            c, p = real[i]
            work = os.path.join(base, f"r{i}")
            sem = write_reference_input(layered.with_strengths(c, p), work, "fa", **ctl)
            env = dict(os.environ, SGEO_TAP_DIR=work, SGEO_TAP_MODE="time")
            running.append((i, work, subprocess.Popen(
                ["bash", "-c", f"ulimit -s unlimited; exec {REF_TAP} {sem}"], cwd=work,
                env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)))
        i, work, pr = running.pop(0)
        pr.wait()
        rows[i] = parse_summary(find_summary(work))
    shutil.rmtree(base)
    offs = np.cumsum([0] + [rows[i].shape[0] for i in range(n)])
    out = dict(mesh_arrays(layered))
    out.update(coh=np.array([r[0][0] for r in real]), phi=np.array([r[1][0] for r in real]),
               rows=np.concatenate([rows[i] for i in range(n)]), offsets=offs,
               control=np.array([ctl["cg_tol"], ctl["cg_maxiter"], ctl["nl_tol"],
                                 ctl["nl_maxiter"]]), srf=np.array(ctl["srf"]))
    np.savez_compressed(os.path.join(HERE, "form_a_8x4x6.npz"), **out)
    print("form_a", n, "realisations;", out["rows"].shape[0], "rows")


def validation1_case():
    """The reference's own example, SRF 1.0 only (the first of its list)."""
    work = tempfile.mkdtemp(prefix="golden_v1_")
    os.makedirs(os.path.join(work, "input"))
    os.makedirs(os.path.join(work, "output"))
    for f in os.listdir(REF_INPUT):
        if f.startswith("validation1") and not f.endswith(".sem") and \
                "phinu" not in f:
            shutil.copy(os.path.join(REF_INPUT, f), os.path.join(work, "input"))
    src = open(os.path.join(REF_INPUT, "validation1.sem")).read()
    src = src.replace("nsrf=9, srf=1.0 1.5 2.0 2.15 2.16 2.17 2.18 2.19 2.20",
                      "nsrf=1, srf=1.0")
    sem = os.path.join(work, "input", "validation1.sem")
    open(sem, "w").write(src)
    run_tap(sem, work, mode="time")
    summ = parse_summary(find_summary(work))
    times = np.loadtxt(os.path.join(work, "pcg_times.txt"), ndmin=2)
    mesh, ctl = read_reference_input(sem)
    out = dict(mesh_arrays(mesh))
    out.update(cg_log=times[:, 1].astype(np.int32), summary=summ,
               control=np.array([ctl["cg_tol"], ctl["cg_maxiter"], ctl["nl_tol"],
                                 ctl["nl_maxiter"]]), srf=np.array(ctl["srf"]))
    np.savez_compressed(os.path.join(HERE, "validation1_srf1.npz"), **out)
    shutil.rmtree(work)
    print("validation1", summ.tolist())


def validation1_phinu_case():
    """The reference's own example input/validation1_phinu (phinu=1: the
    Poisson's ratio is corrected whenever sin(phi_f) < 1 - 2 nu, and the
    stiffness rebuilt, src/plastic_library.f90:72-86, src/slope.F90:259-270)
    at SRF 1.0 and 1.05, each of which sets its own nu; timing tap only."""
    work = tempfile.mkdtemp(prefix="golden_v1p_")
    os.makedirs(os.path.join(work, "input"))
    os.makedirs(os.path.join(work, "output"))
    for f in os.listdir(REF_INPUT):
        if f.startswith("validation1") and not f.endswith(".sem") and \
                not f.endswith(".psem"):
            shutil.copy(os.path.join(REF_INPUT, f), os.path.join(work, "input"))
    src = open(os.path.join(REF_INPUT, "validation1_phinu.sem")).read()
    old = "nsrf=9, srf=1.0 1.5 2.0 2.15 2.16 2.17 2.18 2.19 2.20, phinu=1"
    assert old in src
    src = src.replace(old, "nsrf=2, srf=1.0 1.05, phinu=1")
    sem = os.path.join(work, "input", "validation1_phinu.sem")
    open(sem, "w").write(src)
    run_tap(sem, work, mode="time")
    summ = parse_summary(find_summary(work))
    times = np.loadtxt(os.path.join(work, "pcg_times.txt"), ndmin=2)
    mesh, ctl = read_reference_input(sem)
    assert ctl.get("phinu"), ctl
    out = dict(mesh_arrays(mesh))
    out.update(cg_log=times[:, 1].astype(np.int32), summary=summ,
               control=np.array([ctl["cg_tol"], ctl["cg_maxiter"], ctl["nl_tol"],
                                 ctl["nl_maxiter"]]), srf=np.array(ctl["srf"]),
               phinu=np.int32(1))
    np.savez_compressed(os.path.join(HERE, "validation1_phinu.npz"), **out)
    shutil.rmtree(work)
    print("validation1_phinu", summ.tolist())


def gll_case():
    txt = subprocess.run([GLL_PROBE], capture_output=True, text=True, check=True).stdout
    vals = [float(t) for t in txt.split() if not t.startswith("#") and
            t[0] in "-0123456789" and ("E" in t or "e" in t)]
    v = np.array(vals)
    np.savez_compressed(os.path.join(HERE, "gll_tables.npz"),
                        zwgljd=v[:6].reshape(3, 2), gll_weights=v[6:33],
                        dlagrange_gll=v[33:33 + 2187], dshape_hex8=v[2220:2220 + 648])
    print("gll tables", v.size)


def fn_case():
    rng = np.random.default_rng(5)
    n = 600
    S = rng.normal(0, 100, (n, 6))
    S[:, :3] -= 150
    S[0] = [-100, -100, -100, 0, 0, 0]      # hydrostatic
    S[1] = [-100, -200, -200, 0, 0, 0]      # triaxial extension
    S[2] = [-200, -100, -100, 0, 0, 0]      # triaxial compression
    P = np.stack([rng.uniform(0, 45, n), rng.uniform(0, 50, n),
                  rng.uniform(0, 20, n)], 1)
    rec = np.concatenate([S, P], 1)
    work = tempfile.mkdtemp(prefix="golden_fn_")
    fin, fout = os.path.join(work, "in.bin"), os.path.join(work, "out.bin")
    with open(fin, "wb") as fh:
        np.array([n], np.int32).tofile(fh)
        rec.astype(np.float64).tofile(fh)
    subprocess.run([FN_PROBE, fin, fout], check=True)
    out = np.fromfile(fout, dtype=np.float64)
    A = out[:n * 115].reshape(n, 115)
    B = out[n * 115:].reshape(-1, 6)
    np.savez_compressed(os.path.join(HERE, "constitutive.npz"), inputs=rec,
                        outputs=A, srf_outputs=B)
    shutil.rmtree(work)
    print("constitutive", A.shape, B.shape)


def excavation_cases():
    soil = Material(gam=20.0, ym=1e5, nu=0.3, phi=25.0, coh=15.0)
    pit = pit_mesh(8, 6, 6, lx=16.0, ly=12.0, lz=12.0, pit=(4.0, 12.0, 3.0, 9.0),
                   nstage=3, stage_depth=2.0, soil=soil)
    excavation_case("pit_8x6x6", pit,
                    dict(cg_tol=1e-8, cg_maxiter=5000, nl_tol=1e-5, nl_maxiter=100,
                         srf=[1.0]),
                    dict(nexcav=3, excavid=[2, 3, 4], ninc=2))


def main():
    base = dict(cg_tol=1e-8, cg_maxiter=5000, nl_tol=5e-4, nl_maxiter=300)
    slope_case("slope_4x2x3", slope_mesh(4, 2, 3),
               dict(base, srf=[1.0, 1.5]), keep_calls=list(range(12)) + [20, 100, 310])
    layered = slope_mesh(8, 4, 6, materials=[
        Material(gam=20.0, ym=1e5, nu=0.3, phi=25.0, coh=8.0),
        Material(gam=20.0, ym=1e5, nu=0.3, phi=30.0, coh=20.0)],
        layer_z=[10.0], jitter=0.02, seed=1)
    slope_case("slope_8x4x6_layered", layered,
               dict(base, nl_maxiter=100, srf=[1.0, 1.2, 1.4, 1.6, 1.8, 2.0]),
               keep_calls=[0, 1, 2])
    validation1_case()
    gll_case()
    fn_case()
    excavation_cases()
    excavation_example_case()
    excavation_example_spread()
    list_format_case()
    output_case()


if __name__ == "__main__":
    if len(sys.argv) > 1:  # single cases: excavation | excavation_example [workdir]
        {"excavation": excavation_cases,       # | excavation_spread
         "excavation_example": lambda: excavation_example_case(
             sys.argv[2] if len(sys.argv) > 2 else None),
         "excavation_spread": excavation_example_spread,
         "list_format": list_format_case, "output": output_case,
         "output_excavation": output_excavation_case,
         "excavation_overburden": excavation_overburden_case,
         "c0": c0_case, "c0_spread": c0_spread_case, "c1_like": c1_like_case,
         "form_a": form_a_case,
         "validation1_phinu": validation1_phinu_case}[sys.argv[1]]()
    else:
        main()
