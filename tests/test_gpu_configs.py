"""The hot path at the benchmark configurations' scale, against the reference.

* c0, the 1,152-element homogeneous slope (24x4x12), its full SRF sweep
  1.00:0.05:2.00 (golden from the compiled reference, tests/golden/
  c0_sweep.npz): the FP64 drop-in with reference-order dots gives every
  PCG call's cg_iter, every converged SRF's row and final displacement and
  plastic-point set; its failing SRF the plastic iteration count, so the
  FOS bracket (1.20, 1.25).
* A reduced mesh with c1's features (two layers, 2% jitter; 30x5x15, 2,250
  elements): its full sweep on the FP64 and the 64-column mixed form.
* The reference's own example input/validation1 at SRF 1.0 on the GPU.
* The 64-column batched engine (k_bmv_mfma<4>, k_bassemble_dot<true>)
  against the reference: a c0 sweep with every SRF in flight, and 64 form-A
  strength realisations of the layered 8x4x6 slope (one per column, golden
  brackets from the compiled reference).
* A trial's bits do not depend on the batch width: one column alone and the
  same trial among 63 others (1-column vs 64-column MFMA product tiles).
* The full 144,000-element c1 mesh: the device stiffness equals the
  oracle's bit for bit (sha256 of 7.6 GB), and 40 CG iterations of its
  elastic solve with reference-order dots equal the oracle's.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle as O
from conftest import control_from_golden, golden, mesh_from_golden
from specfem3d_geotech_amd.batch import (Analysis, BatchEngine, device_mesh,
                                         run_analyses, slope3d_mixed)
from specfem3d_geotech_amd.slope import slope3d
from specfem3d_geotech_amd.solver import (REDUCE_SEQUENTIAL, REDUCE_TREE, Context,
                                          DeviceMesh, pcg_solver)
from specfem3d_geotech_amd.spec import hex2spec

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def bracket_of(summary, nl_maxiter):
    if summary[-1, 2] != nl_maxiter:
        return None
    return (float(summary[-2, 0]) if summary.shape[0] > 1 else None, float(summary[-1, 0]))


@pytest.fixture(scope="module")
def ctx_seq():
    c = Context(0)
    c.set_reduction(REDUCE_SEQUENTIAL)
    yield c
    c.close()


def test_c0_fp64_sweep_matches_reference(ctx, ctx_seq):
    """Converged SRFs with reference-order dots: the plastic iteration count
    and the displacement to 1e-10 relative L2 -- or, where larger, twice the
    reference's own spread when its libm moves by one ulp on a fraction of
    the arguments (x_spread: the device sin/cos/tan/asin differ from the
    host's in the last bit, and a long viscoplastic chain carries that into
    x).  Where no such libm change moves the reference's cg_tot, every PCG
    call's cg_iter equals the reference's; where it does, cg_tot stays within
    twice the reference's own movement.  The failing SRF (300 plastic
    iterations of a non-converging flow) runs on the fast path: its nl_tot,
    hence the bracket (1.20, 1.25)."""
    z = golden("c0_sweep")
    mesh = mesh_from_golden(z)
    ctl = control_from_golden(z)
    ref = z["summary"]
    nconv = int(np.count_nonzero(ref[:, 2] < ctl["nl_maxiter"]))
    got = {}
    res = slope3d(ctx_seq, mesh, on_srf=lambda row, x, fl: got.__setitem__(
        row["srf"], (x.copy(), fl.copy())), **dict(ctl, srf=list(ref[:nconv, 0])))
    assert res.fos_bracket is None and len(res.rows) == nconv
    dcg = np.abs(z["spread_rows"][:, :, 1] - ref[None, :nconv, 1]).max(axis=0)
    bad = []
    ncall = 0
    for s, (r, rr) in enumerate(zip(res.rows, ref)):
        n = int(rr[2])
        x, fl = got[r["srf"]]
        e = rel_l2(x, z["x_final"][s])
        bar = max(1e-10, 2.0 * float(z["x_spread"][:, s].max()))
        calls_equal = res.cg_log[ncall:ncall + n] == list(z["cg_log"][ncall:ncall + n])
        ncall += n
        print(f"c0 SRF {rr[0]}: nl {r['nl_tot']}/{int(rr[2])}, cg_tot {r['cg_tot']}/"
              f"{int(rr[1])} (reference's own libm spread {int(dcg[s])}), every call "
              f"equal {calls_equal}; x vs reference {e:.3e} (bar {bar:.3e})", flush=True)
        if r["srf"] != rr[0] or r["nl_tot"] != n:
            bad.append((rr[0], "nl_tot"))
        if dcg[s] == 0 and (not calls_equal or r["cg_tot"] != rr[1]):
            bad.append((rr[0], "cg_iter"))
        if abs(r["cg_tot"] - rr[1]) > 2 * dcg[s]:
            bad.append((rr[0], "cg_tot"))
        if e > bar:
            bad.append((rr[0], "x"))
    # the failing SRF, fast reductions
    fail = slope3d(ctx, mesh, **dict(ctl, srf=[float(ref[nconv, 0])]))
    print("c0 failing SRF", fail.rows[0], ref[nconv].tolist())
    assert not bad, bad
    assert fail.rows[0]["nl_tot"] == ctl["nl_maxiter"]
    assert bracket_of(ref, ctl["nl_maxiter"]) == (float(ref[nconv - 1, 0]),
                                                  float(ref[nconv, 0]))


def test_c1_like_sweeps_match_reference(ctx, ctx_seq):
    """The c1 features (two layers, 2% jitter) at 2,250 elements (30x5x15),
    its full sweep (golden c1_like_sweep, bracket (1.55, 1.60)): the FP64
    drop-in on the fast path gives every SRF's nl_tot, so the bracket, and
    cg_tot within 1%; with reference-order dots its first three SRFs give
    every cg_iter and row; the 64-column mixed form gives the bracket and
    every converged SRF's nl_tot."""
    z = golden("c1_like_sweep")
    mesh = mesh_from_golden(z)
    ctl = control_from_golden(z)
    ref = z["summary"]
    br = bracket_of(ref, ctl["nl_maxiter"])
    fast = slope3d(ctx, mesh, **ctl)
    assert fast.fos_bracket == br and len(fast.rows) == ref.shape[0]
    for r, rr in zip(fast.rows, ref):
        print("c1-like fp64", r, rr.tolist(), flush=True)
        assert r["nl_tot"] == rr[2]
        if rr[2] < ctl["nl_maxiter"]:
            assert abs(r["cg_tot"] - rr[1]) <= 0.01 * rr[1]
            assert abs(r["uxmax"] - rr[3]) <= 1e-8 * rr[3]
    seq = slope3d(ctx_seq, mesh, **dict(ctl, srf=list(ref[:3, 0])))
    ncall = int(ref[:3, 2].sum())
    assert seq.cg_log == list(z["cg_log"][:ncall])
    for s, (r, rr) in enumerate(zip(seq.rows, ref)):
        assert r["cg_tot"] == rr[1] and r["nl_tot"] == rr[2]
        assert abs(r["uxmax"] - rr[3]) <= 1e-10 * rr[3]
    mix = slope3d_mixed(ctx, mesh, ncol=64, **ctl)
    print("c1-like mixed", mix.rows, mix.extra["stats"])
    assert mix.fos_bracket == br
    for r, rr in zip(mix.rows, ref):
        if rr[2] < ctl["nl_maxiter"]:
            assert r["nl_tot"] == rr[2]


def test_validation1_srf1_on_gpu(ctx_seq):
    """The reference's own example (input/validation1, 1,670 elements) at
    SRF 1.0: every cg_iter and the summary row."""
    z = golden("validation1_srf1")
    mesh = mesh_from_golden(z)
    ctl = control_from_golden(z)
    res = slope3d(ctx_seq, mesh, **ctl)
    r, rr = res.rows[0], z["summary"][0]
    print("validation1", r, rr.tolist())
    assert res.cg_log == list(z["cg_log"])
    assert r["nl_tot"] == rr[2] and r["cg_tot"] == rr[1]
    assert abs(r["uxmax"] - rr[3]) <= 1e-10 * rr[3]
    assert abs(r["umax"] - rr[4]) <= 1e-10 * rr[4]
    assert abs(r["fmax"] - rr[5]) <= 1e-8 * max(1.0, abs(rr[5]))


def test_c0_mixed_sweep_64_columns(ctx):
    """The mixed-precision sweep with 64 columns, every SRF of the list in
    flight at once (the B = 64 MFMA product and assembly): the reference's
    bracket, its plastic iteration count at every converged SRF, and at
    every converged SRF a displacement that differs from the reference's by
    less than the reference's own distance from the exactly solved problem
    (each of its PCG solves stops at a relative update of 1e-8, and the
    viscoplastic loop carries that truncation into x)."""
    z = golden("c0_sweep")
    mesh = mesh_from_golden(z)
    ctl = control_from_golden(z)
    ref = z["summary"]
    got = {}
    res = slope3d_mixed(ctx, mesh, ncol=64, on_srf=lambda row, x, fl: got.__setitem__(
        row["srf"], x.copy()), **ctl)
    print("stats", res.extra["stats"])
    assert res.fos_bracket == bracket_of(ref, ctl["nl_maxiter"])
    assert len(res.rows) == ref.shape[0]
    conv = [float(s) for s, nl in zip(ref[:, 0], ref[:, 2]) if nl < ctl["nl_maxiter"]]
    # the exactly solved problem: the FP64 drop-in with cg_tol = 1e-13
    tight = {}
    slope3d(ctx, mesh, on_srf=lambda row, x, fl: tight.__setitem__(row["srf"], x.copy()),
            **dict(ctl, srf=conv, cg_tol=1e-13))
    for s, (r, rr) in enumerate(zip(res.rows, ref)):
        print("c0 mixed", r, rr.tolist(), flush=True)
        assert r["srf"] == rr[0] and r["nl_tot"] == rr[2]
        if rr[2] < ctl["nl_maxiter"]:
            e = rel_l2(got[r["srf"]], z["x_final"][s])
            e_ref = rel_l2(z["x_final"][s], tight[r["srf"]])
            print(f"  x vs reference {e:.3e}; reference vs exact solves {e_ref:.3e}")
            assert e <= max(1e-8, e_ref)


def test_form_a_64_realisations_match_reference(ctx):
    """64 form-A strength realisations of the layered 8x4x6 slope, one per
    column of a 64-column engine: each realisation's FOS bracket and
    per-SRF plastic iteration counts are the compiled reference's."""
    z = golden("form_a_8x4x6")
    mesh = mesh_from_golden(z)
    ctl = control_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    n = z["coh"].size
    an = [Analysis(coh=np.full_like(coh, z["coh"][i]), phi=np.full_like(phi, z["phi"][i]),
                   psi=psi.copy(), srf=list(ctl["srf"]), tag=i) for i in range(n)]
    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        eng = BatchEngine(dm, ext, 64)
        try:
            run_analyses(eng, an, ym=ym, nu=nu, cg_tol=ctl["cg_tol"],
                         cg_maxiter=ctl["cg_maxiter"], nl_tol=ctl["nl_tol"],
                         nl_maxiter=ctl["nl_maxiter"])
            print("stats", eng.stats())
        finally:
            eng.close()
    finally:
        dm.close()
    off = z["offsets"]
    same_nl = total = 0
    for i, a in enumerate(an):
        rows = z["rows"][off[i]:off[i + 1]]
        br = bracket_of(rows, ctl["nl_maxiter"])
        assert a.bracket == br, (i, a.bracket, br)
        for k, rr in enumerate(rows):
            if rr[2] < ctl["nl_maxiter"]:
                total += 1
                same_nl += a.rows[k]["nl_tot"] == rr[2]
                assert abs(a.rows[k]["nl_tot"] - rr[2]) <= 1
    print(f"{same_nl} of {total} converged SRFs with the reference's nl_tot")


def test_trial_bits_same_in_1_and_64_columns(ctx):
    """A trial run alone on a 1-column engine and the same trial on a
    64-column engine among 40 other analyses that start at other times:
    identical rows, displacement and plastic-flag bits."""
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    kw = dict(ym=ym, nu=nu, cg_tol=ctl["cg_tol"], cg_maxiter=ctl["cg_maxiter"],
              nl_tol=ctl["nl_tol"], nl_maxiter=40, keep_state=True)
    srf = [1.0, 1.4, 1.8]
    rng = np.random.default_rng(5)

    def analyses():
        out = [Analysis(coh=coh * f, phi=phi * g, psi=psi.copy(), srf=list(srf), tag=i)
               for i, (f, g) in enumerate(zip(rng.uniform(0.8, 1.2, 41),
                                              rng.uniform(0.9, 1.1, 41)))]
        return out
    mix = analyses()
    rng = np.random.default_rng(5)
    alone = analyses()[17]
    dm, ext = device_mesh(ctx, mesh, spec)
    try:
        eng = BatchEngine(dm, ext, 1)
        try:
            run_analyses(eng, [alone], **kw)
        finally:
            eng.close()
        eng = BatchEngine(dm, ext, 64)
        try:
            run_analyses(eng, mix, srf_parallel=2, **kw)
        finally:
            eng.close()
    finally:
        dm.close()
    got = mix[17]
    print("alone", alone.rows, "\nmix  ", got.rows)
    assert got.rows == alone.rows and got.bracket == alone.bracket
    assert np.array_equal(got.x_last, alone.x_last)
    assert np.array_equal(got.flags_last, alone.flags_last)


def test_c1_full_size_stiffness_and_capped_pcg(ctx_seq):
    """The 144,000-element c1 mesh (3,489,480 equations): the device K equals
    the oracle's (which equals the reference's stiffness_bodyload on every
    golden mesh) bit for bit, and 40 unconverged CG iterations of the elastic
    solve with reference-order dots give the oracle's iterate bit for bit."""
    from specfem3d_geotech_amd import workloads
    os.environ.setdefault("OMP_NUM_THREADS", "16")
    mesh = workloads.get("c1").make()
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    km, dp, ext = O.stiffness_bodyload(spec, gam, nu, ym)
    sha_o = hashlib.sha256(km.tobytes()).digest()
    dinv = np.zeros_like(dp)
    dinv[1:] = 1.0 / dp[1:]
    dm = DeviceMesh(ctx_seq, spec, len(gam))
    try:
        ext_d, dinv_d = dm.build_stiffness(gam, ym, nu)
        k = dm.download_stiffness()
        assert hashlib.sha256(k.tobytes()).digest() == sha_o
        del k
        assert np.array_equal(dinv_d, dinv) and np.array_equal(ext_d, ext)
    finally:
        dm.close()
    u0 = np.zeros(spec.neq + 1)
    u_g, it_g, ec_g, _ = pcg_solver(ctx_seq, spec.neq, spec.nelmt, km, u0, ext, dinv,
                                    spec.gdof_elmt, 1e-300, 40)
    u_o, it_o, ec_o = O.pcg_solver(spec.neq, km, u0, ext, dinv, spec.gdof_elmt, 1e-300, 40)
    print(f"c1: neq {spec.neq}, cg_iter {it_g} / {it_o}, "
          f"bitwise {np.array_equal(u_g, u_o)}, rel {rel_l2(u_g, u_o):.3e}")
    assert (it_g, ec_g) == (it_o, ec_o) == (41, -1)
    assert np.array_equal(u_g, u_o)


def test_validation1_phinu_rebuilds_stiffness_per_srf(ctx_seq):
    """The reference's own input/validation1_phinu (phinu=1) at SRF 1.0 and
    1.05: each corrects the Poisson's ratio to its own value (sin(phi_f) <
    1 - 2 nu, src/plastic_library.f90:72-86), so the stiffness and its
    diagonal are rebuilt on the device before each SRF (src/slope.F90:
    259-270).  Every PCG call's cg_iter and both summary rows match the
    reference; the mixed form's entry point routes such a sweep to the FP64
    drop-in and gives the same rows."""
    from specfem3d_geotech_amd.plastic import strength_reduction
    z = golden("validation1_phinu")
    mesh = mesh_from_golden(z)
    ctl = control_from_golden(z)
    _, ym, nu, phi, coh, psi = mesh.material_arrays()
    for s in ctl["srf"]:
        assert strength_reduction(s, True, coh, nu, phi, psi)[4] == 1
    res = slope3d(ctx_seq, mesh, phinu=True, **ctl)
    ncall = 0
    for r, rr in zip(res.rows, z["summary"]):
        print("validation1_phinu", r, rr.tolist())
        n = int(rr[2])
        assert r["nl_tot"] == n and r["cg_tot"] == rr[1]
        assert res.cg_log[ncall:ncall + n] == list(z["cg_log"][ncall:ncall + n])
        ncall += n
        assert abs(r["uxmax"] - rr[3]) <= 1e-10 * rr[3]
    mix = slope3d_mixed(ctx_seq, mesh, phinu=True, ncol=2, **ctl)
    assert "fp64" in mix.extra.get("engine", "")
    assert mix.rows == res.rows
