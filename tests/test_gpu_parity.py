"""Parity of the HIP path (through the C ABI) with the reference's outputs
and the oracle.

Two reduction orders are tested:
* SGEO_REDUCE_SEQUENTIAL sums the PCG dot products in the reference's order;
  then every PCG call reproduces the reference bit for bit (same u bits,
  same cg_iter), because the element products, the assembly and all vector
  updates already evaluate the reference's operations in its order.
* SGEO_REDUCE_TREE (the fast path) sums them as blocked trees.  Its
  tolerance is the reference's own sensitivity to the order of its sums,
  measured in-test with the oracle summing in descending order (an equally
  valid sequential order): relative L2 <= max(1e-10, 2 x that spread).
"""
import numpy as np
import pytest

import oracle as O
from conftest import control_from_golden, golden, mesh_from_golden
from specfem3d_geotech_amd.slope import slope3d
from specfem3d_geotech_amd.solver import (K_F64, REDUCE_SEQUENTIAL, Context,
                                          DeviceMesh, pcg_solver,
                                          pcg_solver_invalidate)
from specfem3d_geotech_amd.spec import hex2spec

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ctx_seq():
    c = Context(0)
    c.set_reduction(REDUCE_SEQUENTIAL)
    yield c
    c.close()


def order_spread(name):
    """Per-call relative L2 between the reference and the oracle run with
    descending-order dot products, over the golden calls."""
    z = golden(name)
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    km, _, _ = O.stiffness_bodyload(spec, gam, nu, ym)
    ctl = control_from_golden(z)
    O.set_dot_mode(1)
    try:
        worst = 0.0
        for j in range(len(z["call_idx"])):
            u, _, _ = O.pcg_solver(spec.neq, km, z["call_uin"][j], z["call_f"][j],
                                   z["dinv"], spec.gdof_elmt, ctl["cg_tol"],
                                   ctl["cg_maxiter"])
            worst = max(worst, rel_l2(u, z["call_uout"][j]))
    finally:
        O.set_dot_mode(0)
    return worst


@pytest.mark.parametrize("name", ["slope_4x2x3", "slope_8x4x6_layered"])
def test_device_stiffness_matches_reference(ctx, name):
    import hashlib
    z = golden(name)
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    dm = DeviceMesh(ctx, spec, len(gam))
    ext, dinv = dm.build_stiffness(gam, ym, nu)
    k = dm.download_stiffness()
    ko, _, _ = O.stiffness_bodyload(spec, gam, nu, ym)
    rel = np.abs(k - ko).max() / np.abs(ko).max()
    print(f"{name}: K max rel diff {rel:.3e}, bitwise {np.array_equal(k, ko)}")
    # the device K is the reference's bit for bit (same operation order,
    # no contraction): the sha256 of the reference's own storekm
    assert hashlib.sha256(np.ascontiguousarray(k).tobytes()).digest() == \
        z["k_sha256"].tobytes()
    assert np.array_equal(k, ko)
    assert np.array_equal(dinv, z["dinv"])
    assert np.array_equal(ext, z["extload"])
    dm.close()


def test_device_matvec_matches_oracle(ctx):
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    km, dp, _ = O.stiffness_bodyload(spec, gam, nu, ym)
    dm = DeviceMesh(ctx, spec, len(gam))
    dm.upload_stiffness(km, z["dinv"])
    p = np.random.default_rng(0).normal(size=spec.neq + 1)
    p[0] = 0.0
    kp = dm.matvec(p)
    ref = O.ebe_matvec(spec.neq, km, spec.gdof_elmt, p)
    print("matvec bitwise:", np.array_equal(kp, ref), "rel", rel_l2(kp, ref))
    assert np.array_equal(kp, ref)
    dm.close()


def test_seam1_pcg_calls_bitwise_in_reference_order(ctx_seq):
    """sgeo_pcg_solver with reference-order dots on the reference's own
    recorded calls: identical bits and iteration counts."""
    z = golden("slope_4x2x3")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    km, _, _ = O.stiffness_bodyload(spec, gam, nu, ym)
    ctl = control_from_golden(z)
    for j, c in enumerate(z["call_idx"]):
        u, it, ec, tag = pcg_solver(ctx_seq, spec.neq, spec.nelmt, km,
                                    z["call_uin"][j], z["call_f"][j], z["dinv"],
                                    spec.gdof_elmt, ctl["cg_tol"], ctl["cg_maxiter"])
        assert ec == 0 and it == z["cg_log"][c], (c, it, z["cg_log"][c])
        assert np.array_equal(u, z["call_uout"][j]), (c, rel_l2(u, z["call_uout"][j]))


def test_seam1_pcg_calls_match_reference(ctx):
    """sgeo_pcg_solver (tree reductions) on the reference's recorded calls."""
    z = golden("slope_4x2x3")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    km, _, _ = O.stiffness_bodyload(spec, gam, nu, ym)
    ctl = control_from_golden(z)
    tol = max(1e-10, 2.0 * order_spread("slope_4x2x3"))
    worst = 0.0
    for j, c in enumerate(z["call_idx"]):
        u, it, ec, tag = pcg_solver(ctx, spec.neq, spec.nelmt, km, z["call_uin"][j],
                                    z["call_f"][j], z["dinv"], spec.gdof_elmt,
                                    ctl["cg_tol"], ctl["cg_maxiter"])
        assert ec == 0 and tag == ""
        assert abs(it - int(z["cg_log"][c])) <= 1, (c, it, z["cg_log"][c])
        e = rel_l2(u, z["call_uout"][j])
        worst = max(worst, e)
        assert e < tol, (c, e, tol)
    print(f"worst per-call relative L2: {worst:.3e} (tolerance {tol:.3e})")


def test_seam1_in_place_edit_after_invalidate(ctx_seq):
    """k rewritten in place (same buffer) at an entry the cache fingerprint
    does not sample: after sgeo_pcg_solver_invalidate the solve uses the
    edited k, as the oracle on the edited k does."""
    z = golden("slope_4x2x3")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    km, _, _ = O.stiffness_bodyload(spec, gam, nu, ym)
    km = np.ascontiguousarray(km, dtype=np.float64)
    ctl = control_from_golden(z)
    args = (z["call_uin"][0], z["call_f"][0], z["dinv"], spec.gdof_elmt,
            ctl["cg_tol"], ctl["cg_maxiter"])
    u0, _, ec0, _ = pcg_solver(ctx_seq, spec.neq, spec.nelmt, km, *args)
    flat = km.reshape(-1)
    # a diagonal entry of a free equation (gdof != 0), off the words the
    # fingerprint samples (8192 evenly spaced values of k); dprecon is
    # passed unchanged, so only k differs
    step = max(1, flat.size // 8192)
    g = np.asarray(spec.gdof_elmt).reshape(spec.nelmt, 81)
    free = [e * 6561 + 82 * c for e in range(spec.nelmt) for c in range(81)
            if g[e, c] != 0 and (e * 6561 + 82 * c) % step != 0]
    q = free[len(free) // 2]
    assert q != flat.size - 1
    flat[q] *= 1.5
    pcg_solver_invalidate(ctx_seq)
    u1, _, ec1, _ = pcg_solver(ctx_seq, spec.neq, spec.nelmt, km, *args)
    u2, _, ec2 = O.pcg_solver(spec.neq, km, *args)
    assert ec0 == ec1 == ec2 == 0
    assert rel_l2(u1, u2) < 1e-12
    assert rel_l2(u1, u0) > 1e-9  # the edit changed the solution


def test_seam1_material_rebuilt_in_place_without_invalidate(ctx_seq):
    """The reference's phinu path (src/slope.F90:259-270) rebuilds storekm
    and dprecon in place, in the same buffers, when a material's nu
    changes, and never tells the solver.  Here the smaller material block of
    the layered slope gets a new nu: k and dprecon are rewritten in place and
    the drop-in is called again without sgeo_pcg_solver_invalidate.  The
    cache key reads every value of dprecon, so the new operator is used:
    the solve equals the oracle's on the new arrays bit for bit."""
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    ctl = control_from_golden(z)

    def operator(nu_):
        km_, dp_, _ = O.stiffness_bodyload(spec, gam, nu_, ym, with_extload=False)
        dinv_ = np.zeros_like(dp_)
        dinv_[1:] = 1.0 / dp_[1:]
        return np.ascontiguousarray(km_, dtype=np.float64), dinv_

    km, dinv = operator(nu)
    assert np.array_equal(dinv, z["dinv"])
    small = int(np.argmin(np.bincount(spec.mat_id)[1:])) + 1
    args = (z["call_uin"][1], z["call_f"][1])
    tail = (ctl["cg_tol"], ctl["cg_maxiter"])
    u0, it0, ec0, _ = pcg_solver(ctx_seq, spec.neq, spec.nelmt, km, *args, dinv,
                                 spec.gdof_elmt, *tail)
    nu2 = nu.copy()
    nu2[small - 1] = 0.45
    km2, dinv2 = operator(nu2)
    km[...] = km2  # in place: the same host buffers as the first call
    dinv[...] = dinv2
    u1, it1, ec1, _ = pcg_solver(ctx_seq, spec.neq, spec.nelmt, km, *args, dinv,
                                 spec.gdof_elmt, *tail)
    u2, it2, ec2 = O.pcg_solver(spec.neq, km2, *args, dinv2, spec.gdof_elmt, *tail)
    print(f"material {small}: nu {nu[small - 1]} -> 0.45; cg_iter {it0} -> {it1} "
          f"(oracle {it2})")
    assert ec0 == ec1 == ec2 == 0
    assert it1 == it2 and np.array_equal(u1, u2)
    assert rel_l2(u1, u0) > 1e-9  # the new operator changed the solution


def test_seam1_nonconvergence_errcode(ctx):
    z = golden("slope_4x2x3")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    km, _, ext = O.stiffness_bodyload(spec, gam, nu, ym)
    u, it, ec, tag = pcg_solver(ctx, spec.neq, spec.nelmt, km,
                                np.zeros(spec.neq + 1), ext, z["dinv"],
                                spec.gdof_elmt, 1e-8, 3)
    assert ec == -1 and it == 4
    assert tag == "ERROR: PCG solver doesn't converge!"
    u2, it2, ec2 = O.pcg_solver(spec.neq, km, np.zeros(spec.neq + 1), ext,
                                   z["dinv"], spec.gdof_elmt, 1e-8, 3)
    assert rel_l2(u, u2) < 1e-12


@pytest.mark.parametrize("name", ["slope_4x2x3", "slope_8x4x6_layered"])
def test_slope_sweep_matches_reference(ctx_seq, name):
    """Full SRF sweep on the device (reference-order dots) vs the reference's
    summary.  Converged SRFs: identical cg_iter of every PCG call, identical
    nl_iter, displacement within 1e-10, identical plastic-point set at the
    final pass.  The failing SRF (nl_iter = nl_maxiter, a non-converging
    viscoplastic flow) amplifies last-bit differences of the device
    transcendentals like any reordering of the reference's own sums does, so
    there only the failure itself -- hence the FOS bracket -- must agree."""
    z = golden(name)
    mesh = mesh_from_golden(z)
    ctl = control_from_golden(z)
    res = slope3d(ctx_seq, mesh, **ctl)
    ref = z["summary"]
    assert len(res.rows) == ref.shape[0]
    ncall = 0
    for s, (r, rr) in enumerate(zip(res.rows, ref)):
        print(name, r, rr.tolist())
        assert r["srf"] == rr[0] and r["nl_tot"] == rr[2]
        flags = res.extra["flags_per_srf"][s].astype(bool)
        ref_f = z["final_pass_f"][s].reshape(-1, 27)
        if rr[2] == ctl["nl_maxiter"]:
            flips = flags != (ref_f >= 0)
            print(f"  failing SRF: cg_tot {r['cg_tot']} vs {int(rr[1])}, "
                  f"{int(flips.sum())} flag flips, max |f| of flipped "
                  f"{np.abs(ref_f[flips]).max() if flips.any() else 0.0:.3e}")
            continue
        n = int(rr[2])
        assert res.cg_log[ncall:ncall + n] == list(z["cg_log"][ncall:ncall + n])
        ncall += n
        assert r["cg_tot"] == rr[1]
        assert abs(r["uxmax"] - rr[3]) <= 1e-10 * rr[3]
        assert abs(r["umax"] - rr[4]) <= 1e-10 * rr[4]
        assert abs(r["fmax"] - rr[5]) <= 1e-8 * max(1.0, abs(rr[5]))
        assert np.array_equal(flags, ref_f >= 0)
    bracket = (float(ref[-2, 0]) if ref.shape[0] > 1 else None, float(ref[-1, 0])) \
        if ref[-1, 2] == ctl["nl_maxiter"] else None
    assert res.fos_bracket == bracket


@pytest.mark.parametrize("name", ["slope_4x2x3", "slope_8x4x6_layered"])
def test_slope_sweep_tree_reductions(ctx, name):
    """Fast path: same FOS bracket and plastic iteration counts; converged
    SRFs within the reference's own order sensitivity."""
    z = golden(name)
    mesh = mesh_from_golden(z)
    ctl = control_from_golden(z)
    res = slope3d(ctx, mesh, **ctl)
    ref = z["summary"]
    tol = max(1e-10, 2.0 * order_spread(name))
    assert len(res.rows) == ref.shape[0]
    for r, rr in zip(res.rows, ref):
        print(name, r, rr.tolist())
        assert r["nl_tot"] == rr[2]
        if rr[2] < ctl["nl_maxiter"]:
            assert abs(r["uxmax"] - rr[3]) <= tol * rr[3]


def test_slope_sweep_matches_oracle_displacement(ctx):
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    ctl = control_from_golden(z)
    res = slope3d(ctx, mesh, spec=spec, **ctl)
    o = O.slope3d(spec, mesh, **ctl)
    e = rel_l2(res.x, o["x"])
    print("final x relative L2 vs oracle:", e)
    assert e < 1e-10


def test_packed_operator_matches_full(ctx):
    """Symmetric-packed FP64 operator: the reference's K_e is symmetric to
    the last bit, so the packed product agrees with the full one to a few
    ulps of the largest term."""
    from specfem3d_geotech_amd.solver import KFMT_PACKED
    z = golden("slope_8x4x6_layered")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    km, _, _ = O.stiffness_bodyload(spec, gam, nu, ym)
    dm = DeviceMesh(ctx, spec, len(gam))
    dm.build_stiffness(gam, ym, nu)
    dm.set_operator(K_F64, KFMT_PACKED)
    p = np.random.default_rng(1).normal(size=spec.neq + 1)
    p[0] = 0.0
    kp = dm.matvec(p)
    ref = O.ebe_matvec(spec.neq, km, spec.gdof_elmt, p)
    e = rel_l2(kp, ref)
    print("packed vs full relative L2:", e)
    assert e < 1e-14
    dm.close()
