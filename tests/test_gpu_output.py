"""The reference's result files written from the GPU sweeps.

Two-SRF sweep of the 4x2x3 slope with every result field the input file
can ask for (displacement, nodal stress, principal stresses, normal
stress, stress concentration factor, plastic strain), against the compiled
reference's files:
* the EnSight case and geometry files, and the constant scf and vmeps
  files, byte for byte;
* FP64 sweep with reference-order reductions: CG and plastic iteration
  counts exactly, displacements within 1e-10 and fmax within 1e-8 (the
  yield checks' device transcendentals differ from the host libm in the
  last bit), so each FP32 displacement and stress-field value equals the
  reference's or lies one FP32 spacing away;
* mixed-precision sweep (both SRFs at once): the same files in SRF order,
  displacements and stress fields within the solver tolerance.
"""
import os

import numpy as np
import pytest

from conftest import golden
from specfem3d_geotech_amd.batch import slope3d_mixed
from specfem3d_geotech_amd.mesh import HexMesh, Material
from specfem3d_geotech_amd.output import SlopeOutput
from specfem3d_geotech_amd.slope import slope3d
from specfem3d_geotech_amd.solver import REDUCE_SEQUENTIAL, Context
from specfem3d_geotech_amd.spec import hex2spec_device

pytestmark = pytest.mark.gpu

CTL = dict(cg_tol=1e-8, cg_maxiter=5000, nl_tol=5e-4, nl_maxiter=300)
SRF = [1.0, 1.2]
SAVE_ALL = ("disp", "stress", "psigma", "nsigma", "scf", "vmeps")
FLOAT_FIELDS = (".dis", ".sig", ".psig", ".nsig")  # computed from the solution


def _mesh(z):
    mats = [Material(gam=r[0], ym=r[1], nu=r[2], phi=r[3], coh=r[4], psi=r[5])
            for r in z["materials"]]
    return HexMesh(coord=z["coord"], conn=z["conn"], mat_id=z["mat_id"], materials=mats,
                   bc={"ux": z["bc_ux"], "uy": z["bc_uy"], "uz": z["bc_uz"]})


def _summary_values(text):
    lines = text.split("\n")
    lines = lines[:[i for i, ln in enumerate(lines) if "ELAPSED TIME" in ln][0]]
    out = []
    for ln in lines:
        for t in ln.split():
            try:
                out.append(float(t))
            except ValueError:
                out.append(t)
    return out


def _fp32_close(mine: bytes, ref: bytes) -> bool:
    """Each FP32 value equal to the reference's or one FP32 spacing away;
    values near zero within 1e-9 of the field's largest."""
    a = np.frombuffer(mine, np.float32).astype(np.float64)
    b = np.frombuffer(ref, np.float32).astype(np.float64)
    tol = np.maximum(np.spacing(np.abs(b).astype(np.float32)).astype(np.float64),
                     1e-9 * np.abs(b).max())
    return a.shape == b.shape and bool(np.all(np.abs(a - b) <= tol))


def test_fp64_sweep_files_match_reference(tmp_path):
    z = golden("output_slope_4x2x3")
    mesh = _mesh(z)
    ctx = Context(0)
    try:
        ctx.set_reduction(REDUCE_SEQUENTIAL)
        spec = hex2spec_device(ctx, mesh)
        out = SlopeOutput(str(tmp_path), "slope_out", mesh, spec, CTL, SRF, save=SAVE_ALL)
        slope3d(ctx, mesh, spec=spec, srf=SRF, output=out, **CTL)
        out.close()
    finally:
        ctx.close()
    names = [str(n) for n in z["file_names"]]
    assert sorted(os.listdir(tmp_path)) == names
    for n in names:
        ref, mine = z["file:" + n].tobytes(), (tmp_path / n).read_bytes()
        if n.endswith("_summary"):
            va, vb = _summary_values(mine.decode()), _summary_values(ref.decode())
            assert len(va) == len(vb)
            nrow = 6 * len(SRF)
            for k, (a, b) in enumerate(zip(va, vb)):
                if len(va) - k <= nrow and (len(va) - k) % 6 == 1:
                    # fmax: the largest yield value of the last pass, as in
                    # the sweep parity tests
                    assert abs(a - b) <= 1e-8 * max(1.0, abs(b)), (a, b)
                elif isinstance(b, float) and b != int(b):
                    assert abs(a - b) <= 1e-10 * abs(b), (a, b)
                else:
                    assert a == b
        elif n.endswith(FLOAT_FIELDS):
            assert mine[:244] == ref[:244]
            assert _fp32_close(mine[244:], ref[244:]), n
        else:
            assert mine == ref, n


def test_mixed_sweep_files(tmp_path):
    z = golden("output_slope_4x2x3")
    mesh = _mesh(z)
    ctx = Context(0)
    try:
        spec = hex2spec_device(ctx, mesh)
        out = SlopeOutput(str(tmp_path), "slope_out", mesh, spec, CTL, SRF, save=SAVE_ALL)
        slope3d_mixed(ctx, mesh, spec=spec, srf=SRF, ncol=2, output=out, **CTL)
        out.close()
    finally:
        ctx.close()
    names = [str(n) for n in z["file_names"]]
    assert sorted(os.listdir(tmp_path)) == names
    for n in names:
        ref, mine = z["file:" + n].tobytes(), (tmp_path / n).read_bytes()
        if n.endswith(FLOAT_FIELDS):
            a = np.frombuffer(mine[244:], np.float32).astype(np.float64)
            b = np.frombuffer(ref[244:], np.float32).astype(np.float64)
            assert mine[:244] == ref[:244]
            assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b), n
        elif n.endswith("_summary"):
            va, vb = _summary_values(mine.decode()), _summary_values(ref.decode())
            assert len(va) == len(vb)
            # rows: srf, cg_tot (mixed counts may differ slightly), nl_tot, uxmax, umax
            rows_a, rows_b = va[-12:], vb[-12:]
            for r in range(2):
                a, b = rows_a[6 * r:6 * r + 6], rows_b[6 * r:6 * r + 6]
                assert a[0] == b[0] and a[2] == b[2]
                assert abs(a[3] - b[3]) <= 1e-7 * b[3] and abs(a[4] - b[4]) <= 1e-7 * b[4]
        else:
            assert mine == ref, n


def test_device_nodal_fields_match_oracle():
    """sgeo_slope_nodal_fields on an injected stress state against the
    oracle's nodal fields: the nodal stress bit for bit (same sums in the
    same order, then one rounding to FP32), the principal and normal
    stresses equal or one FP32 spacing away (device sin), on a mesh with
    every node valency a slope mesh has and with hydrostatic elements
    (dsbar at its floor, Lode angle 0)."""
    import oracle as O
    from specfem3d_geotech_amd.mesh import slope_mesh
    from specfem3d_geotech_amd.slope import DeviceSlope
    from specfem3d_geotech_amd.solver import K_F64, DeviceMesh

    mesh = slope_mesh(10, 3, 6)
    ctx = Context(0)
    try:
        spec = hex2spec_device(ctx, mesh)
        gam, ym, nu, phi, coh, psi = mesh.material_arrays()
        rng = np.random.default_rng(3)
        stress = rng.normal(size=(spec.nelmt, 27, 6)) * 100.0
        hyd = rng.random(spec.nelmt) < 0.2
        stress[hyd, :, 3:] = 0.0
        stress[hyd, :, 1] = stress[hyd, :, 0]
        stress[hyd, :, 2] = stress[hyd, :, 0]
        dm = DeviceMesh(ctx, spec, len(gam))
        try:
            ext, _ = dm.build_stiffness(gam, ym, nu, kprec=K_F64)
            st = DeviceSlope(dm, ext)
            try:
                st.begin_increment(ym, nu, phi, coh, psi, 1.0, stress0=stress)
                f = st.nodal_fields()
            finally:
                st.close()
        finally:
            dm.close()
    finally:
        ctx.close()
    o = O.nodal_fields(spec, stress)
    print("nnode", spec.nnode, "hydrostatic elements", int(hyd.sum()), flush=True)
    assert np.array_equal(f["stress"], o["stress"].astype(np.float32))
    for k in ("psigma", "nsigma"):
        r = o[k].astype(np.float32)
        assert np.all(np.abs(f[k].astype(np.float64) - r) <=
                      np.spacing(np.abs(r)).astype(np.float64)), k
    assert not np.any(f["vmeps"])


def _compare_run_files(out_dir, z, nrow_fields=6):
    """Every file of a run against the reference's: names, case/geometry
    bytes, FP32 fields within one spacing, summary text before the rows
    and the rows' numbers within the FP64 sweep bars."""
    names = [str(n) for n in z["file_names"]]
    assert sorted(os.listdir(out_dir)) == names
    for n in names:
        ref, mine = z["file:" + n].tobytes(), open(os.path.join(out_dir, n), "rb").read()
        if n.endswith("_summary"):
            va, vb = _summary_values(mine.decode()), _summary_values(ref.decode())
            assert len(va) == len(vb)
            for a, b in zip(va, vb):
                if isinstance(b, float) and b != int(b):
                    assert abs(a - b) <= 1e-8 * max(1.0, abs(b)), (a, b)
                else:
                    assert a == b, (a, b)
        elif n.endswith(".geo") or n.endswith(".case"):
            assert mine == ref, n
        else:
            assert mine[:244] == ref[:244], n
            assert _fp32_close(mine[244:], ref[244:]), n


def test_run_input_files_match_reference(tmp_path):
    """driver.run_input on reference input files written for the two golden
    runs (the 4x2x3 slope sweep and the two-stage pit, every field saved):
    the result files land under the file's out_path and head and match the
    compiled reference's."""
    from specfem3d_geotech_amd.driver import run_input
    from specfem3d_geotech_amd.mesh import write_reference_input

    ctx = Context(0)
    try:
        ctx.set_reduction(REDUCE_SEQUENTIAL)
        z = golden("output_slope_4x2x3")
        sem = write_reference_input(_mesh(z), str(tmp_path / "s"), "slope_out",
                                    out_path=str(tmp_path / "s" / "output"), srf=SRF,
                                    save=SAVE_ALL, **CTL)
        r = run_input(sem, ctx)
        print("slope", r.rows, flush=True)
        _compare_run_files(str(tmp_path / "s" / "output"), z)
        z = golden("output_pit_6x4x4")
        nexcav, ninc, *excavid = [int(v) for v in z["excav"]]
        sem = write_reference_input(
            _mesh(z), str(tmp_path / "e"), "pit_out", out_path=str(tmp_path / "e" / "output"),
            cg_tol=1e-8, cg_maxiter=5000, nl_tol=1e-5, nl_maxiter=100, srf=[1.0],
            extra_control=f", nexcav={nexcav}, excavid={' '.join(map(str, excavid))}, "
                          f"ninc={ninc}", save=SAVE_ALL)
        r = run_input(sem, ctx)
        print("pit", r.rows, flush=True)
        _compare_run_files(str(tmp_path / "e" / "output"), z)
    finally:
        ctx.close()
