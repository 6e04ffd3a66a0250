"""Host-side pieces (no GPU): numbering, inverse map, input format round
trip, per-trial scalars, and the C-ABI library's exports."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, golden, mesh_from_golden
from specfem3d_geotech_amd import plastic
from specfem3d_geotech_amd.mesh import (Material, read_reference_input, slope_mesh,
                                        write_reference_input)
from specfem3d_geotech_amd.spec import hex2spec


def test_inverse_map_reproduces_element_order():
    z = golden("slope_8x4x6_layered")
    spec = hex2spec(mesh_from_golden(z))
    rowptr, slots = spec.inverse_map()
    flat = spec.gdof_elmt.reshape(-1)
    for g in (1, 2, 17, spec.neq // 2, spec.neq):
        s = slots[rowptr[g]:rowptr[g + 1]]
        assert (flat[s] == g).all()
        assert (np.diff(s // 81) > 0).all()  # ascending element order
    assert rowptr[-1] == flat.size


def test_reference_input_round_trip(tmp_path):
    m = slope_mesh(4, 2, 3, materials=[Material(20, 1e5, 0.3, 25, 8),
                                       Material(20, 1e5, 0.3, 30, 20)],
                   layer_z=[10.0], jitter=0.02, seed=7)
    sem = write_reference_input(m, str(tmp_path), "rt", cg_tol=1e-8,
                                cg_maxiter=5000, nl_tol=5e-4, nl_maxiter=300,
                                srf=[1.0, 1.25])
    m2, ctl = read_reference_input(sem)
    assert np.array_equal(m2.coord, m.coord)
    assert np.array_equal(m2.conn, m.conn)
    assert np.array_equal(m2.mat_id, m.mat_id)
    for k in ("ux", "uy", "uz"):
        assert np.array_equal(m2.bc[k], m.bc[k])
    assert ctl["srf"] == [1.0, 1.25] and ctl["nl_maxiter"] == 300
    assert np.array_equal(hex2spec(m2).gdof_elmt, hex2spec(m).gdof_elmt)


def test_trial_scalars_match_oracle():
    """The product's host restatement of strength_reduction / dt_viscoplas
    against the oracle (itself pinned bit for bit to the reference)."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = 3
        coh, nu = rng.uniform(1, 40, n), rng.uniform(0.1, 0.45, n)
        phi, psi = rng.uniform(0, 45, n), rng.uniform(0, 15, n)
        ym = rng.uniform(1e4, 1e6, n)
        srf = float(rng.choice([1.0, rng.uniform(0.8, 3.0)]))
        for phinu in (False, True):
            a = plastic.strength_reduction(srf, phinu, coh, nu, phi, psi)
            b = O.strength_reduction(srf, phinu, coh, nu, phi, psi)
            for x, y in zip(a[:4], b[:4]):
                assert np.array_equal(x, y)
            assert a[4] == b[4]
            assert plastic.dt_viscoplas(a[1], a[2], ym) == O.dt_viscoplas(b[1], b[2], ym)


def _declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "sgeo.h")).read()
    return sorted(set(re.findall(r"\b(sgeo_[a-z0-9_]+)\s*\(", hdr)))


def test_capi_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, "specfem3d_geotech_amd", "libsgeo.so")
    if not os.path.exists(lib):
        pytest.fail("libsgeo.so not built (python -m specfem3d_geotech_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (sgeo_[a-z0-9_]+)", out))
    missing = [s for s in _declared_symbols() if s not in exported]
    assert not missing, missing
    # the ctypes table covers the header too
    from specfem3d_geotech_amd._lib import SIGNATURES
    assert set(_declared_symbols()) <= set(SIGNATURES), \
        set(_declared_symbols()) - set(SIGNATURES)


def test_capi_library_loads_without_gpu():
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "specfem3d_geotech_amd", "libsgeo.so"))
    assert hasattr(lib, "sgeo_pcg_solver")


def test_overburden_stress_matches_oracle():
    """The excavation driver's s0_type 1 initial stress (host, closed form)
    equals the oracle's restatement bit for bit."""
    from specfem3d_geotech_amd.excavation import overburden_stress
    z = golden("pit_overburden_6x4x4")
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam = mesh.material_arrays()[0]
    z0, s0, k0 = (float(v) for v in z["stress0"])
    assert np.array_equal(overburden_stress(spec, gam, z0, s0, k0),
                          O.overburden_stress(spec, gam, z0, s0, k0))


def test_reference_input_save_and_stress0_keys(tmp_path):
    """The reader takes the save: keys and the stress0 line as the
    reference's parser does (src/read_input.f90:236-250, 380-403)."""
    m = slope_mesh(2, 1, 2)
    sem = write_reference_input(m, str(tmp_path), "k", cg_tol=1e-8, cg_maxiter=50,
                                nl_tol=1e-4, nl_maxiter=10, srf=[1.0],
                                extra_lines="stress0: type=1, z0=8.0, s0=-5.0\n",
                                save=("disp", "psigma", "vmeps"))
    _, ctl = read_reference_input(sem)
    assert ctl["save"] == ["disp", "psigma", "vmeps"]
    assert ctl["s0_type"] == 1 and ctl["z0"] == 8.0 and ctl["s0"] == -5.0
    assert ctl["k0"] == 0.0
    assert ctl["out_path"].rstrip("/").endswith("output")


def test_reference_input_numpy_scalars(tmp_path):
    """Numpy scalars (as meshes rebuilt from .npz fixtures hold) are written
    as plain numbers the reference's parser reads."""
    m = slope_mesh(2, 1, 2, materials=[Material(np.float64(20.0), np.float64(1e5),
                                                np.float64(0.3), np.float64(25.0),
                                                np.float64(10.0))])
    sem = write_reference_input(m, str(tmp_path), "n", cg_tol=np.float64(1e-8),
                                cg_maxiter=np.int64(50), nl_tol=1e-4, nl_maxiter=10,
                                srf=[np.float64(1.0)])
    assert "np." not in open(sem).read()
    m2, ctl = read_reference_input(sem)
    assert m2.materials[0].ym == 1e5 and ctl["cg_tol"] == 1e-8
    assert "np." not in open(str(tmp_path / "n_material_list")).read()
