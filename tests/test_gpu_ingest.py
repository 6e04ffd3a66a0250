"""GLL numbering on the device (sgeo_hex2spec) against the reference.

* The reference's own numbering (gdof_elmt and neq captured by the pcg_solver
  tap of the compiled reference) for the golden slope meshes.
* The host restatement (spec.hex2spec, itself pinned to the reference) on
  meshes with layers, jitter, pits and a sloping face, element for element:
  g_num, gdof, gdof_elmt, corners; the GLL point coordinates bit for bit
  against the reference's accumulation order (src/mesh_spec.f90:90-94).
"""
import time

import numpy as np
import pytest

from conftest import golden, mesh_from_golden
from specfem3d_geotech_amd.mesh import Material, pit_mesh, slope_mesh
from specfem3d_geotech_amd.spec import _shape_hex8_at_gll, hex2spec, hex2spec_device

pytestmark = pytest.mark.gpu


def _points_reference_order(mesh, spec):
    """x = 0 + s1*x1 + ... + s8*x8 per GLL point (src/mesh_spec.f90:90-94),
    stored at its node (the last point of a node wins, :119-136)."""
    sh = _shape_hex8_at_gll()                      # (8, 27)
    corners = np.asarray(mesh.coord)[np.asarray(mesh.conn) - 1]   # (nel, 8, 3)
    pts = np.zeros((corners.shape[0], 27, 3))
    for c in range(8):
        pts = pts + sh[c][None, :, None] * corners[:, c, None, :]
    out = np.empty((spec.nnode, 3))
    out[np.asarray(spec.g_num).reshape(-1)] = pts.reshape(-1, 3)
    return out


def _same(dev, host, mesh):
    assert dev.nnode == host.nnode and dev.neq == host.neq
    assert np.array_equal(dev.g_num, host.g_num)
    assert np.array_equal(dev.gdof, host.gdof)
    assert np.array_equal(dev.gdof_elmt, host.gdof_elmt)
    assert np.array_equal(dev.corners, host.corners)
    assert np.array_equal(dev.g_coord, _points_reference_order(mesh, host))


@pytest.mark.parametrize("name", ["slope_4x2x3", "slope_8x4x6_layered"])
def test_device_numbering_matches_reference(ctx, name):
    z = golden(name)
    mesh = mesh_from_golden(z)
    spec = hex2spec_device(ctx, mesh)
    assert spec.neq == int(z["neq"])
    assert np.array_equal(spec.gdof_elmt, z["gdof_elmt"])
    _same(spec, hex2spec(mesh), mesh)


def _meshes():
    two = [Material(gam=20.0, ym=1e5, nu=0.3, phi=25.0, coh=8.0),
           Material(gam=20.0, ym=1e5, nu=0.3, phi=30.0, coh=20.0)]
    soil = Material(gam=20.0, ym=1e5, nu=0.3, phi=30.0, coh=20.0)
    return {
        "c0": slope_mesh(24, 4, 12),
        "jittered_layers": slope_mesh(16, 6, 10, materials=two, layer_z=[10.0],
                                      jitter=0.02, seed=5),
        "steep_face": slope_mesh(10, 3, 7, angle_deg=41.0),
        "pit": pit_mesh(10, 8, 6, lx=20.0, ly=16.0, lz=12.0, pit=(5.0, 15.0, 4.0, 12.0),
                        nstage=3, stage_depth=2.0, soil=soil),
    }


@pytest.mark.parametrize("name", ["c0", "jittered_layers", "steep_face", "pit"])
def test_device_numbering_matches_host(ctx, name):
    mesh = _meshes()[name]
    _same(hex2spec_device(ctx, mesh), hex2spec(mesh), mesh)


def test_reference_example_excavation_numbering(ctx):
    """The reference's example input/excavation_3d mesh (1,700 elements)."""
    mesh = mesh_from_golden(golden("excavation_3d"))
    _same(hex2spec_device(ctx, mesh), hex2spec(mesh), mesh)


def test_device_numbering_large_mesh(ctx):
    """144,000 elements (the c1 mesh): equal to the host numbering; the
    device time is printed beside the host time."""
    mesh = slope_mesh(120, 20, 60, materials=[
        Material(gam=20.0, ym=1e5, nu=0.3, phi=25.0, coh=8.0),
        Material(gam=20.0, ym=1e5, nu=0.3, phi=30.0, coh=20.0)],
        layer_z=[10.0], jitter=0.02, seed=1)
    hex2spec_device(ctx, slope_mesh(4, 2, 3))  # first call: module load
    t0 = time.perf_counter()
    dev = hex2spec_device(ctx, mesh)
    t1 = time.perf_counter()
    host = hex2spec(mesh)
    t2 = time.perf_counter()
    print(f"numbering of 144,000 elements: device {t1 - t0:.3f} s, host {t2 - t1:.3f} s")
    assert dev.neq == 3489480
    assert np.array_equal(dev.g_num, host.g_num)
    assert np.array_equal(dev.gdof_elmt, host.gdof_elmt)
