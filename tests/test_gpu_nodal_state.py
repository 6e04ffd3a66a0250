"""State rules of the excavation stage's nodal accumulators.

The accumulators (per node: summed plastic strain, minimum stress
concentration factor) are sized for the mesh's node map.  With both the
single-trial and the batched engine:
* an accumulation that continues a stage (first=0) before one that starts
  it is refused with an error, not summed onto uninitialised memory;
* a node map replaced with a different node count (sgeo_mesh_set_nodes)
  re-sizes the accumulators at the next accumulation; until then the
  nodal fields carry no accumulated values, and the first accumulation on
  the new map must start the stage again.
No plastic iteration runs: begin_increment zeroes the per-point plastic
strain, so every accumulated vmeps is exactly zero.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import golden, mesh_from_golden
from specfem3d_geotech_amd import _lib as L
from specfem3d_geotech_amd.batch import BatchEngine
from specfem3d_geotech_amd.slope import DeviceSlope
from specfem3d_geotech_amd.solver import DeviceMesh, nodal_fields_call
from specfem3d_geotech_amd.spec import hex2spec

pytestmark = pytest.mark.gpu


def _mesh(ctx):
    mesh = mesh_from_golden(golden("pit_8x6x6"))
    spec = hex2spec(mesh)
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    dm = DeviceMesh(ctx, spec, len(gam))
    dm.build_stiffness(gam, ym, nu)
    mats = (ym, nu, np.radians(phi), coh, np.radians(psi))
    return dm, spec, mats


def _set_nodes(dm, spec, nnode):
    g = np.ascontiguousarray(spec.g_num, dtype=np.int32)
    L.check(L.lib().sgeo_mesh_set_nodes(dm.h, nnode, L.ptr(g, C.c_int32)),
            "sgeo_mesh_set_nodes")


def _check_rules(dm, spec, accumulate, fields_fn, fields_args):
    nn = spec.nnode
    with pytest.raises(L.SgeoError, match="first=0"):
        accumulate(False)
    accumulate(True)
    accumulate(False)
    f = nodal_fields_call(fields_fn, "nodal fields", nn, *fields_args)
    assert np.all(f["vmeps"] == 0.0) and np.all(np.isfinite(f["scf"]))
    # one more node, referenced by no element
    _set_nodes(dm, spec, nn + 1)
    f = nodal_fields_call(fields_fn, "nodal fields", nn + 1, *fields_args)
    assert np.all(f["scf"] == np.float32(1e32)) and np.all(f["vmeps"] == 0.0)
    with pytest.raises(L.SgeoError, match="first=0"):
        accumulate(False)
    accumulate(True)
    f = nodal_fields_call(fields_fn, "nodal fields", nn + 1, *fields_args)
    assert np.all(f["vmeps"][:nn] == 0.0) and np.all(np.isfinite(f["scf"][:nn]))
    assert f["scf"][nn] == np.float32(1e32)  # no point lowers it
    _set_nodes(dm, spec, nn)  # the map the host objects were built with


def test_slope_accumulator_rules(ctx):
    dm, spec, mats = _mesh(ctx)
    t = DeviceSlope(dm, np.zeros(spec.neq + 1))
    try:
        t.begin_increment(*mats, 1.0, stress0=np.zeros((spec.nelmt, 27, 6)))
        dm.ensure_nodes()
        _check_rules(dm, spec, t.accumulate_nodal, L.lib().sgeo_slope_nodal_fields, (t.h,))
    finally:
        t.close()
        dm.close()


def test_batch_accumulator_rules(ctx):
    dm, spec, mats = _mesh(ctx)
    eng = BatchEngine(dm, np.zeros(spec.neq + 1), 2)
    try:
        for col in (0, 1):
            eng.begin_increment(col, *mats, 1.0, stress0=np.zeros((spec.nelmt, 27, 6)))
        dm.ensure_nodes()
        _check_rules(dm, spec, lambda first: eng.accumulate_nodal(1, first),
                     L.lib().sgeo_batch_nodal_fields, (eng.h, 1))
        # column 0 never accumulated on this map: no accumulated values
        f = nodal_fields_call(L.lib().sgeo_batch_nodal_fields, "nodal fields", spec.nnode,
                              eng.h, 0)
        assert np.all(f["scf"] == np.float32(1e32))
    finally:
        eng.close()
        dm.close()
