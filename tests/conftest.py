import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def mesh_from_golden(z):
    from specfem3d_geotech_amd.mesh import HexMesh, Material
    mats = [Material(gam=r[0], ym=r[1], nu=r[2], phi=r[3], coh=r[4], psi=r[5])
            for r in z["materials"]]
    return HexMesh(coord=z["coord"], conn=z["conn"], mat_id=z["mat_id"],
                   materials=mats, bc={"ux": z["bc_ux"], "uy": z["bc_uy"],
                                       "uz": z["bc_uz"]})


def control_from_golden(z):
    c = z["control"]
    return dict(cg_tol=float(c[0]), cg_maxiter=int(c[1]), nl_tol=float(c[2]),
                nl_maxiter=int(c[3]), srf=[float(s) for s in z["srf"]])


@pytest.fixture(scope="session")
def ctx():
    from specfem3d_geotech_amd.solver import Context
    c = Context(0)
    yield c
    c.close()
