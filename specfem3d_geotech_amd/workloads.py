"""Seeded synthetic workloads of the benchmark configurations.

c0  1,152 elements (24x4x12), homogeneous 2H:1V slope, c=10 kPa, phi=20 deg,
    psi=0, gamma=20 kN/m3, E=1e5 kPa, nu=0.3; base fixed, side rollers.
c1  144,000 elements (120x20x60), same geometry, two layers split at
    z=10 m (upper c=8 kPa, phi=25 deg; lower c=20 kPa, phi=30 deg), interior
    nodes jittered by U[-0.02,0.02] x local spacing (seed 1).
c2  400,000 elements (200x20x100), the layers and jitter of c1, slope
    angle drawn from U[26.6, 45] deg (numpy default_rng(2): 31.41 deg), the
    mixed-precision configuration.
Control of all: SRF 1.00:0.05:2.00, cg_tol 1e-8, cg_maxiter 5000,
nl_tol 5e-4, nl_maxiter 300.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Dict

from .mesh import Material, slope_mesh
from .spec import hex2spec

SRF_SWEEP = [round(1.0 + 0.05 * i, 2) for i in range(21)]
CONTROL = dict(cg_tol=1e-8, cg_maxiter=5000, nl_tol=5e-4, nl_maxiter=300)


@dataclass
class Workload:
    name: str
    description: str
    make: Callable
    control: Dict = field(default_factory=lambda: dict(CONTROL, srf=SRF_SWEEP))

    def build(self):
        mesh = self.make()
        return mesh, hex2spec(mesh)


def _c0():
    return slope_mesh(24, 4, 12, materials=[
        Material(gam=20.0, ym=1e5, nu=0.3, phi=20.0, coh=10.0, psi=0.0)])


def _c1():
    return c1_like(120, 20, 60)


def c1_like(nx: int, ny: int, nz: int):
    """The features of c1 (two layers split at z = 10 m, 2% interior jitter,
    seed 1) on a coarser nx x ny x nz grid of the same 2H:1V slope."""
    return slope_mesh(nx, ny, nz, materials=[
        Material(gam=20.0, ym=1e5, nu=0.3, phi=25.0, coh=8.0, psi=0.0),
        Material(gam=20.0, ym=1e5, nu=0.3, phi=30.0, coh=20.0, psi=0.0)],
        layer_z=[10.0], jitter=0.02, seed=1)


def c2_angle() -> float:
    import numpy as np
    return float(np.random.default_rng(2).uniform(26.6, 45.0))


def _c2():
    return slope_mesh(200, 20, 100, angle_deg=c2_angle(), materials=[
        Material(gam=20.0, ym=1e5, nu=0.3, phi=25.0, coh=8.0, psi=0.0),
        Material(gam=20.0, ym=1e5, nu=0.3, phi=30.0, coh=20.0, psi=0.0)],
        layer_z=[10.0], jitter=0.02, seed=1)


WORKLOADS = {
    "c0": Workload("c0", "slope 24x4x12 (1152 el), homogeneous, FP64", _c0),
    "c1": Workload("c1", "slope 120x20x60 (144000 el, 3489480 eq), two layers, "
                         "2% jitter seed 1", _c1),
    "c2": Workload("c2", "slope 200x20x100 (400000 el), face angle 31.41 deg "
                         "(U[26.6,45] seed 2), two layers, 2% jitter seed 1", _c2),
}


def get(name: str) -> Workload:
    return WORKLOADS[name]


def form_a(n: int, nmat: int, seed: int = 3, c_mean: float = 10.0, c_cov: float = 0.3,
           phi_mean_deg: float = 20.0, tan_cov: float = 0.2):
    """Strength realisations of the batched configuration, form A: lognormal
    c (mean 10 kPa, CoV 0.3) and lognormal tan(phi) (mean tan 20 deg, CoV
    0.2), spatially uniform per trial (every material of the mesh takes the
    trial's c and phi).  Returns [(coh[nmat], phi_deg[nmat])] * n."""
    import numpy as np
    # one row of two standard normals per trial: trial i is the same for
    # every n
    z = np.random.default_rng(seed).standard_normal((n, 2))

    def lognormal(mean, cov, zz):
        s2 = np.log1p(cov * cov)
        return np.exp(np.log(mean) - 0.5 * s2 + np.sqrt(s2) * zz)

    c = lognormal(c_mean, c_cov, z[:, 0])
    t = lognormal(np.tan(np.radians(phi_mean_deg)), tan_cov, z[:, 1])
    phi = np.degrees(np.arctan(t))
    return [(np.full(nmat, c[i]), np.full(nmat, phi[i])) for i in range(n)]


def form_b(n: int = 64, lo: float = 1.0, hi: float = 2.0):
    """Form B: n SRF values on [lo, hi] for one strength set."""
    import numpy as np
    return [float(s) for s in np.linspace(lo, hi, n)]
