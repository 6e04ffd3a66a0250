"""Host mirror of `excavation3d` (src/excavation.F90:13-700) over the device
path.

The reference's stage logic stays on the host in its order: intact and void
element sets, intact nodes, the renumbered equations (integer work,
src/excavation_library.f90:17-111), the nodal displacement sum and the
stage summary row, and the overburden initial stress of s0_type 1 (a
closed form per GLL point, src/postprocess.f90:10-33).  Everything else
numeric runs on the GPU through libsgeo.so:
- the initial gravity solve and the elastic stress (s0_type 0);
- each stage's excavation load, built from the removed elements' stress;
- each stage's stiffness, built on the intact elements with the reduced nu;
- every load increment's plastic loop, with its PCG and the element x GLL
  body, starting from the in-situ stress.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from .batch import DEFAULT_INNER_TOL
from .mesh import HexMesh
from .plastic import dt_viscoplas, strength_reduction
from .slope import DeviceSlope, PCGNotConverged
from .solver import K_F64, KFMT_FULL, Context, DeviceMesh
from .spec import SpecMesh, hex2spec_device


def intact_void(mat_id: np.ndarray, ismat: np.ndarray, removed: Sequence[int]):
    """intact_void_elmt (src/excavation_library.f90:17-53): elements of the
    materials removed now are void; elements of still-active materials are
    intact (earlier removals are neither); both in element order."""
    off = np.zeros(ismat.size, bool)
    off[np.asarray(removed) - 1] = True
    m = np.asarray(mat_id) - 1
    void = np.nonzero(off[m])[0].astype(np.int32)
    intact = np.nonzero(~off[m] & ismat[m])[0].astype(np.int32)
    return intact, void


def renumber(gdof_node: np.ndarray, alive: np.ndarray):
    """modify_gdof (src/excavation_library.f90:90-111): dofs of dead nodes
    fixed, the free ones renumbered node-major, component fastest."""
    g = np.where(alive[:, None], gdof_node, 0)
    free = g.reshape(-1) != 0
    out = np.zeros(g.size, np.int64)
    out[free] = np.arange(1, int(free.sum()) + 1)
    return out.reshape(g.shape), int(free.sum())


class _MixedStage:
    """One excavation stage's trial on a one-column mixed-precision engine
    (batch.BatchEngine), with the calls the stage loop makes on a
    DeviceSlope.  DeviceSlope's download gives x and stress; here the
    engine runs each increment's whole plastic loop (src/excavation.F90:
    517-627) and reports its plastic iterations, CG iterations and fmax."""

    def __init__(self, dmesh: DeviceMesh, extload: np.ndarray, inner_tol: float):
        from .batch import BatchEngine
        self.eng = BatchEngine(dmesh, extload, 1)
        self.inner_tol = inner_tol

    def begin_increment(self, ym, nuf, phif, cohf, psif, dt, stress0=None):
        self.eng.begin_increment(0, ym, nuf, phif, cohf, psif, dt, stress0)

    def run(self, cg_tol, cg_maxiter, nl_tol, nl_maxiter):
        while True:
            r = self.eng.advance(cg_tol, cg_maxiter, self.inner_tol, nl_tol, nl_maxiter)
            if r["ended"][0]:
                break
        if r["errcode"][0] != 0:
            raise PCGNotConverged("ERROR: PCG solver doesn't converge!")
        return int(r["nl_tot"][0]), int(r["cg_tot"][0]), float(r["fmax"][0])

    def x(self):
        return self.eng.download(0)[0]

    def stress(self):
        return self.eng.download_stress(0)

    def accumulate_nodal(self, first):
        self.eng.accumulate_nodal(0, first)

    def nodal_fields(self):
        return self.eng.nodal_fields(0)

    def close(self):
        self.eng.close()


def overburden_stress(spec: SpecMesh, gam: np.ndarray, z0: float, p0: float,
                      k0: float) -> np.ndarray:
    """Initial stress of s0_type 1 (src/postprocess.f90:10-33): at GLL point
    i of element e, with z the z coordinate of its node, szz = p0 - gam |z -
    z0| (compression negative), sxx = syy = k0 szz, shears zero; (E, 27, 6)
    in the reference's stress order."""
    z = np.asarray(spec.g_coord, dtype=np.float64)[np.asarray(spec.g_num), 2]
    g = np.asarray(gam, dtype=np.float64)[np.asarray(spec.mat_id) - 1][:, None]
    szz = p0 - g * np.abs(z - z0)
    st = np.zeros((spec.nelmt, 27, 6))
    st[:, :, 2] = szz
    st[:, :, 0] = k0 * szz
    st[:, :, 1] = k0 * szz
    return st


def sub_spec(spec: SpecMesh, elmt: np.ndarray, gdof_node: np.ndarray,
             gdof_elmt: np.ndarray, neq: int) -> SpecMesh:
    return SpecMesh(nelmt=int(elmt.size), nnode=spec.nnode, neq=neq,
                    g_num=np.asarray(spec.g_num)[elmt], g_coord=spec.g_coord,
                    gdof=gdof_node, gdof_elmt=np.ascontiguousarray(gdof_elmt[elmt]),
                    corners=np.ascontiguousarray(np.asarray(spec.corners)[elmt]),
                    mat_id=np.ascontiguousarray(np.asarray(spec.mat_id)[elmt]))


@dataclass
class ExcavationResult:
    rows: List[Dict[str, float]]      # (stage, cg_tot, nl_tot, uxmax, umax, fmax)
    cg_log: List[int]
    nodalu: np.ndarray                # (nnode, 3)
    stress: np.ndarray                # (nelmt, 27, 6) after the last stage
    extra: Dict[str, object] = field(default_factory=dict)


def excavation3d(ctx: Context, mesh: HexMesh, *, cg_tol: float, cg_maxiter: int,
                 nl_tol: float, nl_maxiter: int, srf: Sequence[float], nexcav: int,
                 nexcavid: Sequence[int], excavid: Sequence[int], ninc: int,
                 phinu: bool = False, s0_type: int = 0, kprec: int = K_F64,
                 kfmt: int = KFMT_FULL, spec: Optional[SpecMesh] = None,
                 progress=None, output=None, z0: float = 0.0, s0: float = 0.0,
                 k0: float = 0.0, form: str = "fp64",
                 inner_tol: float = DEFAULT_INNER_TOL,
                 iter_progress=None) -> ExcavationResult:
    """Staged excavation on the GPU for one SRF.  The reference runs
    excavation with nsrf = 1 only (it reallocates and closes per-SRF state
    inside its SRF loop); the first SRF is used.  `output`
    (output.ExcavationOutput) receives the in-situ step and every stage's
    row, displacement and nodal fields (computed on the device) for the
    reference's result files; `progress` is called with each stage's row and
    `iter_progress` with (stage, increment, nl_iter, cg_iter) after each
    plastic iteration of the FP64 form.  form "fp64" runs each increment's plastic
    loop on the FP64 drop-in (operator kprec/kfmt), "mixed" on a one-column
    mixed-precision engine (FP32 inner CG, FP64 residuals to cg_tol, the
    elastic memo serving every increment after a stage's first); the
    in-situ gravity solve is FP64 in both.  s0_type 0 takes the in-situ stress from a
    gravity solve, s0_type 1 from the overburden relation with datum z0,
    stress s0 there and lateral ratio k0 (the input's stress0 line,
    src/read_input.f90:236-250; k0 defaults to 0 as the reference leaves it
    unset)."""
    if s0_type not in (0, 1):
        raise ValueError(f"s0_type {s0_type} not supported (src/excavation.F90:242)")
    if form not in ("fp64", "mixed"):
        raise ValueError(f"unknown form {form!r}")
    spec = spec or hex2spec_device(ctx, mesh)
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    nmat, E = len(gam), spec.nelmt
    g_num = np.asarray(spec.g_num)
    mat = np.asarray(spec.mat_id)
    cg_log: List[int] = []
    # initial stress: the gravity solve on the whole mesh (s0_type 0,
    # src/excavation.F90:180-237) or the overburden relation (s0_type 1,
    # :238-241)
    dm = DeviceMesh(ctx, spec, nmat)
    try:
        if s0_type == 0:
            ext, _ = dm.build_stiffness(gam, ym, nu, kprec=K_F64)
            dm.set_operator(kprec, kfmt)
            x, it, ec = dm.solve(ext, None, cg_tol, cg_maxiter)
            if ec != 0:
                raise PCGNotConverged("ERROR: PCG solver doesn't converge!")
            cg_log.append(it)
            x[0] = 0.0
            stress = dm.elastic_stress(ym, nu, x)
        else:
            stress = overburden_stress(spec, gam, z0, s0, k0)
        if output is not None:  # step 0 (src/excavation.F90:296-323)
            output.in_situ(fields=lambda: dm.nodal_fields(stress))
    finally:
        dm.close()
    s = float(srf[0])
    gdof_node = np.asarray(spec.gdof).copy()
    nodalu = np.zeros((spec.nnode, 3))
    ismat = np.ones(nmat, bool)
    rows: List[Dict[str, float]] = []
    for i_ex in range(nexcav):
        id0 = i_ex * nexcavid[i_ex]  # src/excavation.F90:339
        removed = list(excavid[id0:id0 + nexcavid[i_ex]])
        ismat[np.asarray(removed) - 1] = False
        intact, void = intact_void(mat, ismat, removed)
        if intact.size == 0:
            break
        alive = np.zeros(spec.nnode, bool)
        alive[g_num[intact].reshape(-1)] = True
        gdof_node, neq = renumber(gdof_node, alive)
        gdof_elmt = gdof_node[g_num].reshape(E, 81).astype(np.int32)
        # excavation load of the removed elements (numbered for this stage)
        extload = np.zeros(neq + 1)
        if void.size:
            dv = DeviceMesh(ctx, sub_spec(spec, void, gdof_node, gdof_elmt, neq), nmat)
            try:
                extload = dv.excavation_load(gam, stress[void])
            finally:
                dv.close()
        cohf, nuf, phif, psif, _ = strength_reduction(s, phinu, coh, nu, phi, psi)
        dt = dt_viscoplas(nuf, phif, ym, ismat)
        di = DeviceMesh(ctx, sub_spec(spec, intact, gdof_node, gdof_elmt, neq), nmat)
        trial = None
        try:
            di.build_stiffness(gam, ym, nuf, kprec=K_F64, want_extload=False)
            if form == "mixed":
                trial = _MixedStage(di, extload / ninc, inner_tol)
            else:
                di.set_operator(kprec, kfmt)
                trial = DeviceSlope(di, extload / ninc)
            cg_tot = nl_tot = 0
            fmax = 0.0
            nl_iter = 0
            for inc in range(ninc):
                trial.begin_increment(ym, nuf, phif, cohf, psif, dt,
                                      stress[intact] if inc == 0 else None)
                if form == "mixed":  # the whole plastic loop on the engine
                    nl_iter, cg_inc, fmax = trial.run(cg_tol, cg_maxiter, nl_tol,
                                                      nl_maxiter)
                    cg_tot += cg_inc
                    x = trial.x()
                else:
                    for nl_iter in range(1, nl_maxiter + 1):
                        r = trial.plastic_iter(cg_tol, cg_maxiter, nl_tol,
                                               nl_iter == nl_maxiter)
                        cg_log.append(r["cg_iter"])
                        if r["errcode"] != 0:
                            raise PCGNotConverged("ERROR: PCG solver doesn't converge!")
                        cg_tot += r["cg_iter"]
                        fmax = r["fmax"]
                        if iter_progress:
                            iter_progress(i_ex + 1, inc + 1, nl_iter, r["cg_iter"])
                        if r["nl_isconv"] or nl_iter == nl_maxiter:
                            break
                    x = trial.x()
                nl_tot += nl_iter
                live = gdof_node != 0
                nodalu[live] += x[gdof_node[live]]
                if output is not None and output.wants_fields:
                    trial.accumulate_nodal(first=inc == 0)
            stress[intact] = trial.stress()
            un = nodalu[alive]
            rows.append(dict(stage=i_ex + 1, cg_tot=cg_tot, nl_tot=nl_tot,
                             uxmax=float(np.abs(un).max()),
                             umax=float(np.sqrt((un * un).sum(axis=1)).max()),
                             fmax=float(fmax)))
            if output is not None:  # src/excavation.F90:641-687
                output.stage_done(rows[-1], nodalu, np.nonzero(alive)[0], intact,
                                  fields=trial.nodal_fields)
        finally:
            if trial is not None:
                trial.close()
            di.close()
        if progress:
            progress(rows[-1])
        if nl_iter == nl_maxiter:  # src/excavation.F90:694
            break
    return ExcavationResult(rows=rows, cg_log=cg_log, nodalu=nodalu, stress=stress)
