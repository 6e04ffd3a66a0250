"""Host mirror of `slope3d` (src/slope.F90:13-450) over the device path.

The strength-reduction loop and its bookkeeping stay on the host exactly as
the reference orders them; each plastic iteration (load assembly, warm-
started PCG, convergence check, element x GLL stress update and body-load
assembly) runs on the GPU through libsgeo.so with all state resident.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib as L
from ._lib import check, f64, i32, ptr
from .mesh import HexMesh
from .plastic import dt_viscoplas, strength_reduction
from .solver import K_F64, Context, DeviceMesh, nodal_fields_call
from .spec import SpecMesh, hex2spec_device


class DeviceSlope:
    """State of one strength-reduction trial on the device (sgeo_slope)."""

    def __init__(self, dmesh: DeviceMesh, extload: np.ndarray):
        self.dmesh = dmesh
        self._ext = f64(extload)
        h = C.c_void_p()
        check(L.lib().sgeo_slope_create(dmesh.h, ptr(self._ext), C.byref(h)),
              "sgeo_slope_create")
        self.h = h

    def begin_srf(self, ym, nuf, phif, cohf, psif, dt: float):
        check(L.lib().sgeo_slope_begin_srf(self.h, ptr(f64(ym)), ptr(f64(nuf)),
                                           ptr(f64(phif)), ptr(f64(cohf)),
                                           ptr(f64(psif)), dt),
              "sgeo_slope_begin_srf")

    def begin_increment(self, ym, nuf, phif, cohf, psif, dt: float, stress0=None):
        """Excavation load increment: state reset, stress kept or set."""
        check(L.lib().sgeo_slope_begin_increment(
            self.h, ptr(f64(ym)), ptr(f64(nuf)), ptr(f64(phif)), ptr(f64(cohf)),
            ptr(f64(psif)), dt, ptr(f64(stress0)) if stress0 is not None else None),
            "sgeo_slope_begin_increment")

    def download_vmeps(self):
        vm = np.zeros((self.dmesh.spec.nelmt, 27))
        check(L.lib().sgeo_slope_download_vmeps(self.h, ptr(vm)),
              "sgeo_slope_download_vmeps")
        return vm

    def plastic_iter(self, cg_tol, cg_maxiter, nl_tol, is_last):
        it, ec, conv = C.c_int32(), C.c_int32(), C.c_int32()
        uerr, fmax = C.c_double(), C.c_double()
        check(L.lib().sgeo_slope_plastic_iter(self.h, cg_tol, cg_maxiter, nl_tol,
                                              int(is_last), C.byref(it), C.byref(ec),
                                              C.byref(uerr), C.byref(conv),
                                              C.byref(fmax)),
              "sgeo_slope_plastic_iter")
        return dict(cg_iter=it.value, errcode=ec.value, uerr=uerr.value,
                    nl_isconv=bool(conv.value), fmax=fmax.value)

    def download(self):
        sp = self.dmesh.spec
        x = np.zeros(sp.neq + 1)
        body = np.zeros(sp.neq + 1)
        evpt = np.zeros((sp.nelmt, 27, 6))
        stress = np.zeros((sp.nelmt, 27, 6))
        flags = np.zeros((sp.nelmt, 27), dtype=np.uint8)
        check(L.lib().sgeo_slope_download(self.h, ptr(x), ptr(body), ptr(evpt),
                                          ptr(stress), ptr(flags, C.c_uint8)),
              "sgeo_slope_download")
        return dict(x=x, bodyload=body, evpt=evpt, stress=stress, flags=flags)

    def accumulate_nodal(self, first: bool):
        """Excavation: gathers this increment's plastic strain and stress
        concentration factor at the nodes (sgeo_slope_accumulate_nodal)."""
        self.dmesh.ensure_nodes()
        check(L.lib().sgeo_slope_accumulate_nodal(self.h, int(first)),
              "sgeo_slope_accumulate_nodal")

    def nodal_fields(self) -> Dict[str, np.ndarray]:
        """The nodal result fields of the trial's stress state, computed on
        the device (sgeo_slope_nodal_fields)."""
        self.dmesh.ensure_nodes()
        return nodal_fields_call(L.lib().sgeo_slope_nodal_fields,
                                 "sgeo_slope_nodal_fields", self.dmesh.spec.nnode, self.h)

    def x(self) -> np.ndarray:
        """The trial's displacement x (neq+1)."""
        x = np.zeros(self.dmesh.spec.neq + 1)
        check(L.lib().sgeo_slope_download(self.h, ptr(x), None, None, None, None),
              "sgeo_slope_download")
        return x

    def stress(self) -> np.ndarray:
        """The trial's stress_elmt (nelmt, 27, 6)."""
        st = np.zeros((self.dmesh.spec.nelmt, 27, 6))
        check(L.lib().sgeo_slope_download(self.h, None, None, None, ptr(st), None),
              "sgeo_slope_download")
        return st

    def close(self):
        if self.h:
            L.lib().sgeo_slope_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class SlopeResult:
    rows: List[Dict[str, float]]          # summary lines (src/slope.F90:415)
    cg_log: List[int]                      # cg_iter of every PCG call
    fos_bracket: Optional[tuple]           # (last stable SRF, first failing SRF)
    x: np.ndarray                          # final displacement of the last SRF
    flags: np.ndarray                      # f>=0 at the last body pass
    extra: Dict[str, object] = field(default_factory=dict)


class PCGNotConverged(RuntimeError):
    """errcode -1 of pcg_solver; the reference stops (error_stop)."""


def uxmax_umax(spec: SpecMesh, x: np.ndarray):
    """src/slope.F90:401-413 on the host copy of x."""
    g = spec.gdof
    u = np.where(g > 0, x[g], 0.0)
    return float(np.abs(u).max()), float(np.sqrt(
        u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1] + u[:, 2] * u[:, 2]).max())


def slope3d(ctx: Context, mesh: HexMesh, *, cg_tol: float, cg_maxiter: int,
            nl_tol: float, nl_maxiter: int, srf: Sequence[float],
            phinu: bool = False, kprec: int = K_F64, kfmt: int = 0,
            spec: Optional[SpecMesh] = None, progress=None,
            iter_progress=None, on_srf=None, output=None) -> SlopeResult:
    """Strength-reduction sweep of src/slope.F90:244-443 on the GPU.

    kprec/kfmt select the operator the PCG streams (FP64 full reproduces
    the reference's arithmetic; packed halves the bytes).  `progress`, if
    given, is called with each summary row, `iter_progress` with (nl_iter,
    cg_iter) after each plastic iteration, `on_srf` with (row, x, plastic
    flags) at the end of each SRF.  `output` (output.SlopeOutput) receives
    each SRF's row, displacement and nodal fields for the reference's
    result files."""
    spec = spec or hex2spec_device(ctx, mesh)
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    dm = DeviceMesh(ctx, spec, len(gam))
    ext, _ = dm.build_stiffness(gam, ym, nu, kprec=K_F64)
    dm.set_operator(kprec, kfmt)
    st = DeviceSlope(dm, ext)
    rows: List[Dict[str, float]] = []
    cg_log: List[int] = []
    bracket = None
    prev = None
    state = None
    flags_per_srf: List[np.ndarray] = []
    try:
        for s in srf:
            cohf, nuf, phif, psif, istat = strength_reduction(s, phinu, coh, nu,
                                                               phi, psi)
            dt = dt_viscoplas(nuf, phif, ym)
            if istat == 1:  # src/slope.F90:259-270
                dm.build_stiffness(gam, ym, nuf, kprec=K_F64, want_extload=False)
                dm.set_operator(kprec, kfmt)
            st.begin_srf(ym, nuf, phif, cohf, psif, dt)
            cg_tot = 0
            fmax = 0.0
            nl_iter = 0
            for nl_iter in range(1, nl_maxiter + 1):
                r = st.plastic_iter(cg_tol, cg_maxiter, nl_tol,
                                    nl_iter == nl_maxiter)
                cg_log.append(r["cg_iter"])
                if iter_progress:
                    iter_progress(nl_iter, r["cg_iter"])
                if r["errcode"] != 0:
                    raise PCGNotConverged("ERROR: PCG solver doesn't converge!")
                cg_tot += r["cg_iter"]
                fmax = r["fmax"]
                if r["nl_isconv"] or nl_iter == nl_maxiter:
                    break
            state = st.download()
            flags_per_srf.append(state["flags"])
            ux, um = uxmax_umax(spec, state["x"])
            rows.append(dict(srf=float(s), cg_tot=cg_tot, nl_tot=nl_iter, uxmax=ux,
                             umax=um, fmax=fmax))
            if progress:
                progress(rows[-1])
            if on_srf:
                on_srf(rows[-1], state["x"], state["flags"])
            if output is not None:
                output.srf_done(len(rows), rows[-1], state["x"], fields=st.nodal_fields)
            if nl_iter == nl_maxiter:  # src/slope.F90:441
                bracket = (prev, float(s))
                break
            prev = float(s)
    finally:
        st.close()
        dm.close()
    return SlopeResult(rows=rows, cg_log=cg_log, fos_bracket=bracket,
                       x=state["x"] if state else None,
                       flags=state["flags"] if state else None,
                       extra=dict(flags_per_srf=flags_per_srf))
