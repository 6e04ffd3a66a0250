"""Host mirror of the reference's `solver` module and the resident-mesh API.

`pcg_solver` keeps the reference's interface and error behaviour
(src/solver.f90:14-79): arrays in the reference layout, warm-started `u`,
`cg_iter`, `errcode` (0 / -1) and `errtag`.  The solve itself runs in
libsgeo.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Tuple

import numpy as np

from . import _lib as L
from ._lib import check, f64, i32, ptr
from .spec import SpecMesh

K_F64 = 0
K_F32 = 1
REDUCE_TREE = 0
REDUCE_SEQUENTIAL = 1
KFMT_FULL = 0
KFMT_PACKED = 1


class Context:
    """One HIP device + stream (sgeo_init)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(L.lib().sgeo_init(device, C.byref(h)), "sgeo_init")
        self.h = h
        self.device = device

    def info(self):
        name = C.create_string_buffer(256)
        ncu = C.c_int()
        hbm = C.c_size_t()
        check(L.lib().sgeo_device_info(self.h, name, 256, C.byref(ncu), C.byref(hbm)),
              "sgeo_device_info")
        return dict(name=name.value.decode(), n_cu=ncu.value, hbm_bytes=hbm.value)

    def synchronize(self):
        check(L.lib().sgeo_synchronize(self.h), "sgeo_synchronize")

    def profile(self, enable: int = -1):
        """(launches, total_ms) of the timed element-product kernel since the
        last reset; enable 1/0 then resets and switches timing on/off."""
        n = C.c_int64()
        ms = C.c_double()
        check(L.lib().sgeo_profile(self.h, enable, C.byref(n), C.byref(ms)),
              "sgeo_profile")
        return n.value, ms.value

    def calibrate_copy(self, nbytes: int = 4 << 30, reps: int = 10) -> float:
        """GB/s (read + write) of a plain device copy of `nbytes`."""
        g = C.c_double()
        check(L.lib().sgeo_calibrate_copy(self.h, nbytes, reps, C.byref(g)),
              "sgeo_calibrate_copy")
        return g.value

    def set_reduction(self, mode: int):
        """REDUCE_TREE (fast) or REDUCE_SEQUENTIAL (reference order)."""
        check(L.lib().sgeo_set_reduction(self.h, mode), "sgeo_set_reduction")

    def close(self):
        if self.h:
            L.lib().sgeo_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pcg_solver_invalidate(ctx: Context) -> None:
    """Drops the device copies pcg_solver caches on the context (keyed on the
    host arrays' addresses and a strided fingerprint of k); needed after k,
    dprecon or gdof_elmt were rewritten in place."""
    check(L.lib().sgeo_pcg_solver_invalidate(ctx.h), "sgeo_pcg_solver_invalidate")


def pcg_solver(ctx: Context, neq: int, nelmt: int, k: np.ndarray, u: np.ndarray,
               f: np.ndarray, dprecon: np.ndarray, gdof_elmt: np.ndarray,
               cg_tol: float, cg_maxiter: int) -> Tuple[np.ndarray, int, int, str]:
    """Drop-in for pcg_solver: returns (u, cg_iter, errcode, errtag).

    `k` is storekm in the reference layout (nelmt x 81 x 81 with
    k[e, c, r] = k(r+1, c+1, e+1), i.e. column-major per element),
    `dprecon` the inverted diagonal, `gdof_elmt` (nelmt x 81)."""
    k = f64(k)
    if k.size != nelmt * 6561:
        raise ValueError("k must hold 81*81*nelmt entries")
    u = f64(u).copy()
    f = f64(f)
    dprecon = f64(dprecon)
    g = i32(gdof_elmt)
    if u.size != neq + 1 or f.size != neq + 1 or dprecon.size != neq + 1:
        raise ValueError("vectors must have neq+1 entries")
    if g.size != nelmt * 81:
        raise ValueError("gdof_elmt must hold 81*nelmt entries")
    it = C.c_int32()
    ec = C.c_int32()
    tag = C.create_string_buffer(251)
    check(L.lib().sgeo_pcg_solver(ctx.h, neq, nelmt, ptr(k), ptr(u), ptr(f),
                                  ptr(dprecon), ptr(g, C.c_int32), cg_tol,
                                  cg_maxiter, C.byref(it), C.byref(ec), tag),
          "sgeo_pcg_solver")
    return u, it.value, ec.value, tag.raw[:250].decode().rstrip()


def nodal_fields_call(fn, what: str, nnode: int, *args,
                      with_scf_vmeps: bool = True) -> Dict[str, np.ndarray]:
    """Runs one of the sgeo_*_nodal_fields entry points: FP32 stress
    (nnode, 6), psigma (nnode, 3), nsigma, and scf and vmeps (nnode,), as
    the reference's EnSight files hold them."""
    sig = np.zeros((6, nnode), np.float32)
    ps = np.zeros((3, nnode), np.float32)
    ns = np.zeros(nnode, np.float32)
    fp = lambda a: ptr(a, C.c_float)
    out = dict(stress=sig.T, psigma=ps.T, nsigma=ns)
    if with_scf_vmeps:
        sc = np.zeros(nnode, np.float32)
        vm = np.zeros(nnode, np.float32)
        check(fn(*args, fp(sig), fp(ps), fp(ns), fp(sc), fp(vm)), what)
        out.update(scf=sc, vmeps=vm)
    else:
        check(fn(*args, fp(sig), fp(ps), fp(ns)), what)
    return out


class DeviceMesh:
    """A mesh resident in HBM (sgeo_mesh): gdof_elmt, corners, materials and
    the element stiffness."""

    def __init__(self, ctx: Context, spec: SpecMesh, nmat: int):
        self.ctx = ctx
        self.spec = spec
        self.nmat = nmat
        h = C.c_void_p()
        self._corners = f64(spec.corners)
        self._gdof = i32(spec.gdof_elmt)
        self._mat = i32(spec.mat_id)
        check(L.lib().sgeo_mesh_create(ctx.h, spec.nelmt, spec.neq, ptr(self._corners),
                                       ptr(self._gdof, C.c_int32),
                                       ptr(self._mat, C.c_int32), nmat, C.byref(h)),
              "sgeo_mesh_create")
        self.h = h
        self.kprec: Optional[int] = None

    def solve(self, f, u0=None, cg_tol=1e-8, cg_maxiter=5000):
        """pcg_solver on the resident operator; returns (u, cg_iter, errcode)."""
        n1 = self.spec.neq + 1
        u = np.zeros(n1) if u0 is None else f64(u0).copy()
        it, ec = C.c_int32(), C.c_int32()
        check(L.lib().sgeo_mesh_solve(self.h, ptr(u), ptr(f64(f)), cg_tol, cg_maxiter,
                                      C.byref(it), C.byref(ec)), "sgeo_mesh_solve")
        return u, it.value, ec.value

    def ensure_nodes(self):
        """The node map of the nodal result fields (built once)."""
        if not getattr(self, "_nodes_set", False):
            g = i32(self.spec.g_num)
            check(L.lib().sgeo_mesh_set_nodes(self.h, self.spec.nnode, ptr(g, C.c_int32)),
                  "sgeo_mesh_set_nodes")
            self._nodes_set = True

    def nodal_fields(self, stress) -> Dict[str, np.ndarray]:
        """Nodal stress, principal and normal stress of a host stress state
        (nelmt, 27, 6) on this mesh, computed on the device."""
        self.ensure_nodes()
        st = f64(stress)
        return nodal_fields_call(L.lib().sgeo_mesh_nodal_fields, "sgeo_mesh_nodal_fields",
                                 self.spec.nnode, self.h, ptr(st), with_scf_vmeps=False)

    def elastic_stress(self, ym, nu, x):
        """(nelmt, 27, 6) elastic stress at the GLL points from x."""
        st = np.zeros((self.spec.nelmt, 27, 6))
        check(L.lib().sgeo_mesh_elastic_stress(self.h, ptr(f64(ym)), ptr(f64(nu)),
                                               ptr(f64(x)), ptr(st)),
              "sgeo_mesh_elastic_stress")
        return st

    def excavation_load(self, gam, stress):
        """Excavation load (neq+1) of this mesh's elements from their stress
        (nelmt, 27, 6)."""
        load = np.zeros(self.spec.neq + 1)
        check(L.lib().sgeo_mesh_excavation_load(self.h, ptr(f64(gam)), ptr(f64(stress)),
                                                ptr(load)), "sgeo_mesh_excavation_load")
        return load

    def build_stiffness(self, gam, ym, nu, kprec: int = K_F64,
                        want_extload: bool = True):
        """Device stiffness_bodyload; returns (extload, dinv)."""
        n1 = self.spec.neq + 1
        ext = np.zeros(n1) if want_extload else None
        dinv = np.zeros(n1)
        check(L.lib().sgeo_mesh_build_stiffness(self.h, ptr(f64(gam)), ptr(f64(ym)),
                                                ptr(f64(nu)), kprec, ptr(ext),
                                                ptr(dinv)),
              "sgeo_mesh_build_stiffness")
        self.kprec = kprec
        return ext, dinv

    def upload_stiffness(self, storekm, dinv, kprec: int = K_F64):
        check(L.lib().sgeo_mesh_upload_stiffness(self.h, ptr(f64(storekm)),
                                                 ptr(f64(dinv)), kprec),
              "sgeo_mesh_upload_stiffness")
        self.kprec = kprec

    def set_operator(self, kprec: int = K_F64, kfmt: int = KFMT_FULL):
        """Operator streamed by the PCG: precision and storage format."""
        check(L.lib().sgeo_mesh_set_operator(self.h, kprec, kfmt),
              "sgeo_mesh_set_operator")
        self.kprec = kprec

    def download_stiffness(self) -> np.ndarray:
        k = np.zeros((self.spec.nelmt, 81, 81))
        check(L.lib().sgeo_mesh_download_stiffness(self.h, ptr(k)),
              "sgeo_mesh_download_stiffness")
        return k

    def matvec(self, p) -> np.ndarray:
        kp = np.zeros(self.spec.neq + 1)
        check(L.lib().sgeo_mesh_matvec(self.h, ptr(f64(p)), ptr(kp)), "sgeo_mesh_matvec")
        return kp

    def nbytes(self) -> int:
        b = C.c_size_t()
        check(L.lib().sgeo_mesh_bytes(self.h, C.byref(b)), "sgeo_mesh_bytes")
        return b.value

    def close(self):
        if self.h:
            L.lib().sgeo_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
