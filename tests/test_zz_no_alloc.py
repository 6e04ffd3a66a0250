"""sgeo_pcg_solver_noalloc: the drop-in PCG without device allocation.

Device arrays in and out, every work array in the library's fixed scratch
array.  Checked against sgeo_pcg_solver (host arrays, cached device copies)
on the same systems, bit for bit: u, cg_iter, errcode, errtag, in both
reduction modes, on two calls in a row where the second is smaller (the
scratch keeps the first call's contents), on a warm-started call of the
reference's own sweep, on a call that runs out of iterations, and on the two
rejections (a size beyond the scratch, an equation number out of range).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import golden, mesh_from_golden
from specfem3d_geotech_amd import _lib as L
from specfem3d_geotech_amd.solver import (REDUCE_SEQUENTIAL, REDUCE_TREE,
                                          pcg_solver)
from specfem3d_geotech_amd.spec import hex2spec

pytestmark = pytest.mark.gpu


def _system(name, call):
    z = golden(name)
    mesh = mesh_from_golden(z)
    spec = hex2spec(mesh)
    gam, ym, nu, *_ = mesh.material_arrays()
    km, _, _ = O.stiffness_bodyload(spec, gam, nu, ym, with_extload=False)
    return dict(neq=spec.neq, nelmt=spec.nelmt, k=km, dinv=z["dinv"],
                g=spec.gdof_elmt.astype(np.int32), u=z["call_uin"][call],
                f=z["call_f"][call])


class _Dev:
    """Device buffers for the test, from the HIP runtime the library itself
    links (the caller's memory: the entry point under test allocates none)."""

    def __init__(self, device):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        assert self.hip.hipSetDevice(device) == 0
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # to host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _noalloc(ctx, s, cg_tol, cg_maxiter, u=None):
    dev = _Dev(ctx.device)
    try:
        u0 = np.ascontiguousarray(s["u"] if u is None else u, dtype=np.float64)
        d = {k: dev.put(s[k]) for k in ("k", "dinv", "g", "f")}
        du = dev.put(u0)
        it, ec = C.c_int32(), C.c_int32()
        tag = C.create_string_buffer(251)
        rc = L.lib().sgeo_pcg_solver_noalloc(
            ctx.h, s["neq"], s["nelmt"], d["k"], du, d["f"], d["dinv"], d["g"], cg_tol,
            cg_maxiter, C.byref(it), C.byref(ec), tag)
        L.check(rc, "sgeo_pcg_solver_noalloc")
        return dev.get(du, u0), it.value, ec.value, tag.raw[:250].decode().rstrip()
    finally:
        dev.free()


@pytest.mark.parametrize("mode", [REDUCE_TREE, REDUCE_SEQUENTIAL])
def test_noalloc_matches_pcg_solver(ctx, mode):
    ctx.set_reduction(mode)
    try:
        # the larger system first, then a smaller one over the same scratch
        for name, call in (("slope_8x4x6_layered", 1), ("slope_4x2x3", 5)):
            s = _system(name, call)
            need = L.lib().sgeo_pcg_solver_noalloc_bytes(s["neq"], s["nelmt"])
            assert 0 < need <= 64 << 20
            want = pcg_solver(ctx, s["neq"], s["nelmt"], s["k"], s["u"], s["f"],
                              s["dinv"], s["g"], 1e-8, 5000)
            for rep in range(2):  # twice: results do not depend on leftovers
                got = _noalloc(ctx, s, 1e-8, 5000)
                print(name, "mode", mode, "rep", rep, "cg_iter", got[1], want[1],
                      flush=True)
                assert got[1:] == want[1:]
                assert np.array_equal(got[0], want[0])
            # a call that runs out of iterations: errcode -1, same iterate
            want = pcg_solver(ctx, s["neq"], s["nelmt"], s["k"], s["u"], s["f"],
                              s["dinv"], s["g"], 1e-8, 3)
            got = _noalloc(ctx, s, 1e-8, 3)
            assert want[2] == -1 and got[1:] == want[1:]
            assert np.array_equal(got[0], want[0])
    finally:
        ctx.set_reduction(REDUCE_TREE)


def test_noalloc_rejects_what_it_cannot_hold(ctx):
    s = _system("slope_4x2x3", 0)
    # a size beyond the scratch is refused before the GPU is touched
    big = L.lib().sgeo_pcg_solver_noalloc_bytes(4_000_000, 200_000)
    assert big > 64 << 20
    it, ec = C.c_int32(), C.c_int32()
    tag = C.create_string_buffer(251)
    dev = _Dev(ctx.device)
    try:
        p = dev.put(np.zeros(8))
        rc = L.lib().sgeo_pcg_solver_noalloc(ctx.h, 4_000_000, 200_000, p, p, p, p, p,
                                             1e-8, 10, C.byref(it), C.byref(ec), tag)
    finally:
        dev.free()
    assert rc == -1 and b"scratch" in L.lib().sgeo_last_error()
    # an equation number outside 0..neq is found on the device and refused
    # before any product reads through it
    bad = dict(s)
    bad["g"] = s["g"].copy()
    bad["g"][3, 7] = s["neq"] + 5
    with pytest.raises(L.SgeoError, match="outside"):
        _noalloc(ctx, bad, 1e-8, 10)
    # and the next call is unaffected
    want = pcg_solver(ctx, s["neq"], s["nelmt"], s["k"], s["u"], s["f"], s["dinv"],
                      s["g"], 1e-8, 5000)
    got = _noalloc(ctx, s, 1e-8, 5000)
    assert got[1:] == want[1:] and np.array_equal(got[0], want[0])
