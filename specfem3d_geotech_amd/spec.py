"""Spectral-element numbering: hex8 -> 27-node GLL elements and equations.

Host-side restatement of the reference's pre-processing so that the arrays
handed to the hot path are the ones the reference builds:

* `hex2spec` (src/mesh_spec.f90:17-143): GLL points of every element in
  local (k, j, i) order, i fastest; coincident points are merged by the
  segment-wise lexicographic sort of `get_global` (src/mesh_spec.f90:146-262,
  tolerance 1e-10*|xmax-xmin|), then renumbered by first appearance
  (`get_global_indirect_addressing`, :357-386).
* `apply_bc` (src/apply_bc.f90:36-137): face BCs zero `gdof`, then the
  remaining dofs are numbered node-major, dof fastest.
* `gdof_elmt(:,e) = reshape(gdof(:,g_num(:,e)), nedof)`
  (src/slope.F90:139-143): element dof order is node-major, x/y/z fastest.

Index 0 of every global vector is the sink for constrained dofs.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import HexMesh

NGLL1 = 3
NGLL = 27
NEDOF = 81

# EXODUS corner -> local GLL node (src/element.f90:32-40), 0-based
HEX8_GNODE = np.array([0, 2, 8, 6, 18, 20, 26, 24], dtype=np.int64)

# GLL nodes of the 6 faces, local 0-based (src/element.f90:64-100):
# 1: j==1, 2: i==ngllx, 3: j==nglly, 4: i==1, 5: k==1, 6: k==ngllz
_I = np.arange(NGLL) % 3
_J = (np.arange(NGLL) // 3) % 3
_K = np.arange(NGLL) // 9
HEXFACE = {1: np.nonzero(_J == 0)[0], 2: np.nonzero(_I == 2)[0],
           3: np.nonzero(_J == 2)[0], 4: np.nonzero(_I == 0)[0],
           5: np.nonzero(_K == 0)[0], 6: np.nonzero(_K == 2)[0]}


def _shape_hex8_at_gll() -> np.ndarray:
    """shape_hex8(8, 27) at GLL points {-1, 0, 1} (src/shape_library.f90:14-98,
    middle abscissa forced to zero as in src/mesh_spec.f90:68-70)."""
    g = np.array([-1.0, 0.0, 1.0])
    sh = np.empty((8, NGLL))
    for n in range(NGLL):
        xi, eta, zeta = g[_I[n]], g[_J[n]], g[_K[n]]
        xip, xim = 1.0 + xi, 1.0 - xi
        etap, etam = 1.0 + eta, 1.0 - eta
        zp, zm = 1.0 + zeta, 1.0 - zeta
        sh[:, n] = 0.125 * np.array([xim * etam * zm, xip * etam * zm,
                                     xip * etap * zm, xim * etap * zm,
                                     xim * etam * zp, xip * etam * zp,
                                     xip * etap * zp, xim * etap * zp])
    return sh


def _group_points(p: np.ndarray, tol: float) -> np.ndarray:
    """Group ids of coincident points, segment-wise lexicographic as in
    get_global (src/mesh_spec.f90:190-240)."""
    npt = p.shape[0]
    seg = np.zeros(npt, dtype=np.int64)
    for d in range(3):
        order = np.lexsort((p[:, d], seg))
        s, v = seg[order], p[order, d]
        brk = np.ones(npt, dtype=bool)
        brk[1:] = (s[1:] != s[:-1]) | (np.abs(np.diff(v)) > tol)
        newseg = np.cumsum(brk) - 1
        seg = np.empty(npt, dtype=np.int64)
        seg[order] = newseg
    return seg


@dataclass
class SpecMesh:
    """Arrays of the SEM discretisation (0-based node ids, 1-based equations
    with 0 = constrained sink, like the reference)."""
    nelmt: int
    nnode: int
    neq: int
    g_num: np.ndarray        # (nelmt, 27) int64 node ids, 0-based
    g_coord: np.ndarray      # (nnode, 3)
    gdof: np.ndarray         # (nnode, 3) int64, 0 = constrained
    gdof_elmt: np.ndarray    # (nelmt, 81) int32 (row e = gdof_elmt(:,e))
    corners: np.ndarray      # (nelmt, 8, 3) hex8 corner coordinates
    mat_id: np.ndarray       # (nelmt,) int32 1-based

    def inverse_map(self):
        """CSR map equation -> element dof slots in ascending element order.

        Returns (rowptr[neq+2], slots[int32]) where slots hold e*81+i; the
        equation g's contributions are slots[rowptr[g]:rowptr[g+1]].  The
        order reproduces the reference's element-order accumulation
        `kp(egdof)=kp(egdof)+...` (src/solver.f90:54-58)."""
        flat = self.gdof_elmt.reshape(-1).astype(np.int64)
        order = np.argsort(flat, kind="stable")
        counts = np.bincount(flat, minlength=self.neq + 1)
        rowptr = np.zeros(self.neq + 2, dtype=np.int64)
        rowptr[1:] = np.cumsum(counts)
        return rowptr.astype(np.int32), order.astype(np.int32)


def hex2spec(mesh: HexMesh) -> SpecMesh:
    """Build the 27-node SEM mesh and equation numbering of `mesh`."""
    coord = np.asarray(mesh.coord, dtype=np.float64)
    conn0 = np.asarray(mesh.conn, dtype=np.int64) - 1
    nel = conn0.shape[0]
    sh = _shape_hex8_at_gll()                           # (8, 27)
    corners = coord[conn0]                              # (nel, 8, 3)
    pts = np.einsum("gn,egd->end", sh, corners).reshape(-1, 3)
    xmin, xmax = coord[:, 0].min(), coord[:, 0].max()
    grp = _group_points(pts, 1e-10 * abs(xmax - xmin))
    # first-appearance renumbering over (element, local node)
    _, first = np.unique(grp, return_index=True)
    rank = np.empty_like(first)
    rank[np.argsort(first, kind="stable")] = np.arange(first.size)
    remap = np.empty(grp.max() + 1, dtype=np.int64)
    remap[np.unique(grp)] = rank
    node_of_pt = remap[grp]
    nnode = int(first.size)
    g_num = node_of_pt.reshape(nel, NGLL)
    g_coord = np.empty((nnode, 3))
    g_coord[node_of_pt] = pts          # later points overwrite, as in :119-136

    gdof = np.ones((nnode, 3), dtype=np.int64)
    for comp, name in enumerate(("ux", "uy", "uz")):
        pairs = np.asarray(mesh.bc[name], dtype=np.int64).reshape(-1, 2)
        for face in range(1, 7):
            sel = pairs[pairs[:, 1] == face, 0] - 1
            if sel.size:
                gdof[g_num[sel][:, HEXFACE[face]].ravel(), comp] = 0
    free = gdof.ravel() != 0
    numbered = np.zeros(gdof.size, dtype=np.int64)
    numbered[free] = np.arange(1, int(free.sum()) + 1)
    gdof = numbered.reshape(nnode, 3)
    neq = int(free.sum())
    gdof_elmt = gdof[g_num].reshape(nel, NEDOF).astype(np.int32)
    return SpecMesh(nelmt=nel, nnode=nnode, neq=neq, g_num=g_num,
                    g_coord=g_coord, gdof=gdof, gdof_elmt=gdof_elmt,
                    corners=np.ascontiguousarray(corners),
                    mat_id=np.asarray(mesh.mat_id, dtype=np.int32))


def hex2spec_device(ctx, mesh: HexMesh) -> SpecMesh:
    """`hex2spec` on the GPU (sgeo_hex2spec): the same numbering, built by
    radix sorts and scans on the device instead of host sorts."""
    import ctypes as C

    from . import _lib as L
    coord = np.ascontiguousarray(mesh.coord, dtype=np.float64)
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    nel = conn.shape[0]
    rows = []
    for comp, name in enumerate(("ux", "uy", "uz")):
        pairs = np.asarray(mesh.bc[name], dtype=np.int32).reshape(-1, 2)
        rows.append(np.column_stack([pairs, np.full(len(pairs), comp, np.int32)]))
    bc = np.ascontiguousarray(np.concatenate(rows), dtype=np.int32)
    h = C.c_void_p()
    nnode, neq = C.c_int32(), C.c_int32()
    L.check(L.lib().sgeo_hex2spec(ctx.h, nel, coord.shape[0], L.ptr(coord),
                                  L.ptr(conn, C.c_int32), len(bc), L.ptr(bc, C.c_int32),
                                  C.byref(h), C.byref(nnode), C.byref(neq)),
            "sgeo_hex2spec")
    try:
        nn = nnode.value
        g_num = np.empty((nel, NGLL), np.int32)
        gdof = np.empty((nn, 3), np.int32)
        gdof_elmt = np.empty((nel, NEDOF), np.int32)
        g_coord = np.empty((nn, 3))
        corners = np.empty((nel, 8, 3))
        L.check(L.lib().sgeo_spec_fetch(h, L.ptr(g_num, C.c_int32), L.ptr(gdof, C.c_int32),
                                        L.ptr(gdof_elmt, C.c_int32), L.ptr(g_coord),
                                        L.ptr(corners)), "sgeo_spec_fetch")
    finally:
        L.lib().sgeo_spec_destroy(h)
    return SpecMesh(nelmt=nel, nnode=nn, neq=neq.value, g_num=g_num.astype(np.int64),
                    g_coord=g_coord, gdof=gdof.astype(np.int64), gdof_elmt=gdof_elmt,
                    corners=corners, mat_id=np.asarray(mesh.mat_id, dtype=np.int32))
