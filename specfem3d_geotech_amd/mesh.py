"""Seeded hex8 mesh generation and the reference's plain-text input format.

The reference reads a `.sem` control file plus one file per mesh array
(src/read_input.f90:123-409 for the control keys, :460-614 for the files):

* ``<head>_coord_x|y|z``  -- ``nnode`` then one coordinate per line
* ``<head>_connectivity`` -- ``nelmt`` then 8 node ids per line (EXODUS order,
  src/element.f90:32-40: bottom face counter-clockwise, then top face)
* ``<head>_material_id``  -- ``nelmt`` then one id per line
* ``<head>_material_list`` -- a comment line, ``nmatblk``, then
  ``id, domain, type, gamma, ym, nu, phi, coh, psi`` per block
* ``<head>_ssbcux|y|z``   -- ``2 0.0`` (face BC, zero value), count, then
  ``elem face`` pairs; faces 1..6 = y-min, x-max, y-max, x-min, z-min, z-max
  of the element's local axes (src/element.f90:69-105, src/apply_bc.f90:36-58)

Everything here is host-side set-up: it produces the same arrays the
reference's driver holds before the hot path starts.
"""
from __future__ import annotations

import dataclasses

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np


@dataclass
class Material:
    """One material block (src/read_input.f90:572-585); phi/psi in degrees."""
    gam: float
    ym: float
    nu: float
    phi: float
    coh: float
    psi: float = 0.0


@dataclass
class HexMesh:
    """A hex8 mesh in the reference's input layout (1-based ids)."""
    coord: np.ndarray            # (nnode, 3) float64
    conn: np.ndarray             # (nelmt, 8) int32, 1-based, EXODUS order
    mat_id: np.ndarray           # (nelmt,) int32, 1-based
    materials: List[Material]
    bc: Dict[str, np.ndarray]    # 'ux'|'uy'|'uz' -> (n, 2) int32 (elem, face)
    meta: Dict[str, object] = field(default_factory=dict)

    @property
    def nelmt(self) -> int:
        return int(self.conn.shape[0])

    @property
    def nnode(self) -> int:
        return int(self.coord.shape[0])

    def material_arrays(self):
        """(gam, ym, nu, phi, coh, psi) float64 arrays over material blocks."""
        m = self.materials
        return tuple(np.array([getattr(b, k) for b in m], dtype=np.float64)
                     for k in ("gam", "ym", "nu", "phi", "coh", "psi"))

    def with_strengths(self, coh, phi) -> "HexMesh":
        """The same mesh with the materials' cohesion and friction angle
        replaced (per material block)."""
        mats = [dataclasses.replace(b, coh=float(coh[i]), phi=float(phi[i]))
                for i, b in enumerate(self.materials)]
        return dataclasses.replace(self, materials=mats)


def slope_surface(x: np.ndarray, *, z_crest: float, z_toe: float,
                  x_crest: float, angle_deg: float) -> np.ndarray:
    """Ground height of a planar slope: crest plateau, face at `angle_deg`,
    toe plateau."""
    if z_crest == z_toe:  # flat ground (box meshes)
        return np.full(np.shape(x), float(z_crest))
    run = (z_crest - z_toe) / math.tan(math.radians(angle_deg))
    t = np.clip((x - x_crest) / run, 0.0, 1.0)
    return z_crest - t * (z_crest - z_toe)


def slope_mesh(nx: int, ny: int, nz: int, *, lx: float = 40.0,
               ly: float = 20.0, z_crest: float = 20.0, z_toe: float = 5.0,
               x_crest: float = 5.0, angle_deg: float = math.degrees(math.atan(0.5)),
               materials: Optional[Sequence[Material]] = None,
               layer_z: Sequence[float] = (),
               jitter: float = 0.0, seed: int = 0) -> HexMesh:
    """Structured hex8 slope mesh with column-wise vertical mapping.

    Node (i, j, k) sits at x = lx*i/nx, y = ly*j/ny, z = ztop(x)*k/nz, where
    ztop is the slope surface.  The default geometry is the 2H:1V slope of
    the plain CPU configuration: crest z=20 for x in [0,5], face down to z=5
    over x in [5,35], toe plateau to x=40, box 40 x 20 m.

    `layer_z` splits materials by element-centroid height: material 1 above
    the highest split, material n+1 below the lowest.  `jitter` moves every
    strictly interior node by U[-jitter, jitter] times the local spacing per
    axis (numpy default_rng(seed)); boundary nodes stay put so the fixed and
    roller faces remain planar.

    Boundary conditions: base (z-min) fixed in x, y, z; x-min/x-max faces
    rollers in x; y-min/y-max faces rollers in y.
    """
    if materials is None:
        materials = [Material(gam=20.0, ym=1e5, nu=0.3, phi=20.0, coh=10.0)]
    materials = list(materials)
    if len(materials) != len(layer_z) + 1:
        raise ValueError("need len(layer_z)+1 materials")

    xi = lx * np.arange(nx + 1) / nx
    yj = ly * np.arange(ny + 1) / ny
    kk = np.arange(nz + 1) / nz
    ztop = slope_surface(xi, z_crest=z_crest, z_toe=z_toe, x_crest=x_crest,
                         angle_deg=angle_deg)
    # node id (i,j,k) -> i + (nx+1)*(j + (ny+1)*k)
    K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1),
                          np.arange(nx + 1), indexing="ij")
    X = xi[I].astype(np.float64)
    Y = yj[J].astype(np.float64)
    Z = (ztop[I] * kk[K]).astype(np.float64)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        interior = ((I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz))
        shape = X.shape
        d = rng.uniform(-jitter, jitter, size=(3,) + shape)
        X = X + np.where(interior, d[0] * (lx / nx), 0.0)
        Y = Y + np.where(interior, d[1] * (ly / ny), 0.0)
        Z = Z + np.where(interior, d[2] * (ztop[I] / nz), 0.0)
    coord = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    ek, ej, ei = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx),
                             indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()
    conn = np.stack([nid(ei, ej, ek), nid(ei + 1, ej, ek),
                     nid(ei + 1, ej + 1, ek), nid(ei, ej + 1, ek),
                     nid(ei, ej, ek + 1), nid(ei + 1, ej, ek + 1),
                     nid(ei + 1, ej + 1, ek + 1), nid(ei, ej + 1, ek + 1)],
                    axis=1).astype(np.int32) + 1

    zc = coord[conn - 1, 2].mean(axis=1)
    mat = np.ones(conn.shape[0], dtype=np.int32)
    for z_split in sorted(layer_z, reverse=True):
        mat += (zc < z_split).astype(np.int32)

    eid = np.arange(conn.shape[0], dtype=np.int32) + 1

    def faces(mask, face):
        sel = eid[mask]
        return np.stack([sel, np.full_like(sel, face)], axis=1)

    base = faces(ek == 0, 5)
    xfaces = np.concatenate([faces(ei == 0, 4), faces(ei == nx - 1, 2)])
    yfaces = np.concatenate([faces(ej == 0, 1), faces(ej == ny - 1, 3)])
    bc = {"ux": np.concatenate([base, xfaces]).astype(np.int32),
          "uy": np.concatenate([base, yfaces]).astype(np.int32),
          "uz": base.astype(np.int32)}
    meta = dict(kind="slope", nx=nx, ny=ny, nz=nz, lx=lx, ly=ly,
                z_crest=z_crest, z_toe=z_toe, x_crest=x_crest,
                angle_deg=angle_deg, layer_z=list(layer_z), jitter=jitter,
                seed=seed)
    return HexMesh(coord=coord, conn=conn, mat_id=mat, materials=materials,
                   bc=bc, meta=meta)


def box_mesh(nx: int, ny: int, nz: int, *, lx: float, ly: float, lz: float,
             materials: Sequence[Material], region_of=None) -> HexMesh:
    """Structured box (base fixed, side rollers).  `region_of(xc, yc, zc)`
    returns a 1-based material id per element centroid (default: all 1)."""
    m = slope_mesh(nx, ny, nz, lx=lx, ly=ly, z_crest=lz, z_toe=lz, x_crest=lx,
                   angle_deg=45.0, materials=[materials[0]])
    m.materials = list(materials)
    if region_of is not None:
        c = m.coord[m.conn - 1].mean(axis=1)
        m.mat_id = np.asarray(region_of(c[:, 0], c[:, 1], c[:, 2]),
                              dtype=np.int32)
    m.meta.update(kind="box", lz=lz)
    return m


def pit_mesh(nx: int, ny: int, nz: int, *, lx: float, ly: float, lz: float,
             pit: Sequence[float], nstage: int, stage_depth: float,
             soil: Material) -> HexMesh:
    """Box with a rectangular pit excavated from the top in `nstage` layers
    of `stage_depth`: elements whose centroid lies in the pit footprint
    pit = (x0, x1, y0, y1) and in layer s (counted from the top) get
    material id s + 2; all others id 1.  Every id carries the same soil, as
    in the reference's excavation examples (input/excavation_3d*).  Stage s
    removes id s + 2 (control keys nexcav = nstage, excavid = 2 .. nstage+1).
    """
    x0, x1, y0, y1 = pit

    def region(xc, yc, zc):
        inside = (xc > x0) & (xc < x1) & (yc > y0) & (yc < y1)
        layer = np.floor((lz - zc) / stage_depth).astype(np.int64)
        return np.where(inside & (layer < nstage), layer + 2, 1)

    m = box_mesh(nx, ny, nz, lx=lx, ly=ly, lz=lz,
                 materials=[soil] * (nstage + 1), region_of=region)
    m.meta.update(kind="pit", nstage=nstage)
    return m


# ----------------------------------------------------------------------------
# reference input format
# ----------------------------------------------------------------------------

def _fmt_list(vals) -> str:
    return " ".join(repr(float(v)) if isinstance(v, float) else str(v)
                    for v in vals)


def write_reference_input(mesh: HexMesh, out_dir: str, head: str, *,
                          cg_tol: float, cg_maxiter: int, nl_tol: float,
                          nl_maxiter: int, srf: Sequence[float],
                          out_path: Optional[str] = None,
                          extra_control: str = "",
                          extra_lines: str = "", save_disp: bool = False,
                          save: Sequence[str] = ()) -> str:
    """Write `mesh` plus a `.sem` control file readable by the reference's
    `semgeotech` (serial); `save_disp` asks it for its EnSight displacement
    output, `save` for any of its result fields by their `save:` keys
    (disp, stress, psigma, nsigma, scf, vmeps; src/read_input.f90:380-403).
    Returns the `.sem` path."""
    os.makedirs(out_dir, exist_ok=True)
    out_path = out_path or os.path.join(out_dir, "output")
    os.makedirs(out_path, exist_ok=True)
    p = lambda s: os.path.join(out_dir, f"{head}_{s}")
    for a, name in enumerate(("coord_x", "coord_y", "coord_z")):
        with open(p(name), "w") as fh:
            fh.write(f"{mesh.nnode}\n")
            fh.write("\n".join("%.17e" % v for v in mesh.coord[:, a]))
            fh.write("\n")
    with open(p("connectivity"), "w") as fh:
        fh.write(f"{mesh.nelmt}\n")
        for row in mesh.conn:
            fh.write(" ".join(str(int(v)) for v in row) + "\n")
    with open(p("material_id"), "w") as fh:
        fh.write(f"{mesh.nelmt}\n")
        fh.write("\n".join(str(int(v)) for v in mesh.mat_id) + "\n")
    with open(p("material_list"), "w") as fh:
        fh.write("# material properties (id,domain,type,gamma,ym,nu,phi,coh,psi)\n")
        fh.write(f"{len(mesh.materials)}\n")
        for i, m in enumerate(mesh.materials):
            v = [float(a) for a in (m.gam, m.ym, m.nu, m.phi, m.coh, m.psi)]
            fh.write(f"{i + 1}, 1, 0, " + ", ".join(repr(a) for a in v) + "\n")
    for comp in ("ux", "uy", "uz"):
        with open(p(f"ssbc{comp}"), "w") as fh:
            pairs = mesh.bc[comp]
            fh.write("2 0.0\n")
            fh.write(f"{len(pairs)}\n")
            for e, f in pairs:
                fh.write(f"{int(e)} {int(f)}\n")
    srf_s = " ".join(repr(float(s)) for s in srf)
    sem = os.path.join(out_dir, f"{head}.sem")
    with open(sem, "w") as fh:
        fh.write("preinfo: method='sem', ngllx=3, nglly=3, ngllz=3, nenod=8, "
                 f"ngnode=8, inp_path='{out_dir}/', out_path='{out_path}/'\n")
        fh.write(f"mesh: xfile='{head}_coord_x', yfile='{head}_coord_y', "
                 f"zfile='{head}_coord_z', &\n"
                 f"confile='{head}_connectivity', idfile='{head}_material_id'\n")
        fh.write(f"bc: uxfile='{head}_ssbcux', uyfile='{head}_ssbcuy', "
                 f"uzfile='{head}_ssbcuz'\n")
        fh.write(f"material: matfile='{head}_material_list', matpath='{out_dir}/'\n")
        fh.write(f"control: cg_tol={float(cg_tol)!r}, cg_maxiter={int(cg_maxiter)}, "
                 f"nl_tol={float(nl_tol)!r}, nl_maxiter={int(nl_maxiter)}, nsrf={len(srf)}, "
                 f"srf={srf_s}{extra_control}\n")
        if extra_lines:
            fh.write(extra_lines)
        keys = list(save) or ["disp"]
        flags = {k: 1 for k in save}
        flags.setdefault("disp", 1 if save_disp else 0)
        fh.write("save: " + ", ".join(f"{k}={flags[k]}" for k in keys) + "\n")
    return sem


def _read_table(path, dtype, ncol=None):
    with open(path) as fh:
        n = int(fh.readline().split()[0])
        data = np.loadtxt(fh, dtype=dtype, max_rows=n, ndmin=1 if ncol is None else 2)
    return data


def read_reference_input(sem_path: str) -> tuple:
    """Read a serial `.sem` file and the mesh files it names.

    Supports the keys the slope/excavation paths use (preinfo, mesh, bc,
    material, control, stress0, save); values follow
    src/read_input.f90:123-409.  Returns (HexMesh, control dict); the
    control dict also holds `save` (the saved fields' keys) and `out_path`
    (as written, relative to the working directory as the reference takes
    it)."""
    base_dir = os.path.dirname(os.path.abspath(sem_path))
    lines: List[str] = []
    with open(sem_path) as fh:
        buf = ""
        for raw in fh:
            s = raw.split("#", 1)[0].rstrip()
            if not s.strip():
                continue
            if s.endswith("&"):
                buf += s[:-1] + " "
                continue
            lines.append(buf + s)
            buf = ""
    kv: Dict[str, Dict[str, str]] = {}
    for ln in lines:
        tok, rest = ln.split(":", 1)
        d: Dict[str, str] = {}
        for part in rest.split(","):
            if "=" in part:
                k, v = part.split("=", 1)
                d[k.strip()] = v.strip().strip("'\"")
        kv[tok.strip()] = d
    pre = kv.get("preinfo", {})
    inp = pre.get("inp_path", "./input")
    if not os.path.isabs(inp):
        inp = os.path.join(base_dir, "..", inp) if not os.path.isdir(
            os.path.join(base_dir, inp)) else os.path.join(base_dir, inp)
    mesh_kv, bc_kv, mat_kv = kv["mesh"], kv["bc"], kv["material"]
    cx = [_read_table(os.path.join(inp, mesh_kv[k]), np.float64)
          for k in ("xfile", "yfile", "zfile")]
    coord = np.stack(cx, axis=1)
    conn = _read_table(os.path.join(inp, mesh_kv["confile"]), np.int32, 8)
    mat_id = _read_table(os.path.join(inp, mesh_kv["idfile"]), np.int32)
    mpath = mat_kv.get("matpath", inp)
    mats: List[Material] = []
    with open(os.path.join(mpath, mat_kv["matfile"])) as fh:
        fh.readline()
        nmat = int(fh.readline().split()[0])
        for _ in range(nmat):
            f = [t.strip() for t in fh.readline().replace(",", " ").split()]
            mats.append(Material(gam=float(f[3]), ym=float(f[4]), nu=float(f[5]),
                                 phi=float(f[6]), coh=float(f[7]), psi=float(f[8])))
    bc = {}
    for comp, key in (("ux", "uxfile"), ("uy", "uyfile"), ("uz", "uzfile")):
        pairs = []
        with open(os.path.join(inp, bc_kv[key])) as fh:
            toks = fh.read().split()
        t = 0
        while t < len(toks):
            bctype = int(toks[t]); t += 2
            if bctype != 2:
                raise ValueError("only face BCs (type 2) are supported")
            n = int(toks[t]); t += 1
            arr = np.array(toks[t:t + 2 * n], dtype=np.int32).reshape(n, 2)
            pairs.append(arr); t += 2 * n
        bc[comp] = np.concatenate(pairs) if pairs else np.zeros((0, 2), np.int32)
    ctl = kv.get("control", {})
    control = dict(cg_tol=float(ctl.get("cg_tol", 1e-12)),
                   cg_maxiter=int(ctl.get("cg_maxiter", 100)),
                   nl_tol=float(ctl.get("nl_tol", 1e-12)),
                   nl_maxiter=int(ctl.get("nl_maxiter", 100)),
                   srf=[float(s) for s in ctl.get("srf", "1.0").split()],
                   phinu=ctl.get("phinu", "0").strip() == "1")
    # excavation keys (src/read_input.f90:351-373): nexcav stages, stage i
    # removes the nexcavid(i) material ids that follow in excavid
    nexcav = int(ctl.get("nexcav", "0"))
    if nexcav > 0:
        nid = [int(v) for v in ctl.get("nexcavid", "").split()] or [1] * nexcav
        control.update(nexcav=nexcav, nexcavid=nid,
                       excavid=[int(v) for v in ctl["excavid"].split()][:sum(nid)],
                       ninc=int(ctl.get("ninc", "1")))
    st0 = kv.get("stress0", {})
    control["s0_type"] = int(float(st0.get("type", "0")))
    if control["s0_type"] == 1:  # src/read_input.f90:243-248 (k0 unset: 0)
        control.update(z0=float(st0["z0"]), s0=float(st0["s0"]),
                       k0=float(st0.get("k0", "0")))
    # result fields asked for (src/read_input.f90:380-403) and where they go
    sv = kv.get("save", {})
    control["save"] = [k for k in ("disp", "stress", "psigma", "nsigma", "scf", "vmeps")
                       if sv.get(k, "0").strip() == "1"]
    control["out_path"] = pre.get("out_path", "./output")
    mesh = HexMesh(coord=coord, conn=conn, mat_id=mat_id, materials=mats, bc=bc,
                   meta=dict(kind="file", path=sem_path))
    return mesh, control
