"""Result files in the reference's formats.

* The summary file (src/semgeotech.F90:317-321, src/slope.F90:214-220,
  414-416, src/semgeotech.F90:341-347): list-directed lines as the
  reference's compiler writes them -- `fortran_list` reproduces its
  records (at most 79 columns) and its number forms.
* EnSight Gold, C Binary (src/semgeotech.F90:138-313, src/visual.f90:17-271,
  src/postprocess.f90:128-248): the `_original` case and geometry of the
  hex8 mesh, the case file of the sweep, the GLL geometry (every 27-node
  element as 8 hexa8 cells) and per SRF one FP32 file per saved field:
  displacement, nodal stress, principal stresses, normal stress, stress
  concentration factor and plastic strain.  The stress-derived fields are
  computed on the device (sgeo_*_nodal_fields); this module only writes.

`SlopeOutput` writes these beside a sweep; file writes run on a background
thread, so they overlap the next SRF's GPU work.
"""
from __future__ import annotations

import math
import os
import queue
import threading
import time
from decimal import Decimal
from typing import Dict, List, Optional, Sequence

import numpy as np

from .mesh import HexMesh
from .spec import SpecMesh

RECORD = 80  # record length of list-directed output; a line ends before column 80


# ----------------------------------------------------------------------------
# list-directed number format
# ----------------------------------------------------------------------------

def _shortest_digits(v: float):
    """Shortest significant digits d1..dn and exponent e with v = 0.d1..dn x
    10**e that convert back to v (v > 0 finite); at each length the
    truncated digits are taken when they convert back, else the digits
    rounded up."""
    _, digs, exp = Decimal(v).as_tuple()
    ds = "".join(map(str, digs))
    e10 = exp + len(ds)  # v = 0.ds x 10**e10, exactly
    for n in range(1, len(ds) + 1):
        lo = ds[:n]
        if float(f"0.{lo}e{e10}") == v:
            return lo.rstrip("0") or "0", e10
        up = str(int(lo) + 1)
        ue = e10 + (len(up) - n)
        if float(f"0.{up}e{ue}") == v:
            return up.rstrip("0") or "0", ue
    return ds.rstrip("0"), e10


def _first_digit_exponent(v: float) -> int:
    """Exponent e of v rounded to one significant digit, as 0.d x 10**e."""
    d = Decimal(v)
    q = Decimal(f"1e{d.adjusted()}")
    r = d.quantize(q, rounding="ROUND_HALF_EVEN")
    return r.adjusted() + 1


def fortran_real(v: float) -> str:
    """One REAL(8) item as list-directed output writes it: the shortest
    digits, F form when the one-digit rounded value lies in [0.1, 1e15),
    else E form with a two-digit (or longer) exponent; no leading zero."""
    if v == 0.0:
        return "-0." if math.copysign(1.0, v) < 0 else "0."
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Inf" if v > 0 else "-Inf"
    sign = "-" if v < 0 else ""
    a = abs(v)
    ds, e10 = _shortest_digits(a)
    e1 = _first_digit_exponent(a)
    if 0 <= e1 <= 15:
        if e10 <= 0:
            return f"{sign}.{'0' * (-e10)}{ds}"
        ip = ds[:e10].ljust(e10, "0")
        return f"{sign}{ip}.{ds[e10:]}"
    x = e10 - 1
    return f"{sign}{ds[0]}.{ds[1:]}E{'-' if x < 0 else '+'}{abs(x):02d}"


def fortran_item(v) -> str:
    if isinstance(v, str):
        return v
    if isinstance(v, (bool, np.bool_)):
        return "T" if v else "F"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return fortran_real(float(v))


def fortran_list(*items) -> str:
    """A list-directed WRITE of `items`: a blank, then the items separated
    by blanks; an item that would reach column 80 starts a new record."""
    lines: List[str] = []
    cur = ""
    for it in items:
        s = fortran_item(it)
        if cur and len(cur) + 1 + len(s) >= RECORD:
            lines.append(cur)
            cur = ""
        cur = cur + " " + s
    lines.append(cur)
    return "\n".join(lines) + "\n"


# ----------------------------------------------------------------------------
# EnSight Gold, C Binary
# ----------------------------------------------------------------------------

def _s80(text: str) -> bytes:
    return text.encode("ascii").ljust(80, b" ")[:80]


def _i32(v: int) -> bytes:
    return np.int32(v).tobytes()


def _geo_header(nnode: int, coord: np.ndarray) -> bytes:
    """write_ensight_geocoord (src/visual.f90:17-72): header, one part,
    FP32 coordinates as x block, y block, z block."""
    c = np.asarray(coord, dtype=np.float64).astype(np.float32)
    return b"".join([_s80("C Binary"), _s80("Created by write_ensight Routine"),
                     _s80("semfem3d"), _s80("node id off"), _s80("element id off"),
                     _s80("part"), _i32(1), _s80("unstructured meshes"),
                     _s80("coordinates"), _i32(nnode),
                     np.ascontiguousarray(c.T).tobytes()])


def geo_original(mesh: HexMesh) -> bytes:
    """The hex8 mesh as given (write_ensight_geo, src/visual.f90:77-146)."""
    conn = np.asarray(mesh.conn, dtype=np.int32)
    return b"".join([_geo_header(mesh.nnode, mesh.coord), _s80("hexa8"),
                     _i32(conn.shape[0]), np.ascontiguousarray(conn).tobytes()])


# local GLL nodes of the 8 hexa8 cells of a 27-node element in EXODUS corner
# order (src/semgeotech.F90:290-309, map2exodus src/element.f90:7)
def _cells27() -> np.ndarray:
    out = []
    for k in range(2):
        for j in range(2):
            for i in range(2):
                n1 = k * 9 + j * 3 + i
                n = [n1, n1 + 1, n1 + 3, n1 + 4, n1 + 9, n1 + 10, n1 + 12, n1 + 13]
                out.append([n[m] for m in (0, 1, 3, 2, 4, 5, 7, 6)])
    return np.array(out, dtype=np.int64)


CELLS27 = _cells27()


def geo_gll(spec: SpecMesh) -> bytes:
    """The GLL mesh: every node, every element as 8 hexa8 cells
    (src/semgeotech.F90:272-313), 1-based node ids."""
    g = np.asarray(spec.g_num, dtype=np.int64)
    cells = (g[:, CELLS27] + 1).astype(np.int32).reshape(-1, 8)
    return b"".join([_geo_header(spec.nnode, spec.g_coord), _s80("hexa8"),
                     _i32(cells.shape[0]), np.ascontiguousarray(cells).tobytes()])


def pernode_vector(destag: str, var: np.ndarray) -> bytes:
    """write_ensight_pernodeVECAS (src/visual.f90:191-231): FP32 components
    as blocks; var is (nnode, ncomp), FP64 (rounded here, as real() does)
    or already FP32."""
    v = np.asarray(var)
    v = v.astype(np.float32) if v.dtype != np.float32 else v
    return b"".join([_s80(destag), _s80("part"), _i32(1), _s80("coordinates"),
                     np.ascontiguousarray(v.T).tobytes()])


def pernode_scalar(destag: str, var: np.ndarray) -> bytes:
    """write_ensight_pernodeSCALAS (src/visual.f90:236-271): one FP32
    value per node."""
    v = np.asarray(var)
    v = v.astype(np.float32) if v.dtype != np.float32 else v
    return b"".join([_s80(destag), _s80("part"), _i32(1), _s80("coordinates"),
                     np.ascontiguousarray(v).tobytes()])


# Result fields of a sweep in the reference's case-file order
# (src/semgeotech.F90:216-249): save key, case line type, variable name,
# file suffix, description written into the file (src/postprocess.f90:185-247).
# porep (no pore pressure here) and maxtau (not settable from the input
# file, src/read_input.f90:380-403) are not written by the reference's
# slope runs either.
FIELDS = [
    ("disp", "vector per node", "displacement", "dis", "Displacement field"),
    ("stress", "tensor symm per node", "stress", "sig", "Effective stress tensor"),
    ("psigma", "vector per node", "principal_stress", "psig", "principal stresses"),
    ("vmeps", "scalar per node", "plastic_strain", "eps",
     "von Mises effective plastic strain"),
    ("scf", "scalar per node", "stress_concentration_factor", "scf",
     "principal stress concentration factor"),
    ("nsigma", "scalar per node", "normal_stress", "nsig", "normal stress"),
]
SAVE_KEYS = tuple(f[0] for f in FIELDS)
INFTOL = 1.0e32  # src/global.f90:22: scf of a slope trial (src/slope.F90:249)


def step_width(nt: int) -> int:
    """twidth = ceiling(log10(real(nt)+1.)) (src/semgeotech.F90:136), in the
    reference's single precision."""
    return int(math.ceil(float(np.log10(np.float32(nt) + np.float32(1.0)))))


def case_original(head: str, ptail: str = "") -> str:
    return ("FORMAT\ntype:  ensight gold\n\nGEOMETRY\n"
            f"model:    {head}_original{ptail}.geo\n\n")


def _e12_5(v: float) -> str:
    """Fortran E12.5: [-]0.dddddE+xx, mantissa in [0.1, 1), right-justified
    in 12 columns."""
    d = Decimal(v)
    if d == 0:
        return "0.00000E+00".rjust(12)
    e = d.adjusted() + 1
    m = abs(d).scaleb(-e).quantize(Decimal("0.00001"), rounding="ROUND_HALF_EVEN")
    if m >= 1:
        e += 1
        m = abs(d).scaleb(-e).quantize(Decimal("0.00001"), rounding="ROUND_HALF_EVEN")
    txt = f"{'-' if v < 0 else ''}{m:.5f}E{'-' if e < 0 else '+'}{abs(e):02d}"
    return txt.rjust(12)


def case_slope(head: str, srf: Sequence[float], disp: bool = True, ptail: str = "",
               save: Optional[Sequence[str]] = None) -> str:
    """The sweep's case file (src/semgeotech.F90:196-266): one variable line
    per saved field, time values are the SRFs.  `save` lists the saved
    fields by key (default: the displacement when `disp`)."""
    save = _save_set(save, disp)
    nt = len(srf)
    tw = step_width(nt)
    s = "FORMAT\ntype:  ensight gold\n\nGEOMETRY\n"
    s += f"model:    {head}{ptail}.geo\n\n"
    s += "VARIABLE\n"
    for key, kind, name, suffix, _ in FIELDS:
        if key in save:
            s += f"{kind}: {1:10d} {name} {head}_step{'*' * tw}{ptail}.{suffix}\n\n"
    s += "TIME\n"
    s += f"time set:{1:10d}\n"
    s += f"number of steps:{nt:10d}\n"
    s += f"filename start number:{1:10d}\n"
    s += f"filename increment:{1:10d}\n"
    s += "time values: "
    s += "".join(_e12_5(float(v)) + "\n" for v in srf)
    return s


def _save_set(save: Optional[Sequence[str]], disp: bool) -> set:
    out = set(save) if save is not None else ({"disp"} if disp else set())
    bad = out - set(SAVE_KEYS)
    if bad:
        raise ValueError(f"unknown save fields {sorted(bad)}; known: {SAVE_KEYS}")
    return out


def nodal_displacement(spec: SpecMesh, x: np.ndarray) -> np.ndarray:
    """nodalu(i, j) = x(gdof(i, j)) where gdof /= 0 (src/slope.F90:400-407)."""
    g = np.asarray(spec.gdof)
    u = np.where(g != 0, np.asarray(x)[g], 0.0)
    return u


# ----------------------------------------------------------------------------
# a sweep's files
# ----------------------------------------------------------------------------

class _Files:
    """Background writer and the parts every run's files share."""

    def __init__(self, out_path: str, head: str, spec: SpecMesh,
                 save: Optional[Sequence[str]], disp: bool):
        self.out_path, self.head, self.spec = out_path, head, spec
        self.save = _save_set(save, disp)
        self.disp = "disp" in self.save
        # fields that need the trial's stress state
        self.wants_fields = bool(self.save & {"stress", "psigma", "nsigma", "scf",
                                              "vmeps"})
        self.t0 = time.perf_counter()
        os.makedirs(out_path, exist_ok=True)
        self._q: "queue.Queue" = queue.Queue()
        self._err: Optional[BaseException] = None
        self._th = threading.Thread(target=self._run, daemon=True)
        self._th.start()
        self._summary = self.path(f"{head}_summary")

    def path(self, name: str) -> str:
        return os.path.join(self.out_path, name)

    def _put(self, path: str, data: bytes, append: bool = False):
        self._q.put((path, data, append))

    def _run(self):
        while True:
            job = self._q.get()
            if job is None:
                return
            path, data, append = job
            try:
                with open(path, "ab" if append else "wb") as fh:
                    fh.write(data)
            except BaseException as e:  # reported at close
                self._err = e

    def _summary_head(self, control: Dict) -> str:
        """src/semgeotech.F90:317-321, then the drivers' control lines
        (src/slope.F90:214-217, src/excavation.F90:262-266)."""
        return (fortran_list("--------------------------------------------")
                + fortran_list("Result summary produced by SPECFEM3D_GEOTECH")
                + fortran_list("--------------------------------------------")
                + fortran_list("CG_MAXITER, CG_TOL, NL_MAXITER, NL_TOL")
                + fortran_list(int(control["cg_maxiter"]), float(control["cg_tol"]),
                               int(control["nl_maxiter"]), float(control["nl_tol"])))

    def _fields_of(self, fields):
        if not self.wants_fields:
            return None
        if fields is None:
            raise ValueError("the saved fields need the trial's nodal fields")
        return fields() if callable(fields) else fields

    def _write_step(self, step: int, tw: int, u, f, sel=None):
        """save_data (src/postprocess.f90:128-248): one file per saved field
        of output step `step`; u (nnode, 3) the nodal displacement, f the
        nodal fields, sel the nodes written (all when None)."""
        name = lambda suffix: self.path(f"{self.head}_step{step:0{tw}d}.{suffix}")
        take = (lambda a: a) if sel is None else (lambda a: np.asarray(a)[sel])
        nn = self.spec.nnode if sel is None else len(sel)
        for key, _, _, suffix, destag in FIELDS:
            if key not in self.save:
                continue
            if key == "disp":
                data = pernode_vector(destag, u)
            elif key in ("stress", "psigma"):
                data = pernode_vector(destag, take(f[key]))
            elif key == "nsigma":
                data = pernode_scalar(destag, take(f["nsigma"]))
            elif key == "scf":
                data = pernode_scalar(destag, take(f["scf"]) if "scf" in f
                                      else np.full(nn, INFTOL))
            else:
                data = pernode_scalar(destag, take(f["vmeps"]) if "vmeps" in f
                                      else np.zeros(nn))
            self._put(name(suffix), data)

    def close(self):
        """Elapsed-time lines (src/semgeotech.F90:341-347), then waits for
        every write."""
        t = time.perf_counter() - self.t0
        w = int(math.ceil(math.log10(t + 1.0))) + 5
        line = " ".join(f"{t:{w}.4f}" for _ in range(3))
        self._put(self._summary, (fortran_list("ELAPSED TIME, MAX ELAPSED TIME, "
                                               "MEAN ELAPSED TIME") + line + "\n").encode(),
                  append=True)
        self._q.put(None)
        self._th.join()
        if self._err is not None:
            raise self._err


class SlopeOutput(_Files):
    """The files the reference writes for an SRF sweep, in `out_path` with
    file head `head`.  `save` lists the result fields as the input file's
    `save:` keys (disp, stress, psigma, nsigma, scf, vmeps); the default is
    the displacement alone (`disp`).  Writes run on a background thread;
    `close` waits for them and appends the elapsed-time lines."""

    def __init__(self, out_path: str, head: str, mesh: HexMesh, spec: SpecMesh,
                 control: Dict, srf: Sequence[float], disp: bool = True,
                 save: Optional[Sequence[str]] = None):
        super().__init__(out_path, head, spec, save, disp)
        self.srf = list(srf)
        self.nl_maxiter = int(control["nl_maxiter"])
        self._pending: Dict[int, tuple] = {}
        self._next = 1
        self._stopped = False
        p = self.path
        self._put(p(f"{head}_original.case"), case_original(head).encode())
        self._put(p(f"{head}_original.geo"), geo_original(mesh))
        self._put(p(f"{head}.case"), case_slope(head, self.srf, save=self.save).encode())
        self._put(p(f"{head}.geo"), geo_gll(spec))
        summ = (self._summary_head(control)
                + fortran_list("Number of SRFs")
                + fortran_list(len(self.srf))
                + fortran_list("SRF, CGITER, NLITER, UXMAX, UMAX, fmax"))
        self._put(self._summary, summ.encode())

    def srf_done(self, i_srf: int, row: Dict, x: Optional[np.ndarray], fields=None):
        """Summary row of SRF number i_srf (1-based) and its field files
        (src/slope.F90:414-439, src/postprocess.f90:128-248).  `fields` is
        the trial's nodal result fields (a dict, or a callable returning
        one, called here only if a saved field needs it).  Trials may end in
        any order (SRFs run in parallel); rows are written in SRF order, and
        nothing after the first SRF whose plastic loop reached nl_maxiter,
        where the reference's sweep stops (src/slope.F90:441)."""
        f = self._fields_of(fields)
        self._pending[i_srf] = (row, x if self.disp else None, f)
        while self._next in self._pending and not self._stopped:
            row, x, f = self._pending.pop(self._next)
            self._put(self._summary, fortran_list(
                float(row["srf"]), int(row["cg_tot"]), int(row["nl_tot"]),
                float(row["uxmax"]), float(row["umax"]), float(row["fmax"])).encode(),
                append=True)
            u = nodal_displacement(self.spec, x) if x is not None else None
            self._write_step(self._next, step_width(len(self.srf)), u, f)
            if int(row["nl_tot"]) >= self.nl_maxiter:
                self._stopped = True
            self._next += 1


def case_excavation(head: str, nexcav: int, save: Sequence[str], ptail: str = "") -> str:
    """The excavation's case file (src/semgeotech.F90:126-266): one geometry
    per step, steps 0 (in situ) .. nexcav, time values 0, 1, ..."""
    nt = nexcav + 1
    tw = step_width(nt)
    s = "FORMAT\ntype:  ensight gold\n\nGEOMETRY\n"
    s += f"model:    {1:10d} {head}_step{'*' * tw}{ptail}.geo\n\n"
    s += "VARIABLE\n"
    for key, kind, name, suffix, _ in FIELDS:
        if key in save:
            s += f"{kind}: {1:10d} {name} {head}_step{'*' * tw}{ptail}.{suffix}\n\n"
    s += "TIME\n"
    s += f"time set:{1:10d}\n"
    s += f"number of steps:{nt:10d}\n"
    s += f"filename start number:{0:10d}\n"
    s += f"filename increment:{1:10d}\n"
    s += "time values: "
    s += "".join(_e12_5(float(i)) + "\n" for i in range(nt))
    return s


def geo_intact(spec: SpecMesh, node_intact: np.ndarray, elmt_intact: np.ndarray) -> bytes:
    """A stage's geometry (src/excavation.F90:397-430): the intact nodes,
    renumbered 1.. in ascending order (nmir), and the intact elements as
    8 hexa8 cells each."""
    node_intact = np.asarray(node_intact, dtype=np.int64)
    nmir = np.zeros(spec.nnode, np.int64)
    nmir[node_intact] = np.arange(1, len(node_intact) + 1)
    g = np.asarray(spec.g_num, dtype=np.int64)[np.asarray(elmt_intact, dtype=np.int64)]
    cells = nmir[g[:, CELLS27]].astype(np.int32).reshape(-1, 8)
    coord = np.asarray(spec.g_coord)[node_intact]
    return b"".join([_geo_header(len(node_intact), coord), _s80("hexa8"),
                     _i32(cells.shape[0]), np.ascontiguousarray(cells).tobytes()])


class ExcavationOutput(_Files):
    """The files the reference writes for a staged excavation (one SRF, the
    only case the reference runs): the hex8 original, the case file, step 0
    (the whole GLL mesh and the in-situ stress) and per stage the intact
    mesh and its fields at the intact nodes, and the summary
    (src/excavation.F90:262-291, 296-323, 397-430, 641-687)."""

    def __init__(self, out_path: str, head: str, mesh: HexMesh, spec: SpecMesh,
                 control: Dict, srf: Sequence[float], nexcav: int,
                 save: Optional[Sequence[str]] = None):
        super().__init__(out_path, head, spec, save, True)
        self.tw = step_width(nexcav + 1)
        p = self.path
        self._put(p(f"{head}_original.case"), case_original(head).encode())
        self._put(p(f"{head}_original.geo"), geo_original(mesh))
        self._put(p(f"{head}.case"), case_excavation(head, nexcav, self.save).encode())
        self._put(p(f"{head}_step{0:0{self.tw}d}.geo"), geo_gll(spec))
        summ = (self._summary_head(control)
                + fortran_list("Number of SRFs")
                + fortran_list(len(srf))
                + fortran_list("SRF")
                + fortran_list(float(srf[0]))
                + fortran_list("Number of excavation stages")
                + fortran_list(int(nexcav))
                + fortran_list("STEP, CGITER, NLITER, UXMAX, UMAX, fmax"))
        self._put(self._summary, summ.encode())

    def in_situ(self, fields=None):
        """Step 0: zero displacement, scf and vmeps, the in-situ stress
        fields at every node (src/excavation.F90:296-323)."""
        f = self._fields_of(fields)
        nn = self.spec.nnode
        if f is not None:
            f = dict(f, scf=np.zeros(nn), vmeps=np.zeros(nn))
        self._write_step(0, self.tw, np.zeros((nn, 3)), f)

    def stage_done(self, row: Dict, nodalu: np.ndarray, node_intact: np.ndarray,
                   elmt_intact: np.ndarray, fields=None):
        """Stage row (src/excavation.F90:646-648), the stage's geometry and
        its fields at the intact nodes (:655-687); nodalu (nnode, 3) is the
        displacement summed over the stages so far, fields the stage's
        nodal fields on the whole node range."""
        k = int(row["stage"])
        f = self._fields_of(fields)
        self._put(self.path(f"{self.head}_step{k:0{self.tw}d}.geo"),
                  geo_intact(self.spec, node_intact, elmt_intact))
        self._put(self._summary, fortran_list(
            k, int(row["cg_tot"]), int(row["nl_tot"]), float(row["uxmax"]),
            float(row["umax"]), float(row["fmax"])).encode(), append=True)
        self._write_step(k, self.tw, np.asarray(nodalu)[node_intact], f, sel=node_intact)
