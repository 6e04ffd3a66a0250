"""A reference input file run end to end on the GPU, as the reference's main
program runs it (src/semgeotech.F90:64-347): the mesh, materials, boundary
conditions and control of the `.sem` file, then the strength-reduction
sweep (nexcav = 0, src/semgeotech.F90:328-329) or the staged excavation
(:330-332), with the result files the file asks for written to its
out_path under the file's base name."""
from __future__ import annotations

import os
from typing import Optional

from .batch import slope3d_mixed
from .excavation import excavation3d
from .mesh import read_reference_input
from .output import ExcavationOutput, SlopeOutput
from .slope import slope3d
from .solver import KFMT_FULL, KFMT_PACKED, Context
from .spec import hex2spec_device


def run_input(sem_path: str, ctx: Optional[Context] = None, form: str = "fp64",
              ncol: int = 8, out_path: Optional[str] = None):
    """Runs `sem_path`.  form "fp64" is the FP64 drop-in on the full
    (reference-layout) stiffness, "fp64packed" the same on the packed one,
    "mixed" the mixed-precision sweep with `ncol` SRFs in flight (slopes
    only).  `out_path` overrides the file's.  Returns the driver's result
    (slope.SlopeResult or excavation.ExcavationResult)."""
    if form not in ("fp64", "fp64packed", "mixed"):
        raise ValueError(f"unknown form {form!r}")
    mesh, ctl = read_reference_input(sem_path)
    head = os.path.splitext(os.path.basename(sem_path))[0]
    out_dir = out_path or ctl["out_path"]
    own = ctx is None
    ctx = ctx or Context(0)
    try:
        spec = hex2spec_device(ctx, mesh)
        run = dict(cg_tol=ctl["cg_tol"], cg_maxiter=ctl["cg_maxiter"],
                   nl_tol=ctl["nl_tol"], nl_maxiter=ctl["nl_maxiter"], srf=ctl["srf"])
        kfmt = KFMT_PACKED if form == "fp64packed" else KFMT_FULL
        if ctl.get("nexcav", 0) > 0:
            if form == "mixed":
                raise ValueError("the mixed form runs slope sweeps only")
            out = ExcavationOutput(out_dir, head, mesh, spec, ctl, ctl["srf"],
                                   ctl["nexcav"], save=ctl["save"])
            try:
                return excavation3d(ctx, mesh, nexcav=ctl["nexcav"],
                                    nexcavid=ctl["nexcavid"], excavid=ctl["excavid"],
                                    ninc=ctl["ninc"], phinu=ctl["phinu"],
                                    s0_type=ctl["s0_type"], z0=ctl.get("z0", 0.0),
                                    s0=ctl.get("s0", 0.0), k0=ctl.get("k0", 0.0),
                                    kfmt=kfmt, spec=spec, output=out, **run)
            finally:
                out.close()
        out = SlopeOutput(out_dir, head, mesh, spec, ctl, ctl["srf"], save=ctl["save"])
        try:
            if form == "mixed":
                return slope3d_mixed(ctx, mesh, ncol=ncol, spec=spec, output=out,
                                     phinu=ctl["phinu"], **run)
            return slope3d(ctx, mesh, phinu=ctl["phinu"], kfmt=kfmt, spec=spec,
                           output=out, **run)
        finally:
            out.close()
    finally:
        if own:
            ctx.close()
