"""ctypes wrapper of the C restatement (liboracle.so) -- the checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; the product (specfem3d_geotech_amd) never
imports it.  Array layouts: see oracle/sgeo_oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
from types import SimpleNamespace

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (make -C oracle lib)")
        _LIB = C.CDLL(path)
        d, i32, u8 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
        _LIB.orc_gll_tables.argtypes = [d, d, d]
        _LIB.orc_set_dot_mode.argtypes = [C.c_int]
        _LIB.orc_compute_cmat.argtypes = [C.c_double, C.c_double, d]
        _LIB.orc_stress_invariant.argtypes = [d, d, d, d]
        _LIB.orc_mohcouf.argtypes = [C.c_double] * 5
        _LIB.orc_mohcouf.restype = C.c_double
        _LIB.orc_mohcouq.argtypes = [C.c_double] * 3 + [d, d, d]
        _LIB.orc_formm.argtypes = [d, d, d, d]
        _LIB.orc_dt_viscoplas.argtypes = [C.c_int, d, d, d]
        _LIB.orc_dt_viscoplas.restype = C.c_double
        _LIB.orc_strength_reduction.argtypes = [C.c_double, C.c_int, C.c_int] + [d] * 8
        _LIB.orc_strength_reduction.restype = C.c_int
        _LIB.orc_stiffness_bodyload.argtypes = [C.c_int32, C.c_int32, d, i32, i32,
                                                d, d, d, d, d, d, d]
        _LIB.orc_ebe_matvec.argtypes = [C.c_int32, C.c_int32, d, i32, d, d]
        _LIB.orc_pcg_solver.argtypes = [C.c_int32, C.c_int32, d, d, d, d, i32,
                                        C.c_double, C.c_int32, i32]
        _LIB.orc_pcg_solver.restype = C.c_int
        _LIB.orc_plastic_body.argtypes = [C.c_int32, C.c_int32, d, i32, i32,
                                          d, d, d, d, d, C.c_double, d, d, d, d,
                                          C.c_int, u8]
        _LIB.orc_plastic_body.restype = C.c_double
        _LIB.orc_plastic_body_excav.argtypes = [C.c_int32, C.c_int32, i32, d, i32,
                                                i32, d, d, d, d, d, C.c_double,
                                                d, d, d, d, C.c_int, i32, d, d]
        _LIB.orc_plastic_body_excav.restype = C.c_double
        _LIB.orc_slope3d.argtypes = ([C.c_int32] * 3 + [d, i32, i32, i32, C.c_int]
                                     + [d] * 6 + [C.c_int, C.c_double, C.c_int32,
                                                  C.c_double, C.c_int32, C.c_int, d,
                                                  i32, i32, d, d, d, d, i32,
                                                  C.c_int32, i32, d])
        _LIB.orc_slope3d.restype = C.c_int
        _LIB.orc_nodal_fields.argtypes = [C.c_int32, C.c_int32, i32, d, d, d, d]
        _LIB.orc_overburden_stress.argtypes = [C.c_int32, i32, i32, d, d, C.c_double,
                                               C.c_double, C.c_double, d]
        _LIB.orc_elastic_stress.argtypes = [C.c_int32, i32, d, i32, i32, d, d, d, d]
        _LIB.orc_excavation_load_nodal.argtypes = [C.c_int32, i32, d, i32, i32, d, d,
                                                   C.c_int32, d]
        _LIB.orc_dt_viscoplas_mask.argtypes = [C.c_int, d, d, d, u8]
        _LIB.orc_dt_viscoplas_mask.restype = C.c_double
    return _LIB


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def set_dot_mode(mode: int):
    """0 reference order; 1 descending; 2 blocked pairwise (sensitivity)."""
    lib().orc_set_dot_mode(int(mode))


def gll_tables():
    w = np.zeros(27); dl = np.zeros(3 * 27 * 27); ds = np.zeros(3 * 8 * 27)
    lib().orc_gll_tables(_p(w), _p(dl), _p(ds))
    return w, dl.reshape(27, 27, 3), ds.reshape(27, 8, 3)


def cmat(E, nu):
    c = np.zeros(36)
    lib().orc_compute_cmat(E, nu, _p(c))
    return c.reshape(6, 6).T.copy()


def stress_invariant(s):
    s = _f64(s); a, b, t = C.c_double(), C.c_double(), C.c_double()
    lib().orc_stress_invariant(_p(s), C.byref(a), C.byref(b), C.byref(t))
    return a.value, b.value, t.value


def mohcouf(phi, c, sigm, dsbar, theta):
    return lib().orc_mohcouf(phi, c, sigm, dsbar, theta)


def mohcouq(psi, dsbar, theta):
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().orc_mohcouq(psi, dsbar, theta, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def formm(s):
    s = _f64(s); m = [np.zeros(36) for _ in range(3)]
    lib().orc_formm(_p(s), *(_p(x) for x in m))
    return [x.reshape(6, 6).T.copy() for x in m]


def dt_viscoplas(nuf, phif, ym):
    nuf, phif, ym = _f64(nuf), _f64(phif), _f64(ym)
    return lib().orc_dt_viscoplas(len(nuf), _p(nuf), _p(phif), _p(ym))


def strength_reduction(srf, phinu, coh, nu, phi, psi):
    coh, nu, phi, psi = map(_f64, (coh, nu, phi, psi))
    out = [np.zeros_like(coh) for _ in range(4)]
    st = lib().orc_strength_reduction(srf, int(phinu), len(coh), _p(coh), _p(nu),
                                      _p(phi), _p(psi), *(_p(o) for o in out))
    return (*out, st)   # cohf, nuf, phif, psif, istat


def stiffness_bodyload(spec, gam, nu, ym, eqk=None, with_extload=True):
    """Returns (storekm (E,81,81) with [e, c, r] = k(r,c,e), dprecon
    (unassembled, not inverted), extload)."""
    E, neq = spec.nelmt, spec.neq
    km = np.zeros((E, 81, 81))
    dp = np.zeros(neq + 1)
    ext = np.zeros(neq + 1) if with_extload else None
    xc = _f64(spec.corners)
    g = _i32(spec.gdof_elmt)
    eq = _f64(eqk) if eqk is not None else None
    lib().orc_stiffness_bodyload(E, neq, _p(xc), _p(g, C.c_int32),
                                 _p(_i32(spec.mat_id), C.c_int32), _p(_f64(gam)),
                                 _p(_f64(nu)), _p(_f64(ym)), _p(km), _p(dp),
                                 _p(ext), _p(eq))
    return km, dp, ext


def ebe_matvec(neq, storekm, gdof, p):
    kp = np.zeros(neq + 1)
    storekm = _f64(storekm); gdof = _i32(gdof); p = _f64(p)
    lib().orc_ebe_matvec(neq, gdof.shape[0], _p(storekm), _p(gdof, C.c_int32),
                         _p(p), _p(kp))
    return kp


def pcg_solver(neq, storekm, u, f, dinv, gdof, cg_tol, cg_maxiter):
    """Returns (u_out, cg_iter, errcode)."""
    u = _f64(u).copy(); storekm = _f64(storekm); gdof = _i32(gdof)
    it = C.c_int32()
    ec = lib().orc_pcg_solver(neq, gdof.shape[0], _p(storekm), _p(u), _p(_f64(f)),
                              _p(_f64(dinv)), _p(gdof, C.c_int32), cg_tol,
                              cg_maxiter, C.byref(it))
    return u, it.value, ec


def plastic_body(spec, ym, nuf, phif, cohf, psif, dt, x, evpt, stress,
                 bodyload, conv_or_last, want_flags=False):
    """In-place on evpt (E,27,6), stress (E,27,6), bodyload (neq+1,);
    returns (fmax, flags or None)."""
    flags = np.zeros((spec.nelmt, 27), dtype=np.uint8) if want_flags else None
    fmax = lib().orc_plastic_body(
        spec.neq, spec.nelmt, _p(_f64(spec.corners)),
        _p(_i32(spec.gdof_elmt), C.c_int32), _p(_i32(spec.mat_id), C.c_int32),
        _p(_f64(ym)), _p(_f64(nuf)), _p(_f64(phif)), _p(_f64(cohf)),
        _p(_f64(psif)), dt, _p(_f64(x)), _p(evpt), _p(stress), _p(bodyload),
        int(conv_or_last), _p(flags, C.c_uint8))
    return fmax, flags


def slope3d(spec, mesh, *, cg_tol, cg_maxiter, nl_tol, nl_maxiter, srf,
            phinu=False, cg_log_cap=100000):
    """Full SRF sweep; returns dict with per-SRF summary rows, the final x
    and stress state (stress_elmt) and the cg_iter of every PCG call."""
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    srf = _f64(srf)
    n = len(srf)
    scg = np.zeros(n, np.int32); snl = np.zeros(n, np.int32)
    ux = np.zeros(n); um = np.zeros(n); fm = np.zeros(n)
    x = np.zeros(spec.neq + 1)
    log = np.zeros(cg_log_cap, np.int32); nlog = C.c_int32()
    stress = np.zeros((spec.nelmt, 27, 6))
    gnode = _i32(spec.gdof)
    rc = lib().orc_slope3d(spec.nelmt, spec.nnode, spec.neq,
                           _p(_f64(spec.corners)),
                           _p(_i32(spec.gdof_elmt), C.c_int32),
                           _p(gnode, C.c_int32), _p(_i32(spec.mat_id), C.c_int32),
                           len(gam), _p(gam), _p(ym), _p(nu), _p(phi), _p(coh),
                           _p(psi), int(phinu), cg_tol, cg_maxiter, nl_tol,
                           nl_maxiter, n, _p(srf), _p(scg, C.c_int32),
                           _p(snl, C.c_int32), _p(ux), _p(um), _p(fm), _p(x),
                           _p(log, C.c_int32), cg_log_cap, C.byref(nlog), _p(stress))
    if rc < 0:
        raise RuntimeError("ERROR: PCG solver doesn't converge!")
    rows = [dict(srf=float(srf[s]), cg_tot=int(scg[s]), nl_tot=int(snl[s]),
                 uxmax=float(ux[s]), umax=float(um[s]), fmax=float(fm[s]))
            for s in range(rc)]
    return dict(rows=rows, x=x, stress=stress,
                cg_log=log[:min(nlog.value, cg_log_cap)].copy())


def nodal_fields(spec, stress):
    """The nodal fields the reference writes for a trial's final stress
    state (src/slope.F90:418-437, src/postprocess.f90:146-183): stress
    (nnode, 6) the mean over the elements at a node, psigma (nnode, 3) the
    principal stresses ascending, nsigma (nnode,)."""
    sig = np.zeros((spec.nnode, 6)); psig = np.zeros((spec.nnode, 3))
    nsig = np.zeros(spec.nnode)
    lib().orc_nodal_fields(spec.nelmt, spec.nnode, _p(_i32(spec.g_num), C.c_int32),
                           _p(_f64(stress)), _p(sig), _p(psig), _p(nsig))
    return dict(stress=sig, psigma=psig, nsigma=nsig)


# ----------------------------------------------------------------------------
# excavation (src/excavation.F90:180-232 initial stress, :333-700 stages)
# ----------------------------------------------------------------------------

def elastic_stress(spec, ym, nu, x, elmt=None):
    """stress (E,27,6) = C B u_e at every GLL point (src/postprocess.f90:40-79)."""
    st = np.zeros((spec.nelmt, 27, 6))
    el = _i32(elmt) if elmt is not None else None
    lib().orc_elastic_stress(len(el) if el is not None else spec.nelmt,
                             _p(el, C.c_int32), _p(_f64(spec.corners)),
                             _p(_i32(spec.gdof_elmt), C.c_int32),
                             _p(_i32(spec.mat_id), C.c_int32), _p(_f64(ym)),
                             _p(_f64(nu)), _p(_f64(x)), _p(st))
    return st


def excavation_load_nodal(spec, gam, stress, elmt_void):
    """excavload (nnode, 3) of the void elements (src/excavation_library.f90:
    186-236)."""
    out = np.zeros((spec.nnode, 3))
    ev = _i32(elmt_void)
    lib().orc_excavation_load_nodal(len(ev), _p(ev, C.c_int32),
                                    _p(_f64(spec.corners)),
                                    _p(_i32(spec.g_num), C.c_int32),
                                    _p(_i32(spec.mat_id), C.c_int32), _p(_f64(gam)),
                                    _p(_f64(stress)), spec.nnode, _p(out))
    return out


def dt_viscoplas_mask(nuf, phif, ym, ismat):
    m = np.ascontiguousarray(ismat, dtype=np.uint8)
    nuf, phif, ym = _f64(nuf), _f64(phif), _f64(ym)
    return lib().orc_dt_viscoplas_mask(len(nuf), _p(nuf), _p(phif), _p(ym),
                                       _p(m, C.c_uint8))


def _stiffness_subset(corners, gdof, mat_id, neq, gam, nu, ym):
    E = len(mat_id)
    km = np.zeros((E, 81, 81)); dp = np.zeros(neq + 1)
    lib().orc_stiffness_bodyload(E, neq, _p(_f64(corners)), _p(_i32(gdof), C.c_int32),
                                 _p(_i32(mat_id), C.c_int32), _p(_f64(gam)),
                                 _p(_f64(nu)), _p(_f64(ym)), _p(km), _p(dp), None, None)
    return km, dp


def renumber_gdof(gdof_node, node_alive):
    """modify_gdof (src/excavation_library.f90:90-111): dead nodes fixed,
    then every free dof renumbered node-major, component fastest."""
    g = np.where(node_alive[:, None], gdof_node, 0)
    free = g.reshape(-1) != 0
    out = np.zeros(g.size, np.int64)
    out[free] = np.arange(1, int(free.sum()) + 1)
    return out.reshape(g.shape), int(free.sum())


def overburden_stress(spec, gam, z0, p0, k0):
    """stress (E,27,6) of s0_type 1 (src/postprocess.f90:10-33)."""
    st = np.zeros((spec.nelmt, 27, 6))
    z = np.ascontiguousarray(np.asarray(spec.g_coord, dtype=np.float64)[:, 2])
    lib().orc_overburden_stress(spec.nelmt, _p(_i32(spec.g_num), C.c_int32),
                                _p(_i32(spec.mat_id), C.c_int32), _p(z), _p(_f64(gam)),
                                float(z0), float(p0), float(k0), _p(st))
    return st


def excavation3d(spec, mesh, *, cg_tol, cg_maxiter, nl_tol, nl_maxiter, srf,
                 nexcav, nexcavid, excavid, ninc, phinu=False, s0_type=0,
                 on_pcg=None, z0=0.0, s0=0.0, k0=0.0):
    """The reference's staged excavation (src/excavation.F90:13-700), one SRF
    (the reference supports only nsrf = 1 there).  Returns dict(rows=[(stage,
    cg_tot, nl_tot, uxmax, umax, fmax)], stress0, nodalu, cg_log, fields):
    fields[k] holds what save_data receives for stage k (0 = the in-situ
    state, src/excavation.F90:296-323; then :655-687): node_intact (0-based
    ids), nodalu (nnode_intact, 3), scf, vmeps (nnode_intact,) and the
    nodal fields of the stress (nodal_fields) at those nodes."""
    if s0_type not in (0, 1):
        raise ValueError(f"s0_type {s0_type} not supported")
    gam, ym, nu, phi, coh, psi = mesh.material_arrays()
    nmat, E = len(gam), spec.nelmt
    g_num = np.asarray(spec.g_num)
    mat = np.asarray(spec.mat_id)
    gdof_node = np.asarray(spec.gdof).copy()
    cg_log = []
    if s0_type == 0:
        # initial stress: elastic solve under gravity on the whole mesh
        km, dp, ext = stiffness_bodyload(spec, gam, nu, ym)
        dinv = np.zeros_like(dp); dinv[1:] = 1.0 / dp[1:]
        x, it, ec = pcg_solver(spec.neq, km, np.zeros(spec.neq + 1), ext, dinv,
                               spec.gdof_elmt, cg_tol, cg_maxiter)
        if ec != 0:
            raise RuntimeError("ERROR: PCG solver doesn't converge!")
        cg_log.append(it)
        x[0] = 0.0
        stress = elastic_stress(spec, ym, nu, x)
        del km
    else:  # overburden (src/excavation.F90:238-241)
        stress = overburden_stress(spec, gam, z0, s0, k0)
    stress0 = stress.copy()
    s = float(srf[0])
    nodalu = np.zeros((spec.nnode, 3))
    ismat = np.ones(nmat, bool)
    evpt = np.zeros((E, 27, 6))
    rows = []
    every = np.arange(spec.nnode)
    fields = [dict(node_intact=every, nodalu=np.zeros((spec.nnode, 3)),
                   scf=np.zeros(spec.nnode), vmeps=np.zeros(spec.nnode),
                   **nodal_fields(spec, stress))]
    for i_ex in range(nexcav):
        id0 = i_ex * nexcavid[i_ex]
        ids = np.asarray(excavid[id0:id0 + nexcavid[i_ex]])
        ismat[ids - 1] = False
        off = np.zeros(nmat, bool); off[ids - 1] = True
        void = np.nonzero(off[mat - 1])[0].astype(np.int32)
        intact = np.nonzero(~off[mat - 1] & ismat[mat - 1])[0].astype(np.int32)
        if intact.size == 0:
            break
        alive = np.zeros(spec.nnode, bool); alive[g_num[intact].reshape(-1)] = True
        gdof_node, neq = renumber_gdof(gdof_node, alive)
        gdof_elmt = gdof_node[g_num].reshape(E, 81).astype(np.int32)
        exl = excavation_load_nodal(spec, gam, stress, void)
        extload = np.zeros(neq + 1)
        extload[gdof_node.reshape(-1)] = exl.reshape(-1)   # distribute2ghosts
        extload[0] = 0.0
        cohf, nuf, phif, psif, _ = strength_reduction(s, phinu, coh, nu, phi, psi)
        dt = dt_viscoplas_mask(nuf, phif, ym, ismat)
        ge = gdof_elmt[intact]
        km, dp = _stiffness_subset(spec.corners[intact], ge, mat[intact], neq, gam,
                                   nuf, ym)
        dinv = np.zeros_like(dp); dinv[1:] = 1.0 / dp[1:]
        extload = extload / ninc
        cg_tot = nl_tot = 0
        fmax = 0.0
        nl_iter = 0
        vmeps = np.zeros(spec.nnode)
        scf = np.full(spec.nnode, 1.0e32)  # inftol (src/global.f90:22)
        g_i32 = _i32(g_num)
        for _ in range(ninc):
            bodyload = np.zeros(neq + 1)
            evpt[:] = 0.0
            x = np.zeros(neq + 1); oldx = np.zeros(neq + 1)
            for nl_iter in range(1, nl_maxiter + 1):
                load = extload + bodyload
                load[0] = 0.0
                x, it, ec = pcg_solver(neq, km, x, load, dinv, ge, cg_tol, cg_maxiter)
                if ec != 0:
                    raise RuntimeError("ERROR: PCG solver doesn't converge!")
                cg_log.append(it)
                if on_pcg:
                    on_pcg(i_ex, it)
                cg_tot += it
                x[0] = 0.0
                uerr = np.abs(x - oldx).max() / np.abs(x).max()
                oldx = x.copy()
                conv = uerr <= nl_tol
                last = conv or nl_iter == nl_maxiter
                fmax = lib().orc_plastic_body_excav(
                    neq, len(intact), _p(intact, C.c_int32), _p(_f64(spec.corners)),
                    _p(ge, C.c_int32), _p(_i32(mat), C.c_int32), _p(_f64(ym)),
                    _p(_f64(nuf)), _p(_f64(phif)), _p(_f64(cohf)), _p(_f64(psif)), dt,
                    _p(x), _p(evpt), _p(stress), _p(bodyload), int(last),
                    _p(g_i32, C.c_int32), _p(vmeps), _p(scf))
                if last:
                    break
            nl_tot += nl_iter
            live = gdof_node != 0
            nodalu[live] += x[gdof_node[live]]
        un = nodalu[alive]
        rows.append((i_ex + 1, cg_tot, nl_tot, float(np.abs(un).max()),
                     float(np.sqrt((un * un).sum(axis=1)).max()), float(fmax)))
        node_intact = np.nonzero(alive)[0]
        sub = SimpleNamespace(nelmt=len(intact), nnode=spec.nnode, g_num=g_num[intact])
        nf = nodal_fields(sub, stress[intact])
        val = np.bincount(g_num[intact].reshape(-1), minlength=spec.nnode)
        fields.append(dict(node_intact=node_intact, nodalu=nodalu[node_intact].copy(),
                           scf=scf[node_intact],
                           vmeps=(vmeps / np.maximum(val, 1))[node_intact],
                           **{k: v[node_intact] for k, v in nf.items()}))
        if nl_iter == nl_maxiter:
            break
    return dict(rows=rows, stress0=stress0, nodalu=nodalu, cg_log=cg_log,
                stress=stress, fields=fields)
