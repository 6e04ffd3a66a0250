/*
 * sgeo.h -- C ABI of the MI355X (gfx950) EBE-PCG + viscoplastic hot path of
 * SPECFEM3D_GEOTECH.
 *
 * Conventions (same as the reference, src/slope.F90:139-199):
 *   * equations are numbered 1..neq; every global vector has neq+1 entries
 *     and index 0 is the sink of constrained dofs (always zero on output);
 *   * gdof_elmt[e*81 + i] is the reference's gdof_elmt(i+1, e+1): element dof
 *     order is node-major (27 GLL nodes in local k,j,i order, i fastest),
 *     x/y/z fastest; 0 = constrained;
 *   * storekm[e*6561 + r + 81*c] is the reference's k(r+1, c+1, e+1)
 *     (column-major 81x81 per element);
 *   * corners[e*24 + 3*g + d] is coordinate d of hex8 corner g (EXODUS order,
 *     src/element.f90:32-40) of element e;
 *   * mat_id[e] is 1-based; per-material arrays are indexed mat_id-1;
 *     angles in degrees.
 * Every function returns 0 on success and a negative SGEO_E* code on a
 * device or argument error; sgeo_last_error() then describes it.  PCG
 * non-convergence is not an error of the call: it is reported through
 * errcode = -1 exactly like the reference's pcg_solver.
 * A context is bound to one HIP device and one stream; calls on one context
 * are not thread-safe, separate contexts may be used from separate threads.
 */
#ifndef SGEO_H
#define SGEO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGEO_OK 0
#define SGEO_EARG (-1)     /* invalid argument / size */
#define SGEO_EHIP (-2)     /* HIP runtime error */
#define SGEO_ENOMEM (-3)   /* device allocation failed */
#define SGEO_ESTATE (-4)   /* call out of order */

/* precision of the stored element stiffness */
#define SGEO_K_F64 0
#define SGEO_K_F32 1

typedef struct sgeo_ctx sgeo_ctx;
typedef struct sgeo_mesh sgeo_mesh;
typedef struct sgeo_slope sgeo_slope;
typedef struct sgeo_batch sgeo_batch;
typedef struct sgeo_spec sgeo_spec;

/* ---- context ---------------------------------------------------------- */
int sgeo_init(int device, sgeo_ctx **ctx);
int sgeo_finalize(sgeo_ctx *ctx);
const char *sgeo_last_error(void);
/* Device name, CU count and HBM bytes of the context's device. */
int sgeo_device_info(sgeo_ctx *ctx, char *name, size_t name_len, int *n_cu,
                     size_t *hbm_bytes);
int sgeo_synchronize(sgeo_ctx *ctx);

/* Summation order of the PCG dot products (no counterpart in the reference,
 * whose dot_product is sequential, src/solver.f90:61-62,72).
 *   SGEO_REDUCE_TREE        blocked tree reductions (default, fast,
 *                           deterministic run to run)
 *   SGEO_REDUCE_SEQUENTIAL  index order, one rounding per term: with it the
 *                           whole PCG reproduces the reference bit for bit;
 *                           its dot products are latency-bound (O(neq) on one
 *                           lane), so it is meant for parity runs. */
#define SGEO_REDUCE_TREE 0
#define SGEO_REDUCE_SEQUENTIAL 1
int sgeo_set_reduction(sgeo_ctx *ctx, int mode);

/* HIP-event timing of the element-product (EBE matvec) kernel inside PCG
 * loops, on the context's stream (no counterpart in the reference).  Reads
 * the launches timed and their summed duration since the last reset; then
 * enable >= 0 resets the counters and turns timing on (1) or off (0);
 * enable < 0 only reads. */
int sgeo_profile(sgeo_ctx *ctx, int enable, int64_t *launches,
                 double *total_ms);

/* ---- Seam 1: drop-in for pcg_solver ------------------------------------
 * Replaces subroutine pcg_solver(neq,nelmt,k,u,f,dprecon,gdof_elmt,cg_iter,
 * errcode,errtag) (src/solver.f90:14-79; called from src/slope.F90:293,
 * src/excavation.F90:231,525).  Host arrays in the reference layout:
 * k (81*81*nelmt), u (neq+1, in/out, warm start), f and dprecon (neq+1,
 * dprecon already inverted with dprecon[0]=0 as src/slope.F90:199 leaves
 * it), gdof_elmt (81*nelmt).  cg_tol / cg_maxiter are the reference's
 * module globals (src/global.f90:96-97).  On return cg_iter, errcode
 * (0 converged, -1 not converged) and errtag (250 chars, blank padded,
 * "ERROR: PCG solver doesn't converge!" on -1) follow the reference.
 * Device copies of k, dprecon and gdof_elmt are cached on the context and
 * reused while the host pointers, sizes and a content fingerprint are
 * unchanged.  The fingerprint reads every value of dprecon and gdof_elmt
 * and 8192 evenly spaced values of k, so a rebuild of K that changes any
 * element's diagonal (any change of E or nu) is seen through dprecon.  A
 * caller that rewrites k in place without changing dprecon (same pointers)
 * calls sgeo_pcg_solver_invalidate before the next solve. */
int sgeo_pcg_solver(sgeo_ctx *ctx, int32_t neq, int32_t nelmt,
                    const double *k, double *u, const double *f,
                    const double *dprecon, const int32_t *gdof_elmt,
                    double cg_tol, int32_t cg_maxiter, int32_t *cg_iter,
                    int32_t *errcode, char *errtag);
/* Drops the cached device copies sgeo_pcg_solver keeps (no reference
 * counterpart: the reference reads its arrays at every call), so the next
 * call uploads k, dprecon and gdof_elmt again. */
int sgeo_pcg_solver_invalidate(sgeo_ctx *ctx);

/* Scratch of sgeo_pcg_solver_noalloc: a fixed device array of this many
 * bytes, one per GPU the library runs on, held until the process ends. */
#define SGEO_NOALLOC_SCRATCH_BYTES (64u << 20)

/* sgeo_pcg_solver without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same solve, same iteration order and results as
 * sgeo_pcg_solver (bit for bit, in either reduction mode), but every array
 * is a DEVICE pointer the caller owns: k (81*81*nelmt, reference layout),
 * u (neq+1, in/out), f, dprecon (neq+1), gdof_elmt (81*nelmt).  The
 * equation -> element-slot map, the work vectors and the reduction partials
 * live in the scratch array, rebuilt at every call; nothing is cached.
 * Returns SGEO_EARG, before touching the GPU, when the call needs more than
 * SGEO_NOALLOC_SCRATCH_BYTES (sgeo_pcg_solver_noalloc_bytes gives the
 * need: about 1 KB per element plus 36 bytes per equation, so up to about
 * 35,000 elements of a hexahedral mesh), and SGEO_EARG when gdof_elmt holds
 * an entry outside 0..neq.  All calls on one GPU share its scratch array:
 * calls that overlap in time on one GPU (two host threads, two contexts)
 * race and must be serialised by the caller. */
int sgeo_pcg_solver_noalloc(sgeo_ctx *ctx, int32_t neq, int32_t nelmt,
                            const double *k, double *u, const double *f,
                            const double *dprecon, const int32_t *gdof_elmt,
                            double cg_tol, int32_t cg_maxiter, int32_t *cg_iter,
                            int32_t *errcode, char *errtag);
/* Scratch bytes sgeo_pcg_solver_noalloc needs for this size (0 if invalid). */
size_t sgeo_pcg_solver_noalloc_bytes(int32_t neq, int32_t nelmt);

/* ---- ingest: GLL numbering on the device ---------------------------------
 * Replaces hex2spec + get_global + get_global_indirect_addressing
 * (src/mesh_spec.f90:17-386) and the equation numbering of apply_bc
 * (src/apply_bc.f90:36-137) for a hex8 mesh.
 *   coord8(3, nnode8), conn(8, nelmt) 1-based EXODUS corner order (the
 *   reference's g_coord / g_num before hex2spec); bc(3, nbc) rows (element
 *   1-based, face 1..6, component 0..2 = ux, uy, uz) of the face conditions.
 * Points are merged with the reference's tolerance 1e-10*|xmax - xmin|
 * (x over the hex8 nodes); nodes are numbered by first appearance, free
 * components node-major (component fastest), 1-based, 0 = fixed.
 * sgeo_spec_fetch copies any of (NULL to skip): g_num(27, nelmt) 0-based
 * node ids, gdof(3, nnode), gdof_elmt(81, nelmt), g_coord(3, nnode),
 * corners(3, 8, nelmt). */
int sgeo_hex2spec(sgeo_ctx *ctx, int32_t nelmt, int32_t nnode8, const double *coord8,
                  const int32_t *conn, int32_t nbc, const int32_t *bc, sgeo_spec **spec,
                  int32_t *nnode, int32_t *neq);
int sgeo_spec_fetch(sgeo_spec *spec, int32_t *g_num, int32_t *gdof, int32_t *gdof_elmt,
                    double *g_coord, double *corners);
int sgeo_spec_destroy(sgeo_spec *spec);

/* ---- resident mesh ------------------------------------------------------
 * A mesh keeps gdof_elmt, corner coordinates, material ids and the element
 * stiffness resident in HBM.  nelmt elements, neq equations. */
int sgeo_mesh_create(sgeo_ctx *ctx, int32_t nelmt, int32_t neq,
                     const double *corners, const int32_t *gdof_elmt,
                     const int32_t *mat_id, int32_t nmat, sgeo_mesh **mesh);
int sgeo_mesh_destroy(sgeo_mesh *mesh);
/* Builds storekm, the assembled diagonal and (if extload != NULL) the
 * gravity load on the device: src/preprocess.f90:74-148 with
 * E = ym[m], nu = nu[m], gamma = gam[m].  kprec selects FP64 or FP32
 * storage (FP32 is rounded from the FP64 element matrices).  extload
 * (host, neq+1, may be NULL) receives the gravity load; dprecon_out (host,
 * neq+1, may be NULL) receives the *inverted* diagonal (dprecon[0]=0). */
int sgeo_mesh_build_stiffness(sgeo_mesh *mesh, const double *gam,
                              const double *ym, const double *nu, int kprec,
                              double *extload, double *dprecon_out);
/* Uploads a host storekm (reference layout, FP64) and inverted dprecon. */
int sgeo_mesh_upload_stiffness(sgeo_mesh *mesh, const double *storekm,
                               const double *dprecon, int kprec);
/* Storage format of the operator the PCG streams. */
#define SGEO_KFMT_FULL 0    /* 81x81 per element, reference layout */
#define SGEO_KFMT_PACKED 1  /* upper triangle, 3328 entries per element */
/* Selects the operator used by the PCG of this mesh: precision (SGEO_K_F64 /
 * SGEO_K_F32) and format.  Derived storages are built on the device from
 * the canonical FP64 matrices (FP32 by rounding; packed keeps the upper
 * triangle k(r,c), r <= c, which the reference's matrices match to the last
 * bit only).  The full FP64 operator reproduces the reference bit for bit;
 * the others trade that for fewer bytes per iteration. */
int sgeo_mesh_set_operator(sgeo_mesh *mesh, int kprec, int kfmt);
/* Achieved HBM bandwidth of a plain device copy of `bytes` (read + write,
 * HIP events on the context's stream), for calibration. */
int sgeo_calibrate_copy(sgeo_ctx *ctx, size_t bytes, int reps, double *gbs);
/* Copies the device storekm back (FP64 meshes only). */
int sgeo_mesh_download_stiffness(sgeo_mesh *mesh, double *storekm);
/* kp = K p on the resident operator (host vectors, neq+1). */
int sgeo_mesh_matvec(sgeo_mesh *mesh, const double *p, double *kp);
/* The device's point functions of the plastic body on n records
 * rec(9, n) = (stress(6), phi, c, psi), phi and psi in degrees:
 * out(7, n) = sigm, dsbar, theta (stress_invariant, src/math_library.f90:
 * 144-186), f (mohcouf, src/plastic_library.f90:95-108), dq1, dq2, dq3
 * (mohcouq, :116-143).  Host arrays; for parity checks of the point math. */
int sgeo_constitutive(sgeo_ctx *ctx, int32_t n, const double *rec, double *out);
/* The device math library's asin, sin, cos or tan (fn 0..3) of n host
 * values (no reference counterpart: compares the device functions the
 * point math calls with the host libm the reference calls). */
int sgeo_math(sgeo_ctx *ctx, int32_t fn, int32_t n, const double *x, double *y);
/* HBM bytes held by the mesh. */
int sgeo_mesh_bytes(sgeo_mesh *mesh, size_t *bytes);
/* Node map for the nodal result fields (sgeo_slope_nodal_fields,
 * sgeo_batch_nodal_fields): g_num(27, nelmt) 0-based node ids, as
 * sgeo_spec_fetch returns them; node_valency is the number of (element,
 * point) pairs at a node (src/slope.F90:203-210).  The reference computes
 * the same in its drivers; it has no separate interface. */
int sgeo_mesh_set_nodes(sgeo_mesh *mesh, int32_t nnode, const int32_t *g_num);
/* The stress, principal and normal stress fields (as sgeo_slope_nodal_fields)
 * of a host stress state stress_elmt (6*27*nelmt, reference layout) on this
 * mesh: the excavation's step 0 plots the in-situ stress this way
 * (src/excavation.F90:303-323). */
int sgeo_mesh_nodal_fields(sgeo_mesh *mesh, const double *stress_elmt, float *stress,
                           float *psigma, float *nsigma);
/* pcg_solver (src/solver.f90:14-79) on a resident mesh: its selected
 * operator and inverted diagonal; u (host, neq+1) warm start, overwritten;
 * f (host, neq+1).  cg_iter and errcode as sgeo_pcg_solver. */
int sgeo_mesh_solve(sgeo_mesh *mesh, double *u, const double *f, double cg_tol,
                    int32_t cg_maxiter, int32_t *cg_iter, int32_t *errcode);
/* Elastic stress at every GLL point of the mesh's elements from the
 * displacement x (host, neq+1): replaces elastic_stress
 * (src/postprocess.f90:40-79) with E = ym[m], nu = nu[m]; stress (host)
 * in the reference layout (6,27,nelmt).  The excavation's in-situ stress. */
int sgeo_mesh_elastic_stress(sgeo_mesh *mesh, const double *ym,
                             const double *nu, const double *x, double *stress);
/* Excavation load of the mesh's elements (the elements a stage removes,
 * numbered with the stage's equations): replaces excavation_load_nodal and
 * the nodal-to-equation copy (src/excavation_library.f90:186-236,
 * src/serial_library.f90:110-130).  stress: host (6,27,nelmt) of these
 * elements; load: host (neq+1), load[0] = 0. */
int sgeo_mesh_excavation_load(sgeo_mesh *mesh, const double *gam,
                              const double *stress, double *load);

/* ---- Seam 2: slope plastic iteration on a resident mesh -----------------
 * State of one strength-reduction trial (src/slope.F90:244-443): x, oldx,
 * bodyload, evpt(6,27,E), stress_elmt(6,27,E) live on the device. */
int sgeo_slope_create(sgeo_mesh *mesh, const double *extload,
                      sgeo_slope **slope);
int sgeo_slope_destroy(sgeo_slope *slope);
/* Resets the trial (src/slope.F90:248-249, 281-282) and sets the factored
 * material properties of this SRF (from strength_reduction, per material)
 * and the pseudo-time step dt (dt_viscoplas). */
int sgeo_slope_begin_srf(sgeo_slope *slope, const double *ym,
                         const double *nuf, const double *phif,
                         const double *cohf, const double *psif, double dt);
/* Load increment of an excavation stage (src/excavation.F90:509-512):
 * bodyload, evpt, x and oldx reset; the stress state is kept, or replaced
 * by stress0 (host, (6,27,nelmt) reference layout) when given -- the
 * in-situ stress of the stage's intact elements.  Later plastic iterations
 * also record the von Mises plastic strain of every point at the final
 * pass (src/excavation.F90:601-602, sgeo_slope_download_vmeps). */
int sgeo_slope_begin_increment(sgeo_slope *slope, const double *ym,
                               const double *nuf, const double *phif,
                               const double *cohf, const double *psif,
                               double dt, const double *stress0);
/* sqrt(2/3 evpt.evpt) per GLL point at the last final pass, (27,nelmt). */
int sgeo_slope_download_vmeps(sgeo_slope *slope, double *vm_pt);
/* One plastic iteration (src/slope.F90:286-391): load = extload+bodyload,
 * PCG (warm start), x(0)=0, uerr, nl_isconv, then the element x GLL body.
 * is_last = (nl_iter == nl_maxiter).  Outputs: cg_iter, errcode of the
 * PCG (as sgeo_pcg_solver), uerr, nl_isconv, fmax. */
int sgeo_slope_plastic_iter(sgeo_slope *slope, double cg_tol,
                            int32_t cg_maxiter, double nl_tol, int is_last,
                            int32_t *cg_iter, int32_t *errcode, double *uerr,
                            int32_t *nl_isconv, double *fmax);
/* Full plastic loop of one SRF (the loop above until nl_isconv or
 * nl_maxiter).  Outputs the summary row of src/slope.F90:415 (cg_tot,
 * nl_tot, uxmax, umax, fmax); errcode as sgeo_pcg_solver. */
int sgeo_slope_run_srf(sgeo_slope *slope, const int32_t *gdof_node,
                       int32_t nnode, double cg_tol, int32_t cg_maxiter,
                       double nl_tol, int32_t nl_maxiter, int32_t *cg_tot,
                       int32_t *nl_tot, double *uxmax, double *umax,
                       double *fmax, int32_t *errcode);
/* Host copies of the trial state: x (neq+1), bodyload (neq+1),
 * evpt / stress_elmt (6*27*nelmt, reference layout), plastic flags
 * (27*nelmt bytes: f>=0 at the last body pass).  Any pointer may be NULL. */
int sgeo_slope_download(sgeo_slope *slope, double *x, double *bodyload,
                        double *evpt, double *stress_elmt, uint8_t *flags);
/* The nodal fields save_data writes for the trial's final stress state
 * (src/slope.F90:418-439, src/excavation.F90:655-687, src/postprocess.f90:
 * 128-248), computed on the device and rounded to FP32 as the EnSight files
 * hold them, in component blocks of nnode values: stress (6 blocks) the
 * mean of stress_elmt over a node's elements summed in element order,
 * psigma (3 blocks) the principal stresses ascending, nsigma = (psigma1 +
 * psigma3) / 2, scf and vmeps the stress concentration factor and the
 * plastic strain.  A slope trial updates neither (scf = 1e32, vmeps = 0,
 * src/slope.F90:249, 366-376); an excavation stage gives what
 * sgeo_slope_accumulate_nodal gathered over its increments.  Any pointer
 * may be NULL; needs sgeo_mesh_set_nodes. */
int sgeo_slope_nodal_fields(sgeo_slope *slope, float *stress, float *psigma,
                            float *nsigma, float *scf, float *vmeps);
/* After each load increment of an excavation stage (src/excavation.F90:
 * 599-609): per node, adds the increment's per-point plastic strain
 * sqrt(2/3 evpt.evpt) and lowers scf to the points' stress concentration
 * factor, in element order; first != 0 starts the stage from 0 and 1e32
 * (src/excavation.F90:335).  Needs an increment begun with
 * sgeo_slope_begin_increment and sgeo_mesh_set_nodes.  The first call, and
 * the first after sgeo_mesh_set_nodes changed the node count, must pass
 * first != 0 (SGEO_ESTATE otherwise). */
int sgeo_slope_accumulate_nodal(sgeo_slope *slope, int32_t first);

/* ---- multi-trial mixed-precision engine ---------------------------------
 * Up to 64 (a power of two) strength-reduction trials on one resident mesh,
 * sharing its elastic stiffness.  Each PCG call of a trial is a Jacobi-PCG
 * whose iteration runs in FP32 (packed FP32 operator on MFMA, FP32 r, p,
 * K p) with FP64 dot products and an FP64 solution x.  The FP64 residual
 * f - K x (packed FP64 operator) starts the solve and replaces the FP32
 * recursive residual whenever r.z has dropped by inner_tol since the last
 * replacement; the search direction is kept across replacements.  The solve
 * stops by the reference's own test |alpha| max|p| / max|x| <= cg_tol
 * (src/solver.f90:65), so cg_iter counts CG iterations as the reference
 * does.  The yield/return-mapping step stays FP64.  Each column's
 * arithmetic is independent of the others and of ncol, so a trial gives the
 * same bits alone and in any batch.  (No reference counterpart: the
 * reference runs one trial at a time.) */
int sgeo_batch_create(sgeo_mesh *mesh, int32_t ncol, const double *extload,
                      sgeo_batch **batch);
int sgeo_batch_destroy(sgeo_batch *batch);
int sgeo_batch_bytes(sgeo_batch *batch, size_t *bytes);
/* Starts a trial in column col (src/slope.F90:248-249, 281-282) with the
 * factored materials of its SRF and dt.  The first PCG call of every trial
 * solves K x = extload from x = 0 (bodyload is zero), the same system for
 * every trial of the mesh: the engine keeps the first such solution and
 * seeds later trials with it (their first call then reports its iteration
 * count).  x0 (may be NULL) seeds x explicitly in the same way, reported
 * with x0_cg_iter iterations. */
int sgeo_batch_begin(sgeo_batch *batch, int32_t col, const double *ym,
                     const double *nuf, const double *phif, const double *cohf,
                     const double *psif, double dt, const double *x0,
                     int32_t x0_cg_iter);
int sgeo_batch_release(sgeo_batch *batch, int32_t col);
/* One plastic iteration of every running column in lockstep
 * (src/slope.F90:286-391); per-column arrays of length ncol (entries of
 * other columns untouched). */
int sgeo_batch_plastic_iter(sgeo_batch *batch, const int32_t *is_last,
                            double cg_tol, int32_t cg_maxiter, double inner_tol,
                            double nl_tol, int32_t *cg_iter, int32_t *errcode,
                            double *uerr, int32_t *nl_isconv, double *fmax);
/* Runs every running column through its own plastic iterations
 * (src/slope.F90:285-392, is_last = the nl_maxiter-th), columns advancing
 * independently: a column whose PCG call finished gets its body pass and
 * starts its next call while the others keep iterating.  Returns when at
 * least one trial ended (nl_isconv, nl_maxiter reached, or a PCG call that
 * did not converge: errcode -1), when the elastic solution has just become
 * available (memo_ready), when no column runs, or once max_steps CG steps
 * ran (never more: max_steps >= 1 is the exact count of a call that none of
 * the other conditions ends early).
 * ended[col] = 1 marks the trials that ended in this call, with their
 * summary row (src/slope.F90:415): nl_tot, cg_tot, fmax; the column keeps
 * its state (sgeo_batch_download) until the next begin. */
int sgeo_batch_advance(sgeo_batch *batch, double cg_tol, int32_t cg_maxiter,
                       double inner_tol, double nl_tol, int32_t nl_maxiter,
                       int64_t max_steps, int32_t *ended, int32_t *nl_tot,
                       int32_t *cg_tot, double *fmax, int32_t *errcode,
                       int32_t *memo_ready);
int sgeo_batch_download(sgeo_batch *batch, int32_t col, double *x,
                        uint8_t *flags);
/* sgeo_slope_nodal_fields for the trial held by column col (an excavation
 * stage's scf and vmeps from sgeo_batch_accumulate_nodal). */
int sgeo_batch_nodal_fields(sgeo_batch *batch, int32_t col, float *stress, float *psigma,
                            float *nsigma, float *scf, float *vmeps);
/* Excavation load increment on column col, the mixed-precision form of
 * sgeo_slope_begin_increment (src/excavation.F90:509-515): materials and
 * dt, evpt, bodyload, x and oldx reset, the stress state set from
 * stress0 (host, 6*27*nelmt, reference layout) or, with NULL, kept from
 * the column's last increment; the body pass then also records each
 * point's plastic strain and stress concentration factor at convergence
 * (src/excavation.F90:599-609).  The increment's first PCG call solves the
 * same system for every increment of a stage, so the engine's elastic
 * memo serves increments after the first. */
int sgeo_batch_begin_increment(sgeo_batch *batch, int32_t col, const double *ym,
                               const double *nuf, const double *phif, const double *cohf,
                               const double *psif, double dt, const double *stress0);
/* Column col's stress_elmt (6*27*nelmt, reference layout) to the host. */
int sgeo_batch_download_stress(sgeo_batch *batch, int32_t col, double *stress_elmt);
/* sgeo_slope_accumulate_nodal for column col's excavation increment (the
 * same first != 0 rule, per column). */
int sgeo_batch_accumulate_nodal(sgeo_batch *batch, int32_t col, int32_t first);
/* Plastic iterations done so far by each column's current trial (ncol
 * values; 0 where no trial runs), as of the last body pass.  Lets a
 * scheduler see a trial heading for nl_maxiter (a failing SRF) before it
 * ends.  No device access. */
int sgeo_batch_nl_iters(sgeo_batch *batch, int32_t *nl_iter);
/* CG steps run as captured HIP graphs of 8, 16 or 32 steps (default, fewer
 * launch submissions for small meshes) or as separate launches (enable 0;
 * HIP-event timing with sgeo_profile always uses separate launches).  The
 * results are the same bits either way. */
int sgeo_batch_set_graphs(sgeo_batch *batch, int enable);
/* Launch-geometry autotuner of the MFMA element product (no counterpart in
 * the reference, which has no GPU product).  Times `reps` launches of each
 * candidate on this engine's mesh and iterating columns -- blocks per CU
 * 4, 3, 2, 1 with the XCD-aware element split, then 4, 3, 2, 1 with a plain
 * grid-stride walk (ms_table[8], in that order, may be NULL) -- and keeps the
 * fastest for later steps.  Needs at least one column iterating
 * (SGEO_ESTATE otherwise).  The geometry never changes a result bit, and
 * tuning between steps does not change the trials' results. */
int sgeo_batch_autotune(sgeo_batch *batch, int32_t reps, int32_t *blocks_per_cu,
                        int32_t *xcd_split, double *ms_best, double *ms_table);
/* Sets that geometry directly (blocks per CU 1..4, XCD split 0/1); the
 * default is 3 blocks per CU with the XCD split. */
int sgeo_batch_set_geometry(sgeo_batch *batch, int32_t blocks_per_cu, int32_t xcd_split);
/* FP64 residuals (the product on the packed FP64 operator that starts a
 * column's PCG call or replaces its residual) are formed on every period-th
 * CG step of the engine only (1, 2, 4 or 8; default 1), so one pass over the
 * FP64 operator serves the columns of several steps.  A starting column
 * waits for the next such step; a column due for a replacement keeps
 * iterating until then.  Results depend on the period (the iterate at which
 * a residual is replaced), never on the batch or on the other columns.
 * Set it before starting trials.  No counterpart in the reference. */
int sgeo_batch_set_residual_period(sgeo_batch *batch, int32_t period);
/* With a residual period P > 1, x is written on every L-th iteration only
 * (L = P, at most 8, or 2 when 8 (neq + 1) ncol reaches 2^32): that
 * iteration applies the FP64 updates of the L - 1 before it in their order,
 * so x keeps its bits, and the other iterations read neither x nor p.
 * enable = 0 writes x on every iteration (default 1).  SGEO_ESTATE when it
 * would change while trials run.  No counterpart in the reference. */
int sgeo_batch_set_deferred_x(sgeo_batch *batch, int32_t enable);
/* Totals since creation, as of the last host synchronisation: CG
 * iterations of all trials (seeded elastic solutions not counted), FP64
 * residuals of all trials, and batched steps (one step = one CG iteration
 * of every running column).  sgeo_profile times the batched element-product
 * launches (k_bmv_mfma) like the single-trial one. */
int sgeo_batch_stats(sgeo_batch *batch, int64_t *inner_iters,
                     int64_t *outer_iters, int64_t *steps);

#ifdef __cplusplus
}
#endif
#endif /* SGEO_H */
