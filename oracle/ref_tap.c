/*
 * Link-time tap around the unmodified reference solver (test infrastructure).
 *
 * Built with -DSGEO_TAP_SOLVER and linked into oracle/_ref/semgeotech_tap with
 *   -Wl,--wrap=_QMsolverPpcg_solver -Wl,--wrap=_QMplastic_libraryPmohcouf
 * so every call the reference makes to pcg_solver (src/solver.f90:14) and
 * mohcouf (src/plastic_library.f90:95) passes through here first.  The
 * reference sources are not edited.
 *
 * Environment:
 *   SGEO_TAP_DIR   directory for the dumps (unset: tap is transparent)
 *   SGEO_TAP_MODE  "full"  dump operands and results of every pcg_solver call
 *                          and the yield value of every mohcouf call
 *                  "time"  only log per-call wall time and cg_iter
 *
 * Files written in SGEO_TAP_DIR:
 *   pcg_static.bin  (full) k(81,81,nelmt), dprecon(0:neq), gdof_elmt(81,nelmt)
 *                   of the first call, re-dumped (appended) when they change
 *   pcg_calls.bin   (full) per call: int32 {neq,nelmt,cg_iter,errcode},
 *                   double u_in(0:neq), f(0:neq), u_out(0:neq)
 *   pcg_times.txt   per call: call index, cg_iter, errcode, seconds
 *   mohcouf.bin     (full) double f per call, in call order
 *
 * read_input (src/read_input.f90:13) is wrapped as well: it tests
 * `present(ispartmesh).and.ispartmesh` (:153, :155, :186), which flang does
 * not short-circuit, so an absent optional argument is dereferenced.  The
 * wrapper passes an explicit .false., which every use in read_input treats
 * exactly like an absent argument (:98-99, :153, :155, :186, :532).
 * It also closes unit 11, which read_input leaves connected to the material
 * list and the driver re-opens with status='replace'
 * (src/semgeotech.F90:143); see oracle/ref_fix_units.f90.
 *
 * The pcg_solver dummy gdof_elmt is assumed-shape (src/solver.f90:23), so
 * flang passes a descriptor; it is forwarded untouched and read through the
 * ISO_Fortran_binding layout (which flang's runtime descriptor follows).
 */
#define _POSIX_C_SOURCE 200809L
#include <ISO_Fortran_binding.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

extern void __real__QMsolverPpcg_solver(int32_t *neq, int32_t *nelmt, double *k,
                                        double *u, double *f, double *dprecon,
                                        CFI_cdesc_t *gdof_elmt, int32_t *cg_iter,
                                        int32_t *errcode, char *errtag,
                                        size_t errtag_len);
extern void __real__QMplastic_libraryPmohcouf(double *phi, double *c,
                                              double *sigm, double *dsbar,
                                              double *theta, double *f);

extern void __real__QMinputPread_input(void *a0, void *a1, void *a2, void *a3,
                                       void *a4, void *a5);

extern void sgeo_ref_close_unit11(void);

/* (inp_fname, errcode, errtag, ispartmesh, hidden lengths...) */
void __wrap__QMinputPread_input(void *a0, void *a1, void *a2, void *a3,
                                void *a4, void *a5) {
  static int32_t logical_false = 0;
  __real__QMinputPread_input(a0, a1, a2, a3 ? a3 : (void *)&logical_false, a4,
                             a5);
  sgeo_ref_close_unit11();
}

#ifdef SGEO_TAP_SOLVER
static int tap_state = -1; /* -1 unknown, 0 off, 1 time, 2 full */
static char tap_dir[1024];
static long ncall = 0;
static uint64_t last_sig = 0;
static FILE *fmoh = NULL;

static void tap_init(void) {
  if (tap_state >= 0) return;
  const char *d = getenv("SGEO_TAP_DIR");
  if (!d || !*d) { tap_state = 0; return; }
  snprintf(tap_dir, sizeof tap_dir, "%s", d);
  const char *m = getenv("SGEO_TAP_MODE");
  tap_state = (m && strcmp(m, "full") == 0) ? 2 : 1;
}

static FILE *tap_open(const char *name, const char *mode) {
  char path[1200];
  snprintf(path, sizeof path, "%s/%s", tap_dir, name);
  FILE *fp = fopen(path, mode);
  if (!fp) { fprintf(stderr, "ref_tap: cannot open %s\n", path); exit(3); }
  return fp;
}

static uint64_t fnv(const void *p, size_t n, uint64_t h) {
  const unsigned char *b = (const unsigned char *)p;
  for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ULL; }
  return h;
}

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void __wrap__QMsolverPpcg_solver(int32_t *neq, int32_t *nelmt, double *k,
                                 double *u, double *f, double *dprecon,
                                 CFI_cdesc_t *gdof_elmt, int32_t *cg_iter,
                                 int32_t *errcode, char *errtag,
                                 size_t errtag_len) {
  tap_init();
  if (tap_state == 0) {
    __real__QMsolverPpcg_solver(neq, nelmt, k, u, f, dprecon, gdof_elmt,
                                cg_iter, errcode, errtag, errtag_len);
    return;
  }
  const int64_t n1 = (int64_t)(*neq) + 1;
  const int64_t ne = *nelmt;
  const int64_t nedof = gdof_elmt->dim[0].extent;
  double *u_in = NULL;
  if (tap_state == 2) {
    /* contiguous copy of the assumed-shape gdof_elmt(:,:) */
    int32_t *g = (int32_t *)malloc(sizeof(int32_t) * nedof * ne);
    const char *base = (const char *)gdof_elmt->base_addr;
    for (int64_t e = 0; e < ne; e++)
      for (int64_t i = 0; i < nedof; i++)
        g[e * nedof + i] = *(const int32_t *)(base + i * gdof_elmt->dim[0].sm +
                                              e * gdof_elmt->dim[1].sm);
    uint64_t sig = 1469598103934665603ULL;
    sig = fnv(k, sizeof(double) * nedof * nedof * ne, sig);
    sig = fnv(dprecon, sizeof(double) * n1, sig);
    sig = fnv(g, sizeof(int32_t) * nedof * ne, sig);
    if (sig != last_sig) {
      FILE *fs = tap_open("pcg_static.bin", "ab");
      int32_t hdr[4] = {*neq, *nelmt, (int32_t)nedof, (int32_t)ncall};
      fwrite(hdr, sizeof hdr, 1, fs);
      fwrite(k, sizeof(double), nedof * nedof * ne, fs);
      fwrite(dprecon, sizeof(double), n1, fs);
      fwrite(g, sizeof(int32_t), nedof * ne, fs);
      fclose(fs);
      last_sig = sig;
    }
    free(g);
    u_in = (double *)malloc(sizeof(double) * n1);
    memcpy(u_in, u, sizeof(double) * n1);
  }
  double t0 = now();
  __real__QMsolverPpcg_solver(neq, nelmt, k, u, f, dprecon, gdof_elmt, cg_iter,
                              errcode, errtag, errtag_len);
  double t1 = now();
  FILE *ft = tap_open("pcg_times.txt", "a");
  fprintf(ft, "%ld %d %d %.6e\n", ncall, *cg_iter, *errcode, t1 - t0);
  fclose(ft);
  if (tap_state == 2) {
    FILE *fc = tap_open("pcg_calls.bin", "ab");
    int32_t hdr[4] = {*neq, *nelmt, *cg_iter, *errcode};
    fwrite(hdr, sizeof hdr, 1, fc);
    fwrite(u_in, sizeof(double), n1, fc);
    fwrite(f, sizeof(double), n1, fc);
    fwrite(u, sizeof(double), n1, fc);
    fclose(fc);
    free(u_in);
  }
  if (fmoh) fflush(fmoh);
  ncall++;
}

void __wrap__QMplastic_libraryPmohcouf(double *phi, double *c, double *sigm,
                                       double *dsbar, double *theta,
                                       double *f) {
  __real__QMplastic_libraryPmohcouf(phi, c, sigm, dsbar, theta, f);
  tap_init();
  if (tap_state == 2) {
    if (!fmoh) fmoh = tap_open("mohcouf.bin", "ab");
    fwrite(f, sizeof(double), 1, fmoh);
  }
}
#endif /* SGEO_TAP_SOLVER */
