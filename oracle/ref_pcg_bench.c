/*
 * Times the reference's own compiled pcg_solver (src/solver.f90:14-79, from
 * the unmodified solver.o that oracle/Makefile builds) on a stored system:
 * the CPU baseline of the EBE-PCG on the large meshes, where running the
 * whole reference (its serial stiffness build alone takes minutes) is not a
 * bounded sample.  Test and benchmark infrastructure only.
 *
 *   ref_pcg_bench DIR NEQ NELMT ITERS
 *
 * DIR holds k.bin (k(81,81,nelmt), the reference layout), f.bin and
 * dinv.bin ((0:neq) doubles, dprecon already inverted) and gdof.bin
 * (gdof_elmt(81,nelmt) int32), memory-mapped read-only so that concurrent
 * processes share one copy.  u starts at zero; cg_tol = 1e-300 keeps the
 * solve from converging, so exactly ITERS iterations run (errcode -1).
 * Prints "cg_iter errcode seconds" of the call.
 */
#define _POSIX_C_SOURCE 200809L
#include <ISO_Fortran_binding.h>
#include <fcntl.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

extern void _QMsolverPpcg_solver(int32_t *neq, int32_t *nelmt, double *k, double *u,
                                 double *f, double *dprecon, CFI_cdesc_t *gdof_elmt,
                                 int32_t *cg_iter, int32_t *errcode, char *errtag,
                                 size_t errtag_len);
/* module globals the solver reads (src/global.f90:96-97) */
extern int32_t _QMglobalEcg_maxiter;
extern double _QMglobalEcg_tol;
extern int32_t _QMglobalEnedof;

static void *map(const char *dir, const char *name, size_t bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  int fd = open(path, O_RDONLY);
  if (fd < 0) { perror(path); exit(2); }
  struct stat st;
  if (fstat(fd, &st) != 0 || (size_t)st.st_size != bytes) {
    fprintf(stderr, "%s: expected %zu bytes\n", path, bytes);
    exit(2);
  }
  void *p = mmap(NULL, bytes, PROT_READ, MAP_SHARED, fd, 0);
  close(fd);
  if (p == MAP_FAILED) { perror("mmap"); exit(2); }
  return p;
}

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s DIR NEQ NELMT ITERS\n", argv[0]);
    return 2;
  }
  int32_t neq = atoi(argv[2]), nelmt = atoi(argv[3]), iters = atoi(argv[4]);
  if (neq <= 0 || nelmt <= 0 || iters <= 0) return 2;
  const size_t n1 = (size_t)neq + 1;
  double *k = map(argv[1], "k.bin", (size_t)nelmt * 6561 * 8);
  double *f = map(argv[1], "f.bin", n1 * 8);
  double *dinv = map(argv[1], "dinv.bin", n1 * 8);
  int32_t *gdof = map(argv[1], "gdof.bin", (size_t)nelmt * 81 * 4);
  double *u = calloc(n1, sizeof(double));
  if (!u) return 3;
  CFI_CDESC_T(2) desc;
  CFI_index_t ext[2] = {81, nelmt};
  if (CFI_establish((CFI_cdesc_t *)&desc, gdof, CFI_attribute_other, CFI_type_int32_t,
                    sizeof(int32_t), 2, ext) != CFI_SUCCESS) {
    fprintf(stderr, "CFI_establish failed\n");
    return 3;
  }
  _QMglobalEcg_maxiter = iters;
  _QMglobalEcg_tol = 1e-300;
  _QMglobalEnedof = 81;
  int32_t cg_iter = 0, errcode = 0;
  char errtag[250];
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  _QMsolverPpcg_solver(&neq, &nelmt, k, u, f, dinv, (CFI_cdesc_t *)&desc, &cg_iter,
                       &errcode, errtag, sizeof errtag);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  printf("%d %d %.6f\n", cg_iter, errcode,
         (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec));
  free(u);
  return 0;
}
