/* Sensitivity probe for golden fixtures (CPU only, never linked into the
 * library): preloaded into a run of the compiled reference, it moves the
 * libm result of asin (SGEO_ULP_FUNCS=asin) or of asin, sin, cos, tan and
 * sincos (SGEO_ULP_FUNCS=all) by one ulp, up or down, for about 1/64 of the
 * arguments, or 1/n with SGEO_ULP_EVERY=n (chosen by a hash of the
 * argument's bits, so a run is deterministic).  The spread of the reference's own results under such
 * last-bit changes is the parity tolerance for results that pass through
 * transcendental functions: a GPU's device math library differs from the
 * host's libm in the last bit on a fraction of arguments.
 *
 *   gcc -O2 -fPIC -shared -o ulp_shim.so ulp_shim.c -ldl -lm
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int mode(void) { /* 1: asin only, 2: all */
  static int m = 0;
  if (!m) {
    const char *s = getenv("SGEO_ULP_FUNCS");
    m = (s && strcmp(s, "all") == 0) ? 2 : 1;
  }
  return m;
}

/* SGEO_ULP_EVERY=n moves one argument in n (n a power of 2, default 64) */
static uint64_t every_mask(void) {
  static uint64_t m = 0;
  if (!m) {
    const char *s = getenv("SGEO_ULP_EVERY");
    const long n = s ? atol(s) : 64;
    m = (n >= 1 && (n & (n - 1)) == 0) ? (uint64_t)n : 64;
  }
  return m - 1;
}

static double bump(double x, double r) {
  uint64_t u;
  memcpy(&u, &x, 8);
  u ^= u >> 29;
  u *= 0x9E3779B97F4A7C15ull;
  u ^= u >> 32;
  const uint64_t m = every_mask();
  if ((u & m) != 0) return r;
  return nextafter(r, (u & (m + 1)) ? INFINITY : -INFINITY);
}

typedef double (*fn1)(double);

#define WRAP(f, need)                                      \
  double f(double x) {                                     \
    static fn1 p;                                          \
    if (!p) p = (fn1)dlsym(RTLD_NEXT, #f);                 \
    const double r = p(x);                                 \
    return mode() >= (need) ? bump(x, r) : r;              \
  }

WRAP(asin, 1)
WRAP(sin, 2)
WRAP(cos, 2)
WRAP(tan, 2)

void sincos(double x, double *s, double *c) {
  static void (*p)(double, double *, double *);
  if (!p) p = (void (*)(double, double *, double *))dlsym(RTLD_NEXT, "sincos");
  p(x, s, c);
  if (mode() >= 2) {
    *s = bump(x, *s);
    *c = bump(-x, *c);
  }
}
