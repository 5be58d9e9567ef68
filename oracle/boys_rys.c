/* Oracle (test infrastructure only): Boys function and Rys quadrature.
 *
 * Restates the quantities libcint computes in
 *   src/gradscf/integrals/_native/vendor/libcint/src/fmt.c:186-226   (gamma_inc_like)
 *   src/gradscf/integrals/_native/vendor/libcint/src/rys_roots.c:57-125 (CINTrys_roots)
 * by a different, textbook route, in long double:
 *   F_m(x)  : power series e^{-x} sum_k (2x)^k / ((2m+1)(2m+3)...(2m+2k+1)) with
 *             downward recursion for x < 45; erf and upward recursion above.
 *   roots   : Gauss quadrature of the weight exp(-x t^2) dt on [0,1], written in
 *             s = t^2.  The measure is discretised by Gauss-Legendre nodes in t
 *             (the discretised Stieltjes procedure), the three-term recurrence
 *             of its orthonormal polynomials is built, the Jacobi-matrix
 *             eigenvalues are found by Sturm-sequence bisection and the weights
 *             from the Christoffel function 1 / sum_j q_j(s_i)^2.
 */
#include <math.h>
#include <pthread.h>
#include <string.h>
#include "oracle.h"

#define PI_L 3.141592653589793238462643383279502884L

void oracle_boys_ld(int mmax, long double x, long double* f) {
  const long double ex = expl(-x);
  if (x < 45.0L) {
    long double term = 1.0L / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 4000; ++k) {
      term *= 2.0L * x / (2 * mmax + 2 * k + 1);
      sum += term;
      if (term < sum * 1e-22L) break;
    }
    f[mmax] = ex * sum;
    for (int m = mmax; m > 0; --m) f[m - 1] = (2.0L * x * f[m] + ex) / (2 * m - 1);
  } else {
    f[0] = 0.5L * sqrtl(PI_L / x) * erfl(sqrtl(x));
    for (int m = 0; m < mmax; ++m) f[m + 1] = ((2 * m + 1) * f[m] - ex) / (2.0L * x);
  }
}

void oracle_boys(int mmax, double x, double* f) {
  long double t[64];
  if (mmax > 63) mmax = 63;
  oracle_boys_ld(mmax, (long double)x, t);
  for (int m = 0; m <= mmax; ++m) f[m] = (double)t[m];
}

/* Gauss-Legendre nodes/weights on [0,1] for M = 32, 64, 128 (Newton on P_M). */
static long double gl_x[3][128], gl_w[3][128];
static const int gl_m[3] = {32, 64, 128};
static pthread_once_t gl_once = PTHREAD_ONCE_INIT;

static void gl_init(void) {
  for (int t = 0; t < 3; ++t) {
    const int M = gl_m[t];
    for (int i = 0; i < M; ++i) {
      long double z = cosl(PI_L * (i + 0.75L) / (M + 0.5L)), dp = 0;
      for (int it = 0; it < 100; ++it) {
        long double p0 = 1.0L, p1 = z;
        for (int k = 2; k <= M; ++k) {
          long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
          p0 = p1; p1 = p2;
        }
        dp = M * (z * p1 - p0) / (z * z - 1.0L);
        long double dz = p1 / dp;
        z -= dz;
        if (fabsl(dz) < 1e-21L) break;
      }
      {
        long double p0 = 1.0L, p1 = z;
        for (int k = 2; k <= M; ++k) {
          long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
          p0 = p1; p1 = p2;
        }
        dp = M * (z * p1 - p0) / (z * z - 1.0L);
      }
      /* map [-1,1] -> [0,1] */
      gl_x[t][i] = 0.5L * (1.0L - z);
      gl_w[t][i] = 1.0L / ((1.0L - z * z) * dp * dp);
    }
  }
}

static int sturm_count(int n, const long double* a, const long double* b2, long double lam) {
  int c = 0;
  long double d = a[0] - lam;
  if (d < 0) ++c;
  for (int i = 1; i < n; ++i) {
    if (d == 0) d = 1e-4000L;
    d = (a[i] - lam) - b2[i] / d;
    if (d < 0) ++c;
  }
  return c;
}

int oracle_rys_roots_ld(int n, long double x, long double* s_out, long double* w_out) {
  if (n < 1 || n > ORC_MAX_ROOTS) return -1;
  pthread_once(&gl_once, gl_init);
  long double b = 1.0L, aeff = x;
  if (x > 90.0L) { b = sqrtl(90.0L / x); aeff = 90.0L; }
  const int t = aeff <= 5.0L ? 0 : (aeff <= 30.0L ? 1 : 2);
  const int M = gl_m[t];
  long double W[128], S[128], p[128], pp[128];
  long double mu0 = 0;
  for (int k = 0; k < M; ++k) {
    const long double tk = b * gl_x[t][k];
    S[k] = tk * tk;
    W[k] = b * gl_w[t][k] * expl(-x * S[k]);
    mu0 += W[k];
  }
  long double a[ORC_MAX_ROOTS], bo[ORC_MAX_ROOTS + 1], b2[ORC_MAX_ROOTS];
  const long double q0 = 1.0L / sqrtl(mu0);
  for (int k = 0; k < M; ++k) { p[k] = q0; pp[k] = 0; }
  bo[0] = 0;
  for (int j = 0; j < n; ++j) {
    long double aj = 0;
    for (int k = 0; k < M; ++k) aj += W[k] * S[k] * p[k] * p[k];
    a[j] = aj;
    if (j == n - 1) break;
    long double nrm = 0;
    for (int k = 0; k < M; ++k) {
      const long double r = (S[k] - aj) * p[k] - bo[j] * pp[k];
      pp[k] = p[k];
      p[k] = r;
      nrm += W[k] * r * r;
    }
    bo[j + 1] = sqrtl(nrm);
    for (int k = 0; k < M; ++k) p[k] /= bo[j + 1];
  }
  b2[0] = 0;
  for (int j = 1; j < n; ++j) b2[j] = bo[j] * bo[j];
  for (int i = 0; i < n; ++i) {
    long double lo = 0.0L, hi = 1.0L;
    for (int it = 0; it < 200; ++it) {
      const long double mid = 0.5L * (lo + hi);
      if (mid == lo || mid == hi) break;
      if (sturm_count(n, a, b2, mid) > i) hi = mid; else lo = mid;
    }
    const long double lam = 0.5L * (lo + hi);
    long double qm = 0, q = q0, sum = q0 * q0;
    for (int j = 0; j + 1 < n; ++j) {
      const long double qn = ((lam - a[j]) * q - bo[j] * qm) / bo[j + 1];
      qm = q; q = qn;
      sum += q * q;
    }
    s_out[i] = lam;
    w_out[i] = 1.0L / sum;
  }
  return 0;
}

int oracle_rys_roots(int n, double x, double* t2, double* w) {
  long double s[ORC_MAX_ROOTS], ww[ORC_MAX_ROOTS];
  int rc = oracle_rys_roots_ld(n, (long double)x, s, ww);
  if (rc) return rc;
  for (int i = 0; i < n; ++i) { t2[i] = (double)s[i]; w[i] = (double)ww[i]; }
  return 0;
}
