/* Oracle (test infrastructure only): restatement of the reference direct-J/K driver.
 *
 *   table validation + AO offsets  src/gradscf/integrals/_native/csrc/tables.h:31-76
 *   Schwarz bounds                 csrc/coulomb_shells.h:60-77
 *   canonical unique-quartet loop  csrc/coulomb_shells.h:79-128
 *   row-ownership workers + 8-fold digestion   csrc/direct_jk.cc:7-95
 *   coordinate VJP (all s1 quartets, int2e_ip1, four permutations)
 *                                  csrc/direct_jk_geometry.cc:68-84,
 *                                  csrc/eri_geometry_entries.h:5-31
 *
 * The ERI blocks come from the oracle's own Rys code (backend 0) or, for
 * validation and as the CPU baseline, from the reference's libcint compiled
 * from its sources into oracle/_ref (backend 1, loaded with dlopen).
 */
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static __thread char err_buf[256];
void oracle_set_error(const char* msg) { snprintf(err_buf, sizeof err_buf, "%s", msg); }
const char* oracle_last_error(void) { return err_buf; }

typedef int (*intor_t)(double*, int*, int*, const int*, int, const int*, int, const double*, void*, double*);
typedef void (*optfn_t)(void**, const int*, int, const int*, int, const double*);
typedef void (*delopt_t)(void**);
static intor_t lc_cart, lc_sph, lc_ip1_cart, lc_ip1_sph;
static optfn_t lc_opt;
static delopt_t lc_del;
static int backend_sel = 0;

int oracle_use_libcint(const char* path) {
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) { oracle_set_error(dlerror()); return -1; }
  lc_cart = (intor_t)dlsym(h, "int2e_cart");
  lc_sph = (intor_t)dlsym(h, "int2e_sph");
  lc_ip1_cart = (intor_t)dlsym(h, "int2e_ip1_cart");
  lc_ip1_sph = (intor_t)dlsym(h, "int2e_ip1_sph");
  lc_opt = (optfn_t)dlsym(h, "int2e_optimizer");
  lc_del = (delopt_t)dlsym(h, "CINTdel_optimizer");
  if (!lc_cart || !lc_sph || !lc_opt || !lc_del || !lc_ip1_cart || !lc_ip1_sph) {
    oracle_set_error("libcint symbols missing");
    return -1;
  }
  return 0;
}

void oracle_select_backend(int backend) { backend_sel = backend; }

typedef struct {
  const int *atm, *bas;
  const double* env;
  int natm, nbas, nenv, cart, n;
  int* ao;
  void* opt;
} tables_t;

static int validate(tables_t* t, const int* atm, int natm, const int* bas, int nbas,
                    const double* env, int nenv, int cart) {
  t->atm = atm; t->bas = bas; t->env = env;
  t->natm = natm; t->nbas = nbas; t->nenv = nenv; t->cart = cart;
  t->opt = NULL; t->ao = NULL;
  if (nenv < 20 || (cart != 0 && cart != 1)) {
    oracle_set_error("Invalid libcint table dimensions or operator");
    return -1;
  }
  for (int i = 0; i < nenv; ++i)
    if (!isfinite(env[i])) { oracle_set_error("Native integral env must be finite"); return -1; }
  if (env[8] != 0.0) {
    oracle_set_error("Range-separated integrals require the full ERI path and nonnegative omega");
    return -1;
  }
  for (int a = 0; a < natm; ++a) {
    const int* row = atm + a * ORC_ATM_SLOTS;
    if (row[1] < 20 || row[1] + 3 > nenv || row[2] != 1) {
      oracle_set_error("Native backend requires valid point nuclei");
      return -1;
    }
  }
  t->ao = (int*)malloc(sizeof(int) * (nbas + 1));
  t->ao[0] = 0;
  for (int s = 0; s < nbas; ++s) {
    const int* row = bas + s * ORC_BAS_SLOTS;
    const int l = row[1], np = row[2], nc = row[3];
    if (row[0] < 0 || row[0] >= natm || l < 0 || l > 12 || np < 1 || np > 64 || nc < 1 || nc > 64 ||
        row[4] != 0 || row[5] < 20 || row[5] + np > nenv || row[6] < 20 || row[6] + np * nc > nenv) {
      oracle_set_error("Invalid or unsupported libcint basis table");
      free(t->ao);
      t->ao = NULL;
      return -1;
    }
    for (int p = 0; p < np; ++p)
      if (env[row[5] + p] <= 0) {
        oracle_set_error("Gaussian exponents must be positive");
        free(t->ao);
        t->ao = NULL;
        return -1;
      }
    t->ao[s + 1] = t->ao[s] + nc * (cart ? (l + 1) * (l + 2) / 2 : 2 * l + 1);
  }
  t->n = t->ao[nbas];
  if (t->n < 1) { oracle_set_error("Invalid AO count"); free(t->ao); t->ao = NULL; return -1; }
  if (backend_sel == 1) lc_opt(&t->opt, atm, natm, bas, nbas, env);
  return 0;
}

static void release(tables_t* t) {
  if (t->opt && lc_del) lc_del(&t->opt);
  free(t->ao);
}

static int block(const tables_t* t, int* sh, double* out, double* cache, int deriv) {
  if (backend_sel == 1) {
    intor_t f = deriv ? (t->cart ? lc_ip1_cart : lc_ip1_sph) : (t->cart ? lc_cart : lc_sph);
    f(out, NULL, sh, t->atm, t->natm, t->bas, t->nbas, t->env, deriv ? NULL : t->opt, cache);
    return 0;
  }
  return oracle_eri_block(sh, t->atm, t->natm, t->bas, t->nbas, t->env, t->cart, deriv, out);
}

static size_t cache_len(const tables_t* t) {
  (void)t;
  return 1u << 22;
}

static size_t max_block(const tables_t* t) {
  int m = 0;
  for (int s = 0; s < t->nbas; ++s) m = t->ao[s + 1] - t->ao[s] > m ? t->ao[s + 1] - t->ao[s] : m;
  return (size_t)m * m * m * m;
}

static int schwarz(const tables_t* t, double* q) {
  double* data = (double*)malloc(sizeof(double) * max_block(t));
  double* cache = backend_sel == 1 ? (double*)malloc(sizeof(double) * cache_len(t)) : NULL;
  int rc = 0;
  for (int i = 0; i < t->nbas && !rc; ++i)
    for (int j = 0; j <= i; ++j) {
      int sh[4] = {i, j, i, j};
      const int di = t->ao[i + 1] - t->ao[i], dj = t->ao[j + 1] - t->ao[j];
      if ((rc = block(t, sh, data, cache, 0))) break;
      double largest = 0;
      for (int p = 0; p < di; ++p)
        for (int qq = 0; qq < dj; ++qq) {
          const double v = fabs(data[(((size_t)qq * di + p) * dj + qq) * di + p]);
          if (v > largest) largest = v;
        }
      q[(size_t)i * t->nbas + j] = sqrt(largest);
    }
  free(data);
  free(cache);
  return rc;
}

int oracle_schwarz(const int* atm, int natm, const int* bas, int nbas, const double* env,
                   int cart, double* q) {
  tables_t t;
  int nenv = 0;
  for (int s = 0; s < nbas; ++s) {
    const int* row = bas + s * ORC_BAS_SLOTS;
    int e = row[6] + row[2] * row[3];
    if (e > nenv) nenv = e;
  }
  for (int a = 0; a < natm; ++a) if (atm[a * ORC_ATM_SLOTS + 1] + 3 > nenv) nenv = atm[a * ORC_ATM_SLOTS + 1] + 3;
  if (validate(&t, atm, natm, bas, nbas, env, nenv, cart)) return -1;
  memset(q, 0, sizeof(double) * nbas * nbas);
  int rc = schwarz(&t, q);
  release(&t);
  return rc;
}

static size_t pair_index(size_t a, size_t b) {
  if (a < b) { size_t t = a; a = b; b = t; }
  return a * (a + 1) / 2 + b;
}

typedef struct {
  const tables_t* t;
  const double* q;
  double qmax, cutoff;
  const double* dm;
  double* out;
  int nb, with_j, with_k, shell_begin, shell_end, rc;
} worker_t;

static void* worker_run(void* arg) {
  worker_t* w = (worker_t*)arg;
  const tables_t* t = w->t;
  const int nbas = t->nbas, n = t->n;
  const size_t nn = (size_t)n * n;
  const int row_begin = t->ao[w->shell_begin], row_end = t->ao[w->shell_end];
  double* data = (double*)malloc(sizeof(double) * max_block(t));
  double* cache = backend_sel == 1 ? (double*)malloc(sizeof(double) * cache_len(t)) : NULL;
#define OWNS_SH(s) (w->shell_begin <= (s) && (s) < w->shell_end)
#define OWNS(r) (row_begin <= (r) && (r) < row_end)
  for (int i = w->shell_begin; i < nbas && !w->rc; ++i)
    for (int j = 0; j <= i && !w->rc; ++j) {
      if (w->cutoff > 0 && w->q[(size_t)i * nbas + j] * w->qmax < w->cutoff) continue;
      const int outer = OWNS_SH(i) || OWNS_SH(j);
      for (int k = outer ? 0 : w->shell_begin; k <= i && !w->rc; ++k) {
        const int owned = outer || OWNS_SH(k);
        const int first = owned ? 0 : w->shell_begin;
        const int last = owned ? k + 1 : (k + 1 < w->shell_end ? k + 1 : w->shell_end);
        for (int l = first; l < last; ++l) {
          if (w->cutoff > 0 && w->q[(size_t)i * nbas + j] * w->q[(size_t)k * nbas + l] < w->cutoff) continue;
          const int di = t->ao[i + 1] - t->ao[i], dj = t->ao[j + 1] - t->ao[j];
          const int dk = t->ao[k + 1] - t->ao[k], dl = t->ao[l + 1] - t->ao[l];
          int sh[4] = {i, j, k, l};
          if ((w->rc = block(t, sh, data, cache, 0))) break;
          for (int p = 0; p < di; ++p)
            for (int qq = 0; qq < dj; ++qq) {
              const int a = t->ao[i] + p, b = t->ao[j] + qq;
              if (a < b) continue;
              const size_t ab = pair_index(a, b);
              for (int r = 0; r < dk; ++r)
                for (int s = 0; s < dl; ++s) {
                  const int c = t->ao[k] + r, e = t->ao[l] + s;
                  if (c < e || ab < pair_index(c, e)) continue;
                  const double v = data[(((size_t)s * dk + r) * dj + qq) * di + p];
                  const size_t ce = pair_index(c, e);
                  const int oa = OWNS(a), ob = OWNS(b), oc = OWNS(c), oe = OWNS(e);
                  for (int bt = 0; bt < w->nb; ++bt) {
                    const double* d = w->dm + (size_t)bt * nn;
                    if (w->with_j) {
                      double* J = w->out + (size_t)bt * nn;
                      if (oa || ob) {
                        const double jab = v * (d[(size_t)c * n + e] + (c != e ? d[(size_t)e * n + c] : 0.));
                        if (oa) J[(size_t)a * n + b] += jab;
                        if (a != b && ob) J[(size_t)b * n + a] += jab;
                      }
                      if (ab != ce && (oc || oe)) {
                        const double jce = v * (d[(size_t)a * n + b] + (a != b ? d[(size_t)b * n + a] : 0.));
                        if (oc) J[(size_t)c * n + e] += jce;
                        if (c != e && oe) J[(size_t)e * n + c] += jce;
                      }
                    }
                    if (w->with_k) {
                      double* K = w->out + (size_t)(w->with_j * w->nb + bt) * nn;
                      if (oa) K[(size_t)a * n + c] += v * d[(size_t)b * n + e];
                      if (c != e && oa) K[(size_t)a * n + e] += v * d[(size_t)b * n + c];
                      if (a != b && ob) K[(size_t)b * n + c] += v * d[(size_t)a * n + e];
                      if (a != b && c != e && ob) K[(size_t)b * n + e] += v * d[(size_t)a * n + c];
                      if (ab != ce) {
                        if (oc) K[(size_t)c * n + a] += v * d[(size_t)e * n + b];
                        if (a != b && oc) K[(size_t)c * n + b] += v * d[(size_t)e * n + a];
                        if (c != e && oe) K[(size_t)e * n + a] += v * d[(size_t)c * n + b];
                        if (a != b && c != e && oe) K[(size_t)e * n + b] += v * d[(size_t)c * n + a];
                      }
                    }
                  }
                }
            }
        }
      }
    }
#undef OWNS
#undef OWNS_SH
  free(data);
  free(cache);
  return NULL;
}

int oracle_direct_jk(const int* atm, int natm, const int* bas, int nbas, const double* env,
                     int nenv, const double* dm, int nb, double* out, int cart, double cutoff,
                     int workers, int with_j, int with_k) {
  if (workers < 1) { oracle_set_error("Direct J/K workers must be positive"); return -1; }
  if ((with_j != 0 && with_j != 1) || (with_k != 0 && with_k != 1) || !(with_j || with_k)) {
    oracle_set_error("Direct J/K requires at least one selected operator");
    return -1;
  }
  if (!isfinite(cutoff) || cutoff < 0) { oracle_set_error("Invalid Schwarz cutoff"); return -1; }
  tables_t t;
  if (validate(&t, atm, natm, bas, nbas, env, nenv, cart)) return -1;
  const int n = t.n;
  const size_t nn = (size_t)n * n;
  for (size_t i = 0; i < (size_t)nb * nn; ++i)
    if (!isfinite(dm[i])) { release(&t); oracle_set_error("Nonfinite density"); return -1; }
  memset(out, 0, sizeof(double) * (size_t)(with_j + with_k) * nb * nn);
  double* q = NULL;
  double qmax = 0;
  if (cutoff > 0) {
    q = (double*)calloc((size_t)nbas * nbas, sizeof(double));
    if (schwarz(&t, q)) { free(q); release(&t); return -1; }
    for (size_t i = 0; i < (size_t)nbas * nbas; ++i) qmax = q[i] > qmax ? q[i] : qmax;
  }
  if (workers > nbas) workers = nbas;
  worker_t* w = (worker_t*)calloc(workers, sizeof(worker_t));
  pthread_t* th = (pthread_t*)calloc(workers, sizeof(pthread_t));
  for (int k = 0; k < workers; ++k) {
    w[k].t = &t; w[k].q = q; w[k].qmax = qmax; w[k].cutoff = cutoff;
    w[k].dm = dm; w[k].out = out; w[k].nb = nb; w[k].with_j = with_j; w[k].with_k = with_k;
    w[k].shell_begin = k * nbas / workers;
    w[k].shell_end = (k + 1) * nbas / workers;
    if (workers == 1) worker_run(&w[k]);
    else pthread_create(&th[k], NULL, worker_run, &w[k]);
  }
  int rc = 0;
  for (int k = 0; k < workers; ++k) {
    if (workers > 1) pthread_join(th[k], NULL);
    rc |= w[k].rc;
  }
  free(w); free(th); free(q);
  release(&t);
  return rc;
}

int oracle_jk_geometry_vjp(const int* atm, int natm, const int* bas, int nbas,
                           const double* env, int nenv, const double* dm, const double* cot_j,
                           const double* cot_k, int cart, double* grad) {
  tables_t t;
  if (validate(&t, atm, natm, bas, nbas, env, nenv, cart)) return -1;
  const int n = t.n;
  memset(grad, 0, sizeof(double) * 3 * natm);
  const size_t mb = max_block(&t);
  double* data = (double*)malloc(sizeof(double) * 3 * mb);
  double* cache = backend_sel == 1 ? (double*)malloc(sizeof(double) * cache_len(&t)) : NULL;
  int rc = 0;
#define W(a, b, c, d) ((cot_j ? cot_j[(size_t)(a) * n + (b)] * dm[(size_t)(c) * n + (d)] : 0.) + \
                       (cot_k ? cot_k[(size_t)(a) * n + (c)] * dm[(size_t)(b) * n + (d)] : 0.))
  for (int i = 0; i < nbas && !rc; ++i)
    for (int j = 0; j < nbas && !rc; ++j)
      for (int k = 0; k < nbas && !rc; ++k)
        for (int l = 0; l < nbas; ++l) {
          int sh[4] = {i, j, k, l};
          if ((rc = block(&t, sh, data, cache, 1))) break;
          const int di = t.ao[i + 1] - t.ao[i], dj = t.ao[j + 1] - t.ao[j];
          const int dk = t.ao[k + 1] - t.ao[k], dl = t.ao[l + 1] - t.ao[l];
          const size_t blk = (size_t)di * dj * dk * dl;
          const int atom = bas[i * ORC_BAS_SLOTS + 0];
          for (int s = 0; s < dl; ++s)
            for (int r = 0; r < dk; ++r)
              for (int qq = 0; qq < dj; ++qq)
                for (int p = 0; p < di; ++p) {
                  const int a = t.ao[i] + p, b = t.ao[j] + qq, c = t.ao[k] + r, d = t.ao[l] + s;
                  const double wsum = W(a, b, c, d) + W(b, a, c, d) + W(c, d, a, b) + W(c, d, b, a);
                  const size_t o = (((size_t)s * dk + r) * dj + qq) * di + p;
                  for (int x = 0; x < 3; ++x) grad[atom * 3 + x] -= data[x * blk + o] * wsum;
                }
        }
#undef W
  free(data);
  free(cache);
  release(&t);
  return rc;
}
