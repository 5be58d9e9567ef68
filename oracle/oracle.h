/* Oracle for the direct-SCF J/K path: a plain-C CPU restatement of the
 * reference algorithm, used ONLY as the checker by tests, smoke() and the
 * CPU-baseline leg of bench.py.  Nothing in the product links or calls it.
 *
 * Pinning of each part (what makes it trustworthy, beyond restating formulas):
 *   Boys F_m          closed forms F_m(0)=1/(2m+1), F_0(x)=sqrt(pi/x)erf(sqrt x)/2,
 *                     downward recursion identity, libcint gamma_inc_like (_ref)
 *   Rys roots/weights moment identities sum_i w_i t_i^{2k} = F_k(x), k<2n,
 *                     libcint CINTrys_roots (_ref)
 *   ERI blocks        (ss|ss) closed form, 8-fold permutation symmetry,
 *                     libcint int2e_cart / int2e_sph blocks (_ref)
 *   Schwarz           Q_ij^2 = max diagonal of the (ij|ij) block rebuilt from own and
 *                     libcint (_ref) blocks; |(ij|kl)| <= Q_ij Q_kl on random quartets
 *   J/K               full-tensor einsum brute force (unscreened); screened builds equal
 *                     the full tensor masked by Q_ij Q_kl >= cutoff (1e-12 and 1e-7) and
 *                     the unmasked one to 1e-10 at 1e-12; the same driver on libcint blocks
 *   one-electron      libcint int1e_ovlp/kin/nuc (_ref), published RHF energies
 *   ERI derivatives   libcint int2e_ip1 (_ref), translation sum rule, finite differences
 */
#ifndef GRADSCF_ORACLE_H
#define GRADSCF_ORACLE_H
#include <stddef.h>

#define ORC_ATM_SLOTS 6
#define ORC_BAS_SLOTS 8
#define ORC_MAX_L 4
#define ORC_MAX_ROOTS 9

/* Boys function F_m(x), m = 0..mmax (long double internally). */
void oracle_boys(int mmax, double x, double* f);
/* n-point Rys quadrature for weight exp(-x t^2) on t in [0,1]:
 * roots returned as t^2 (ascending), weights sum to F_0(x). */
int oracle_rys_roots(int n, double x, double* t2, double* w);

/* Contracted ERI block for shells (i,j,k,l) in libcint layout
 * out[((l*dk + k)*dj + j)*di + i] (first index fastest), d* = nctr*ncomp.
 * deriv: 0 plain; 1 = d/dA_{x,y,z} of the first centre (three blocks, like
 * libcint int2e_ip1: component-major, out[comp*block + ...]). Returns 0 ok. */
int oracle_eri_block(const int* shls, const int* atm, int natm, const int* bas, int nbas,
                     const double* env, int cart, int deriv, double* out);

/* Select the ERI backend used by oracle_direct_jk: 0 own Rys, 1 libcint _ref. */
int oracle_use_libcint(const char* path);
void oracle_select_backend(int backend);

/* Schwarz bounds Q[i*nbas+j] (i>=j) = sqrt(max_{p in i, q in j} |(pq|pq)|). */
int oracle_schwarz(const int* atm, int natm, const int* bas, int nbas, const double* env,
                   int cart, double* q);

/* Direct J/K with the reference's driver semantics:
 * out[(op*nb + batch)*n*n + ...], op 0 = J (if with_j), then K. */
int oracle_direct_jk(const int* atm, int natm, const int* bas, int nbas, const double* env,
                     int nenv, const double* dm, int nb, double* out, int cart, double cutoff,
                     int workers, int with_j, int with_k);

/* Nuclear ERI-gradient contraction (no screening, all s1 quartets):
 * grad[a*3+x] = sum_abcd d(ab|cd)/dR_{atom(a),x} (cj_ab D_cd + ck_ac D_bd),
 * atoms are the atm rows that shells sit on. */
int oracle_jk_geometry_vjp(const int* atm, int natm, const int* bas, int nbas,
                           const double* env, int nenv, const double* dm, const double* cot_j,
                           const double* cot_k, int cart, double* grad);

/* One-electron S, T, V (full n x n, libcint AO order). */
int oracle_int1e(const int* atm, int natm, const int* bas, int nbas, const double* env,
                 int cart, double* s, double* t, double* v);

const char* oracle_last_error(void);
int oracle_nao(const int* bas, int nbas, int cart);
#endif
