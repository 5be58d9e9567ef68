/* Oracle (test infrastructure only): contracted Rys-quadrature ERI blocks,
 * their first-centre derivative, and one-electron S/T/V.
 *
 * Restates the arithmetic of libcint's two-electron driver
 *   CINTinit_int2e_EnvVars  src/gradscf/integrals/_native/vendor/libcint/src/g2e.c:21-172
 *   CINTg0_2e (roots, B00/B10/B01/C00/C0p)  g2e.c:4425-4569
 *   2D recurrence / horizontal transfer      g2e.c:272-690
 *   gout assembly + primitive contraction    cint2e.c:660-758, 961
 *   cart -> real-spherical (d shells)        cart2sph.c:5324 (c2s_sph_2e1)
 *   s/p common factors                       g1e.c:566-573
 *   first-centre derivative (int2e_ip1)      autocode/grad2.c
 * written from the Rys-quadrature mathematics rather than from that code:
 *   [ab|cd] = 2 pi^{5/2} / (p q sqrt(p+q)) K_AB K_CD sum_r w_r Ix Iy Iz
 * with the vertical recurrences of the Obara-Saika relation in u = t^2 and the
 * horizontal transfer written as an explicit binomial sum (x-B)^b =
 * sum_k C(b,k) (x-A)^k (A-B)^{b-k}.  Block layout follows libcint:
 * out[((l*dk + k)*dj + j)*di + i], each shell index = ctr*ncomp + comp, Cartesian
 * components in descending-x then descending-y order (cint_bas.c:166).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

int oracle_rys_roots_ld(int n, long double x, long double* s, long double* w);
void oracle_boys_ld(int mmax, long double x, long double* f);
void oracle_set_error(const char* msg);

#define LR (ORC_MAX_L + 1)          /* highest raised angular momentum */
#define GD (2 * LR + 1)             /* 2D-table extent per index */
#define PI 3.14159265358979323846

typedef struct {
  int l, np, nc;
  const double* a;
  const double* c;
  double r[3];
} shl_t;

static void load_shell(shl_t* s, int sh, const int* atm, const int* bas, const double* env) {
  const int* row = bas + sh * ORC_BAS_SLOTS;
  s->l = row[1];
  s->np = row[2];
  s->nc = row[3];
  s->a = env + row[5];
  s->c = env + row[6];
  const double* r = env + atm[row[0] * ORC_ATM_SLOTS + 1];
  s->r[0] = r[0]; s->r[1] = r[1]; s->r[2] = r[2];
}

static double fac_sp(int l) {
  if (l == 0) return 0.5 / sqrt(PI);
  if (l == 1) return sqrt(3.0 / (4.0 * PI));
  return 1.0;
}

static int cart_list(int l, int (*xyz)[3]) {
  int n = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      xyz[n][0] = lx; xyz[n][1] = ly; xyz[n][2] = l - lx - ly;
      ++n;
    }
  return n;
}

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static double binom(int n, int k) {
  double r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

/* H[a][b][c][d] (a<=la, b<=lb, c<=lc, d<=ld) from G[n][m] with the binomial transfer. */
static void transfer(int la, int lb, int lc, int ld, double AB, double CD,
                     double G[GD][GD], double* H) {
  double pab[LR + 1], pcd[LR + 1];
  pab[0] = pcd[0] = 1;
  for (int i = 1; i <= LR; ++i) { pab[i] = pab[i - 1] * AB; pcd[i] = pcd[i - 1] * CD; }
  for (int a = 0; a <= la; ++a)
    for (int b = 0; b <= lb; ++b)
      for (int c = 0; c <= lc; ++c)
        for (int d = 0; d <= ld; ++d) {
          double s = 0;
          for (int k = 0; k <= b; ++k)
            for (int k2 = 0; k2 <= d; ++k2)
              s += binom(b, k) * pab[b - k] * binom(d, k2) * pcd[d - k2] * G[a + k][c + k2];
          H[((a * (lb + 1) + b) * (lc + 1) + c) * (ld + 1) + d] = s;
        }
}

/* Primitive Cartesian block(s).  deriv=1 adds the electron-coordinate gradient on
 * the first function: out[comp][...] with (a_x I(a-1) - 2 alpha I(a+1)). */
static void prim_quartet(const shl_t* S[4], const int ip[4], int deriv, double* prim) {
  const double ai = S[0]->a[ip[0]], aj = S[1]->a[ip[1]], ak = S[2]->a[ip[2]], al = S[3]->a[ip[3]];
  const int la = S[0]->l, lb = S[1]->l, lc = S[2]->l, ld = S[3]->l;
  const int lar = la + deriv;
  const double p = ai + aj, q = ak + al;
  double P[3], Q[3], AB[3], CD[3], rab = 0, rcd = 0, rpq = 0;
  for (int x = 0; x < 3; ++x) {
    P[x] = (ai * S[0]->r[x] + aj * S[1]->r[x]) / p;
    Q[x] = (ak * S[2]->r[x] + al * S[3]->r[x]) / q;
    AB[x] = S[0]->r[x] - S[1]->r[x];
    CD[x] = S[2]->r[x] - S[3]->r[x];
    rab += AB[x] * AB[x];
    rcd += CD[x] * CD[x];
    rpq += (P[x] - Q[x]) * (P[x] - Q[x]);
  }
  const double rho = p * q / (p + q);
  const double T = rho * rpq;
  const double pref = 2.0 * pow(PI, 2.5) / (p * q * sqrt(p + q)) *
                      exp(-ai * aj / p * rab - ak * al / q * rcd);
  const int nroots = (lar + lb + lc + ld) / 2 + 1;
  long double t2l[ORC_MAX_ROOTS], wl[ORC_MAX_ROOTS];
  oracle_rys_roots_ld(nroots, (long double)T, t2l, wl);
  const int nmax = lar + lb, mmax = lc + ld;
  int ci[64][3], cj[64][3], ck[64][3], cl[64][3];
  const int ni = cart_list(la, ci), nj = cart_list(lb, cj), nk = cart_list(lc, ck), nl = cart_list(ld, cl);
  const int hs = (lar + 1) * (lb + 1) * (lc + 1) * (ld + 1);
  double* H = (double*)malloc(sizeof(double) * 3 * hs);
  const int nblk = ni * nj * nk * nl;
  for (int r = 0; r < nroots; ++r) {
    const double u = (double)t2l[r];
    const double B00 = 0.5 * u / (p + q);
    const double B10 = 0.5 / p * (1.0 - rho / p * u);
    const double B01 = 0.5 / q * (1.0 - rho / q * u);
    for (int x = 0; x < 3; ++x) {
      double G[GD][GD];
      const double c00 = (P[x] - S[0]->r[x]) - rho / p * (P[x] - Q[x]) * u;
      const double c0p = (Q[x] - S[2]->r[x]) + rho / q * (P[x] - Q[x]) * u;
      G[0][0] = x == 2 ? (double)wl[r] * pref : 1.0;
      for (int n = 0; n < nmax; ++n)
        G[n + 1][0] = c00 * G[n][0] + (n ? n * B10 * G[n - 1][0] : 0.0);
      for (int m = 0; m < mmax; ++m)
        for (int n = 0; n <= nmax; ++n)
          G[n][m + 1] = c0p * G[n][m] + (m ? m * B01 * G[n][m - 1] : 0.0) +
                        (n ? n * B00 * G[n - 1][m] : 0.0);
      transfer(lar, lb, lc, ld, AB[x], CD[x], G, H + x * hs);
    }
#define HI(x, a, b, c, d) H[(x) * hs + (((a) * (lb + 1) + (b)) * (lc + 1) + (c)) * (ld + 1) + (d)]
    for (int l_ = 0; l_ < nl; ++l_)
      for (int k_ = 0; k_ < nk; ++k_)
        for (int j_ = 0; j_ < nj; ++j_)
          for (int i_ = 0; i_ < ni; ++i_) {
            const int idx = ((l_ * nk + k_) * nj + j_) * ni + i_;
            double f[3];
            for (int x = 0; x < 3; ++x)
              f[x] = HI(x, ci[i_][x], cj[j_][x], ck[k_][x], cl[l_][x]);
            if (!deriv) {
              prim[idx] += f[0] * f[1] * f[2];
            } else {
              for (int x = 0; x < 3; ++x) {
                const int a = ci[i_][x];
                double dv = -2.0 * ai * HI(x, a + 1, cj[j_][x], ck[k_][x], cl[l_][x]);
                if (a) dv += a * HI(x, a - 1, cj[j_][x], ck[k_][x], cl[l_][x]);
                double v = dv;
                for (int y = 0; y < 3; ++y)
                  if (y != x) v *= f[y];
                prim[x * nblk + idx] += v;
              }
            }
          }
#undef HI
  }
  free(H);
}

/* d-shell real-spherical transform coefficients (rows m=-2..2: xy, yz, z2, xz, x2-y2;
 * columns xx, xy, xz, yy, yz, zz), from the real spherical harmonics Y_2m r^2. */
static void c2s_d(double m[5][6]) {
  memset(m, 0, sizeof(double) * 30);
  const double a = sqrt(15.0 / (4.0 * PI));
  const double b = sqrt(5.0 / (16.0 * PI));
  const double c = sqrt(15.0 / (16.0 * PI));
  m[0][1] = a;
  m[1][4] = a;
  m[2][0] = -b; m[2][3] = -b; m[2][5] = 2.0 * b;
  m[3][2] = a;
  m[4][0] = c; m[4][3] = -c;
}

/* Transform index `axis` (0..3) of a block from Cartesian to spherical for l=2. */
static void sph_axis(double* in, int dims[4], int axis, int nc, double* out) {
  double m[5][6];
  c2s_d(m);
  int od[4] = {dims[0], dims[1], dims[2], dims[3]};
  od[axis] = nc * 5;
  int st[4] = {1, dims[0], dims[0] * dims[1], dims[0] * dims[1] * dims[2]};
  int ost[4] = {1, od[0], od[0] * od[1], od[0] * od[1] * od[2]};
  const int total = od[0] * od[1] * od[2] * od[3];
  for (int o = 0; o < total; ++o) {
    int idx[4], rem = o;
    idx[3] = rem / ost[3]; rem %= ost[3];
    idx[2] = rem / ost[2]; rem %= ost[2];
    idx[1] = rem / ost[1]; idx[0] = rem % ost[1];
    const int ctr = idx[axis] / 5, ms = idx[axis] % 5;
    double s = 0;
    for (int cc = 0; cc < 6; ++cc) {
      if (m[ms][cc] == 0) continue;
      int base = 0;
      for (int t = 0; t < 4; ++t) base += (t == axis ? ctr * 6 + cc : idx[t]) * st[t];
      s += m[ms][cc] * in[base];
    }
    out[o] = s;
  }
  dims[axis] = od[axis];
}

int oracle_eri_block(const int* shls, const int* atm, int natm, const int* bas, int nbas,
                     const double* env, int cart, int deriv, double* out) {
  (void)natm; (void)nbas;
  shl_t sh[4];
  const shl_t* S[4];
  for (int t = 0; t < 4; ++t) {
    load_shell(&sh[t], shls[t], atm, bas, env);
    S[t] = &sh[t];
    if (sh[t].l > ORC_MAX_L - (t == 0 ? deriv : 0)) {
      oracle_set_error("oracle: angular momentum too high");
      return -1;
    }
    if (!cart && sh[t].l > 2) {
      oracle_set_error("oracle: spherical transform implemented for l<=2");
      return -1;
    }
  }
  const int ni = ncart(sh[0].l), nj = ncart(sh[1].l), nk = ncart(sh[2].l), nl = ncart(sh[3].l);
  const int nblk = ni * nj * nk * nl;
  const int ncomp = deriv ? 3 : 1;
  const int di = ni * sh[0].nc, dj = nj * sh[1].nc, dk = nk * sh[2].nc, dl = nl * sh[3].nc;
  const size_t cblk = (size_t)di * dj * dk * dl;
  double* cartb = (double*)calloc(cblk * ncomp, sizeof(double));
  double* prim = (double*)malloc(sizeof(double) * nblk * ncomp);
  const double fsp = fac_sp(sh[0].l) * fac_sp(sh[1].l) * fac_sp(sh[2].l) * fac_sp(sh[3].l);
  int ip[4];
  for (ip[3] = 0; ip[3] < sh[3].np; ++ip[3])
    for (ip[2] = 0; ip[2] < sh[2].np; ++ip[2])
      for (ip[1] = 0; ip[1] < sh[1].np; ++ip[1])
        for (ip[0] = 0; ip[0] < sh[0].np; ++ip[0]) {
          memset(prim, 0, sizeof(double) * nblk * ncomp);
          prim_quartet(S, ip, deriv, prim);
          for (int c3 = 0; c3 < sh[3].nc; ++c3)
            for (int c2 = 0; c2 < sh[2].nc; ++c2)
              for (int c1 = 0; c1 < sh[1].nc; ++c1)
                for (int c0 = 0; c0 < sh[0].nc; ++c0) {
                  const double cf = fsp * sh[0].c[c0 * sh[0].np + ip[0]] * sh[1].c[c1 * sh[1].np + ip[1]] *
                                    sh[2].c[c2 * sh[2].np + ip[2]] * sh[3].c[c3 * sh[3].np + ip[3]];
                  for (int comp = 0; comp < ncomp; ++comp)
                    for (int l_ = 0; l_ < nl; ++l_)
                      for (int k_ = 0; k_ < nk; ++k_)
                        for (int j_ = 0; j_ < nj; ++j_)
                          for (int i_ = 0; i_ < ni; ++i_) {
                            const size_t o = (((size_t)(c3 * nl + l_) * dk + (c2 * nk + k_)) * dj +
                                              (c1 * nj + j_)) * di + (c0 * ni + i_);
                            cartb[comp * cblk + o] +=
                                cf * prim[comp * nblk + ((l_ * nk + k_) * nj + j_) * ni + i_];
                          }
                }
        }
  free(prim);
  if (cart) {
    memcpy(out, cartb, sizeof(double) * cblk * ncomp);
  } else {
    int odims[4] = {(sh[0].l == 2 ? 5 : ni) * sh[0].nc, (sh[1].l == 2 ? 5 : nj) * sh[1].nc,
                    (sh[2].l == 2 ? 5 : nk) * sh[2].nc, (sh[3].l == 2 ? 5 : nl) * sh[3].nc};
    const size_t oblk = (size_t)odims[0] * odims[1] * odims[2] * odims[3];
    double* a = (double*)malloc(sizeof(double) * cblk);
    double* b = (double*)malloc(sizeof(double) * cblk);
    for (int comp = 0; comp < ncomp; ++comp) {
      int dims[4] = {di, dj, dk, dl};
      memcpy(a, cartb + comp * cblk, sizeof(double) * cblk);
      for (int t = 0; t < 4; ++t)
        if (sh[t].l == 2) {
          sph_axis(a, dims, t, sh[t].nc, b);
          double* tmp = a; a = b; b = tmp;
        }
      memcpy(out + comp * oblk, a, sizeof(double) * oblk);
    }
    free(a);
    free(b);
  }
  free(cartb);
  return 0;
}

/* ---------------- one-electron integrals ---------------- */

/* 1D overlap table S[a][b], a <= amax, b <= bmax, Obara-Saika. */
static void overlap_1d(double ai, double aj, double A, double B, int amax, int bmax,
                       double S[LR + 3][LR + 3]) {
  const double p = ai + aj, P = (ai * A + aj * B) / p;
  S[0][0] = sqrt(PI / p) * exp(-ai * aj / p * (A - B) * (A - B));
  for (int a = 0; a <= amax; ++a)
    for (int b = 0; b <= bmax; ++b) {
      if (a == 0 && b == 0) continue;
      if (a > 0) {
        double v = (P - A) * S[a - 1][b];
        if (a > 1) v += (a - 1) * S[a - 2][b] / (2 * p);
        if (b > 0) v += b * S[a - 1][b - 1] / (2 * p);
        S[a][b] = v;
      } else {
        double v = (P - B) * S[a][b - 1];
        if (b > 1) v += (b - 1) * S[a][b - 2] / (2 * p);
        S[a][b] = v;
      }
    }
}

int oracle_nao(const int* bas, int nbas, int cart) {
  int n = 0;
  for (int s = 0; s < nbas; ++s) {
    const int l = bas[s * ORC_BAS_SLOTS + 1], nc = bas[s * ORC_BAS_SLOTS + 3];
    n += nc * (cart ? ncart(l) : 2 * l + 1);
  }
  return n;
}

/* Cartesian 1e block (ncart_i*nc_i x ncart_j*nc_j, i fastest) for S, T, V. */
static void int1e_block(const shl_t* A, const shl_t* B, const int* atm, int natm, const double* env,
                        double* s, double* t, double* v) {
  int ci[64][3], cj[64][3];
  const int ni = cart_list(A->l, ci), nj = cart_list(B->l, cj);
  const int di = ni * A->nc, dj = nj * B->nc;
  memset(s, 0, sizeof(double) * di * dj);
  memset(t, 0, sizeof(double) * di * dj);
  memset(v, 0, sizeof(double) * di * dj);
  const double fsp = fac_sp(A->l) * fac_sp(B->l);
  for (int ip = 0; ip < A->np; ++ip)
    for (int jp = 0; jp < B->np; ++jp) {
      const double ai = A->a[ip], aj = B->a[jp], p = ai + aj;
      double S1[3][LR + 3][LR + 3];
      for (int x = 0; x < 3; ++x) overlap_1d(ai, aj, A->r[x], B->r[x], A->l, B->l + 2, S1[x]);
      /* nuclear attraction via Rys with a point charge */
      double P[3], rab = 0;
      for (int x = 0; x < 3; ++x) {
        P[x] = (ai * A->r[x] + aj * B->r[x]) / p;
        rab += (A->r[x] - B->r[x]) * (A->r[x] - B->r[x]);
      }
      const double kab = exp(-ai * aj / p * rab);
      double pv[64 * 64];
      memset(pv, 0, sizeof(pv));
      const int nroots = (A->l + B->l) / 2 + 1;
      for (int c = 0; c < natm; ++c) {
        const int Z = atm[c * ORC_ATM_SLOTS + 0];
        if (!Z) continue;
        const double* C = env + atm[c * ORC_ATM_SLOTS + 1];
        double pc2 = 0;
        for (int x = 0; x < 3; ++x) pc2 += (P[x] - C[x]) * (P[x] - C[x]);
        long double t2[ORC_MAX_ROOTS], w[ORC_MAX_ROOTS];
        oracle_rys_roots_ld(nroots, (long double)(p * pc2), t2, w);
        for (int r = 0; r < nroots; ++r) {
          const double u = (double)t2[r];
          double H[3][LR + 1][LR + 1];
          for (int x = 0; x < 3; ++x) {
            double G[GD];
            const double c00 = (P[x] - A->r[x]) - (P[x] - C[x]) * u;
            const double b10 = 0.5 / p * (1.0 - u);
            G[0] = x == 2 ? -Z * 2.0 * PI / p * kab * (double)w[r] : 1.0;
            for (int n = 0; n < A->l + B->l; ++n)
              G[n + 1] = c00 * G[n] + (n ? n * b10 * G[n - 1] : 0.0);
            const double ab = A->r[x] - B->r[x];
            for (int a = 0; a <= A->l; ++a)
              for (int b = 0; b <= B->l; ++b) {
                double sum = 0;
                for (int k = 0; k <= b; ++k) sum += binom(b, k) * pow(ab, b - k) * G[a + k];
                H[x][a][b] = sum;
              }
          }
          for (int j_ = 0; j_ < nj; ++j_)
            for (int i_ = 0; i_ < ni; ++i_)
              pv[j_ * ni + i_] += H[0][ci[i_][0]][cj[j_][0]] * H[1][ci[i_][1]][cj[j_][1]] *
                                  H[2][ci[i_][2]][cj[j_][2]];
        }
      }
      for (int c1 = 0; c1 < B->nc; ++c1)
        for (int c0 = 0; c0 < A->nc; ++c0) {
          const double cf = fsp * A->c[c0 * A->np + ip] * B->c[c1 * B->np + jp];
          for (int j_ = 0; j_ < nj; ++j_)
            for (int i_ = 0; i_ < ni; ++i_) {
              double sv = 1, tv = 0;
              for (int x = 0; x < 3; ++x) sv *= S1[x][ci[i_][x]][cj[j_][x]];
              for (int x = 0; x < 3; ++x) {
                const int a = ci[i_][x], b = cj[j_][x];
                double tx = -2.0 * aj * aj * S1[x][a][b + 2] + aj * (2 * b + 1) * S1[x][a][b];
                if (b > 1) tx -= 0.5 * b * (b - 1) * S1[x][a][b - 2];
                double prod = tx;
                for (int y = 0; y < 3; ++y)
                  if (y != x) prod *= S1[y][ci[i_][y]][cj[j_][y]];
                tv += prod;
              }
              const size_t o = (size_t)(c1 * nj + j_) * di + (c0 * ni + i_);
              s[o] += cf * sv;
              t[o] += cf * tv;
              v[o] += cf * pv[j_ * ni + i_];
            }
        }
    }
}

static void sph_1e(const double* in, int li, int nci, int lj, int ncj, double* out) {
  /* out dims: (nsph_i*nci) x (nsph_j*ncj), i fastest; l<=1 identity, l=2 c2s */
  double m[5][6];
  c2s_d(m);
  const int ni = ncart(li), nj = ncart(lj);
  const int si = li == 2 ? 5 : ni, sj = lj == 2 ? 5 : nj;
  const int di = ni * nci;
  for (int c1 = 0; c1 < ncj; ++c1)
    for (int b = 0; b < sj; ++b)
      for (int c0 = 0; c0 < nci; ++c0)
        for (int a = 0; a < si; ++a) {
          double sum = 0;
          for (int cb = 0; cb < nj; ++cb) {
            const double fb = lj == 2 ? m[b][cb] : (cb == b);
            if (fb == 0) continue;
            for (int ca = 0; ca < ni; ++ca) {
              const double fa = li == 2 ? m[a][ca] : (ca == a);
              if (fa == 0) continue;
              sum += fa * fb * in[(size_t)(c1 * nj + cb) * di + (c0 * ni + ca)];
            }
          }
          out[(size_t)(c1 * sj + b) * (si * nci) + (c0 * si + a)] = sum;
        }
}

int oracle_int1e(const int* atm, int natm, const int* bas, int nbas, const double* env,
                 int cart, double* s, double* t, double* v) {
  const int n = oracle_nao(bas, nbas, cart);
  int* ao = (int*)malloc(sizeof(int) * (nbas + 1));
  ao[0] = 0;
  for (int i = 0; i < nbas; ++i) {
    const int l = bas[i * ORC_BAS_SLOTS + 1], nc = bas[i * ORC_BAS_SLOTS + 3];
    if (l > ORC_MAX_L || (!cart && l > 2)) {
      free(ao);
      oracle_set_error("oracle: unsupported angular momentum");
      return -1;
    }
    ao[i + 1] = ao[i] + nc * (cart ? ncart(l) : 2 * l + 1);
  }
  double *bs = malloc(sizeof(double) * 36 * 36 * 4), *bt = malloc(sizeof(double) * 36 * 36 * 4),
         *bv = malloc(sizeof(double) * 36 * 36 * 4), *tmp = malloc(sizeof(double) * 36 * 36 * 4);
  for (int i = 0; i < nbas; ++i)
    for (int j = 0; j < nbas; ++j) {
      shl_t A, B;
      load_shell(&A, i, atm, bas, env);
      load_shell(&B, j, atm, bas, env);
      int1e_block(&A, &B, atm, natm, env, bs, bt, bv);
      const int di = ao[i + 1] - ao[i], dj = ao[j + 1] - ao[j];
      double* srcs[3] = {bs, bt, bv};
      double* dsts[3] = {s, t, v};
      for (int w = 0; w < 3; ++w) {
        const double* blk = srcs[w];
        if (!cart) {
          sph_1e(srcs[w], A.l, A.nc, B.l, B.nc, tmp);
          blk = tmp;
        }
        for (int q = 0; q < dj; ++q)
          for (int p = 0; p < di; ++p) dsts[w][(size_t)(ao[i] + p) * n + ao[j] + q] = blk[q * di + p];
      }
    }
  free(bs); free(bt); free(bv); free(tmp);
  free(ao);
  return 0;
}
