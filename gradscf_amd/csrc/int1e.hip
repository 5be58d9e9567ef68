// One-electron overlap, kinetic and nuclear-attraction integrals (FP64) for the
// SCF core Hamiltonian: the int1e_ovlp / int1e_kin / int1e_nuc evaluation the
// reference performs on the CPU before its SCF loop (src/gradscf/integrals/
// _native/csrc/ffi.cc:11-86 via libcint cint1e.c and autocode/intor1.c).
//
// One lane per segmented shell pair (all classes with l <= 2).  Overlap and
// kinetic from 1D Obara-Saika tables; nuclear attraction by Rys quadrature with
// a point charge, (a|1/r_C|b) = 2 pi/p K_AB sum_r w_r Ix Iy Iz, with
// C00 = (P - A) - (P - C) u and B10 = (1 - u)/(2p).
#include "gscf_device.h"

namespace gscf {

struct Int1eArgs {
  const Pair* pairs;
  const PrimPair* prims;
  const ShellXYZ* shells;
  const double* nuclei;   // [nnuc][5]: x, y, z, Z, molecule
  int nnuc;
  int npairs;
  const MolInfo* mols;    // Cartesian matrix of each molecule
  double* s;
  double* t;
  double* v;
};

template <int LA, int LB>
__global__ __launch_bounds__(256) void int1e_kernel(Int1eArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB);
  constexpr int NR = (LA + LB) / 2 + 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.npairs) return;
  const Pair pr = a.pairs[i];
  const ShellXYZ A = a.shells[pr.sa], B = a.shells[pr.sb];
  const double Ar[3] = {A.x, A.y, A.z}, Br[3] = {B.x, B.y, B.z};
  double accS[NA][NB], accT[NA][NB], accV[NA][NB];
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) accS[ia][ib] = accT[ia][ib] = accV[ia][ib] = 0.0;
  const double pi = 3.14159265358979323846;
  const double s2p = sqrt(2.0 * pi * pi * sqrt(pi));
  for (int ip = 0; ip < pr.np; ++ip) {
    const PrimPair P = a.prims[pr.poff + ip];
    const double p = P.p, ip2 = 0.5 / P.p;
    const double Pr[3] = {P.px, P.py, P.pz};
    const double ckab = P.e * p / s2p;  // c_a c_b K_AB
    const double aj = P.ab;
    // 1D overlap tables S[x][a][b], a <= LA, b <= LB + 2 (unit base; (pi/p)^{3/2} applied below)
    double S1[3][LA + 1][LB + 3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      const double pa = Pr[x] - Ar[x], pb = Pr[x] - Br[x];
#pragma unroll
      for (int ia = 0; ia <= LA; ++ia)
#pragma unroll
        for (int ib = 0; ib <= LB + 2; ++ib) {
          double val;
          if (ia == 0 && ib == 0) {
            val = 1.0;
          } else if (ia > 0) {
            val = pa * S1[x][ia - 1][ib];
            if (ia > 1) val += (ia - 1) * ip2 * S1[x][ia - 2][ib];
            if (ib > 0) val += ib * ip2 * S1[x][ia - 1][ib - 1];
          } else {
            val = pb * S1[x][0][ib - 1];
            if (ib > 1) val += (ib - 1) * ip2 * S1[x][0][ib - 2];
          }
          S1[x][ia][ib] = val;
        }
    }
    const double sfac = ckab * (pi / p) * sqrt(pi / p);
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
        int ea[3] = {cart_comp(LA, ia, 0), cart_comp(LA, ia, 1), cart_comp(LA, ia, 2)};
        int eb[3] = {cart_comp(LB, ib, 0), cart_comp(LB, ib, 1), cart_comp(LB, ib, 2)};
        const double sx = S1[0][ea[0]][eb[0]], sy = S1[1][ea[1]][eb[1]], sz = S1[2][ea[2]][eb[2]];
        double tk = 0.0;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          const int bx = eb[x];
          double tx = -2.0 * aj * aj * S1[x][ea[x]][bx + 2] + aj * (2 * bx + 1) * S1[x][ea[x]][bx];
          if (bx > 1) tx -= 0.5 * bx * (bx - 1) * S1[x][ea[x]][bx - 2];
          tk += tx * (x == 0 ? sy * sz : (x == 1 ? sx * sz : sx * sy));
        }
        accS[ia][ib] += sfac * sx * sy * sz;
        accT[ia][ib] += sfac * tk;
      }
    // nuclear attraction
    for (int c = 0; c < a.nnuc; ++c) {
      const double* C = a.nuclei + 5 * c;
      if ((int)C[4] != A.mol) continue;
      const double pc[3] = {Pr[0] - C[0], Pr[1] - C[1], Pr[2] - C[2]};
      const double T = p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
      double u[NR], w[NR];
      rys_roots<NR>(T, u, w);
      const double vf = -C[3] * 2.0 * pi / p * ckab;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const double b10 = ip2 * (1.0 - u[r]);
        double h[3][LA + 1][LB + 1];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          double g[LA + LB + 1];
          const double c00 = (Pr[x] - Ar[x]) - pc[x] * u[r];
          g[0] = x == 2 ? vf * w[r] : 1.0;
#pragma unroll
          for (int m = 0; m < LA + LB; ++m) g[m + 1] = c00 * g[m] + (m ? m * b10 * g[m - 1] : 0.0);
          hrr_bra1<LA, LB>(g, Ar[x] - Br[x], h[x]);
        }
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
          for (int ib = 0; ib < NB; ++ib)
            accV[ia][ib] += h[0][cart_comp(LA, ia, 0)][cart_comp(LB, ib, 0)] *
                            h[1][cart_comp(LA, ia, 1)][cart_comp(LB, ib, 1)] *
                            h[2][cart_comp(LA, ia, 2)][cart_comp(LB, ib, 2)];
      }
    }
  }
  const MolInfo mi = a.mols[A.mol];
  const int nc = mi.n;
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
      const size_t x = mi.base + (size_t)(A.ao + ia) * nc + B.ao + ib;
      const size_t y = mi.base + (size_t)(B.ao + ib) * nc + A.ao + ia;
      a.s[x] = accS[ia][ib]; a.s[y] = accS[ia][ib];
      a.t[x] = accT[ia][ib]; a.t[y] = accT[ia][ib];
      a.v[x] = accV[ia][ib]; a.v[y] = accV[ia][ib];
    }
}

void launch_int1e_class(int la, int lb, const Pair* pairs, int npairs, const PrimPair* prims,
                        const ShellXYZ* shells, const double* nuclei, int nnuc, const MolInfo* mols, double* s,
                        double* t, double* v, hipStream_t st) {
  if (npairs <= 0) return;
  Int1eArgs a{pairs, prims, shells, nuclei, nnuc, npairs, mols, s, t, v};
  const dim3 grid((npairs + 255) / 256), block(256);
#define L1E(A, B)                                                          \
  if (la == A && lb == B) {                                                \
    hipLaunchKernelGGL((int1e_kernel<A, B>), grid, block, 0, st, a);       \
    return;                                                                \
  }
  L1E(0, 0) L1E(1, 0) L1E(1, 1) L1E(2, 0) L1E(2, 1) L1E(2, 2)
#undef L1E
}

}  // namespace gscf
