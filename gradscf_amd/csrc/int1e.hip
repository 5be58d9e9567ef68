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
  const double* nuclei;   // [nnuc][6]: x, y, z, Z, molecule, atm row
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
      const double* C = a.nuclei + 6 * c;
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

// ---------------------------------------------------------------------------
// One-electron part of the RHF nuclear gradient, contracted on the fly:
//   g[R] += sum_{mu nu} D_{mu nu} d(T + V)_{mu nu}/dR - W_{mu nu} dS_{mu nu}/dR
// (the reference's int1e_ipovlp / int1e_ipkin / int1e_ipnuc + int1e_iprinv
// contractions of its SCF gradient, src/gradscf/integrals/_native/csrc/geometry.cc,
// via libcint autocode/grad1.c).  Centre derivatives of a Cartesian Gaussian,
//   d/dA_x g_a = 2 alpha_a g_{a+1_x} - a_x g_{a-1_x},
// give the A derivatives of S, T and V; two-centre S and T are translation
// invariant (d/dB = -d/dA); for V the B derivative uses the raised b and the
// derivative with respect to the nucleus C follows from
// d/dA + d/dB + d/dC = 0 for each nucleus's term (the Hellmann-Feynman force).
// One lane per segmented shell pair (S >= T); pair factor 2 for S != T with
// symmetrised D and W.  Per-atom sums in LDS, one flush per workgroup.
// ---------------------------------------------------------------------------
struct Grad1eArgs {
  const Pair* pairs;
  const PrimPair* prims;
  const ShellXYZ* shells;
  const int* seg_atom;    // atm row of each segmented shell's centre
  const double* nuclei;   // [nnuc][6]: x, y, z, Z, molecule, atm row
  int nnuc;
  int npairs;
  const MolInfo* mols;
  const double* dens;     // packed Cartesian D
  const double* wdens;    // packed Cartesian energy-weighted density W
  int natm;
  double* grad;           // [natm][3]
};

template <int LA, int LB>
__global__ __launch_bounds__(256) void grad1e_kernel(Grad1eArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB);
  constexpr int NR = (LA + LB + 1) / 2 + 1;
  extern __shared__ double gl[];
  for (int i = threadIdx.x; i < 3 * a.natm; i += blockDim.x) gl[i] = 0.0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.npairs) {
    const Pair pr = a.pairs[i];
    const ShellXYZ A = a.shells[pr.sa], B = a.shells[pr.sb];
    const double Ar[3] = {A.x, A.y, A.z}, Br[3] = {B.x, B.y, B.z};
    const MolInfo mi = a.mols[A.mol];
    const int nc = mi.n;
    const double* Dm = a.dens + mi.base;
    const double* Wm = a.wdens + mi.base;
    const double fpair = pr.sa == pr.sb ? 1.0 : 2.0;
    double wd[NA][NB], ww[NA][NB];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
        const size_t x = (size_t)(A.ao + ia) * nc + B.ao + ib, y = (size_t)(B.ao + ib) * nc + A.ao + ia;
        wd[ia][ib] = 0.5 * fpair * (Dm[x] + Dm[y]);
        ww[ia][ib] = 0.5 * fpair * (Wm[x] + Wm[y]);
      }
    const double pi = 3.14159265358979323846;
    const double s2p = sqrt(2.0 * pi * pi * sqrt(pi));
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
    for (int ip = 0; ip < pr.np; ++ip) {
      const PrimPair P = a.prims[pr.poff + ip];
      const double p = P.p, ip2 = 0.5 / P.p;
      const double Pr[3] = {P.px, P.py, P.pz};
      const double ckab = P.e * p / s2p;
      const double ai = P.aa, aj = P.ab;
      // 1D overlaps S1[x][a][b], a <= LA + 1, b <= LB + 2
      double S1[3][LA + 2][LB + 3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        const double pa = Pr[x] - Ar[x], pb = Pr[x] - Br[x];
#pragma unroll
        for (int ia = 0; ia <= LA + 1; ++ia)
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
      // 1D kinetic t(a, b) = -2 aj^2 s(a, b+2) + aj (2b+1) s(a, b) - b(b-1)/2 s(a, b-2)
      auto t1 = [&](int x, int ia, int ib) {
        double t = -2.0 * aj * aj * S1[x][ia][ib + 2] + aj * (2 * ib + 1) * S1[x][ia][ib];
        if (ib > 1) t -= 0.5 * ib * (ib - 1) * S1[x][ia][ib - 2];
        return t;
      };
#pragma unroll
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
          const int ea[3] = {cart_comp(LA, ia, 0), cart_comp(LA, ia, 1), cart_comp(LA, ia, 2)};
          const int eb[3] = {cart_comp(LB, ib, 0), cart_comp(LB, ib, 1), cart_comp(LB, ib, 2)};
          double s[3], t[3], ds[3], dt[3];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            s[x] = S1[x][ea[x]][eb[x]];
            t[x] = t1(x, ea[x], eb[x]);
            ds[x] = 2.0 * ai * S1[x][ea[x] + 1][eb[x]] - (ea[x] ? ea[x] * S1[x][ea[x] - 1][eb[x]] : 0.0);
            dt[x] = 2.0 * ai * t1(x, ea[x] + 1, eb[x]) - (ea[x] ? ea[x] * t1(x, ea[x] - 1, eb[x]) : 0.0);
          }
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            const int y = (x + 1) % 3, z = (x + 2) % 3;
            const double dS = ds[x] * s[y] * s[z];
            const double dT = dt[x] * s[y] * s[z] + ds[x] * (t[y] * s[z] + s[y] * t[z]);
            const double v = sfac * (wd[ia][ib] * dT - ww[ia][ib] * dS);
            gA[x] += v;
            gB[x] -= v;
          }
        }
      // nuclear attraction: A and B derivatives per nucleus, C by invariance
      for (int c = 0; c < a.nnuc; ++c) {
        const double* C = a.nuclei + 6 * c;
        if ((int)C[4] != A.mol) continue;
        const double pc[3] = {Pr[0] - C[0], Pr[1] - C[1], Pr[2] - C[2]};
        const double T = p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
        double u[NR], w[NR];
        rys_roots<NR>(T, u, w);
        const double vf = -C[3] * 2.0 * pi / p * ckab;
        double cA[3] = {0, 0, 0}, cB[3] = {0, 0, 0};
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const double b10 = ip2 * (1.0 - u[r]);
          double hA[3][LA + 2][LB + 1], hB[3][LA + 1][LB + 2];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            double g[LA + LB + 2];
            const double c00 = (Pr[x] - Ar[x]) - pc[x] * u[r];
            g[0] = x == 2 ? vf * w[r] : 1.0;
#pragma unroll
            for (int m = 0; m < LA + LB + 1; ++m) g[m + 1] = c00 * g[m] + (m ? m * b10 * g[m - 1] : 0.0);
            hrr_bra1<LA + 1, LB>(g, Ar[x] - Br[x], hA[x]);
            hrr_bra1<LA, LB + 1>(g, Ar[x] - Br[x], hB[x]);
          }
#pragma unroll
          for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
              const int ea[3] = {cart_comp(LA, ia, 0), cart_comp(LA, ia, 1), cart_comp(LA, ia, 2)};
              const int eb[3] = {cart_comp(LB, ib, 0), cart_comp(LB, ib, 1), cart_comp(LB, ib, 2)};
              double v0[3], dA[3], dB[3];
#pragma unroll
              for (int x = 0; x < 3; ++x) {
                v0[x] = hA[x][ea[x]][eb[x]];
                dA[x] = 2.0 * ai * hA[x][ea[x] + 1][eb[x]] - (ea[x] ? ea[x] * hA[x][ea[x] - 1][eb[x]] : 0.0);
                dB[x] = 2.0 * aj * hB[x][ea[x]][eb[x] + 1] - (eb[x] ? eb[x] * hB[x][ea[x]][eb[x] - 1] : 0.0);
              }
              const double wv = wd[ia][ib];
#pragma unroll
              for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                cA[x] += wv * dA[x] * v0[y] * v0[z];
                cB[x] += wv * dB[x] * v0[y] * v0[z];
              }
            }
        }
        const int atc = (int)C[5];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          gA[x] += cA[x];
          gB[x] += cB[x];
          atomicAdd(gl + 3 * atc + x, -(cA[x] + cB[x]));
        }
      }
    }
    const int atA = a.seg_atom[pr.sa], atB = a.seg_atom[pr.sb];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      atomicAdd(gl + 3 * atA + x, gA[x]);
      atomicAdd(gl + 3 * atB + x, gB[x]);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 3 * a.natm; k += blockDim.x) {
    const double v = gl[k];
    if (v != 0.0) atomicAdd(a.grad + k, v);
  }
}

void launch_grad1e_class(int la, int lb, const Pair* pairs, int npairs, const PrimPair* prims,
                         const ShellXYZ* shells, const int* seg_atom, const double* nuclei, int nnuc,
                         const MolInfo* mols, const double* dens, const double* wdens, int natm, double* grad,
                         hipStream_t st) {
  if (npairs <= 0) return;
  Grad1eArgs a{pairs, prims, shells, seg_atom, nuclei, nnuc, npairs, mols, dens, wdens, natm, grad};
  const dim3 grid((npairs + 255) / 256), block(256);
  const size_t lds = sizeof(double) * 3 * natm;
#define G1E(A, B)                                                          \
  if (la == A && lb == B) {                                                \
    hipLaunchKernelGGL((grad1e_kernel<A, B>), grid, block, lds, st, a);    \
    return;                                                                \
  }
  G1E(0, 0) G1E(1, 0) G1E(1, 1) G1E(2, 0) G1E(2, 1) G1E(2, 2)
#undef G1E
}

// ---------------------------------------------------------------------------
// Basis-parameter derivatives of E1 = sum D (T + V) - W S (symmetrised D, W):
// per primitive pair (i, j) of a segmented shell pair,
//   dE1/dc_i += [coefficient-free value] c_j,
//   dE1/d(alpha_i) += -sum_x value(a + 2_x) c_i c_j   (d/d alpha g = -|r-A|^2 g),
// and alike for j, with the raised overlaps/kinetic from 1D tables two units
// higher and the raised nuclear attraction from Rys tables two units higher.
// ---------------------------------------------------------------------------
struct Param1eArgs {
  const Pair* pairs;
  const PrimPair* prims;
  const PrimAux* aux;
  const ShellXYZ* shells;
  const int* seg_prim0;
  const double* nuclei;
  int nnuc;
  int npairs;
  const MolInfo* mols;
  const double* dens;
  const double* wdens;
  int nprim;
  double* gcoef;
  double* gexp;
};

template <int LA, int LB>
__global__ __launch_bounds__(256) void param1e_kernel(Param1eArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB);
  constexpr int NR = (LA + LB + 2) / 2 + 1;
  extern __shared__ double pl[];
  double* ycoef = pl;
  double* yexp = pl + a.nprim;
  for (int i = threadIdx.x; i < 2 * a.nprim; i += blockDim.x) pl[i] = 0.0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.npairs) {
    const Pair pr = a.pairs[i];
    const ShellXYZ A = a.shells[pr.sa], B = a.shells[pr.sb];
    const double Ar[3] = {A.x, A.y, A.z}, Br[3] = {B.x, B.y, B.z};
    const MolInfo mi = a.mols[A.mol];
    const int nc = mi.n;
    const double* Dm = a.dens + mi.base;
    const double* Wm = a.wdens + mi.base;
    const double fpair = pr.sa == pr.sb ? 1.0 : 2.0;
    double wd[NA][NB], ww[NA][NB];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
        const size_t x = (size_t)(A.ao + ia) * nc + B.ao + ib, y = (size_t)(B.ao + ib) * nc + A.ao + ia;
        wd[ia][ib] = 0.5 * fpair * (Dm[x] + Dm[y]);
        ww[ia][ib] = 0.5 * fpair * (Wm[x] + Wm[y]);
      }
    const double pi = 3.14159265358979323846;
    const double s2p = sqrt(2.0 * pi * pi * sqrt(pi));
    const int pA = a.seg_prim0[pr.sa], pB = a.seg_prim0[pr.sb];
    for (int ip = 0; ip < pr.np; ++ip) {
      const PrimPair P = a.prims[pr.poff + ip];
      const PrimAux X = a.aux[pr.poff + ip];
      const double p = P.p, ip2 = 0.5 / P.p;
      const double Pr[3] = {P.px, P.py, P.pz};
      const double kab = X.ek * p / s2p;  // K_AB (coefficient-free)
      const double ai = P.aa, aj = P.ab;
      double S1[3][LA + 3][LB + 5];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        const double pa = Pr[x] - Ar[x], pb = Pr[x] - Br[x];
#pragma unroll
        for (int ia = 0; ia <= LA + 2; ++ia)
#pragma unroll
          for (int ib = 0; ib <= LB + 4; ++ib) {
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
      const double sfac = kab * (pi / p) * sqrt(pi / p);
      auto t1 = [&](int x, int ia, int ib) {
        double t = -2.0 * aj * aj * S1[x][ia][ib + 2] + aj * (2 * ib + 1) * S1[x][ia][ib];
        if (ib > 1) t -= 0.5 * ib * (ib - 1) * S1[x][ia][ib - 2];
        return t;
      };
      double val = 0.0, rA = 0.0, rB = 0.0;
#pragma unroll
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
          const int ea[3] = {cart_comp(LA, ia, 0), cart_comp(LA, ia, 1), cart_comp(LA, ia, 2)};
          const int eb[3] = {cart_comp(LB, ib, 0), cart_comp(LB, ib, 1), cart_comp(LB, ib, 2)};
          double s[3], t[3];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            s[x] = S1[x][ea[x]][eb[x]];
            t[x] = t1(x, ea[x], eb[x]);
          }
          const double sv = s[0] * s[1] * s[2];
          const double tv = t[0] * s[1] * s[2] + s[0] * t[1] * s[2] + s[0] * s[1] * t[2];
          val += sfac * (wd[ia][ib] * tv - ww[ia][ib] * sv);
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            const int y = (x + 1) % 3, z = (x + 2) % 3;
            // a raised by two on axis x
            const double sa2 = S1[x][ea[x] + 2][eb[x]], ta2 = t1(x, ea[x] + 2, eb[x]);
            rA += sfac * (wd[ia][ib] * (ta2 * s[y] * s[z] + sa2 * (t[y] * s[z] + s[y] * t[z])) -
                          ww[ia][ib] * sa2 * s[y] * s[z]);
            // b raised by two on axis x
            const double sb2 = S1[x][ea[x]][eb[x] + 2], tb2 = t1(x, ea[x], eb[x] + 2);
            rB += sfac * (wd[ia][ib] * (tb2 * s[y] * s[z] + sb2 * (t[y] * s[z] + s[y] * t[z])) -
                          ww[ia][ib] * sb2 * s[y] * s[z]);
          }
        }
      for (int c = 0; c < a.nnuc; ++c) {
        const double* C = a.nuclei + 6 * c;
        if ((int)C[4] != A.mol) continue;
        const double pc[3] = {Pr[0] - C[0], Pr[1] - C[1], Pr[2] - C[2]};
        const double T = p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
        double u[NR], w[NR];
        rys_roots<NR>(T, u, w);
        const double vf = -C[3] * 2.0 * pi / p * kab;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const double b10 = ip2 * (1.0 - u[r]);
          double hA[3][LA + 3][LB + 1], hB[3][LA + 1][LB + 3];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            double g[LA + LB + 3];
            const double c00 = (Pr[x] - Ar[x]) - pc[x] * u[r];
            g[0] = x == 2 ? vf * w[r] : 1.0;
#pragma unroll
            for (int m = 0; m < LA + LB + 2; ++m) g[m + 1] = c00 * g[m] + (m ? m * b10 * g[m - 1] : 0.0);
            hrr_bra1<LA + 2, LB>(g, Ar[x] - Br[x], hA[x]);
            hrr_bra1<LA, LB + 2>(g, Ar[x] - Br[x], hB[x]);
          }
#pragma unroll
          for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
              const int ea[3] = {cart_comp(LA, ia, 0), cart_comp(LA, ia, 1), cart_comp(LA, ia, 2)};
              const int eb[3] = {cart_comp(LB, ib, 0), cart_comp(LB, ib, 1), cart_comp(LB, ib, 2)};
              double v[3];
#pragma unroll
              for (int x = 0; x < 3; ++x) v[x] = hA[x][ea[x]][eb[x]];
              const double o[3] = {v[1] * v[2], v[0] * v[2], v[0] * v[1]};
              const double wv = wd[ia][ib];
              val += wv * v[0] * o[0];
#pragma unroll
              for (int x = 0; x < 3; ++x) {
                rA += wv * hA[x][ea[x] + 2][eb[x]] * o[x];
                rB += wv * hB[x][ea[x]][eb[x] + 2] * o[x];
              }
            }
        }
      }
      atomicAdd(ycoef + pA + X.ia, val * X.cb);
      atomicAdd(ycoef + pB + X.ib, val * X.ca);
      atomicAdd(yexp + pA + X.ia, -rA * X.ca * X.cb);
      atomicAdd(yexp + pB + X.ib, -rB * X.ca * X.cb);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < a.nprim; k += blockDim.x) {
    const double vc = ycoef[k], ve = yexp[k];
    if (vc != 0.0) atomicAdd(a.gcoef + k, vc);
    if (ve != 0.0) atomicAdd(a.gexp + k, ve);
  }
}

void launch_param1e_class(int la, int lb, const Pair* pairs, int npairs, const PrimPair* prims, const PrimAux* aux,
                          const ShellXYZ* shells, const int* seg_prim0, const double* nuclei, int nnuc,
                          const MolInfo* mols, const double* dens, const double* wdens, int nprim, double* gcoef,
                          double* gexp, hipStream_t st) {
  if (npairs <= 0) return;
  Param1eArgs a{pairs, prims, aux, shells, seg_prim0, nuclei, nnuc, npairs, mols, dens, wdens, nprim, gcoef, gexp};
  const dim3 grid((npairs + 255) / 256), block(256);
  const size_t lds = sizeof(double) * 2 * nprim;
#define P1E(A, B)                                                          \
  if (la == A && lb == B) {                                                \
    hipLaunchKernelGGL((param1e_kernel<A, B>), grid, block, lds, st, a);   \
    return;                                                                \
  }
  P1E(0, 0) P1E(1, 0) P1E(1, 1) P1E(2, 0) P1E(2, 1) P1E(2, 2)
#undef P1E
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
