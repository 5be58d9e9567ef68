// Basis-parameter VJP of direct J/K: derivatives of
//   E = sum(cot_J * J(D)) + sum(cot_K * K(D))
// with respect to every primitive exponent and every (normalised) contraction
// coefficient in env, on canonical quartets.
//
// Reference semantics: the reference differentiates J/K through its JAX
// normalisation chain for coefficients and through d/d(alpha) g = -|r - A|^2 g
// for exponents (src/gradscf/integrals/_native/csrc/exponents.cc:1-3,
// src/gradscf/integrals/backends/native/jk.py:167-249).  Here, per primitive
// quartet (i j | k l) with the 8-permutation weight W of the canonical quartet
// (as in the coordinate VJP):
//   coefficient:  dE/de_i += W . [i j|k l]' e_j e_k e_l   (e-free integral ',
//                 so a zero coefficient still gets its derivative)
//   exponent:     dE/d(alpha_i) += -W . sum_x [(i + 2_x) j|k l] e_i e_j e_k e_l
// and alike for the j, k, l slots; the raised integrals come from Rys tables
// with two more units of angular momentum on the bra and on the ket.  One lane
// per (quartet item, Cartesian ket component pair); per-primitive sums are
// accumulated in LDS and flushed once per workgroup.
#pragma once
#include "grad_kernels.h"

namespace gscf {

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void param_kernel(ParamArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NKET = NC * ND;
  constexpr int NM = LA + LB + 2, MM = LC + LD + 2;  // raised by two on each side
  constexpr int NR = (LA + LB + LC + LD + 2) / 2 + 1;
  extern __shared__ double pl[];
  double* ycoef = pl;               // [nprim] dE/de (coefficient-free integrals)
  double* yexp = pl + a.nprim;      // [nprim] dE/d alpha
  for (int i = threadIdx.x; i < 2 * a.nprim; i += blockDim.x) pl[i] = 0.0;
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t item = t / NKET;
  if (item < a.nitems) {
    const int kc = (int)(t - item * NKET);
    const int ic = kc / ND, id = kc - (kc / ND) * ND;
    int64_t g;
    int s0, S;
    {
      int lo = 0, hi = a.ntasks;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.tasks[mid].item0 <= item) lo = mid; else hi = mid;
      }
      const SubTask tk = a.tasks[lo];
      const int64_t r = item - tk.item0;
      s0 = (int)(r / tk.nq);
      g = tk.q0 + (r - (int64_t)s0 * tk.nq);
      S = tk.S;
    }
    const int2 qk = a.quartets[g];
    const Pair bp = a.pairs[qk.x];
    const Pair kp = a.pairs[qk.y];
    const ShellXYZ sA = a.shells[bp.sa], sB = a.shells[bp.sb], sC = a.shells[kp.sa], sD = a.shells[kp.sb];
    QuartetGeom geo;
    double AB[3], CD[3];
    load_geom(sA, sB, sC, sD, geo, AB, CD);
    double f = 1.0;
    if (bp.sa == bp.sb) f *= 0.5;
    if (kp.sa == kp.sb) f *= 0.5;
    if (qk.x == qk.y) f *= 0.5;
    const MolInfo mi = a.mols[sA.mol];
    const int n = mi.n;
    const int oa = sA.ao, ob = sB.ao;
    const int c = sC.ao + ic, d = sD.ao + id;
    const double* Dm = a.dens + mi.base;
    const double* CJ = a.cj ? a.cj + mi.base : nullptr;
    const double* CK = a.ck ? a.ck + mi.base : nullptr;
    double W[NA][NB];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
        const int p = oa + ia, q = ob + ib;
        double w = 0.0;
        if (CJ) {
          w += (CJ[(size_t)p * n + q] + CJ[(size_t)q * n + p]) * (Dm[(size_t)c * n + d] + Dm[(size_t)d * n + c]);
          w += (CJ[(size_t)c * n + d] + CJ[(size_t)d * n + c]) * (Dm[(size_t)p * n + q] + Dm[(size_t)q * n + p]);
        }
        if (CK) {
          const double* K = CK;
          w += K[(size_t)p * n + c] * Dm[(size_t)q * n + d] + K[(size_t)q * n + c] * Dm[(size_t)p * n + d] +
               K[(size_t)p * n + d] * Dm[(size_t)q * n + c] + K[(size_t)q * n + d] * Dm[(size_t)p * n + c] +
               K[(size_t)c * n + p] * Dm[(size_t)d * n + q] + K[(size_t)d * n + p] * Dm[(size_t)c * n + q] +
               K[(size_t)c * n + q] * Dm[(size_t)d * n + p] + K[(size_t)d * n + q] * Dm[(size_t)c * n + p];
        }
        W[ia][ib] = f * w;
      }
    int ec[3], ed[3];
    cart_xyz<LC>(ic, ec);
    cart_xyz<LD>(id, ed);
    // ket weights: plain, c raised by two, d raised by two (per axis)
    double w0[3][MM + 1], wc[3][MM + 1], wd[3][MM + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      ket_weights<MM>(ec[x], ed[x], CD[x], w0[x]);
      ket_weights<MM>(ec[x] + 2, ed[x], CD[x], wc[x]);
      ket_weights<MM>(ec[x], ed[x] + 2, CD[x], wd[x]);
    }
    const int pA = a.seg_prim0[bp.sa], pB = a.seg_prim0[bp.sb];
    const int pC = a.seg_prim0[kp.sa], pD = a.seg_prim0[kp.sb];
    for (int ip = s0; ip < bp.np; ip += S) {
      PrimPair P = a.prims[bp.poff + ip];
      const PrimAux xP = a.aux[bp.poff + ip];
      P.e = xP.ek;  // coefficient-free bra factor
      for (int kq = 0; kq < kp.np; ++kq) {
        PrimPair Q = a.prims[kp.poff + kq];
        const PrimAux xQ = a.aux[kp.poff + kq];
        Q.e = xQ.ek;
        PrimQ<NR> q;
        prim_setup<NR>(P, Q, geo, q);
        double sV = 0.0, sA = 0.0, sB = 0.0, sC = 0.0, sD = 0.0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const double u = q.s[r];
          const double b00 = 0.5 * u * q.ipq;
          const double b10 = 0.5 * q.ip * (1.0 - q.qf * u);
          const double b01 = 0.5 * q.iq * (1.0 - q.pf * u);
          double hA[3][LA + 3][LB + 1], hB[3][LA + 1][LB + 3], hC[3][LA + 1][LB + 1], hD[3][LA + 1][LB + 1];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            double G[NM + 1][MM + 1];
            const double c00 = q.PA[x] - q.qf * q.PQ[x] * u;
            const double c0p = q.QC[x] + q.pf * q.PQ[x] * u;
            vrr<NM, MM>(c00, c0p, b10, b01, b00, x == 2 ? q.w[r] * q.pref : 1.0, G);
            double g0[NM + 1], gc[LA + LB + 1], gd[LA + LB + 1];
#pragma unroll
            for (int nn = 0; nn <= NM; ++nn) {
              double s_0 = 0.0, s_c = 0.0, s_d = 0.0;
#pragma unroll
              for (int m = 0; m <= MM; ++m) {
                s_0 += w0[x][m] * G[nn][m];
                s_c += wc[x][m] * G[nn][m];
                s_d += wd[x][m] * G[nn][m];
              }
              g0[nn] = s_0;
              if (nn <= LA + LB) {
                gc[nn] = s_c;
                gd[nn] = s_d;
              }
            }
            hrr_bra1<LA + 2, LB>(g0, AB[x], hA[x]);
            hrr_bra1<LA, LB + 2>(g0, AB[x], hB[x]);
            hrr_bra1<LA, LB>(gc, AB[x], hC[x]);
            hrr_bra1<LA, LB>(gd, AB[x], hD[x]);
          }
#pragma unroll
          for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
              const int e_a[3] = {cart_comp(LA, ia, 0), cart_comp(LA, ia, 1), cart_comp(LA, ia, 2)};
              const int e_b[3] = {cart_comp(LB, ib, 0), cart_comp(LB, ib, 1), cart_comp(LB, ib, 2)};
              double v[3];
#pragma unroll
              for (int x = 0; x < 3; ++x) v[x] = hA[x][e_a[x]][e_b[x]];
              const double o[3] = {v[1] * v[2], v[0] * v[2], v[0] * v[1]};
              double rA = 0.0, rB = 0.0, rC = 0.0, rD = 0.0;
#pragma unroll
              for (int x = 0; x < 3; ++x) {
                rA += hA[x][e_a[x] + 2][e_b[x]] * o[x];
                rB += hB[x][e_a[x]][e_b[x] + 2] * o[x];
                rC += hC[x][e_a[x]][e_b[x]] * o[x];
                rD += hD[x][e_a[x]][e_b[x]] * o[x];
              }
              const double w = W[ia][ib];
              sV += w * v[0] * o[0];
              sA += w * rA;
              sB += w * rB;
              sC += w * rC;
              sD += w * rD;
            }
        }
        const double ca = xP.ca, cb = xP.cb, cc = xQ.ca, cd = xQ.cb;
        const double all = ca * cb * cc * cd;
        atomicAdd(ycoef + pA + xP.ia, sV * cb * cc * cd);
        atomicAdd(ycoef + pB + xP.ib, sV * ca * cc * cd);
        atomicAdd(ycoef + pC + xQ.ia, sV * ca * cb * cd);
        atomicAdd(ycoef + pD + xQ.ib, sV * ca * cb * cc);
        atomicAdd(yexp + pA + xP.ia, -sA * all);
        atomicAdd(yexp + pB + xP.ib, -sB * all);
        atomicAdd(yexp + pC + xQ.ia, -sC * all);
        atomicAdd(yexp + pD + xQ.ib, -sD * all);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < a.nprim; i += blockDim.x) {
    const double vc = ycoef[i], ve = yexp[i];
    if (vc != 0.0) atomicAdd(a.gcoef + i, vc);
    if (ve != 0.0) atomicAdd(a.gexp + i, ve);
  }
}

template <int LA, int LB, int LC, int LD>
void launch_param_t(const ParamArgs& a, hipStream_t s) {
  constexpr int NKET = ncart(LC) * ncart(LD);
  if (a.nitems <= 0) return;
  const int64_t threads = a.nitems * NKET;
  const int64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL((param_kernel<LA, LB, LC, LD>), dim3((unsigned)blocks), dim3(256),
                     sizeof(double) * 2 * a.nprim, s, a);
}

}  // namespace gscf
