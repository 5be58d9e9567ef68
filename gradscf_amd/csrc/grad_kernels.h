// Coordinate VJP of direct J/K: contraction of first ERI derivatives with
// the density and the J/K cotangents, on canonical quartets.
//
// Reference semantics (src/gradscf/integrals/_native/csrc/direct_jk_geometry.cc:68-84
// with csrc/eri_geometry_entries.h:5-31): over all s1 quartets,
//   grad[R] += d(ab|cd)/dR * (cJ_ab D_cd + cK_ac D_bd).
// Here each canonical quartet carries the 8-permutation sum of that weight,
//   W = f [ (cJ_ab + cJ_ba)(D_cd + D_dc) + (cJ_cd + cJ_dc)(D_ab + D_ba)
//           + cK_ac D_bd + cK_bc D_ad + cK_ad D_bc + cK_bd D_ac
//           + cK_ca D_db + cK_da D_cb + cK_cb D_da + cK_db D_ca ],
// and the derivative of the (permutation-invariant) integral with respect to
// the centres of A, B and C is formed from Rys tables with one extra unit of
// angular momentum on the bra and on the ket:
//   d/dA_x [a b|c d] = 2 alpha_a [a+1_x b|c d] - a_x [a-1_x b|c d]   (B, C alike),
// the D derivative by translational invariance, d/dD = -(d/dA + d/dB + d/dC).
// One lane per (quartet item, Cartesian ket component pair); per-atom sums are
// accumulated in LDS and flushed once per workgroup.
#pragma once
#include "jk_kernels.h"

namespace gscf {


__device__ __forceinline__ void map_item_g(const GradArgs& a, int64_t item, int64_t& g, int& s, int& S) {
  int lo = 0, hi = a.ntasks;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a.tasks[mid].item0 <= item) lo = mid; else hi = mid;
  }
  const SubTask t = a.tasks[lo];
  const int64_t r = item - t.item0;
  s = (int)(r / t.nq);
  g = t.q0 + (r - (int64_t)s * t.nq);
  S = t.S;
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void grad_kernel(GradArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NKET = NC * ND;
  constexpr int NM = LA + LB + 1, MM = LC + LD + 1;
  constexpr int NR = (LA + LB + LC + LD + 1) / 2 + 1;
  extern __shared__ double gl[];
  for (int i = threadIdx.x; i < 3 * a.natm; i += blockDim.x) gl[i] = 0.0;
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t item = t / NKET;
  if (item < a.nitems) {
    const int kc = (int)(t - item * NKET);
    const int ic = kc / ND, id = kc - (kc / ND) * ND;
    int64_t g;
    int s0, S;
    map_item_g(a, item, g, s0, S);
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
    // 8-permutation weight of every bra component for this ket component
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
    double w0[3][MM + 1], wp[3][MM + 1], wm[3][MM + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      ket_weights<MM>(ec[x], ed[x], CD[x], w0[x]);
      ket_weights<MM>(ec[x] + 1, ed[x], CD[x], wp[x]);
      if (ec[x] > 0) {
        ket_weights<MM>(ec[x] - 1, ed[x], CD[x], wm[x]);
      } else {
#pragma unroll
        for (int m = 0; m <= MM; ++m) wm[x][m] = 0.0;
      }
    }
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
    for (int ip = s0; ip < bp.np; ip += S) {
      const PrimPair P = a.prims[bp.poff + ip];
      for (int kq = 0; kq < kp.np; ++kq) {
        const PrimPair Q = a.prims[kp.poff + kq];
        PrimQ<NR> q;
        prim_setup<NR>(P, Q, geo, q);
        const double ta = 2.0 * P.aa, tb = 2.0 * P.ab, tc = 2.0 * Q.aa;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const double u = q.s[r];
          const double b00 = 0.5 * u * q.ipq;
          const double b10 = 0.5 * q.ip * (1.0 - q.qf * u);
          const double b01 = 0.5 * q.iq * (1.0 - q.pf * u);
          double hA[3][LA + 2][LB + 1], hB[3][LA + 1][LB + 2], hP[3][LA + 1][LB + 1], hM[3][LA + 1][LB + 1];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            double G[NM + 1][MM + 1];
            const double c00 = q.PA[x] - q.qf * q.PQ[x] * u;
            const double c0p = q.QC[x] + q.pf * q.PQ[x] * u;
            vrr<NM, MM>(c00, c0p, b10, b01, b00, x == 2 ? q.w[r] * q.pref : 1.0, G);
            double g0[NM + 1], gp[NM + 1], gm[NM + 1];
#pragma unroll
            for (int nn = 0; nn <= NM; ++nn) {
              double s_0 = 0.0, s_p = 0.0, s_m = 0.0;
#pragma unroll
              for (int m = 0; m <= MM; ++m) {
                s_0 += w0[x][m] * G[nn][m];
                s_p += wp[x][m] * G[nn][m];
                s_m += wm[x][m] * G[nn][m];
              }
              g0[nn] = s_0;
              gp[nn] = s_p;
              gm[nn] = s_m;
            }
            hrr_bra1<LA + 1, LB>(g0, AB[x], hA[x]);
            double g0b[LA + LB + 2];
#pragma unroll
            for (int nn = 0; nn <= NM; ++nn) g0b[nn] = g0[nn];
            hrr_bra1<LA, LB + 1>(g0b, AB[x], hB[x]);
            double gpc[LA + LB + 1], gmc[LA + LB + 1];
#pragma unroll
            for (int nn = 0; nn <= LA + LB; ++nn) {
              gpc[nn] = gp[nn];
              gmc[nn] = gm[nn];
            }
            hrr_bra1<LA, LB>(gpc, AB[x], hP[x]);
            hrr_bra1<LA, LB>(gmc, AB[x], hM[x]);
          }
#pragma unroll
          for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
              const int e_a[3] = {cart_comp(LA, ia, 0), cart_comp(LA, ia, 1), cart_comp(LA, ia, 2)};
              const int e_b[3] = {cart_comp(LB, ib, 0), cart_comp(LB, ib, 1), cart_comp(LB, ib, 2)};
              double v[3], dA[3], dB[3], dC[3];
#pragma unroll
              for (int x = 0; x < 3; ++x) {
                v[x] = hA[x][e_a[x]][e_b[x]];
                dA[x] = ta * hA[x][e_a[x] + 1][e_b[x]] - (e_a[x] ? e_a[x] * hA[x][e_a[x] - 1][e_b[x]] : 0.0);
                dB[x] = tb * hB[x][e_a[x]][e_b[x] + 1] - (e_b[x] ? e_b[x] * hB[x][e_a[x]][e_b[x] - 1] : 0.0);
                dC[x] = tc * hP[x][e_a[x]][e_b[x]] - (ec[x] ? ec[x] * hM[x][e_a[x]][e_b[x]] : 0.0);
              }
              const double w = W[ia][ib];
              const double vyz = v[1] * v[2], vxz = v[0] * v[2], vxy = v[0] * v[1];
              gA[0] += w * dA[0] * vyz; gA[1] += w * dA[1] * vxz; gA[2] += w * dA[2] * vxy;
              gB[0] += w * dB[0] * vyz; gB[1] += w * dB[1] * vxz; gB[2] += w * dB[2] * vxy;
              gC[0] += w * dC[0] * vyz; gC[1] += w * dC[1] * vxz; gC[2] += w * dC[2] * vxy;
            }
        }
      }
    }
    const int atA = a.seg_atom[bp.sa], atB = a.seg_atom[bp.sb], atC = a.seg_atom[kp.sa], atD = a.seg_atom[kp.sb];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      atomicAdd(gl + 3 * atA + x, gA[x]);
      atomicAdd(gl + 3 * atB + x, gB[x]);
      atomicAdd(gl + 3 * atC + x, gC[x]);
      atomicAdd(gl + 3 * atD + x, -(gA[x] + gB[x] + gC[x]));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * a.natm; i += blockDim.x) {
    const double v = gl[i];
    if (v != 0.0) atomicAdd(a.grad + i, v);
  }
}

template <int LA, int LB, int LC, int LD>
void launch_grad_t(const GradArgs& a, hipStream_t s) {
  constexpr int NKET = ncart(LC) * ncart(LD);
  if (a.nitems <= 0) return;
  const int64_t threads = a.nitems * NKET;
  const int64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL((grad_kernel<LA, LB, LC, LD>), dim3((unsigned)blocks), dim3(256),
                     sizeof(double) * 3 * a.natm, s, a);
}

}  // namespace gscf
