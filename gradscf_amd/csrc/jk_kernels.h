// Rys-quadrature shell-quartet kernels with fused J/K digestion (FP64), as templates;
// instantiated one class per translation unit by jk_class.hip.
//
// One template instantiation per canonical angular-momentum class
// (LA>=LB, LC>=LD, pair class AB >= CD), 21 classes for s/p/d.  Two layouts:
//   * full:  one lane = one shell quartet; all Cartesian components of the
//            block live in registers (small classes);
//   * split: one lane = (shell quartet, one Cartesian ket component pair);
//            the ket transfer is applied as a weight vector on the vertical
//            table, the bra transfer and the block slice stay in registers
//            (large classes).
// Digestion follows the 8-fold permutational sum of csrc/direct_jk.cc:45-76,
// folded into partial matrices: Jp[a,b] += f v (D_cd + D_dc),
// Jp[c,d] += f v (D_ab + D_ba), Kp[a,c] += f v D_bd, Kp[a,d] += f v D_bc,
// Kp[b,c] += f v D_ad, Kp[b,d] += f v D_ac, with f = 1/2 per coincident
// shell / pair; J = Jp + Jp^T and K = Kp(D) + Kp(D^T)^T are formed afterwards.
#pragma once
#include <type_traits>

#include "gscf_device.h"
#include "gscf_args.h"

namespace gscf {

// Work item -> (quartet index, primitive split index s, split count S).
__device__ __forceinline__ void map_item(const JKArgs& a, int64_t item, int64_t& g, int& s, int& S) {
  int lo = 0, hi = a.ntasks;  // tasks[lo].item0 <= item < tasks[hi].item0
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
struct Class {
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NM = LA + LB, MM = LC + LD;
  static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
  static constexpr int NBRA = NA * NB, NKET = NC * ND;
  static constexpr bool kSplit = NBRA * NKET > 36;
  static constexpr int SPLIT = kSplit ? NKET : 1;
};

struct QuartetGeom {
  double A[3], B[3], C[3], D[3];
};

// Per primitive quartet: roots and the per-root recurrence coefficients.
template <int NR>
struct PrimQ {
  double s[NR], w[NR];
  double pref, ip, iq, ipq, qf, pf;
  double PA[3], QC[3], PQ[3];
};

template <int NR, bool MOMENTS = (NR <= 2)>
__device__ __forceinline__ void prim_setup(const PrimPair& P, const PrimPair& Q, const QuartetGeom& g,
                                           PrimQ<NR>& o, const double* __restrict__ tab) {
  const double pq = P.p + Q.p;
  const double rs = rsqrt_pos(pq);  // 1/sqrt(p+q)
  o.ipq = rs * rs;
  o.ip = P.ip;
  o.iq = Q.ip;
  o.PQ[0] = P.px - Q.px;
  o.PQ[1] = P.py - Q.py;
  o.PQ[2] = P.pz - Q.pz;
  o.PA[0] = P.px - g.A[0];
  o.PA[1] = P.py - g.A[1];
  o.PA[2] = P.pz - g.A[2];
  o.QC[0] = Q.px - g.C[0];
  o.QC[1] = Q.py - g.C[1];
  o.QC[2] = Q.pz - g.C[2];
  o.qf = Q.p * o.ipq;  // rho / p
  o.pf = P.p * o.ipq;  // rho / q
  const double rho = P.p * Q.p * o.ipq;
  const double T = rho * (o.PQ[0] * o.PQ[0] + o.PQ[1] * o.PQ[1] + o.PQ[2] * o.PQ[2]);
  o.pref = P.e * Q.e * rs;
  rys_roots_tab<NR, MOMENTS>(tab, T, o.s, o.w);
}

template <int NR>
__device__ __forceinline__ void prim_setup(const PrimPair& P, const PrimPair& Q, const QuartetGeom& g,
                                           PrimQ<NR>& o) {
  prim_setup<NR>(P, Q, g, o, rys_tab + rys_table_offset[NR - 1]);
}

// Full block of one primitive quartet accumulated into acc[NA][NB][NC][ND].
// The per-root coefficients are formed in FP64 and the recurrences, transfers
// and accumulation run in R (double, or float on the mixed-precision path).
template <int LA, int LB, int LC, int LD, typename R = double>
__device__ __forceinline__ void accumulate_full(const PrimQ<Class<LA, LB, LC, LD>::NR>& q, const R AB[3],
                                                const R CD[3],
                                                R (&acc)[ncart(LA)][ncart(LB)][ncart(LC)][ncart(LD)]) {
  using K = Class<LA, LB, LC, LD>;
#pragma unroll
  for (int r = 0; r < K::NR; ++r) {
    const double u = q.s[r];
    const R b00 = (R)(0.5 * u * q.ipq);
    const R b10 = (R)(0.5 * q.ip * (1.0 - q.qf * u));
    const R b01 = (R)(0.5 * q.iq * (1.0 - q.pf * u));
    R H[3][LA + 1][LB + 1][LC + 1][LD + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      R G[K::NM + 1][K::MM + 1];
      const R c00 = (R)(q.PA[x] - q.qf * q.PQ[x] * u);
      const R c0p = (R)(q.QC[x] + q.pf * q.PQ[x] * u);
      vrr<K::NM, K::MM, R>(c00, c0p, b10, b01, b00, (R)(x == 2 ? q.w[r] * q.pref : 1.0), G);
      hrr_full<LA, LB, LC, LD, R>(G, AB[x], CD[x], H[x]);
    }
#pragma unroll
    for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < K::NB; ++ib)
#pragma unroll
        for (int ic = 0; ic < K::NC; ++ic)
#pragma unroll
          for (int id = 0; id < K::ND; ++id) {
            const R vx = H[0][cart_comp(LA, ia, 0)][cart_comp(LB, ib, 0)][cart_comp(LC, ic, 0)][cart_comp(LD, id, 0)];
            const R vy = H[1][cart_comp(LA, ia, 1)][cart_comp(LB, ib, 1)][cart_comp(LC, ic, 1)][cart_comp(LD, id, 1)];
            const R vz = H[2][cart_comp(LA, ia, 2)][cart_comp(LB, ib, 2)][cart_comp(LC, ic, 2)][cart_comp(LD, id, 2)];
            acc[ia][ib][ic][id] += vx * vy * vz;
          }
  }
}

// One primitive quartet from a staged bra record (p, P - A, e, 1/p) and a
// ket primitive pair (with QA = Q - A, QC = Q - C), added into blk: the
// one-root s-ket classes in the Obara-Saika form from F_0 (and F_1),
//   (s|ss) = pref F0,  (p_x|ss) = pref (PA_x F0 - (q/(p+q)) PQ_x F1),
// everything else by the Rys quadrature (accumulate_full).
template <int LA, int LB, int LC, int LD, int MOMENTS, typename R>
__device__ __forceinline__ void staged_quartet(const double* __restrict__ r, const PrimPair& Q, const double (&QA)[3],
                                               const double (&QC)[3], const R AB[3], const R CD[3],
                                               const double* __restrict__ tab,
                                               R (&blk)[ncart(LA)][ncart(LB)][ncart(LC)][ncart(LD)]) {
  using K = Class<LA, LB, LC, LD>;
  const double PA[3] = {r[1], r[2], r[3]};
  const double PQ[3] = {PA[0] - QA[0], PA[1] - QA[1], PA[2] - QA[2]};
  const double pq = r[0] + Q.p;
  const double d2 = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
  if constexpr (K::NR == 1 && LB == 0 && std::is_same<R, double>::value) {
    // X = p q |PQ|^2 = (p + q) T.  With pref = e_ab e_cd / sqrt(p + q):
    //   T >= TMAX:  pref F0 = (sqrt(pi)/2) e_ab e_cd / sqrt(X),
    //               pref (q/(p+q)) F1 = pref F0 q / (2 X)      (F1 = F0 / (2T)),
    //   T <  TMAX:  T = X / (p + q), F0 and F1 from the grid.
    // The branch test needs no division (X >= TMAX (p + q)), and both
    // branches take one reciprocal square root, of X or of p + q.
    const double X = r[0] * Q.p * d2;
    const bool asym = X >= BOYS_TMAX * pq;
    const double y = rsqrt_pos(asym ? X : pq);
    double af, bf = 0.0;
    if (asym) {
      af = 0.88622692545275801365 * r[4] * Q.e * y;  // sqrt(pi) / 2
      if constexpr (LA == 1) bf = 0.5 * af * Q.p * (y * y);
    } else {
      const double ipq = y * y;
      double F[LA + 1];
      boys_taylor<LA + 1>(X * ipq, F);
      const double pref = r[4] * Q.e * y;
      af = pref * F[0];
      if constexpr (LA == 1) bf = pref * Q.p * ipq * F[1];
    }
    if constexpr (LA == 0) {
      blk[0][0][0][0] += af;
    } else {
#pragma unroll
      for (int x = 0; x < 3; ++x) blk[x][0][0][0] += PA[x] * af - PQ[x] * bf;
    }
  } else {
    const double rs = rsqrt_pos(pq);
    const double ipq = rs * rs;
    const double T = r[0] * Q.p * ipq * d2;
    const double pref = r[4] * Q.e * rs;
    PrimQ<K::NR> q;
    q.ipq = ipq;
    q.ip = r[5];
    q.iq = Q.ip;
    q.qf = Q.p * ipq;
    q.pf = r[0] * ipq;
    q.pref = pref;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      q.PA[x] = PA[x];
      q.QC[x] = QC[x];
      q.PQ[x] = PQ[x];
    }
    rys_roots_tab<K::NR, MOMENTS>(tab, T, q.s, q.w);
    accumulate_full<LA, LB, LC, LD, R>(q, AB, CD, blk);
  }
}

// Runtime Cartesian exponents of component i of shell l.
template <int L>
__device__ __forceinline__ void cart_xyz(int i, int (&e)[3]) {
  e[0] = e[1] = e[2] = 0;
#pragma unroll
  for (int k = 0; k < ncart(L); ++k)
    if (k == i) {
      e[0] = cart_comp(L, k, 0);
      e[1] = cart_comp(L, k, 1);
      e[2] = cart_comp(L, k, 2);
    }
}

// Ket transfer weights for a fixed ket component pair (c, d) on one axis:
// I(n; c, d) = sum_m wk[m] G(n, m), wk[m] = C(d, m-c) CD^{d-(m-c)}.
template <int MM, typename R = double>
__device__ __forceinline__ void ket_weights(int c, int d, R cd, R (&wk)[MM + 1]) {
#pragma unroll
  for (int m = 0; m <= MM; ++m) {
    const int k = m - c;
    R v = R(0);
    if (k >= 0 && k <= d) {
      R pw = R(1);
      for (int t = 0; t < d - k; ++t) pw *= cd;
      v = (R)binom(d, k) * pw;
    }
    wk[m] = v;
  }
}

// Block slice for one ket component pair accumulated into acc[NA][NB].
template <int LA, int LB, int LC, int LD, typename R = double>
__device__ __forceinline__ void accumulate_split(const PrimQ<Class<LA, LB, LC, LD>::NR>& q, const R AB[3],
                                                 const R (&wk)[3][LC + LD + 1],
                                                 R (&acc)[ncart(LA)][ncart(LB)]) {
  using K = Class<LA, LB, LC, LD>;
#pragma unroll
  for (int r = 0; r < K::NR; ++r) {
    const double u = q.s[r];
    const R b00 = (R)(0.5 * u * q.ipq);
    const R b10 = (R)(0.5 * q.ip * (1.0 - q.qf * u));
    const R b01 = (R)(0.5 * q.iq * (1.0 - q.pf * u));
    R h[3][LA + 1][LB + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      R G[K::NM + 1][K::MM + 1];
      const R c00 = (R)(q.PA[x] - q.qf * q.PQ[x] * u);
      const R c0p = (R)(q.QC[x] + q.pf * q.PQ[x] * u);
      vrr<K::NM, K::MM, R>(c00, c0p, b10, b01, b00, (R)(x == 2 ? q.w[r] * q.pref : 1.0), G);
      R g1[K::NM + 1];
#pragma unroll
      for (int n = 0; n <= K::NM; ++n) {
        R s = R(0);
#pragma unroll
        for (int m = 0; m <= K::MM; ++m) s += wk[x][m] * G[n][m];
        g1[n] = s;
      }
      hrr_bra1<LA, LB, R>(g1, AB[x], h[x]);
    }
#pragma unroll
    for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < K::NB; ++ib)
        acc[ia][ib] += h[0][cart_comp(LA, ia, 0)][cart_comp(LB, ib, 0)] *
                       h[1][cart_comp(LA, ia, 1)][cart_comp(LB, ib, 1)] *
                       h[2][cart_comp(LA, ia, 2)][cart_comp(LB, ib, 2)];
  }
}

__device__ __forceinline__ void load_geom(const ShellXYZ& a, const ShellXYZ& b, const ShellXYZ& c,
                                          const ShellXYZ& d, QuartetGeom& g, double AB[3], double CD[3]) {
  g.A[0] = a.x; g.A[1] = a.y; g.A[2] = a.z;
  g.B[0] = b.x; g.B[1] = b.y; g.B[2] = b.z;
  g.C[0] = c.x; g.C[1] = c.y; g.C[2] = c.z;
  g.D[0] = d.x; g.D[1] = d.y; g.D[2] = d.z;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    AB[x] = g.A[x] - g.B[x];
    CD[x] = g.C[x] - g.D[x];
  }
}

// ----------------------------------------------------------------------------
// Digestion adds.  FP64 form: atomic adds (LDS or device scope).
// Deterministic form: the value, already multiplied by its molecule's
// power-of-two scale, is rounded to a 64-bit integer and added as one;
// integer sums do not depend on the order of the additions, so every build
// of the same density gives the same bits.  Buffers hold the integers until
// a conversion pass turns them back into FP64.
// ----------------------------------------------------------------------------
template <bool DET>
__device__ __forceinline__ void acc_add(double* p, double v) {
  if constexpr (DET)
    atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(__double2ll_rn(v)));
  else
    atomicAdd(p, v);
}

// workgroup-scope add into the LDS ket window
template <bool DET>
__device__ __forceinline__ void acc_add_wg(double* p, double v) {
  if constexpr (DET)
    __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(p),
                           static_cast<unsigned long long>(__double2ll_rn(v)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  else
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// an accumulated value (FP64, or the deterministic form's integer) is zero
template <bool DET>
__device__ __forceinline__ bool acc_zero(const double* p) {
  if constexpr (DET)
    return *reinterpret_cast<const unsigned long long*>(p) == 0ull;
  else
    return *p == 0.0;
}

// adds an accumulated value into global memory with a device-scope atomic
template <bool DET>
__device__ __forceinline__ void acc_flush(double* dst, const double* src) {
  if constexpr (DET)
    atomicAdd(reinterpret_cast<unsigned long long*>(dst), *reinterpret_cast<const unsigned long long*>(src));
  else
    atomicAdd(dst, *src);
}

// ----------------------------------------------------------------------------
// J/K kernels
// ----------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, bool DET = false>
__global__ __launch_bounds__(256) void jk_full_kernel(JKArgs a) {
  using K = Class<LA, LB, LC, LD>;
  const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= a.nitems) return;
  int64_t g;
  int s0, S;
  map_item(a, item, g, s0, S);
  const int2 qk = a.quartets[g];
  const Pair bp = a.pairs[qk.x];
  const Pair kp = a.pairs[qk.y];
  const ShellXYZ sA = a.shells[bp.sa], sB = a.shells[bp.sb], sC = a.shells[kp.sa], sD = a.shells[kp.sb];
  QuartetGeom geo;
  double AB[3], CD[3];
  load_geom(sA, sB, sC, sD, geo, AB, CD);
  double acc[K::NA][K::NB][K::NC][K::ND];
#pragma unroll
  for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < K::NB; ++ib)
#pragma unroll
      for (int ic = 0; ic < K::NC; ++ic)
#pragma unroll
        for (int id = 0; id < K::ND; ++id) acc[ia][ib][ic][id] = 0.0;
  for (int ip = s0; ip < bp.np; ip += S) {
    const PrimPair P = a.prims[bp.poff + ip];
    for (int kq = 0; kq < kp.np; ++kq) {
      const PrimPair Q = a.prims[kp.poff + kq];
      PrimQ<K::NR> q;
      prim_setup<K::NR>(P, Q, geo, q);
      accumulate_full<LA, LB, LC, LD>(q, AB, CD, acc);
    }
  }
  double f = DET ? a.det_scale[2 * sA.mol] : 1.0;
  if (bp.sa == bp.sb) f *= 0.5;
  if (kp.sa == kp.sb) f *= 0.5;
  if (qk.x == qk.y) f *= 0.5;
  const MolInfo mi = a.mols[sA.mol];
  const int n = mi.n;
  const int oa = sA.ao, ob = sB.ao, oc = sC.ao, od = sD.ao;
  for (int bt = 0; bt < a.nbd; ++bt) {
    const double* __restrict__ Dm = a.dens + (size_t)bt * a.bstride + mi.base;
    if (a.with_j && bt < a.nbj) {
      double* J = a.jp + (size_t)bt * a.bstride + mi.base;
      double dcd[K::NC][K::ND], dab[K::NA][K::NB];
#pragma unroll
      for (int ic = 0; ic < K::NC; ++ic)
#pragma unroll
        for (int id = 0; id < K::ND; ++id)
          dcd[ic][id] = Dm[(size_t)(oc + ic) * n + od + id] + Dm[(size_t)(od + id) * n + oc + ic];
#pragma unroll
      for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < K::NB; ++ib)
          dab[ia][ib] = Dm[(size_t)(oa + ia) * n + ob + ib] + Dm[(size_t)(ob + ib) * n + oa + ia];
#pragma unroll
      for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < K::NB; ++ib) {
          double s = 0.0;
#pragma unroll
          for (int ic = 0; ic < K::NC; ++ic)
#pragma unroll
            for (int id = 0; id < K::ND; ++id) s += acc[ia][ib][ic][id] * dcd[ic][id];
          acc_add<DET>(J + (size_t)(oa + ia) * n + ob + ib, f * s);
        }
#pragma unroll
      for (int ic = 0; ic < K::NC; ++ic)
#pragma unroll
        for (int id = 0; id < K::ND; ++id) {
          double s = 0.0;
#pragma unroll
          for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < K::NB; ++ib) s += acc[ia][ib][ic][id] * dab[ia][ib];
          acc_add<DET>(J + (size_t)(oc + ic) * n + od + id, f * s);
        }
    }
    if (a.with_k) {
      double* Kp = a.kp + (size_t)bt * a.bstride + mi.base;
      double dbd[K::NB][K::ND], dbc[K::NB][K::NC], dad[K::NA][K::ND], dac[K::NA][K::NC];
#pragma unroll
      for (int ib = 0; ib < K::NB; ++ib) {
#pragma unroll
        for (int id = 0; id < K::ND; ++id) dbd[ib][id] = Dm[(size_t)(ob + ib) * n + od + id];
#pragma unroll
        for (int ic = 0; ic < K::NC; ++ic) dbc[ib][ic] = Dm[(size_t)(ob + ib) * n + oc + ic];
      }
#pragma unroll
      for (int ia = 0; ia < K::NA; ++ia) {
#pragma unroll
        for (int id = 0; id < K::ND; ++id) dad[ia][id] = Dm[(size_t)(oa + ia) * n + od + id];
#pragma unroll
        for (int ic = 0; ic < K::NC; ++ic) dac[ia][ic] = Dm[(size_t)(oa + ia) * n + oc + ic];
      }
#pragma unroll
      for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
        for (int ic = 0; ic < K::NC; ++ic) {
          double s = 0.0;
#pragma unroll
          for (int ib = 0; ib < K::NB; ++ib)
#pragma unroll
            for (int id = 0; id < K::ND; ++id) s += acc[ia][ib][ic][id] * dbd[ib][id];
          acc_add<DET>(Kp + (size_t)(oa + ia) * n + oc + ic, f * s);
        }
#pragma unroll
      for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
        for (int id = 0; id < K::ND; ++id) {
          double s = 0.0;
#pragma unroll
          for (int ib = 0; ib < K::NB; ++ib)
#pragma unroll
            for (int ic = 0; ic < K::NC; ++ic) s += acc[ia][ib][ic][id] * dbc[ib][ic];
          acc_add<DET>(Kp + (size_t)(oa + ia) * n + od + id, f * s);
        }
#pragma unroll
      for (int ib = 0; ib < K::NB; ++ib)
#pragma unroll
        for (int ic = 0; ic < K::NC; ++ic) {
          double s = 0.0;
#pragma unroll
          for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
            for (int id = 0; id < K::ND; ++id) s += acc[ia][ib][ic][id] * dad[ia][id];
          acc_add<DET>(Kp + (size_t)(ob + ib) * n + oc + ic, f * s);
        }
#pragma unroll
      for (int ib = 0; ib < K::NB; ++ib)
#pragma unroll
        for (int id = 0; id < K::ND; ++id) {
          double s = 0.0;
#pragma unroll
          for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
            for (int ic = 0; ic < K::NC; ++ic) s += acc[ia][ib][ic][id] * dac[ia][ic];
          acc_add<DET>(Kp + (size_t)(ob + ib) * n + od + id, f * s);
        }
    }
  }
}

template <int LA, int LB, int LC, int LD, bool DET = false>
__global__ __launch_bounds__(256) void jk_split_kernel(JKArgs a) {
  using K = Class<LA, LB, LC, LD>;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t item = t / K::SPLIT;
  if (item >= a.nitems) return;
  const int kc = (int)(t - item * K::SPLIT);
  const int ic = kc / K::ND, id = kc - (kc / K::ND) * K::ND;
  int64_t g;
  int s0, S;
  map_item(a, item, g, s0, S);
  const int2 qk = a.quartets[g];
  const Pair bp = a.pairs[qk.x];
  const Pair kp = a.pairs[qk.y];
  const ShellXYZ sA = a.shells[bp.sa], sB = a.shells[bp.sb], sC = a.shells[kp.sa], sD = a.shells[kp.sb];
  QuartetGeom geo;
  double AB[3], CD[3];
  load_geom(sA, sB, sC, sD, geo, AB, CD);
  int ec[3], ed[3];
  cart_xyz<LC>(ic, ec);
  cart_xyz<LD>(id, ed);
  double wk[3][K::MM + 1];
#pragma unroll
  for (int x = 0; x < 3; ++x) ket_weights<K::MM>(ec[x], ed[x], CD[x], wk[x]);
  double acc[K::NA][K::NB];
#pragma unroll
  for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < K::NB; ++ib) acc[ia][ib] = 0.0;
  for (int ip = s0; ip < bp.np; ip += S) {
    const PrimPair P = a.prims[bp.poff + ip];
    for (int kq = 0; kq < kp.np; ++kq) {
      const PrimPair Q = a.prims[kp.poff + kq];
      PrimQ<K::NR> q;
      prim_setup<K::NR>(P, Q, geo, q);
      accumulate_split<LA, LB, LC, LD>(q, AB, wk, acc);
    }
  }
  double f = DET ? a.det_scale[2 * sA.mol] : 1.0;
  if (bp.sa == bp.sb) f *= 0.5;
  if (kp.sa == kp.sb) f *= 0.5;
  if (qk.x == qk.y) f *= 0.5;
  const MolInfo mi = a.mols[sA.mol];
  const int n = mi.n;
  const int oa = sA.ao, ob = sB.ao;
  const int c = sC.ao + ic, d = sD.ao + id;
  for (int bt = 0; bt < a.nbd; ++bt) {
    const double* __restrict__ Dm = a.dens + (size_t)bt * a.bstride + mi.base;
    if (a.with_j && bt < a.nbj) {
      double* J = a.jp + (size_t)bt * a.bstride + mi.base;
      const double dcd = Dm[(size_t)c * n + d] + Dm[(size_t)d * n + c];
      double s2 = 0.0;
#pragma unroll
      for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < K::NB; ++ib) {
          acc_add<DET>(J + (size_t)(oa + ia) * n + ob + ib, f * acc[ia][ib] * dcd);
          s2 += acc[ia][ib] * (Dm[(size_t)(oa + ia) * n + ob + ib] + Dm[(size_t)(ob + ib) * n + oa + ia]);
        }
      acc_add<DET>(J + (size_t)c * n + d, f * s2);
    }
    if (a.with_k) {
      double* Kp = a.kp + (size_t)bt * a.bstride + mi.base;
      double db_d[K::NB], db_c[K::NB];
#pragma unroll
      for (int ib = 0; ib < K::NB; ++ib) {
        db_d[ib] = Dm[(size_t)(ob + ib) * n + d];
        db_c[ib] = Dm[(size_t)(ob + ib) * n + c];
      }
#pragma unroll
      for (int ia = 0; ia < K::NA; ++ia) {
        double sc = 0.0, sd = 0.0;
#pragma unroll
        for (int ib = 0; ib < K::NB; ++ib) {
          sc += acc[ia][ib] * db_d[ib];
          sd += acc[ia][ib] * db_c[ib];
        }
        acc_add<DET>(Kp + (size_t)(oa + ia) * n + c, f * sc);
        acc_add<DET>(Kp + (size_t)(oa + ia) * n + d, f * sd);
      }
      double da_d[K::NA], da_c[K::NA];
#pragma unroll
      for (int ia = 0; ia < K::NA; ++ia) {
        da_d[ia] = Dm[(size_t)(oa + ia) * n + d];
        da_c[ia] = Dm[(size_t)(oa + ia) * n + c];
      }
#pragma unroll
      for (int ib = 0; ib < K::NB; ++ib) {
        double sc = 0.0, sd = 0.0;
#pragma unroll
        for (int ia = 0; ia < K::NA; ++ia) {
          sc += acc[ia][ib] * da_d[ia];
          sd += acc[ia][ib] * da_c[ia];
        }
        acc_add<DET>(Kp + (size_t)(ob + ib) * n + c, f * sc);
        acc_add<DET>(Kp + (size_t)(ob + ib) * n + d, f * sd);
      }
    }
  }
}

// A value every lane of the wavefront holds alike, moved to scalar registers.
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Element idx (unsigned, < 2^29) of a chunk-uniform base through a 32-bit
// byte offset, which lets the load use the scalar-base + vector-offset form.
__device__ __forceinline__ double ld_u(const double* __restrict__ base, unsigned idx) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + (idx << 3));
}

// Sum of v[0..N) over the S adjacent lanes of a lane group (S a power of two
// <= 64, groups aligned to S): butterfly exchange, every lane ends with the sum.
template <int N, typename R>
__device__ __forceinline__ void group_sum(R* v, int S) {
  for (int m = 1; m < S; m <<= 1) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += __shfl_xor(v[i], m);
  }
}

// ----------------------------------------------------------------------------
// Column layout of a quartet class on a contracted pair set: shell X carries
// CX contraction columns and its "wide" block WX = CX * ncart(lX) components
// in AO order (columns contiguous), so digestion addresses AO oX + wide index
// and the exchange slabs, bra J blocks and ket J windows simply use the wide
// widths.  CX = 1 everywhere is the segmented (one-column) case.
template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD>
struct Wide {
  using K = Class<LA, LB, LC, LD>;
  static constexpr int WA = CA * K::NA, WB = CB * K::NB, WC = CC * K::NC, WD = CD * K::ND;
  static constexpr bool GEN = CA * CB * CC * CD > 1;
  static constexpr int NBRA = WA * WB, NKET = WC * WD;
  static constexpr bool kSplit = NBRA * NKET > 36;
  static constexpr int SPLIT = kSplit ? NKET : 1;
};

// One work item of the bra-major kernel in real type R: evaluate the quartet
// (or its wide ket-component slice) over bra primitive pairs s0, s0+S, ... and
// digest it (exchange rows / bra J block into LDS, ket J block into the LDS
// window or to global).  With general contraction each primitive quartet is
// evaluated once, coefficient-free, and added to every column combination
// with its coefficient product c_A c_B c_C c_D.
// MODE 1 also stores the block (after the lane-group sum) at cq, MODE 2 reads
// it from cq instead of evaluating it (in-core integrals).
// DET: deterministic digestion (integer adds of values scaled by dscale).
template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD, bool MOMENTS, typename R, int MODE = 0,
          bool DET = false>
__device__ __forceinline__ void bra_item(const BraArgs& a, const MolInfo& mi, const Pair& bp, const Pair& kp,
                                         const int2 qk, int slot, int s0, int S, int kc, double* __restrict__ krow,
                                         double* __restrict__ jrow, double* __restrict__ jket, int kwin0,
                                         const double* __restrict__ tab, double* __restrict__ cq,
                                         double dscale = 1.0, const double* __restrict__ bpl = nullptr,
                                         const double* __restrict__ dabs = nullptr) {
  using K = Class<LA, LB, LC, LD>;
  using W = Wide<LA, LB, LC, LD, CA, CB, CC, CD>;
  constexpr int WA = W::WA, WB = W::WB, WC = W::WC, WD = W::WD;
  constexpr int ROWS = WA + WB;
  const int ns = a.n;  // LDS row length
  const ShellXYZ sA = a.shells[bp.sa], sB = a.shells[bp.sb], sC = a.shells[kp.sa], sD = a.shells[kp.sb];
  // digestion addresses: molecule bases are chunk-uniform (scalar), element
  // offsets 32-bit (n <= 16384 Cartesian AOs per molecule, checked at plan
  // creation)
  const unsigned n = (unsigned)mi.n;
  // bra J copy of this lane: the digesting lanes are the group leaders
  // (multiples of S), so consecutive groups take consecutive copies
  const int jcopy = (threadIdx.x >> (__ffs(S) - 1)) & (kJBraCopies - 1);
  QuartetGeom geo;
  double ABd[3], CDd[3];
  load_geom(sA, sB, sC, sD, geo, ABd, CDd);
  const R AB[3] = {(R)ABd[0], (R)ABd[1], (R)ABd[2]};
  const R CD_[3] = {(R)CDd[0], (R)CDd[1], (R)CDd[2]};
  double f = DET ? dscale : 1.0;
  if (bp.sa == bp.sb) f *= 0.5;
  if (kp.sa == kp.sb) f *= 0.5;
  if (qk.x == qk.y) f *= 0.5;
  const int oa = sA.ao, ob = sB.ao;
  if constexpr (!W::kSplit) {
    R acc[WA][WB][WC][WD];
    if constexpr (MODE == 2) {
#pragma unroll
      for (int e = 0; e < WA * WB * WC * WD; ++e) (&acc[0][0][0][0])[e] = cq[e];
    } else {
#pragma unroll
    for (int ia = 0; ia < WA; ++ia)
#pragma unroll
      for (int ib = 0; ib < WB; ++ib)
#pragma unroll
        for (int ic = 0; ic < WC; ++ic)
#pragma unroll
          for (int id = 0; id < WD; ++id) acc[ia][ib][ic][id] = R(0);
    const bool staged = bpl != nullptr;
    // staged bra primitives: ket primitives outermost (one memory read of
    // each per lane), the chunk's bra primitives from LDS inside
    for (int kq = 0; staged && kq < kp.np; ++kq) {
      const PrimPair Q = a.prims[kp.poff + kq];
      // ket primitive relative to the bra's first centre and its own
      const double QA[3] = {Q.px - geo.A[0], Q.py - geo.A[1], Q.pz - geo.A[2]};
      const double QC[3] = {Q.px - geo.C[0], Q.py - geo.C[1], Q.pz - geo.C[2]};
      if constexpr (!W::GEN) {
        for (int ip = s0; ip < bp.np; ip += S)
          staged_quartet<LA, LB, LC, LD, MOMENTS, R>(bpl + ip * kStagedPrim, Q, QA, QC, AB, CD_, tab, acc);
      } else {
        // general contraction: coefficient-free blocks summed over the bra
        // primitives with the bra column products (record: ..., c_A[2],
        // c_B[2]), then into every ket column with the ket's products
        constexpr int REC = kStagedPrim + (CA * CB > 1 ? 4 : 0);
        R tmp[WA][WB][K::NC][K::ND];
#pragma unroll
        for (int e = 0; e < WA * WB * K::NC * K::ND; ++e) (&tmp[0][0][0][0])[e] = R(0);
        for (int ip = s0; ip < bp.np; ip += S) {
          const double* __restrict__ r = bpl + ip * REC;
          R blk[K::NA][K::NB][K::NC][K::ND];
#pragma unroll
          for (int e = 0; e < K::NA * K::NB * K::NC * K::ND; ++e) (&blk[0][0][0][0])[e] = R(0);
          staged_quartet<LA, LB, LC, LD, MOMENTS, R>(r, Q, QA, QC, AB, CD_, tab, blk);
#pragma unroll
          for (int x = 0; x < CA; ++x)
#pragma unroll
            for (int y = 0; y < CB; ++y) {
              const R cf = CA * CB > 1 ? (R)(r[kStagedPrim + x] * r[kStagedPrim + 2 + y]) : R(1);
#pragma unroll
              for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
                for (int ib = 0; ib < K::NB; ++ib)
#pragma unroll
                  for (int ic = 0; ic < K::NC; ++ic)
#pragma unroll
                    for (int id = 0; id < K::ND; ++id)
                      tmp[x * K::NA + ia][y * K::NB + ib][ic][id] += cf * blk[ia][ib][ic][id];
            }
        }
        PrimCoef cQ = {{1.0, 0.0}, {1.0, 0.0}};
        if constexpr (CC * CD > 1) cQ = a.coef[kp.poff + kq];
#pragma unroll
        for (int z = 0; z < CC; ++z)
#pragma unroll
          for (int w = 0; w < CD; ++w) {
            const R cf = (R)(cQ.ca[z] * cQ.cb[w]);
#pragma unroll
            for (int ia = 0; ia < WA; ++ia)
#pragma unroll
              for (int ib = 0; ib < WB; ++ib)
#pragma unroll
                for (int ic = 0; ic < K::NC; ++ic)
#pragma unroll
                  for (int id = 0; id < K::ND; ++id)
                    acc[ia][ib][z * K::NC + ic][w * K::ND + id] += cf * tmp[ia][ib][ic][id];
          }
      }
    }
    if (!staged)
    for (int ip = s0; ip < bp.np; ip += S) {
      const PrimPair P = a.prims[bp.poff + ip];
      // a side with one column already carries its coefficients in e
      PrimCoef cP = {{1.0, 0.0}, {1.0, 0.0}};
      if constexpr (CA * CB > 1) cP = a.coef[bp.poff + ip];
      const PrimPair* __restrict__ kpr = a.prims + kp.poff;
      for (int kq = 0; kq < kp.np; ++kq) {
        const PrimPair Q = kpr[kq];
        PrimQ<K::NR> q;
        prim_setup<K::NR, MOMENTS>(P, Q, geo, q, tab);
        if constexpr (!W::GEN) {
          accumulate_full<LA, LB, LC, LD, R>(q, AB, CD_, acc);
        } else {
          PrimCoef cQ = {{1.0, 0.0}, {1.0, 0.0}};
          if constexpr (CC * CD > 1) cQ = a.coef[kp.poff + kq];
          R blk[K::NA][K::NB][K::NC][K::ND];
#pragma unroll
          for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < K::NB; ++ib)
#pragma unroll
              for (int ic = 0; ic < K::NC; ++ic)
#pragma unroll
                for (int id = 0; id < K::ND; ++id) blk[ia][ib][ic][id] = R(0);
          accumulate_full<LA, LB, LC, LD, R>(q, AB, CD_, blk);
#pragma unroll
          for (int x = 0; x < CA; ++x)
#pragma unroll
            for (int y = 0; y < CB; ++y)
#pragma unroll
              for (int z = 0; z < CC; ++z)
#pragma unroll
                for (int w = 0; w < CD; ++w) {
                  const R cf = (R)(cP.ca[x] * cP.cb[y] * cQ.ca[z] * cQ.cb[w]);
#pragma unroll
                  for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
                    for (int ib = 0; ib < K::NB; ++ib)
#pragma unroll
                      for (int ic = 0; ic < K::NC; ++ic)
#pragma unroll
                        for (int id = 0; id < K::ND; ++id)
                          acc[x * K::NA + ia][y * K::NB + ib][z * K::NC + ic][w * K::ND + id] +=
                              cf * blk[ia][ib][ic][id];
                }
        }
      }
    }
    group_sum<WA * WB * WC * WD>(&acc[0][0][0][0], S);
    if (s0 != 0) return;
    if constexpr (MODE == 1) {
#pragma unroll
      for (int e = 0; e < WA * WB * WC * WD; ++e) cq[e] = (double)(&acc[0][0][0][0])[e];
    }
    }
    const unsigned oc = sC.ao, od = sD.ao;
    // LDS exchange-row columns of C and D (compact over the class's ket mask)
    const unsigned kc0 = a.shell_cofs ? (unsigned)a.shell_cofs[kp.sa * 8 + a.cmask] : oc;
    const unsigned kd0 = a.shell_cofs ? (unsigned)a.shell_cofs[kp.sb * 8 + a.cmask] : od;
    const unsigned ra = (unsigned)oa * n, rb = (unsigned)ob * n, rc = oc * n, rd = od * n;
    for (int bt = 0; bt < a.nbd; ++bt) {
      const double* __restrict__ Dm = a.dens + (size_t)bt * a.bstride + mi.base;
      if (a.with_j && bt < a.nbj) {
        double* jr = jrow + ((bt * a.nbra_max + slot) * kJBraCopies + jcopy) * W::NBRA;
        double dcd[WC][WD], dab[WA][WB];
#pragma unroll
        for (int ic = 0; ic < WC; ++ic)
#pragma unroll
          for (int id = 0; id < WD; ++id)
            dcd[ic][id] = ld_u(Dm, rc + ic * n + od + id) + ld_u(Dm, rd + id * n + oc + ic);
        // D_ab + D_ba of the bra block: staged per slot in LDS
        const double* __restrict__ dsl = dabs + (bt * a.nbra_max + slot) * W::NBRA;
#pragma unroll
        for (int ia = 0; ia < WA; ++ia)
#pragma unroll
          for (int ib = 0; ib < WB; ++ib) dab[ia][ib] = dsl[ia * WB + ib];
#pragma unroll
        for (int ia = 0; ia < WA; ++ia)
#pragma unroll
          for (int ib = 0; ib < WB; ++ib) {
            double sj = 0.0;
#pragma unroll
            for (int ic = 0; ic < WC; ++ic)
#pragma unroll
              for (int id = 0; id < WD; ++id) sj += (double)acc[ia][ib][ic][id] * dcd[ic][id];
            acc_add<DET>(jr + ia * WB + ib, f * sj);
          }
#pragma unroll
        for (int ic = 0; ic < WC; ++ic)
#pragma unroll
          for (int id = 0; id < WD; ++id) {
            double sj = 0.0;
#pragma unroll
            for (int ia = 0; ia < WA; ++ia)
#pragma unroll
              for (int ib = 0; ib < WB; ++ib) sj += (double)acc[ia][ib][ic][id] * dab[ia][ib];
            // (workgroup scope for the LDS window keeps the two atomics from
            // being merged into one generic-address atomic)
            if (jket)
              acc_add_wg<DET>(jket + (bt * a.nket_max + (qk.y - kwin0)) * W::NKET + ic * WD + id, f * sj);
            else
              acc_add<DET>(a.jket_pad + 8 * ((size_t)bt * a.bstride + mi.base) + 8u * (rc + ic * n + od + id), f * sj);
          }
      }
      if (a.with_k) {
        double* kr = krow + (bt * a.nbra_max + slot) * ROWS * ns;
        double dbd[WB][WD], dbc[WB][WC], dad[WA][WD], dac[WA][WC];
#pragma unroll
        for (int ib = 0; ib < WB; ++ib) {
#pragma unroll
          for (int id = 0; id < WD; ++id) dbd[ib][id] = ld_u(Dm, rb + ib * n + od + id);
#pragma unroll
          for (int ic = 0; ic < WC; ++ic) dbc[ib][ic] = ld_u(Dm, rb + ib * n + oc + ic);
        }
#pragma unroll
        for (int ia = 0; ia < WA; ++ia) {
#pragma unroll
          for (int id = 0; id < WD; ++id) dad[ia][id] = ld_u(Dm, ra + ia * n + od + id);
#pragma unroll
          for (int ic = 0; ic < WC; ++ic) dac[ia][ic] = ld_u(Dm, ra + ia * n + oc + ic);
        }
#pragma unroll
        for (int ia = 0; ia < WA; ++ia) {
#pragma unroll
          for (int ic = 0; ic < WC; ++ic) {
            double sk = 0.0;
#pragma unroll
            for (int ib = 0; ib < WB; ++ib)
#pragma unroll
              for (int id = 0; id < WD; ++id) sk += (double)acc[ia][ib][ic][id] * dbd[ib][id];
            acc_add<DET>(kr + ia * ns + kc0 + ic, f * sk);
          }
#pragma unroll
          for (int id = 0; id < WD; ++id) {
            double sk = 0.0;
#pragma unroll
            for (int ib = 0; ib < WB; ++ib)
#pragma unroll
              for (int ic = 0; ic < WC; ++ic) sk += (double)acc[ia][ib][ic][id] * dbc[ib][ic];
            acc_add<DET>(kr + ia * ns + kd0 + id, f * sk);
          }
        }
#pragma unroll
        for (int ib = 0; ib < WB; ++ib) {
#pragma unroll
          for (int ic = 0; ic < WC; ++ic) {
            double sk = 0.0;
#pragma unroll
            for (int ia = 0; ia < WA; ++ia)
#pragma unroll
              for (int id = 0; id < WD; ++id) sk += (double)acc[ia][ib][ic][id] * dad[ia][id];
            acc_add<DET>(kr + (WA + ib) * ns + kc0 + ic, f * sk);
          }
#pragma unroll
          for (int id = 0; id < WD; ++id) {
            double sk = 0.0;
#pragma unroll
            for (int ia = 0; ia < WA; ++ia)
#pragma unroll
              for (int ic = 0; ic < WC; ++ic) sk += (double)acc[ia][ib][ic][id] * dac[ia][ic];
            acc_add<DET>(kr + (WA + ib) * ns + kd0 + id, f * sk);
          }
        }
      }
    }
  } else {
    // this lane: wide ket component kc = (column z, component ic) x (column w, component id)
    const int icw = kc / WD, idw = kc - (kc / WD) * WD;
    const int zc = icw / K::NC, ic = icw - zc * K::NC;
    const int wcol = idw / K::ND, id = idw - wcol * K::ND;
    int ec[3], ed[3];
    cart_xyz<LC>(ic, ec);
    cart_xyz<LD>(id, ed);
    R wk[3][K::MM + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) ket_weights<K::MM, R>(ec[x], ed[x], CD_[x], wk[x]);
    R acc[WA][WB];
    if constexpr (MODE == 2) {
#pragma unroll
      for (int e = 0; e < WA * WB; ++e) (&acc[0][0])[e] = cq[e];
    } else {
#pragma unroll
    for (int ia = 0; ia < WA; ++ia)
#pragma unroll
      for (int ib = 0; ib < WB; ++ib) acc[ia][ib] = R(0);
    for (int ip = s0; ip < bp.np; ip += S) {
      const PrimPair P = a.prims[bp.poff + ip];
      PrimCoef cP = {{1.0, 0.0}, {1.0, 0.0}};
      if constexpr (CA * CB > 1) cP = a.coef[bp.poff + ip];
      const PrimPair* __restrict__ kpr = a.prims + kp.poff;
      for (int kq = 0; kq < kp.np; ++kq) {
        const PrimPair Q = kpr[kq];
        PrimQ<K::NR> q;
        prim_setup<K::NR, MOMENTS>(P, Q, geo, q, tab);
        if constexpr (!W::GEN) {
          accumulate_split<LA, LB, LC, LD, R>(q, AB, wk, acc);
        } else {
          PrimCoef cQ = {{1.0, 0.0}, {1.0, 0.0}};
          if constexpr (CC * CD > 1) cQ = a.coef[kp.poff + kq];
          R blk[K::NA][K::NB];
#pragma unroll
          for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < K::NB; ++ib) blk[ia][ib] = R(0);
          accumulate_split<LA, LB, LC, LD, R>(q, AB, wk, blk);
          const double ck = cQ.ca[zc] * cQ.cb[wcol];
#pragma unroll
          for (int x = 0; x < CA; ++x)
#pragma unroll
            for (int y = 0; y < CB; ++y) {
              const R cf = (R)(cP.ca[x] * cP.cb[y] * ck);
#pragma unroll
              for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
                for (int ib = 0; ib < K::NB; ++ib) acc[x * K::NA + ia][y * K::NB + ib] += cf * blk[ia][ib];
            }
        }
      }
    }
    group_sum<WA * WB>(&acc[0][0], S);
    if (s0 != 0) return;
    if constexpr (MODE == 1) {
#pragma unroll
      for (int e = 0; e < WA * WB; ++e) cq[e] = (double)(&acc[0][0])[e];
    }
    }
    const unsigned c = sC.ao + icw, d = sD.ao + idw;
    // LDS exchange-row columns of this lane's C and D components
    const unsigned xc = (a.shell_cofs ? (unsigned)a.shell_cofs[kp.sa * 8 + a.cmask] : sC.ao) + icw;
    const unsigned xd = (a.shell_cofs ? (unsigned)a.shell_cofs[kp.sb * 8 + a.cmask] : sD.ao) + idw;
    const unsigned ra = (unsigned)oa * n, rb = (unsigned)ob * n;
    for (int bt = 0; bt < a.nbd; ++bt) {
      const double* __restrict__ Dm = a.dens + (size_t)bt * a.bstride + mi.base;
      if (a.with_j && bt < a.nbj) {
        double* jr = jrow + ((bt * a.nbra_max + slot) * kJBraCopies + jcopy) * W::NBRA;
        const double dcd = ld_u(Dm, c * n + d) + ld_u(Dm, d * n + c);
        // D_ab + D_ba of the bra block: staged per slot in LDS
        const double* __restrict__ dsl = dabs + (bt * a.nbra_max + slot) * W::NBRA;
        double s2 = 0.0;
#pragma unroll
        for (int ia = 0; ia < WA; ++ia)
#pragma unroll
          for (int ib = 0; ib < WB; ++ib) {
            acc_add<DET>(jr + ia * WB + ib, f * (double)acc[ia][ib] * dcd);
            s2 += (double)acc[ia][ib] * dsl[ia * WB + ib];
          }
        if (jket)
          acc_add_wg<DET>(jket + (bt * a.nket_max + (qk.y - kwin0)) * W::NKET + kc, f * s2);
        else
          acc_add<DET>(a.jket_pad + 8 * ((size_t)bt * a.bstride + mi.base) + 8u * (c * n + d), f * s2);
      }
      if (a.with_k) {
        double* kr = krow + (bt * a.nbra_max + slot) * ROWS * ns;
        double db_d[WB], db_c[WB];
#pragma unroll
        for (int ib = 0; ib < WB; ++ib) {
          db_d[ib] = ld_u(Dm, rb + ib * n + d);
          db_c[ib] = ld_u(Dm, rb + ib * n + c);
        }
#pragma unroll
        for (int ia = 0; ia < WA; ++ia) {
          double sc = 0.0, sd = 0.0;
#pragma unroll
          for (int ib = 0; ib < WB; ++ib) {
            sc += (double)acc[ia][ib] * db_d[ib];
            sd += (double)acc[ia][ib] * db_c[ib];
          }
          acc_add<DET>(kr + ia * ns + xc, f * sc);
          acc_add<DET>(kr + ia * ns + xd, f * sd);
        }
        double da_d[WA], da_c[WA];
#pragma unroll
        for (int ia = 0; ia < WA; ++ia) {
          da_d[ia] = ld_u(Dm, ra + ia * n + d);
          da_c[ia] = ld_u(Dm, ra + ia * n + c);
        }
#pragma unroll
        for (int ib = 0; ib < WB; ++ib) {
          double sc = 0.0, sd = 0.0;
#pragma unroll
          for (int ia = 0; ia < WA; ++ia) {
            sc += (double)acc[ia][ib] * da_d[ia];
            sd += (double)acc[ia][ib] * da_c[ia];
          }
          acc_add<DET>(kr + (WA + ib) * ns + xc, f * sc);
          acc_add<DET>(kr + (WA + ib) * ns + xd, f * sd);
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------
// Bra-major J/K kernel: one workgroup per chunk of quartets sharing up to
// nbra_max consecutive bra pairs.  All exchange terms of the 8-fold sum land in
// rows of the bra shells (Kp[a,c], Kp[a,d], Kp[b,c], Kp[b,d]), so they are
// accumulated with LDS atomics in per-bra row slabs [rows of A, rows of B] x n
// and flushed to global memory once per workgroup; the bra-side J block is
// accumulated in LDS the same way.  Only the ket-side J terms go straight to
// global atomics.  Global FP64 atomics are executed at the memory side on this
// platform, so their count, not FLOPs, is what this layout minimises.
// With a.mixed set, a quartet whose Q_ij Q_kl max|D| (max over the six density
// blocks it touches) is below a.thr may be evaluated in FP32 (roots, weights
// and prefactors stay FP64, digestion accumulates in FP64); the choice is made
// per wavefront (all lanes below the threshold), never per lane.
// ----------------------------------------------------------------------------
// With TLDS the Rys table slice of the class (rys_slice_size(NR) doubles) is
// copied to LDS after the slabs, so the per-lane interval gathers of the root
// evaluation read LDS instead of global memory.
// MODE (in-core integrals): 0 evaluates, 1 evaluates and stores the blocks,
// 2 digests stored blocks (one lane per quartet component: no primitive split).
// DET: deterministic digestion -- every LDS and global accumulator holds a
// 64-bit integer (value x the molecule's scale, a.det_scale[2 mol]) added by
// integer atomics, so the result does not depend on the order of the adds.
template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD, bool TLDS, bool MIX, int MODE = 0,
          bool DET = false>
__global__ __launch_bounds__(512) void jk_bra_kernel(BraArgs a) {
  static_assert(!(MIX && DET), "the deterministic form is FP64 only");
  using K = Class<LA, LB, LC, LD>;
  using W = Wide<LA, LB, LC, LD, CA, CB, CC, CD>;
  constexpr int ROWS = W::WA + W::WB;
  extern __shared__ double lds[];
  const int ns = a.n;  // LDS row length (widest molecule)
  const BraChunk ch = a.chunks[blockIdx.x];
  // every quartet of a chunk belongs to one molecule (runs never cross
  // molecules); its layout is workgroup-uniform, kept in scalar registers
  MolInfo mi;
  double dscale = 1.0;
  {
    const int mol = a.shells[a.pairs[ch.bra0].sa].mol;
    if constexpr (DET) dscale = a.det_scale[2 * mol];
    const MolInfo m = a.mols[mol];
    mi.base = uniform_i64(m.base);
    mi.dbase = uniform_i64(m.dbase);
    mi.n = __builtin_amdgcn_readfirstlane(m.n);
    mi.nseg = __builtin_amdgcn_readfirstlane(m.nseg);
  }
  const int krow_size = a.nbd * a.nbra_max * ROWS * ns;
  double* krow = lds;                      // [nbd][slot][ROWS][n]
  double* jrow = lds + krow_size;          // [nbj][slot][copy][NA][NB]
  const int jrow_size = a.nbj * a.nbra_max * kJBraCopies * W::NBRA;
  // ket-side J of the chunk's ket window: [nbj][nket_max][NC*ND]
  double* jket = (a.nket_max > 0 && ch.nket > 0) ? jrow + jrow_size : nullptr;
  const int jket_size = a.nbj * a.nket_max * W::NKET;
  const int zero_size = krow_size + jrow_size + (jket ? jket_size : 0);
  for (int i = threadIdx.x; i < zero_size; i += blockDim.x) lds[i] = 0.0;
  // D_ab + D_ba of every slot's bra block, per Coulomb density: read by the
  // bra-side and ket-side J digestion of every quartet of the slot
  const int dab_size = a.with_j ? a.nbj * a.nbra_max * W::NBRA : 0;
  double* dabs = lds + krow_size + jrow_size + jket_size;
  for (int i = threadIdx.x; i < dab_size; i += blockDim.x) {
    const int e = i % W::NBRA;
    const int slot = (i / W::NBRA) % a.nbra_max;
    const int bt = i / (W::NBRA * a.nbra_max);
    double v = 0.0;
    if (slot < ch.nbra) {
      const Pair bp = a.pairs[ch.bra0 + slot];
      const unsigned oa = a.shells[bp.sa].ao + e / W::WB, ob = a.shells[bp.sb].ao + e % W::WB;
      const double* __restrict__ Dm = a.dens + (size_t)bt * a.bstride + mi.base;
      v = Dm[(size_t)oa * mi.n + ob] + Dm[(size_t)ob * mi.n + oa];
    }
    dabs[i] = v;
  }
  // staged bra primitive pairs [records][kStagedPrim] and the first record
  // of every slot (full-block segmented classes that evaluate)
  constexpr bool kStage = !W::kSplit && MODE != 2;
  // record doubles: p, P - A, e, 1/p (+ the bra's column coefficients)
  constexpr int kRec = kStagedPrim + (CA * CB > 1 ? 4 : 0);
  const int stage_base = (krow_size + jrow_size + jket_size + dab_size + 1) & ~1;  // 16-byte aligned
  const int stage_size = kStage && a.bra_stage ? kRec * a.bra_stage + (a.nbra_max + 2) / 2 : 0;
  double* bstage = lds + stage_base;
  int* bslot = reinterpret_cast<int*>(bstage + kRec * a.bra_stage);
  if constexpr (kStage) {
    if (a.bra_stage) {
      if (threadIdx.x == 0) {
        int o = 0;
        for (int j = 0; j < ch.nbra; ++j) {
          bslot[j] = o;
          o += a.pairs[ch.bra0 + j].np;
        }
      }
      __syncthreads();
      for (int j = 0; j < ch.nbra; ++j) {
        const Pair bp = a.pairs[ch.bra0 + j];
        double* dst = bstage + kRec * bslot[j];
        for (int k = threadIdx.x; k < bp.np; k += blockDim.x) {
          const PrimPair P = a.prims[bp.poff + k];
          const ShellXYZ A = a.shells[bp.sa];
          double* rec = dst + kRec * k;
          rec[0] = P.p;
          rec[1] = P.px - A.x;
          rec[2] = P.py - A.y;
          rec[3] = P.pz - A.z;
          rec[4] = P.e;
          rec[5] = P.ip;
          if constexpr (CA * CB > 1) {
            const PrimCoef c = a.coef[bp.poff + k];
            rec[6] = c.ca[0];
            rec[7] = c.ca[1];
            rec[8] = c.cb[0];
            rec[9] = c.cb[1];
          }
        }
      }
    }
  }
  const double* tab = rys_tab + rys_table_offset[K::NR - 1];
  // one or two roots: the table staged in LDS, or the Boys-moment quadrature
  // (global Boys grid) when the table stays in global memory
  constexpr bool kMoments = K::NR <= 2 && !TLDS;
  if constexpr (TLDS && MODE != 2) {
    double* ltab = lds + ((stage_base + stage_size + 1) & ~1);  // 16-byte aligned
    for (int i = threadIdx.x; i < rys_slice_size(K::NR); i += blockDim.x) ltab[i] = tab[i];
    tab = ltab;
  }
  __syncthreads();
  // item -> (quartet, ket component kc, bra-primitive slice s0), s0 fastest:
  // the S slices of one quartet component occupy S adjacent lanes of one
  // wavefront (S is a power of two <= 64 and blockDim a multiple of 64), which
  // add their partial integrals by cross-lane exchange before the s0 = 0 lane
  // digests them alone.
  //
  // The chunk's quartets are visited from a workgroup-dependent starting
  // point: chunks of neighbouring bras share their kets, and workgroups that
  // walk the kets in the same order make their ket-side J atomics collide on
  // the same addresses at the same time (the memory-side atomic unit then
  // serialises them).
  const int S = MODE == 2 ? 1 : ch.S;
  const int64_t total = (int64_t)ch.nq * S * W::SPLIT;
  const int64_t rot = ((int64_t)blockIdx.x * 40503) % ch.nq;
  constexpr int kBlock = W::kSplit ? W::NBRA : W::NBRA * W::NKET;
  for (int64_t it = threadIdx.x; it < total; it += blockDim.x) {
    const int s0 = (int)(it & (S - 1));
    const int64_t rest = it / S;
    const int kc = (int)(rest % W::SPLIT);
    int64_t qr = rest / W::SPLIT + rot;
    if (qr >= ch.nq) qr -= ch.nq;
    const int64_t g = ch.q0 + qr;
    const int2 qk = a.quartets[g];
    const int slot = qk.x - ch.bra0;
    const Pair bp = a.pairs[qk.x];
    const Pair kp = a.pairs[qk.y];
    double* cq = nullptr;
    if constexpr (MODE != 0) cq = a.cache + ((size_t)g * W::SPLIT + kc) * kBlock;
    bool lowp = false;
    if constexpr (MIX) {
      const int la = a.shells[bp.sa].lseg, lb = a.shells[bp.sb].lseg;
      const int lc = a.shells[kp.sa].lseg, ld = a.shells[kp.sb].lseg;
      const double* dm = a.dmax + mi.dbase;
      const int nsg = mi.nseg;
      const double d6 = fmax(fmax(fmax(dm[la * nsg + lb], dm[lc * nsg + ld]),
                                  fmax(dm[la * nsg + lc], dm[la * nsg + ld])),
                             fmax(dm[lb * nsg + lc], dm[lb * nsg + ld]));
      lowp = a.pair_q[qk.x] * a.pair_q[qk.y] * d6 < a.thr;
      // wave-uniform: FP32 only when every active lane of the wavefront
      // qualifies, so a wave never executes both precisions (neighbouring
      // lanes hold Schwarz-adjacent quartets and mostly agree)
      lowp = __all(lowp);
    }
    if constexpr (MIX) {
      if (lowp) {
        bra_item<LA, LB, LC, LD, CA, CB, CC, CD, kMoments, float>(
            a, mi, bp, kp, qk, slot, s0, S, kc, krow, jrow, jket, ch.ket0, tab, cq, 1.0,
            (kStage && a.bra_stage) ? bstage + kRec * bslot[slot] : nullptr, dabs);
        continue;
      }
    }
    const double* bpl = (kStage && a.bra_stage) ? bstage + kRec * bslot[slot] : nullptr;
    bra_item<LA, LB, LC, LD, CA, CB, CC, CD, kMoments, double, MODE, DET>(a, mi, bp, kp, qk, slot, s0, S, kc, krow,
                                                                           jrow, jket, ch.ket0, tab, cq, dscale, bpl,
                                                                           dabs);
  }
  __syncthreads();
  // the chunk's molecule and the columns its exchange terms can reach (the
  // shells whose l is in the class's ket mask)
  const bool compact = a.shell_cofs != nullptr;
  const int2 cspace = ((a.part_mode & 3) || compact) ? a.cspace[a.shells[a.pairs[ch.bra0].sa].mol * 8 + a.cmask]
                                                     : make_int2(0, mi.n);
  if (a.part_mode & 1) {
    // exchange rows to the chunk's partial slab with plain stores, over the
    // mask's columns in compact order (zeros included: the reduction reads
    // every entry)
    const int nm = cspace.y;
    const int* __restrict__ cmap = a.colmap + cspace.x;
    const int cnt = a.nbd * ch.nbra * ROWS * nm;
    double* __restrict__ dst = a.part + ch.poff;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
      const int cc = i % nm;
      const int rs = i / nm;  // (bt, slot, row)
      const int r = rs % ROWS;
      const int slot = (rs / ROWS) % ch.nbra;
      const int bt = rs / (ROWS * ch.nbra);
      dst[(size_t)bt * a.part_stride + ((size_t)slot * ROWS + r) * nm + cc] =
          krow[((bt * a.nbra_max + slot) * ROWS + r) * ns + (compact ? cc : cmap[cc])];
    }
  } else
  // flush: exchange rows and bra J blocks of every slot (skip untouched zeros)
  for (int i = threadIdx.x; i < krow_size; i += blockDim.x) {
    if (acc_zero<DET>(krow + i)) continue;
    int col = i % ns;
    const int r = (i / ns) % ROWS;
    const int slot = (i / (ns * ROWS)) % a.nbra_max;
    const int bt = i / (ns * ROWS * a.nbra_max);
    if (slot >= ch.nbra) continue;
    if (compact) {
      if (col >= cspace.y) continue;
      col = a.colmap[cspace.x + col];
    }
    const Pair bp = a.pairs[ch.bra0 + slot];
    const MolInfo mi = a.mols[a.shells[bp.sa].mol];
    const int row = r < W::WA ? a.shells[bp.sa].ao + r : a.shells[bp.sb].ao + (r - W::WA);
    acc_flush<DET>(a.kp + (size_t)bt * a.bstride + mi.base + (size_t)row * mi.n + col, krow + i);
  }
  for (int i = threadIdx.x; i < jrow_size / kJBraCopies; i += blockDim.x) {
    const int e = i % W::NBRA;
    const int slot = (i / W::NBRA) % a.nbra_max;
    const int bt = i / (W::NBRA * a.nbra_max);
    if (slot >= ch.nbra) continue;
    double v = 0.0;
    if constexpr (DET) {
      // the copies' integers add with wrap-around, like the atomics that filled them
      unsigned long long u = 0;
#pragma unroll
      for (int c = 0; c < kJBraCopies; ++c)
        u += reinterpret_cast<const unsigned long long*>(jrow)[((bt * a.nbra_max + slot) * kJBraCopies + c) *
                                                                W::NBRA + e];
      v = __longlong_as_double((long long)u);
    } else {
#pragma unroll
      for (int c = 0; c < kJBraCopies; ++c) v += jrow[((bt * a.nbra_max + slot) * kJBraCopies + c) * W::NBRA + e];
    }
    if (acc_zero<DET>(&v)) continue;
    const Pair bp = a.pairs[ch.bra0 + slot];
    const MolInfo mi = a.mols[a.shells[bp.sa].mol];
    const int row = a.shells[bp.sa].ao + e / W::WB, col = a.shells[bp.sb].ao + e % W::WB;
    acc_flush<DET>(a.jp + (size_t)bt * a.bstride + mi.base + (size_t)row * mi.n + col, &v);
  }
  if (jket && (a.part_mode & 2)) {
    // ket window to the partial slab after the exchange rows
    const int nk = a.nbj * ch.nket * W::NKET;
    // (the chunk's exchange rows precede its window whether this build
    // writes them or not)
    double* __restrict__ dst = a.part + ch.poff + (size_t)ch.nbra * ROWS * cspace.y;
    for (int i = threadIdx.x; i < nk; i += blockDim.x) {
      const int ew = i % (ch.nket * W::NKET);  // (window entry, element)
      const int bt = i / (ch.nket * W::NKET);
      dst[(size_t)bt * a.part_stride + ew] = jket[(size_t)bt * a.nket_max * W::NKET + ew];
    }
  } else if (jket) {
    const int nk = a.nbj * ch.nket * W::NKET;
    for (int i = threadIdx.x; i < nk; i += blockDim.x) {
      const int e = i % W::NKET;
      const int w = (i / W::NKET) % ch.nket;
      const int bt = i / (W::NKET * ch.nket);
      const double* src = jket + ((size_t)bt * a.nket_max + w) * W::NKET + e;
      if (acc_zero<DET>(src)) continue;
      const Pair kp = a.pairs[ch.ket0 + w];
      const MolInfo mi = a.mols[a.shells[kp.sa].mol];
      const int row = a.shells[kp.sa].ao + e / W::WD, col = a.shells[kp.sb].ao + e % W::WD;
      acc_flush<DET>(a.jket_pad + 8 * ((size_t)bt * a.bstride + mi.base + (size_t)row * mi.n + col), src);
    }
  }
}

// ----------------------------------------------------------------------------
// Block kernel (explicit quartet list; one lane per (task, ket component pair))
// ----------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void blocks_kernel(BlockArgs a) {
  using K = Class<LA, LB, LC, LD>;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t g = t / K::NKET;
  if (g >= a.ntasks) return;
  const int kc = (int)(t - g * K::NKET);
  const int ic = kc / K::ND, id = kc - (kc / K::ND) * K::ND;
  const BlockTask task = a.tasks[g];
  const Pair bp = a.pairs[task.bra];
  const Pair kp = a.pairs[task.ket];
  const ShellXYZ sA = a.shells[bp.sa], sB = a.shells[bp.sb], sC = a.shells[kp.sa], sD = a.shells[kp.sb];
  QuartetGeom geo;
  double AB[3], CD[3];
  load_geom(sA, sB, sC, sD, geo, AB, CD);
  int ec[3], ed[3];
  cart_xyz<LC>(ic, ec);
  cart_xyz<LD>(id, ed);
  double wk[3][K::MM + 1];
#pragma unroll
  for (int x = 0; x < 3; ++x) ket_weights<K::MM>(ec[x], ed[x], CD[x], wk[x]);
  double acc[K::NA][K::NB];
#pragma unroll
  for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < K::NB; ++ib) acc[ia][ib] = 0.0;
  for (int ip = 0; ip < bp.np; ++ip) {
    const PrimPair P = a.prims[bp.poff + ip];
    for (int kq = 0; kq < kp.np; ++kq) {
      const PrimPair Q = a.prims[kp.poff + kq];
      PrimQ<K::NR> q;
      prim_setup<K::NR>(P, Q, geo, q);
      accumulate_split<LA, LB, LC, LD>(q, AB, wk, acc);
    }
  }
  const int64_t o = task.base + (int64_t)(task.coff[2] + ic) * task.stride[2] +
                    (int64_t)(task.coff[3] + id) * task.stride[3];
#pragma unroll
  for (int ia = 0; ia < K::NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < K::NB; ++ib)
      a.out[o + (int64_t)(task.coff[0] + ia) * task.stride[0] + (int64_t)(task.coff[1] + ib) * task.stride[1]] =
          acc[ia][ib];
}

// ----------------------------------------------------------------------------
// Launchers (instantiated per class in jk_class.hip)
// ----------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
void launch_jk_t(const JKArgs& a, hipStream_t s) {
  using K = Class<LA, LB, LC, LD>;
  if (a.nitems <= 0) return;
  const int64_t threads = a.nitems * K::SPLIT;
  const int64_t blocks = (threads + 255) / 256;
  if constexpr (K::kSplit) {
    if (a.det_scale)
      hipLaunchKernelGGL((jk_split_kernel<LA, LB, LC, LD, true>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((jk_split_kernel<LA, LB, LC, LD>), dim3((unsigned)blocks), dim3(256), 0, s, a);
  } else {
    if (a.det_scale)
      hipLaunchKernelGGL((jk_full_kernel<LA, LB, LC, LD, true>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((jk_full_kernel<LA, LB, LC, LD>), dim3((unsigned)blocks), dim3(256), 0, s, a);
  }
}

template <int LA, int LB, int LC, int LD>
void launch_blocks_t(const BlockArgs& a, hipStream_t s) {
  using K = Class<LA, LB, LC, LD>;
  if (a.ntasks <= 0) return;
  const int64_t threads = a.ntasks * K::NKET;
  const int64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL((blocks_kernel<LA, LB, LC, LD>), dim3((unsigned)blocks), dim3(256), 0, s, a);
}

// Dynamic LDS above 64 KB (one-density launches of wide molecules) is
// requested explicitly; a runtime that needs no request and rejects it leaves
// the launch itself to report a size it cannot give.
template <typename Kernel>
void request_lds(Kernel kernel, size_t bytes) {
  if (bytes <= 65536) return;
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)bytes) != hipSuccess)
    (void)hipGetLastError();
}

template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD>
void launch_jk_bra_t(const BraArgs& a, size_t lds_bytes, hipStream_t s) {
  if (a.nchunks <= 0) return;
  const dim3 block(a.threads), grid((unsigned)a.nchunks);
#define GSCF_BRA_LAUNCH(...)                                  \
  do {                                                        \
    request_lds(__VA_ARGS__, lds_bytes);                      \
    hipLaunchKernelGGL((__VA_ARGS__), grid, block, lds_bytes, s, a); \
  } while (0)
  // in-core integrals (FP64 blocks; segmented patterns only)
  if constexpr (CA * CB * CC * CD == 1) {
    if (a.mode == 1) {
      GSCF_BRA_LAUNCH(jk_bra_kernel<LA, LB, LC, LD, CA, CB, CC, CD, false, false, 1>);
      return;
    }
    if (a.mode == 2) {
      if (a.det_scale)
        GSCF_BRA_LAUNCH(jk_bra_kernel<LA, LB, LC, LD, CA, CB, CC, CD, false, false, 2, true>);
      else
        GSCF_BRA_LAUNCH(jk_bra_kernel<LA, LB, LC, LD, CA, CB, CC, CD, false, false, 2>);
      return;
    }
  }
  // deterministic digestion: FP64, Rys table from global memory
  if (a.det_scale) {
    GSCF_BRA_LAUNCH(jk_bra_kernel<LA, LB, LC, LD, CA, CB, CC, CD, false, false, 0, true>);
    return;
  }
  // variants: Rys table in LDS or global; FP64 only, or with the FP32 quartet path
  if (a.tab_lds && a.mixed)
    GSCF_BRA_LAUNCH(jk_bra_kernel<LA, LB, LC, LD, CA, CB, CC, CD, true, true>);
  else if (a.tab_lds)
    GSCF_BRA_LAUNCH(jk_bra_kernel<LA, LB, LC, LD, CA, CB, CC, CD, true, false>);
  else if (a.mixed)
    GSCF_BRA_LAUNCH(jk_bra_kernel<LA, LB, LC, LD, CA, CB, CC, CD, false, true>);
  else
    GSCF_BRA_LAUNCH(jk_bra_kernel<LA, LB, LC, LD, CA, CB, CC, CD, false, false>);
#undef GSCF_BRA_LAUNCH
}

// Column-pattern dispatch of the bra-major kernel (patterns: columns of the
// pair's shells (A, B) = (1 + (p & 1), 1 + (p >> 1))).
template <int LA, int LB, int LC, int LD, int PB, int PK>
bool launch_jk_bra_if(const BraArgs& a, int pb, int pk, size_t lds, hipStream_t s) {
  if constexpr (pattern_valid(LA, LB, PB) && pattern_valid(LC, LD, PK)) {
    if (pb == PB && pk == PK) {
      launch_jk_bra_t<LA, LB, LC, LD, 1 + (PB & 1), 1 + (PB >> 1), 1 + (PK & 1), 1 + (PK >> 1)>(a, lds, s);
      return true;
    }
  }
  return false;
}

template <int LA, int LB, int LC, int LD>
bool launch_jk_bra_patterns(const BraArgs& a, int pb, int pk, size_t lds, hipStream_t s) {
  return launch_jk_bra_if<LA, LB, LC, LD, 0, 0>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 0, 1>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 0, 2>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 0, 3>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 1, 0>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 1, 1>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 1, 3>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 2, 0>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 2, 1>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 2, 2>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 2, 3>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 3, 0>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 3, 1>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 3, 3>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 1, 2>(a, pb, pk, lds, s) ||
         launch_jk_bra_if<LA, LB, LC, LD, 3, 2>(a, pb, pk, lds, s);
}

}  // namespace gscf
