// Device-side building blocks of the MI355X direct-J/K engine (gfx950, wave64).
//
// Arithmetic follows the Rys-quadrature form of the reference's ERI kernels
// (libcint, src/gradscf/integrals/_native/vendor/libcint/src/g2e.c:4425-4569 for
// the root/weight and B00/B10/B01/C00/C0p set-up, g2e.c:272-690 for the 2D
// recurrences and transfers, cint2e.c:660-758,961 for assembly/contraction),
// restructured for the GPU: one angular-momentum class per template
// instantiation, everything unrolled into registers, and the density
// digestion of csrc/direct_jk.cc:45-76 fused into the same kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "_generated/rys_table.inc"
#include "gscf_args.h"

namespace gscf {

constexpr int kMaxL = 2;  // shells up to d in the J/K kernels

__device__ static const double rys_tab[] = {
#include "_generated/rys_table_body.inc"
};

// F_0..F_{BOYS_NB-1} at T_i = i BOYS_H, i < BOYS_NPT (BOYS_NB per point)
__device__ static const double boys_tab[] = {
#include "_generated/boys_table_body.inc"
};

// ---------------------------------------------------------------------------
// Rys roots (s = t^2) and weights for weight exp(-x t^2) on [0,1]
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Reciprocal and reciprocal square root of finite, positive, normal FP64
// arguments (exponent sums, Boys arguments): the hardware estimates refined
// by two Newton-Raphson steps, r <- r + r (1 - x r) and
// y <- y + y (1/2 - x y^2 / 2); each step squares the relative error, so two
// steps from the estimate reach FP64 rounding (about 1 ulp).  The general
// IEEE sequences handle zeros, denormals and infinities these arguments
// never take, at three to four times the instruction count.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rcp_pos(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

__device__ __forceinline__ double rsqrt_pos(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  double t = fma(-hx * y, y, 0.5);
  y = fma(y, t, y);
  t = fma(-hx * y, y, 0.5);
  return fma(y, t, y);
}

// Boys functions F_0..F_{M-1}(x), 0 <= x < BOYS_TMAX, by a BOYS_K-term Taylor
// expansion about the nearest grid point T_i = i BOYS_H (dF_m/dT = -F_{m+1}):
//   F_m(T_i - d) = sum_{k < K} F_{m+k}(T_i) d^k / k!,   |d| <= BOYS_H / 2,
// nested as f_m + (d/1)(f_{m+1} + (d/2)(f_{m+2} + ...)).  Truncation below
// 2e-17 relative; every function shares the d/k factors and one grid row.
template <int M>
__device__ __forceinline__ void boys_taylor(double x, double (&F)[M]) {
  static_assert(M + BOYS_K - 1 <= BOYS_NB, "Boys grid row too short");
  const int i = (int)(x * (1.0 / BOYS_H) + 0.5);
  const double d = i * BOYS_H - x;
  const double* __restrict__ f = boys_tab + i * BOYS_NB;
  double dk[BOYS_K];
#pragma unroll
  for (int k = 1; k < BOYS_K; ++k) dk[k] = d * (1.0 / k);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    double r = f[m + BOYS_K - 1];
#pragma unroll
    for (int k = BOYS_K - 1; k >= 1; --k) r = fma(r, dk[k], f[m + k - 1]);
    F[m] = r;
  }
}

// Rys roots s_i = t_i^2 and weights of one or two points from the Boys
// moments mu_k = F_k(x) = sum_i w_i s_i^k (exact for k < 2N):
//   N = 1:  w = F_0, s = F_1 / F_0;
//   N = 2:  s_1 < s_2 are the zeros of s^2 + a1 s + a0 with
//           [mu0 mu1; mu1 mu2] (a0, a1) = -(mu2, mu3), the larger root from
//           the quadratic formula without cancellation (-a1 > 0), the smaller
//           as a0 / s_2; w_1 = (s_2 mu0 - mu1) / (s_2 - s_1), w_2 = mu0 - w_1.
// Relative error below 4e-15 for x < BOYS_TMAX (tested against 60-digit values).
template <int N>
__device__ __forceinline__ void rys_roots_boys(double x, double* __restrict__ s, double* __restrict__ w) {
  static_assert(N == 1 || N == 2, "closed-form quadrature for one or two roots");
  if constexpr (N == 1) {
    double F[2];
    boys_taylor<2>(x, F);
    w[0] = F[0];
    s[0] = F[1] * rcp_pos(F[0]);
  } else {
    double F[4];
    boys_taylor<4>(x, F);
    const double iD = rcp_pos(F[0] * F[2] - F[1] * F[1]);
    const double a1 = (F[1] * F[2] - F[0] * F[3]) * iD;
    const double a0 = (F[1] * F[3] - F[2] * F[2]) * iD;
    const double disc = a1 * a1 - 4.0 * a0;
    const double s2 = 0.5 * (disc * rsqrt_pos(disc) - a1);
    const double s1 = a0 * rcp_pos(s2);
    const double w1 = (s2 * F[0] - F[1]) * rcp_pos(s2 - s1);
    s[0] = s1;
    s[1] = s2;
    w[0] = w1;
    w[1] = F[0] - w1;
  }
}

// `tab` is the N-root slice of the table (rys_tab + rys_table_offset[N - 1],
// or a copy of it staged in LDS): RYS_NINT intervals of (RYS_DEG + 1) x 2N
// monomial coefficients in y (degree-major), roots then weights.  With
// MOMENTS (one or two roots) the quadrature comes from the Boys moments and
// the table is not read: fewer memory reads, more arithmetic; which one wins
// depends on where the table lives (LDS or global) and is chosen per class.
template <int N, bool MOMENTS = (N <= 2)>
__device__ __forceinline__ void rys_roots_tab(const double* __restrict__ tab, double x, double* __restrict__ s,
                                              double* __restrict__ w) {
  static_assert(!MOMENTS || N <= 2, "moment quadrature for one or two roots");
  if constexpr (MOMENTS) {
    if (x < BOYS_TMAX) {
      rys_roots_boys<N>(x, s, w);
      return;
    }
  } else {
    if (x < RYS_XMAX) {
      const int k = (int)(x * (1.0 / RYS_H));
      const double y = x - (RYS_H * k + 0.5 * RYS_H);
      const double* __restrict__ c = tab + k * ((RYS_DEG + 1) * 2 * N);
      // Horner in y on [-1, 1): one FMA per degree and function
      double acc[2 * N];
#pragma unroll
      for (int f = 0; f < 2 * N; ++f) acc[f] = c[RYS_DEG * 2 * N + f];
#pragma unroll
      for (int j = RYS_DEG - 1; j >= 0; --j)
#pragma unroll
        for (int f = 0; f < 2 * N; ++f) acc[f] = fma(acc[f], y, c[j * 2 * N + f]);
#pragma unroll
      for (int f = 0; f < N; ++f) {
        s[f] = acc[f];
        w[f] = acc[N + f];
      }
      return;
    }
  }
  // large x: Gauss-Hermite limit, exact to double precision above the switch
  const double sq = rsqrt_pos(x);
  const double ix = sq * sq;
  constexpr int base = N * (N - 1) / 2;
#pragma unroll
  for (int f = 0; f < N; ++f) {
    s[f] = rys_asym_root_c[base + f] * ix;
    w[f] = rys_asym_weight_c[base + f] * sq;
  }
}

template <int N>
__device__ __forceinline__ void rys_roots(double x, double* __restrict__ s, double* __restrict__ w) {
  rys_roots_tab<N>(rys_tab + rys_table_offset[N - 1], x, s, w);
}

// Doubles in the N-root slice of the table.
constexpr int rys_slice_size(int n) { return RYS_NINT * (RYS_DEG + 1) * 2 * n; }

// ---------------------------------------------------------------------------
// Cartesian components in libcint order (descending x, then descending y)
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

__host__ __device__ constexpr int cart_comp(int l, int i, int axis) {
  int n = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      if (n == i) return axis == 0 ? lx : (axis == 1 ? ly : l - lx - ly);
      ++n;
    }
  return 0;
}

__host__ __device__ constexpr int binom(int n, int k) {
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// ---------------------------------------------------------------------------
// Vertical recurrence for one axis and one root:
//   G(n+1,0) = C00 G(n,0) + n B10 G(n-1,0)
//   G(n,m+1) = C0p G(n,m) + m B01 G(n,m-1) + n B00 G(n-1,m)
// ---------------------------------------------------------------------------
// (terms with a zero factor are left out rather than added as zeros, which
// the compiler may not drop: x + 0 is not x for x = -0)
template <int NM, int MM, typename R = double>
__device__ __forceinline__ void vrr(R c00, R c0p, R b10, R b01, R b00, R base, R (&G)[NM + 1][MM + 1]) {
  G[0][0] = base;
#pragma unroll
  for (int n = 0; n < NM; ++n) {
    R t = c00 * G[n][0];
    if (n) t += R(n) * b10 * G[n - 1][0];
    G[n + 1][0] = t;
  }
#pragma unroll
  for (int m = 0; m < MM; ++m)
#pragma unroll
    for (int n = 0; n <= NM; ++n) {
      R t = c0p * G[n][m];
      if (m) t += R(m) * b01 * G[n][m - 1];
      if (n) t += R(n) * b00 * G[n - 1][m];
      G[n][m + 1] = t;
    }
}

// Horizontal transfer (x - B) = (x - A) + (A - B) on the bra index of a 1D vector.
template <int LA, int LB, typename R = double>
__device__ __forceinline__ void hrr_bra1(const R (&g)[LA + LB + 1], R ab, R (&h)[LA + 1][LB + 1]) {
  R t[LB + 1][LA + LB + 1];
#pragma unroll
  for (int n = 0; n <= LA + LB; ++n) t[0][n] = g[n];
#pragma unroll
  for (int b = 1; b <= LB; ++b)
#pragma unroll
    for (int a = 0; a <= LA + LB - b; ++a) t[b][a] = t[b - 1][a + 1] + ab * t[b - 1][a];
#pragma unroll
  for (int a = 0; a <= LA; ++a)
#pragma unroll
    for (int b = 0; b <= LB; ++b) h[a][b] = t[b][a];
}

// Full 4-index transfer of one axis: H[a][b][c][d] from G[n][m].
template <int LA, int LB, int LC, int LD, typename R = double>
__device__ __forceinline__ void hrr_full(const R (&G)[LA + LB + 1][LC + LD + 1], R ab, R cd,
                                         R (&H)[LA + 1][LB + 1][LC + 1][LD + 1]) {
  R X[LA + 1][LB + 1][LC + LD + 1];
#pragma unroll
  for (int m = 0; m <= LC + LD; ++m) {
    R g[LA + LB + 1];
#pragma unroll
    for (int n = 0; n <= LA + LB; ++n) g[n] = G[n][m];
    R h[LA + 1][LB + 1];
    hrr_bra1<LA, LB, R>(g, ab, h);
#pragma unroll
    for (int a = 0; a <= LA; ++a)
#pragma unroll
      for (int b = 0; b <= LB; ++b) X[a][b][m] = h[a][b];
  }
#pragma unroll
  for (int a = 0; a <= LA; ++a)
#pragma unroll
    for (int b = 0; b <= LB; ++b) {
      R g[LC + LD + 1];
#pragma unroll
      for (int m = 0; m <= LC + LD; ++m) g[m] = X[a][b][m];
      R h[LC + 1][LD + 1];
      hrr_bra1<LC, LD, R>(g, cd, h);
#pragma unroll
      for (int c = 0; c <= LC; ++c)
#pragma unroll
        for (int d = 0; d <= LD; ++d) H[a][b][c][d] = h[c][d];
    }
}

}  // namespace gscf
