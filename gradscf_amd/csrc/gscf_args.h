// Kernel argument blocks shared by the host plan code and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gscf {

// ---------------------------------------------------------------------------
// Plan data (device, structure-of-arrays where lanes gather)
// ---------------------------------------------------------------------------
struct PrimPair {
  double p;        // alpha_a + alpha_b
  double px, py, pz;
  double e;        // c_a c_b K_AB sqrt(2 pi^{5/2}) / p
  double aa;       // alpha_a (first-centre exponent)
  double ab;       // alpha_b
  double pad;
};

struct Pair {
  int sa, sb;      // segmented shells, l(sa) >= l(sb)
  int poff, np;    // primitive-pair range
};

struct ShellXYZ {
  double x, y, z;
  int ao;          // Cartesian AO offset
  int l;
};


// One canonical quartet class: bra pairs [bra0, bra0+nbra) of pair class P, ket
// pairs [ket0, ...) of class Q <= P.  Bra b owns kets [0, qoff[b+1]-qoff[b]).
struct JKArgs {
  const Pair* pairs;
  const PrimPair* prims;
  const ShellXYZ* shells;
  const int64_t* qoff;
  int bra0, ket0, nbra, same_class;
  int64_t nq;
  const double* dens;   // [nbd][n][n] Cartesian densities
  double* jp;           // [nbj][n][n] partial J (J = Jp + Jp^T)
  double* kp;           // [nbd][n][n] partial K
  int n, nbd, nbj, with_j, with_k;
};

// Explicit quartet list writing contracted Cartesian blocks:
// out[base + sum_slot (coff[slot] + comp[slot]) * stride[slot]] for class-order slots A,B,C,D.
struct BlockTask {
  int bra, ket;
  int64_t base;
  int stride[4];
  int coff[4];
};

struct BlockArgs {
  const Pair* pairs;
  const PrimPair* prims;
  const ShellXYZ* shells;
  const BlockTask* tasks;
  int64_t ntasks;
  double* out;
};

// Pair classes for l <= 2: index la*(la+1)/2 + lb with la >= lb.
constexpr int kNumPairClasses = 6;
__host__ __device__ constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }

void launch_jk_class(int la, int lb, int lc, int ld, const JKArgs& a, hipStream_t s);
void launch_blocks_class(int la, int lb, int lc, int ld, const BlockArgs& a, hipStream_t s);

}  // namespace gscf
