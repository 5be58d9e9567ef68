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
  double ip;       // 1 / p
};

// Per primitive pair, for the basis-parameter derivatives only: primitive
// indices inside the two segmented shells, their coefficients (s/p factor
// included) and the coefficient-free factor K_AB sqrt(2 pi^{5/2}) / p.
struct PrimAux {
  int ia, ib;
  double ca, cb;
  double ek;
};

// Column coefficients (s/p factor included) of a primitive pair of a
// contracted pair: primitive i of shell A in columns 0/1, primitive j of shell
// B in columns 0/1 (zero where a shell has one column).  The primitive pair's
// e then carries no coefficients.
struct PrimCoef {
  double ca[2];
  double cb[2];
};

// Column pattern of a pair: columns of (A, B) = (1 + (p & 1), 1 + (p >> 1)).
__host__ __device__ constexpr bool pattern_valid(int la, int lb, int p) {
  return p == 0 || (la == 0 && lb == 0 && (p == 1 || p == 3)) || (la > 0 && lb == 0 && p == 2);
}

struct Pair {
  int sa, sb;      // segmented shells, l(sa) >= l(sb)
  int poff, np;    // primitive-pair range
};

struct ShellXYZ {
  double x, y, z;
  int ao;          // Cartesian AO offset inside its molecule
  int l;
  int mol;         // molecule of a batched plan
  int lseg;        // segmented-shell index inside its molecule
};

// Per-molecule layout of a (batched) plan: Cartesian matrices of molecule m
// live at base .. base + n*n of every density/J/K batch; dbase/nseg address
// its shell-block density maxima.
struct MolInfo {
  int64_t base;
  int64_t dbase;
  int32_t n;
  int32_t nseg;
};


// Work of one quartet class is a list of canonical quartets (bra pair, ket
// pair), grouped into sub-tasks of similar primitive counts.  Sub-task t owns
// quartets [q0, q0+nq) and work items [item0, item0 + S*nq): item
// item0 + s*nq + g processes quartet q0+g for bra primitive pairs s, s+S, ...
struct SubTask {
  int64_t item0;
  int64_t q0;
  int64_t nq;
  int32_t S;
  int32_t pad;
};

struct JKArgs {
  const Pair* pairs;
  const PrimPair* prims;
  const ShellXYZ* shells;
  const int2* quartets;
  const SubTask* tasks;
  int ntasks;
  int64_t nitems;
  const double* dens;   // [nbd][sum_m n_m^2] Cartesian densities
  double* jp;           // [nbj][...] partial J (J = Jp + Jp^T)
  double* kp;           // [nbd][...] partial K
  const MolInfo* mols;
  int64_t bstride;      // elements per density batch
  int n, nbd, nbj, with_j, with_k;
  // deterministic digestion: [nmol][2] (scale, 1/scale); NULL for FP64 atomics
  const double* det_scale;
};

// Bra-major chunk: quartets [q0, q0+nq) of one sub-task whose bras are the
// consecutive pairs [bra0, bra0+nbra); one workgroup per chunk accumulates the
// exchange rows of those bras (and their J blocks) in LDS.
struct BraChunk {
  int64_t q0;
  int32_t nq;
  int32_t S;
  int32_t bra0;
  int32_t nbra;
  int32_t ket0;   // first ket pair of the chunk's ket window
  int32_t nket;   // window length when the ket-side J is accumulated in LDS (0: global atomics)
  // partial-sum slab of the chunk (per density): its exchange rows
  // [nbra][ROWS][n_m], then its ket window [nket][NKET]
  int64_t poff;
};

// Copies of each bra slot's J block in LDS: lane l adds into copy l % 8, so
// the lanes of a wavefront (which mostly share one bra) do not serialise on
// the same LDS addresses; the copies are summed at the flush.
constexpr int kJBraCopies = 8;

struct BraArgs {
  const Pair* pairs;
  const PrimPair* prims;
  const ShellXYZ* shells;
  const int2* quartets;
  const BraChunk* chunks;
  int nchunks;
  int nbra_max;         // LDS slots per workgroup
  const double* dens;   // [nbd][sum_m n_m^2]
  double* jp;
  double* kp;
  const MolInfo* mols;
  int64_t bstride;
  int n, nbd, nbj, with_j, with_k;  // n: widest molecule (LDS row length)
  // mixed precision: FP32 for quartets with pair_q[bra] pair_q[ket] dmax6 < thr
  int mixed;
  double thr;
  const double* pair_q;  // Schwarz bound per pair
  const double* dmax;    // per molecule [nseg][nseg] max |D| over shell blocks (all batches)
  const PrimCoef* coef;  // column coefficients of a contracted pair set (general patterns)
  int tab_lds;           // stage the class's Rys table slice in LDS (after the slabs)
  int nket_max;          // LDS ket-J window capacity in ket pairs (0: ket-side J by global atomics)
  int threads;           // workgroup size (256 or 512)
  // ket-side J with one 64-byte line per element (element i of jp at
  // jket_pad[8 i]), so atomics to neighbouring J elements do not share a line
  double* jket_pad;
  // in-core integrals: 0 evaluates quartets, 1 evaluates and stores each
  // contracted Cartesian block (the slice of one ket component for split
  // classes) at cache[(quartet * SPLIT + kc) * block], 2 reads them back
  int mode;
  double* cache;
  // partial sums instead of flush atomics: bit 0 stores the exchange rows, bit
  // 1 the ket window, into part[bt * part_stride + chunk.poff + ...] with plain
  // stores; reduction kernels add them into K and J afterwards
  int part_mode;
  int64_t part_stride;
  double* part;
  // exchange-row slab columns: per (molecule, ket mask) the first entry in
  // colmap and the count of the molecule's Cartesian columns whose shell l is
  // in the mask; this class's mask (1 << lc | 1 << ld)
  const int* colmap;
  const int2* cspace;
  int cmask;
  // LDS exchange rows over the same compact columns (segmented sets): per
  // shell and mask its first compact column; NULL: rows over all n columns
  const int* shell_cofs;
  // deterministic digestion: [nmol][2] (scale, 1/scale) of the build; NULL
  // for FP64 atomics (tab_lds 0, mixed 0 and part_mode 0 with it)
  const double* det_scale;
  // bra primitive pairs of the chunk staged in LDS (full-block segmented
  // classes): capacity in records of kStagedPrim doubles, 0 = not staged
  int bra_stage;
};

// Doubles per staged primitive pair: p, P - A (A: the pair's first centre),
// e, 1/p; a bra with general-contraction columns adds c_A[2], c_B[2].
constexpr int kStagedPrim = 6;
constexpr int kStagedPrimCoef = kStagedPrim + 4;

struct ParamArgs {
  const Pair* pairs;
  const PrimPair* prims;
  const PrimAux* aux;
  const ShellXYZ* shells;
  const int* seg_prim0;  // first primitive slot of each segmented shell
  const int2* quartets;
  const SubTask* tasks;
  int ntasks;
  int64_t nitems;
  const double* dens;
  const double* cj;
  const double* ck;
  const MolInfo* mols;
  int nprim;             // primitive slots over all segmented shells
  double* gcoef;         // [nprim] dE/d(coefficient incl. s/p factor)
  double* gexp;          // [nprim] dE/d(exponent) at fixed coefficients
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
void launch_jk_bra_class(int la, int lb, int lc, int ld, const BraArgs& a, size_t lds_bytes, hipStream_t s);
struct GradArgs {
  const Pair* pairs;
  const PrimPair* prims;
  const ShellXYZ* shells;
  const int* seg_atom;   // atm row of each segmented shell
  const int2* quartets;
  const SubTask* tasks;
  int ntasks;
  int64_t nitems;
  const double* dens;    // Cartesian D, cot_J, cot_K (per molecule n x n); cot may be null
  const double* cj;
  const double* ck;
  const MolInfo* mols;
  int natm;
  double* grad;          // [natm][3]
};

void launch_grad_class(int la, int lb, int lc, int ld, const GradArgs& a, hipStream_t s);
void launch_blocks_class(int la, int lb, int lc, int ld, const BlockArgs& a, hipStream_t s);

}  // namespace gscf
