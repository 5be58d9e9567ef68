// Host side of the MI355X direct-J/K engine: table validation, device plan,
// Schwarz bounds, class task lists, density / J / K transforms, C ABI.
//
// Reference semantics followed:
//   table validation and AO offsets      src/gradscf/integrals/_native/csrc/tables.h:31-76
//   Schwarz bounds Q_ij                   csrc/coulomb_shells.h:60-77
//   pair prune / quartet skip tests       csrc/coulomb_shells.h:98-110
//   operator selection, shape, finite-density checks, output layout
//                                         csrc/direct_jk.cc:15-28,51,64,97-102
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "_generated/rys_table.inc"
#include "gradscf_jk.h"
#include "gscf_args.h"

namespace gscf {

// ----------------------------------------------------------------------------
// errors
// ----------------------------------------------------------------------------
static thread_local std::string g_error;

struct InvalidArgument : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct Internal : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIP_OK(expr)                                                                     \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      throw Internal(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr); \
  } while (0)

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return GSCF_OK;
  } catch (const InvalidArgument& e) {
    g_error = e.what();
    return GSCF_INVALID_ARGUMENT;
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return GSCF_INVALID_ARGUMENT;
  } catch (const std::exception& e) {
    g_error = e.what();
    return GSCF_INTERNAL;
  }
}

// libcint table slots
constexpr int ATM_SLOTS = 6, BAS_SLOTS = 8;
constexpr int CHARGE_OF = 0, PTR_COORD = 1, NUC_MOD_OF = 2;
constexpr int ATOM_OF = 0, ANG_OF = 1, NPRIM_OF = 2, NCTR_OF = 3, KAPPA_OF = 4, PTR_EXP = 5, PTR_COEFF = 6;
constexpr int PTR_ENV_START = 20, PTR_RANGE_OMEGA = 8;
constexpr int kMaxL = 2;

static int ncart_h(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian -> real spherical coefficients for l <= 2 (rows: spherical m in
// PySCF order; columns: Cartesian components in libcint order).
// Real solid harmonics r^l Y_lm of l >= 3 (unit-sphere normalisation,
// m = -l..l) as polynomials in the Cartesian monomials, in libcint's
// component order (descending x, then descending y): the closed form of
// Helgaker, Jorgensen and Olsen, Molecular Electronic-Structure Theory,
// eqs. (6.4.47)-(6.4.48), scaled from Racah's normalisation by
// sqrt((2l+1)/(4 pi)).  (The d rows below are the same polynomials.)
static void solid_harmonic_rows(int l, std::vector<std::vector<double>>& m) {
  const double pi = 3.14159265358979323846;
  auto fact = [](int k) {
    double r = 1.0;
    for (int i = 2; i <= k; ++i) r *= i;
    return r;
  };
  auto choose = [&](int a, int b) { return (b < 0 || b > a) ? 0.0 : fact(a) / (fact(b) * fact(a - b)); };
  auto comp = [&](int lx, int ly) {
    int idx = 0;
    for (int x = l; x >= 0; --x)
      for (int y = l - x; y >= 0; --y) {
        if (x == lx && y == ly) return idx;
        ++idx;
      }
    return -1;
  };
  m.assign(2 * l + 1, std::vector<double>(ncart_h(l), 0.0));
  const double unit = std::sqrt((2.0 * l + 1.0) / (4.0 * pi));
  for (int mm = -l; mm <= l; ++mm) {
    const int am = std::abs(mm);
    const int half = mm < 0 ? 1 : 0;  // v runs over half-integers (2v odd) for m < 0
    const double norm = std::sqrt(2.0 * fact(l + am) * fact(l - am) / (mm == 0 ? 2.0 : 1.0)) /
                        (std::pow(2.0, am) * fact(l));
    for (int t = 0; t <= (l - am) / 2; ++t)
      for (int u = 0; u <= t; ++u)
        for (int twov = half; twov <= (half ? 2 * ((am - 1) / 2) + 1 : 2 * (am / 2)); twov += 2) {
          // (-1)^(t + v - v_m): v - v_m = (twov - half) / 2
          const double sign = ((t + (twov - half) / 2) & 1) ? -1.0 : 1.0;
          const double c = sign * std::pow(0.25, t) * choose(l, t) * choose(l - t, am + t) * choose(t, u) *
                           choose(am, twov);
          const int px = 2 * t + am - 2 * u - twov, py = 2 * u + twov;
          if (px < 0 || c == 0.0) continue;
          m[mm + l][comp(px, py)] += norm * c * unit;
        }
  }
}

static void c2s_rows(int l, std::vector<std::vector<double>>& m) {
  const double pi = 3.14159265358979323846;
  if (l >= 3) {
    solid_harmonic_rows(l, m);
    return;
  }
  if (l <= 1) {
    m.assign(ncart_h(l), std::vector<double>(ncart_h(l), 0.0));
    for (int i = 0; i < ncart_h(l); ++i) m[i][i] = 1.0;
    return;
  }
  const double a = std::sqrt(15.0 / (4.0 * pi)), b = std::sqrt(5.0 / (16.0 * pi)), c = std::sqrt(15.0 / (16.0 * pi));
  m.assign(5, std::vector<double>(6, 0.0));
  m[0][1] = a;                                   // xy
  m[1][4] = a;                                   // yz
  m[2][0] = -b; m[2][3] = -b; m[2][5] = 2.0 * b;  // z^2
  m[3][2] = a;                                   // xz
  m[4][0] = c; m[4][3] = -c;                      // x^2 - y^2
}

template <class T>
static T* dev_upload(const std::vector<T>& v, hipStream_t s) {
  if (v.empty()) return nullptr;
  T* p = nullptr;
  HIP_OK(hipMalloc(&p, sizeof(T) * v.size()));
  HIP_OK(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  return p;
}

struct Seg {
  int parent, ctr, l, np;
  double r[3];
  int ao_cart;               // Cartesian AO offset inside its molecule
  int mol, lseg;             // molecule, segment index inside the molecule
  std::vector<double> a, c;  // exponents, coefficients * s/p factor
};

// Per-molecule AO layout on the device for the transform kernels: matrices of
// molecule m start at sbase (requested basis), cbase (Cartesian) and dbase
// (shell-block maxima) of every batch; ao0/aoc0 index the global AO maps.
struct MolAO {
  int64_t sbase, cbase, dbase;
  int32_t n, nc, ao0, aoc0, seg0, nseg;
};

// Per-class schedule parameters (see gscf_class_tuning); plans copy the
// process-wide table when they are created.
constexpr int kNumQuartetClasses = kNumPairClasses * (kNumPairClasses + 1) / 2;
static gscf_class_tuning g_tuning[kNumQuartetClasses] = {};
static bool g_tuning_init = false;

static gscf_class_tuning default_tuning() {
  gscf_class_tuning t;
  t.chunk_items = 4096;   // lane-items per workgroup before a chunk is closed
  t.bras_max = 16;        // bra pairs (LDS slots) per chunk
  t.lds_pack = 32 * 1024; // LDS bytes the slots of one workgroup may use
  t.rys_lds = 0;          // stage the Rys table slice in LDS
  t.target_prim = 48;     // primitive quartets per work item (primitive split)
  t.threads = 256;        // workgroup size
  t.ket_lds = 0;          // accumulate the ket-side J of a chunk's ket window in LDS
  return t;
}

static void ensure_tuning() {
  if (g_tuning_init) return;
  for (auto& t : g_tuning) t = default_tuning();
  g_tuning_init = true;
}

// defaults: segmented general contractions (faster on the measured cc-pVDZ
// workloads; a tuning table may switch it on), four launch streams
static gscf_options g_options = {0, 4, 1, 0, 0, 1};
constexpr int kMaxMoleculeCartAO = 16384;  // per molecule (32-bit in-molecule offsets)
constexpr int kMaxStreams = 4;

static size_t rys_slice_bytes(int nr) {
  return sizeof(double) * ((size_t)RYS_NINT * (RYS_DEG + 1) * 2 * nr + 2);  // + alignment pad
}

struct QuartetClass {
  int la, lb, lc, ld;
  int pb = 0, pk = 0;       // column patterns of bra and ket pairs (contracted sets)
  int tuning_index = 0;
  int tab_lds = 0;          // Rys table slice staged in LDS
  size_t tab_bytes = 0;
  int nket_max = 0;         // LDS ket-J window capacity (ket pairs), 0 = global atomics
  int stage_cap = 0;        // staged bra primitive pairs per chunk (full-block segmented classes)
  int threads = 256;
  size_t per_slot = 0;      // LDS bytes of one bra slot for one density
  int row_len = 0;          // LDS exchange-row length (columns of the ket mask, widest molecule)
  int64_t nq = 0, nitems = 0;
  double cost = 0.0;        // estimated work, for spreading launches over streams
  std::vector<SubTask> tasks;
  SubTask* d_tasks = nullptr;
  int2* d_quartets = nullptr;
  // bra-major chunks for the LDS-accumulating kernel (built for one density)
  std::vector<BraChunk> chunks;
  BraChunk* d_chunks = nullptr;
  int nbra_max = 0;
  size_t cache_off = 0, cache_doubles = 0;  // in-core integrals of this class (doubles)
  int64_t part_beg = 0, part_end = 0;       // partial-slab range of this class's chunks (per density)
  bool part = false;                        // chunks flush into partial slabs
};

// LDS of the bra-major kernel per workgroup: slots are packed up to the
// class's lds_pack so that several workgroups stay resident per CU, within
// kLdsBudget for the densities of one launch.  Exchange rows span the widest
// molecule, so a wide molecule's slot can exceed that: such a class runs one
// density per launch, and a slot may then take up to the device's LDS per
// workgroup (160 KB on gfx950, one workgroup per CU) before the class falls
// back to the global-atomic kernels.
constexpr size_t kLdsBudget = 64 * 1024;
constexpr size_t kLdsDeviceMax = 160 * 1024;
constexpr size_t kKetLdsBytes = 16 * 1024;  // LDS ket-J window per density
// Partial-sum slabs per density at most; a plan whose slabs would be larger
// flushes with atomics instead.
constexpr int64_t kPartCapBytes = int64_t(2) << 30;

// The shell pairs quartet kernels run over.  Segmented sets hold one
// contraction column per shell; contracted sets (J/K only) hold parent shells
// whose general-contraction columns are evaluated from one primitive
// integral each (column pattern per pair: (1,1), (2,1), (2,2) for ss pairs,
// (1,2) for ps/ds pairs with a general s shell).
struct PairSet {
  bool contracted = false;
  bool built = false;
  ShellXYZ* d_shells = nullptr;   // owned only by contracted sets
  PrimPair* d_prims = nullptr;    // owned only by contracted sets
  PrimCoef* d_coef = nullptr;     // contracted sets: column coefficients per primitive pair
  int* d_blk = nullptr;           // [nshell][2]: AO offset and width of each shell's density block
  std::vector<Pair> pairs;        // retained, class/pattern/molecule/bucket sorted
  std::vector<double> pair_q;
  Pair* d_pairs = nullptr;
  double* d_pair_q = nullptr;
  std::vector<QuartetClass> classes;
  int64_t nquartets = 0;
  int64_t nprims = 0;             // primitive pairs owned by a contracted set
  // partial-sum slabs of the bra-major chunks (segmented sets): doubles per
  // density, and plan-time lists of the slab rows that add into each
  // Cartesian K row (krp/kent, row-major CSR over all molecules' rows) and
  // of the window entries that add into each ket pair's J block (jrp/jent)
  int64_t part_total = 0;
  int64_t* d_krp = nullptr;
  int64_t* d_kent = nullptr;
  int64_t* d_jrp = nullptr;
  int64_t* d_jent = nullptr;
  std::vector<std::pair<int64_t, int64_t>> h_kent;  // (row, slab offset) of every exchange-row slab row
  std::vector<int> h_kent_cls;                      // its class
  int64_t* d_row_dst = nullptr;   // per global Cartesian row: first K element and length
  int* d_row_n = nullptr;
  int* d_row_mol = nullptr;       // molecule of every global Cartesian row
  // exchange-row slab columns (segmented sets): a class's slab rows hold the
  // Cartesian AO columns of its molecule's shells whose l is in the class's
  // ket mask (1 << lc | 1 << ld), the only columns its exchange terms reach.
  // Per (molecule, mask): (first entry in colmap, count); cinv maps a
  // molecule's column to its compact index in the mask's list, or -1.
  std::vector<int2> h_cspace;
  int* d_shell_cofs = nullptr;    // per segmented shell and mask: first compact column (-1: l not in mask)
  int* d_colmap = nullptr;
  int2* d_cspace = nullptr;
  int64_t* d_cinv_off = nullptr;
  int* d_cinv = nullptr;
  int64_t nrows = 0;
  int nc_max = 0;
};

struct PlanData {
  gscf_class_tuning tuning[kNumQuartetClasses];
  gscf_options options;
  size_t lds_budget = kLdsBudget;  // LDS of a launch over all densities (packing target)
  size_t lds_max = kLdsBudget;     // LDS of a one-density launch (device limit)
  int cart = 1, n = 0, n_cart = 0, nbas = 0, natm = 0;
  double cutoff = 0.0, plan_ms = 0.0;
  std::vector<int> atm, bas, ao, ao_cart;  // ao/ao_cart: global AO offsets of every shell
  std::vector<double> env;
  // molecules: atm rows [mol_atm[m], mol_atm[m+1]), bas rows [mol_bas[m], mol_bas[m+1])
  int nmol = 1;
  std::vector<int> mol_atm, mol_bas, mol_seg;  // (nmol + 1) offsets
  std::vector<MolAO> mols;                     // nmol + 1 (the last holds the totals)
  std::vector<int64_t> mol_pair0;              // first all_pairs index of each molecule
  int nc_max = 0;                              // widest molecule (Cartesian AOs)
  int64_t nn_tot = 0, ncc_tot = 0, nd_tot = 0; // packed matrix sizes of one batch
  std::vector<Seg> seg;
  std::vector<PrimPair> prims;
  std::vector<PrimAux> aux;          // aligned with prims
  std::vector<int> seg_prim0;        // primitive slots of each segmented shell (nseg + 1)
  // all intra-molecule segmented pairs (unscreened) for block queries
  std::vector<Pair> all_pairs;
  std::vector<int> all_class;
  // J/K pair set and the segmented set of the derivative kernels (the same
  // object when J/K runs on segmented pairs; built on first use otherwise)
  PairSet jk;
  PairSet seg_set;
  std::vector<double> seg_pair_q;  // parent-shell Schwarz bound of every all_pairs entry
  std::vector<double> qmax_mol;    // largest bound per molecule
  PairSet& segq() { return jk.contracted ? seg_set : jk; }
  // device
  ShellXYZ* d_shells = nullptr;
  PrimPair* d_prims = nullptr;
  Pair* d_all_pairs = nullptr;
  int* d_c2s_idx = nullptr;     // [n_cart][2] spherical AO feeding a Cartesian AO
  double* d_c2s_coef = nullptr;
  int* d_s2c_idx = nullptr;     // [n][3] Cartesian AOs feeding a spherical AO
  double* d_s2c_coef = nullptr;
  MolInfo* d_molinfo = nullptr; // kernel view of the molecule layout
  MolAO* d_molao = nullptr;
  // workspace
  double* d_dc = nullptr;
  double* d_jp = nullptr;
  double* d_jpad = nullptr;     // ket-side J, one 64-byte line per element
  double* d_part = nullptr;     // partial-sum slabs of the bra-major chunks, [nbd][part_total]
  // per launch stream: row lists of the exchange-row slabs of its classes, so
  // that each stream folds its own slabs as soon as its classes are done
  std::vector<int64_t*> d_lane_krp, d_lane_kent;
  size_t part_nbd = 0;
  double* d_incore = nullptr;   // in-core integrals: stored Cartesian blocks of the J/K quartets
  size_t incore_doubles = 0;

  double* d_kp = nullptr;
  size_t ws_nbd = 0;
  std::vector<double*> retired;  // outgrown workspaces, freed with the plan
  int* d_flag = nullptr;
  int* h_flag = nullptr;
  // one-electron: all segmented pairs sorted by class, nuclei (x, y, z, Z, molecule)
  Pair* d_pairs1e = nullptr;
  int p1e_begin[kNumPairClasses + 1] = {0};
  double* d_nuclei = nullptr;
  int nnuc = 0;
  // parent-shell Schwarz bounds, one nbas_m x nbas_m block per molecule
  // (lower triangle, local shell indices; empty when cutoff == 0)
  std::vector<double> parent_q;
  std::vector<int64_t> qbase;      // start of each molecule's parent_q block
  // mixed precision: per-build density block maxima over the J/K shell blocks
  double* d_dmax = nullptr;
  // deterministic digestion: per density-block Cartesian Schwarz weight
  // Q(S,T) nS nT [nd_tot], per molecule (Qmax, max_S sum_T Q(S,T) nT), and
  // the fixed-point (scale, 1/scale) of the current build, per molecule
  double* d_det_qblk = nullptr;
  double* d_det_mol = nullptr;
  double* d_det_scale = nullptr;
  int* d_seg_atom = nullptr; // atm row of each segmented shell
  // basis-parameter derivatives
  PrimAux* d_aux = nullptr;
  int* d_seg_prim0 = nullptr;
  int2* d_prim_slots = nullptr;  // per primitive slot: env index of its exponent, of its coefficient
  double* d_prim_fsp = nullptr;  // s/p factor folded into the slot's coefficient
  double* d_gparam = nullptr;    // [2][nprim] scratch sums
  // optional per-class kernel timing with HIP events on the launch stream
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct ProfRec { int cls; size_t ev; int64_t build; int qci; };
  std::vector<ProfRec> prof_recs;  // (canonical class, event pair offset, build number, class list index)
  int64_t prof_builds = 0;
  // class launches of one build spread over the caller's stream and side
  // streams (default priority, non-blocking), joined back by events
  std::vector<hipStream_t> side;
  hipEvent_t fork_ev = nullptr;
  std::vector<hipEvent_t> join_ev;
  std::vector<std::vector<int>> lane_classes;  // class indices per stream (0 = caller's)
};

// Per-class launchers, one translation unit each (jk_class.hip).
#define GSCF_CLASSES(X)                                                               \
  X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) \
  X(2, 0, 0, 0) X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) \
  X(2, 1, 1, 1) X(2, 1, 2, 0) X(2, 1, 2, 1) X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1) \
  X(2, 2, 2, 0) X(2, 2, 2, 1) X(2, 2, 2, 2)
#define X(A, B, C, D)                                                       \
  void launch_jk_c##A##B##C##D(const JKArgs& a, hipStream_t s);            \
  void launch_blocks_c##A##B##C##D(const BlockArgs& a, hipStream_t s);
GSCF_CLASSES(X)
#undef X

// density-fitting block kernels, one translation unit per auxiliary l (df_class.hip)
bool launch_df_blocks_c0(int la, int lb, const BlockArgs& a, hipStream_t s);
bool launch_df_blocks_c1(int la, int lb, const BlockArgs& a, hipStream_t s);
bool launch_df_blocks_c2(int la, int lb, const BlockArgs& a, hipStream_t s);
bool launch_df_blocks_c3(int la, int lb, const BlockArgs& a, hipStream_t s);
bool launch_df_blocks_c4(int la, int lb, const BlockArgs& a, hipStream_t s);

void launch_int1e_class(int la, int lb, const Pair* pairs, int npairs, const PrimPair* prims,
                        const ShellXYZ* shells, const double* nuclei, int nnuc, const MolInfo* mols, double* s,
                        double* t, double* v, hipStream_t st);
void launch_param1e_class(int la, int lb, const Pair* pairs, int npairs, const PrimPair* prims, const PrimAux* aux,
                          const ShellXYZ* shells, const int* seg_prim0, const double* nuclei, int nnuc,
                          const MolInfo* mols, const double* dens, const double* wdens, int nprim, double* gcoef,
                          double* gexp, hipStream_t st);
void launch_grad1e_class(int la, int lb, const Pair* pairs, int npairs, const PrimPair* prims,
                         const ShellXYZ* shells, const int* seg_atom, const double* nuclei, int nnuc,
                         const MolInfo* mols, const double* dens, const double* wdens, int natm, double* grad,
                         hipStream_t st);

#define X(A, B, C, D) bool launch_jk_bra_c##A##B##C##D(const BraArgs& a, int pb, int pk, size_t lds, hipStream_t s);
GSCF_CLASSES(X)
#undef X
#define X(A, B, C, D) void launch_grad_c##A##B##C##D(const GradArgs& a, hipStream_t s);
GSCF_CLASSES(X)
#undef X

#define X(A, B, C, D) void launch_param_c##A##B##C##D(const ParamArgs& a, hipStream_t s);
GSCF_CLASSES(X)
#undef X

void launch_param_class(int la, int lb, int lc, int ld, const ParamArgs& a, hipStream_t s) {
#define X(A, B, C, D)                                      \
  if (la == A && lb == B && lc == C && ld == D) {          \
    launch_param_c##A##B##C##D(a, s);                      \
    return;                                                \
  }
  GSCF_CLASSES(X)
#undef X
  throw Internal("no kernel for this angular-momentum class");
}

void launch_grad_class(int la, int lb, int lc, int ld, const GradArgs& a, hipStream_t s) {
#define X(A, B, C, D)                                      \
  if (la == A && lb == B && lc == C && ld == D) {          \
    launch_grad_c##A##B##C##D(a, s);                       \
    return;                                                \
  }
  GSCF_CLASSES(X)
#undef X
  throw Internal("no kernel for this angular-momentum class");
}

void launch_jk_bra_class(int la, int lb, int lc, int ld, int pb, int pk, const BraArgs& a, size_t lds,
                         hipStream_t s) {
#define X(A, B, C, D)                                                                 \
  if (la == A && lb == B && lc == C && ld == D) {                                     \
    if (!launch_jk_bra_c##A##B##C##D(a, pb, pk, lds, s))                              \
      throw Internal("no kernel for this column pattern");                            \
    return;                                                                           \
  }
  GSCF_CLASSES(X)
#undef X
  throw Internal("no kernel for this angular-momentum class");
}

void launch_jk_class(int la, int lb, int lc, int ld, const JKArgs& a, hipStream_t s) {
#define X(A, B, C, D)                                      \
  if (la == A && lb == B && lc == C && ld == D) {          \
    launch_jk_c##A##B##C##D(a, s);                         \
    return;                                                \
  }
  GSCF_CLASSES(X)
#undef X
  throw Internal("no kernel for this angular-momentum class");
}

void launch_blocks_class(int la, int lb, int lc, int ld, const BlockArgs& a, hipStream_t s) {
#define X(A, B, C, D)                                      \
  if (la == A && lb == B && lc == C && ld == D) {          \
    launch_blocks_c##A##B##C##D(a, s);                     \
    return;                                                \
  }
  GSCF_CLASSES(X)
#undef X
  throw Internal("no kernel for this angular-momentum class");
}

}  // namespace gscf

struct gscf_plan {
  gscf::PlanData d;
};

namespace gscf {

// ----------------------------------------------------------------------------
// small device kernels
// ----------------------------------------------------------------------------
// FP64 FMA issue-rate probe: eight independent FMA chains per lane (enough
// to cover the FMA latency), the result stored so nothing is eliminated.
__global__ __launch_bounds__(256) void fma_probe_kernel(int64_t iters, double* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = 1e-3 * (double)(t + k);
  const double m = 0.9999999, c = 1e-9;
  for (int64_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = fma(a[k], m, c);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k];
  out[t] = s;
}

__global__ void fill_kernel(double* p, size_t count, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// Flags and counters are cleared with a kernel, not hipMemsetAsync, so that
// every launch sequence of a build can be captured into a HIP graph.
__global__ void fill_int_kernel(int* p, int count, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

__global__ void finite_kernel(const double* p, size_t count, int* flag) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count && !isfinite(p[i])) atomicOr(flag, 1);
}

// Molecule owning element r of a packed batch: the last m with base(m) <= r,
// where base is sbase (which = 0), cbase (1) or dbase (2).
__device__ __forceinline__ int find_mol(const MolAO* __restrict__ M, int nmol, int64_t r, int which) {
  int lo = 0, hi = nmol;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    const int64_t b = which == 0 ? M[mid].sbase : (which == 1 ? M[mid].cbase : M[mid].dbase);
    if (b <= r) lo = mid; else hi = mid;
  }
  return lo;
}

// Dc[b][P][Q] = sum T[p,P] T[q,Q] D[src b][p][q] (or D^T when transpose), per
// molecule; the maps hold molecule-local indices at global AO positions.
__global__ void dens_to_cart_kernel(const double* __restrict__ D, int64_t nn_tot, int nb, double* __restrict__ Dc,
                                    int64_t ncc_tot, const MolAO* __restrict__ M, int nmol,
                                    const int* __restrict__ idx, const double* __restrict__ coef, int nbd) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)nbd * ncc_tot;
  if (i >= total) return;
  const int bt = (int)(i / ncc_tot);
  const int64_t r = (int64_t)(i - (size_t)bt * ncc_tot);
  const MolAO m = M[find_mol(M, nmol, r, 1)];
  const int rem = (int)(r - m.cbase);
  const int P = rem / m.nc, Q = rem - (rem / m.nc) * m.nc;
  const bool tr = bt >= nb;
  const int n = m.n;
  const double* Db = D + (size_t)(tr ? bt - nb : bt) * nn_tot + m.sbase;
  const int gP = m.aoc0 + P, gQ = m.aoc0 + Q;
  double s = 0.0;
  for (int e1 = 0; e1 < 2; ++e1) {
    const double c1 = coef[gP * 2 + e1];
    if (c1 == 0.0) continue;
    const int p = idx[gP * 2 + e1];
    for (int e2 = 0; e2 < 2; ++e2) {
      const double c2 = coef[gQ * 2 + e2];
      if (c2 == 0.0) continue;
      const int q = idx[gQ * 2 + e2];
      s += c1 * c2 * (tr ? Db[(size_t)q * n + p] : Db[(size_t)p * n + q]);
    }
  }
  Dc[i] = s;
}

// out[op][b][p][q]: J = T (Jp + Jp^T) T^T, K = T (Kp_b + Kp_b'^T) T^T, per molecule
// K rows += their partial exchange rows: block (x: global Cartesian row R,
// y: 64-column range, z: density).  The slab rows that add into row R are
// listed per ket mask (rp over keys R * 8 + mask, block-uniform scalar loads):
// a slab row of mask `mask` holds the molecule's columns of that mask in
// compact order, so every thread looks up its column's compact index once
// per mask (cinv, -1 when the mask does not contain the column) and then
// reads eight slab rows per step so that many loads are in flight.
// row_dst[R] / row_n[R] / row_mol[R]: first element, length and molecule of
// row R of K.
__global__ __launch_bounds__(64) void kpart_reduce_kernel(const double* __restrict__ part, int64_t stride,
                                                          const int64_t* __restrict__ rp,
                                                          const int64_t* __restrict__ ent,
                                                          const int64_t* __restrict__ row_dst,
                                                          const int* __restrict__ row_n,
                                                          const int* __restrict__ row_mol,
                                                          const int64_t* __restrict__ cinv_off,
                                                          const int* __restrict__ cinv, int64_t ncc_tot,
                                                          int atomic, double* __restrict__ kp) {
  const int64_t R = blockIdx.x;
  const int bt = blockIdx.z;
  const int col = blockIdx.y * 64 + threadIdx.x;
  const int n = row_n[R];
  if (col >= n) return;
  const int m = row_mol[R];
  const double* __restrict__ pb = part + (size_t)bt * stride;
  double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  bool any = false;
  for (int mask = 1; mask < 8; ++mask) {
    const int64_t e0 = rp[R * 8 + mask], e1 = rp[R * 8 + mask + 1];
    if (e0 == e1) continue;
    const int j = cinv[cinv_off[m * 8 + mask] + col];
    if (j < 0) continue;
    any = true;
    int64_t e = e0;
    for (; e + 8 <= e1; e += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] += pb[ent[e + u] + j];
    }
    for (int u = 0; e < e1; ++e, ++u) v[u] += pb[ent[e] + j];
  }
  const double sum = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
  double* dst = kp + (size_t)bt * ncc_tot + row_dst[R] + col;
  if (!atomic) {
    *dst += sum;
  } else if (any) {
    atomicAdd(dst, sum);  // another stream may fold into the same row
  }
}

// Ket-side J blocks from the partial ket windows: one wavefront per (ket
// pair, density), lane e < block size; every pair owns its J block, so the
// adds do not race.
__global__ __launch_bounds__(64) void jpart_reduce_kernel(const double* __restrict__ part, int64_t stride,
                                                          const int64_t* __restrict__ rp,
                                                          const int64_t* __restrict__ ent,
                                                          const Pair* __restrict__ pairs,
                                                          const ShellXYZ* __restrict__ shells,
                                                          const MolInfo* __restrict__ mols, int64_t ncc_tot,
                                                          double* __restrict__ jpad) {
  const int64_t k = blockIdx.x;
  const int bt = blockIdx.y;
  const int e = threadIdx.x;
  const int64_t e0 = rp[k], e1 = rp[k + 1];
  if (e0 == e1) return;
  const Pair pr = pairs[k];
  const ShellXYZ C = shells[pr.sa], D = shells[pr.sb];
  const int wd = (D.l + 1) * (D.l + 2) / 2;
  if (e >= (C.l + 1) * (C.l + 2) / 2 * wd) return;
  const double* __restrict__ pb = part + (size_t)bt * stride + e;
  double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int64_t q = e0;
  for (; q + 8 <= e1; q += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] += pb[ent[q + u]];
  }
  for (int u = 0; q < e1; ++q, ++u) v[u] += pb[ent[q]];
  const MolInfo mi = mols[C.mol];
  jpad[8 * ((size_t)bt * ncc_tot + mi.base + (size_t)(C.ao + e / wd) * mi.n + D.ao + e % wd)] +=
      ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

__global__ void finalize_kernel(const double* __restrict__ jp, const double* __restrict__ jpad,
                                const double* __restrict__ kp, int64_t ncc_tot,
                                int64_t nn_tot, int nb, int sym, int with_j, int with_k,
                                const MolAO* __restrict__ M, int nmol, const int* __restrict__ idx,
                                const double* __restrict__ coef, double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nops = with_j + with_k;
  const size_t total = (size_t)nops * nb * nn_tot;
  if (i >= total) return;
  const int opb = (int)(i / nn_tot);
  const int64_t r = (int64_t)(i - (size_t)opb * nn_tot);
  const MolAO m = M[find_mol(M, nmol, r, 0)];
  const int rem = (int)(r - m.sbase);
  const int p = rem / m.n, q = rem - (rem / m.n) * m.n;
  const int op = opb / nb, b = opb - (opb / nb) * nb;
  const bool isj = with_j && op == 0;
  const double* A = (isj ? jp + (size_t)b * ncc_tot : kp + (size_t)b * ncc_tot) + m.cbase;
  const double* B = isj ? A : kp + (size_t)(sym ? b : nb + b) * ncc_tot + m.cbase;
  // ket-side J partials, one element per 64-byte line
  const double* X = isj ? jpad + 8 * ((size_t)b * ncc_tot + m.cbase) : nullptr;
  const int nc = m.nc, gp = m.ao0 + p, gq = m.ao0 + q;
  double s = 0.0;
  for (int e1 = 0; e1 < 3; ++e1) {
    const double c1 = coef[gp * 3 + e1];
    if (c1 == 0.0) continue;
    const int P = idx[gp * 3 + e1];
    for (int e2 = 0; e2 < 3; ++e2) {
      const double c2 = coef[gq * 3 + e2];
      if (c2 == 0.0) continue;
      const int Q = idx[gq * 3 + e2];
      double v = A[(size_t)P * nc + Q] + B[(size_t)Q * nc + P];
      if (X) v += X[8 * ((size_t)P * nc + Q)] + X[8 * ((size_t)Q * nc + P)];
      s += c1 * c2 * v;
    }
  }
  out[i] = s;
}

// out[p][q] = sum T[p,P] T[q,Q] M[P][Q] (Cartesian -> requested AO basis), per molecule
__global__ void sandwich_kernel(const double* __restrict__ Mc, int64_t nn_tot, const MolAO* __restrict__ M, int nmol,
                                const int* __restrict__ idx, const double* __restrict__ coef,
                                double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nn_tot) return;
  const MolAO m = M[find_mol(M, nmol, (int64_t)i, 0)];
  const int rem = (int)((int64_t)i - m.sbase);
  const int p = rem / m.n, q = rem - (rem / m.n) * m.n;
  const double* A = Mc + m.cbase;
  const int gp = m.ao0 + p, gq = m.ao0 + q;
  double s = 0.0;
  for (int e1 = 0; e1 < 3; ++e1) {
    const double c1 = coef[gp * 3 + e1];
    if (c1 == 0.0) continue;
    const int P = idx[gp * 3 + e1];
    for (int e2 = 0; e2 < 3; ++e2) {
      const double c2 = coef[gq * 3 + e2];
      if (c2 == 0.0) continue;
      s += c1 * c2 * A[(size_t)P * m.nc + idx[gq * 3 + e2]];
    }
  }
  out[i] = s;
}

// Quartet list of one sub-task: g -> (bra0 + b, ket0 + g - qoff[b]) with qoff[b] <= g < qoff[b+1].
// Quartets of one (bra run, ket run) sub-task: bra b keeps the ket-run prefix
// [0, qoff[b+1] - qoff[b]) (Schwarz-descending run).  The prefix is cut into
// segments of seg kets (the ket windows of single-bra chunks) and each segment
// is written in the run's primitive-count order: groups[k] = (begin, count) of
// group k in order[], whose entries are run positions, ascending inside a
// group, groups largest primitive count first.  Adjacent lanes then take kets
// of near-equal primitive counts, so their loop trip counts agree, and every
// segment keeps its set of kets.
__global__ void expand_kernel(const int64_t* __restrict__ qoff, int nbra, int bra0, int ket0, int64_t nq,
                              int seg, const int* __restrict__ order, const int2* __restrict__ groups,
                              int ngroups, int2* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nq) return;
  int lo = 0, hi = nbra;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (qoff[mid] <= g) lo = mid; else hi = mid;
  }
  const int cnt = (int)(qoff[lo + 1] - qoff[lo]);
  const int r = (int)(g - qoff[lo]);
  const int s0 = r - r % seg, s1 = (cnt - s0 > seg) ? s0 + seg : cnt;
  int rr = r - s0, ket = r;
  // the groups partition the run, so their members in [s0, s1) number exactly
  // s1 - s0 and rr < s1 - s0 lands in one of them
  for (int k = 0; k < ngroups; ++k) {
    const int2 gr = groups[k];
    int a = 0, b = gr.y;
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (order[gr.x + mid] < s0) a = mid + 1; else b = mid;
    }
    int c = a;
    b = gr.y;
    while (c < b) {
      const int mid = (c + b) >> 1;
      if (order[gr.x + mid] < s1) c = mid + 1; else b = mid;
    }
    if (rr < c - a) {
      ket = order[gr.x + a + rr];
      break;
    }
    rr -= c - a;
  }
  out[g] = make_int2(bra0 + lo, ket0 + ket);
}

// Order of the quartets inside every bra-major chunk: by ket primitive-pair
// count (descending), then by the squared distance between the bra and ket
// pair midpoints (ascending).  Lanes of a wavefront then run the same number
// of ket primitive iterations, and their Boys arguments T = rho |P - Q|^2
// stay close, so the wavefront mostly takes one branch of the Boys
// evaluation (grid expansion below the switch, asymptotic form above) and
// reads neighbouring grid rows.  One workgroup per chunk: 64-bit keys
// (count, distance bits, position) bitonic-sorted in LDS, then the quartets
// gathered through a scratch copy.  nq of a chunk is at most kSortMax.
constexpr int kSortMax = 8192;
__global__ __launch_bounds__(256) void sort_chunk_kernel(const BraChunk* __restrict__ chunks,
                                                         const Pair* __restrict__ pairs,
                                                         const ShellXYZ* __restrict__ shells,
                                                         const int2* __restrict__ src, int2* __restrict__ dst) {
  __shared__ uint64_t key[kSortMax];
  const BraChunk c = chunks[blockIdx.x];
  const int nq = c.nq;
  int np2 = 1;
  while (np2 < nq) np2 <<= 1;
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    uint64_t k = ~0ull;
    if (i < nq) {
      const int2 q = src[c.q0 + i];
      const Pair b = pairs[q.x], t = pairs[q.y];
      const ShellXYZ A = shells[b.sa], B = shells[b.sb], C = shells[t.sa], D = shells[t.sb];
      const double dx = 0.5 * (A.x + B.x - C.x - D.x), dy = 0.5 * (A.y + B.y - C.y - D.y),
                   dz = 0.5 * (A.z + B.z - C.z - D.z);
      const float d2 = (float)(dx * dx + dy * dy + dz * dz);
      const uint32_t fb = __float_as_uint(d2) >> 1;  // non-negative: bits order as the values
      const uint32_t npk = 65535u - (uint32_t)min(t.np, 65535);
      k = ((uint64_t)npk << 45) | ((uint64_t)fb << 14) | (uint64_t)i;
    }
    key[i] = k;
  }
  __syncthreads();
  for (int size = 2; size <= np2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < np2 / 2; i += blockDim.x) {
        const int lo = 2 * i - (i & (stride - 1));
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t a = key[lo], b = key[hi];
        if ((a > b) == up) {
          key[lo] = b;
          key[hi] = a;
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < nq; i += blockDim.x) dst[c.q0 + i] = src[c.q0 + (int)(key[i] & 0x3fff)];
}

// dmax[m][S][T] = max over batches and over the Cartesian block (S, T) of |D_m|
// (S, T segment indices inside molecule m).
__global__ void dmax_kernel(const double* __restrict__ D, int64_t ncc_tot, int nbd, const MolAO* __restrict__ M,
                            int nmol, int64_t nd_tot, const int* __restrict__ seg_ao, double* __restrict__ dmax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nd_tot) return;
  const MolAO mo = M[find_mol(M, nmol, i, 2)];
  const int r = (int)(i - mo.dbase);
  const int S = mo.seg0 + r / mo.nseg, T = mo.seg0 + r % mo.nseg;
  const int oS = seg_ao[2 * S], nS = seg_ao[2 * S + 1], oT = seg_ao[2 * T], nT = seg_ao[2 * T + 1];
  const int nc = mo.nc;
  double m = 0.0;
  for (int b = 0; b < nbd; ++b) {
    const double* Db = D + (size_t)b * ncc_tot + mo.cbase;
    for (int p = 0; p < nS; ++p)
      for (int q = 0; q < nT; ++q) m = fmax(m, fabs(Db[(size_t)(oS + p) * nc + oT + q]));
  }
  dmax[i] = m;
}

// Fixed-point scale of a deterministic build, one workgroup per molecule m:
// every J/K partial element of m is bounded by
//   B_J = Qmax sum_{S,T} Q(S,T) nS nT max|D_ST|       (Coulomb rows)
//   B_K = (max_S sum_T Q(S,T) nT)^2 max|D|            (exchange rows)
// (Cauchy-Schwarz |(ab|cd)| <= Q_ab Q_cd over Cartesian blocks); with 2^e >
// 8 (B_J + B_K) the scale 2^(60-e) keeps every final integer below 2^60.  The
// sums run in a fixed order (strided partial sums, then a fixed tree), so the
// scale, like the build, does not depend on scheduling.
__global__ __launch_bounds__(256) void det_scale_kernel(const double* __restrict__ dmax,
                                                        const double* __restrict__ qblk,
                                                        const MolAO* __restrict__ M,
                                                        const double* __restrict__ mol,
                                                        double* __restrict__ scale) {
  __shared__ double sj[256], sm[256];
  const int m = blockIdx.x;
  const MolAO mo = M[m];
  const int64_t cnt = (int64_t)mo.nseg * mo.nseg;
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < cnt; i += blockDim.x) {
    const double d = dmax[mo.dbase + i];
    a += qblk[mo.dbase + i] * d;
    b = fmax(b, d);
  }
  sj[threadIdx.x] = a;
  sm[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sj[threadIdx.x] += sj[threadIdx.x + w];
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double bound = 8.0 * (mol[2 * m] * sj[0] + mol[2 * m + 1] * mol[2 * m + 1] * sm[0]);
    int e = (bound > 0.0 && isfinite(bound)) ? ilogb(bound) + 1 : 0;
    e = max(-900, min(900, e));
    scale[2 * m] = ldexp(1.0, 60 - e);
    scale[2 * m + 1] = ldexp(1.0, e - 60);
  }
}

// Deterministic builds: p[i] (a 64-bit integer, element `i` of a packed
// buffer whose density stride is ncc_tot, every `step`-th slot used) back to
// FP64 with its molecule's 1/scale.
__global__ void fixed_to_double_kernel(double* __restrict__ p, int64_t count, int step, int64_t ncc_tot,
                                       const MolAO* __restrict__ M, int nmol, const double* __restrict__ scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t e = (i % (ncc_tot * step)) / step;
  const int m = find_mol(M, nmol, e, 1);
  const long long v = __double_as_longlong(p[i]);
  p[i] = (double)v * scale[2 * m + 1];
}

static unsigned nblocks(size_t count) { return (unsigned)((count + 255) / 256); }

// env_grad[exponent slot] += gexp; env_grad[coefficient slot] += fsp * gcoef
// (the exponent slot is shared by the contraction columns of a shell).
__global__ void param_to_env_kernel(int nprim, const int2* __restrict__ slots, const double* __restrict__ fsp,
                                    const double* __restrict__ gcoef, const double* __restrict__ gexp,
                                    double* __restrict__ env_grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nprim) return;
  const int2 sl = slots[i];
  if (gexp[i] != 0.0) atomicAdd(env_grad + sl.x, gexp[i]);
  env_grad[sl.y] += fsp[i] * gcoef[i];
}

// ----------------------------------------------------------------------------
// plan construction
// ----------------------------------------------------------------------------
// mol_atm / mol_bas (nmol + 1 offsets each, NULL for one molecule) split the
// concatenated tables into molecules; every shell must sit on an atom of its
// own molecule.
static void validate_tables(PlanData& P, const int32_t* atm, int natm, const int32_t* bas, int nbas,
                            const double* env, int nenv, int cart, int nmol = 1, const int32_t* mol_atm = nullptr,
                            const int32_t* mol_bas = nullptr, int max_l = kMaxL) {
  if (!atm || !bas || !env || natm < 0 || nbas < 0 || nenv < 20 || (cart != 0 && cart != 1))
    throw InvalidArgument("Invalid libcint table dimensions or operator");
  ensure_tuning();
  std::copy(g_tuning, g_tuning + kNumQuartetClasses, P.tuning);
  P.options = g_options;
  if (nmol < 1 || (nmol > 1 && (!mol_atm || !mol_bas)))
    throw InvalidArgument("A batched plan needs per-molecule atm and bas offsets");
  P.nmol = nmol;
  P.mol_atm.assign(nmol + 1, 0);
  P.mol_bas.assign(nmol + 1, 0);
  if (mol_atm && mol_bas) {
    P.mol_atm.assign(mol_atm, mol_atm + nmol + 1);
    P.mol_bas.assign(mol_bas, mol_bas + nmol + 1);
  } else {
    P.mol_atm[1] = natm;
    P.mol_bas[1] = nbas;
  }
  if (P.mol_atm[0] != 0 || P.mol_bas[0] != 0 || P.mol_atm[nmol] != natm || P.mol_bas[nmol] != nbas)
    throw InvalidArgument("Molecule offsets must start at 0 and end at natm / nbas");
  for (int m = 0; m < nmol; ++m)
    if (P.mol_atm[m + 1] < P.mol_atm[m] || P.mol_bas[m + 1] <= P.mol_bas[m])
      throw InvalidArgument("Every molecule needs a non-decreasing atom range and at least one shell");
  P.natm = natm;
  P.nbas = nbas;
  P.cart = cart;
  P.atm.assign(atm, atm + (size_t)natm * ATM_SLOTS);
  P.bas.assign(bas, bas + (size_t)nbas * BAS_SLOTS);
  P.env.assign(env, env + nenv);
  for (double v : P.env)
    if (!std::isfinite(v)) throw InvalidArgument("Native integral env must be finite");
  if (P.env[PTR_RANGE_OMEGA] > std::sqrt(std::numeric_limits<double>::max()) / 2.)
    throw InvalidArgument("omega exceeds the safe float64 squared range");
  if (P.env[PTR_RANGE_OMEGA] != 0.)
    throw InvalidArgument("Range-separated integrals require the full ERI path and nonnegative omega");
  auto inside = [&](int ptr, int64_t count) {
    return ptr >= PTR_ENV_START && count > 0 && int64_t(ptr) + count <= int64_t(nenv);
  };
  for (int a = 0; a < natm; ++a) {
    const int* row = P.atm.data() + a * ATM_SLOTS;
    if (!inside(row[PTR_COORD], 3) || row[NUC_MOD_OF] != 1)
      throw InvalidArgument("Native backend requires valid point nuclei");
  }
  P.ao.assign(nbas + 1, 0);
  P.ao_cart.assign(nbas + 1, 0);
  for (int m = 0; m < nmol; ++m)
    for (int s = P.mol_bas[m]; s < P.mol_bas[m + 1]; ++s) {
      const int* row = P.bas.data() + s * BAS_SLOTS;
      const int l = row[ANG_OF], np = row[NPRIM_OF], nc = row[NCTR_OF];
      if (row[ATOM_OF] < P.mol_atm[m] || row[ATOM_OF] >= P.mol_atm[m + 1] || l < 0 || l > 12 || np < 1 ||
          np > 64 || nc < 1 || nc > 64 || row[KAPPA_OF] != 0 || !inside(row[PTR_EXP], np) ||
          !inside(row[PTR_COEFF], int64_t(np) * nc))
        throw InvalidArgument("Invalid or unsupported libcint basis table");
      for (int p = 0; p < np; ++p)
        if (P.env[row[PTR_EXP] + p] <= 0.) throw InvalidArgument("Gaussian exponents must be positive");
      if (l > max_l)
        throw InvalidArgument(max_l == kMaxL ? "The MI355X J/K engine supports shells up to d (l <= 2)"
                                             : "Auxiliary shells are supported up to g (l <= 4)");
      P.ao[s + 1] = P.ao[s] + nc * (cart ? ncart_h(l) : 2 * l + 1);
      P.ao_cart[s + 1] = P.ao_cart[s] + nc * ncart_h(l);
    }
  P.n = P.ao.back();
  P.n_cart = P.ao_cart.back();
  if (P.n < 1) throw InvalidArgument("Invalid AO count");
  // per-molecule matrix layout (segment fields are filled by build_segments)
  P.mols.assign(nmol + 1, MolAO{});
  int64_t sb = 0, cb = 0;
  P.nc_max = 0;
  for (int m = 0; m < nmol; ++m) {
    MolAO& M = P.mols[m];
    M.ao0 = P.ao[P.mol_bas[m]];
    M.aoc0 = P.ao_cart[P.mol_bas[m]];
    M.n = P.ao[P.mol_bas[m + 1]] - M.ao0;
    M.nc = P.ao_cart[P.mol_bas[m + 1]] - M.aoc0;
    M.sbase = sb;
    M.cbase = cb;
    sb += (int64_t)M.n * M.n;
    cb += (int64_t)M.nc * M.nc;
    P.nc_max = std::max(P.nc_max, M.nc);
  }
  // the quartet kernels address one molecule's matrices with 32-bit offsets
  if (P.nc_max > kMaxMoleculeCartAO)
    throw InvalidArgument("A molecule has more than 16384 Cartesian AOs; the J/K engine addresses one "
                          "molecule's matrices with 32-bit offsets");
  P.mols[nmol].sbase = sb;
  P.mols[nmol].cbase = cb;
  P.nn_tot = sb;
  P.ncc_tot = cb;
}

static void build_segments(PlanData& P) {
  const double pi = 3.14159265358979323846;
  P.mol_seg.assign(P.nmol + 1, 0);
  int64_t db = 0;
  for (int m = 0; m < P.nmol; ++m) {
    P.mol_seg[m] = (int)P.seg.size();
    for (int s = P.mol_bas[m]; s < P.mol_bas[m + 1]; ++s) {
      const int* row = P.bas.data() + s * BAS_SLOTS;
      const int l = row[ANG_OF], np = row[NPRIM_OF], nc = row[NCTR_OF];
      const double fsp = l == 0 ? 0.5 / std::sqrt(pi) : (l == 1 ? std::sqrt(3.0 / (4.0 * pi)) : 1.0);
      const double* r = P.env.data() + P.atm[row[ATOM_OF] * ATM_SLOTS + PTR_COORD];
      for (int c = 0; c < nc; ++c) {
        Seg g;
        g.parent = s;
        g.ctr = c;
        g.l = l;
        g.np = np;
        g.r[0] = r[0]; g.r[1] = r[1]; g.r[2] = r[2];
        g.ao_cart = P.ao_cart[s] - P.mols[m].aoc0 + c * ncart_h(l);
        g.mol = m;
        g.lseg = (int)P.seg.size() - P.mol_seg[m];
        for (int p = 0; p < np; ++p) {
          g.a.push_back(P.env[row[PTR_EXP] + p]);
          g.c.push_back(P.env[row[PTR_COEFF] + c * np + p] * fsp);
        }
        P.seg.push_back(std::move(g));
      }
    }
    MolAO& M = P.mols[m];
    M.seg0 = P.mol_seg[m];
    M.nseg = (int)P.seg.size() - M.seg0;
    M.dbase = db;
    db += (int64_t)M.nseg * M.nseg;
  }
  P.mol_seg[P.nmol] = (int)P.seg.size();
  P.mols[P.nmol].seg0 = (int)P.seg.size();
  P.mols[P.nmol].dbase = db;
  P.nd_tot = db;
}

// all_pairs index of the intra-molecule segment pair (S, T), S >= T.
static int64_t pair_idx(const PlanData& P, int S, int T) {
  const int m = P.seg[S].mol;
  const int64_t s = S - P.mol_seg[m], t = T - P.mol_seg[m];
  return P.mol_pair0[m] + s * (s + 1) / 2 + t;
}

// Primitive pairs of segmented shells (A, B) with l(A) >= l(B).  Pairs whose
// Gaussian-product factor exp(-mu R^2) < e^{-60} are dropped (the primitive
// screening threshold libcint applies by default, g2e.c:57-62).
static void make_pair(PlanData& P, int sa, int sb, Pair& out) {
  const Seg& A = P.seg[sa];
  const Seg& B = P.seg[sb];
  const double pi = 3.14159265358979323846;
  const double s2p = std::sqrt(2.0 * std::pow(pi, 2.5));
  double ab2 = 0;
  for (int x = 0; x < 3; ++x) ab2 += (A.r[x] - B.r[x]) * (A.r[x] - B.r[x]);
  out.sa = sa;
  out.sb = sb;
  out.poff = (int)P.prims.size();
  out.np = 0;
  for (int i = 0; i < A.np; ++i)
    for (int j = 0; j < B.np; ++j) {
      const double ai = A.a[i], aj = B.a[j], p = ai + aj;
      const double mu = ai * aj / p * ab2;
      if (mu > 60.0) continue;
      PrimPair pp;
      pp.p = p;
      pp.px = (ai * A.r[0] + aj * B.r[0]) / p;
      pp.py = (ai * A.r[1] + aj * B.r[1]) / p;
      pp.pz = (ai * A.r[2] + aj * B.r[2]) / p;
      pp.e = A.c[i] * B.c[j] * std::exp(-mu) * s2p / p;
      pp.aa = ai;
      pp.ab = aj;
      pp.ip = 1.0 / p;
      P.prims.push_back(pp);
      PrimAux x;
      x.ia = i;
      x.ib = j;
      x.ca = A.c[i];
      x.cb = B.c[j];
      x.ek = std::exp(-mu) * s2p / p;
      P.aux.push_back(x);
      ++out.np;
    }
}

// Runs the block kernel over tasks grouped by class; tasks must be class-sorted.
static void run_blocks(PlanData& P, const Pair* d_pairs, const std::vector<BlockTask>& tasks,
                       const std::vector<std::array<int, 4>>& task_class, double* d_out, hipStream_t s) {
  if (tasks.empty()) return;
  BlockTask* d_tasks = dev_upload(tasks, s);
  size_t i = 0;
  while (i < tasks.size()) {
    size_t j = i;
    while (j < tasks.size() && task_class[j] == task_class[i]) ++j;
    BlockArgs a;
    a.pairs = d_pairs;
    a.prims = P.d_prims;
    a.shells = P.d_shells;
    a.tasks = d_tasks + i;
    a.ntasks = (int64_t)(j - i);
    a.out = d_out;
    launch_blocks_class(task_class[i][0], task_class[i][1], task_class[i][2], task_class[i][3], a, s);
    i = j;
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(s));
  HIP_OK(hipFree(d_tasks));
}

static void compute_schwarz(PlanData& P, std::vector<double>& seg_q, hipStream_t s,
                            std::vector<double>* cart_q = nullptr) {
  const int npair = (int)P.all_pairs.size();
  // tasks sorted by pair class
  std::vector<int> order(npair);
  for (int i = 0; i < npair; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return P.all_class[x] < P.all_class[y]; });
  std::vector<BlockTask> tasks;
  std::vector<std::array<int, 4>> tcls;
  std::vector<int64_t> base(npair);
  int64_t total = 0;
  for (int i : order) {
    const Pair& pr = P.all_pairs[i];
    const int la = P.seg[pr.sa].l, lb = P.seg[pr.sb].l;
    const int na = ncart_h(la), nb = ncart_h(lb);
    BlockTask t;
    t.bra = t.ket = i;
    t.base = total;
    base[i] = total;
    t.stride[0] = 1;
    t.stride[1] = na;
    t.stride[2] = na * nb;
    t.stride[3] = na * nb * na;
    for (int k = 0; k < 4; ++k) t.coff[k] = 0;
    tasks.push_back(t);
    tcls.push_back({la, lb, la, lb});
    total += (int64_t)na * nb * na * nb;
  }
  double* d_out = nullptr;
  HIP_OK(hipMalloc(&d_out, sizeof(double) * std::max<int64_t>(total, 1)));
  run_blocks(P, P.d_all_pairs, tasks, tcls, d_out, s);
  std::vector<double> h(total);
  HIP_OK(hipMemcpy(h.data(), d_out, sizeof(double) * total, hipMemcpyDeviceToHost));
  HIP_OK(hipFree(d_out));
  seg_q.assign(npair, 0.0);
  if (cart_q) cart_q->assign(npair, 0.0);
  for (int i = 0; i < npair; ++i) {
    const Pair& pr = P.all_pairs[i];
    const int la = P.seg[pr.sa].l, lb = P.seg[pr.sb].l;
    const int na = ncart_h(la), nb = ncart_h(lb);
    const double* B = h.data() + base[i];
    auto at = [&](int a, int b, int c, int d) { return B[((d * na + c) * nb + b) * na + a]; };
    double largest = 0.0;
    if (cart_q) {
      double lc = 0.0;
      for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) lc = std::max(lc, std::fabs(at(a, b, a, b)));
      (*cart_q)[i] = lc;
    }
    if (P.cart) {
      for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) largest = std::max(largest, std::fabs(at(a, b, a, b)));
    } else {
      std::vector<std::vector<double>> ta, tb;
      c2s_rows(la, ta);
      c2s_rows(lb, tb);
      for (size_t sa = 0; sa < ta.size(); ++sa)
        for (size_t sb = 0; sb < tb.size(); ++sb) {
          double v = 0.0;
          for (int a1 = 0; a1 < na; ++a1) {
            if (ta[sa][a1] == 0) continue;
            for (int a2 = 0; a2 < na; ++a2) {
              if (ta[sa][a2] == 0) continue;
              for (int b1 = 0; b1 < nb; ++b1) {
                if (tb[sb][b1] == 0) continue;
                for (int b2 = 0; b2 < nb; ++b2) {
                  if (tb[sb][b2] == 0) continue;
                  v += ta[sa][a1] * ta[sa][a2] * tb[sb][b1] * tb[sb][b2] * at(a1, b1, a2, b2);
                }
              }
            }
          }
          largest = std::max(largest, std::fabs(v));
        }
    }
    seg_q[i] = largest;
  }
}

static void build_c2s_maps(PlanData& P, hipStream_t s) {
  std::vector<int> cidx((size_t)P.n_cart * 2, 0), sidx((size_t)P.n * 3, 0);
  std::vector<double> ccoef((size_t)P.n_cart * 2, 0.0), scoef((size_t)P.n * 3, 0.0);
  // positions are global AO indices, stored indices are local to the molecule
  for (int mo = 0; mo < P.nmol; ++mo)
  for (int sh = P.mol_bas[mo]; sh < P.mol_bas[mo + 1]; ++sh) {
    const int l = P.bas[sh * BAS_SLOTS + ANG_OF], nc = P.bas[sh * BAS_SLOTS + NCTR_OF];
    const int aoc0 = P.mols[mo].aoc0, ao0 = P.mols[mo].ao0;
    const int nf = ncart_h(l);
    const int ns = P.cart ? nf : 2 * l + 1;
    std::vector<std::vector<double>> T;
    if (P.cart) {
      T.assign(nf, std::vector<double>(nf, 0.0));
      for (int i = 0; i < nf; ++i) T[i][i] = 1.0;
    } else {
      c2s_rows(l, T);
    }
    for (int c = 0; c < nc; ++c) {
      const int cb = P.ao_cart[sh] + c * nf, sbase = P.ao[sh] + c * ns;
      for (int m = 0; m < ns; ++m) {
        int k = 0;
        for (int f = 0; f < nf; ++f)
          if (T[m][f] != 0.0) {
            if (k >= 3) throw Internal("c2s map overflow");
            sidx[(size_t)(sbase + m) * 3 + k] = cb + f - aoc0;
            scoef[(size_t)(sbase + m) * 3 + k] = T[m][f];
            ++k;
          }
      }
      for (int f = 0; f < nf; ++f) {
        int k = 0;
        for (int m = 0; m < ns; ++m)
          if (T[m][f] != 0.0) {
            if (k >= 2) throw Internal("c2s map overflow");
            cidx[(size_t)(cb + f) * 2 + k] = sbase + m - ao0;
            ccoef[(size_t)(cb + f) * 2 + k] = T[m][f];
            ++k;
          }
      }
    }
  }
  P.d_c2s_idx = dev_upload(cidx, s);
  P.d_c2s_coef = dev_upload(ccoef, s);
  P.d_s2c_idx = dev_upload(sidx, s);
  P.d_s2c_coef = dev_upload(scoef, s);
}

static int bucket_of(int np) {
  int b = 0;
  while ((2 << b) <= np) ++b;
  return b;
}

struct PairCand {
  Pair pr;
  int cls, pat, mol;
  double q;
};

// Screen candidate pairs (pair prune Q_ij Qmax < cutoff), sort them by (pair
// class, molecule, column pattern, primitive-count bucket, Schwarz
// descending) and organise canonical quartets: for every (bra run, ket run)
// of one molecule a sub-task whose surviving kets form a prefix of the ket
// run; one QuartetClass (one launch) per (bra class, ket class, bra pattern,
// ket pattern); bra chunks for the LDS kernel when with_chunks.
static void build_pair_set(PlanData& P, PairSet& S, std::vector<PairCand>& cand, bool with_chunks, hipStream_t s) {
  const double cutoff = P.cutoff;
  // slab column spaces per (molecule, ket l mask); segmented sets only (the
  // contracted sets flush exchange rows by atomics)
  std::vector<int> colmap, cinv, shell_cofs;
  std::vector<int64_t> cinv_off((size_t)P.nmol * 8, 0);
  S.h_cspace.assign((size_t)P.nmol * 8, make_int2(0, 0));
  if (!S.contracted) {
    shell_cofs.assign(P.seg.size() * 8, -1);
    for (int m = 0; m < P.nmol; ++m)
      for (int mask = 1; mask < 8; ++mask) {
        const int nc = P.mols[m].nc;
        S.h_cspace[m * 8 + mask].x = (int)colmap.size();
        cinv_off[m * 8 + mask] = (int64_t)cinv.size();
        cinv.resize(cinv.size() + nc, -1);
        int cnt = 0;
        for (int g = P.mol_seg[m]; g < P.mol_seg[m + 1]; ++g) {
          const Seg& sg = P.seg[g];
          if (!((mask >> sg.l) & 1)) continue;
          shell_cofs[(size_t)g * 8 + mask] = cnt;
          for (int c = 0; c < ncart_h(sg.l); ++c) {
            colmap.push_back(sg.ao_cart + c);
            cinv[cinv_off[m * 8 + mask] + sg.ao_cart + c] = cnt++;
          }
        }
        S.h_cspace[m * 8 + mask].y = cnt;
      }
  }
  std::vector<int> keep;
  for (int i = 0; i < (int)cand.size(); ++i) {
    if (cand[i].pr.np == 0) continue;
    if (cutoff > 0 && cand[i].q * P.qmax_mol[cand[i].mol] < cutoff) continue;
    keep.push_back(i);
  }
  std::stable_sort(keep.begin(), keep.end(), [&](int x, int y) {
    const PairCand& a = cand[x];
    const PairCand& b = cand[y];
    if (a.cls != b.cls) return a.cls < b.cls;
    if (a.mol != b.mol) return a.mol < b.mol;
    if (a.pat != b.pat) return a.pat < b.pat;
    const int bx = bucket_of(a.pr.np), by = bucket_of(b.pr.np);
    if (bx != by) return bx > by;
    return a.q > b.q;
  });
  S.pairs.clear();
  S.pair_q.clear();
  for (int i : keep) {
    S.pairs.push_back(cand[i].pr);
    S.pair_q.push_back(cand[i].q);
  }
  struct Run { int cls, mol, pat, bkt, beg, end, npmin, npmax; };
  std::vector<Run> runs;
  for (int k = 0; k < (int)keep.size();) {
    const PairCand& c0 = cand[keep[k]];
    Run r{c0.cls, c0.mol, c0.pat, bucket_of(c0.pr.np), k, k, 1 << 30, 0};
    while (k < (int)keep.size()) {
      const PairCand& c = cand[keep[k]];
      if (c.cls != r.cls || c.mol != r.mol || c.pat != r.pat || bucket_of(c.pr.np) != r.bkt) break;
      r.npmin = std::min(r.npmin, c.pr.np);
      r.npmax = std::max(r.npmax, c.pr.np);
      ++k;
    }
    r.end = k;
    runs.push_back(r);
  }
  // ket order of every run for the quartet expansion: run positions grouped by
  // primitive count (1/16-octave groups, largest first), ascending inside a
  // group, so that every group keeps the run's Schwarz order and a bra's
  // surviving prefix is a prefix of each group
  std::vector<int> run_order(S.pairs.size());
  std::vector<int2> run_groups;
  std::vector<int> run_g0(runs.size() + 1, 0);
  for (size_t ri = 0; ri < runs.size(); ++ri) {
    const Run& r = runs[ri];
    const int len = r.end - r.beg;
    std::vector<int> key(len);
    for (int j = 0; j < len; ++j) key[j] = (int)std::floor(16.0 * std::log2((double)S.pairs[r.beg + j].np));
    int* ord = run_order.data() + r.beg;
    for (int j = 0; j < len; ++j) ord[j] = j;
    std::stable_sort(ord, ord + len, [&](int x, int y) { return key[x] > key[y]; });
    run_g0[ri] = (int)run_groups.size();
    for (int j = 0; j < len;) {
      int e = j;
      while (e < len && key[ord[e]] == key[ord[j]]) ++e;
      run_groups.push_back(make_int2(j, e - j));
      j = e;
    }
  }
  run_g0[runs.size()] = (int)run_groups.size();
  if (run_order.empty()) run_order.push_back(0);
  if (run_groups.empty()) run_groups.push_back(make_int2(0, 0));
  // runs of each (class, molecule) in list order, for the intra-molecule pairing
  std::map<std::pair<int, int>, std::vector<int>> runs_of;
  for (int i = 0; i < (int)runs.size(); ++i) runs_of[{runs[i].cls, runs[i].mol}].push_back(i);
  S.d_pairs = dev_upload(S.pairs, s);
  int* d_run_order = dev_upload(run_order, s);
  int2* d_run_groups = dev_upload(run_groups, s);
  {
    std::vector<double> pq = S.pair_q;
    if (pq.empty()) pq.push_back(0.0);
    S.d_pair_q = dev_upload(pq, s);
  }
  static const int pl[kNumPairClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
  S.nquartets = 0;
  S.part_total = 0;
  std::vector<std::pair<int64_t, int64_t>> kent, jent;  // (key, slab offset)
  std::vector<int> kent_cls;                              // class of each kent entry
  for (int cb = 0; cb < kNumPairClasses; ++cb)
    for (int ck = 0; ck <= cb; ++ck)
      for (int pb = 0; pb < 4; ++pb)
        for (int pk = 0; pk < 4; ++pk) {
          if (!pattern_valid(pl[cb][0], pl[cb][1], pb) || !pattern_valid(pl[ck][0], pl[ck][1], pk)) continue;
          if (!S.contracted && (pb || pk)) continue;
          QuartetClass qc;
          qc.la = pl[cb][0]; qc.lb = pl[cb][1]; qc.lc = pl[ck][0]; qc.ld = pl[ck][1];
          qc.pb = pb;
          qc.pk = pk;
          qc.tuning_index = cb * (cb + 1) / 2 + ck;
          const gscf_class_tuning& tu = P.tuning[qc.tuning_index];
          const int64_t kTargetPrim = tu.target_prim, kChunkItems = tu.chunk_items;
          const int kChunkBrasMax = tu.bras_max;
          qc.nq = 0;
          qc.nitems = 0;
          struct Pending { std::vector<int64_t> qoff; int bra0, ket0, nbra; int64_t q0, nq; int rk, seg; };
          std::vector<Pending> pend;
          for (int rbi = 0; rbi < (int)runs.size(); ++rbi) {
            const Run& rb = runs[rbi];
            if (rb.cls != cb || rb.pat != pb) continue;
            auto it = runs_of.find({ck, rb.mol});
            if (it == runs_of.end()) continue;
            for (int rki : it->second) {
              const Run& rk = runs[rki];
              if (rk.pat != pk) continue;
              // canonical: within one class each unordered pair of runs once (bra run
              // not after the ket run in list order), a run with itself as a triangle
              if (cb == ck && rbi > rki) continue;
              const bool same_list = rbi == rki;
              const int nbra = rb.end - rb.beg, nket = rk.end - rk.beg;
              std::vector<int64_t> qoff(nbra + 1, 0);
              for (int b = 0; b < nbra; ++b) {
                int cnt = nket;
                if (cutoff > 0) {
                  const double qb = S.pair_q[rb.beg + b];
                  int lo = 0, hi = nket;
                  while (lo < hi) {
                    const int mid = (lo + hi) / 2;
                    if (qb * S.pair_q[rk.beg + mid] < cutoff) hi = mid; else lo = mid + 1;
                  }
                  cnt = lo;
                }
                if (same_list) cnt = std::min(cnt, b + 1);
                qoff[b + 1] = qoff[b] + cnt;
              }
              const int64_t nq = qoff[nbra];
              if (nq == 0) continue;
              SubTask t;
              t.q0 = qc.nq;
              t.nq = nq;
              const int64_t work = (int64_t)rb.npmax * rk.npmax;
              // primitive split: a power of two <= 64 (the bra-major kernel
              // sums the S slices of a quartet across adjacent lanes)
              const int64_t sreq = std::max<int64_t>(1, std::min<int64_t>(rb.npmin, (work + kTargetPrim - 1) / kTargetPrim));
              t.S = 1;
              while (t.S < 64 && 2 * t.S <= sreq) t.S *= 2;
              t.pad = 0;
              t.item0 = qc.nitems;
              qc.nitems += (int64_t)t.S * nq;
              {
                // rough work model: primitive quartets weighted by the class's
                // Cartesian block size, plus a per-quartet digestion term
                const double nblk = (double)ncart_h(qc.la) * ncart_h(qc.lb) * ncart_h(qc.lc) * ncart_h(qc.ld);
                const double cols = (double)(1 + (pb & 1)) * (1 + (pb >> 1)) * (1 + (pk & 1)) * (1 + (pk >> 1));
                const double npq = 0.25 * (double)(rb.npmin + rb.npmax) * (rk.npmin + rk.npmax);
                qc.cost += (double)nq * (npq * (8.0 + 0.5 * nblk + 0.25 * nblk * (cols - 1.0)) + 4.0 * nblk * cols + 64.0);
              }
              qc.nq += nq;
              qc.tasks.push_back(t);
              pend.push_back({std::move(qoff), rb.beg, rk.beg, nbra, t.q0, nq, rki, std::numeric_limits<int>::max()});
            }
          }
          if (qc.nq == 0) continue;
          if (with_chunks) {
            // bra-major chunks: consecutive bras of one sub-task, bounded LDS slots and lane-items;
            // widths count every contraction column of a shell
            const int wa = ncart_h(qc.la) * (1 + (pb & 1)), wb = ncart_h(qc.lb) * (1 + (pb >> 1));
            const int wc = ncart_h(qc.lc) * (1 + (pk & 1)), wd = ncart_h(qc.ld) * (1 + (pk >> 1));
            const int rows = wa + wb;
            const int split = (wa * wb * wc * wd > 36) ? wc * wd : 1;
            // full-block classes stage their chunk's bra primitive pairs in
            // LDS (kStagedPrim doubles each, kStagedPrimCoef with bra columns)
            const bool stage = split == 1;
            const int rec = pb ? kStagedPrimCoef : kStagedPrim;
            int bra_np_max = 0;
            if (stage)
              for (const Run& r : runs)
                if (r.cls == cb && r.pat == pb) bra_np_max = std::max(bra_np_max, r.npmax);
            // exchange rows in LDS span the columns the class's kets reach
            // (segmented sets; contracted sets keep whole rows)
            int row_len = P.nc_max;
            if (!S.contracted) {
              const int kmask = (1 << pl[ck][0]) | (1 << pl[ck][1]);
              row_len = 0;
              for (int m = 0; m < P.nmol; ++m) row_len = std::max(row_len, S.h_cspace[m * 8 + kmask].y);
            }
            qc.row_len = row_len;
            const size_t per_slot = sizeof(double) * ((size_t)rows * row_len + (size_t)wa * wb * (kJBraCopies + 1) +
                                                      (size_t)rec * bra_np_max + 1);
            const int nr = (qc.la + qc.lb + qc.lc + qc.ld) / 2 + 1;
            size_t pack = std::min<size_t>((size_t)tu.lds_pack, P.lds_budget);
            if (tu.rys_lds && per_slot + rys_slice_bytes(nr) <= P.lds_budget) {
              qc.tab_lds = 1;
              qc.tab_bytes = rys_slice_bytes(nr);
              pack = pack > qc.tab_bytes + per_slot ? pack - qc.tab_bytes : per_slot;
            }
            qc.per_slot = per_slot;
            // a slot wider than the packing budget still gets chunks of one
            // bra when one density's slot fits the device's LDS
            qc.nbra_max = per_slot > P.lds_max ? 0
                          : (int)std::max<size_t>(1, std::min<size_t>((size_t)kChunkBrasMax, pack / per_slot));
            for (size_t ti = 0; qc.nbra_max >= 1 && ti < pend.size(); ++ti) {
              Pending& pd = pend[ti];
              const SubTask& t = qc.tasks[ti];
              const int64_t per_q = (int64_t)t.S * split;
              // a bra longer than one chunk is cut into windows of step kets;
              // shorter bras fit one window (cnt <= step)
              pd.seg = (int)std::min<int64_t>(std::numeric_limits<int>::max(),
                                              std::max<int64_t>(1, kChunkItems / per_q));
              int b = 0;
              while (b < pd.nbra) {
                const int64_t cnt = pd.qoff[b + 1] - pd.qoff[b];
                if (cnt == 0) { ++b; continue; }
                if (cnt * per_q > kChunkItems) {
                  // one long bra: several single-bra chunks
                  const int64_t step = std::max<int64_t>(1, kChunkItems / per_q);
                  for (int64_t k0 = 0; k0 < cnt; k0 += step) {
                    BraChunk c;
                    c.q0 = t.q0 + pd.qoff[b] + k0;
                    c.nq = (int32_t)std::min<int64_t>(step, cnt - k0);
                    c.S = t.S;
                    c.bra0 = pd.bra0 + b;
                    c.nbra = 1;
                    c.ket0 = pd.ket0 + (int32_t)k0;
                    c.nket = c.nq;
                    qc.chunks.push_back(c);
                  }
                  ++b;
                  continue;
                }
                BraChunk c;
                c.q0 = t.q0 + pd.qoff[b];
                c.S = t.S;
                c.bra0 = pd.bra0 + b;
                int64_t items = 0;
                int e = b;
                while (e < pd.nbra && e - b < qc.nbra_max) {
                  const int64_t ce = pd.qoff[e + 1] - pd.qoff[e];
                  if (items > 0 && (items + ce * per_q > kChunkItems || ce * per_q > kChunkItems)) break;
                  items += ce * per_q;
                  ++e;
                }
                c.nbra = e - b;
                c.nq = (int32_t)(pd.qoff[e] - pd.qoff[b]);
                c.ket0 = pd.ket0;
                c.nket = 0;
                for (int bb = b; bb < e; ++bb)
                  c.nket = std::max<int32_t>(c.nket, (int32_t)(pd.qoff[bb + 1] - pd.qoff[bb]));
                qc.chunks.push_back(c);
                b = e;
              }
            }
            // LDS is laid out for nbra_max slots in every workgroup: size it
            // for the widest chunk formed (the item limit often closes
            // chunks first), not for the cap
            if (qc.nbra_max >= 1 && !qc.chunks.empty()) {
              int widest = 1;
              for (const BraChunk& c : qc.chunks) widest = std::max(widest, (int)c.nbra);
              qc.nbra_max = std::min(qc.nbra_max, widest);
            }
            if (stage)
              for (const BraChunk& c : qc.chunks) {
                int cap = 0;
                for (int j = 0; j < c.nbra; ++j) cap += S.pairs[c.bra0 + j].np;
                qc.stage_cap = std::max(qc.stage_cap, cap);
              }
            // ket-side J in LDS for chunks whose ket window fits kKetLdsBytes per density
            qc.threads = tu.threads;
            if (tu.ket_lds) {
              const int cap = (int)(kKetLdsBytes / (sizeof(double) * wc * wd));
              for (BraChunk& c : qc.chunks) {
                if (c.nket > cap) c.nket = 0;
                qc.nket_max = std::max(qc.nket_max, c.nket);
              }
            } else {
              for (BraChunk& c : qc.chunks) c.nket = 0;
            }
            // partial slabs (segmented sets): each chunk's exchange rows
            // [nbra][rows][columns of the class's ket mask], then its ket
            // window [nket][wc*wd]; row-list keys are (global row, mask)
            const int pmode = P.options.partial_sums;
            if (pmode && !S.contracted && qc.nbra_max >= 1 && !qc.pb && !qc.pk) {
              qc.part = true;
              qc.part_beg = S.part_total;
              const int kmask = (1 << qc.lc) | (1 << qc.ld);
              for (BraChunk& c : qc.chunks) {
                const int m = P.seg[S.pairs[c.bra0].sa].mol;
                const int nm = S.h_cspace[m * 8 + kmask].y;
                c.poff = S.part_total;
                if (pmode & 1)
                  for (int sl = 0; sl < c.nbra; ++sl) {
                    const Pair& pr = S.pairs[c.bra0 + sl];
                    for (int r = 0; r < rows; ++r) {
                      const int ao = r < wa ? P.seg[pr.sa].ao_cart + r : P.seg[pr.sb].ao_cart + (r - wa);
                      kent.push_back({((int64_t)P.mols[m].aoc0 + ao) * 8 + kmask,
                                      S.part_total + ((int64_t)sl * rows + r) * nm});
                      kent_cls.push_back((int)S.classes.size());
                    }
                  }
                S.part_total += (int64_t)c.nbra * rows * nm;
                if (pmode & 2) {
                  for (int w = 0; w < c.nket; ++w)
                    jent.push_back({(int64_t)c.ket0 + w, S.part_total + (int64_t)w * wc * wd});
                  S.part_total += (int64_t)c.nket * wc * wd;
                }
              }
              qc.part_end = S.part_total;
            }
            if (qc.nbra_max >= 1) qc.d_chunks = dev_upload(qc.chunks, s);
          }
          HIP_OK(hipMalloc(&qc.d_quartets, sizeof(int2) * qc.nq));
          for (const Pending& pd : pend) {
            int64_t* d_qoff = dev_upload(pd.qoff, s);
            hipLaunchKernelGGL(expand_kernel, dim3(nblocks((size_t)pd.nq)), dim3(256), 0, s, d_qoff, pd.nbra,
                               pd.bra0, pd.ket0, pd.nq, pd.seg, d_run_order + pd.ket0, d_run_groups + run_g0[pd.rk],
                               run_g0[pd.rk + 1] - run_g0[pd.rk], qc.d_quartets + pd.q0);
            HIP_OK(hipGetLastError());
            HIP_OK(hipStreamSynchronize(s));
            HIP_OK(hipFree(d_qoff));
          }
          int64_t nq_chunk_max = 0;
          for (const BraChunk& c : qc.chunks) nq_chunk_max = std::max<int64_t>(nq_chunk_max, c.nq);
          // (tuned chunk sizes above kSortMax keep the expansion order)
          if (qc.d_chunks && !qc.chunks.empty() && nq_chunk_max <= kSortMax) {
            int2* tmp = nullptr;
            HIP_OK(hipMalloc(&tmp, sizeof(int2) * qc.nq));
            HIP_OK(hipMemcpyAsync(tmp, qc.d_quartets, sizeof(int2) * qc.nq, hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(sort_chunk_kernel, dim3((unsigned)qc.chunks.size()), dim3(256), 0, s, qc.d_chunks,
                               S.d_pairs, S.d_shells, tmp, qc.d_quartets);
            HIP_OK(hipGetLastError());
            HIP_OK(hipStreamSynchronize(s));
            HIP_OK(hipFree(tmp));
          }
          qc.d_tasks = dev_upload(qc.tasks, s);
          S.nquartets += qc.nq;
          S.classes.push_back(std::move(qc));
        }
  HIP_OK(hipFree(d_run_order));
  HIP_OK(hipFree(d_run_groups));
  if (S.part_total * (int64_t)sizeof(double) > kPartCapBytes) {
    S.part_total = 0;
    for (QuartetClass& qc : S.classes) qc.part = false;
  }
  if (S.part_total > 0) {
    // CSR lists: K by global Cartesian row, J by ket pair
    auto csr = [&](std::vector<std::pair<int64_t, int64_t>>& ent, int64_t nkeys, int64_t*& d_rp, int64_t*& d_ent) {
      std::vector<int64_t> rp(nkeys + 1, 0), off(std::max<size_t>(1, ent.size()));
      for (const auto& e : ent) ++rp[e.first + 1];
      for (int64_t k = 0; k < nkeys; ++k) rp[k + 1] += rp[k];
      std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
      for (const auto& e : ent) off[fill[e.first]++] = e.second;
      d_rp = dev_upload(rp, s);
      d_ent = dev_upload(off, s);
    };
    int64_t nrows = 0;
    for (int m = 0; m < P.nmol; ++m) nrows = std::max<int64_t>(nrows, (int64_t)P.mols[m].aoc0 + P.mols[m].nc);
    std::vector<int64_t> row_dst(nrows, 0);
    std::vector<int> row_n(nrows, 0), row_mol(nrows, 0);
    for (int m = 0; m < P.nmol; ++m)
      for (int a = 0; a < P.mols[m].nc; ++a) {
        row_dst[P.mols[m].aoc0 + a] = P.mols[m].cbase + (int64_t)a * P.mols[m].nc;
        row_n[P.mols[m].aoc0 + a] = P.mols[m].nc;
        row_mol[P.mols[m].aoc0 + a] = m;
        S.nc_max = std::max(S.nc_max, P.mols[m].nc);
      }
    S.nrows = nrows;
    S.d_row_dst = dev_upload(row_dst, s);
    S.d_row_n = dev_upload(row_n, s);
    S.d_row_mol = dev_upload(row_mol, s);
    csr(kent, nrows * 8, S.d_krp, S.d_kent);
    csr(jent, (int64_t)S.pairs.size(), S.d_jrp, S.d_jent);
    S.h_kent = std::move(kent);
    S.h_kent_cls = std::move(kent_cls);
  }
  if (!colmap.empty()) {
    S.d_colmap = dev_upload(colmap, s);
    S.d_cinv = dev_upload(cinv, s);
    S.d_shell_cofs = dev_upload(shell_cofs, s);
  }
  S.d_cspace = dev_upload(S.h_cspace, s);
  S.d_cinv_off = dev_upload(cinv_off, s);
  S.built = true;
}

static void free_pair_set(PairSet& S) {
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  if (S.contracted) {
    f(S.d_shells);
    f(S.d_prims);
  }
  f(S.d_coef); f(S.d_blk); f(S.d_pairs); f(S.d_pair_q);
  f(S.d_krp); f(S.d_kent); f(S.d_jrp); f(S.d_jent); f(S.d_row_dst); f(S.d_row_n); f(S.d_row_mol);
  f(S.d_colmap); f(S.d_cspace); f(S.d_cinv_off); f(S.d_cinv); f(S.d_shell_cofs);
  for (auto& c : S.classes) {
    f(c.d_quartets);
    f(c.d_tasks);
    f(c.d_chunks);
  }
  S = PairSet();
}

// Segmented pair set over every intra-molecule segment pair; its density
// blocks are the segments.
static void build_segmented_set(PlanData& P, PairSet& S, bool with_chunks, hipStream_t s) {
  S.contracted = false;
  S.d_shells = P.d_shells;
  S.d_prims = P.d_prims;
  std::vector<PairCand> cand;
  cand.reserve(P.all_pairs.size());
  for (size_t i = 0; i < P.all_pairs.size(); ++i)
    cand.push_back({P.all_pairs[i], P.all_class[i], 0, P.seg[P.all_pairs[i].sa].mol, P.seg_pair_q[i]});
  if (&S == &P.jk) {
    std::vector<int> blk;
    for (const Seg& g : P.seg) {
      blk.push_back(g.ao_cart);
      blk.push_back(ncart_h(g.l));
    }
    S.d_blk = dev_upload(blk, s);
  }
  build_pair_set(P, S, cand, with_chunks, s);
}

// The derivative kernels run over segmented pairs (no chunks needed).
static PairSet& ensure_seg_set(PlanData& P, hipStream_t s) {
  PairSet& S = P.segq();
  if (!S.built) {
    build_segmented_set(P, S, false, s);
    HIP_OK(hipStreamSynchronize(s));
  }
  return S;
}

// General contraction is evaluated natively for s shells with two columns
// (cc-pVXZ-type 1s/2s); other generally contracted shells use segments.
static bool contracted_eligible(const PlanData& P) {
  bool any = false;
  for (int sh = 0; sh < P.nbas; ++sh) {
    const int l = P.bas[sh * BAS_SLOTS + ANG_OF], nc = P.bas[sh * BAS_SLOTS + NCTR_OF];
    if (nc > 1) {
      if (nc > 2 || l != 0) return false;
      any = true;
    }
  }
  return any;
}

// Contracted J/K pair set over parent shells.  Returns false (and leaves the
// set empty) when a general-pattern class cannot hold one bra's exchange rows
// for one density in LDS; the caller then uses segmented pairs.
static bool build_contracted_set(PlanData& P, hipStream_t s) {
  PairSet& S = P.jk;
  S.contracted = true;
  const double pi = 3.14159265358979323846;
  const double s2p = std::sqrt(2.0 * std::pow(pi, 2.5));
  std::vector<ShellXYZ> sh(P.nbas);
  std::vector<int> blk;
  std::vector<MolAO> mols = P.mols;
  int64_t db = 0;
  for (int m = 0; m < P.nmol; ++m) {
    mols[m].seg0 = P.mol_bas[m];
    mols[m].nseg = P.mol_bas[m + 1] - P.mol_bas[m];
    mols[m].dbase = db;
    db += (int64_t)mols[m].nseg * mols[m].nseg;
    for (int b = P.mol_bas[m]; b < P.mol_bas[m + 1]; ++b) {
      const int* row = P.bas.data() + b * BAS_SLOTS;
      const double* r = P.env.data() + P.atm[row[ATOM_OF] * ATM_SLOTS + PTR_COORD];
      ShellXYZ& x = sh[b];
      x.x = r[0]; x.y = r[1]; x.z = r[2];
      x.ao = P.ao_cart[b] - P.mols[m].aoc0;
      x.l = row[ANG_OF];
      x.mol = m;
      x.lseg = b - P.mol_bas[m];
      blk.push_back(x.ao);
      blk.push_back(row[NCTR_OF] * ncart_h(x.l));
    }
  }
  mols[P.nmol].seg0 = P.nbas;
  mols[P.nmol].dbase = db;
  // parent pairs with their primitive pairs and column coefficients
  std::vector<PrimPair> prims;
  std::vector<PrimCoef> coef;
  std::vector<PairCand> cand;
  auto fsp = [&](int l) { return l == 0 ? 0.5 / std::sqrt(pi) : (l == 1 ? std::sqrt(3.0 / (4.0 * pi)) : 1.0); };
  for (int m = 0; m < P.nmol; ++m) {
    const int nb = P.mol_bas[m + 1] - P.mol_bas[m];
    for (int Sb = P.mol_bas[m]; Sb < P.mol_bas[m + 1]; ++Sb)
      for (int Tb = P.mol_bas[m]; Tb <= Sb; ++Tb) {
        int a = Sb, b = Tb;
        const int* ra = P.bas.data() + a * BAS_SLOTS;
        const int* rb = P.bas.data() + b * BAS_SLOTS;
        // l(A) >= l(B); for equal l the generally contracted shell first
        if (rb[ANG_OF] > ra[ANG_OF] || (rb[ANG_OF] == ra[ANG_OF] && rb[NCTR_OF] > ra[NCTR_OF])) {
          std::swap(a, b);
          std::swap(ra, rb);
        }
        const int la = ra[ANG_OF], lb = rb[ANG_OF], nca = ra[NCTR_OF], ncb = rb[NCTR_OF];
        const int pat = (nca - 1) + 2 * (ncb - 1);
        if (!pattern_valid(la, lb, pat)) throw Internal("unsupported general-contraction pattern");
        const double* A = P.env.data() + P.atm[ra[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        const double* B = P.env.data() + P.atm[rb[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        double ab2 = 0;
        for (int x = 0; x < 3; ++x) ab2 += (A[x] - B[x]) * (A[x] - B[x]);
        Pair pr;
        pr.sa = a;
        pr.sb = b;
        pr.poff = (int)prims.size();
        pr.np = 0;
        const int npa = ra[NPRIM_OF], npb = rb[NPRIM_OF];
        const double fa = fsp(la), fb = fsp(lb);
        for (int i = 0; i < npa; ++i)
          for (int j = 0; j < npb; ++j) {
            const double ai = P.env[ra[PTR_EXP] + i], aj = P.env[rb[PTR_EXP] + j], p = ai + aj;
            const double mu = ai * aj / p * ab2;
            if (mu > 60.0) continue;
            PrimCoef c;
            for (int k = 0; k < 2; ++k) {
              c.ca[k] = k < nca ? P.env[ra[PTR_COEFF] + k * npa + i] * fa : 0.0;
              c.cb[k] = k < ncb ? P.env[rb[PTR_COEFF] + k * npb + j] * fb : 0.0;
            }
            PrimPair pp;
            pp.p = p;
            pp.px = (ai * A[0] + aj * B[0]) / p;
            pp.py = (ai * A[1] + aj * B[1]) / p;
            pp.pz = (ai * A[2] + aj * B[2]) / p;
            const double ek = std::exp(-mu) * s2p / p;
            pp.e = pat == 0 ? c.ca[0] * c.cb[0] * ek : ek;
            pp.aa = ai;
            pp.ab = aj;
            pp.ip = 1.0 / p;
            prims.push_back(pp);
            coef.push_back(c);
            ++pr.np;
          }
        const int ps = Sb - P.mol_bas[m], pt = Tb - P.mol_bas[m];
        const double qv = P.parent_q.empty() ? 1.0
                          : P.parent_q[(size_t)P.qbase[m] + (size_t)std::max(ps, pt) * nb + std::min(ps, pt)];
        cand.push_back({pr, pair_class(la, lb), pat, m, qv});
      }
  }
  if (prims.empty()) {
    prims.push_back(PrimPair{});
    coef.push_back(PrimCoef{});
  }
  S.d_shells = dev_upload(sh, s);
  S.d_prims = dev_upload(prims, s);
  S.d_coef = dev_upload(coef, s);
  S.nprims = (int64_t)prims.size();
  S.d_blk = dev_upload(blk, s);
  build_pair_set(P, S, cand, true, s);
  for (const QuartetClass& qc : S.classes)
    if ((qc.pb || qc.pk) && qc.nbra_max < 1) {
      free_pair_set(S);
      return false;
    }
  P.mols = mols;
  P.nd_tot = db;
  return true;
}

// Spread the class launches of a build over the plan's streams: longest
// class first onto the least-loaded stream (stream 0 is the caller's), by
// qc.cost (the plan-time work model, or measured durations after
// balance_lanes), and build each stream's exchange-row reduction lists.
static void assign_lanes(PlanData& P) {
  const int ns = (int)P.side.size() + 1;
  for (int64_t* p : P.d_lane_krp) HIP_OK(hipFree(p));
  for (int64_t* p : P.d_lane_kent) HIP_OK(hipFree(p));
  P.d_lane_krp.clear();
  P.d_lane_kent.clear();
  P.lane_classes.assign(ns, {});
  std::vector<int> order(P.jk.classes.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return P.jk.classes[x].cost > P.jk.classes[y].cost; });
  std::vector<double> load(ns, 0.0);
  for (int ci : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[l] += P.jk.classes[ci].cost;
    P.lane_classes[l].push_back(ci);
  }
  // per-stream row lists of the exchange-row slabs
  const PairSet& S = P.jk;
  if (ns > 1 && S.part_total > 0 && !S.h_kent.empty()) {
    std::vector<int> lane_of(S.classes.size(), 0);
    for (int l = 0; l < ns; ++l)
      for (int ci : P.lane_classes[l]) lane_of[ci] = l;
    for (int l = 0; l < ns; ++l) {
      const int64_t nkeys = S.nrows * 8;  // (row, ket mask)
      std::vector<int64_t> rp(nkeys + 1, 0), off;
      for (size_t i = 0; i < S.h_kent.size(); ++i)
        if (lane_of[S.h_kent_cls[i]] == l) ++rp[S.h_kent[i].first + 1];
      for (int64_t k = 0; k < nkeys; ++k) rp[k + 1] += rp[k];
      off.assign(std::max<int64_t>(1, rp[nkeys]), 0);
      std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
      for (size_t i = 0; i < S.h_kent.size(); ++i)
        if (lane_of[S.h_kent_cls[i]] == l) off[fill[S.h_kent[i].first]++] = S.h_kent[i].second;
      P.d_lane_krp.push_back(dev_upload(rp, nullptr));
      P.d_lane_kent.push_back(dev_upload(off, nullptr));
    }
    // the lists are read from the plan's non-blocking streams
    HIP_OK(hipDeviceSynchronize());
  }
}

// Side streams and fork/join events of the class launches, then the lanes.
static void assign_streams(PlanData& P) {
  const int ns = std::max(1, std::min<int>(kMaxStreams, P.options.streams));
  for (int l = 1; l < ns; ++l) {
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    P.side.push_back(st);
    hipEvent_t ev;
    HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    P.join_ev.push_back(ev);
  }
  if (ns > 1) HIP_OK(hipEventCreateWithFlags(&P.fork_ev, hipEventDisableTiming));
  assign_lanes(P);
}

// Schwarz tables of the deterministic scale (det_scale_kernel): Cartesian
// bounds Q(S,T) = sqrt(max |(st|st)|) over the J/K shell set's blocks (a
// contracted set's parent shell takes the maximum over its segments).
static void build_det_tables(PlanData& P, hipStream_t s) {
  std::vector<double> seg_q, cart_q;
  compute_schwarz(P, seg_q, s, &cart_q);
  const bool con = P.jk.contracted;
  std::vector<double> qblk((size_t)std::max<int64_t>(P.nd_tot, 1), 0.0), mol(2 * (size_t)P.nmol, 0.0);
  std::vector<std::vector<double>> q(P.nmol);
  for (int m = 0; m < P.nmol; ++m) q[m].assign((size_t)P.mols[m].nseg * P.mols[m].nseg, 0.0);
  for (size_t i = 0; i < P.all_pairs.size(); ++i) {
    const Pair& pr = P.all_pairs[i];
    const int m = P.seg[pr.sa].mol;
    const MolAO& mo = P.mols[m];
    const int a = (con ? P.seg[pr.sa].parent : pr.sa) - mo.seg0;
    const int b = (con ? P.seg[pr.sb].parent : pr.sb) - mo.seg0;
    const double v = std::sqrt(cart_q[i]);
    double& x = q[m][(size_t)a * mo.nseg + b];
    double& y = q[m][(size_t)b * mo.nseg + a];
    x = std::max(x, v);
    y = std::max(y, v);
  }
  for (int m = 0; m < P.nmol; ++m) {
    const MolAO& mo = P.mols[m];
    std::vector<double> w(mo.nseg);
    for (int k = 0; k < mo.nseg; ++k) {
      if (con) {
        const int* row = &P.bas[(size_t)(mo.seg0 + k) * BAS_SLOTS];
        w[k] = (double)ncart_h(row[ANG_OF]) * row[NCTR_OF];
      } else {
        w[k] = (double)ncart_h(P.seg[mo.seg0 + k].l);
      }
    }
    double qmax = 0.0, rmax = 0.0;
    for (int a = 0; a < mo.nseg; ++a) {
      double row = 0.0;
      for (int b = 0; b < mo.nseg; ++b) {
        const double v = q[m][(size_t)a * mo.nseg + b];
        qblk[(size_t)mo.dbase + (size_t)a * mo.nseg + b] = v * w[a] * w[b];
        qmax = std::max(qmax, v);
        row += v * w[b];
      }
      rmax = std::max(rmax, row);
    }
    mol[2 * m] = qmax;
    mol[2 * m + 1] = rmax;
  }
  P.d_det_qblk = dev_upload(qblk, s);
  P.d_det_mol = dev_upload(mol, s);
  HIP_OK(hipMalloc(&P.d_det_scale, sizeof(double) * 2 * P.nmol));
}

static void build_plan(PlanData& P, double cutoff, hipStream_t s) {
  if (!std::isfinite(cutoff) || cutoff < 0) throw InvalidArgument("Invalid Schwarz cutoff");
  P.cutoff = cutoff;
  {
    // LDS a one-density launch may take: the device's limit per workgroup
    // (at most kLdsDeviceMax), lowered by the lds_limit option
    int dev = 0, lmax = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    P.lds_max = std::min<size_t>(kLdsDeviceMax, (size_t)std::max(lmax, 0));
    if (P.options.lds_limit > 0) P.lds_max = std::min<size_t>(P.lds_max, (size_t)P.options.lds_limit);
    P.lds_budget = std::min(kLdsBudget, P.lds_max);
  }
  build_segments(P);
  const int nseg = (int)P.seg.size();
  // shells on device
  std::vector<ShellXYZ> sh(nseg);
  for (int i = 0; i < nseg; ++i) {
    sh[i].x = P.seg[i].r[0];
    sh[i].y = P.seg[i].r[1];
    sh[i].z = P.seg[i].r[2];
    sh[i].ao = P.seg[i].ao_cart;
    sh[i].l = P.seg[i].l;
    sh[i].mol = P.seg[i].mol;
    sh[i].lseg = P.seg[i].lseg;
  }
  P.d_shells = dev_upload(sh, s);
  // all intra-molecule segmented pairs S >= T, molecule by molecule
  P.mol_pair0.assign(P.nmol + 1, 0);
  for (int m = 0; m < P.nmol; ++m) {
    P.mol_pair0[m] = (int64_t)P.all_pairs.size();
    for (int S = P.mol_seg[m]; S < P.mol_seg[m + 1]; ++S)
      for (int T = P.mol_seg[m]; T <= S; ++T) {
        int a = S, b = T;
        if (P.seg[b].l > P.seg[a].l) std::swap(a, b);
        Pair pr;
        make_pair(P, a, b, pr);
        P.all_pairs.push_back(pr);
        P.all_class.push_back(pair_class(P.seg[a].l, P.seg[b].l));
      }
  }
  P.mol_pair0[P.nmol] = (int64_t)P.all_pairs.size();
  if (P.prims.empty()) {
    P.prims.push_back(PrimPair{});
    P.aux.push_back(PrimAux{});
  }
  P.d_prims = dev_upload(P.prims, s);
  P.d_aux = dev_upload(P.aux, s);
  P.d_all_pairs = dev_upload(P.all_pairs, s);
  {
    // primitive slots of every segmented shell and where they live in env
    const double pi = 3.14159265358979323846;
    P.seg_prim0.assign(nseg + 1, 0);
    std::vector<int2> slots;
    std::vector<double> fsp;
    for (int i = 0; i < nseg; ++i) {
      const Seg& g = P.seg[i];
      P.seg_prim0[i + 1] = P.seg_prim0[i] + g.np;
      const int* row = P.bas.data() + g.parent * BAS_SLOTS;
      const double f = g.l == 0 ? 0.5 / std::sqrt(pi) : (g.l == 1 ? std::sqrt(3.0 / (4.0 * pi)) : 1.0);
      for (int p = 0; p < g.np; ++p) {
        slots.push_back(make_int2(row[PTR_EXP] + p, row[PTR_COEFF] + g.ctr * g.np + p));
        fsp.push_back(f);
      }
    }
    P.d_seg_prim0 = dev_upload(P.seg_prim0, s);
    P.d_prim_slots = dev_upload(slots, s);
    P.d_prim_fsp = dev_upload(fsp, s);
  }
  // Schwarz bounds (parent-shell maxima, as the reference defines them per shell pair)
  const int npair = (int)P.all_pairs.size();
  std::vector<double>& q = P.seg_pair_q;
  q.assign(npair, 1.0);
  P.qmax_mol.assign(P.nmol, 0.0);
  P.qbase.assign(P.nmol + 1, 0);
  for (int m = 0; m < P.nmol; ++m) {
    const int64_t nb = P.mol_bas[m + 1] - P.mol_bas[m];
    P.qbase[m + 1] = P.qbase[m] + nb * nb;
  }
  if (cutoff > 0) {
    std::vector<double> seg_q;
    compute_schwarz(P, seg_q, s);
    // parent_q[qbase[m] + i*nb + j], i >= j local parent shells: max over segment pairs
    P.parent_q.assign((size_t)P.qbase[P.nmol], 0.0);
    auto pslot = [&](int S, int T) -> double& {
      const int m = P.seg[S].mol;
      const int nb = P.mol_bas[m + 1] - P.mol_bas[m];
      const int ps = P.seg[S].parent - P.mol_bas[m], pt = P.seg[T].parent - P.mol_bas[m];
      return P.parent_q[(size_t)P.qbase[m] + (size_t)std::max(ps, pt) * nb + std::min(ps, pt)];
    };
    for (int m = 0; m < P.nmol; ++m)
      for (int S = P.mol_seg[m]; S < P.mol_seg[m + 1]; ++S)
        for (int T = P.mol_seg[m]; T <= S; ++T) {
          double& v = pslot(S, T);
          v = std::max(v, seg_q[pair_idx(P, S, T)]);
        }
    for (double& v : P.parent_q) v = std::sqrt(v);
    for (int m = 0; m < P.nmol; ++m)
      for (int S = P.mol_seg[m]; S < P.mol_seg[m + 1]; ++S)
        for (int T = P.mol_seg[m]; T <= S; ++T) q[pair_idx(P, S, T)] = pslot(S, T);
  }
  for (int i = 0; i < npair; ++i) {
    const int m = P.seg[P.all_pairs[i].sa].mol;
    P.qmax_mol[m] = std::max(P.qmax_mol[m], q[i]);
  }
  {
    std::vector<int> at;
    for (const Seg& g : P.seg) at.push_back(P.bas[g.parent * BAS_SLOTS + ATOM_OF]);
    P.d_seg_atom = dev_upload(at, s);
  }
  // J/K pair set: parent shells with their general-contraction columns where
  // the basis has them (and the exchange slabs fit), segmented shells otherwise
  bool done = false;
  if (P.options.general_contraction && contracted_eligible(P)) done = build_contracted_set(P, s);
  if (!done) build_segmented_set(P, P.jk, true, s);
  assign_streams(P);
  // molecule tables (the density-maximum blocks follow the J/K shell set)
  {
    std::vector<MolInfo> mi(P.nmol);
    for (int m = 0; m < P.nmol; ++m) {
      mi[m].base = P.mols[m].cbase;
      mi[m].dbase = P.mols[m].dbase;
      mi[m].n = P.mols[m].nc;
      mi[m].nseg = P.mols[m].nseg;
    }
    P.d_molinfo = dev_upload(mi, s);
    P.d_molao = dev_upload(P.mols, s);
    HIP_OK(hipMalloc(&P.d_dmax, sizeof(double) * std::max<int64_t>(P.nd_tot, 1)));
  }
  if (P.options.deterministic) build_det_tables(P, s);
  build_c2s_maps(P, s);
  {
    std::vector<int> order(npair);
    for (int i = 0; i < npair; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return P.all_class[x] < P.all_class[y]; });
    std::vector<Pair> p1e;
    int k = 0;
    for (int c = 0; c < kNumPairClasses; ++c) {
      P.p1e_begin[c] = k;
      while (k < npair && P.all_class[order[k]] == c) p1e.push_back(P.all_pairs[order[k++]]);
    }
    P.p1e_begin[kNumPairClasses] = k;
    P.d_pairs1e = dev_upload(p1e, s);
    std::vector<double> nuc;
    for (int m = 0; m < P.nmol; ++m)
      for (int a = P.mol_atm[m]; a < P.mol_atm[m + 1]; ++a) {
        const int z = P.atm[a * ATM_SLOTS + CHARGE_OF];
        if (z == 0) continue;
        const double* r = P.env.data() + P.atm[a * ATM_SLOTS + PTR_COORD];
        nuc.insert(nuc.end(), {r[0], r[1], r[2], (double)z, (double)m, (double)a});
      }
    P.nnuc = (int)(nuc.size() / 6);
    if (nuc.empty()) nuc.assign(6, -1.0);
    P.d_nuclei = dev_upload(nuc, s);
  }
  HIP_OK(hipMalloc(&P.d_flag, sizeof(int)));
  HIP_OK(hipHostMalloc(&P.h_flag, sizeof(int)));
  HIP_OK(hipStreamSynchronize(s));
}

// Device bytes a plan holds, with the J/K workspace for nbd densities and
// quartet-list / chunk sizes as given (exact after build_plan).
static int64_t plan_bytes(const PlanData& P, int64_t npairs, int64_t nq, int64_t ntasks, int64_t nchunks,
                          int64_t nbd, int64_t extra, int64_t part_doubles) {
  int64_t b = 0;
  // partial-sum slabs per density, their row lists (one entry per slab row
  // and window entry at most) and the per-row tables
  b += part_doubles * (int64_t)sizeof(double) * nbd;
  b += std::min<int64_t>(part_doubles, nq * 48) * (int64_t)sizeof(int64_t) +
       (int64_t)(P.n_cart + npairs + 2) * (int64_t)(2 * sizeof(int64_t) + sizeof(int));
  b += (int64_t)P.seg.size() * (sizeof(ShellXYZ) + 3 * sizeof(int));         // shells, seg_ao, seg_atom
  b += (int64_t)std::max<size_t>(P.prims.size(), 1) * (sizeof(PrimPair) + sizeof(PrimAux));
  b += (int64_t)(P.seg.size() + 1) * sizeof(int);
  for (const Seg& g : P.seg) b += (int64_t)g.np * (sizeof(int2) + 3 * sizeof(double));  // slots, fsp, sums
  b += (int64_t)P.all_pairs.size() * 2 * sizeof(Pair);                        // all pairs, 1e pairs
  b += npairs * (sizeof(Pair) + sizeof(double));                              // screened pairs, pair_q
  b += extra;                                                                 // contracted J/K shells and primitives
  b += (int64_t)P.n_cart * 2 * (sizeof(int) + sizeof(double)) + (int64_t)P.n * 3 * (sizeof(int) + sizeof(double));
  b += (int64_t)P.nmol * sizeof(MolInfo) + (int64_t)(P.nmol + 1) * sizeof(MolAO);
  b += std::max<int64_t>(P.nd_tot, 1) * sizeof(double);                       // shell-block density maxima
  b += (int64_t)5 * sizeof(double) * std::max(P.natm, 1);                     // nuclei
  b += nq * sizeof(int2) + ntasks * sizeof(SubTask) + nchunks * sizeof(BraChunk);
  b += 3 * P.ncc_tot * (int64_t)sizeof(double) * nbd;                         // Cartesian D, Jp, Kp
  b += 8 * P.ncc_tot * (int64_t)sizeof(double) * nbd;                         // ket-side J, one line per element
  return b;
}

static int64_t plan_bytes(const PlanData& P, int64_t nbd) {
  int64_t np = 0, nq = 0, nt = 0, nc = 0, extra = 0;
  for (const PairSet* S : {&P.jk, &P.seg_set}) {
    np += (int64_t)S->pairs.size();
    for (const QuartetClass& c : S->classes) {
      nq += c.nq;
      nt += (int64_t)c.tasks.size();
      nc += (int64_t)c.chunks.size();
    }
    if (S->contracted)
      extra += (int64_t)P.nbas * (sizeof(ShellXYZ) + 2 * sizeof(int)) +
               S->nprims * (int64_t)(sizeof(PrimPair) + sizeof(PrimCoef));
  }
  return plan_bytes(P, np, nq, nt, nc, nbd, extra, P.options.partial_sums ? P.jk.part_total : 0);
}

// Host-only upper bound of the device bytes of a plan plus one build of nb
// densities (nonsymmetric K doubles the workspace): the exact size of every
// table, with every intra-molecule canonical pair quartet retained (no
// Schwarz screening is evaluated), at most one chunk per quartet and one
// sub-task per pair of (class, bucket) runs.  Transient buffers (Schwarz blocks, one-electron staging) are freed
// before a build and do not add to it.
static int64_t forecast_bytes(PlanData& P, int nb) {
  build_segments(P);
  P.mol_pair0.assign(P.nmol + 1, 0);
  int64_t nq = 0, npairs = 0, part = 0;
  for (int m = 0; m < P.nmol; ++m) {
    int64_t live = 0;
    for (int S = P.mol_seg[m]; S < P.mol_seg[m + 1]; ++S)
      for (int T = P.mol_seg[m]; T <= S; ++T) {
        int a = S, c = T;
        if (P.seg[c].l > P.seg[a].l) std::swap(a, c);
        Pair pr;
        make_pair(P, a, c, pr);
        P.all_pairs.push_back(pr);
        if (pr.np > 0) ++live;
      }
    nq += live * (live + 1) / 2;
    npairs += live;
    // slab rows: at most one bra slot (<= 12 rows of n_m) per quartet
    part += live * (live + 1) / 2 * 12 * P.mols[m].nc;
  }
  part = P.options.partial_sums ? std::min<int64_t>(part, kPartCapBytes / (int64_t)sizeof(double)) : 0;
  // sub-tasks pair a bra run with a ket run of one molecule: at most
  // (pair classes x column patterns x primitive-count buckets)^2 per molecule;
  // a contracted J/K set adds parent shells and at most as many primitive
  // pairs (with column coefficients) as the segmented set holds
  const int64_t runs = (int64_t)kNumPairClasses * 2 * 8;
  const int64_t extra = (int64_t)P.nbas * (sizeof(ShellXYZ) + 2 * sizeof(int)) +
                        (int64_t)P.prims.size() * (sizeof(PrimPair) + sizeof(PrimCoef));
  return plan_bytes(P, npairs, nq, std::min<int64_t>(nq, runs * runs * P.nmol), nq, 2 * (int64_t)nb, extra, part);
}

static void free_plan(PlanData& P) {
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  free_pair_set(P.jk);
  free_pair_set(P.seg_set);
  f(P.d_shells); f(P.d_prims); f(P.d_all_pairs);
  f(P.d_c2s_idx); f(P.d_c2s_coef); f(P.d_s2c_idx); f(P.d_s2c_coef);
  f(P.d_dc); f(P.d_jp); f(P.d_jpad); f(P.d_kp); f(P.d_flag); f(P.d_incore); f(P.d_part);
  f(P.d_pairs1e); f(P.d_nuclei);
  f(P.d_dmax); f(P.d_seg_atom);
  f(P.d_det_qblk); f(P.d_det_mol); f(P.d_det_scale);
  f(P.d_molinfo); f(P.d_molao);
  f(P.d_aux); f(P.d_seg_prim0); f(P.d_prim_slots); f(P.d_prim_fsp); f(P.d_gparam);
  for (double* p : P.retired) f(p);
  for (hipEvent_t e : P.ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : P.join_ev) (void)hipEventDestroy(e);
  if (P.fork_ev) (void)hipEventDestroy(P.fork_ev);
  for (hipStream_t st : P.side) (void)hipStreamDestroy(st);
  for (int64_t* p : P.d_lane_krp) f(p);
  for (int64_t* p : P.d_lane_kent) f(p);
  if (P.h_flag) (void)hipHostFree(P.h_flag);
}

// Workspace for nbd densities.  A grown workspace does not free the previous
// buffers: a HIP graph captured from an earlier build still addresses them, so
// they are retired and released with the plan.
static void ensure_workspace(PlanData& P, size_t nbd) {
  if (P.ws_nbd >= nbd) return;
  for (double** p : {&P.d_dc, &P.d_jp, &P.d_jpad, &P.d_kp}) {
    if (*p) P.retired.push_back(*p);
    *p = nullptr;
  }
  const size_t nn = (size_t)P.ncc_tot;
  HIP_OK(hipMalloc(&P.d_dc, sizeof(double) * nn * nbd));
  HIP_OK(hipMalloc(&P.d_jp, sizeof(double) * nn * nbd));
  HIP_OK(hipMalloc(&P.d_jpad, sizeof(double) * 8 * nn * nbd));
  HIP_OK(hipMalloc(&P.d_kp, sizeof(double) * nn * nbd));
  P.ws_nbd = nbd;
}

// Partial slabs for nbd densities; a grown buffer is retired like the workspace.
static void ensure_part(PlanData& P, size_t nbd) {
  if (P.part_nbd >= nbd || P.jk.part_total == 0 || P.options.partial_sums == 0) return;
  if (P.d_part) P.retired.push_back(P.d_part);
  P.d_part = nullptr;
  HIP_OK(hipMalloc(&P.d_part, sizeof(double) * (size_t)P.jk.part_total * nbd));
  P.part_nbd = nbd;
}

// Device buffers of one build: Cartesian densities, partial J and K, the
// padded ket-side J and the partial slabs (NULL: flushes by atomics).
struct Workspace {
  double* dc;
  double* jp;
  double* jpad;
  double* kp;
  double* part;
};

// Workspace doubles of a build of nb densities without partial slabs, laid out
// [dc | jp | kp] over nbd densities and jpad over nb (8 slots per element).
static size_t workspace_doubles(const PlanData& P, int nb, int nbd) {
  return (size_t)P.ncc_tot * (3 * (size_t)nbd + 8 * (size_t)nb);
}

// fixed: buffers supplied by the caller (no allocation in the build); NULL
// grows the plan's own workspace and slabs as needed.
static void jk_build(PlanData& P, const double* density, int nb, double* out, int with_j, int with_k,
                     uint32_t flags, double mixed_threshold, hipStream_t s, const Workspace* fixed = nullptr) {
  const bool mixed = (flags & GSCF_FLAG_MIXED_PRECISION) != 0;
  if (mixed && !(std::isfinite(mixed_threshold) && mixed_threshold >= 0))
    throw InvalidArgument("mixed-precision threshold must be finite and non-negative");
  if ((with_j != 0 && with_j != 1) || (with_k != 0 && with_k != 1) || !(with_j || with_k))
    throw InvalidArgument("Direct J/K requires at least one selected operator");
  if (nb < 1) throw InvalidArgument("Direct J/K shape mismatch");
  if (!density || !out) throw InvalidArgument("Direct J/K requires device density and output buffers");
  // deterministic digestion (engine option chosen at plan creation): integer
  // accumulation in FP64 only, every flush by atomics, no partial slabs
  const bool det = P.options.deterministic != 0;
  if (det && mixed) throw InvalidArgument("mixed precision is not available in deterministic mode");
  const size_t nn = (size_t)P.nn_tot, ncc = (size_t)P.ncc_tot;
  if (!(flags & GSCF_FLAG_NO_VALIDATE)) {
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, P.d_flag, 1, 0);
    hipLaunchKernelGGL(finite_kernel, dim3(nblocks(nn * nb)), dim3(256), 0, s, density, nn * nb, P.d_flag);
    HIP_OK(hipMemcpyAsync(P.h_flag, P.d_flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (*P.h_flag) throw InvalidArgument("Nonfinite density");
  }
  const bool sym = (flags & GSCF_FLAG_SYMMETRIC_DENSITY) != 0;
  const int nbd = (with_k && !sym) ? 2 * nb : nb;
  Workspace W;
  if (fixed) {
    W = *fixed;
  } else {
    ensure_workspace(P, nbd);
    ensure_part(P, nbd);
    W = {P.d_dc, P.d_jp, P.d_jpad, P.d_kp, P.d_part};
  }
  hipLaunchKernelGGL(dens_to_cart_kernel, dim3(nblocks(ncc * nbd)), dim3(256), 0, s, density, P.nn_tot, nb, W.dc,
                     P.ncc_tot, P.d_molao, P.nmol, P.d_c2s_idx, P.d_c2s_coef, nbd);
  if (mixed || det)
    hipLaunchKernelGGL(dmax_kernel, dim3(nblocks((size_t)P.nd_tot)), dim3(256), 0, s, W.dc, P.ncc_tot, nbd,
                       P.d_molao, P.nmol, P.nd_tot, P.jk.d_blk, P.d_dmax);
  if (det)
    hipLaunchKernelGGL(det_scale_kernel, dim3((unsigned)P.nmol), dim3(256), 0, s, P.d_dmax, P.d_det_qblk,
                       P.d_molao, P.d_det_mol, P.d_det_scale);
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks(ncc * nbd)), dim3(256), 0, s, W.jp, ncc * nbd, 0.0);
  if (with_j)
    hipLaunchKernelGGL(fill_kernel, dim3(nblocks(8 * ncc * nb)), dim3(256), 0, s, W.jpad, 8 * ncc * nb, 0.0);
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks(ncc * nbd)), dim3(256), 0, s, W.kp, ncc * nbd, 0.0);
  const PairSet& S = P.jk;
  if (P.prof) ++P.prof_builds;
  // Class launches: in list order on the caller's stream while profiling (so
  // that the per-class events time one kernel each), otherwise spread over
  // the plan's streams, forked from and joined back into the caller's stream.
  const bool fork = !P.prof && !P.side.empty();
  // with several streams each folds the exchange-row slabs of its own
  // classes right after them (device-scope adds into K), overlapping the
  // other streams' class kernels; otherwise one reduction after the join
  const int psums = (det || !W.part) ? 0 : P.options.partial_sums;
  const bool lane_reduce = fork && with_k && (psums & 1) && S.part_total > 0 &&
                           P.d_lane_krp.size() == P.lane_classes.size();
  if (fork) {
    HIP_OK(hipEventRecord(P.fork_ev, s));
    for (hipStream_t st : P.side) HIP_OK(hipStreamWaitEvent(st, P.fork_ev, 0));
  }
  std::vector<std::pair<int, hipStream_t>> launches;
  if (fork) {
    for (size_t l = 0; l < P.lane_classes.size(); ++l)
      for (int ci : P.lane_classes[l]) launches.push_back({ci, l == 0 ? s : P.side[l - 1]});
  } else {
    for (size_t ci = 0; ci < S.classes.size(); ++ci) launches.push_back({(int)ci, s});
  }
  for (const auto& [ci, ls] : launches) {
    const QuartetClass& qc = S.classes[ci];
    hipStream_t s = ls;  // this class's launch stream
    if (P.prof) {
      if (P.ev_used + 2 > P.ev_pool.size()) {
        for (int e = 0; e < 64; ++e) {
          hipEvent_t ev;
          HIP_OK(hipEventCreate(&ev));
          P.ev_pool.push_back(ev);
        }
      }
      P.prof_recs.push_back({qc.tuning_index, P.ev_used, P.prof_builds, ci});
      HIP_OK(hipEventRecord(P.ev_pool[P.ev_used], s));
    }
    // widths count every contraction column of a shell
    const int wa = ncart_h(qc.la) * (1 + (qc.pb & 1)), wb = ncart_h(qc.lb) * (1 + (qc.pb >> 1));
    const int wc = ncart_h(qc.lc) * (1 + (qc.pk & 1)), wd = ncart_h(qc.ld) * (1 + (qc.pk >> 1));
    BraArgs b;
    std::memset(&b, 0, sizeof(b));
    b.det_scale = det ? P.d_det_scale : nullptr;
    b.mode = 0;
    b.cache = nullptr;
    b.part_mode = 0;
    b.part_stride = S.part_total;
    b.part = W.part;
    b.colmap = S.d_colmap;
    b.cspace = S.d_cspace;
    b.cmask = (1 << qc.lc) | (1 << qc.ld);
    if (P.d_incore && qc.cache_doubles) {
      b.mode = 2;
      b.cache = P.d_incore + qc.cache_off;
    }
    b.pairs = S.d_pairs;
    b.prims = S.d_prims;
    b.shells = S.d_shells;
    b.coef = S.d_coef;
    b.quartets = qc.d_quartets;
    b.chunks = qc.d_chunks;
    b.nchunks = (int)qc.chunks.size();
    b.nbra_max = qc.nbra_max;
    b.mols = P.d_molinfo;
    b.n = qc.row_len ? qc.row_len : P.nc_max;
    b.shell_cofs = S.d_shell_cofs;
    b.with_j = with_j;
    b.with_k = with_k;
    b.mixed = mixed ? 1 : 0;
    b.thr = mixed_threshold;
    b.pair_q = S.d_pair_q;
    b.dmax = P.d_dmax;
    b.threads = qc.threads;
    // LDS of one launch over nbd_l exchange densities and nbj_l Coulomb
    // densities; the ket window and the Rys table only where they fit `cap`
    // staged bra primitive pairs (evaluating launches of full-block segmented classes)
    const int stage_cap = b.mode == 2 ? 0 : qc.stage_cap;
    const int stage_rec = qc.pb ? kStagedPrimCoef : kStagedPrim;
    auto layout = [&](int nbd_l, int nbj_l, size_t cap, size_t& lds, int& nket_max, int& tab_lds) {
      lds = sizeof(double) * ((size_t)nbd_l * qc.nbra_max * (wa + wb) * b.n +
                              (size_t)nbj_l * qc.nbra_max * wa * wb * kJBraCopies);
      nket_max = with_j ? qc.nket_max : 0;
      const size_t jket_bytes = sizeof(double) * (size_t)nbj_l * nket_max * wc * wd;
      if (nket_max && lds + jket_bytes <= cap) lds += jket_bytes; else nket_max = 0;
      // the slots' bra density blocks D_ab + D_ba of every Coulomb density
      if (with_j) lds += sizeof(double) * (size_t)nbj_l * qc.nbra_max * wa * wb;
      // (the kernel places the staging area 16-byte aligned after the slabs)
      if (stage_cap) lds = ((lds + 15) & ~(size_t)15) + sizeof(double) * ((size_t)stage_rec * stage_cap + (qc.nbra_max + 2) / 2);
      tab_lds = (b.mode == 2 || det) ? 0 : qc.tab_lds;  // stored blocks need no Rys table
      // (the kernel places the table 16-byte aligned after the staging area)
      const size_t tab_at = (lds + 15) & ~(size_t)15;
      if (tab_lds && tab_at + qc.tab_bytes <= cap) lds = tab_at + qc.tab_bytes; else tab_lds = 0;
    };
    const int nbd_eff = with_k ? nbd : nb;
    size_t lds;
    int nket_max, tab_lds;
    layout(nbd_eff, nb, P.lds_budget, lds, nket_max, tab_lds);
    // every density in one launch within the packing budget, else within the
    // device's LDS; otherwise one launch per density (up to the device's
    // LDS); otherwise, for segmented classes only, the global-atomic kernels
    bool bra_path = qc.d_chunks && lds <= P.lds_budget;
    if (!bra_path && qc.d_chunks && nbd_eff > 1) {
      // several densities: one launch with up to the device's LDS (every
      // quartet evaluated once, at lower occupancy) before one per density
      size_t lds_all;
      int nket_all, tab_all;
      layout(nbd_eff, nb, P.lds_max, lds_all, nket_all, tab_all);
      if (lds_all <= P.lds_max) {
        bra_path = true;
        lds = lds_all;
        nket_max = nket_all;
        tab_lds = tab_all;
      }
    }
    size_t lds1 = 0;
    int nket1 = 0, tab1 = 0;
    if (!bra_path && qc.d_chunks) layout(1, 1, P.lds_max, lds1, nket1, tab1);
    const bool per_density = !bra_path && qc.d_chunks && lds1 <= P.lds_max;
    // partial slabs: a slab range this build does not write (ket windows not
    // used, or the class not run by the bra-major kernel) is cleared first,
    // since the reductions read every listed slab entry
    if (qc.part) {
      const int pm = psums;
      const bool need_k = with_k && (pm & 1), need_j = with_j && (pm & 2);
      if (bra_path) b.part_mode = (need_k ? 1 : 0) | ((need_j && nket_max > 0) ? 2 : 0);
      if ((need_k && !(b.part_mode & 1)) || (need_j && !(b.part_mode & 2)))
        for (int bt = 0; bt < nbd_eff; ++bt)
          hipLaunchKernelGGL(fill_kernel, dim3(nblocks((size_t)(qc.part_end - qc.part_beg))), dim3(256), 0, s,
                             W.part + (size_t)bt * S.part_total + qc.part_beg,
                             (size_t)(qc.part_end - qc.part_beg), 0.0);
    }
    if (bra_path) {
      b.dens = W.dc;
      b.jp = W.jp;
      b.jket_pad = W.jpad;
      b.kp = W.kp;
      b.bstride = P.ncc_tot;
      b.nbd = nbd_eff;
      b.nbj = nb;
      b.tab_lds = tab_lds;
      b.nket_max = nket_max;
      b.bra_stage = stage_cap;
      launch_jk_bra_class(qc.la, qc.lb, qc.lc, qc.ld, qc.pb, qc.pk, b, lds, s);
    } else if (per_density) {
      // one launch per density (J and K for D_b, K only for D_b^T of a
      // nonsymmetric K), each within the device's LDS per workgroup
      for (int bt = 0; bt < nbd_eff; ++bt) {
        b.dens = W.dc + (size_t)bt * ncc;
        b.jp = W.jp + (size_t)bt * ncc;
        b.jket_pad = W.jpad + 8 * (size_t)bt * ncc;
        b.kp = W.kp + (size_t)bt * ncc;
        b.bstride = P.ncc_tot;
        b.nbd = 1;
        b.nbj = 1;
        b.with_j = with_j && bt < nb;
        b.tab_lds = tab1;
        b.nket_max = b.with_j ? nket1 : 0;
        b.bra_stage = stage_cap;
        if (!b.with_j && !b.with_k) continue;
        launch_jk_bra_class(qc.la, qc.lb, qc.lc, qc.ld, qc.pb, qc.pk, b, lds1, s);
      }
    } else if (!qc.pb && !qc.pk) {
      JKArgs a;
      std::memset(&a, 0, sizeof(a));
      a.det_scale = det ? P.d_det_scale : nullptr;
      a.pairs = S.d_pairs;
      a.prims = S.d_prims;
      a.shells = S.d_shells;
      a.quartets = qc.d_quartets;
      a.tasks = qc.d_tasks;
      a.ntasks = (int)qc.tasks.size();
      a.nitems = qc.nitems;
      a.dens = W.dc;
      a.jp = W.jp;
      a.kp = W.kp;
      a.mols = P.d_molinfo;
      a.bstride = P.ncc_tot;
      a.n = P.nc_max;
      a.nbd = nbd_eff;
      a.nbj = nb;
      a.with_j = with_j;
      a.with_k = with_k;
      launch_jk_class(qc.la, qc.lb, qc.lc, qc.ld, a, s);
    } else {
      // general-contraction classes have no global-atomic kernel (the plan
      // keeps them segmented when one density's slot does not fit)
      throw Internal("general-contraction class does not fit in LDS");
    }
    if (P.prof) {
      HIP_OK(hipEventRecord(P.ev_pool[P.ev_used + 1], s));
      P.ev_used += 2;
    }
  }
  if (lane_reduce)
    for (size_t l = 0; l < P.lane_classes.size(); ++l)
      hipLaunchKernelGGL(kpart_reduce_kernel, dim3((unsigned)S.nrows, (unsigned)((S.nc_max + 63) / 64),
                         (unsigned)nbd), dim3(64), 0, l == 0 ? s : P.side[l - 1], W.part, S.part_total,
                         P.d_lane_krp[l], P.d_lane_kent[l], S.d_row_dst, S.d_row_n, S.d_row_mol, S.d_cinv_off,
                         S.d_cinv, P.ncc_tot, 1, W.kp);
  if (fork) {
    for (size_t l = 0; l < P.side.size(); ++l) {
      HIP_OK(hipEventRecord(P.join_ev[l], P.side[l]));
      HIP_OK(hipStreamWaitEvent(s, P.join_ev[l], 0));
    }
  }
  if (S.part_total > 0) {
    const int nbd_k = with_k ? nbd : nb;
    if (with_k && (psums & 1) && !lane_reduce)
      hipLaunchKernelGGL(kpart_reduce_kernel, dim3((unsigned)S.nrows, (unsigned)((S.nc_max + 63) / 64),
                         (unsigned)nbd_k), dim3(64), 0, s, W.part, S.part_total, S.d_krp, S.d_kent,
                         S.d_row_dst, S.d_row_n, S.d_row_mol, S.d_cinv_off, S.d_cinv, P.ncc_tot, 0, W.kp);
    if (with_j && (psums & 2))
      hipLaunchKernelGGL(jpart_reduce_kernel, dim3((unsigned)S.pairs.size(), (unsigned)nb), dim3(64), 0, s,
                         W.part, S.part_total, S.d_jrp, S.d_jent, S.d_pairs, S.d_shells, P.d_molinfo,
                         P.ncc_tot, W.jpad);
  }
  if (det) {
    // integers back to FP64 (Jp / Kp over nbd densities, ket-side J over nb)
    hipLaunchKernelGGL(fixed_to_double_kernel, dim3(nblocks(ncc * nbd)), dim3(256), 0, s, W.jp,
                       (int64_t)(ncc * nbd), 1, P.ncc_tot, P.d_molao, P.nmol, P.d_det_scale);
    hipLaunchKernelGGL(fixed_to_double_kernel, dim3(nblocks(ncc * nbd)), dim3(256), 0, s, W.kp,
                       (int64_t)(ncc * nbd), 1, P.ncc_tot, P.d_molao, P.nmol, P.d_det_scale);
    if (with_j)
      hipLaunchKernelGGL(fixed_to_double_kernel, dim3(nblocks(8 * ncc * nb)), dim3(256), 0, s, W.jpad,
                         (int64_t)(8 * ncc * nb), 8, P.ncc_tot, P.d_molao, P.nmol, P.d_det_scale);
  }
  hipLaunchKernelGGL(finalize_kernel, dim3(nblocks((size_t)(with_j + with_k) * nb * nn)), dim3(256), 0, s,
                     W.jp, W.jpad, W.kp, P.ncc_tot, P.nn_tot, nb, sym ? 1 : 0, with_j, with_k, P.d_molao,
                     P.nmol, P.d_s2c_idx, P.d_s2c_coef, out);
  HIP_OK(hipGetLastError());
}

// Stream balance from measured durations: one untimed build (code objects,
// workspace), then one build with every class launch timed alone by HIP
// events on the caller's stream; the classes are then spread over the
// streams by those times.  The plan-time work model weighs primitive
// quartets and block sizes and misses what bounds the classes with few,
// long work items (latency of the per-lane item chains), which left streams
// idle while others still ran such classes back to back.  The probe density
// is the identity per molecule (the kernels' work does not depend on the
// density's values).
static void balance_lanes(PlanData& P, hipStream_t s) {
  if (P.side.empty() || !P.options.balance || P.jk.classes.empty()) return;
  const size_t nn = (size_t)P.nn_tot;
  std::vector<double> eye(nn, 0.0);
  for (int m = 0; m < P.nmol; ++m)
    for (int i = 0; i < P.mols[m].n; ++i) eye[(size_t)P.mols[m].sbase + (size_t)i * P.mols[m].n + i] = 1.0;
  double* d_dens = dev_upload(eye, s);
  double* d_out = nullptr;
  HIP_OK(hipMalloc(&d_out, sizeof(double) * 2 * nn));
  P.prof = true;
  P.prof_recs.clear();
  P.prof_builds = 0;
  P.ev_used = 0;
  std::vector<double> t(P.jk.classes.size(), 0.0);
  try {
    jk_build(P, d_dens, 1, d_out, 1, 1, 0u, 0.0, s);
    P.prof_recs.clear();
    P.ev_used = 0;
    jk_build(P, d_dens, 1, d_out, 1, 1, 0u, 0.0, s);
    for (const auto& r : P.prof_recs) {
      HIP_OK(hipEventSynchronize(P.ev_pool[r.ev + 1]));
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, P.ev_pool[r.ev], P.ev_pool[r.ev + 1]));
      t[r.qci] += ms;
    }
    HIP_OK(hipStreamSynchronize(s));
  } catch (...) {
    P.prof = false;
    (void)hipFree(d_dens);
    (void)hipFree(d_out);
    throw;
  }
  P.prof = false;
  P.prof_recs.clear();
  P.prof_builds = 0;
  P.ev_used = 0;
  HIP_OK(hipFree(d_dens));
  HIP_OK(hipFree(d_out));
  for (size_t i = 0; i < t.size(); ++i) P.jk.classes[i].cost = t[i];
  assign_lanes(P);
}

// In-core integrals: the contracted Cartesian blocks of every quartet of the
// segmented bra-major classes, stored once (quartet-major per class, a split
// class's block as its ket-component slices) so that later builds digest
// them instead of evaluating them.  Classes without bra-major chunks or with
// general-contraction patterns keep evaluating.
static size_t incore_layout(PlanData& P) {
  size_t total = 0;
  for (QuartetClass& qc : P.jk.classes) {
    qc.cache_off = total;
    qc.cache_doubles = 0;
    if (qc.pb || qc.pk || !qc.d_chunks || qc.nbra_max < 1) continue;
    qc.cache_doubles = (size_t)qc.nq * ncart_h(qc.la) * ncart_h(qc.lb) * ncart_h(qc.lc) * ncart_h(qc.ld);
    total += qc.cache_doubles;
  }
  return total;
}

static void incore_set(PlanData& P, int enable, hipStream_t s) {
  if (enable != 0 && enable != 1) throw InvalidArgument("in-core flag must be 0 or 1");
  if (!enable) {
    if (P.d_incore) {
      HIP_OK(hipStreamSynchronize(s));
      HIP_OK(hipFree(P.d_incore));
    }
    P.d_incore = nullptr;
    P.incore_doubles = 0;
    for (QuartetClass& qc : P.jk.classes) qc.cache_doubles = 0;
    return;
  }
  if (P.d_incore) return;
  const size_t total = incore_layout(P);
  if (total == 0) throw InvalidArgument("no quartet class of this plan can keep its integrals in core");
  double* buf = nullptr;
  if (hipMalloc(&buf, sizeof(double) * total) != hipSuccess) {
    (void)hipGetLastError();
    for (QuartetClass& qc : P.jk.classes) qc.cache_doubles = 0;
    throw InvalidArgument("in-core integrals need " + std::to_string(sizeof(double) * total) +
                          " bytes of device memory, which could not be allocated");
  }
  const PairSet& S = P.jk;
  for (const QuartetClass& qc : S.classes) {
    if (!qc.cache_doubles) continue;
    BraArgs b;
    std::memset(&b, 0, sizeof(b));
    b.mode = 1;
    b.cache = buf + qc.cache_off;
    b.pairs = S.d_pairs;
    b.prims = S.d_prims;
    b.shells = S.d_shells;
    b.coef = S.d_coef;
    b.quartets = qc.d_quartets;
    b.chunks = qc.d_chunks;
    b.nchunks = (int)qc.chunks.size();
    b.nbra_max = qc.nbra_max;
    b.mols = P.d_molinfo;
    b.n = P.nc_max;
    b.pair_q = S.d_pair_q;
    b.threads = qc.threads;
    b.bstride = P.ncc_tot;
    // no densities: the pass only evaluates and stores; it stages the bra
    // primitive pairs like the evaluating builds, so that both sum every
    // quartet's primitives in one order (the deterministic mode's in-core
    // builds equal its direct builds bit for bit)
    b.bra_stage = qc.stage_cap;
    const size_t lds = qc.stage_cap ? sizeof(double) * ((size_t)(qc.pb ? kStagedPrimCoef : kStagedPrim) * qc.stage_cap +
                                                        (qc.nbra_max + 2) / 2) : 0;
    launch_jk_bra_class(qc.la, qc.lb, qc.lc, qc.ld, qc.pb, qc.pk, b, lds, s);
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(s));
  P.d_incore = buf;
  P.incore_doubles = total;
}

static void int1e(PlanData& P, double* s_out, double* t_out, double* v_out, hipStream_t st) {
  if (!s_out || !t_out || !v_out) throw InvalidArgument("int1e requires three device output buffers");
  const size_t ncc = (size_t)P.ncc_tot;
  double* tmp = nullptr;
  HIP_OK(hipMalloc(&tmp, sizeof(double) * 3 * ncc));
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks(3 * ncc)), dim3(256), 0, st, tmp, 3 * ncc, 0.0);
  static const int pl[kNumPairClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
  for (int c = 0; c < kNumPairClasses; ++c) {
    const int b = P.p1e_begin[c], e = P.p1e_begin[c + 1];
    launch_int1e_class(pl[c][0], pl[c][1], P.d_pairs1e + b, e - b, P.d_prims, P.d_shells, P.d_nuclei, P.nnuc,
                       P.d_molinfo, tmp, tmp + ncc, tmp + 2 * ncc, st);
  }
  double* outs[3] = {s_out, t_out, v_out};
  for (int k = 0; k < 3; ++k)
    hipLaunchKernelGGL(sandwich_kernel, dim3(nblocks((size_t)P.nn_tot)), dim3(256), 0, st, tmp + k * ncc,
                       P.nn_tot, P.d_molao, P.nmol, P.d_s2c_idx, P.d_s2c_coef, outs[k]);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(st));
  HIP_OK(hipFree(tmp));
}

// One-electron gradient sum D d(T+V)/dR - W dS/dR into grad[natm][3] (+=).
static void int1e_gradient(PlanData& P, const double* dens, const double* wdens, double* grad, hipStream_t s) {
  if (!dens || !wdens || !grad) throw InvalidArgument("int1e gradient needs D, W and a gradient buffer");
  if ((size_t)3 * P.natm * sizeof(double) > kLdsBudget) throw InvalidArgument("too many atoms for the gradient kernel");
  const size_t ncc = (size_t)P.ncc_tot;
  ensure_workspace(P, 2);
  const double* src[2] = {dens, wdens};
  for (int k = 0; k < 2; ++k)
    hipLaunchKernelGGL(dens_to_cart_kernel, dim3(nblocks(ncc)), dim3(256), 0, s, src[k], P.nn_tot, 1,
                       P.d_dc + k * ncc, P.ncc_tot, P.d_molao, P.nmol, P.d_c2s_idx, P.d_c2s_coef, 1);
  static const int pl[kNumPairClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
  for (int c = 0; c < kNumPairClasses; ++c) {
    const int b = P.p1e_begin[c], e = P.p1e_begin[c + 1];
    launch_grad1e_class(pl[c][0], pl[c][1], P.d_pairs1e + b, e - b, P.d_prims, P.d_shells, P.d_seg_atom, P.d_nuclei,
                        P.nnuc, P.d_molinfo, P.d_dc, P.d_dc + ncc, P.natm, grad, s);
  }
  HIP_OK(hipGetLastError());
}

static void geometry_vjp(PlanData& P, const double* dens, const double* cj, const double* ck, double* grad,
                         hipStream_t s) {
  if (P.cutoff > 0) throw InvalidArgument("Screened direct J/K geometry AD requires cutoff=0");
  if (!dens || !grad || (!cj && !ck)) throw InvalidArgument("geometry VJP needs a density, a gradient buffer and a cotangent");
  const size_t ncc = (size_t)P.ncc_tot;
  ensure_workspace(P, 3);
  const double* src[3] = {dens, cj, ck};
  for (int k = 0; k < 3; ++k)
    if (src[k])
      hipLaunchKernelGGL(dens_to_cart_kernel, dim3(nblocks(ncc)), dim3(256), 0, s, src[k], P.nn_tot, 1,
                         P.d_dc + k * ncc, P.ncc_tot, P.d_molao, P.nmol, P.d_c2s_idx, P.d_c2s_coef, 1);
  const PairSet& S = ensure_seg_set(P, s);
  for (const QuartetClass& qc : S.classes) {
    GradArgs a;
    a.pairs = S.d_pairs;
    a.prims = P.d_prims;
    a.shells = P.d_shells;
    a.seg_atom = P.d_seg_atom;
    a.quartets = qc.d_quartets;
    a.tasks = qc.d_tasks;
    a.ntasks = (int)qc.tasks.size();
    a.nitems = qc.nitems;
    a.dens = P.d_dc;
    a.cj = cj ? P.d_dc + ncc : nullptr;
    a.ck = ck ? P.d_dc + 2 * ncc : nullptr;
    a.mols = P.d_molinfo;
    a.natm = P.natm;
    a.grad = grad;
    if ((size_t)3 * P.natm * sizeof(double) > kLdsBudget) throw InvalidArgument("too many atoms for the gradient kernel");
    launch_grad_class(qc.la, qc.lb, qc.lc, qc.ld, a, s);
  }
  HIP_OK(hipGetLastError());
}

// Basis-parameter VJP of J/K into env_grad[nenv] (+=): exponent and
// coefficient slots of every shell.
static void jk_param_vjp(PlanData& P, const double* dens, const double* cj, const double* ck, double* env_grad,
                         hipStream_t s) {
  if (P.cutoff > 0) throw InvalidArgument("Screened direct J/K basis-parameter AD requires cutoff=0");
  if (!dens || !env_grad || (!cj && !ck))
    throw InvalidArgument("basis-parameter VJP needs a density, an env gradient buffer and a cotangent");
  const int nprim = P.seg_prim0.back();
  if ((size_t)2 * nprim * sizeof(double) > kLdsBudget)
    throw InvalidArgument("too many primitives for the basis-parameter kernel");
  const size_t ncc = (size_t)P.ncc_tot;
  ensure_workspace(P, 3);
  if (!P.d_gparam) HIP_OK(hipMalloc(&P.d_gparam, sizeof(double) * 2 * std::max(nprim, 1)));
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks(2 * (size_t)nprim)), dim3(256), 0, s, P.d_gparam,
                     2 * (size_t)nprim, 0.0);
  const double* src[3] = {dens, cj, ck};
  for (int k = 0; k < 3; ++k)
    if (src[k])
      hipLaunchKernelGGL(dens_to_cart_kernel, dim3(nblocks(ncc)), dim3(256), 0, s, src[k], P.nn_tot, 1,
                         P.d_dc + k * ncc, P.ncc_tot, P.d_molao, P.nmol, P.d_c2s_idx, P.d_c2s_coef, 1);
  const PairSet& S = ensure_seg_set(P, s);
  for (const QuartetClass& qc : S.classes) {
    ParamArgs a;
    a.pairs = S.d_pairs;
    a.prims = P.d_prims;
    a.aux = P.d_aux;
    a.shells = P.d_shells;
    a.seg_prim0 = P.d_seg_prim0;
    a.quartets = qc.d_quartets;
    a.tasks = qc.d_tasks;
    a.ntasks = (int)qc.tasks.size();
    a.nitems = qc.nitems;
    a.dens = P.d_dc;
    a.cj = cj ? P.d_dc + ncc : nullptr;
    a.ck = ck ? P.d_dc + 2 * ncc : nullptr;
    a.mols = P.d_molinfo;
    a.nprim = nprim;
    a.gcoef = P.d_gparam;
    a.gexp = P.d_gparam + nprim;
    launch_param_class(qc.la, qc.lb, qc.lc, qc.ld, a, s);
  }
  hipLaunchKernelGGL(param_to_env_kernel, dim3(nblocks((size_t)nprim)), dim3(256), 0, s, nprim, P.d_prim_slots,
                     P.d_prim_fsp, P.d_gparam, P.d_gparam + nprim, env_grad);
  HIP_OK(hipGetLastError());
}

// Basis-parameter derivatives of sum D (T + V) - W S into env_grad (+=).
static void int1e_param_vjp(PlanData& P, const double* dens, const double* wdens, double* env_grad, hipStream_t s) {
  if (!dens || !wdens || !env_grad) throw InvalidArgument("int1e parameter VJP needs D, W and an env gradient buffer");
  const int nprim = P.seg_prim0.back();
  if ((size_t)2 * nprim * sizeof(double) > kLdsBudget)
    throw InvalidArgument("too many primitives for the basis-parameter kernel");
  const size_t ncc = (size_t)P.ncc_tot;
  ensure_workspace(P, 2);
  if (!P.d_gparam) HIP_OK(hipMalloc(&P.d_gparam, sizeof(double) * 2 * std::max(nprim, 1)));
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks(2 * (size_t)nprim)), dim3(256), 0, s, P.d_gparam,
                     2 * (size_t)nprim, 0.0);
  const double* src[2] = {dens, wdens};
  for (int k = 0; k < 2; ++k)
    hipLaunchKernelGGL(dens_to_cart_kernel, dim3(nblocks(ncc)), dim3(256), 0, s, src[k], P.nn_tot, 1,
                       P.d_dc + k * ncc, P.ncc_tot, P.d_molao, P.nmol, P.d_c2s_idx, P.d_c2s_coef, 1);
  static const int pl[kNumPairClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
  for (int c = 0; c < kNumPairClasses; ++c) {
    const int b = P.p1e_begin[c], e = P.p1e_begin[c + 1];
    launch_param1e_class(pl[c][0], pl[c][1], P.d_pairs1e + b, e - b, P.d_prims, P.d_aux, P.d_shells,
                         P.d_seg_prim0, P.d_nuclei, P.nnuc, P.d_molinfo, P.d_dc, P.d_dc + ncc, nprim, P.d_gparam,
                         P.d_gparam + nprim, s);
  }
  hipLaunchKernelGGL(param_to_env_kernel, dim3(nblocks((size_t)nprim)), dim3(256), 0, s, nprim, P.d_prim_slots,
                     P.d_prim_fsp, P.d_gparam, P.d_gparam + nprim, env_grad);
  HIP_OK(hipGetLastError());
}

static void eri_blocks(PlanData& P, const int32_t* shls, int64_t nq, const int64_t* offsets, double* out,
                       hipStream_t s) {
  if (nq < 0 || (nq > 0 && (!shls || !offsets || !out))) throw InvalidArgument("Invalid ERI block request");
  struct Item {
    std::array<int, 4> cls;
    BlockTask t;
  };
  std::vector<Item> items;
  std::vector<int> seg_first(P.nbas + 1, 0);
  for (size_t i = 0; i < P.seg.size(); ++i) seg_first[P.seg[i].parent + 1] = (int)i + 1;
  for (int i = 0; i < P.nbas; ++i) seg_first[i + 1] = std::max(seg_first[i + 1], seg_first[i]);
  std::vector<int64_t> offs(offsets, offsets + nq);
  for (int64_t q = 0; q < nq; ++q) {
    int sh[4], dims[4];
    for (int k = 0; k < 4; ++k) {
      sh[k] = shls[q * 4 + k];
      if (sh[k] < 0 || sh[k] >= P.nbas) throw InvalidArgument("ERI block shell index out of range");
      if (P.seg[seg_first[sh[k]]].mol != P.seg[seg_first[sh[0]]].mol)
        throw InvalidArgument("ERI block shells must belong to one molecule");
      const int l = P.bas[sh[k] * BAS_SLOTS + ANG_OF], nc = P.bas[sh[k] * BAS_SLOTS + NCTR_OF];
      dims[k] = nc * ncart_h(l);
    }
    const int stride_req[4] = {1, dims[0], dims[0] * dims[1], dims[0] * dims[1] * dims[2]};
    const int nc0 = P.bas[sh[0] * BAS_SLOTS + NCTR_OF], nc1 = P.bas[sh[1] * BAS_SLOTS + NCTR_OF];
    const int nc2 = P.bas[sh[2] * BAS_SLOTS + NCTR_OF], nc3 = P.bas[sh[3] * BAS_SLOTS + NCTR_OF];
    for (int c0 = 0; c0 < nc0; ++c0)
      for (int c1 = 0; c1 < nc1; ++c1)
        for (int c2 = 0; c2 < nc2; ++c2)
          for (int c3 = 0; c3 < nc3; ++c3) {
            const int sg[4] = {seg_first[sh[0]] + c0, seg_first[sh[1]] + c1, seg_first[sh[2]] + c2,
                               seg_first[sh[3]] + c3};
            const int cf[4] = {c0, c1, c2, c3};
            // positions of each requested slot's segment inside its pair
            auto pair_of = [&](int x, int y, int& idx, int& first_pos, int& second_pos) {
              const int S = std::max(sg[x], sg[y]), T = std::min(sg[x], sg[y]);
              idx = (int)pair_idx(P, S, T);
              const Pair& pr = P.all_pairs[idx];
              if (pr.sa == sg[x] && (pr.sb == sg[y])) {
                first_pos = x;
                second_pos = y;
              } else {
                first_pos = y;
                second_pos = x;
              }
            };
            int i01, a01, b01, i23, a23, b23;
            pair_of(0, 1, i01, a01, b01);
            pair_of(2, 3, i23, a23, b23);
            int bra = i01, ket = i23, pos[4] = {a01, b01, a23, b23};
            if (P.all_class[i23] > P.all_class[i01]) {
              bra = i23;
              ket = i01;
              pos[0] = a23; pos[1] = b23; pos[2] = a01; pos[3] = b01;
            }
            Item it;
            it.t.bra = bra;
            it.t.ket = ket;
            it.t.base = offs[q];
            for (int k = 0; k < 4; ++k) {
              const int l = P.seg[sg[pos[k]]].l;
              it.t.stride[k] = stride_req[pos[k]];
              it.t.coff[k] = cf[pos[k]] * ncart_h(l);
            }
            const Pair& bp = P.all_pairs[bra];
            const Pair& kp = P.all_pairs[ket];
            it.cls = {P.seg[bp.sa].l, P.seg[bp.sb].l, P.seg[kp.sa].l, P.seg[kp.sb].l};
            items.push_back(it);
          }
  }
  std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cls < y.cls; });
  std::vector<BlockTask> tasks;
  std::vector<std::array<int, 4>> cls;
  for (auto& it : items) {
    tasks.push_back(it.t);
    cls.push_back(it.cls);
  }
  run_blocks(P, P.d_all_pairs, tasks, cls, out, s);
}

// ----------------------------------------------------------------------------
// Packed ERIs and packed J/K (the reference's compact-integral and packed J/K
// targets, csrc/packed_eri.cc:5-14, csrc/compact.cc:5-17, csrc/packed_jk.cc:6-79)
// ----------------------------------------------------------------------------
struct PackTask {
  int64_t off;   // first requested-basis element of the block (element enumeration)
  int64_t coff;  // first element of the staged Cartesian block
  int ao[4];     // first AO of each shell (requested basis) and its width
  int d[4];
  int cao[4];    // first Cartesian AO of each shell and its width
  int cd[4];
};

__device__ __forceinline__ int64_t tri_row(int64_t idx) {
  // largest r with r (r + 1) / 2 <= idx
  int64_t r = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > idx) --r;
  while ((r + 1) * (r + 2) / 2 <= idx) ++r;
  return r;
}

// One thread per block element in the requested basis: unique AO quartets
// (a >= b, c >= d, ab >= cd) go to the packed layout (0: pair matrix, both
// triangles; 1: s8 triangle), transformed from the staged Cartesian block by
// the plan's spherical-to-Cartesian map (up to 3 terms per AO; identity for
// Cartesian plans).
__global__ void pack_eri_kernel(const double* __restrict__ blocks, const PackTask* __restrict__ tasks,
                                int64_t ntask, int64_t total, int layout, int64_t np,
                                const int* __restrict__ s2c_idx, const double* __restrict__ s2c_coef,
                                double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int64_t lo = 0, hi = ntask;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (tasks[mid].off <= e) lo = mid; else hi = mid;
  }
  const PackTask t = tasks[lo];
  int64_t r = e - t.off;  // libcint layout: first index fastest
  const int p = (int)(r % t.d[0]);
  r /= t.d[0];
  const int q = (int)(r % t.d[1]);
  r /= t.d[1];
  const int u = (int)(r % t.d[2]);
  const int w = (int)(r / t.d[2]);
  const int64_t a = t.ao[0] + p, b = t.ao[1] + q, c = t.ao[2] + u, d = t.ao[3] + w;
  if (a < b || c < d) return;
  const int64_t ab = a * (a + 1) / 2 + b, cd = c * (c + 1) / 2 + d;
  if (ab < cd) return;
  const int64_t g[4] = {a, b, c, d};
  int ci[4][3];
  double cc[4][3];
  for (int x = 0; x < 4; ++x)
    for (int m = 0; m < 3; ++m) {
      ci[x][m] = s2c_idx[g[x] * 3 + m] - t.cao[x];
      cc[x][m] = s2c_coef[g[x] * 3 + m];
    }
  const double* __restrict__ blk = blocks + t.coff;
  double v = 0.0;
  for (int m3 = 0; m3 < 3; ++m3) {
    if (cc[3][m3] == 0.0) continue;
    for (int m2 = 0; m2 < 3; ++m2) {
      if (cc[2][m2] == 0.0) continue;
      for (int m1 = 0; m1 < 3; ++m1) {
        if (cc[1][m1] == 0.0) continue;
        const int64_t base = ci[1][m1] + (int64_t)t.cd[1] * (ci[2][m2] + (int64_t)t.cd[2] * ci[3][m3]);
        const double f = cc[1][m1] * cc[2][m2] * cc[3][m3];
        for (int m0 = 0; m0 < 3; ++m0)
          if (cc[0][m0] != 0.0) v += cc[0][m0] * f * blk[ci[0][m0] + (int64_t)t.cd[0] * base];
      }
    }
  }
  if (layout == 0) {
    out[ab * np + cd] = v;
    out[cd * np + ab] = v;
  } else {
    out[ab * (ab + 1) / 2 + cd] = v;
  }
}

// One thread per stored element (ab >= cd) of the s8 triangle or of the lower
// triangle of the pair matrix: the 8-fold digestion of the reference's packed
// J/K into out[op][b][n][n] with device-scope atomics.
__global__ void packed_jk_kernel(const double* __restrict__ eri, int s8, int64_t np, const double* __restrict__ D,
                                 int nb, int n, int with_j, int with_k, double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= np * (np + 1) / 2) return;
  const int64_t ab = tri_row(idx), cd = idx - ab * (ab + 1) / 2;
  const double v = s8 ? eri[idx] : eri[ab * np + cd];
  if (v == 0.0) return;
  const int a = (int)tri_row(ab), b = (int)(ab - (int64_t)a * (a + 1) / 2);
  const int c = (int)tri_row(cd), d = (int)(cd - (int64_t)c * (c + 1) / 2);
  const size_t nn = (size_t)n * n;
  for (int bt = 0; bt < nb; ++bt) {
    const double* dm = D + (size_t)bt * nn;
    if (with_j) {
      double* j = out + (size_t)bt * nn;
      const double jab = v * (dm[(size_t)c * n + d] + (c != d ? dm[(size_t)d * n + c] : 0.0));
      atomicAdd(j + (size_t)a * n + b, jab);
      if (a != b) atomicAdd(j + (size_t)b * n + a, jab);
      if (ab != cd) {
        const double jcd = v * (dm[(size_t)a * n + b] + (a != b ? dm[(size_t)b * n + a] : 0.0));
        atomicAdd(j + (size_t)c * n + d, jcd);
        if (c != d) atomicAdd(j + (size_t)d * n + c, jcd);
      }
    }
    if (with_k) {
      double* k = out + (size_t)(with_j * nb + bt) * nn;
      atomicAdd(k + (size_t)a * n + c, v * dm[(size_t)b * n + d]);
      if (c != d) atomicAdd(k + (size_t)a * n + d, v * dm[(size_t)b * n + c]);
      if (a != b) atomicAdd(k + (size_t)b * n + c, v * dm[(size_t)a * n + d]);
      if (a != b && c != d) atomicAdd(k + (size_t)b * n + d, v * dm[(size_t)a * n + c]);
      if (ab != cd) {
        atomicAdd(k + (size_t)c * n + a, v * dm[(size_t)d * n + b]);
        if (a != b) atomicAdd(k + (size_t)c * n + b, v * dm[(size_t)d * n + a]);
        if (c != d) atomicAdd(k + (size_t)d * n + a, v * dm[(size_t)c * n + b]);
        if (a != b && c != d) atomicAdd(k + (size_t)d * n + b, v * dm[(size_t)c * n + a]);
      }
    }
  }
}

// Every unique AO quartet of the plan's molecule (no screening), staged shell
// quartet by shell quartet in the reference's visiting order (i, j <= i,
// k <= i, l <= k; csrc/coulomb_shells.h:98-125) through the block kernels.
static void packed_eri(PlanData& P, int layout, double* out, int64_t size, hipStream_t s) {
  if (layout != 0 && layout != 1) throw Internal("Unknown compact integral layout");
  if (P.nmol != 1) throw Internal("Packed ERIs need a single-molecule plan");
  const int64_t n = P.n, np = n * (n + 1) / 2;
  if (!out || size != (layout == 0 ? np * np : np * (np + 1) / 2)) throw Internal("Packed ERI shape mismatch");
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks((size_t)size)), dim3(256), 0, s, out, (size_t)size, 0.0);
  constexpr int64_t kStage = int64_t(1) << 24;  // doubles of staged blocks per pass
  double* stage = nullptr;
  HIP_OK(hipMalloc(&stage, sizeof(double) * kStage));
  std::vector<int32_t> shls;
  std::vector<int64_t> offs;
  std::vector<PackTask> tasks;
  int64_t tot = 0, ctot = 0;  // requested-basis elements, staged Cartesian doubles
  auto flush = [&]() {
    if (tasks.empty()) return;
    eri_blocks(P, shls.data(), (int64_t)tasks.size(), offs.data(), stage, s);
    PackTask* d_tasks = dev_upload(tasks, s);
    hipLaunchKernelGGL(pack_eri_kernel, dim3(nblocks((size_t)tot)), dim3(256), 0, s, stage, d_tasks,
                       (int64_t)tasks.size(), tot, layout, np, P.d_s2c_idx, P.d_s2c_coef, out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipFree(d_tasks));
    shls.clear();
    offs.clear();
    tasks.clear();
    tot = 0;
    ctot = 0;
  };
  const int nbas = P.nbas;
  for (int i = 0; i < nbas; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= i; ++k)
        for (int l = 0; l <= k; ++l) {
          const int sh[4] = {i, j, k, l};
          PackTask t;
          int64_t sz = 1, csz = 1;
          for (int x = 0; x < 4; ++x) {
            t.ao[x] = P.ao[sh[x]];
            t.d[x] = P.ao[sh[x] + 1] - P.ao[sh[x]];
            t.cao[x] = P.ao_cart[sh[x]];
            t.cd[x] = P.ao_cart[sh[x] + 1] - P.ao_cart[sh[x]];
            sz *= t.d[x];
            csz *= t.cd[x];
          }
          if (ctot + csz > kStage) flush();
          t.off = tot;
          t.coff = ctot;
          for (int x = 0; x < 4; ++x) shls.push_back(sh[x]);
          offs.push_back(ctot);
          tasks.push_back(t);
          tot += sz;
          ctot += csz;
        }
  flush();
  HIP_OK(hipFree(stage));
}

// ----------------------------------------------------------------------------
// Density-fitting integrals: the reference's compact layouts 2 (two-centre
// metric (P|Q), m x m) and 3 (three-centre (ab|P), m x n(n+1)/2 over AO pairs
// a >= b at pair_index(a, b)), csrc/density_fitting.cc:5-34.  The tables hold
// orbital shells [0, split) and auxiliary shells [split, nbas).  A one-centre
// factor P(r) is evaluated as the pair of P with a unit function (one
// exponent-0 s primitive, coefficient 1, at P's centre), so the two- and
// three-centre integrals run through the four-centre block kernels with the
// libcint conventions of the two- and three-centre drivers.
// ----------------------------------------------------------------------------
constexpr int kMaxAuxL = 4;
constexpr int kDfMaxTerms = 6;  // Cartesian terms of one real solid harmonic up to g

struct DfPack {
  int64_t off;   // first requested-basis element of the block
  int64_t coff;  // first element of the staged Cartesian block
  int ao[3];     // first AO (requested basis) and width of the three slots
  int d[3];
  int cao[3];    // first Cartesian AO and width
  int cd[3];
};

static void launch_df_blocks(int la, int lb, int lc, const BlockArgs& a, hipStream_t s) {
  bool ok = false;
  switch (lc) {
    case 0: ok = launch_df_blocks_c0(la, lb, a, s); break;
    case 1: ok = launch_df_blocks_c1(la, lb, a, s); break;
    case 2: ok = launch_df_blocks_c2(la, lb, a, s); break;
    case 3: ok = launch_df_blocks_c3(la, lb, a, s); break;
    case 4: ok = launch_df_blocks_c4(la, lb, a, s); break;
    default: break;
  }
  if (!ok) throw Internal("no density-fitting kernel for this angular-momentum class");
}

// One thread per requested-basis element of a staged block: the solid-harmonic
// (or identity) transform of every slot, then the scatter into the layout.
// layout 3: out[(c - n) np + a (a + 1) / 2 + b] for a >= b (slots: a, b, c);
// layout 2: out[(a - n) m + (c - n)] and its transpose for a >= c (slots a, -, c).
__global__ void df_pack_kernel(const double* __restrict__ stage, const DfPack* __restrict__ tasks, int64_t ntask,
                               int64_t total, int layout, int64_t n, int64_t m, int64_t np,
                               const int* __restrict__ s2c_idx, const double* __restrict__ s2c_coef,
                               double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int64_t lo = 0, hi = ntask;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (tasks[mid].off <= e) lo = mid; else hi = mid;
  }
  const DfPack t = tasks[lo];
  int64_t r = e - t.off;  // first slot fastest
  const int p = (int)(r % t.d[0]);
  r /= t.d[0];
  const int q = (int)(r % t.d[1]);
  const int u = (int)(r / t.d[1]);
  const int64_t a = t.ao[0] + p, b = t.ao[1] + q, c = t.ao[2] + u;
  if (layout == 3 && a < b) return;
  if (layout == 2 && a < c) return;
  const double* __restrict__ blk = stage + t.coff;
  const int64_t g[3] = {a, layout == 3 ? b : -1, c};
  double v = 0.0;
  for (int k2 = 0; k2 < kDfMaxTerms; ++k2) {
    const double c2 = s2c_coef[g[2] * kDfMaxTerms + k2];
    if (c2 == 0.0) continue;
    const int64_t i2 = s2c_idx[g[2] * kDfMaxTerms + k2] - t.cao[2];
    for (int k1 = 0; k1 < (g[1] < 0 ? 1 : kDfMaxTerms); ++k1) {
      double c1 = 1.0;
      int64_t i1 = 0;
      if (g[1] >= 0) {
        c1 = s2c_coef[g[1] * kDfMaxTerms + k1];
        if (c1 == 0.0) continue;
        i1 = s2c_idx[g[1] * kDfMaxTerms + k1] - t.cao[1];
      }
      const double f = c1 * c2;
      const int64_t base = (int64_t)t.cd[0] * (i1 + (int64_t)t.cd[1] * i2);
      for (int k0 = 0; k0 < kDfMaxTerms; ++k0) {
        const double c0 = s2c_coef[g[0] * kDfMaxTerms + k0];
        if (c0 != 0.0) v += c0 * f * blk[base + s2c_idx[g[0] * kDfMaxTerms + k0] - t.cao[0]];
      }
    }
  }
  if (layout == 3) {
    out[(c - n) * np + a * (a + 1) / 2 + b] = v;
  } else {
    out[(a - n) * m + (c - n)] = v;
    out[(c - n) * m + (a - n)] = v;
  }
}

static void density_fitting(const int32_t* atm, int natm, const int32_t* bas, int nbas, const double* env, int nenv,
                            int cart, int layout, int split, double* out, int64_t size, hipStream_t s) {
  if (layout != 2 && layout != 3) throw Internal("Unknown compact integral layout");
  PlanData P;
  validate_tables(P, atm, natm, bas, nbas, env, nenv, cart, 1, nullptr, nullptr, kMaxAuxL);
  if (split < 1 || split >= nbas) throw Internal("Invalid orbital/auxiliary shell split");
  for (int sh = 0; sh < split; ++sh)
    if (P.bas[(size_t)sh * BAS_SLOTS + ANG_OF] > kMaxL)
      throw Internal("Density-fitting orbital shells are supported up to d (l <= 2)");
  const int64_t n = P.ao[split], m = P.n - n, np = n * (n + 1) / 2;
  if (!out || size != (layout == 2 ? m * m : m * np)) throw Internal("Auxiliary integral shape mismatch");
  build_segments(P);
  // unit functions: one exponent-0 s segment at the centre of every auxiliary segment
  const int nseg0 = (int)P.seg.size();
  std::vector<int> unit_of(nseg0, -1);
  for (int i = 0; i < nseg0; ++i) {
    if (P.seg[i].parent < split) continue;
    Seg g;
    g.parent = -1;
    g.ctr = 0;
    g.l = 0;
    g.np = 1;
    for (int x = 0; x < 3; ++x) g.r[x] = P.seg[i].r[x];
    g.ao_cart = 0;
    g.mol = 0;
    g.lseg = 0;
    g.a = {0.0};
    g.c = {1.0};
    unit_of[i] = (int)P.seg.size();
    P.seg.push_back(std::move(g));
  }
  std::vector<int> seg_first(nbas, 0);
  for (int i = nseg0 - 1; i >= 0; --i) seg_first[P.seg[i].parent] = i;
  // pairs: the orbital segment pairs and every auxiliary segment with its unit
  // function; make_pair wants the higher angular momentum first
  std::vector<Pair> pairs;
  std::map<std::pair<int, int>, int> pair_of;
  auto get_pair = [&](int a, int b) {
    if (P.seg[b].l > P.seg[a].l) std::swap(a, b);
    auto it = pair_of.find({a, b});
    if (it != pair_of.end()) return it->second;
    Pair pr;
    make_pair(P, a, b, pr);
    pairs.push_back(pr);
    pair_of[{a, b}] = (int)pairs.size() - 1;
    return (int)pairs.size() - 1;
  };
  struct Item {
    std::array<int, 3> cls;
    BlockTask t;
    int blk;  // index of the requested block (DfPack)
  };
  std::vector<Item> items;
  std::vector<DfPack> packs;
  auto ncomp = [&](int sh) { return P.ao_cart[sh + 1] - P.ao_cart[sh]; };
  auto add_block = [&](int i, int j, int k) {
    // requested block (i, j | k) (layout 3) or (i | k) (layout 2, j < 0)
    DfPack pk;
    const int sh[3] = {i, j, k};
    for (int x = 0; x < 3; ++x) {
      if (sh[x] < 0) {
        pk.ao[x] = -1; pk.d[x] = 1; pk.cao[x] = 0; pk.cd[x] = 1;
        continue;
      }
      pk.ao[x] = P.ao[sh[x]];
      pk.d[x] = P.ao[sh[x] + 1] - P.ao[sh[x]];
      pk.cao[x] = P.ao_cart[sh[x]];
      pk.cd[x] = ncomp(sh[x]);
    }
    packs.push_back(pk);
    const int bi = (int)packs.size() - 1;
    const int st[3] = {1, pk.cd[0], pk.cd[0] * pk.cd[1]};
    const int nci = P.bas[(size_t)i * BAS_SLOTS + NCTR_OF], ncj = j < 0 ? 1 : P.bas[(size_t)j * BAS_SLOTS + NCTR_OF];
    const int nck = P.bas[(size_t)k * BAS_SLOTS + NCTR_OF];
    for (int ci = 0; ci < nci; ++ci)
      for (int cj = 0; cj < ncj; ++cj)
        for (int ck = 0; ck < nck; ++ck) {
          const int si = seg_first[i] + ci, sk = seg_first[k] + ck;
          const int sj = j < 0 ? unit_of[si] : seg_first[j] + cj;
          Item it;
          it.blk = bi;
          const int bra = get_pair(si, sj);
          const int ket = get_pair(sk, unit_of[sk]);
          it.t.bra = bra;
          it.t.ket = ket;
          it.t.base = 0;  // set per staging pass
          const bool swapped = pairs[bra].sa != si;
          const int off_i = ci * ncart_h(P.seg[si].l), off_j = j < 0 ? 0 : cj * ncart_h(P.seg[sj].l);
          it.t.stride[0] = swapped ? (j < 0 ? 0 : st[1]) : st[0];
          it.t.coff[0] = swapped ? off_j : off_i;
          it.t.stride[1] = swapped ? st[0] : (j < 0 ? 0 : st[1]);
          it.t.coff[1] = swapped ? off_i : off_j;
          it.t.stride[2] = st[2];
          it.t.coff[2] = ck * ncart_h(P.seg[sk].l);
          it.t.stride[3] = 0;
          it.t.coff[3] = 0;
          it.cls = {P.seg[pairs[bra].sa].l, P.seg[pairs[bra].sb].l, P.seg[sk].l};
          items.push_back(it);
        }
  };
  if (layout == 3) {
    for (int i = 0; i < split; ++i)
      for (int j = 0; j <= i; ++j)
        for (int k = split; k < nbas; ++k) add_block(i, j, k);
  } else {
    for (int i = split; i < nbas; ++i)
      for (int k = split; k <= i; ++k) add_block(i, -1, k);
  }
  // device tables
  std::vector<ShellXYZ> sh(P.seg.size());
  for (size_t i = 0; i < P.seg.size(); ++i) {
    sh[i].x = P.seg[i].r[0];
    sh[i].y = P.seg[i].r[1];
    sh[i].z = P.seg[i].r[2];
    sh[i].ao = P.seg[i].ao_cart;
    sh[i].l = P.seg[i].l;
    sh[i].mol = 0;
    sh[i].lseg = 0;
  }
  // solid-harmonic (or identity) map of every requested AO: up to six Cartesian terms
  std::vector<int> sidx((size_t)P.n * kDfMaxTerms, 0);
  std::vector<double> scoef((size_t)P.n * kDfMaxTerms, 0.0);
  for (int shl = 0; shl < nbas; ++shl) {
    const int l = P.bas[(size_t)shl * BAS_SLOTS + ANG_OF], nc = P.bas[(size_t)shl * BAS_SLOTS + NCTR_OF];
    std::vector<std::vector<double>> rows;
    if (cart) {
      rows.assign(ncart_h(l), std::vector<double>(ncart_h(l), 0.0));
      for (int k = 0; k < ncart_h(l); ++k) rows[k][k] = 1.0;
    } else {
      c2s_rows(l, rows);
    }
    const int w = (int)rows.size();
    for (int c = 0; c < nc; ++c)
      for (int r = 0; r < w; ++r) {
        const size_t ao = (size_t)P.ao[shl] + (size_t)c * w + r;
        int t = 0;
        for (int k = 0; k < ncart_h(l); ++k) {
          if (rows[r][k] == 0.0) continue;
          if (t == kDfMaxTerms) throw Internal("solid harmonic with more Cartesian terms than the map holds");
          sidx[ao * kDfMaxTerms + t] = P.ao_cart[shl] + c * ncart_h(l) + k;
          scoef[ao * kDfMaxTerms + t] = rows[r][k];
          ++t;
        }
      }
  }
  ShellXYZ* d_shells = dev_upload(sh, s);
  PrimPair* d_prims = dev_upload(P.prims, s);
  Pair* d_pairs = dev_upload(pairs, s);
  int* d_sidx = dev_upload(sidx, s);
  double* d_scoef = dev_upload(scoef, s);
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks((size_t)size)), dim3(256), 0, s, out, (size_t)size, 0.0);
  // staging passes of at most kStage doubles, blocks in class order inside a pass
  constexpr int64_t kStage = int64_t(1) << 24;
  double* stage = nullptr;
  HIP_OK(hipMalloc(&stage, sizeof(double) * kStage));
  std::vector<int64_t> blk_coff(packs.size(), -1);
  size_t ib = 0;  // items are grouped by block, in block order
  while (ib < items.size()) {
    // blocks [b0, b1) of this pass
    const int b0 = items[ib].blk;
    int64_t ctot = 0;
    size_t ie = ib;
    while (ie < items.size()) {
      const DfPack& pk = packs[items[ie].blk];
      const int64_t csz = (int64_t)pk.cd[0] * pk.cd[1] * pk.cd[2];
      if (ctot > 0 && ctot + csz > kStage) break;
      if (csz > kStage) throw Internal("auxiliary integral block larger than the staging buffer");
      const int bk = items[ie].blk;
      blk_coff[bk] = ctot;
      ctot += csz;
      while (ie < items.size() && items[ie].blk == bk) ++ie;
    }
    const int b1 = items[ie - 1].blk + 1;
    std::vector<Item> pass(items.begin() + ib, items.begin() + ie);
    for (Item& it : pass) it.t.base = blk_coff[it.blk];
    std::stable_sort(pass.begin(), pass.end(), [](const Item& x, const Item& y) { return x.cls < y.cls; });
    std::vector<BlockTask> tasks;
    for (const Item& it : pass) tasks.push_back(it.t);
    BlockTask* d_tasks = dev_upload(tasks, s);
    size_t i = 0;
    while (i < pass.size()) {
      size_t j = i;
      while (j < pass.size() && pass[j].cls == pass[i].cls) ++j;
      BlockArgs a;
      a.pairs = d_pairs;
      a.prims = d_prims;
      a.shells = d_shells;
      a.tasks = d_tasks + i;
      a.ntasks = (int64_t)(j - i);
      a.out = stage;
      launch_df_blocks(pass[i].cls[0], pass[i].cls[1], pass[i].cls[2], a, s);
      i = j;
    }
    std::vector<DfPack> pp(packs.begin() + b0, packs.begin() + b1);
    int64_t tot = 0;
    for (int b = b0; b < b1; ++b) {
      pp[b - b0].coff = blk_coff[b];
      pp[b - b0].off = tot;
      tot += (int64_t)pp[b - b0].d[0] * pp[b - b0].d[1] * pp[b - b0].d[2];
    }
    DfPack* d_pp = dev_upload(pp, s);
    hipLaunchKernelGGL(df_pack_kernel, dim3(nblocks((size_t)tot)), dim3(256), 0, s, stage, d_pp, (int64_t)pp.size(),
                       tot, layout, n, m, np, d_sidx, d_scoef, out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipFree(d_tasks));
    HIP_OK(hipFree(d_pp));
    ib = ie;
  }
  HIP_OK(hipFree(stage));
  for (void* p : {(void*)d_shells, (void*)d_prims, (void*)d_pairs, (void*)d_sidx, (void*)d_scoef}) (void)hipFree(p);
}

static void packed_jk(const double* eri, int64_t eri_size, int ndim, const double* density, int nb, int n,
                      double* out, int with_j, int with_k, hipStream_t s) {
  if ((with_j != 0 && with_j != 1) || (with_k != 0 && with_k != 1) || !(with_j || with_k))
    throw Internal("Packed J/K requires at least one selected operator");
  const int64_t np = (int64_t)n * (n + 1) / 2;
  if (n < 1 || nb < 1 || (ndim != 1 && ndim != 2) || eri_size != (ndim == 1 ? np * (np + 1) / 2 : np * np))
    throw Internal("Packed J/K dimensions do not match");
  if (!eri || !density || !out) throw Internal("Packed J/K requires device buffers");
  const size_t total = (size_t)(with_j + with_k) * nb * n * n;
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks(total)), dim3(256), 0, s, out, total, 0.0);
  hipLaunchKernelGGL(packed_jk_kernel, dim3(nblocks((size_t)(np * (np + 1) / 2))), dim3(256), 0, s, eri,
                     ndim == 1 ? 1 : 0, np, density, nb, n, with_j, with_k, out);
  HIP_OK(hipGetLastError());
}

}  // namespace gscf

// ----------------------------------------------------------------------------
// C ABI
// ----------------------------------------------------------------------------
using namespace gscf;

extern "C" {

const char* gscf_last_error(void) { return g_error.c_str(); }

int gscf_plan_create(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                     int32_t nenv, int32_t cart, double cutoff, void* stream, gscf_plan** plan) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("plan output pointer is NULL");
    *plan = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    gscf_plan* p = new gscf_plan();
    try {
      validate_tables(p->d, atm, natm, bas, nbas, env, nenv, cart);
      build_plan(p->d, cutoff, (hipStream_t)stream);
      balance_lanes(p->d, (hipStream_t)stream);
    } catch (...) {
      free_plan(p->d);
      delete p;
      throw;
    }
    p->d.plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *plan = p;
  });
}

int gscf_plan_create_batch(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                           int32_t nenv, int32_t nmol, const int32_t* mol_atm, const int32_t* mol_bas, int32_t cart,
                           double cutoff, void* stream, gscf_plan** plan) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("plan output pointer is NULL");
    *plan = nullptr;
    if (nmol < 1 || !mol_atm || !mol_bas) throw InvalidArgument("A batched plan needs per-molecule atm and bas offsets");
    auto t0 = std::chrono::steady_clock::now();
    gscf_plan* p = new gscf_plan();
    try {
      validate_tables(p->d, atm, natm, bas, nbas, env, nenv, cart, nmol, mol_atm, mol_bas);
      build_plan(p->d, cutoff, (hipStream_t)stream);
      balance_lanes(p->d, (hipStream_t)stream);
    } catch (...) {
      free_plan(p->d);
      delete p;
      throw;
    }
    p->d.plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *plan = p;
  });
}

int gscf_plan_get_molecules(const gscf_plan* plan, int32_t max_mol, int32_t* n, int64_t* offset) {
  return guarded([&] {
    if (!plan || !n || !offset) throw InvalidArgument("NULL plan or output");
    const PlanData& P = plan->d;
    if (max_mol < P.nmol) throw InvalidArgument("molecule arrays too small");
    for (int m = 0; m < P.nmol; ++m) {
      n[m] = P.mols[m].n;
      offset[m] = P.mols[m].sbase;
    }
  });
}

int gscf_workspace_bytes(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                         int32_t nenv, int32_t cart, double cutoff, int32_t nb, int64_t* bytes) {
  return guarded([&] {
    if (!bytes) throw InvalidArgument("NULL output");
    if (nb < 1) throw InvalidArgument("Direct J/K shape mismatch");
    if (!std::isfinite(cutoff) || cutoff < 0) throw InvalidArgument("Invalid Schwarz cutoff");
    PlanData P;
    validate_tables(P, atm, natm, bas, nbas, env, nenv, cart);
    *bytes = forecast_bytes(P, nb);
  });
}

int gscf_tuning_set(int32_t cls, const gscf_class_tuning* t) {
  return guarded([&] {
    if (!t) throw InvalidArgument("NULL tuning");
    if (cls < -1 || cls >= kNumQuartetClasses) throw InvalidArgument("tuning class index out of range");
    if (t->chunk_items < 64 || t->chunk_items > (1 << 20) || t->bras_max < 1 || t->bras_max > 256 ||
        t->lds_pack < 1024 || t->lds_pack > (int)kLdsBudget || (t->rys_lds != 0 && t->rys_lds != 1) ||
        t->target_prim < 1 || t->target_prim > 4096 || (t->threads != 256 && t->threads != 512) ||
        (t->ket_lds != 0 && t->ket_lds != 1))
      throw InvalidArgument("tuning parameter out of range");
    ensure_tuning();
    for (int c = 0; c < kNumQuartetClasses; ++c)
      if (cls < 0 || c == cls) g_tuning[c] = *t;
  });
}

int gscf_options_set(const gscf_options* o) {
  return guarded([&] {
    if (!o) throw InvalidArgument("NULL options");
    if (o->general_contraction != 0 && o->general_contraction != 1)
      throw InvalidArgument("general_contraction must be 0 or 1");
    if (o->streams < 1 || o->streams > kMaxStreams) throw InvalidArgument("streams must be in 1..4");
    if (o->partial_sums < 0 || o->partial_sums > 3) throw InvalidArgument("partial_sums must be in 0..3");
    if (o->lds_limit != 0 && (o->lds_limit < 4096 || o->lds_limit > (int32_t)kLdsDeviceMax))
      throw InvalidArgument("lds_limit must be 0 (the device's limit) or in 4096..163840 bytes");
    if (o->deterministic != 0 && o->deterministic != 1) throw InvalidArgument("deterministic must be 0 or 1");
    if (o->balance != 0 && o->balance != 1) throw InvalidArgument("balance must be 0 or 1");
    g_options = *o;
  });
}

int gscf_options_get(gscf_options* o) {
  return guarded([&] {
    if (!o) throw InvalidArgument("NULL options");
    *o = g_options;
  });
}

int gscf_fma_probe(int32_t blocks, int64_t iters, double* out, void* stream) {
  return guarded([&] {
    if (blocks < 1 || iters < 1 || !out) throw InvalidArgument("fma probe needs blocks >= 1, iters >= 1 and an output");
    hipLaunchKernelGGL(fma_probe_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, iters, out);
    HIP_OK(hipGetLastError());
  });
}

int gscf_tuning_reset(void) {
  return guarded([&] {
    g_tuning_init = false;
    ensure_tuning();
  });
}

int gscf_tuning_get(int32_t cls, gscf_class_tuning* t) {
  return guarded([&] {
    if (!t || cls < 0 || cls >= kNumQuartetClasses) throw InvalidArgument("tuning class index out of range");
    ensure_tuning();
    *t = g_tuning[cls];
  });
}

int gscf_plan_destroy(gscf_plan* plan) {
  return guarded([&] {
    if (!plan) return;
    free_plan(plan->d);
    delete plan;
  });
}

int gscf_plan_get_info(const gscf_plan* plan, gscf_plan_info* info) {
  return guarded([&] {
    if (!plan || !info) throw InvalidArgument("NULL plan or info");
    const PlanData& P = plan->d;
    info->n = P.n;
    info->n_cart = P.n_cart;
    info->nbas = P.nbas;
    info->nseg = (int32_t)P.seg.size();
    info->npairs = (int64_t)P.jk.pairs.size();
    info->nquartets = P.jk.nquartets;
    info->cutoff = P.cutoff;
    info->plan_ms = P.plan_ms;
    info->nmol = P.nmol;
    info->nc_max = P.nc_max;
    info->nn_total = P.nn_tot;
    info->device_bytes = plan_bytes(P, (int64_t)P.ws_nbd) + (int64_t)(sizeof(double) * P.incore_doubles);
    info->lds_max = (int64_t)P.lds_max;
    int fallback = 0;
    for (const QuartetClass& qc : P.jk.classes) fallback += qc.d_chunks ? 0 : 1;
    info->classes_global_atomic = fallback;
  });
}

int gscf_jk_build(gscf_plan* plan, const double* density, int32_t nb, double* out, int32_t with_j, int32_t with_k,
                  uint32_t flags, double mixed_threshold, void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    jk_build(plan->d, density, nb, out, with_j, with_k, flags, mixed_threshold, (hipStream_t)stream);
  });
}

// ----------------------------------------------------------------------------
// Allocation-free builds: the workspace lives in a file-scope device array
// (64 MiB, held by every GPU the library runs on until the process ends).
// Host code addresses it only through the per-device pointer that
// hipGetSymbolAddress returns (the array's name is a host-side stand-in).
// ----------------------------------------------------------------------------
constexpr size_t kScratchDoubles = size_t(8) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;  // per device

static double* scratch_ptr() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Internal("no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

int gscf_jk_build_noalloc(gscf_plan* plan, const double* density, int32_t nb, double* out, int32_t with_j,
                          int32_t with_k, uint32_t flags, double mixed_threshold, void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    PlanData& P = plan->d;
    if (nb < 1) throw InvalidArgument("Direct J/K shape mismatch");
    if (P.prof) throw InvalidArgument("the allocation-free build does not run while the plan is profiling");
    const bool sym = (flags & GSCF_FLAG_SYMMETRIC_DENSITY) != 0;
    const int nbd = (with_k && !sym) ? 2 * nb : nb;
    const size_t need = workspace_doubles(P, nb, nbd);
    if (need > kScratchDoubles)
      throw InvalidArgument("the allocation-free J/K build needs " + std::to_string(sizeof(double) * need) +
                            " bytes of scratch for " + std::to_string(nb) +
                            " densities; its scratch holds " + std::to_string(sizeof(double) * kScratchDoubles));
    double* base = scratch_ptr();
    const size_t c = (size_t)P.ncc_tot * nbd;
    // [dc | jp | kp] over nbd densities, then the padded ket-side J over nb
    const Workspace W = {base, base + c, base + 3 * c, base + 2 * c, nullptr};
    jk_build(P, density, nb, out, with_j, with_k, flags, mixed_threshold, (hipStream_t)stream, &W);
  });
}

int gscf_direct_jk(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                   int32_t nenv, const double* density, int32_t nb, int32_t n, double* out, int32_t cart,
                   double cutoff, int32_t workers, int32_t with_j, int32_t with_k, void* stream) {
  // the reference target reports every failure, argument checks included, as Internal
  // (direct_jk.cc:83,93): collapse the InvalidArgument / Internal split here
  const int rc = guarded([&] {
    if (workers < 1) throw InvalidArgument("Direct J/K workers must be positive");
    if ((with_j != 0 && with_j != 1) || (with_k != 0 && with_k != 1) || !(with_j || with_k))
      throw InvalidArgument("Direct J/K requires at least one selected operator");
    gscf_plan p;
    try {
      validate_tables(p.d, atm, natm, bas, nbas, env, nenv, cart);
      if (n != p.d.n || nb < 1) throw InvalidArgument("Direct J/K shape mismatch");
      build_plan(p.d, cutoff, (hipStream_t)stream);
      jk_build(p.d, density, nb, out, with_j, with_k, 0u, 0.0, (hipStream_t)stream);
      HIP_OK(hipStreamSynchronize((hipStream_t)stream));
    } catch (...) {
      free_plan(p.d);
      throw;
    }
    free_plan(p.d);
  });
  return rc == GSCF_OK ? GSCF_OK : GSCF_INTERNAL;
}

int gscf_plan_get_schwarz(const gscf_plan* plan, double* q) {
  return guarded([&] {
    if (!plan || !q) throw InvalidArgument("NULL plan or output");
    const PlanData& P = plan->d;
    if (P.parent_q.empty()) throw InvalidArgument("Schwarz bounds are computed only for a positive cutoff");
    std::copy(P.parent_q.begin(), P.parent_q.end(), q);
  });
}

int gscf_packed_eri(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                    int32_t nenv, int32_t cart, int32_t layout, double* out, int64_t size, void* stream) {
  // every failure is Internal, as in the reference's compact target (compact.cc:16)
  gscf_plan* plan = nullptr;
  const int rc = gscf_plan_create(atm, natm, bas, nbas, env, nenv, cart, 0.0, stream, &plan);
  if (rc != GSCF_OK) return GSCF_INTERNAL;
  const int rc2 = guarded([&] { packed_eri(plan->d, layout, out, size, (hipStream_t)stream); });
  if (rc2 != GSCF_OK) {
    const std::string msg = g_error;
    (void)gscf_plan_destroy(plan);
    g_error = msg;
    return GSCF_INTERNAL;
  }
  return gscf_plan_destroy(plan);
}

int gscf_density_fitting(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                         int32_t nenv, int32_t cart, int32_t layout, int32_t split, double* out, int64_t size,
                         void* stream) {
  // the compact target reports every failure as Internal (compact.cc:16)
  const int rc = guarded([&] {
    density_fitting(atm, natm, bas, nbas, env, nenv, cart, layout, split, out, size, (hipStream_t)stream);
  });
  return rc == GSCF_OK ? rc : GSCF_INTERNAL;
}

int gscf_compact(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                 int32_t nenv, int32_t layout, int32_t cart, int32_t split, double* out, int64_t size, void* stream) {
  if (layout < 0 || layout > 3) {
    g_error = "Unknown compact integral layout";
    return GSCF_INTERNAL;
  }
  if (layout == 2 || layout == 3)
    return gscf_density_fitting(atm, natm, bas, nbas, env, nenv, cart, layout, split, out, size, stream);
  return gscf_packed_eri(atm, natm, bas, nbas, env, nenv, cart, layout, out, size, stream);
}

int gscf_packed_jk(const double* eri, int64_t eri_size, int32_t ndim, const double* density, int32_t nb, int32_t n,
                   double* out, int32_t with_j, int32_t with_k, void* stream) {
  return guarded([&] { packed_jk(eri, eri_size, ndim, density, nb, n, out, with_j, with_k, (hipStream_t)stream); });
}

int gscf_plan_incore_bytes(const gscf_plan* plan, int64_t* bytes) {
  return guarded([&] {
    if (!plan || !bytes) throw InvalidArgument("NULL plan or output");
    size_t total = 0;
    for (const QuartetClass& qc : plan->d.jk.classes)
      if (!qc.pb && !qc.pk && qc.d_chunks && qc.nbra_max >= 1)
        total += (size_t)qc.nq * ncart_h(qc.la) * ncart_h(qc.lb) * ncart_h(qc.lc) * ncart_h(qc.ld);
    *bytes = (int64_t)(sizeof(double) * total);
  });
}

int gscf_plan_incore(gscf_plan* plan, int32_t enable, void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    incore_set(plan->d, enable, (hipStream_t)stream);
  });
}

int gscf_plan_profile(gscf_plan* plan, int32_t enable) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    PlanData& P = plan->d;
    P.prof = enable != 0;
    P.prof_recs.clear();
    P.prof_builds = 0;
    P.ev_used = 0;
  });
}

int gscf_plan_class_times(gscf_plan* plan, int32_t max_classes, int32_t* classes, double* ms,
                          int64_t* launches, int32_t* count) {
  return guarded([&] {
    if (!plan || !count) throw InvalidArgument("NULL plan or output");
    PlanData& P = plan->d;
    // one entry per angular-momentum class (its column-pattern launches summed);
    // launches = builds in which the class ran
    std::vector<int> order;
    std::map<int, int> slot;
    for (const QuartetClass& qc : P.jk.classes)
      if (!slot.count(qc.tuning_index)) {
        slot[qc.tuning_index] = (int)order.size();
        order.push_back(qc.tuning_index);
      }
    const int nc = (int)order.size();
    *count = nc;
    if (max_classes < nc || !classes || !ms || !launches) throw InvalidArgument("class arrays too small");
    for (const QuartetClass& qc : P.jk.classes) {
      const int i = slot[qc.tuning_index];
      classes[4 * i + 0] = qc.la;
      classes[4 * i + 1] = qc.lb;
      classes[4 * i + 2] = qc.lc;
      classes[4 * i + 3] = qc.ld;
    }
    std::vector<int64_t> last(nc, -1);
    for (int i = 0; i < nc; ++i) {
      ms[i] = 0.0;
      launches[i] = 0;
    }
    for (const auto& r : P.prof_recs) {
      HIP_OK(hipEventSynchronize(P.ev_pool[r.ev + 1]));
      float t = 0.f;
      HIP_OK(hipEventElapsedTime(&t, P.ev_pool[r.ev], P.ev_pool[r.ev + 1]));
      const int i = slot[r.cls];
      ms[i] += t;
      if (last[i] != r.build) {
        launches[i] += 1;
        last[i] = r.build;
      }
    }
  });
}

int gscf_eri_blocks(gscf_plan* plan, const int32_t* shls, int64_t nq, const int64_t* offsets, double* out,
                    void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    eri_blocks(plan->d, shls, nq, offsets, out, (hipStream_t)stream);
  });
}

int gscf_int1e(gscf_plan* plan, double* s, double* t, double* v, void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    int1e(plan->d, s, t, v, (hipStream_t)stream);
  });
}

int gscf_int1e_gradient(gscf_plan* plan, const double* density, const double* wdensity, double* grad,
                        void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    int1e_gradient(plan->d, density, wdensity, grad, (hipStream_t)stream);
  });
}

int gscf_jk_param_vjp(gscf_plan* plan, const double* density, const double* cot_j, const double* cot_k,
                      double* env_grad, void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    jk_param_vjp(plan->d, density, cot_j, cot_k, env_grad, (hipStream_t)stream);
  });
}

int gscf_int1e_param_vjp(gscf_plan* plan, const double* density, const double* wdensity, double* env_grad,
                         void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    int1e_param_vjp(plan->d, density, wdensity, env_grad, (hipStream_t)stream);
  });
}

int gscf_jk_geometry_vjp(gscf_plan* plan, const double* density, const double* cot_j, const double* cot_k,
                         double* grad, void* stream) {
  return guarded([&] {
    if (!plan) throw InvalidArgument("NULL plan");
    geometry_vjp(plan->d, density, cot_j, cot_k, grad, (hipStream_t)stream);
  });
}

}  // extern "C"
