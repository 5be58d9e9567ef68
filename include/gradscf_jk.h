/* gradscf_jk.h -- C ABI of the MI355X (gfx950) direct-SCF J/K engine.
 *
 * Drop-in boundary for the reference's native direct-J/K FFI target
 *   "gradscf_direct_jk_cpu_v1" -> GradSCFDirectJK
 *   (src/gradscf/integrals/_native/csrc/direct_jk.cc:7-102, binding :97-102:
 *    args atm S32[natm,6], bas S32[nbas,8], env F64[nenv], density F64[nb,n,n];
 *    result F64[nops,nb,n,n] (J block first, then K); attrs cart, cutoff,
 *    workers, with_j, with_k)
 * and for its geometry VJP target
 *   "gradscf_direct_jk_geometry_vjp_cpu_v2" -> GradSCFDirectJKGeometryVJPParallel
 *   (src/gradscf/integrals/_native/csrc/direct_jk_geometry.cc:122-201).
 *
 * Tables (atm/bas/env) are host arrays in libcint layout.  Densities, results
 * and gradients are DEVICE pointers (hipMalloc'd, row-major float64).  Streams
 * are hipStream_t passed as void* (NULL = default stream).  Every call
 * validates its arguments on the host before any launch and returns
 *   GSCF_OK, GSCF_INVALID_ARGUMENT (-> ffi::Error::InvalidArgument) or
 *   GSCF_INTERNAL (-> ffi::Error::Internal), with the message in
 *   gscf_last_error() (thread-local).  The drop-ins follow their reference
 * targets' error classes: gscf_direct_jk, gscf_packed_eri and gscf_packed_jk
 * report every failure as GSCF_INTERNAL (direct_jk.cc:78-93, compact.cc:16,
 * packed_jk.cc); gscf_jk_geometry_vjp keeps the split of the geometry target
 * (std::invalid_argument -> InvalidArgument, else Internal,
 * direct_jk_geometry.cc:155-179); the plan-based calls, which have no
 * reference counterpart, use the same split.
 */
#ifndef GRADSCF_JK_H
#define GRADSCF_JK_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSCF_OK 0
#define GSCF_INVALID_ARGUMENT 1
#define GSCF_INTERNAL 2

/* jk flags */
#define GSCF_FLAG_SYMMETRIC_DENSITY 1u  /* caller promises D == D^T (halves K digestion) */
#define GSCF_FLAG_MIXED_PRECISION 2u    /* FP32 quartets where Q_ij Q_kl max|D| < mixed_threshold */
#define GSCF_FLAG_NO_VALIDATE 4u        /* skip the device finite-density check (no host sync) */

typedef struct gscf_plan gscf_plan;

/* Schedule parameters of one canonical quartet class (no counterpart in the
 * reference; set by the autotuner).  Class index c = cb*(cb+1)/2 + ck over
 * pair classes ss, ps, pp, ds, dp, dd = 0..5 with ck <= cb (0 = ssss ...
 * 20 = dddd).  Plans copy the process-wide table when they are created. */
typedef struct {
  int32_t chunk_items;  /* lane-items per workgroup before a bra chunk is closed (64..2^20) */
  int32_t bras_max;     /* bra pairs (LDS exchange slots) per chunk (1..256) */
  int32_t lds_pack;     /* LDS bytes the slots of one workgroup may use (1024..65536) */
  int32_t rys_lds;      /* 1: stage the class's Rys table slice in LDS; 0: global table
                           (one or two roots: Boys-moment quadrature instead) */
  int32_t target_prim;  /* primitive quartets per work item (primitive split), 1..4096 */
  int32_t threads;      /* workgroup size, 256 or 512 */
  int32_t ket_lds;      /* 1: accumulate the ket-side J of each chunk's ket window in LDS */
} gscf_class_tuning;

typedef struct {
  int32_t n;            /* AO count (spherical or Cartesian, as requested) */
  int32_t n_cart;       /* Cartesian AO count used internally */
  int32_t nbas;         /* shells in the tables */
  int32_t nseg;         /* segmented (single-contraction) shells */
  int64_t npairs;       /* retained segmented shell pairs */
  int64_t nquartets;    /* retained canonical segmented shell quartets */
  double cutoff;        /* Schwarz cutoff */
  double plan_ms;       /* wall time of gscf_plan_create */
  int32_t nmol;         /* molecules (1 unless built by gscf_plan_create_batch) */
  int32_t nc_max;       /* Cartesian AO count of the widest molecule */
  int64_t nn_total;     /* elements of one packed density: sum over molecules of n_m^2 */
  int64_t device_bytes; /* device memory held by the plan, its current workspace and
                           in-core integrals */
  int64_t lds_max;      /* LDS per workgroup a one-density launch may use */
  int32_t classes_global_atomic; /* quartet classes whose exchange rows exceed lds_max even for
                                    one density (run by the global-atomic kernels) */
} gscf_plan_info;

/* Device-memory forecast before any GPU work (the reference's
 * native_jk_workspace_bytes, backends/native/jk.py:14-68, which the SCF input
 * assembly uses to reject a direct SCF up front, scf/inputs/assembly.py:145-152):
 * validates the tables like gscf_plan_create and returns in *bytes an upper
 * bound of the device memory of a plan plus one gscf_jk_build of nb densities
 * (every canonical pair quartet counted as retained; cutoff is only
 * validated).  Host-only: callable without a GPU. */
int gscf_workspace_bytes(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas,
                         const double* env, int32_t nenv, int32_t cart, double cutoff, int32_t nb,
                         int64_t* bytes);

/* Process-wide tuning table: set one class (cls >= 0) or all (cls = -1),
 * read one class, or restore the defaults.  Not thread-safe against
 * concurrent plan creation. */
int gscf_tuning_set(int32_t cls, const gscf_class_tuning* t);
int gscf_tuning_get(int32_t cls, gscf_class_tuning* t);
int gscf_tuning_reset(void);

/* Process-wide engine options (no counterpart in the reference), copied by
 * plans when they are created, like the tuning table. */
typedef struct {
  int32_t general_contraction; /* 1: two-column s shells (cc-pVXZ 1s/2s) evaluated once per
                                  primitive quartet for both columns; 0 (default): as two segments */
  int32_t streams;             /* 1..4 HIP streams the class launches of one J/K build are
                                  spread over (fork/join on the caller's stream; default 4) */
  int32_t partial_sums;        /* bit 0: a workgroup hands its exchange rows, bit 1 its ket-side
                                  J window, to a reduction pass as plain stores into per-build
                                  slabs instead of device-scope atomics (0..3; default 1) */
  int32_t lds_limit;           /* bytes of LDS a workgroup of the bra-major J/K kernels may use:
                                  0 (default) = the device's limit per workgroup (160 KB on
                                  gfx950); lower values (4096..) force the per-density and
                                  global-atomic routes that wide molecules take */
  int32_t deterministic;       /* 1: J/K builds are reproducible bit for bit -- every digestion
                                  add is a 64-bit integer add of the value times a per-molecule
                                  power-of-two scale (from Cauchy-Schwarz bounds of the density
                                  and the Schwarz table, so every result stays below 2^60), which
                                  integer atomics sum independently of their order (FP64 only;
                                  slower: no partial slabs, Rys table read from global memory);
                                  0 (default): FP64 atomics, equal to rounding */
  int32_t balance;             /* 1 (default): plan creation times every quartet class of one
                                  build alone (two builds of an identity density, HIP events)
                                  and spreads the classes over the streams by those times;
                                  0: by the plan-time work model (no probe builds) */
} gscf_options;
int gscf_options_set(const gscf_options* o);
int gscf_options_get(gscf_options* o);

/* FP64 FMA issue-rate probe (no counterpart in the reference; measures the
 * device's vector FP64 peak for the roofline): launches blocks x 256 lanes,
 * each running 8 independent FMA chains of iters steps (16 * iters FLOP per
 * lane), and writes one double per lane to the device buffer out. */
int gscf_fma_probe(int32_t blocks, int64_t iters, double* out, void* stream);

/* Last error message of the calling thread. */
const char* gscf_last_error(void);

/* Build a device plan: validates tables exactly like IntegralTables
 * (csrc/tables.h:31-76), segments generally contracted shells, computes the
 * Schwarz bounds Q_ij = sqrt(max |(pq|pq)|) on the GPU when cutoff > 0
 * (csrc/coulomb_shells.h:60-77) and builds Schwarz-sorted pair lists per
 * angular-momentum class.  Replaces the per-call table copy, libcint optimizer
 * and bounds of direct_jk.cc:18-33; the plan may be reused for any density.
 * A molecule may have at most 16384 Cartesian AOs (larger ones are rejected
 * with GSCF_INVALID_ARGUMENT: the kernels use 32-bit in-molecule offsets). */
int gscf_plan_create(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas,
                     const double* env, int32_t nenv, int32_t cart, double cutoff,
                     void* stream, gscf_plan** plan);
int gscf_plan_destroy(gscf_plan* plan);
int gscf_plan_get_info(const gscf_plan* plan, gscf_plan_info* info);

/* Batched plan over nmol independent molecules (no counterpart in the
 * reference, whose batch rule maps molecules one call at a time,
 * backends/native/jk.py:253-317).  The tables are the molecules' libcint
 * tables concatenated; molecule m owns atm rows [mol_atm[m], mol_atm[m+1]) and
 * bas rows [mol_bas[m], mol_bas[m+1]) (nmol + 1 offsets each, starting at 0),
 * and its shells must sit on its own atoms.  Pairs, quartets, screening
 * (Schwarz maximum per molecule) and AO blocks never cross molecules, so the
 * result for every molecule equals a single-molecule plan of its tables.
 * Densities, results, S/T/V and cotangents of a batched plan are PACKED:
 * molecule m's n_m x n_m matrix starts at element offset[m] (from
 * gscf_plan_get_molecules) of each density, and one density has nn_total
 * elements.  A single-molecule plan is the case nmol = 1 (offset 0, n x n). */
int gscf_plan_create_batch(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas,
                           const double* env, int32_t nenv, int32_t nmol, const int32_t* mol_atm,
                           const int32_t* mol_bas, int32_t cart, double cutoff, void* stream,
                           gscf_plan** plan);
/* AO count n[m] and packed element offset offset[m] of every molecule. */
int gscf_plan_get_molecules(const gscf_plan* plan, int32_t max_mol, int32_t* n, int64_t* offset);

/* Parent-shell Schwarz bounds q[i*nbas + j] (i >= j) used by the plan
 * (csrc/coulomb_shells.h:60-77); only for plans built with cutoff > 0.  For a
 * batched plan, one nbas_m x nbas_m block per molecule, in molecule order,
 * with local shell indices. */
int gscf_plan_get_schwarz(const gscf_plan* plan, double* q);

/* In-core integrals: the plan-resident form of the reference's packed-ERI
 * route (producer csrc/packed_eri.cc:5-14 + consumer csrc/packed_jk.cc:6-79,
 * selected for small molecules at scf/rks.py:323-350).  gscf_plan_incore(plan,
 * 1, stream) evaluates the contracted Cartesian block of every quartet the
 * plan screens in once and keeps it in device memory (stored in the order the
 * J/K kernels visit quartets); until gscf_plan_incore(plan, 0, stream)
 * releases it, every gscf_jk_build of the plan digests stored blocks instead
 * of evaluating them (FP64 values; GSCF_FLAG_MIXED_PRECISION has no effect on
 * stored classes).  Classes the plan runs with general contraction keep
 * evaluating.  gscf_plan_incore_bytes gives the device bytes enabling would
 * allocate; enabling fails with InvalidArgument when they cannot be
 * allocated.  Enabling synchronises `stream` (not graph-capturable); builds
 * afterwards are capturable as before. */
/* Packed ERIs of every unique AO quartet, unscreened, for one molecule: the
 * reference's compact-integral target GradSCFCompact with layout 0 or 1
 * (csrc/compact.cc:5-21 -> packed_eri, csrc/packed_eri.cc:5-14).  out is a
 * DEVICE buffer of `size` doubles: layout 1 the s8 triangle, np (np + 1) / 2
 * values at pair_index(ab, cd), ab >= cd; layout 0 the symmetric np x np pair
 * matrix; np = n (n + 1) / 2, ab = pair_index(a, b) = a (a + 1) / 2 + b for
 * a >= b.  Cartesian or spherical AOs per `cart`.  Synchronises `stream`.
 * Every failure (tables included) is GSCF_INTERNAL, as in the reference. */
int gscf_packed_eri(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas,
                    const double* env, int32_t nenv, int32_t cart, int32_t layout, double* out,
                    int64_t size, void* stream);

/* J/K from packed ERIs: the reference's GradSCFPackedJK
 * (csrc/packed_jk.cc:6-85): eri F64 s8 triangle (ndim 1) or np x np pair
 * matrix (ndim 2, lower triangle read), density F64[nb,n,n] (any symmetry)
 * -> out F64[nops,nb,n,n], nops = with_j + with_k (J first).  All DEVICE
 * pointers; asynchronous on `stream`.  Errors are GSCF_INTERNAL with the
 * reference's messages. */
int gscf_packed_jk(const double* eri, int64_t eri_size, int32_t ndim, const double* density,
                   int32_t nb, int32_t n, double* out, int32_t with_j, int32_t with_k, void* stream);

int gscf_plan_incore_bytes(const gscf_plan* plan, int64_t* bytes);
int gscf_plan_incore(gscf_plan* plan, int32_t enable, void* stream);

/* Measurement hooks (no counterpart in the reference): while enabled, each
 * gscf_jk_build records a HIP event pair around every class kernel on the
 * launch stream; gscf_plan_class_times returns the summed elapsed ms and
 * launch counts per class (classes[i*4..] = la, lb, lc, ld). */
int gscf_plan_profile(gscf_plan* plan, int32_t enable);
int gscf_plan_class_times(gscf_plan* plan, int32_t max_classes, int32_t* classes, double* ms,
                          int64_t* launches, int32_t* count);

/* J/K for nb densities: density F64[nb,n,n] -> out F64[nops,nb,n,n],
 * nops = with_j + with_k (J block first), on `stream`.  Asynchronous.
 * For a batched plan, n*n reads nn_total (packed molecules).  With
 * GSCF_FLAG_NO_VALIDATE the call makes no host synchronisation and issues
 * kernel launches only (no memset or copy), so it may be captured into a HIP
 * graph once the workspace for nb exists (after one uncaptured call). */
int gscf_jk_build(gscf_plan* plan, const double* density, int32_t nb, double* out,
                  int32_t with_j, int32_t with_k, uint32_t flags, double mixed_threshold,
                  void* stream);

/* Allocation-free variant of gscf_jk_build.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results; the build's workspace (Cartesian densities, partial
 * J/K, padded ket-side J: 8 * n_cc * (3 nbd + 8 nb) bytes, n_cc = sum of the
 * molecules' squared Cartesian AO counts, nbd = 2 nb for a nonsymmetric K,
 * else nb) lives in one 64 MiB device array per GPU, held until the process
 * ends; larger requests return GSCF_INVALID_ARGUMENT.  Exchange rows are
 * flushed by atomics (no partial slabs), so results equal gscf_jk_build's to
 * FP64 rounding.  Calls on one GPU share that array: calls that overlap on one
 * GPU (different streams or threads) race and must be serialised by the
 * caller.  The plan must not be profiling.  Without GSCF_FLAG_NO_VALIDATE the
 * finite-density check synchronises the stream (pinned host flag, no
 * allocation). */
int gscf_jk_build_noalloc(gscf_plan* plan, const double* density, int32_t nb, double* out,
                          int32_t with_j, int32_t with_k, uint32_t flags, double mixed_threshold,
                          void* stream);

/* Density-fitting integrals: the reference's compact target layouts 2 and 3
 * (GradSCFCompact -> density_fitting, csrc/compact.cc:5-21,
 * csrc/density_fitting.cc:5-34).  The tables hold the orbital shells
 * [0, split) (up to d) followed by the auxiliary shells [split, nbas) (up to
 * g); n / m are their AO counts (Cartesian or spherical as cart says).
 *   layout 2: out[m][m] (device), the two-centre metric (P|Q);
 *   layout 3: out[m][n(n+1)/2] (device), the three-centre (ab|P) at
 *             pair index a(a+1)/2 + b for a >= b.
 * size must equal the element count.  Every failure is GSCF_INTERNAL, as in
 * the reference target.  Synchronises `stream`. */
int gscf_density_fitting(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas,
                         const double* env, int32_t nenv, int32_t cart, int32_t layout, int32_t split,
                         double* out, int64_t size, void* stream);

/* The whole compact target with its argument list (csrc/compact.cc:5-21):
 * layouts 0 and 1 go to gscf_packed_eri (split ignored), 2 and 3 to
 * gscf_density_fitting; any other layout is GSCF_INTERNAL ("Unknown compact
 * integral layout"). */
int gscf_compact(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                 int32_t nenv, int32_t layout, int32_t cart, int32_t split, double* out, int64_t size,
                 void* stream);

/* One-shot form with the argument list of GradSCFDirectJK (direct_jk.cc:97-102):
 * builds a plan, runs gscf_jk_build (nonfinite densities rejected, as
 * direct_jk.cc:24), synchronises `stream`, frees the plan.  `workers` is
 * validated (>= 1) and otherwise unused on the GPU.  Every failure --
 * argument checks, table validation, shape mismatch, HIP errors -- returns
 * GSCF_INTERNAL with the reference's message, as direct_jk.cc:83,93 does. */
int gscf_direct_jk(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas,
                   const double* env, int32_t nenv, const double* density, int32_t nb, int32_t n,
                   double* out, int32_t cart, double cutoff, int32_t workers, int32_t with_j,
                   int32_t with_k, void* stream);

/* Contracted ERI blocks for an explicit list of shell quartets (shls int32[nq,4],
 * host), written in libcint layout (first index fastest) at out + offsets[q]
 * (device).  Cartesian components, whatever the plan's basis.  The reference's
 * int2e_cart/int2e_sph shell-block call (csrc/coulomb_shells.h:32-58). */
int gscf_eri_blocks(gscf_plan* plan, const int32_t* shls, int64_t nq, const int64_t* offsets,
                    double* out, void* stream);

/* One-electron S, T, V (each F64[n,n], device) in the plan's AO basis: the
 * reference's int1e_ovlp / int1e_kin / int1e_nuc evaluation used for the SCF
 * core Hamiltonian (csrc/ffi.cc:11-86). */
int gscf_int1e(gscf_plan* plan, double* s, double* t, double* v, void* stream);

/* One-electron part of the SCF nuclear gradient (the reference's
 * int1e_ipovlp / int1e_ipkin / int1e_ipnuc + int1e_iprinv contractions,
 * csrc/geometry.cc via libcint autocode/grad1.c):
 *   grad[natm,3] (device) += sum D d(T+V)/dR - W dS/dR
 * for every atm row (basis-centre rows and nuclear rows, the latter carrying
 * the Hellmann-Feynman term), with D, W F64[n,n] device matrices in the plan's
 * AO basis (packed for a batched plan), symmetrised inside. */
int gscf_int1e_gradient(gscf_plan* plan, const double* density, const double* wdensity, double* grad,
                        void* stream);

/* Basis-parameter VJPs (exponents and contraction coefficients): env_grad
 * F64[nenv] (device) += dE/d env at the exponent and coefficient slots of
 * every shell (PTR_EXP + p, PTR_COEFF + c*nprim + p), coefficients taken as
 * the normalised values stored in env (the caller chains through its
 * normalisation).  gscf_jk_param_vjp: E = sum(cot_j*J(D)) + sum(cot_k*K(D))
 * (plan cutoff must be 0; cot_j / cot_k may be NULL); gscf_int1e_param_vjp:
 * E = sum D (T+V) - W S.  Exponent derivatives use d/d alpha g = -|r-A|^2 g
 * (the reference's exponent route, csrc/exponents.cc:1-3). */
int gscf_jk_param_vjp(gscf_plan* plan, const double* density, const double* cot_j, const double* cot_k,
                      double* env_grad, void* stream);
int gscf_int1e_param_vjp(gscf_plan* plan, const double* density, const double* wdensity, double* env_grad,
                         void* stream);

/* Coordinate VJP of J/K (direct_jk_geometry.cc:122-182, cutoff must be 0 in the
 * reference): grad[natm,3] (device) += d/dR sum(cot_j*J(D)) + sum(cot_k*K(D))
 * for the atm-row coordinates shells sit on.  cot_j / cot_k may be NULL. */
int gscf_jk_geometry_vjp(gscf_plan* plan, const double* density, const double* cot_j,
                         const double* cot_k, double* grad, void* stream);

#ifdef __cplusplus
}
#endif
#endif
