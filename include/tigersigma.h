/*
 * tigersigma -- MI355X (gfx950) sigma-vector build sigma = H.c and the
 * Davidson-Liu solver around it, for local MRSDCI with Cholesky-decomposed
 * two-electron integrals (closed-shell reference, valence + N-1 + N-2 CSF
 * space, local pair domains).
 *
 * C ABI: plain pointers and sizes, int status (0 = ok), text of the last
 * error from tci_last_error().  No exception crosses this boundary.
 * Vectors passed to tci_sigma_apply / tci_diagonal live in device memory;
 * tci_htimesc takes host vectors.
 *
 * CI vector layout (one entry per CSF, 0-based), walks in the reference's
 * lexical DRT order:
 *   [0]                    valence (closed-shell reference)
 *   [1 + (n_int-1-i)*n_v + a]  N-1 walk of hole i, virtual a (walks by
 *                          descending hole, the DRT walk order of a
 *                          closed-shell reference)
 *   per strong pair walk P = (i,j), i >= j, in the order given (the
 *   reference's order is i descending, then j descending):
 *       aa block  d_P entries                       (virtual doubly occupied)
 *       ab block  nsp * d_P (d_P-1)/2 entries       (nsp = 1 if i==j else 2:
 *                 singlet-coupled block, then triplet-coupled block; packed
 *                 pair order (2,1),(3,1),(3,2),... over domain positions)
 * This is the layout documented at graph_mod.f90:1041-1092 with the packing
 * of scepmaker_mod.f90:84-96.
 */
#ifndef TIGERSIGMA_H
#define TIGERSIGMA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum tci_status {
  TCI_OK = 0,
  TCI_ERR_ARG = 1,     /* invalid argument or inconsistent sizes */
  TCI_ERR_HIP = 2,     /* HIP runtime failure */
  TCI_ERR_OOM = 3,     /* device allocation failed */
  TCI_ERR_NUMERIC = 4  /* self-test mismatch / non-finite result */
};

enum tci_dtype { TCI_F64 = 0, TCI_F32 = 1 };

/* arithmetic of the pair contractions (4-external ladder) in a sigma build */
enum tci_precision {
  TCI_SIGMA_FP64 = 0,    /* FP64 MFMA throughout */
  TCI_SIGMA_BF16X3 = 1,  /* ladder as hi*hi + hi*lo + lo*hi BF16 MFMA, FP32 accumulate */
  TCI_SIGMA_FP16X3 = 2   /* the same split in FP16 with power-of-two operand scales, FP32 accumulate */
};

/* flag bits for tci_sigma_apply (low byte = tci_precision) */
#define TCI_ACCUMULATE 0x100 /* sigma += H c (htimesc semantics) instead of sigma = H c */
#define TCI_ADD_EREF   0x200 /* include E_ref on the diagonal (bare H) */
/* Run the build as one HIP graph launch.  The first call with a given
 * precision/ACCUMULATE/ADD_EREF combination captures the kernel sequence
 * on a private stream; later calls replay it on `stream` for
 * any civec/sigmavec pair.  Calls on one plan must be ordered on one stream.
 * With profiling on, only the whole build is timed (no ladder time). */
#define TCI_SIGMA_GRAPH 0x400
/* tci_davidson_params.precision only: run each Davidson iteration's device
 * work (residuals and their norms, the preconditioner, Gram-Schmidt, the
 * sigma builds of the new vectors, the new columns of the subspace matrix)
 * as one HIP graph launch, with one host synchronisation per iteration for
 * the subspace eigenproblem (tci_davidson_run and tci_davidson_gen_run).
 * Implies TCI_SIGMA_GRAPH for the sigma builds outside the iteration (the
 * guess and collapses). */
#define TCI_DAVIDSON_GRAPH 0x800

/* Problem description.  All pointers are host pointers, read during
 * tci_plan_create only.  Orbitals: internal 0..n_int-1, virtual
 * n_int..n_int+n_v-1.  `chol` holds one record of n_chol values per orbital
 * pair p>=q, record index p(p+1)/2+q (the MO Cholesky record layout of
 * ThreeIndexWriter.cpp:18-35 / cholesky_structs.f90:35-39, 0-based).
 * One-electron part: either `h`, the one-electron integrals in the packed
 * lower-triangle order of the reference's pp array (element (p,q), p >= q, at
 * p(p+1)/2+q; getpp, get_integrals_mod.f90:61-82, index2 of utilities_mod.f90),
 * from which the plan builds the closed-shell Fock matrix f = h + G on the
 * device (localized orbitals: off-diagonal f_ij, f_ia and f_ab all enter
 * sigma); or, when h is NULL, `eps`, orbital energies of canonical orbitals
 * (the core Hamiltonian is then the one whose closed-shell Fock matrix is
 * diag(eps)).  Domains are sorted virtual indices (0-based over the virtual
 * space) per strong pair walk (allowed virtuals,
 * allowed_virtuals_mod.f90:163-224); pairs left out carry no doubles. */
typedef struct tci_problem {
  int64_t n_int;
  int64_t n_v;
  int64_t n_chol;
  const double* eps;       /* n_int + n_v orbital energies (ignored when h is given) */
  const void* chol;        /* (n_orb(n_orb+1)/2) x n_chol, dtype chol_dtype */
  int32_t chol_dtype;      /* TCI_F64 or TCI_F32 */
  int32_t ladder_dtype;    /* device storage of L^J_ab for the ladder: TCI_F64 or TCI_F32 */
  int64_t n_pairs;         /* strong pair walks */
  const int32_t* pairs;    /* n_pairs x 2: (i, j), i >= j */
  const int64_t* dom_off;  /* n_pairs + 1 */
  const int32_t* dom_idx;  /* dom_off[n_pairs] entries */
  const double* h;         /* NULL, or n_orb(n_orb+1)/2 packed one-electron integrals */
} tci_problem;

typedef struct tci_plan tci_plan;

typedef struct tci_plan_info {
  int64_t n_csf;           /* length of the CI vector */
  int64_t n_ladder_items;  /* work items of the grouped ladder launch */
  double e_ref;            /* closed-shell reference energy */
  double flops_ladder;     /* algorithmic flops of the ladder per sigma: sum_P 4 d_P^3 n_chol */
  double flops_sigma;      /* algorithmic flops of one FP64 sigma (all contractions) */
  double bytes_ladder;     /* algorithmic HBM bytes of the ladder per sigma */
  int64_t device_bytes;    /* device memory held by the plan */
} tci_plan_info;

/* Build a plan: uploads L^J, assembles (ia|jb), (ij|ab), (ij|kl), (ij|ka)
 * and the exact diagonal on the device, and builds the grouped launch
 * tables.  Replaces the setup that htimesc performs on its first call
 * (htimesc.f90:77-82 first_time; makeIntegrals make_all_integrals.f90:23;
 * record dry runs three_and_four_seg_mod_1.f90:3142-3238,
 * three_and_four_seg_mod_2.f90:1095-1195).  Uses the current HIP device. */
int tci_plan_create(const tci_problem* prob, tci_plan** out);
void tci_plan_destroy(tci_plan* plan);
int tci_plan_get_info(const tci_plan* plan, tci_plan_info* info);

/* sigma = (H - E_ref) c  (or += with TCI_ACCUMULATE; bare H with TCI_ADD_EREF).
 * c and sigma are device arrays of n = n_csf doubles; stream may be NULL.
 * Replaces htimesc(civec, sigmavec, data) (htimesc.f90:46-147). */
int tci_sigma_apply(tci_plan* plan, const double* c, double* sigma, int64_t n,
                    int32_t flags, void* stream);

/* nvec sigma builds at once: c and sigma hold nvec vectors of n doubles back to
 * back (device memory, not overlapping); flags as tci_sigma_apply.  In a split
 * precision with per-walk ladder programs and resident L^J blocks the ladder
 * of all vectors runs as one batched pass that reads every L^J block once per
 * batch; otherwise the vectors are built one after another.  No graph replay.
 * No reference counterpart (the reference builds one vector per call). */
int tci_sigma_apply_batch(tci_plan* plan, const double* c, double* sigma, int64_t n, int32_t nvec, int32_t flags,
                          void* stream);

/* Host-vector drop-in with htimesc semantics: sigmavec += H civec (bare H,
 * FP64).  Copies over PCIe; synchronous. */
int tci_htimesc(tci_plan* plan, const double* civec, double* sigmavec, int64_t n);

/* tci_htimesc without device allocation: the same semantics (sigmavec +=
 * H civec, bare H, FP64, host vectors, synchronous), with civec and sigmavec
 * staged through a fixed 64 MiB device array that the library holds on every
 * GPU this is called on, for the life of the process.  n is limited to
 * 4,194,304 CSFs (TCI_ERR_ARG beyond).  Calls on one GPU share that array and
 * are serialised by the library.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). */
int tci_htimesc_noalloc(tci_plan* plan, const double* civec, double* sigmavec, int64_t n);

/* diag[k] = H_kk - E_ref (device array of n doubles); the values
 * diag_element (diag_element_mod_2.f90:83) writes to iodiagstore. */
int tci_diagonal(tci_plan* plan, double* diag, int64_t n, void* stream);

typedef struct tci_davidson_params {
  int32_t roots;          /* number of lowest roots */
  int32_t maxit;          /* 0 -> 50*roots (davidson.f90:80) */
  double energy_tol;      /* |dE| convergence (ENERGY TOLERANCE, default 1e-6) */
  double norm_tol;        /* residual norm convergence (RESIDUAL TOLERANCE) */
  int32_t precision;      /* tci_precision of the expansion sigmas, | TCI_SIGMA_GRAPH to replay
                             every sigma build from a captured HIP graph, | TCI_DAVIDSON_GRAPH
                             to replay whole iterations */
  double switch_tol;      /* mixed: switch to FP64 sigma once every root has |dE| < switch_tol
                             or a residual below norm_tol (forced at maxit/2) */
  int32_t max_vectors;    /* subspace collapse budget (0 -> maxit*roots + roots, capped at 32 GB of vectors) */
  int32_t verbose;        /* 1: print the reference-style iteration table */
} tci_davidson_params;

typedef struct tci_davidson_result {
  int32_t iterations;
  int32_t sigma_fp64;     /* number of FP64 sigma builds */
  int32_t sigma_mixed;    /* number of mixed-precision sigma builds */
  int32_t converged;      /* all roots met a convergence criterion before maxit */
  double seconds;         /* wall time of the solve */
  double seconds_sigma;   /* of which inside sigma builds (in iteration graphs: by the device clock) */
  int32_t collapses;      /* subspace collapses on the vector budget (restarts) */
} tci_davidson_result;

/* Davidson-Liu for the lowest roots (davidson_liu, davidson.f90:47-239).
 * energies: host array of `roots` total energies (E_ref included).
 * vectors: device array roots x n_csf for the converged vectors, or NULL. */
int tci_davidson_run(tci_plan* plan, const tci_davidson_params* params,
                     double* energies, double* vectors, tci_davidson_result* result);

/* ACPF-type functionals (acpf_flag of the reference input, gtimesc.f90
 * get_gvalues) */
enum tci_functional {
  TCI_ACPF = 1,        /* g3 = g4 = g5 = 2/n */
  TCI_ACPF2 = 2,       /* g3 = g4 = (4/n)(1 - 1/(2(n-1))), g5 = 2/n */
  TCI_MRAQCC = 3,      /* g3 = g4 = g5 = (4/n)(1 - 1/(2(n-1))) */
  TCI_ACPF_CUSTOM = 4, /* g3 = g4 = custom_g, g5 = 2/n */
  TCI_ACPF_TEST = 9    /* all g = 1 (the standard problem) */
};

/* The five g values (g1, g2 = 1 for the reference and other valence CSFs;
 * g3 valence CSFs with inactive holes; g4 N-1 walks; g5 N-2 walks) of a
 * functional for n_corr_elec correlated electrons (get_gvalues,
 * gtimesc.f90).  Host arithmetic only. */
int tci_acpf_gvalues(int32_t functional, int32_t n_corr_elec, double custom_g, double* g_values);

/* Davidson-Liu for the generalized problem (H - E_ref) c = E G c with the
 * diagonal G of the g values per CSF class (davidson_liu_gen_eig,
 * davidson.f90:253-469; gtimesc / build_g_diagonal, gtimesc.f90).  In the
 * closed-shell single-reference space the valence CSF takes g1, the N-1
 * walks g4 and the N-2 walks g5.  Same parameters, precision modes and
 * results as tci_davidson_run; energies are E_ref + E. */
int tci_davidson_gen_run(tci_plan* plan, const tci_davidson_params* params, const double* g_values,
                         double* energies, double* vectors, tci_davidson_result* result);

/* Device memory for callers without a GPU runtime of their own (a Fortran or
 * C Davidson driver).  No reference counterpart. */
int tci_device_alloc(int64_t bytes, void** ptr);
int tci_device_free(void* ptr);
int tci_memcpy_h2d(void* dst, const void* src, int64_t bytes);
int tci_memcpy_d2h(void* dst, const void* src, int64_t bytes);
int tci_memcpy_d2d(void* dst, const void* src, int64_t bytes);
int tci_device_synchronize(void);

/* Kernel timing with HIP events on the stream of each sigma build: when
 * enabled, every sigma build records the duration of its ladder launch and of
 * the whole build, summed per tci_precision; tci_profile_read returns the sums
 * for one precision since enabling.  No reference counterpart (the
 * reference's per-phase wall clocks live in time_var_mod.f90). */
int tci_profile_enable(tci_plan* plan, int32_t on);
int tci_profile_read(const tci_plan* plan, int32_t precision, double* ladder_ms,
                     int64_t* ladder_launches, double* sigma_ms, int64_t* sigma_calls);

/* Per-kernel event timing of the same profiled sigma builds: one text line per
 * kernel role ("name|launches|total_ms|algorithmic_flops_per_launch"), for the
 * builds of one precision since profiling was enabled.  The algorithmic flops
 * of a launch count 2 m n k over the valid (unpadded) extents of its
 * products.  Copies at most len-1 characters plus a terminating 0.  No
 * reference counterpart. */
int tci_profile_kernels(const tci_plan* plan, int32_t precision, char* buf, int64_t len);

/* Launch schedules the plan picked by timing candidates on the device: K-split
 * fill targets and workgroup orders of the ladder programs (FP64 at plan
 * creation, split precisions when first used), one text line per ladder.
 * TCI_AUTOTUNE=0 in the environment keeps fixed defaults instead.  Copies at
 * most len-1 characters plus a terminating 0.  No reference counterpart. */
int tci_plan_tuning(const tci_plan* plan, char* buf, int64_t len);

/* Checks the MFMA operand/result lane layouts the kernels assume against a
 * scalar product on the device; 0 when they agree.  No reference counterpart. */
int tci_selftest_mfma(void);

/* Internal walk table of the CSF space: every walk of the internal DRT that
 * ends with n_elec - n_ext electrons (n_ext = 0, 1, 2: valence, N-1, N-2
 * walks), in the order of the reference's lexical walk index, with the keep
 * flag of its tree search.  Replaces graph_border and drt_maker
 * (F/graph_mod.f90:101-380), the walk enumeration of tree_searchX
 * (F/new_tree_search.f90:45-202) with index = 1 + sum of arc weights
 * (F/locist_mod.f90:394), and the excitation-level and frozen-level tests of
 * tree_search_internal (F/graph_mod.f90:700-745,963-1028).  references:
 * n_ref x n_int occupations (0, 1, 2) of the internal levels; the first
 * n_frozen levels are frozen.  occ (capacity x n_int) and kept (capacity) are
 * host arrays; *n_walks returns the walk count, and capacity 0 only queries
 * it.  The walks are generated on the GPU (synchronous). */
int tci_internal_walks(int32_t n_int, int32_t n_v, int32_t n_elec, int32_t spin_mult, int32_t n_ref,
                       const int8_t* references, int32_t n_frozen, int32_t n_ext, int64_t capacity,
                       int8_t* occ, uint8_t* kept, int64_t* n_walks);

/* Allowed virtuals of internal walks under the weak-pairs and truncation-of-
 * virtuals approximations: find_hole_pairs and find_allowed_virtuals
 * (F/locist_mod.f90:453-607), WP_throw_out (:210-245) and the table of
 * distinct lists of add_new_virtual_list (F/allowed_virtuals_mod.f90:163-219).
 * Inputs (host): n_ref <= 16 references and n_walks walks as internal
 * occupations (n_ref x n_int, n_walks x n_int), wp[i][j] (n_int x n_int) and
 * tov[i][j][a] (n_int x n_int x n_v), nonzero where are_interacting_WP(i, j) /
 * are_interacting_TOV(i, j, a) (F/wp_tov_mod.f90:254-323) hold, read for hole
 * pairs i <= j.  Outputs (host): status[w] = 0 reference, 1 single excitation
 * (every virtual), 2 double excitation, 3 double excitation removed by WP
 * (empty list); list_of_walk[w] = index of its list (-1 for a reference); the
 * distinct lists in order of first appearance as list_off[0..*n_lists] (room
 * for n_walks + 1) into list_virt (sorted 0-based virtuals, list_capacity
 * entries).  A walk that is not a reference, single or double excitation of
 * the references is an error, as in the reference.  Computed on the GPU. */
int tci_allowed_virtuals(int32_t n_int, int32_t n_v, int32_t n_ref, const int8_t* references, int64_t n_walks,
                         const int8_t* walk_occ, const uint8_t* wp, const uint8_t* tov, int32_t* status,
                         int64_t* list_of_walk, int64_t* list_off, int32_t* list_virt, int64_t list_capacity,
                         int64_t* n_lists);

/* AO -> MO transformation of Cholesky vectors: first_half_transform,
 * second_half_transform and write_MO_cho_vec (F/cholesky_transform.f90:305-411).
 * ao_pairs: n_ao_pairs pairs (mu, nu) of 0-based AO indices, each unordered
 * pair at most once; ao_vec: n_chol x n_ao_pairs, vector J's values on those
 * pairs (L^J is symmetric); mo_coef: n_ao x n_mo column-major (C[mu + n_ao p],
 * the reference's molecular_orbitals(mu, p)).  Output: mo_records, the
 * n_mo (n_mo + 1) / 2 records of n_chol values in the layout tci_problem.chol
 * reads (record p(p+1)/2 + q, p >= q, holds sum_{mu nu} C[mu][q] C[nu][p]
 * L^J(mu, nu) over J), and, if norms is not NULL, the 2-norm of every record
 * (cho_norms).  Host arrays; the transform runs on the GPU (synchronous). */
int tci_cholesky_ao_to_mo(int32_t n_ao, int32_t n_mo, int64_t n_chol, int64_t n_ao_pairs, const int32_t* ao_pairs,
                          const double* ao_vec, const double* mo_coef, double* mo_records, double* norms);

/* Spin functions of n_open open shells (2..12) with multiplicity spin_mult,
 * in the reference's order, and the symmetric-group representation matrices
 * built on them: branching_diagram and sft_graph (F/spin_mod.f90:99-521),
 * make_cycles (:523-698) and make_transpositions (:700-860).  With nf =
 * fsn(n_open, spin_mult), host outputs (each may be NULL):
 *   sft[(i-1) nf + (w-1)], i = 1..n_open-1: the reference's sft(i, w), the
 *     signed 1-based index of spin function w under the elementary
 *     transposition (i, i+1) (0: w changes sign only);
 *   spin_matrix[v nf + (w-1)], v = 0..n_open: 2S+1 after the shells above v;
 *   cycles[c nf nf + r nf + s]: matrix of the cycle (l..k), l < k, at label
 *     c = (k-1)(k-2)/2 + l - 1 (row r, column s, 0-based), (-1)^P included;
 *   transpositions: matrix of the transposition (i, j), i > j, at label
 *     (i-1)(i-2)/2 + j - 1, same layout.
 * Computed on the GPU (synchronous). */
int tci_spin_tables(int32_t n_open, int32_t spin_mult, int32_t* sft, int32_t* spin_matrix, double* cycles,
                    double* transpositions);

/* Contracted cycles (l..k)(j..i)|d1,d2 between configurations of `left` and
 * `right` open shells: contracted_cycle_stateless (F/spin_mod.f90:861-1312),
 * the spin part of the loop coupling coefficients Px / Py
 * (F/three_and_four_seg_mod_2.f90:638-997,2267-2383).  records: n_rec x 8
 * int32 (i, j, k, l, d1, d2, left, right) with the reference's 1-based shell
 * indices, left <= n_open (<= 12), right <= left; record r writes its
 * fsn(left) x fsn(right) matrix (row-major) at out + out_off[r], blocks in
 * record order and not overlapping.  Singlet spin functions (spin_mult 1).
 * Host arrays; computed on the GPU (synchronous). */
int tci_contracted_cycles(int32_t n_open, int32_t spin_mult, int64_t n_rec, const int32_t* records,
                          const int64_t* out_off, double* out);

/* Ladder loop records of the reference's first sigma: the iteration-1 dry run
 * of all_external_3_4_loops_cho_2 with build_ext_34seg_dryrun
 * (F/three_and_four_seg_mod_1.f90:318-469,3142-3238) and the record of
 * abcd_seg_writeData (F/abcd_seg_info_replacement.f90:42-62).  For every
 * virtual pair a >= c (a outer, both ascending) that is not ignorable, and
 * every one of the n_walks kept N-2 walks (in the caller's order, the tree
 * search's) whose allowed-virtual list holds both: one record of 7 int32
 * (a_g (a_g - 1) / 2 + c_g with orbital indices a_g = n_int + 1 + a,
 * head_weight, fsn(s), fsn(s - 2) (0 for s < 2), fsn(s + 2), 1-based position
 * of a and of c in the walk's list), s the walk's open shells num_singles[w],
 * fsn of multiplicity spin_mult.  Lists: sorted 0-based virtuals at
 * list_virt[list_off[w] .. list_off[w+1]) (as tci_allowed_virtuals returns
 * them per walk); ignorable: n_v x n_v flags [c][a] of ignorable_pair, or
 * NULL for none.  *n_records returns the count; capacity 0 only queries it.
 * Host arrays; computed on the GPU (synchronous). */
int tci_ladder_records(int32_t n_int, int32_t n_v, int32_t spin_mult, int64_t n_walks, const int64_t* head_weight,
                       const int32_t* num_singles, const int64_t* list_off, const int32_t* list_virt,
                       const uint8_t* ignorable, int64_t capacity, int32_t* records, int64_t* n_records);

const char* tci_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TIGERSIGMA_H */
