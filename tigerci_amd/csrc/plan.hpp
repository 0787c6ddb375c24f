// Plan object behind the C ABI (see plan.hip).
#pragma once

#include <map>

#include "common.hpp"
#include "ladder.hpp"
#include "ladder_fused.hpp"
#include "terms.hpp"

struct tci_plan {
  // sizes
  int n_int = 0, n_v = 0, n_pairs = 0, dmax = 0;
  int64_t n_chol = 0, n_orb = 0, npair_all = 0, n_csf = 0, sum_d2 = 0;
  double e_ref = 0.0;
  bool lvv_f32 = false;
  // host tables
  std::vector<int32_t> h_pairs, h_dom_idx;
  std::vector<int64_t> h_dom_off, h_cd_off, h_aa_off, h_ab_off;
  std::vector<double> h_eps;
  std::vector<int32_t> pairid;   // [i*n_int+j] -> strong walk index or -1 (both orders)
  // device tables
  tci::DevBuf pairs, dom_off, dom_idx, cd_off, aa_off, ab_off, eps;
  tci::DevBuf pos;               // [Q][n_v] position of virtual in D_Q or -1
  tci::DevBuf xmap;              // composite maps D_P -> positions in D_Q
  tci::DevBuf isect_off, isect, inv_off, inv_q, inv_pos, gpart_off;
  // integrals (device)
  tci::DevBuf Lvv, Lov, Kov, Jvv, Mss, O4, O3, vvd, vvx, diag;
  // explicit one-electron integrals: closed-shell Fock matrix F = h + G
  // (n_orb x n_orb), its virtual block off the diagonal (n_v x n_v) and its
  // occupied-virtual block (n_int x n_v); the occupied off-diagonal elements
  // act through O4, the occupied-virtual ones also through O3
  bool explicit_h = false, has_fvv = false, has_fov = false;
  tci::DevBuf Fm, Fvvp, Fov;
  tci::GemmProgram fockvv;         // R_P += F'[D,D] C_P + C_P F'[D,D]
  // work buffers
  tci::DevBuf C, Ct, R, s, sT, g, gss, gW, Y, W, gpart, part, sig0, lad_part;
  int n_split = 1;
  int lad_chunks = 1;
  // programs
  tci::GemmProgram ring, progY, progW, progGW;
  // pair terms as dense products over embedded pair blocks (pairs_full.hip)
  bool dense_pairs = false;
  tci::DevBuf pairid_d, Jbig, CRb, CtRb, CCb, Zb, Z3b, Cemb, CTemb, Remb, Yt, holeA, LovT;
  tci::GemmProgram ringZ, ringZ3, holeG, progWd, progU;
  // Kov = Lov^T Lov: the C~R Kov part of Z runs as U = Yt + C~R Lov^T, then
  // U Lov (2 N^2 n_chol flops in place of 2 N^3); TCI_Z_LOWRANK=0 keeps C~R Kov
  bool z_lowrank = false;
  tci::LadderLaunch ladder;        // fused single-kernel ladder (TCI_LADDER=fused)
  bool ladder_fused = false;
  // two-phase ladder: per J-chunk of length Lc, T_P = L[D,D] C_P then R_P += sum T L
  struct LadderPhase {
    int64_t Lc = 0;                 // chunk length these tables are built for
    int nsplit = 1;                 // K-splits of phase 2 (partial R buffers)
    tci::GemmProgram p1, p2;        // dense programs (FP64, or split for the mixed ladder)
    std::vector<tci::GemmProgram> p1x, p2x;  // split per-walk programs, one per tile shape
    std::vector<tci::GemmProgram> pp1, pp2;  // one program per walk (TCI_LADDER_PER_PAIR=1)
    tci::DevBuf lp_off;             // packed-L offsets of the local-domain pairs
  };
  tci::DevBuf Lp, pairs_local;      // packed L^J[D,D] of local-domain pairs (one chunk, or all J when lp_all)
  bool lp_all = false;              // Lp holds every J (packed once per plan)
  tci::DevBuf lp_off_all;           // offsets of the local walks in an all-J Lp
  int n_local = 0;
  std::vector<LadderPhase> lphase;  // one per distinct chunk length
  std::vector<std::pair<int64_t, int>> lchunks;  // (J0, index into lphase)
  tci::DevBuf Tbuf, Rpart;
  int rpart_n = 1;
  // every walk has the full domain: both phases run as single products over all
  // walks (C blocks side by side in Cw)
  bool ladder_merged = false;
  tci::DevBuf Cw;
  // split-precision (bf16x3) two-phase ladder, built on first use:
  //   phase 1 T'_P[(jj,a)][c] = sum_k L^J[a][k] C_P[c][k]   (T' = L C^T, split rows)
  //   phase 2 R_P[n][a]     += sum_jj sum_c L^J[n][c] T'_P[jj][a][c]
  // every operand k-contiguous in split-row storage (L^J symmetric)
  struct MixedLadder {
    bool ready = false;
    tci::DevBuf Lx;                 // split rows of L^J[v][v] for all J (full-domain pairs present)
    tci::DevBuf Lpx;                // split rows of L^J[D,D] of local-domain pairs (one chunk)
    tci::DevBuf pairs_local;        // indices of the local-domain pairs
    int n_local = 0;
    bool prepacked = false;         // Lpx holds every J (packed once per plan)
    int prec = -1;                  // split precision the copies are made for
    // scales (fp16x3; all 1 for bf16x3): [0] constants (L, Kov, Jbig, Lov,
    // the 4-internal coefficients),
    // [1] C of the ladder, [2] ladder intermediate, [3] phase-1 output,
    // [4] phase-2 output, [5] pair-product output, [6] C and Yt of the pair
    // products, [8] max of the constants; maxbits: [0] max|C|, [1] max|Yt|,
    // [2] max over the constants (float bits)
    tci::DevBuf scales, maxbits;
    tci::DevBuf lp_off_all;         // offsets of the local pairs in a prepacked Lpx
    tci::DevBuf Cx, cx_off;         // split rows of the C_P blocks (of C_P^T when merged)
    bool merged = false;            // all walks full-domain: single products over all walks
    bool per_pair = false;          // one launch per walk and phase (comparison mode)
    tci::DevBuf Tx;                 // T' of one chunk
    tci::DevBuf Rpart;              // partial R buffers of phase 2 K-splits
    int rpart_n = 1;
    // dense pair terms in split precision (full domains)
    tci::DevBuf Kx, Jx, CRx, CtRx, CCx, Lovx, LovTx, Ytx;
    tci::GemmProgram ringZx, ringZ3x, progWx;
    // bf16x3 with the low-rank Z: U (into Yt) and W as one launch before Z
    bool z_lowrank = false;
    tci::GemmProgram progUx, pairsUWx;
    // 4-internal product in split precision: the coefficient matrix and the
    // embedded [C; C^T] blocks as split rows over k = walk
    tci::DevBuf holeAx, embTx;
    tci::GemmProgram holeGx;
    // Z, Z3, W and the 4-internal product as one task list (128 x 128 tiles)
    tci::GemmProgram pairsx;
    bool pairs_merged = false;
    // tile configurations of the split pair products, tuned on the first split
    // preparation and kept for both split precisions (-1: not yet tuned)
    int tuned_pairs_z = -1, tuned_pairs_uw = -1;
    std::vector<LadderPhase> phase;
    std::vector<std::pair<int64_t, int>> chunks;
    int64_t bytes = 0;
    // fused ladder (walks of 129-192 with prepacked L): task table, on or off
    bool fused = false;
    std::vector<tci::FusedTask> fused_h;
    tci::DevBuf fused_d;
    int fused_remap = 1;            // XCD-contiguous: the panels of one walk and J share L2
    int tuned_fused_order = 1;
    bool tuned_fused = false;
    int64_t tuned_fused_target = 0;
    // fused ladder of whole walks of 65-128 (prepacked L), by descending domain
    bool fused2 = false;
    std::vector<tci::FusedTask> fused2_h;
    tci::DevBuf fused2_d;
    int fused2_remap = 0;
    int fused2_waves = 4;           // workgroup of the 65-128 kernel: 4 or 8 waves
    int tuned_fused2_waves = 4;
    int tuned_fused2_order = 0;
    int64_t tuned_fused2_target = 0;
    // fused ladder of whole walks of 129-192 (prepacked L, C in LDS), by descending domain
    bool fused3 = false;
    std::vector<tci::FusedTask> fused3_h;
    tci::DevBuf fused3_d;
    int fused3_remap = 0;
    int tuned_fused3_order = 0;
    int64_t tuned_fused3_target = 0;
    // batched programs for `nvec` vectors (local domains, two-phase): C_v split
    // rows per walk [Q][v][c][k], T' of one chunk [Q][jj][v][a][c], R_v back to
    // back and their K-split partials
    struct Batch {
      int nvec = 0;
      tci::DevBuf Cxb, Txb, Rb, Rpartb;
      int rpart_n = 1;
      std::vector<LadderPhase> phase;
      std::vector<std::pair<int64_t, int>> chunks;
    } bat;
    int64_t tuned_target = -1;      // autotuned K-split fill target (kept across precision switches)
    bool tuned_big = false;         // autotuned: phase 2 of domains 129-192 on 192 x 192 tiles
    bool tuned_dma = false;         // autotuned: those 192 x 192 tiles fed by direct global-to-LDS loads
    int tuned_dma1 = 0;             // autotuned: stages of direct loads for the 128 x 192 phase-1 tiles (0: off)
    std::vector<int> tuned_orders;  // autotuned workgroup order per program role
  } mx;
  // ladder from the resident virtual block (ab|cd) (ladder_vv.hip): V+- once
  // per plan, every sigma one product S = X V over all walks
  struct VVLadder {
    bool built = false;             // FP64 V+- resident
    bool fp64 = false;              // the FP64 sigma uses it
    bool split = false;             // the split sigma of precision `prec` uses it
    int prec = -1;                  // split copies made for this precision
    int64_t np = 0, nm = 0;         // packed pair counts n_v(n_v+1)/2, n_v(n_v-1)/2
    int ncm = 0;                    // walks with i != j (rows of X-)
    tci::DevBuf Vp, Vm;             // FP64 [np][np], [nm][nm]
    tci::DevBuf Vpx, Vmx;           // split rows of V+-
    tci::DevBuf Xp, Xm, Xpx, Xmx;   // X+- (FP64 rows, split rows)
    tci::DevBuf Sp, Sm;             // S+- partials of the K-splits
    // split product run transposed, S^T = V X^T (pairs x walks, the walks as
    // the tiles' N side), then transposed back into Sp / Sm
    bool xtrans = false;
    tci::DevBuf STp, STm;
    tci::DevBuf colm;               // [n_pairs] row of the walk in X- or -1
    int ks = 1, ksx = 1;            // K-splits of the FP64 / split products
    tci::GemmProgram g64, gx;       // the products (both parts in one launch)
    // split schedule chosen on the first split preparation (kept for both
    // split precisions, so a plan's split sigma does not change): tile,
    // tile_n, dma (wave count for 256 tiles), K-split, order; use: -1 not yet
    // decided, 0 Cholesky-direct, 1 resident
    int split_tile = 0, split_tile_n = 0, split_dma = 0, split_ks = 1, split_order = 0, split_use = -1;
    bool split_trans = false;
  } vv;
  tci::PairTab tab{};
  // accounting
  double flops_ladder = 0.0, flops_sigma = 0.0, bytes_ladder = 0.0;
  int64_t device_bytes = 0;
  std::string tuning;               // autotuning record (tci_plan_tuning)
  // event timing (tci_profile_enable)
  bool profiling = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double prof_ladder_ms[3] = {0, 0, 0}, prof_sigma_ms[3] = {0, 0, 0};
  int64_t prof_ladder_n[3] = {0, 0, 0}, prof_sigma_n[3] = {0, 0, 0};
  // sigma builds captured as HIP graphs (TCI_SIGMA_GRAPH), one executable
  // graph per flag set; the graphs read x and y from the device slot `io`
  tci::DevBuf io;
  hipStream_t gstream = nullptr;    // capture stream (default priority)
  std::map<int, hipGraphExec_t> graphs;
  bool capturing = false;
  // whole Davidson iterations captured as HIP graphs (davidson.hip): one
  // executable per (precision, active roots, new vectors, basis slots), the
  // device block their replays read the iteration's state from and write
  // their results to, its pinned host mirror, and the rate of the device
  // clock that times the captured sigma builds
  std::map<int64_t, hipGraphExec_t> iter_graphs;
  tci::DevBuf iter_buf;
  void* iter_host = nullptr;
  size_t iter_host_bytes = 0;
  int iter_clock_khz = 0;
  bool timing() const { return profiling && !capturing; }
  // per-kernel event timing of profiled eager builds (tci_profile_kernels):
  // kt_begin records a start event (-1 when not timing), kt_end the stop event
  // of one launch; kt_collect reads the pairs once the build has synchronised
  struct KStat {
    int64_t n = 0;
    double ms = 0.0, flops = 0.0;  // launches, total milliseconds, total algorithmic flops
  };
  struct KRec {
    std::string name;
    double flops;
    int e0, e1, prec;
  };
  std::map<std::string, KStat> kstats[3];
  std::vector<hipEvent_t> kev;
  std::vector<KRec> kpend;
  size_t kev_used = 0;
  int kt_begin(hipStream_t st);
  void kt_end(int e0, const std::string& name, double flops, int prec, hipStream_t st);
  void kt_collect();
  // time one program launch under its role name and the program's algorithmic flops
  template <class F>
  void kt_run(const std::string& role, const tci::GemmProgram& g, int prec, hipStream_t st, F&& launch) {
    const int e0 = kt_begin(st);
    launch();
    if (e0 >= 0) kt_end(e0, role + " " + g.kernel_name(), g.alg_flops, prec, st);
  }
  void drop_graphs();
  ~tci_plan();

  void build(const tci_problem& p);
  void sigma(const double* x, double* y, int flags, hipStream_t st);
  void sigma_body(const double* x, double* y, const tci::IoSlot* io, int flags, hipStream_t st);
  void sigma_rest(const double* x, double* y, const tci::IoSlot* io, int flags, hipStream_t st);
  // several vectors at once: the split ladder of local domains runs as one
  // batched pass (L^J read once per batch), every other term per vector
  void sigma_batch(const double* const* x, double* const* y, int nvec, int flags, hipStream_t st);
  void sigma_graph(const double* x, double* y, int flags, hipStream_t st);
  void unpack_and_ladder(const double* x, const tci::IoSlot* io, int precision, hipStream_t st);
  void build_two_phase_ladder(hipStream_t st);
  void two_phase_ladder(hipStream_t st);
  void ladder_chunk(const std::pair<int64_t, int>& ch, hipStream_t st);
};

namespace tci {
// split-precision ladder (ladder_mixed.hip): R (zeroed inside) = sum_J L^J C L^J
// from the FP64 C blocks of the plan; records ev[3] before the final reduction
// when profiling
void mixed_prepare(tci_plan& P, int precision, hipStream_t st);
void mixed_ladder(tci_plan& P, int precision, hipStream_t st);
void mixed_ladder_chunk(tci_plan& P, int precision, const std::pair<int64_t, int>& ch, hipStream_t st);
// the split ladder of one sigma after preparation (C split, chunks, fused walks, reduction)
void mixed_ladder_run(tci_plan& P, int precision, hipStream_t st);
// batched split ladder (ladder_mixed.hip): programs for nvec vectors, the split
// of vector v's C (after unpack) and the ladder pass into bat.Rb
bool mixed_batch_supported(const tci_plan& P);
void mixed_batch_prepare(tci_plan& P, int precision, int nvec, hipStream_t st);
void mixed_batch_split_c(tci_plan& P, int precision, int v, hipStream_t st);
// fp16x3: max |C| over the batch (call after each vector's unpack), then the scales
void mixed_batch_absmax(tci_plan& P, bool first, hipStream_t st);
void mixed_batch_set_scales(tci_plan& P, hipStream_t st);
void mixed_batch_ladder(tci_plan& P, int precision, hipStream_t st);
// fp16x3 scales: mx[slot] = max |x[0..n)|, then set_scales mode (1: ladder, 2: pair products)
void launch_split_scales(int mode, const double* x, int64_t n, unsigned* mx, int slot, float* sc, float dmax,
                         hipStream_t st);
// ladder from the resident (ab|cd) block (ladder_vv.hip)
bool vv_fits(const tci_plan& P, double* bytes_out);
double vv_flops(const tci_plan& P);
void vv_build(tci_plan& P, hipStream_t st);
void vv_set_schedule(tci_plan& P, int precision, int tile, int tile_n, int dma, int ks, hipStream_t st,
                     bool trans = false);
void vv_prepare_split(tci_plan& P, int precision, hipStream_t st);
void vv_ladder(tci_plan& P, int precision, hipStream_t st);
double vv_tune(tci_plan& P, int precision, hipStream_t st, std::string& log);
int64_t vv_device_bytes(const tci_plan& P);
// dense pair terms (pairs_full.hip)
bool dense_pairs_choose(const tci_plan& P, double f_gather);
void dense_pairs_build(tci_plan& P, const std::vector<double>& o4, hipStream_t st);
// cfg_z (Z, Z3, 4-internal) and cfg_uw (U, W): 0 = 128 x 128 tiles, 1 / 2 =
// 256 x 256 with 4 / 8 waves, -1 = the TCI_PAIRS_TILE choice
void dense_pairs_build_split(tci_plan& P, hipStream_t st, int cfg_z = -1, int cfg_uw = -1);
void dense_pairs_terms(tci_plan& P, int precision, hipStream_t st);
}  // namespace tci
