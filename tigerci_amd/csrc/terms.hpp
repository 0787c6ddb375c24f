// Elementwise / gather kernels of the sigma build (see terms.hip).
#pragma once

#include "common.hpp"

namespace tci {

// Device views of the per-walk tables.
struct PairTab {
  int32_t n_pairs;
  int32_t n_int;
  int32_t n_v;
  const int32_t* pairs;    // [P][2] = (i, j), i >= j
  const int64_t* dom_off;  // [P+1]
  const int32_t* dom_idx;
  const int64_t* cd_off;   // dense block offsets (d_P^2 each)
  const int64_t* aa_off;   // CSF offsets
  const int64_t* ab_off;
  const double* eps;       // [n_int + n_v]
};

// Device slot holding the CSF vectors of a sigma build.  A sigma replayed from
// a HIP graph reads x and y from here, so one captured graph serves any pair
// of vectors; the eager path passes io = nullptr and the pointers directly.
struct IoSlot {
  const double* x;
  double* y;
};

// x (CSF vector) -> dense C, Ct = 2C - C^T blocks, singles s = x_s / sqrt2
void launch_unpack(const PairTab& T, int dmax, const double* x, double* C, double* Ct, double* s,
                   hipStream_t st, const IoSlot* io = nullptr);

// R += c0 K^{ij} - sum_k [(ki|bj) s_k^a + (kj|ai) s_k^b]   (c0 read from x[0])
void launch_rterms(const PairTab& T, int dmax, const double* x, const double* s, const double* Kov,
                   const double* O3, double* R, hipStream_t st, const IoSlot* io = nullptr);

// R_P += sum_Q [(ki|lj) C_Q(pos) + (li|kj) C_Q(pos)^T] over domain intersections
struct HoleTab {
  const int64_t* isect_off;   // [P*n_pairs + Q] -> start in isect (n_pairs^2 + 1)
  const int32_t* isect;       // pairs (pos in P, pos in Q)
};
void launch_hole(const PairTab& T, const HoleTab& H, int dmax, const double* C, const double* O4,
                 double* R, hipStream_t st);

// per-pair <K^{ij}, Ct^{ij}> (x2 for i>j), summed into *sig0; pair_part holds
// ref_dot_slices() partials per pair
int ref_dot_slices();
void launch_ref_dot(const PairTab& T, const double* Kov, const double* Ct, double* pair_part,
                    double* sig0, hipStream_t st);

// out[c][r] = in[r][c] for a small rows x cols matrix
void launch_transpose_small(const double* in, int rows, int cols, double* out, hipStream_t st);

// singles <- doubles through (ji|kb): gpart[(2Q+o)][i][pos] = -sum_c Ct^{jk}[pos][c] (ji|kb)
void launch_s3int_singles(const PairTab& T, int dmax, const int64_t* gpart_off, const double* Ct,
                          const double* O3, double* gpart, hipStream_t st);

// g[i][a] = gss[i][a] + sum_s gW[s][i][a] + sum_{(Q,o,pos): D_Q[pos]=a} gpart[(2Q+o)][i][pos]
struct InvTab {
  const int64_t* inv_off;  // [n_v+1]
  const int32_t* inv_q;    // pair Q
  const int32_t* inv_pos;  // position of a in D_Q
};
void launch_singles_reduce(const PairTab& T, const InvTab& I, const int64_t* gpart_off,
                           const double* gpart, const double* gss, const double* gW, int n_split,
                           double* g, hipStream_t st);

// sigma = S^T R (with CSF factors) + Fock diagonal + optional E_ref; flags as tci_sigma_apply
void launch_pack(const PairTab& T, int dmax, const double* x, const double* R, const double* g,
                 const double* sig0, double e_ref, int flags, double* sigma, hipStream_t st,
                 const IoSlot* io = nullptr);

// io->x = x, io->y = y, in stream order (before a graph replay)
void launch_set_io(IoSlot* io, const double* x, double* y, hipStream_t st);

// exact diagonal H_kk - E_ref
void launch_diag(const PairTab& T, int dmax, const double* Kov, const double* Jvv, const double* O4,
                 const double* vv_diag, const double* vv_exch, double* diag, hipStream_t st);

// plan-time layout kernels
void launch_gather_Lvv(const double* Lpm, int64_t n_chol, int n_int, int n_v, void* Lvv, bool f32,
                       hipStream_t st);
void launch_gather_Lov(const double* Lpm, int64_t n_chol, int n_int, int n_v, double* Lov,
                       hipStream_t st);
void launch_vv_exch(const void* Lvv, bool f32, int64_t n_chol, int n_v, double* out, hipStream_t st);
void launch_build_mss(const double* Kov, const double* Jvv, int n_int, int n_v, double* Mss,
                      hipStream_t st);
void launch_fill(double* p, int64_t n, double v, hipStream_t st);
// explicit one-electron integrals: F = h + G (closed-shell Fock matrix, n_orb x
// n_orb) from the packed h and the assembled integral blocks
void launch_fock(const double* h_packed, const double* O4, const double* O3, const double* Kov, const double* Jvv,
                 int n_int, int n_v, double* F, hipStream_t st);
// O3[k][k][q][b] -= f_qb and Mss += delta_ij f'_ab - delta_ab f'_ij (f' = F off its diagonal)
void launch_fold_fock(const double* F, int n_int, int n_v, double* O3, double* Mss, hipStream_t st);
// g += c0 f_ov, *sig0 += 2 <f_ov, s>  (c0 read from x[0])
void launch_fov_terms(const double* x, const double* s, const double* Fov, int64_t niv, double* g, double* sig0,
                      hipStream_t st, const IoSlot* io = nullptr);
// y = A x, A n x n row-major
void launch_gemv(const double* A, int64_t n, const double* x, double* y, hipStream_t st);
// Cw[k][P d + c] = C_P[k][c] for n_pairs blocks of size d x d stored back to back
void launch_concat_cols(const double* C, int n_pairs, int d, double* Cw, hipStream_t st);
// Lp[P][jj][r][c] = L^{J0+jj}[D_P r][D_P c] for the local-domain pairs of a chunk
struct PackTab {
  const int32_t* pairs_local;  // local-domain pair indices
  const int64_t* lp_off;       // [n_local+1] element offsets into Lp (per chunk length)
  int32_t n_local;
};
void launch_pack_L(const PairTab& T, const PackTab& K, const void* Lvv, bool f32, int64_t J0, int64_t len,
                   void* Lp, hipStream_t st);

}  // namespace tci
