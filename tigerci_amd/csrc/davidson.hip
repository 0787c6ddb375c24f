// Davidson-Liu on device (davidson_liu, davidson.f90:47-239).
//
// The expansion vectors c_k and sigma_k = H c_k stay in HBM (the reference
// keeps them in files, davidson_io.f90:26-52).  Per iteration, for the roots
// that still take new vectors together (every pass over the basis serves up
// to four of them, the vector work being HBM-bound):
//   h(k,idx) = <c_k, sigma_idx>                  one multi-dot pass
//   r = sum_k y_k (sigma_k - E c_k)              one multi-axpy pass
//   r <- -r / (H_ii - E), 0 where H_ii - E == 0  (build_next_vector_std_eig)
//   orthonormalise against all c_k              classical Gram-Schmidt, a
//                                                second sweep when the first
//                                                removed over half of a
//                                                vector's squared norm, then
//                                                against the new
//                                                vectors before it, then
//                                                normalise; zero if the norm
//                                                < 1e-14.  The reference's
//                                                modified GS with
//                                                renormalisation after every
//                                                projection spans the same vector.
// The m x m subspace problem (m <= 50 roots) is solved on the host (Jacobi).
// With TCI_DAVIDSON_GRAPH (standard problem) the device work of an iteration,
// from the residuals to the new columns of h, is one HIP graph replay: the
// kernels read the basis size, the Ritz coefficients and the energies from a
// device block the host uploads, the second Gram-Schmidt sweep is switched on
// the device, and a root whose residual norm is below norm_tol has its new
// vector zeroed in the graph and dropped by the host afterwards (its sigma is
// built and discarded).  The host synchronises once per iteration, for the
// subspace problem and the convergence bookkeeping.  The generalized problem
// adds its metric: G in the residuals and the preconditioner, and the new
// columns <c_q, G c_k> beside those of h.
// Mixed precision: expansion sigmas use the split ladder until every root's
// |dE| < switch_tol; then the subspace collapses to the current Ritz vectors,
// their sigmas are rebuilt in FP64 and the solve finishes in FP64.  The
// switch is not a restart: in the iteration after it, dE of a root is the
// difference between its FP64 Ritz value in the collapsed space and its last
// mixed-precision one (the Rayleigh-quotient error of the split sigma), and
// the root converges there when that dE < energy_tol or its FP64 residual
// norm < norm_tol.  A budget collapse is a restart: neither that iteration
// nor the next counts as converged (davidson_restart.f90).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "plan.hpp"

namespace tci {
namespace {

constexpr int MD_GROUP = 8;

// partial[k][b] = sum over block b's elements of X_k . v
__global__ void mdot_kernel(const double* const* __restrict__ X, int m, const double* __restrict__ v, int64_t n,
                            double* __restrict__ partial, int nblk) {
  const int k0 = blockIdx.y * MD_GROUP;
  double acc[MD_GROUP];
#pragma unroll
  for (int q = 0; q < MD_GROUP; ++q) acc[q] = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)nblk * blockDim.x) {
    const double ve = v[e];
#pragma unroll
    for (int q = 0; q < MD_GROUP; ++q)
      if (k0 + q < m) acc[q] += X[k0 + q][e] * ve;
  }
  __shared__ double red[MD_GROUP][256];
#pragma unroll
  for (int q = 0; q < MD_GROUP; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < MD_GROUP; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < MD_GROUP && k0 + (int)threadIdx.x < m) partial[(int64_t)(k0 + threadIdx.x) * nblk + blockIdx.x] = red[threadIdx.x][0];
}

__global__ void mdot_reduce_kernel(const double* __restrict__ partial, int nblk, int m, double* __restrict__ out) {
  const int k = blockIdx.x;
  if (k >= m) return;
  __shared__ double red[256];
  double a = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) a += partial[(int64_t)k * nblk + b];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// Several vectors at once: partial[(k R + j) nblk + b] = block b's share of
// X_k . V_j for R <= 4 vectors V_j, so one pass over the X_k serves all of
// them (each V_j is re-read once per group of MD_GROUP X's).  256 threads.
template <int R>
__global__ __launch_bounds__(256) void mdotm_kernel(const double* const* __restrict__ X, int m,
                                                    const double* const* __restrict__ V, int64_t n,
                                                    double* __restrict__ partial, int nblk) {
  const int k0 = blockIdx.y * MD_GROUP;
  const int g = min(MD_GROUP, m - k0);
  const double* x[MD_GROUP];
#pragma unroll
  for (int q = 0; q < MD_GROUP; ++q) x[q] = X[k0 + min(q, g - 1)];
  const double* v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = V[j];
  double acc[MD_GROUP][R];
#pragma unroll
  for (int q = 0; q < MD_GROUP; ++q)
#pragma unroll
    for (int j = 0; j < R; ++j) acc[q][j] = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)nblk * blockDim.x) {
    double ve[R];
#pragma unroll
    for (int j = 0; j < R; ++j) ve[j] = v[j][e];
#pragma unroll
    for (int q = 0; q < MD_GROUP; ++q)
      if (q < g) {
        const double xe = x[q][e];
#pragma unroll
        for (int j = 0; j < R; ++j) acc[q][j] += xe * ve[j];
      }
  }
  // wave sums by lane exchange, then the four waves through LDS
  __shared__ double red[4][MD_GROUP * R];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < MD_GROUP; ++q)
#pragma unroll
    for (int j = 0; j < R; ++j) {
      double a = acc[q][j];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
      if (lane == 0) red[w][q * R + j] = a;
    }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < MD_GROUP * R && t / R < g)
    partial[((int64_t)(k0 + t / R) * R + t % R) * nblk + blockIdx.x] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
}

// out_j = beta out_j + sum_k coef[j m + k] X_k for R <= 4 outputs: one pass
// over the X_k for all of them
template <int R>
__global__ void lincombm_kernel(double* const* __restrict__ out, double beta, const double* const* __restrict__ X,
                                const double* __restrict__ coef, int m, int64_t n) {
  double* o[R];
#pragma unroll
  for (int j = 0; j < R; ++j) o[j] = out[j];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double a[R];
#pragma unroll
    for (int j = 0; j < R; ++j) a[j] = beta == 0.0 ? 0.0 : beta * o[j][e];
    for (int k = 0; k < m; ++k) {
      const double xe = X[k][e];
#pragma unroll
      for (int j = 0; j < R; ++j) a[j] += coef[j * m + k] * xe;
    }
#pragma unroll
    for (int j = 0; j < R; ++j) o[j][e] = a[j];
  }
}

// out = beta*out + sum_k coef[k] X_k
__global__ void lincomb_kernel(double* __restrict__ out, double beta, const double* const* __restrict__ X,
                               const double* __restrict__ coef, int m, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double a = beta == 0.0 ? 0.0 : beta * out[e];
    for (int k = 0; k < m; ++k) a += coef[k] * X[k][e];
    out[e] = a;
  }
}

__global__ void precond_kernel(double* __restrict__ r, const double* __restrict__ diag, double E, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double p = diag[e] - E;
    r[e] = (p == 0.0) ? 0.0 : -r[e] / p;
  }
}

// generalized problem (H - E_ref) c = E G c, G diagonal
__global__ void precond_gen_kernel(double* __restrict__ r, const double* __restrict__ diag,
                                   const double* __restrict__ g, double E, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double p = diag[e] - E * g[e];
    r[e] = (p == 0.0) ? 0.0 : -r[e] / p;
  }
}

// out = g o x (elementwise)
__global__ void gmul_kernel(double* __restrict__ out, const double* __restrict__ g, const double* __restrict__ x,
                            int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = g[e] * x[e];
}

// r += a (g o x)
__global__ void gaxpy_kernel(double* __restrict__ r, double a, const double* __restrict__ g,
                             const double* __restrict__ x, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    r[e] += a * g[e] * x[e];
}

// diagonal of G by CSF class (build_g_diagonal, gtimesc.f90): the valence
// (reference) CSF g1, the N-1 walks g4, the N-2 walks g5
__global__ void gdiag_kernel(double* __restrict__ g, int64_t n, int64_t n_singles, double g1, double g4, double g5) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    g[e] = e == 0 ? g1 : (e <= n_singles ? g4 : g5);
}

__global__ void scale_kernel(double* __restrict__ v, double a, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    v[e] *= a;
}

// (value, index) order: by value, ties to the lower index
__device__ __host__ inline bool vi_less(double v, int64_t i, double bv, int64_t bi) {
  return bi < 0 || v < bv || (v == bv && i < bi);
}

// Block b's smallest (value, index) pair of d that comes after (pv, pi) in
// that order (index -1: none).  Called once per root with the previous pick,
// it lists the lowest diagonal elements in order.  256 threads.
__global__ __launch_bounds__(256) void argmin_after_kernel(const double* __restrict__ d, int64_t n, double pv,
                                                           int64_t pi, double* __restrict__ bv,
                                                           int64_t* __restrict__ bi) {
  double best = 0.0;
  int64_t bidx = -1;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double v = d[e];
    const bool after = v > pv || (v == pv && e > pi);
    if (after && vi_less(v, e, best, bidx)) {
      best = v;
      bidx = e;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(best, off);
    const long long oi = __shfl_xor((long long)bidx, off);
    if (oi >= 0 && vi_less(ov, oi, best, bidx)) {
      best = ov;
      bidx = oi;
    }
  }
  __shared__ double sv[4];
  __shared__ int64_t si[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sv[w] = best;
    si[w] = bidx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; ++q)
      if (si[q] >= 0 && vi_less(sv[q], si[q], best, bidx)) {
        best = sv[q];
        bidx = si[q];
      }
    bv[blockIdx.x] = best;
    bi[blockIdx.x] = bidx;
  }
}

__global__ void unit_kernel(double* __restrict__ v, int64_t n, int64_t k) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    v[e] = (e == k) ? 1.0 : 0.0;
}

// ---- kernels of the graph-captured iteration ----
// A replay serves every iteration with the same numbers of active roots and
// new vectors, so the kernels read what changes between iterations from a
// device block: the basis size m (vector slots are addressed relative to it),
// the store bases, the residual coefficients and the energies.

struct IterState {
  double* V;  // basis store: slot k at V + k n
  double* S;  // sigma store, same slots
  const double* G;  // diagonal metric of the generalized problem, null for the standard one
  int64_t n;
  int m;      // basis size when the iteration starts
  int pad;
  double norm_tol;
};

// slot mmul*m + off of the basis (sel 0) or sigma (sel 1) store
struct SlotSpec {
  int sel, mmul, off;
};
// a count of mmul*m + off vectors
struct CountSpec {
  int mmul, off;
};

__device__ __forceinline__ double* islot(const IterState& s, SlotSpec q, int k) {
  return (q.sel ? s.S : s.V) + (int64_t)(q.mmul * s.m + q.off + k) * s.n;
}

// mdotm_kernel over slot ranges: partial[(k R + j) nblk + b] = block b's share
// of X_k . W_j for k < count.  A set gate that reads 0 skips the launch's work
// (the second Gram-Schmidt sweep is switched on the device).  256 threads.
// (with wg, W_j weighted by the metric: X_k . G W_j)
template <int R>
__global__ __launch_bounds__(256) void sdots_kernel(const IterState* __restrict__ ps, SlotSpec xs, CountSpec xc,
                                                    SlotSpec ws, double* __restrict__ partial, int nblk,
                                                    const int* __restrict__ gate, int wg) {
  if (gate && *gate == 0) return;
  const IterState s = *ps;
  const int m = xc.mmul * s.m + xc.off;
  const int k0 = blockIdx.y * MD_GROUP;
  if (k0 >= m) return;
  const int g = min(MD_GROUP, m - k0);
  const double* x[MD_GROUP];
#pragma unroll
  for (int q = 0; q < MD_GROUP; ++q) x[q] = islot(s, xs, k0 + min(q, g - 1));
  const double* v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = islot(s, ws, j);
  double acc[MD_GROUP][R];
#pragma unroll
  for (int q = 0; q < MD_GROUP; ++q)
#pragma unroll
    for (int j = 0; j < R; ++j) acc[q][j] = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < s.n; e += (int64_t)nblk * blockDim.x) {
    double ve[R];
    const double ge = wg ? s.G[e] : 1.0;
#pragma unroll
    for (int j = 0; j < R; ++j) ve[j] = wg ? ge * v[j][e] : v[j][e];
#pragma unroll
    for (int q = 0; q < MD_GROUP; ++q)
      if (q < g) {
        const double xe = x[q][e];
#pragma unroll
        for (int j = 0; j < R; ++j) acc[q][j] += xe * ve[j];
      }
  }
  __shared__ double red[4][MD_GROUP * R];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < MD_GROUP; ++q)
#pragma unroll
    for (int j = 0; j < R; ++j) {
      double a = acc[q][j];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
      if (lane == 0) red[w][q * R + j] = a;
    }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < MD_GROUP * R && t / R < g)
    partial[((int64_t)(k0 + t / R) * R + t % R) * nblk + blockIdx.x] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
}

// out[k R + j] = sum over blocks of partial, k < count.  256 threads.
__global__ void sreduce_kernel(const IterState* __restrict__ ps, CountSpec xc, int R, const double* __restrict__ partial,
                               int nblk, double* __restrict__ out, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int k = blockIdx.x;
  if (k >= (xc.mmul * ps->m + xc.off) * R) return;
  __shared__ double red[256];
  double a = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) a += partial[(int64_t)k * nblk + b];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// W_j -= sum_{k < count} d[k R + j] X_k (d as sreduce_kernel writes it)
template <int R>
__global__ void sproject_kernel(const IterState* __restrict__ ps, SlotSpec ws, SlotSpec xs, CountSpec xc,
                                const double* __restrict__ d, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const IterState s = *ps;
  const int m = xc.mmul * s.m + xc.off;
  double* o[R];
#pragma unroll
  for (int j = 0; j < R; ++j) o[j] = islot(s, ws, j);
  const double* x0 = islot(s, xs, 0);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < s.n; e += (int64_t)gridDim.x * blockDim.x) {
    double a[R];
#pragma unroll
    for (int j = 0; j < R; ++j) a[j] = o[j][e];
    for (int k = 0; k < m; ++k) {
      const double xe = x0[(int64_t)k * s.n + e];
#pragma unroll
      for (int j = 0; j < R; ++j) a[j] += (-d[k * R + j]) * xe;
    }
#pragma unroll
    for (int j = 0; j < R; ++j) o[j][e] = a[j];
  }
}

// residuals into basis slots m + j0 + j, from c rows of length ld:
//   standard:     r_j = sum_{q < m} c[j][2q] sigma_q + c[j][2q+1] c_q, c[j][2q+1] = -E_j y_q
//   generalized:  r_j = sum_q c[j][2q] sigma_q - E_j G (sum_q c[j][2q+1] c_q), c[j][2q+1] = y_q
template <int R>
__global__ void sres_kernel(const IterState* __restrict__ ps, int j0, const double* __restrict__ coef, int ld,
                            const double* __restrict__ eres) {
  const IterState s = *ps;
  double* o[R];
#pragma unroll
  for (int j = 0; j < R; ++j) o[j] = s.V + (int64_t)(s.m + j0 + j) * s.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < s.n; e += (int64_t)gridDim.x * blockDim.x) {
    double a[R], b[R];
#pragma unroll
    for (int j = 0; j < R; ++j) a[j] = b[j] = 0.0;
    if (!s.G) {
      for (int q = 0; q < s.m; ++q) {
        const double se = s.S[(int64_t)q * s.n + e];
#pragma unroll
        for (int j = 0; j < R; ++j) a[j] += coef[(j0 + j) * ld + 2 * q] * se;
        const double ce = s.V[(int64_t)q * s.n + e];
#pragma unroll
        for (int j = 0; j < R; ++j) a[j] += coef[(j0 + j) * ld + 2 * q + 1] * ce;
      }
    } else {
      for (int q = 0; q < s.m; ++q) {
        const double se = s.S[(int64_t)q * s.n + e], ce = s.V[(int64_t)q * s.n + e];
#pragma unroll
        for (int j = 0; j < R; ++j) {
          a[j] += coef[(j0 + j) * ld + 2 * q] * se;
          b[j] += coef[(j0 + j) * ld + 2 * q + 1] * ce;
        }
      }
      const double ge = s.G[e];
#pragma unroll
      for (int j = 0; j < R; ++j) a[j] += -eres[j0 + j] * ge * b[j];
    }
#pragma unroll
    for (int j = 0; j < R; ++j) o[j][e] = a[j];
  }
}

// dst[j0 + j] = d[j R + j], the squared norms among a group's dots
__global__ void sdiag_kernel(const double* __restrict__ d, int R, int j0, double* __restrict__ dst) {
  const int j = threadIdx.x;
  if (j < R) dst[j0 + j] = d[j * R + j];
}

// residual norm test of the new vectors: mask[j] = 1 where the root may stop
// on its residual norm (rnok) and |r_j| < norm_tol
__global__ void smask_kernel(const IterState* __restrict__ ps, int ne, const double* __restrict__ r2,
                             const int* __restrict__ rnok, double* __restrict__ mask) {
  const int j = threadIdx.x;
  if (j >= ne) return;
  const double rn = sqrt(fmax(r2[j], 0.0));
  mask[j] = rnok[j] && rn < ps->norm_tol ? 1.0 : 0.0;
}

// r <- -r / (H_ii - E_j g_i) (g = 1 for the standard problem; 0 where that
// is 0) on slot m + j; zero if masked
__global__ void sprecond_kernel(const IterState* __restrict__ ps, int j, const double* __restrict__ diag,
                                const double* __restrict__ eng, const double* __restrict__ mask) {
  const IterState s = *ps;
  double* r = s.V + (int64_t)(s.m + j) * s.n;
  const double E = eng[j];
  const bool off = mask[j] != 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < s.n; e += (int64_t)gridDim.x * blockDim.x) {
    const double p = s.G ? diag[e] - E * s.G[e] : diag[e] - E;
    r[e] = (off || p == 0.0) ? 0.0 : -r[e] / p;
  }
}

// second-sweep test of one group after the first sweep's dots d (k R + j, k <
// m): set *flag when some unmasked W_j lost more than half of its squared
// norm w2 (|W_j'|^2 = w2 - sum_k d^2 for orthonormal X)
__global__ void sagain_kernel(const IterState* __restrict__ ps, const double* __restrict__ d, int R, int j0,
                              const double* __restrict__ w2, const double* __restrict__ mask, int* __restrict__ flag) {
  const int j = threadIdx.x;
  if (j >= R || mask[j0 + j] != 0.0) return;
  const int m = ps->m;
  double proj = 0.0;
  for (int k = 0; k < m; ++k) proj += d[k * R + j] * d[k * R + j];
  const double w = w2[j0 + j];
  if (!(w - proj > 0.5 * w)) flag[0] = 1;
}

__global__ void sfill_int_kernel(int* __restrict__ p, int v) {
  if (threadIdx.x == 0) p[0] = v;
}

// fac[j] = 1/|W_j| from its squared norm d[0], 0 below 1e-14
__global__ void snorm_kernel(const double* __restrict__ d, int j, double* __restrict__ fac) {
  if (threadIdx.x != 0) return;
  const double nr = sqrt(fmax(d[0], 0.0));
  fac[j] = nr < 1e-14 ? 0.0 : 1.0 / nr;
}

__global__ void sscale_kernel(const IterState* __restrict__ ps, int j, const double* __restrict__ fac) {
  const IterState s = *ps;
  double* v = s.V + (int64_t)(s.m + j) * s.n;
  const double a = fac[j];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < s.n; e += (int64_t)gridDim.x * blockDim.x)
    v[e] *= a;
}

// constant-rate device clock, taken once the preceding kernels have finished
// (brackets the sigma builds of a replay)
__global__ void sstamp_kernel(long long* __restrict__ t) {
  if (threadIdx.x == 0) t[0] = wall_clock64();
}

// the sigma build of new vector j reads c and writes sigma through the slot
__global__ void sset_io_kernel(const IterState* __restrict__ ps, int j, IoSlot* __restrict__ io) {
  if (threadIdx.x != 0) return;
  const IterState s = *ps;
  io->x = s.V + (int64_t)(s.m + j) * s.n;
  io->y = s.S + (int64_t)(s.m + j) * s.n;
}

// cyclic Jacobi eigen-solver for a small symmetric matrix (ascending eigenvalues)
void jacobi_eigh(int m, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) V[(size_t)i * m + i] = 1.0;
  auto a = [&](int i, int j) -> double& { return A[(size_t)i * m + j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        tot += a(i, j) * a(i, j);
        if (i != j) off += a(i, j) * a(i, j);
      }
    if (off <= 1e-30 * std::max(tot, 1e-300)) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = a(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          const double vkp = V[(size_t)k * m + p], vkq = V[(size_t)k * m + q];
          V[(size_t)k * m + p] = c * vkp - s * vkq;
          V[(size_t)k * m + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(m);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  w.resize(m);
  std::vector<double> Vs((size_t)m * m);
  for (int c = 0; c < m; ++c) {
    w[c] = a(ord[c], ord[c]);
    for (int k = 0; k < m; ++k) Vs[(size_t)k * m + c] = V[(size_t)k * m + ord[c]];
  }
  V.swap(Vs);
}

// h y = w g y for symmetric h and positive definite g (dsygv_wrapper role,
// math_utils.f90:417-479): g = L L^T, Jacobi on L^-1 h L^-T, y = L^-T z, so
// the columns of y are g-normalised
void gen_eigh(int m, const std::vector<double>& h, const std::vector<double>& g, std::vector<double>& w,
              std::vector<double>& Y) {
  std::vector<double> L((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = g[(size_t)i * m + j];
      for (int k = 0; k < j; ++k) s -= L[(size_t)i * m + k] * L[(size_t)j * m + k];
      if (i == j) {
        TCI_REQUIRE(s > 0.0, "subspace metric G is not positive definite");
        L[(size_t)i * m + i] = std::sqrt(s);
      } else {
        L[(size_t)i * m + j] = s / L[(size_t)j * m + j];
      }
    }
  // B = L^-1 h (forward substitution per column), A = B L^-T
  std::vector<double> B((size_t)m * m), A((size_t)m * m);
  for (int c = 0; c < m; ++c)
    for (int i = 0; i < m; ++i) {
      double s = h[(size_t)i * m + c];
      for (int k = 0; k < i; ++k) s -= L[(size_t)i * m + k] * B[(size_t)k * m + c];
      B[(size_t)i * m + c] = s / L[(size_t)i * m + i];
    }
  for (int r = 0; r < m; ++r)
    for (int i = 0; i < m; ++i) {
      double s = B[(size_t)r * m + i];
      for (int k = 0; k < i; ++k) s -= A[(size_t)r * m + k] * L[(size_t)i * m + k];
      A[(size_t)r * m + i] = s / L[(size_t)i * m + i];
    }
  for (int i = 0; i < m; ++i)
    for (int j = i + 1; j < m; ++j) A[(size_t)i * m + j] = A[(size_t)j * m + i] = 0.5 * (A[(size_t)i * m + j] + A[(size_t)j * m + i]);
  std::vector<double> Z;
  jacobi_eigh(m, A, w, Z);
  // Y = L^-T Z (back substitution per column)
  Y.assign((size_t)m * m, 0.0);
  for (int c = 0; c < m; ++c)
    for (int i = m - 1; i >= 0; --i) {
      double s = Z[(size_t)i * m + c];
      for (int k = i + 1; k < m; ++k) s -= L[(size_t)k * m + i] * Y[(size_t)k * m + c];
      Y[(size_t)i * m + c] = s / L[(size_t)i * m + i];
    }
}

struct Dav {
  tci_plan& P;
  hipStream_t st = nullptr;
  int64_t n;
  int nblk = 1024;
  DevBuf ptrs, coefs, partial, dots;
  std::vector<double*> hptr;
  // pointer table: X vectors from 0, the second operand set from XMAX
  static constexpr int XMAX = 2048, RB = 4;
  explicit Dav(tci_plan& p) : P(p), n(p.n_csf) {
    nblk = (int)std::min<int64_t>(1024, std::max<int64_t>(1, cdiv(n, 256 * 4)));
    ptrs.alloc((XMAX + RB) * sizeof(double*));
    coefs.alloc((size_t)RB * XMAX * sizeof(double));
    partial.alloc((size_t)RB * XMAX * nblk * sizeof(double));
    dots.alloc((size_t)RB * XMAX * sizeof(double));
  }
  // out[k R + j] = X_k . V_j (R = V.size()): one pass over X per group of RB
  // vectors V_j
  void mdotm(const std::vector<double*>& X, const std::vector<double*>& V, std::vector<double>& out) {
    const int m = (int)X.size(), R = (int)V.size();
    out.assign((size_t)m * R, 0.0);
    if (m == 0 || R == 0) return;
    TCI_REQUIRE(m <= XMAX, "too many vectors");
    TCI_HIP(hipMemcpyAsync(ptrs.p, X.data(), m * sizeof(double*), hipMemcpyHostToDevice, st));
    std::vector<double> part;
    for (int j0 = 0; j0 < R; j0 += RB) {
      const int r = std::min(RB, R - j0);
      double** vp = static_cast<double**>(ptrs.p) + XMAX;
      TCI_HIP(hipMemcpyAsync(vp, V.data() + j0, r * sizeof(double*), hipMemcpyHostToDevice, st));
      const dim3 grid(nblk, (unsigned)cdiv(m, MD_GROUP));
      const auto* xp = (const double* const*)ptrs.p;
      const auto* vq = (const double* const*)vp;
      switch (r) {
        case 1: hipLaunchKernelGGL(mdotm_kernel<1>, grid, dim3(256), 0, st, xp, m, vq, n, partial.as<double>(), nblk); break;
        case 2: hipLaunchKernelGGL(mdotm_kernel<2>, grid, dim3(256), 0, st, xp, m, vq, n, partial.as<double>(), nblk); break;
        case 3: hipLaunchKernelGGL(mdotm_kernel<3>, grid, dim3(256), 0, st, xp, m, vq, n, partial.as<double>(), nblk); break;
        default: hipLaunchKernelGGL(mdotm_kernel<4>, grid, dim3(256), 0, st, xp, m, vq, n, partial.as<double>(), nblk); break;
      }
      TCI_HIP(hipGetLastError());
      hipLaunchKernelGGL(mdot_reduce_kernel, dim3(m * r), dim3(256), 0, st, partial.as<double>(), nblk, m * r,
                         dots.as<double>());
      TCI_HIP(hipGetLastError());
      part.resize((size_t)m * r);
      TCI_HIP(hipMemcpyAsync(part.data(), dots.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, st));
      TCI_HIP(hipStreamSynchronize(st));
      for (int k = 0; k < m; ++k)
        for (int j = 0; j < r; ++j) out[(size_t)k * R + j0 + j] = part[(size_t)k * r + j];
    }
  }
  // outs_j = beta outs_j + sum_k c[j m + k] X_k (m = X.size()): one pass over
  // X per group of RB outputs
  void lincombm(const std::vector<double*>& outs, double beta, const std::vector<double*>& X,
                const std::vector<double>& c) {
    const int m = (int)X.size(), R = (int)outs.size();
    if (R == 0) return;
    TCI_REQUIRE(m <= XMAX, "too many vectors");
    TCI_REQUIRE(c.size() == (size_t)m * R, "coefficient count");
    if (m > 0) TCI_HIP(hipMemcpyAsync(ptrs.p, X.data(), m * sizeof(double*), hipMemcpyHostToDevice, st));
    const unsigned grid = (unsigned)std::min<int64_t>(cdiv(n, 256), 8192);
    for (int j0 = 0; j0 < R; j0 += RB) {
      const int r = std::min(RB, R - j0);
      double** op = static_cast<double**>(ptrs.p) + XMAX;
      TCI_HIP(hipMemcpyAsync(op, outs.data() + j0, r * sizeof(double*), hipMemcpyHostToDevice, st));
      if (m > 0)
        TCI_HIP(hipMemcpyAsync(coefs.p, c.data() + (size_t)j0 * m, (size_t)r * m * sizeof(double),
                               hipMemcpyHostToDevice, st));
      const auto* xp = (const double* const*)ptrs.p;
      const double* cp = coefs.as<double>();
      switch (r) {
        case 1: hipLaunchKernelGGL(lincombm_kernel<1>, dim3(grid), dim3(256), 0, st, op, beta, xp, cp, m, n); break;
        case 2: hipLaunchKernelGGL(lincombm_kernel<2>, dim3(grid), dim3(256), 0, st, op, beta, xp, cp, m, n); break;
        case 3: hipLaunchKernelGGL(lincombm_kernel<3>, dim3(grid), dim3(256), 0, st, op, beta, xp, cp, m, n); break;
        default: hipLaunchKernelGGL(lincombm_kernel<4>, dim3(grid), dim3(256), 0, st, op, beta, xp, cp, m, n); break;
      }
      TCI_HIP(hipGetLastError());
      // the pointer/coef tables are reused by the next launch
      TCI_HIP(hipStreamSynchronize(st));
    }
  }
  // out[k] = X_k . v for k < m
  void mdot(const std::vector<double*>& X, const double* v, std::vector<double>& out) {
    const int m = (int)X.size();
    out.assign(m, 0.0);
    if (m == 0) return;
    TCI_REQUIRE(m <= 1024, "too many vectors");
    TCI_HIP(hipMemcpyAsync(ptrs.p, X.data(), m * sizeof(double*), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(mdot_kernel, dim3(nblk, (unsigned)cdiv(m, MD_GROUP)), dim3(256), 0, st,
                       (const double* const*)ptrs.p, m, v, n, partial.as<double>(), nblk);
    TCI_HIP(hipGetLastError());
    hipLaunchKernelGGL(mdot_reduce_kernel, dim3(m), dim3(256), 0, st, partial.as<double>(), nblk, m,
                       dots.as<double>());
    TCI_HIP(hipGetLastError());
    TCI_HIP(hipMemcpyAsync(out.data(), dots.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
    TCI_HIP(hipStreamSynchronize(st));
  }
  void lincomb(double* out, double beta, const std::vector<double*>& X, const std::vector<double>& c) {
    const int m = (int)X.size();
    TCI_REQUIRE(m <= 1024, "too many vectors");
    if (m > 0) {
      TCI_HIP(hipMemcpyAsync(ptrs.p, X.data(), m * sizeof(double*), hipMemcpyHostToDevice, st));
      TCI_HIP(hipMemcpyAsync(coefs.p, c.data(), m * sizeof(double), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(lincomb_kernel, dim3((unsigned)std::min<int64_t>(cdiv(n, 256), 8192)), dim3(256), 0, st, out, beta,
                       (const double* const*)ptrs.p, coefs.as<double>(), m, n);
    TCI_HIP(hipGetLastError());
    // the pointer/coef tables are reused by the next call: keep ordering simple
    TCI_HIP(hipStreamSynchronize(st));
  }
  // (norms and the projections of orthonormalize() go through mdotm, whose
  // summation order the graph-captured iteration reproduces)
  double norm(const double* v) {
    std::vector<double> o;
    mdotm({const_cast<double*>(v)}, {const_cast<double*>(v)}, o);
    return std::sqrt(std::max(o[0], 0.0));
  }
  void scale(double* v, double a) {
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)std::min<int64_t>(cdiv(n, 256), 8192)), dim3(256), 0, st, v, a, n);
    TCI_HIP(hipGetLastError());
  }
  // orthonormalise v against X (CGS2); returns false if v vanished
  bool orthonormalize(double* v, const std::vector<double*>& X) {
    std::vector<double> ov;
    for (int pass = 0; pass < 2 && !X.empty(); ++pass) {
      mdotm(X, {v}, ov);
      for (double& o : ov) o = -o;
      lincomb(v, 1.0, X, ov);
    }
    const double nr = norm(v);
    if (nr < 1e-14) {
      scale(v, 0.0);
      return false;
    }
    scale(v, 1.0 / nr);
    return true;
  }
  // orthonormalise each of W against X and the W before it: CGS2 against X
  // with one pass over X per sweep for all of W, then orthonormalize() against
  // the preceding W.  W_j spans what orthonormalising W_0, W_1, ... one at a
  // time against X and their predecessors spans; a vanished W_j is zeroed.
  // The second sweep runs only when the first removed more than half of some
  // W_j's squared norm (|W_j'|^2 = |W_j|^2 - sum_k <X_k, W_j>^2 for
  // orthonormal X): otherwise one classical sweep is already orthogonal to
  // working precision (the twice-is-enough test), and the pass over X is saved.
  void orthonormalize_block(const std::vector<double*>& W, const std::vector<double*>& X) {
    const int K = (int)W.size(), m = (int)X.size();
    if (K == 0) return;
    std::vector<double> ov, wn, c((size_t)K * m);
    if (m > 0) mdotm(W, W, wn);
    for (int pass = 0; pass < 2 && m > 0; ++pass) {
      mdotm(X, W, ov);
      for (int j = 0; j < K; ++j)
        for (int k = 0; k < m; ++k) c[(size_t)j * m + k] = -ov[(size_t)k * K + j];
      lincombm(W, 1.0, X, c);
      if (pass == 0) {
        bool again = false;
        for (int j = 0; j < K; ++j) {
          double proj = 0.0;
          for (int k = 0; k < m; ++k) proj += ov[(size_t)k * K + j] * ov[(size_t)k * K + j];
          const double w2 = wn[(size_t)j * K + j];
          if (!(w2 - proj > 0.5 * w2)) again = true;
        }
        if (!again) break;
      }
    }
    for (int j = 0; j < K; ++j) orthonormalize(W[j], std::vector<double*>(W.begin(), W.begin() + j));
  }
};

// Byte layout of the plan's iteration block for S basis slots: the part the
// host uploads before a replay (state, energies, norm-test switches, residual
// coefficients), the part copied back after it (squared residual norms,
// vanished-vector flags, device clock stamps, new subspace columns: group g of
// four new vectors at hd + 4 S g doubles, entry q R + j, and for the
// generalized problem the new metric columns at gsd, laid out the same), then
// device scratch.  The host mirror
// (pinned) uses the same offsets.
struct IterLayout {
  static constexpr int RMAX = 64;  // roots
  int S, ld;
  size_t eng, eres, rnok, cres, up_end, rd, mask, ts, hd, gsd, out_end, w2, fac, flag, dots, partial, total;
  IterLayout(int S_, int nblk) : S(S_), ld(2 * S_) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    eng = 256;  // IterState at 0
    eres = al(eng + RMAX * 8);
    rnok = al(eres + RMAX * 8);
    cres = al(rnok + RMAX * 4);
    up_end = al(cres + (size_t)RMAX * ld * 8);
    rd = up_end;
    mask = al(rd + RMAX * 8);
    ts = al(mask + RMAX * 8);  // device clock before and after the sigma builds
    hd = al(ts + 16);
    gsd = al(hd + (size_t)RMAX * S * 8);  // metric columns <c_q, G c_k>, laid out as hd
    out_end = al(gsd + (size_t)RMAX * S * 8);
    w2 = out_end;
    fac = al(w2 + RMAX * 8);
    flag = al(fac + RMAX * 8);
    dots = al(flag + 256);
    partial = al(dots + (size_t)4 * (S + RMAX) * 8);
    total = al(partial + (size_t)4 * (S + RMAX) * nblk * 8);
  }
};

void drop_iter_graphs(tci_plan& P) {
  if (P.iter_graphs.empty()) return;
  TCI_HIP(hipDeviceSynchronize());
  for (auto& kv : P.iter_graphs) (void)hipGraphExecDestroy(kv.second);
  P.iter_graphs.clear();
}

// the block and its host mirror for S slots (graphs that point into smaller
// ones are dropped first)
void iter_ensure(tci_plan& P, const IterLayout& L) {
  if (P.iter_buf.bytes < L.total) {
    drop_iter_graphs(P);
    P.iter_buf.alloc(L.total);
  }
  if (P.iter_host_bytes < L.out_end) {
    drop_iter_graphs(P);
    if (P.iter_host) (void)hipHostFree(P.iter_host);
    P.iter_host = nullptr;
    P.iter_host_bytes = 0;
    TCI_HIP(hipHostMalloc(&P.iter_host, L.out_end, hipHostMallocDefault));
    P.iter_host_bytes = L.out_end;
  }
  if (P.iter_clock_khz <= 0) {
    int dev = 0;
    TCI_HIP(hipGetDevice(&dev));
    TCI_HIP(hipDeviceGetAttribute(&P.iter_clock_khz, hipDeviceAttributeWallClockRate, dev));
  }
  if (!P.io.p) P.io.alloc(sizeof(IoSlot));
  if (!P.gstream) TCI_HIP(hipStreamCreateWithFlags(&P.gstream, hipStreamNonBlocking));
}

// Record one iteration of the standard problem with na active roots, the
// first ne of which take a new vector (residual slots m .. m + na - 1, the
// new vectors first), on stream cs:
//   residuals and their squared norms; the norm test; the preconditioner;
//   one classical Gram-Schmidt sweep against the basis and a second one that
//   runs only when the first removed more than half of some vector's squared
//   norm (decided on the device); CGS2 within the new block and
//   normalisation (zero below 1e-14); the sigma builds of the new vectors
//   (between two reads of the device clock); the new columns <c_q, sigma_k>, q < m + ne;
//   one copy of the results to the host mirror.
void iter_record(tci_plan& P, const IterLayout& L, int nblk, int precision, int na, int ne, bool gen, hipStream_t cs) {
  char* base = static_cast<char*>(P.iter_buf.p);
  const IterState* ps = reinterpret_cast<const IterState*>(base);
  const double* eng = reinterpret_cast<const double*>(base + L.eng);
  const double* eres = reinterpret_cast<const double*>(base + L.eres);
  const int* rnok = reinterpret_cast<const int*>(base + L.rnok);
  const double* cres = reinterpret_cast<const double*>(base + L.cres);
  double* rd = reinterpret_cast<double*>(base + L.rd);
  double* mask = reinterpret_cast<double*>(base + L.mask);
  double* hd = reinterpret_cast<double*>(base + L.hd);
  double* w2 = reinterpret_cast<double*>(base + L.w2);
  double* fac = reinterpret_cast<double*>(base + L.fac);
  int* flag = reinterpret_cast<int*>(base + L.flag);
  double* dots = reinterpret_cast<double*>(base + L.dots);
  double* partial = reinterpret_cast<double*>(base + L.partial);
  const int64_t n = P.n_csf;
  const unsigned gx = (unsigned)std::min<int64_t>(cdiv(n, 256), 8192);
  const double* diag = P.diag.as<double>();
  // dots of X (count xc, at most xmax) with R vectors W, reduced into out
  auto dotsR = [&](SlotSpec xs, CountSpec xc, int xmax, SlotSpec ws, int R, double* out, const int* gate, int wg = 0) {
    const dim3 grid(nblk, (unsigned)cdiv(xmax, MD_GROUP));
    switch (R) {
      case 1: hipLaunchKernelGGL(sdots_kernel<1>, grid, dim3(256), 0, cs, ps, xs, xc, ws, partial, nblk, gate, wg); break;
      case 2: hipLaunchKernelGGL(sdots_kernel<2>, grid, dim3(256), 0, cs, ps, xs, xc, ws, partial, nblk, gate, wg); break;
      case 3: hipLaunchKernelGGL(sdots_kernel<3>, grid, dim3(256), 0, cs, ps, xs, xc, ws, partial, nblk, gate, wg); break;
      default: hipLaunchKernelGGL(sdots_kernel<4>, grid, dim3(256), 0, cs, ps, xs, xc, ws, partial, nblk, gate, wg); break;
    }
    TCI_HIP(hipGetLastError());
    hipLaunchKernelGGL(sreduce_kernel, dim3((unsigned)(xmax * R)), dim3(256), 0, cs, ps, xc, R, partial, nblk, out,
                       gate);
    TCI_HIP(hipGetLastError());
  };
  auto project = [&](SlotSpec ws, SlotSpec xs, CountSpec xc, int R, const int* gate) {
    switch (R) {
      case 1: hipLaunchKernelGGL(sproject_kernel<1>, dim3(gx), dim3(256), 0, cs, ps, ws, xs, xc, dots, gate); break;
      case 2: hipLaunchKernelGGL(sproject_kernel<2>, dim3(gx), dim3(256), 0, cs, ps, ws, xs, xc, dots, gate); break;
      case 3: hipLaunchKernelGGL(sproject_kernel<3>, dim3(gx), dim3(256), 0, cs, ps, ws, xs, xc, dots, gate); break;
      default: hipLaunchKernelGGL(sproject_kernel<4>, dim3(gx), dim3(256), 0, cs, ps, ws, xs, xc, dots, gate); break;
    }
    TCI_HIP(hipGetLastError());
  };
  const SlotSpec basis{0, 0, 0};
  const CountSpec all_m{1, 0};
  hipLaunchKernelGGL(sfill_int_kernel, dim3(1), dim3(64), 0, cs, flag, 0);
  for (int j0 = 0; j0 < na; j0 += 4) {
    const int R = std::min(4, na - j0);
    switch (R) {
      case 1: hipLaunchKernelGGL(sres_kernel<1>, dim3(gx), dim3(256), 0, cs, ps, j0, cres, L.ld, eres); break;
      case 2: hipLaunchKernelGGL(sres_kernel<2>, dim3(gx), dim3(256), 0, cs, ps, j0, cres, L.ld, eres); break;
      case 3: hipLaunchKernelGGL(sres_kernel<3>, dim3(gx), dim3(256), 0, cs, ps, j0, cres, L.ld, eres); break;
      default: hipLaunchKernelGGL(sres_kernel<4>, dim3(gx), dim3(256), 0, cs, ps, j0, cres, L.ld, eres); break;
    }
    TCI_HIP(hipGetLastError());
  }
  for (int j0 = 0; j0 < na; j0 += 4) {
    const int R = std::min(4, na - j0);
    dotsR({0, 1, j0}, {0, R}, R, {0, 1, j0}, R, dots, nullptr);
    hipLaunchKernelGGL(sdiag_kernel, dim3(1), dim3(64), 0, cs, dots, R, j0, rd);
  }
  hipLaunchKernelGGL(smask_kernel, dim3(1), dim3(64), 0, cs, ps, ne, rd, rnok, mask);
  for (int j = 0; j < ne; ++j) hipLaunchKernelGGL(sprecond_kernel, dim3(gx), dim3(256), 0, cs, ps, j, diag, eng, mask);
  TCI_HIP(hipGetLastError());
  for (int j0 = 0; j0 < ne; j0 += 4) {
    const int R = std::min(4, ne - j0);
    dotsR({0, 1, j0}, {0, R}, R, {0, 1, j0}, R, dots, nullptr);
    hipLaunchKernelGGL(sdiag_kernel, dim3(1), dim3(64), 0, cs, dots, R, j0, w2);
  }
  for (int sweep = 0; sweep < 2; ++sweep)
    for (int j0 = 0; j0 < ne; j0 += 4) {
      const int R = std::min(4, ne - j0);
      const int* gate = sweep ? flag : nullptr;
      dotsR(basis, all_m, L.S, {0, 1, j0}, R, dots, gate);
      if (sweep == 0) hipLaunchKernelGGL(sagain_kernel, dim3(1), dim3(64), 0, cs, ps, dots, R, j0, w2, mask, flag);
      project({0, 1, j0}, basis, all_m, R, gate);
    }
  for (int j = 0; j < ne; ++j) {
    for (int pass = 0; pass < 2 && j > 0; ++pass) {
      dotsR({0, 1, 0}, {0, j}, j, {0, 1, j}, 1, dots, nullptr);
      project({0, 1, j}, {0, 1, 0}, {0, j}, 1, nullptr);
    }
    dotsR({0, 1, j}, {0, 1}, 1, {0, 1, j}, 1, dots, nullptr);
    hipLaunchKernelGGL(snorm_kernel, dim3(1), dim3(64), 0, cs, dots, j, fac);
    hipLaunchKernelGGL(sscale_kernel, dim3(gx), dim3(256), 0, cs, ps, j, fac);
  }
  TCI_HIP(hipGetLastError());
  long long* ts = reinterpret_cast<long long*>(base + L.ts);
  hipLaunchKernelGGL(sstamp_kernel, dim3(1), dim3(64), 0, cs, ts);
  IoSlot* io = P.io.as<IoSlot>();
  for (int j = 0; j < ne; ++j) {
    hipLaunchKernelGGL(sset_io_kernel, dim3(1), dim3(64), 0, cs, ps, j, io);
    TCI_HIP(hipGetLastError());
    P.sigma_body(nullptr, nullptr, io, precision, cs);
  }
  hipLaunchKernelGGL(sstamp_kernel, dim3(1), dim3(64), 0, cs, ts + 1);
  TCI_HIP(hipGetLastError());
  for (int j0 = 0, g = 0; j0 < ne; j0 += 4, ++g) {
    const int R = std::min(4, ne - j0);
    dotsR(basis, {1, ne}, L.S, {1, 1, j0}, R, hd + (size_t)4 * L.S * g, nullptr);
    if (gen)  // metric columns <c_q, G c_k>
      dotsR(basis, {1, ne}, L.S, {0, 1, j0}, R, reinterpret_cast<double*>(base + L.gsd) + (size_t)4 * L.S * g, nullptr,
            1);
  }
  const size_t out_bytes = (gen ? L.gsd : L.hd) - L.rd + (size_t)cdiv(ne, 4) * 4 * L.S * 8;
  TCI_HIP(hipMemcpyAsync(static_cast<char*>(P.iter_host) + L.rd, base + L.rd, out_bytes, hipMemcpyDeviceToHost, cs));
}

hipGraphExec_t iter_graph(tci_plan& P, const IterLayout& L, int nblk, int precision, int na, int ne, bool gen) {
  const int64_t key =
      ((int64_t)L.S << 25) | ((int64_t)gen << 24) | ((int64_t)precision << 16) | ((int64_t)na << 8) | ne;
  auto it = P.iter_graphs.find(key);
  if (it != P.iter_graphs.end()) return it->second;
  hipStream_t cs = P.gstream;
  hipGraph_t g = nullptr;
  TCI_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  P.capturing = true;
  try {
    iter_record(P, L, nblk, precision, na, ne, gen, cs);
  } catch (...) {
    P.capturing = false;
    (void)hipStreamEndCapture(cs, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  P.capturing = false;
  TCI_HIP(hipStreamEndCapture(cs, &g));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  TCI_HIP(e);
  P.iter_graphs.emplace(key, ex);
  return ex;
}

}  // namespace
}  // namespace tci

using namespace tci;

namespace {

// gvals: null for the standard problem (H - E_ref) c = E c, else the five g
// values of the generalized problem (H - E_ref) c = E G c
int davidson_impl(tci_plan* plan, const tci_davidson_params* prm, const double* gvals, double* energies,
                  double* vectors, tci_davidson_result* result) {
  try {
    TCI_REQUIRE(plan && prm && energies, "null argument");
    const int roots = prm->roots;
    TCI_REQUIRE(roots >= 1 && roots <= 64 && roots <= plan->n_csf, "roots out of range");
    const int maxit = prm->maxit > 0 ? prm->maxit : 50 * roots;
    // default subspace budget: every iteration's vectors, capped so that the
    // basis and its sigma vectors take at most 32 GB (the reference's disk
    // budget rule, davidson_restart.f90:26-60, with device memory as the store)
    // The basis and sigma stores hold mv + 2*roots vectors: the collapse
    // builds its (at most 2*roots) new vectors in the slots beyond the
    // current basis before moving them to the front, and a collapse can fall
    // due at any m <= mv.
    int mv = prm->max_vectors;
    if (mv <= 0) {
      const int64_t cap = (int64_t)(32.0e9 / (16.0 * (double)plan->n_csf)) - 2 * roots;
      mv = (int)std::max<int64_t>(5 * roots, std::min<int64_t>((int64_t)maxit * roots + roots, cap));
    }
    mv = std::min(mv, 1000);
    TCI_REQUIRE(mv >= 5 * roots, "max_vectors must be at least 5*roots");
    const int mslots = mv + 2 * roots;
    // low byte: tci_precision of the expansion sigmas; TCI_SIGMA_GRAPH replays
    // every sigma build from a captured HIP graph, TCI_DAVIDSON_GRAPH every
    // iteration of the standard problem (and the builds outside it)
    TCI_REQUIRE((prm->precision & ~(0xff | TCI_SIGMA_GRAPH | TCI_DAVIDSON_GRAPH)) == 0, "unknown bits in precision");
    const int graph_flag = prm->precision & (TCI_SIGMA_GRAPH | TCI_DAVIDSON_GRAPH) ? TCI_SIGMA_GRAPH : 0;
    const bool igraph = (prm->precision & TCI_DAVIDSON_GRAPH) != 0;
    const int prec0 = prm->precision & 0xff;
    TCI_REQUIRE(prec0 == TCI_SIGMA_FP64 || prec0 == TCI_SIGMA_BF16X3 || prec0 == TCI_SIGMA_FP16X3, "unknown precision");
    const int64_t n = plan->n_csf;
    Dav D(*plan);
    hipStream_t st = D.st;
    auto t_start = std::chrono::steady_clock::now();
    double t_sigma = 0.0;
    int n64 = 0, nmix = 0, n_collapse = 0;
    DevBuf Vb, Sb, work;
    Vb.alloc((size_t)mslots * n * 8);
    Sb.alloc((size_t)mslots * n * 8);
    work.alloc((size_t)n * 8);
    double* Vp = Vb.as<double>();
    double* Sp = Sb.as<double>();
    auto Vk = [&](int k) { return Vp + (int64_t)k * n; };
    auto Sk = [&](int k) { return Sp + (int64_t)k * n; };
    double* tmp = work.as<double>();  // G c of the generalized residual
    const double* diag = plan->diag.as<double>();
    const unsigned gxn = (unsigned)std::min<int64_t>(cdiv(n, 256), 8192);
    // generalized problem: G on the device, its subspace matrix gs beside h
    const bool gen = gvals != nullptr;
    DevBuf gbuf;
    const double* gd = nullptr;
    if (gen) {
      for (int q = 0; q < 5; ++q) TCI_REQUIRE(std::isfinite(gvals[q]) && gvals[q] > 0.0, "g values must be positive");
      gbuf.alloc((size_t)n * 8);
      hipLaunchKernelGGL(gdiag_kernel, dim3(gxn), dim3(256), 0, st, gbuf.as<double>(), n,
                         (int64_t)plan->n_int * plan->n_v, gvals[0], gvals[3], gvals[4]);
      TCI_HIP(hipGetLastError());
      gd = gbuf.as<double>();
    }
    std::vector<double> gs;
    // column k of the subspace metric: gs(q,k) = <c_q, G c_k> for q < m
    auto g_column = [&](int k, int m) {
      hipLaunchKernelGGL(gmul_kernel, dim3(gxn), dim3(256), 0, st, tmp, gd, Vk(k), n);
      TCI_HIP(hipGetLastError());
      std::vector<double*> X;
      for (int q = 0; q < m; ++q) X.push_back(Vk(q));
      std::vector<double> gv;
      D.mdot(X, tmp, gv);
      return gv;
    };
    int precision = prec0;
    // TCI_DAVIDSON_BATCH=1: the sigma builds of one step go as one batch (one
    // ladder pass for all of them in split precision); off by default, the
    // batched ladder measures slower per vector than single builds
    const char* benv = std::getenv("TCI_DAVIDSON_BATCH");
    const bool batch = benv && benv[0] == '1';
    auto sigma_many = [&](const std::vector<int>& ks) {
      if (ks.empty()) return;
      auto t0 = std::chrono::steady_clock::now();
      if (batch) {
        std::vector<const double*> xs;
        std::vector<double*> ys;
        for (int k : ks) {
          xs.push_back(Vk(k));
          ys.push_back(Sk(k));
        }
        plan->sigma_batch(xs.data(), ys.data(), (int)ks.size(), precision | graph_flag, st);
      } else {
        for (int k : ks) plan->sigma(Vk(k), Sk(k), precision | graph_flag, st);
      }
      TCI_HIP(hipStreamSynchronize(st));
      t_sigma += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (precision == TCI_SIGMA_FP64) n64 += (int)ks.size();
      else nmix += (int)ks.size();
    };
    // initial guess: unit vectors at the lowest diagonal elements
    // (init_from_diag), ties to the lower index; one device pass per root
    std::vector<int64_t> idx;
    {
      const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, cdiv(n, 256 * 4)));
      DevBuf sel;
      sel.alloc((size_t)nb * 16);
      double* bv = sel.as<double>();
      int64_t* bi = reinterpret_cast<int64_t*>(bv + nb);
      std::vector<double> hv(nb);
      std::vector<int64_t> hi(nb);
      double pv = -INFINITY;
      int64_t pi = -1;
      for (int r = 0; r < roots; ++r) {
        hipLaunchKernelGGL(argmin_after_kernel, dim3(nb), dim3(256), 0, st, diag, n, pv, pi, bv, bi);
        TCI_HIP(hipGetLastError());
        TCI_HIP(hipMemcpyAsync(hv.data(), bv, nb * sizeof(double), hipMemcpyDeviceToHost, st));
        TCI_HIP(hipMemcpyAsync(hi.data(), bi, nb * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        TCI_HIP(hipStreamSynchronize(st));
        double best = 0.0;
        int64_t bidx = -1;
        for (int b = 0; b < nb; ++b)
          if (hi[b] >= 0 && vi_less(hv[b], hi[b], best, bidx)) {
            best = hv[b];
            bidx = hi[b];
          }
        TCI_REQUIRE(bidx >= 0, "fewer ordered diagonal elements than roots");
        idx.push_back(bidx);
        pv = best;
        pi = bidx;
      }
    }
    int m = 0;
    for (int r = 0; r < roots; ++r) {
      hipLaunchKernelGGL(unit_kernel, dim3((unsigned)std::min<int64_t>(cdiv(n, 256), 8192)), dim3(256), 0, st, Vk(m), n,
                         idx[r]);
      TCI_HIP(hipGetLastError());
      ++m;
    }
    std::vector<int> root2idx(roots);
    std::iota(root2idx.begin(), root2idx.end(), 0);
    std::vector<char> conv(roots, 0), met(roots, 0), ready(roots, 0);
    std::vector<double> energy(roots, 0.0), dE(roots, 0.0);
    std::vector<double> h((size_t)mv * mv, 0.0);  // row-major, leading dim mv
    if (gen) gs.assign((size_t)mv * mv, 0.0);
    // subspace eigenpairs of the leading m x m blocks (ascending)
    auto subspace_solve = [&](int m, std::vector<double>& w_, std::vector<double>& Y_) {
      std::vector<double> hm((size_t)m * m), gm;
      for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) hm[(size_t)a * m + b] = h[(size_t)a * mv + b];
      if (!gen) {
        jacobi_eigh(m, hm, w_, Y_);
        return;
      }
      gm.resize((size_t)m * m);
      for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) gm[(size_t)a * m + b] = 0.5 * (gs[(size_t)a * mv + b] + gs[(size_t)b * mv + a]);
      gen_eigh(m, hm, gm, w_, Y_);
    };
    std::vector<double> w, Y, Yprev;
    std::vector<char> restarts;
    int it = 0;
    bool all_conv = false;
    bool sig_done = false;  // the last iteration graph built the new vectors' sigmas and h columns
    bool sigma_timed = true;
    const IterLayout IL(mslots, D.nblk);
    if (igraph) {
      TCI_REQUIRE(roots <= IterLayout::RMAX, "too many roots for iteration graphs");
      iter_ensure(*plan, IL);
    }
    for (it = 1; it <= maxit; ++it) {
      if (!sig_done) {
        // new vectors of the unsettled roots: their sigmas, then
        // h(q,k) = <c_q, sigma_k> for q <= k in one pass over the basis
        std::vector<int> ks;
        for (int r = 0; r < roots; ++r)
          if (!conv[r] && !ready[r]) ks.push_back(root2idx[r]);
        sigma_many(ks);
        if (!ks.empty()) {
          const int kmax = *std::max_element(ks.begin(), ks.end());
          std::vector<double*> X, Ss;
          for (int q = 0; q <= kmax; ++q) X.push_back(Vk(q));
          for (int k : ks) Ss.push_back(Sk(k));
          std::vector<double> hv;
          D.mdotm(X, Ss, hv);
          const int R = (int)ks.size();
          for (int j = 0; j < R; ++j) {
            const int k = ks[j];
            for (int q = 0; q <= k; ++q) h[(size_t)q * mv + k] = h[(size_t)k * mv + q] = hv[(size_t)q * R + j];
          }
        }
        if (gen)
          for (int k : ks) {
            const std::vector<double> gv = g_column(k, k + 1);
            for (int q = 0; q <= k; ++q) gs[(size_t)q * mv + k] = gs[(size_t)k * mv + q] = gv[q];
          }
      }
      sig_done = false;
      subspace_solve(m, w, Y);
      for (int r = 0; r < roots; ++r) {
        dE[r] = std::fabs(w[r] - energy[r]);
        energy[r] = w[r];
      }
      bool restarted = false;
      // mixed -> FP64 switch
      bool do_switch = false;
      // in the mixed phase no root is declared converged: a root that meets
      // the convergence rules with the split sigma (|dE| < switch_tol or a
      // residual below norm_tol) is settled and gets no new vectors; once every
      // root has settled the subspace collapses to the Ritz vectors and the
      // solve continues with FP64 sigma; half of maxit forces the switch
      if (precision != TCI_SIGMA_FP64 && it > 1) {
        do_switch = true;
        for (int r = 0; r < roots; ++r)
          if (!conv[r] && !ready[r]) do_switch = false;
        if (it >= std::max(2, maxit / 2)) do_switch = true;
      }
      const bool budget = (m > 4 * roots && m + 3 * roots > mv);
      if (do_switch || budget) {
        // collapse to the current (and, for a budget restart, previous) Ritz vectors
        int newm = 0;
        // build new basis in tmp slots at the end of V/S, then move
        std::vector<std::vector<double>> cols;
        for (int r = 0; r < roots; ++r) {
          std::vector<double> c(m);
          for (int a = 0; a < m; ++a) c[a] = Y[(size_t)a * m + r];
          cols.push_back(c);
          if (budget && !do_switch && !Yprev.empty()) {
            const int mp = (int)std::lround(std::sqrt((double)Yprev.size()));
            std::vector<double> cp(m, 0.0);
            for (int a = 0; a < std::min(m, mp); ++a) cp[a] = Yprev[(size_t)a * mp + r];
            cols.push_back(cp);
          }
        }
        // new vectors go to slots m .. m+ncols-1 (m <= mv, ncols <= 2*roots)
        TCI_REQUIRE(m + (int)cols.size() <= mslots, "subspace store too small for a collapse");
        std::vector<double*> Xv, Xs;
        for (int q = 0; q < m; ++q) {
          Xv.push_back(Vk(q));
          Xs.push_back(Sk(q));
        }
        std::vector<int> slots;
        for (auto& c : cols) {
          const int slot = m + (int)slots.size();
          D.lincomb(Vk(slot), 0.0, Xv, c);
          std::vector<double*> prevV;
          for (int sidx : slots) prevV.push_back(Vk(sidx));
          if (!do_switch) D.lincomb(Sk(slot), 0.0, Xs, c);
          // orthonormalise against already kept vectors (and carry sigma along)
          bool ok = true;
          if (!prevV.empty()) {
            std::vector<double> ov;
            D.mdot(prevV, Vk(slot), ov);
            for (double& o : ov) o = -o;
            D.lincomb(Vk(slot), 1.0, prevV, ov);
            if (!do_switch) {
              std::vector<double*> prevS;
              for (int sidx : slots) prevS.push_back(Sk(sidx));
              D.lincomb(Sk(slot), 1.0, prevS, ov);
            }
          }
          const double nr = D.norm(Vk(slot));
          if (nr < 1e-8) ok = false;
          if (ok) {
            D.scale(Vk(slot), 1.0 / nr);
            if (!do_switch) D.scale(Sk(slot), 1.0 / nr);
            slots.push_back(slot);
          }
        }
        // move kept vectors to the front
        for (size_t q = 0; q < slots.size(); ++q) {
          TCI_HIP(hipMemcpyAsync(Vk((int)q), Vk(slots[q]), n * 8, hipMemcpyDeviceToDevice, st));
          if (!do_switch) TCI_HIP(hipMemcpyAsync(Sk((int)q), Sk(slots[q]), n * 8, hipMemcpyDeviceToDevice, st));
        }
        newm = (int)slots.size();
        m = newm;
        if (do_switch) {
          precision = TCI_SIGMA_FP64;
          std::fill(ready.begin(), ready.end(), 0);
          std::vector<int> all(m);
          std::iota(all.begin(), all.end(), 0);
          sigma_many(all);
        }
        for (int a = 0; a < m; ++a) {
          std::vector<double*> X;
          for (int q = 0; q < m; ++q) X.push_back(Vk(q));
          std::vector<double> hv;
          D.mdot(X, Sk(a), hv);
          for (int q = 0; q < m; ++q) h[(size_t)q * mv + a] = hv[q];
        }
        for (int a = 0; a < m; ++a)
          for (int b = a + 1; b < m; ++b) {
            const double s2 = 0.5 * (h[(size_t)a * mv + b] + h[(size_t)b * mv + a]);
            h[(size_t)a * mv + b] = h[(size_t)b * mv + a] = s2;
          }
        if (gen)
          for (int a = 0; a < m; ++a) {
            const std::vector<double> gv = g_column(a, m);
            for (int q = 0; q < m; ++q) gs[(size_t)q * mv + a] = gv[q];
          }
        subspace_solve(m, w, Y);
        // after the switch |dE| compares the FP64 Ritz values of the collapsed
        // space with the mixed ones: the Rayleigh-quotient error of the split
        // sigma (1e-10 Eh scale), so the energy criterion stays meaningful
        for (int r = 0; r < roots; ++r) {
          if (do_switch) dE[r] = std::fabs(w[r] - energy[r]);
          energy[r] = w[r];
        }
        restarted = !do_switch;
        if (restarted) ++n_collapse;
      }
      restarts.push_back(restarted ? 1 : 0);
      Yprev = Y;
      if (prm->verbose) std::printf("  Iteration %3d has finished (m=%d, %s)\n", it, m, precision == TCI_SIGMA_FP64 ? "fp64" : "mixed");
      const int mcur = m;
      std::vector<double*> Xv, Xs;
      for (int q = 0; q < mcur; ++q) {
        Xv.push_back(Vk(q));
        Xs.push_back(Sk(q));
      }
      // residuals r = sum_k y_k (sigma_k - E G c_k) of the unsettled roots
      // (build_residual_gen_eig for the generalized problem, G = 1 otherwise)
      // go straight into the free basis slots m, m + 1, ...: the restart rule
      // above keeps m + roots <= mv.  For the standard problem one pass over
      // the basis and its sigmas builds all of them.
      std::vector<int> act;
      for (int r = 0; r < roots; ++r)
        if (!conv[r] && !ready[r]) act.push_back(r);
      const int na = (int)act.size();
      TCI_REQUIRE(m + na <= mv, "Davidson subspace budget exhausted");
      if (igraph && na > 0 && it < maxit) {
        // the rules the host can apply before the residuals exist (dE); the
        // residual-norm rule runs inside the graph
        const bool in_restart = restarts.back() || (restarts.size() >= 2 && restarts[restarts.size() - 2]);
        const double etol = precision != TCI_SIGMA_FP64 ? prm->switch_tol : prm->energy_tol;
        std::vector<int> order, stop;  // indices into act: new vectors first, then roots that stop on dE
        for (int i = 0; i < na; ++i) (!in_restart && it > 1 && dE[act[i]] < etol ? stop : order).push_back(i);
        const int ne = (int)order.size();
        order.insert(order.end(), stop.begin(), stop.end());
        if (ne > 0) {
          if (precision != TCI_SIGMA_FP64) mixed_prepare(*plan, precision, st);
          iter_ensure(*plan, IL);
          char* hb = static_cast<char*>(plan->iter_host);
          IterState hs{Vp, Sp, gen ? gd : nullptr, n, m, 0, prm->norm_tol};
          std::memcpy(hb, &hs, sizeof(hs));
          double* heng = reinterpret_cast<double*>(hb + IL.eng);
          double* heres = reinterpret_cast<double*>(hb + IL.eres);
          int* hrnok = reinterpret_cast<int*>(hb + IL.rnok);
          double* hcres = reinterpret_cast<double*>(hb + IL.cres);
          for (int t = 0; t < ne; ++t) {
            heng[t] = energy[act[order[t]]];
            hrnok[t] = in_restart ? 0 : 1;
          }
          for (int i = 0; i < na; ++i) {
            const int r = act[order[i]];
            heres[i] = energy[r];
            for (int q = 0; q < m; ++q) {
              const double y = Y[(size_t)q * m + r];
              hcres[(size_t)i * IL.ld + 2 * q] = y;
              hcres[(size_t)i * IL.ld + 2 * q + 1] = gen ? y : -energy[r] * y;
            }
          }
          hipGraphExec_t ex = iter_graph(*plan, IL, D.nblk, precision, na, ne, gen);
          TCI_HIP(hipMemcpyAsync(plan->iter_buf.p, hb, IL.cres + (size_t)na * IL.ld * 8, hipMemcpyHostToDevice, st));
          TCI_HIP(hipGraphLaunch(ex, st));
          TCI_HIP(hipStreamSynchronize(st));
          const long long* hts = reinterpret_cast<const long long*>(hb + IL.ts);
          if (plan->iter_clock_khz > 0 && hts[1] >= hts[0])
            t_sigma += (double)(hts[1] - hts[0]) / (1e3 * plan->iter_clock_khz);
          else
            sigma_timed = false;
          if (precision == TCI_SIGMA_FP64) n64 += ne;
          else nmix += ne;
          const double* hrd = reinterpret_cast<const double*>(hb + IL.rd);
          const double* hmask = reinterpret_cast<const double*>(hb + IL.mask);
          const double* hhd = reinterpret_cast<const double*>(hb + IL.hd);
          std::vector<int> pos(na);
          for (int i = 0; i < na; ++i) pos[order[i]] = i;
          if (prm->verbose)
            for (int i = 0; i < na; ++i) {
              const int r = act[i];
              std::printf("  %5d %5d %22.12f %15.8e %15.8e\n", it, r + 1, energy[r] + plan->e_ref,
                          std::sqrt(std::max(hrd[pos[i]], 0.0)), dE[r]);
            }
          // roots that stop: on dE (host) or on the residual norm (masked in the graph)
          std::vector<int> kept;  // new-vector indices t that join the basis
          for (int i = 0; i < na; ++i) {
            const int r = act[order[i]];
            if (i < ne && hmask[i] == 0.0) {
              kept.push_back(i);
              continue;
            }
            if (precision != TCI_SIGMA_FP64) {
              ready[r] = 1;
            } else {
              met[r] = 1;
              conv[r] = 1;
            }
          }
          // kept vectors to consecutive slots (a masked vector is zero, its sigma too)
          for (int u = 0; u < (int)kept.size(); ++u)
            if (kept[u] != u) {
              TCI_HIP(hipMemcpyAsync(Vk(m + u), Vk(m + kept[u]), n * 8, hipMemcpyDeviceToDevice, st));
              TCI_HIP(hipMemcpyAsync(Sk(m + u), Sk(m + kept[u]), n * 8, hipMemcpyDeviceToDevice, st));
            }
          // h(q, k) = <c_q, sigma_k> for the new k, q <= k
          const double* hgsd = reinterpret_cast<const double*>(hb + IL.gsd);
          auto col = [&](const double* blk, int q_old, int t) {
            const int g = t / 4, R = std::min(4, ne - 4 * g);
            return blk[(size_t)4 * IL.S * g + (size_t)q_old * R + t % 4];
          };
          for (int u = 0; u < (int)kept.size(); ++u) {
            const int k = m + u, t = kept[u];
            for (int q = 0; q <= k; ++q) {
              const int q_old = q < m ? q : m + kept[q - m];
              h[(size_t)q * mv + k] = h[(size_t)k * mv + q] = col(hhd, q_old, t);
              if (gen) gs[(size_t)q * mv + k] = gs[(size_t)k * mv + q] = col(hgsd, q_old, t);
            }
            root2idx[act[order[t]]] = k;
          }
          m += (int)kept.size();
          sig_done = true;
          all_conv = std::all_of(conv.begin(), conv.end(), [](char c) { return c != 0; });
          if (all_conv) break;
          continue;
        }
      }
      std::vector<double*> rs;
      for (int i = 0; i < na; ++i) rs.push_back(Vk(m + i));
      if (gen) {
        for (int i = 0; i < na; ++i) {
          const int r = act[i];
          std::vector<double> c(mcur);
          for (int q = 0; q < mcur; ++q) c[q] = Y[(size_t)q * mcur + r];
          D.lincomb(rs[i], 0.0, Xs, c);
          D.lincomb(tmp, 0.0, Xv, c);
          hipLaunchKernelGGL(gaxpy_kernel, dim3(gxn), dim3(256), 0, st, rs[i], -energy[r], gd, tmp, n);
          TCI_HIP(hipGetLastError());
        }
      } else if (na > 0) {
        std::vector<double*> X;
        for (int q = 0; q < mcur; ++q) {
          X.push_back(Sk(q));
          X.push_back(Vk(q));
        }
        std::vector<double> c((size_t)na * 2 * mcur);
        for (int i = 0; i < na; ++i)
          for (int q = 0; q < mcur; ++q) {
            const double y = Y[(size_t)q * mcur + act[i]];
            c[(size_t)i * 2 * mcur + 2 * q] = y;
            c[(size_t)i * 2 * mcur + 2 * q + 1] = -energy[act[i]] * y;
          }
        D.lincombm(rs, 0.0, X, c);
      }
      std::vector<double> rdots;
      D.mdotm(rs, rs, rdots);
      std::vector<int> kept;  // indices into act of the roots that get a new vector
      for (int i = 0; i < na; ++i) {
        const int r = act[i];
        const double rn = std::sqrt(std::max(rdots[(size_t)i * na + i], 0.0));
        if (prm->verbose)
          std::printf("  %5d %5d %22.12f %15.8e %15.8e\n", it, r + 1, energy[r] + plan->e_ref, rn, dE[r]);
        // is_converged (davidson.f90:564-602)
        bool halt = false;
        const bool in_restart = restarts.back() || (restarts.size() >= 2 && restarts[restarts.size() - 2]);
        if (precision != TCI_SIGMA_FP64) {
          if (!in_restart && (rn < prm->norm_tol || (dE[r] < prm->switch_tol && it > 1))) {
            ready[r] = 1;
            continue;
          }
        } else if (!in_restart) {
          if (rn < prm->norm_tol) halt = true;
          else if (dE[r] < prm->energy_tol && it > 1) halt = true;
        }
        if (halt) met[r] = 1;
        if (it == maxit) halt = true;
        if (halt) {
          conv[r] = 1;
          continue;
        }
        kept.push_back(i);
      }
      // the kept residuals move to the front of the free slots (in root
      // order, so a source slot is never overwritten before it is read) and
      // are preconditioned in place: r_i <- -r_i / (H_ii - E_ref - E g_i), 0
      // where that is 0 (build_next_vector_std_eig / build_next_vector_gen_eig)
      std::vector<double*> W;
      for (int t = 0; t < (int)kept.size(); ++t) {
        const int i = kept[t], r = act[i];
        if (i != t) TCI_HIP(hipMemcpyAsync(Vk(m + t), Vk(m + i), n * 8, hipMemcpyDeviceToDevice, st));
        if (gen)
          hipLaunchKernelGGL(precond_gen_kernel, dim3(gxn), dim3(256), 0, st, Vk(m + t), diag, gd, energy[r], n);
        else
          hipLaunchKernelGGL(precond_kernel, dim3(gxn), dim3(256), 0, st, Vk(m + t), diag, energy[r], n);
        TCI_HIP(hipGetLastError());
        W.push_back(Vk(m + t));
      }
      if (!W.empty()) {
        std::vector<double*> prev;
        for (int q = 0; q < m; ++q) prev.push_back(Vk(q));
        D.orthonormalize_block(W, prev);
        for (int t = 0; t < (int)kept.size(); ++t) root2idx[act[kept[t]]] = m + t;
        m += (int)kept.size();
      }
      all_conv = std::all_of(conv.begin(), conv.end(), [](char c) { return c != 0; });
      if (all_conv) break;
    }
    // final vectors
    for (int r = 0; r < roots; ++r) energies[r] = energy[r] + plan->e_ref;
    if (vectors) {
      const int mcur = (int)std::lround(std::sqrt((double)Y.size()));
      std::vector<double*> X;
      for (int q = 0; q < mcur; ++q) X.push_back(Vk(q));
      for (int r = 0; r < roots; ++r) {
        std::vector<double> c(mcur);
        for (int q = 0; q < mcur; ++q) c[q] = Y[(size_t)q * mcur + r];
        double* out = vectors + (int64_t)r * n;
        D.lincomb(out, 0.0, X, c);
        const double nr = D.norm(out);
        if (nr > 0) D.scale(out, 1.0 / nr);
      }
    }
    TCI_HIP(hipStreamSynchronize(st));
    if (result) {
      result->iterations = std::min(it, maxit);
      result->sigma_fp64 = n64;
      result->sigma_mixed = nmix;
      result->converged = std::all_of(met.begin(), met.end(), [](char c) { return c != 0; }) ? 1 : 0;
      result->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
      result->seconds_sigma = sigma_timed ? t_sigma : -1.0;
      result->collapses = n_collapse;
    }
    return TCI_OK;
  } catch (const tci::Error& e) {
    tci::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    tci::set_last_error(e.what());
    return TCI_ERR_ARG;
  }
}

}  // namespace

extern "C" int tci_davidson_run(tci_plan* plan, const tci_davidson_params* prm, double* energies, double* vectors,
                                tci_davidson_result* result) {
  return davidson_impl(plan, prm, nullptr, energies, vectors, result);
}

extern "C" int tci_davidson_gen_run(tci_plan* plan, const tci_davidson_params* prm, const double* g_values,
                                    double* energies, double* vectors, tci_davidson_result* result) {
  if (!g_values) {
    tci::set_last_error("null g_values");
    return TCI_ERR_ARG;
  }
  return davidson_impl(plan, prm, g_values, energies, vectors, result);
}

// get_gvalues (gtimesc.f90): g1..g5 of the ACPF-type functionals from the
// number of correlated electrons n (electrons minus 2 per frozen orbital)
extern "C" int tci_acpf_gvalues(int32_t functional, int32_t n_corr_elec, double custom_g, double* g) {
  if (!g || n_corr_elec < 2) {
    tci::set_last_error("g must be non-null and n_corr_elec >= 2");
    return TCI_ERR_ARG;
  }
  const double n = (double)n_corr_elec;
  const double acpf = 2.0 / n, acpf2 = (4.0 / n) * (1.0 - 1.0 / (2.0 * (n - 1.0)));
  g[0] = g[1] = 1.0;
  switch (functional) {
    case TCI_ACPF: g[2] = g[3] = g[4] = acpf; break;
    case TCI_ACPF2: g[2] = g[3] = acpf2; g[4] = acpf; break;
    case TCI_MRAQCC: g[2] = g[3] = g[4] = acpf2; break;
    case TCI_ACPF_CUSTOM: g[2] = g[3] = custom_g; g[4] = acpf; break;
    case TCI_ACPF_TEST: g[2] = g[3] = g[4] = 1.0; break;
    default:
      tci::set_last_error("unknown functional");
      return TCI_ERR_ARG;
  }
  return TCI_OK;
}
