// Elementwise / gather kernels of the sigma build.
//
// Dense pair blocks: for strong walk P = (i,j), i >= j, with domain D (sorted
// virtuals, d = |D|), C^{ij} is a d x d row-major block whose row index pairs
// with internal orbital i.  From the CSF vector (scepper, scepmaker_mod.f90:
// 121-212):  i > j:  C = S+(x_ab+, x_aa)/2 + S-(x_ab-)/(2 sqrt3)
//            i = j:  C = S+(x_ab, x_aa)/sqrt2
// S+ puts v at (r,c) and (c,r) and sqrt2*w on the diagonal; S- puts +v above
// and -v below the diagonal.  Packing applies the adjoints (unscepper,
// scepmaker_mod.f90:633-733) with factors 1 (i>j, singlet), sqrt3 (i>j,
// triplet), 1/sqrt2 (i=j).  Packed index of (hi, lo), hi > lo: hi(hi-1)/2+lo.
#include "terms.hpp"

#include <cmath>

namespace tci {

namespace {

constexpr double SQRT2 = 1.41421356237309504880;
constexpr double INV_SQRT2 = 0.70710678118654752440;
constexpr double SQRT3 = 1.73205080756887729353;
constexpr double INV_2SQRT3 = 0.28867513459481288225;  // 1/(2 sqrt3)

__device__ __forceinline__ void tri_decode(int64_t idx, int& hi, int& lo) {
  // largest hi with hi(hi-1)/2 <= idx
  int h = (int)((1.0 + sqrt(1.0 + 8.0 * (double)idx)) * 0.5);
  while ((int64_t)h * (h - 1) / 2 > idx) --h;
  while ((int64_t)(h + 1) * h / 2 <= idx) ++h;
  hi = h;
  lo = (int)(idx - (int64_t)h * (h - 1) / 2);
}

// CSF index of the singles element e = i n_v + a: the N-1 walks are laid out
// in the reference's lexical walk order, which for a closed-shell reference
// runs over the hole orbital in descending order (index_vector_internal,
// graph_mod.f90:1032-1589, with the DRT weights of drt_maker), virtuals
// ascending within a walk
__device__ __forceinline__ int64_t single_csf(const PairTab& T, int64_t e) {
  const int64_t i = e / T.n_v, a = e - i * T.n_v;
  return 1 + (T.n_int - 1 - i) * T.n_v + a;
}

__global__ void unpack_kernel(PairTab T, const double* __restrict__ x, double* __restrict__ C,
                              double* __restrict__ Ct, double* __restrict__ s, const IoSlot* __restrict__ io) {
  if (io) x = io->x;
  const int P = blockIdx.y;
  if (P == T.n_pairs) {  // singles: s[i][a] from the N-1 walk of hole i (walks by descending hole)
    const int64_t n = (int64_t)T.n_int * T.n_v;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
      s[e] = x[single_csf(T, e)] * INV_SQRT2;
    return;
  }
  const int i = T.pairs[2 * P], j = T.pairs[2 * P + 1];
  const int d = (int)(T.dom_off[P + 1] - T.dom_off[P]);
  const int64_t dd = (int64_t)d * d;
  const int64_t m2 = (int64_t)d * (d - 1) / 2;
  const double* xa = x + T.aa_off[P];
  const double* xb = x + T.ab_off[P];
  double* Cb = C + T.cd_off[P];
  double* Ctb = Ct + T.cd_off[P];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < dd; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / d), c = (int)(e - (int64_t)r * d);
    double crc, ccr;
    // i > j: the doubly occupied (aa) and singlet-coupled (ab) CSFs enter
    // with a minus sign, the phase of the reference's SGGA spin functions
    // relative to the valence CSF (coupling coefficients (-1, -sqrt3) of the
    // valence <-> N-2 loop, three_and_four_seg_mod_2.f90:1524-1533,8007-8162)
    if (r == c) {
      crc = ccr = (i == j) ? xa[r] : -xa[r] * INV_SQRT2;
    } else {
      const int hi = r > c ? r : c, lo = r > c ? c : r;
      const int64_t idx = (int64_t)hi * (hi - 1) / 2 + lo;
      if (i == j) {
        crc = ccr = xb[idx] * INV_SQRT2;
      } else {
        const double vs = -0.5 * xb[idx];
        const double q = xb[m2 + idx] * INV_2SQRT3;
        crc = r < c ? vs + q : vs - q;
        ccr = r < c ? vs - q : vs + q;
      }
    }
    Cb[e] = crc;
    Ctb[e] = 2.0 * crc - ccr;
  }
}

// R_P[r][c] += c0 (ia|jb) - sum_k [(ki|bj) s_k^a + (kj|ai) s_k^b] at
// a = D_P[r], b = D_P[c] (O3[k][p][q][b] = (kp|bq)).  A workgroup takes a
// strip of RT_ROWS rows of one pair block: the strip's row factors
// s_k^a and (kj|ai) go through LDS, RT_K values of k at a time, and each
// thread keeps its column's factors (ki|bj) and s_k^b in registers while it
// accumulates all rows of the strip, in the same order over k as a
// per-element loop.
constexpr int RT_ROWS = 16, RT_K = 32, RT_THREADS = 128;
__global__ __launch_bounds__(RT_THREADS) void rterms_kernel(PairTab T, const double* __restrict__ x,
                                                             const double* __restrict__ s,
                                                             const double* __restrict__ Kov,
                                                             const double* __restrict__ O3, double* __restrict__ R,
                                                             const IoSlot* __restrict__ io) {
  if (io) x = io->x;
  const int P = blockIdx.y;
  const int64_t off = T.dom_off[P];
  const int d = (int)(T.dom_off[P + 1] - off);
  const int r0 = blockIdx.x * RT_ROWS;
  if (r0 >= d) return;  // whole workgroup: no barrier is skipped by part of it
  const int nr = min(RT_ROWS, d - r0);
  const int i = T.pairs[2 * P], j = T.pairs[2 * P + 1];
  const int n = T.n_int, nv = T.n_v;
  const int64_t niv = (int64_t)n * nv;
  const double c0 = x[0];
  double* Rb = R + T.cd_off[P];
  __shared__ double sa[RT_K][RT_ROWS], oa[RT_K][RT_ROWS];
  __shared__ int arow[RT_ROWS];
  if ((int)threadIdx.x < RT_ROWS) arow[threadIdx.x] = T.dom_idx[off + r0 + min((int)threadIdx.x, nr - 1)];
  __syncthreads();
  for (int cb = 0; cb < d; cb += RT_THREADS) {
    const int c = cb + (int)threadIdx.x;
    const bool col = c < d;
    const int b = col ? T.dom_idx[off + c] : 0;
    double acc[RT_ROWS];
#pragma unroll
    for (int rr = 0; rr < RT_ROWS; ++rr) acc[rr] = 0.0;
    for (int k0 = 0; k0 < n; k0 += RT_K) {
      const int kn = min(RT_K, n - k0);
      __syncthreads();  // the previous factors are consumed
      for (int t = threadIdx.x; t < RT_K * RT_ROWS; t += RT_THREADS) {
        const int kk = t / RT_ROWS, rr = t - kk * RT_ROWS;
        double vs = 0.0, vo = 0.0;
        if (kk < kn && rr < nr) {
          const int k = k0 + kk, a = arow[rr];
          vs = s[(int64_t)k * nv + a];
          vo = O3[(((int64_t)k * n + j) * n + i) * nv + a];
        }
        sa[kk][rr] = vs;
        oa[kk][rr] = vo;
      }
      __syncthreads();
      if (col)
        for (int kk = 0; kk < kn; ++kk) {
          const int k = k0 + kk;
          const double ob = O3[(((int64_t)k * n + i) * n + j) * nv + b];
          const double sb = s[(int64_t)k * nv + b];
#pragma unroll
          for (int rr = 0; rr < RT_ROWS; ++rr) {
            acc[rr] += ob * sa[kk][rr];
            acc[rr] += oa[kk][rr] * sb;
          }
        }
    }
    if (col)
#pragma unroll
      for (int rr = 0; rr < RT_ROWS; ++rr)
        if (rr < nr) {
          const int a = arow[rr];
          const int64_t e = (int64_t)(r0 + rr) * d + c;
          Rb[e] += c0 * Kov[((int64_t)i * nv + a) * niv + (int64_t)j * nv + b] - acc[rr];
        }
  }
}

// block (P, row chunk of 16 rows); loops over all strong Q
__global__ void hole_kernel(PairTab T, HoleTab H, const double* __restrict__ C, const double* __restrict__ O4,
                            double* __restrict__ R) {
  const int P = blockIdx.y;
  const int i = T.pairs[2 * P], j = T.pairs[2 * P + 1];
  const int d = (int)(T.dom_off[P + 1] - T.dom_off[P]);
  const int r0 = blockIdx.x * 16;
  if (r0 >= d) return;
  const int n = T.n_int;
  double* Rb = R + T.cd_off[P];
  for (int Q = 0; Q < T.n_pairs; ++Q) {
    const int k = T.pairs[2 * Q], l = T.pairs[2 * Q + 1];
    const double cA = O4[(((int64_t)k * n + i) * n + l) * n + j];                 // (ki|lj)
    const double cB = (k != l) ? O4[(((int64_t)l * n + i) * n + k) * n + j] : 0.0;  // (li|kj)
    const int64_t s0 = H.isect_off[(int64_t)P * T.n_pairs + Q];
    const int m = (int)(H.isect_off[(int64_t)P * T.n_pairs + Q + 1] - s0);
    if (m == 0) continue;
    const int dq = (int)(T.dom_off[Q + 1] - T.dom_off[Q]);
    const double* Cq = C + T.cd_off[Q];
    const int32_t* is = H.isect + 2 * s0;
    // rows of the intersection whose P-position lies in [r0, r0+16)
    for (int e = threadIdx.x; e < 16 * m; e += blockDim.x) {
      const int xr = e / m, y = e - xr * m;
      // find intersection entry with P-position r0 + xr: entries are sorted by P-position
      // (binary search over the m entries)
      int lo = 0, hi = m;
      const int target = r0 + xr;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (is[2 * mid] < target) lo = mid + 1; else hi = mid;
      }
      if (lo >= m || is[2 * lo] != target) continue;
      const int qx = is[2 * lo + 1];
      const int py = is[2 * y], qy = is[2 * y + 1];
      Rb[(int64_t)target * d + py] += cA * Cq[(int64_t)qx * dq + qy] + cB * Cq[(int64_t)qy * dq + qx];
    }
    __syncthreads();
  }
}

// Row slices of a pair block per ref_dot block: each of a block's four waves
// takes rows y*4 + w, stepping by the slices' 32 rows, lanes over columns
// (coalesced reads of C and of the gathered row of Kov).
constexpr int kRefSlices = 8;

__global__ __launch_bounds__(256) void ref_dot_kernel(PairTab T, const double* __restrict__ Kov,
                                                      const double* __restrict__ Ct, double* __restrict__ part) {
  const int P = blockIdx.x, y = blockIdx.y;
  const int i = T.pairs[2 * P], j = T.pairs[2 * P + 1];
  const int64_t off = T.dom_off[P];
  const int d = (int)(T.dom_off[P + 1] - off);
  const int nv = T.n_v;
  const int64_t niv = (int64_t)T.n_int * nv;
  const double* Cb = Ct + T.cd_off[P];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int r = y * 4 + wv; r < d; r += kRefSlices * 4) {
    const double* krow = Kov + ((int64_t)i * nv + T.dom_idx[off + r]) * niv + (int64_t)j * nv;
    const double* crow = Cb + (int64_t)r * d;
    for (int c = lane; c < d; c += 64) acc += krow[T.dom_idx[off + c]] * crow[c];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  __shared__ double red[4];
  if (lane == 0) red[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    part[(int64_t)P * kRefSlices + y] = (i == j ? 1.0 : 2.0) * ((red[0] + red[1]) + (red[2] + red[3]));
}

__global__ void sum_kernel(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) acc += part[e];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// block (Q, o): ordered pair (j,k) = o==0 ? (qi,qj) : (qj,qi)
// out[i][r] = -sum_c Ct'[r][c] W[i][c], W[i][c] = (jj i|kk D[c]) (a row of O3),
// Ct' = Ct (o = 0) or Ct^T (o = 1).  Block (Q, o, 64 rows of r); c in chunks of
// 32 staged in LDS with coalesced reads of Ct, i in groups of 64 (16 per thread).
constexpr int S3R = 64, S3C = 32;

__global__ __launch_bounds__(256) void s3int_singles_kernel(PairTab T, const int64_t* __restrict__ gpart_off,
                                                            const double* __restrict__ Ct,
                                                            const double* __restrict__ O3,
                                                            double* __restrict__ gpart) {
  const int Q = blockIdx.x, o = blockIdx.y, r0 = blockIdx.z * S3R;
  const int qi = T.pairs[2 * Q], qj = T.pairs[2 * Q + 1];
  const int64_t off = T.dom_off[Q];
  const int d = (int)(T.dom_off[Q + 1] - off);
  if ((o == 1 && qi == qj) || r0 >= d) return;
  const int jj = o == 0 ? qi : qj, kk = o == 0 ? qj : qi;
  const int n = T.n_int, nv = T.n_v;
  const double* Cb = Ct + T.cd_off[Q];
  double* out = gpart + gpart_off[2 * Q + o];  // [n_int][d]
  __shared__ double sC[S3R][S3C + 1];
  __shared__ double sW[64][S3C + 1];
  const int t = threadIdx.x, r = t & 63, ig = t >> 6;
  for (int i0 = 0; i0 < n; i0 += 64) {
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    for (int c0 = 0; c0 < d; c0 += S3C) {
      __syncthreads();
      for (int e = t; e < S3R * S3C; e += 256) {
        int rr, cc;
        if (o == 0) {
          rr = e / S3C;
          cc = e % S3C;
        } else {
          rr = e % S3R;
          cc = e / S3R;
        }
        const int gr = r0 + rr, gc = c0 + cc;
        sC[rr][cc] = (gr < d && gc < d) ? (o == 0 ? Cb[(int64_t)gr * d + gc] : Cb[(int64_t)gc * d + gr]) : 0.0;
      }
      for (int e = t; e < 64 * S3C; e += 256) {
        const int ii = e / S3C, cc = e % S3C;
        const int gi = i0 + ii, gc = c0 + cc;
        sW[ii][cc] = (gi < n && gc < d) ? O3[(((int64_t)jj * n + gi) * n + kk) * nv + T.dom_idx[off + gc]] : 0.0;
      }
      __syncthreads();
      // rows i >= n hold zeros: skip them (ig is the wave index, so the test
      // is uniform over a wave)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        if (i0 + ig + 4 * q >= n) continue;
        double a = 0.0;
#pragma unroll 8
        for (int cc = 0; cc < S3C; ++cc) a += sW[ig + 4 * q][cc] * sC[r][cc];
        acc[q] += a;
      }
    }
    if (r0 + r < d)
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int i = i0 + ig + 4 * q;
        if (i < n) out[(int64_t)i * d + r0 + r] = -acc[q];
      }
  }
}

// g[i][a] = gss + sum of the K-split partials of gW + every gpart entry of a
// walk whose domain holds a.  One block per virtual a; thread (i, slice) sums
// the partials and inverse-map entries of its slice for internal i (64 at a
// time), and the sixteen slices meet in LDS (fixed order: deterministic).
// Sixteen slices keep each thread's chain of dependent loads (inverse map,
// walk tables, then the partial) short: a launch has only n_v blocks.
constexpr int kRedSlices = 16;

__global__ __launch_bounds__(64 * kRedSlices) void singles_reduce_kernel(PairTab T, InvTab I,
                                                                         const int64_t* __restrict__ gpart_off,
                                                                         const double* __restrict__ gpart,
                                                                         const double* __restrict__ gss,
                                                                         const double* __restrict__ gW, int n_split,
                                                                         double* __restrict__ g) {
  const int nv = T.n_v, no = T.n_int, a = blockIdx.x;
  const int64_t n = (int64_t)no * nv;
  const int il = threadIdx.x & 63, sl = threadIdx.x >> 6;
  __shared__ double part[kRedSlices][64];
  const int64_t q0 = I.inv_off[a], q1 = I.inv_off[a + 1];
  for (int i0 = 0; i0 < no; i0 += 64) {
    const int i = i0 + il;
    double v = 0.0;
    if (i < no) {
      const int64_t e = (int64_t)i * nv + a;
      if (sl == 0) v = gss[e];
      for (int sp = sl; sp < n_split; sp += kRedSlices) v += gW[(int64_t)sp * n + e];
      for (int64_t q = q0 + sl; q < q1; q += kRedSlices) {
        const int Q = I.inv_q[q], pos = I.inv_pos[q];
        const int d = (int)(T.dom_off[Q + 1] - T.dom_off[Q]);
        v += gpart[gpart_off[2 * Q] + (int64_t)i * d + pos];
        if (T.pairs[2 * Q] != T.pairs[2 * Q + 1]) v += gpart[gpart_off[2 * Q + 1] + (int64_t)i * d + pos];
      }
    }
    part[sl][il] = v;
    __syncthreads();
    if (sl == 0 && i < no) {
      double s = part[0][il];
      for (int k = 1; k < kRedSlices; ++k) s += part[k][il];
      g[(int64_t)i * nv + a] = s;
    }
    __syncthreads();
  }
}

__global__ void set_io_kernel(IoSlot* io, const double* x, double* y) {
  if (threadIdx.x == 0) {
    io->x = x;
    io->y = y;
  }
}

__global__ void pack_kernel(PairTab T, const double* __restrict__ x, const double* __restrict__ R,
                            const double* __restrict__ g, const double* __restrict__ sig0, double e_ref,
                            int flags, double* __restrict__ sigma, const IoSlot* __restrict__ io) {
  if (io) {
    x = io->x;
    sigma = io->y;
  }
  const bool accum = (flags & TCI_ACCUMULATE) != 0;
  const double eadd = (flags & TCI_ADD_EREF) ? e_ref : 0.0;
  const int P = blockIdx.y;
  const int no = T.n_int, nv = T.n_v;
  if (P == T.n_pairs) {
    const int64_t n = 1 + (int64_t)no * nv;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
      double v;
      if (k == 0) {
        v = *sig0 + eadd * x[0];
      } else {
        const int64_t w = (k - 1) / nv;  // N-1 walk rank: hole no - 1 - w
        const int i = no - 1 - (int)w, a = (int)(k - 1 - w * nv);
        const int64_t e = (int64_t)i * nv + a;
        v = SQRT2 * g[e] + (T.eps[no + a] - T.eps[i] + eadd) * x[k];
      }
      sigma[k] = accum ? sigma[k] + v : v;
    }
    return;
  }
  const int i = T.pairs[2 * P], j = T.pairs[2 * P + 1];
  const int64_t off = T.dom_off[P];
  const int d = (int)(T.dom_off[P + 1] - off);
  const int64_t m2 = (int64_t)d * (d - 1) / 2;
  const int nsp = (i == j) ? 1 : 2;
  const int64_t cnt = d + nsp * m2;
  const int64_t base = T.aa_off[P];   // ab block follows the aa block
  const double* Rb = R + T.cd_off[P];
  const double ei = T.eps[i], ej = T.eps[j];
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < cnt; q += (int64_t)gridDim.x * blockDim.x) {
    double v, fock;
    if (q < d) {
      const int r = (int)q;
      const double ea = T.eps[no + T.dom_idx[off + r]];
      v = (i == j) ? Rb[(int64_t)r * d + r] : -SQRT2 * Rb[(int64_t)r * d + r];
      fock = 2.0 * ea - ei - ej;
    } else {
      int64_t idx = q - d;
      const bool trip = idx >= m2;
      if (trip) idx -= m2;
      int hi, lo;
      tri_decode(idx, hi, lo);
      const double up = Rb[(int64_t)lo * d + hi], dn = Rb[(int64_t)hi * d + lo];
      if (i == j) v = (up + dn) * INV_SQRT2;
      else v = trip ? SQRT3 * (up - dn) : -(up + dn);   // i > j: aa and singlet CSFs with a minus sign
      fock = T.eps[no + T.dom_idx[off + hi]] + T.eps[no + T.dom_idx[off + lo]] - ei - ej;
    }
    const int64_t k = base + q;
    v += (fock + eadd) * x[k];
    sigma[k] = accum ? sigma[k] + v : v;
  }
}

__global__ void diag_kernel(PairTab T, const double* __restrict__ Kov, const double* __restrict__ Jvv,
                            const double* __restrict__ O4, const double* __restrict__ vvd,
                            const double* __restrict__ vvx, double* __restrict__ diag) {
  const int P = blockIdx.y;
  const int no = T.n_int, nv = T.n_v;
  const int64_t niv = (int64_t)no * nv;
  // (ma|ma) and (mm|aa)
  auto Kd = [&](int m, int a) { return Kov[((int64_t)m * nv + a) * niv + (int64_t)m * nv + a]; };
  auto Jd = [&](int m, int a) {
    const int64_t pm = (int64_t)m * (m + 1) / 2 + m;
    return Jvv[pm * nv * nv + (int64_t)a * nv + a];
  };
  auto o4 = [&](int p, int q, int r, int s) { return O4[(((int64_t)p * no + q) * no + r) * no + s]; };
  if (P == T.n_pairs) {
    const int64_t n = 1 + niv;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
      if (k == 0) { diag[0] = 0.0; continue; }
      const int64_t w = (k - 1) / nv;  // N-1 walk rank: hole no - 1 - w
      const int i = no - 1 - (int)w, a = (int)(k - 1 - w * nv);
      diag[k] = T.eps[no + a] - T.eps[i] + 2.0 * Kd(i, a) - Jd(i, a);
    }
    return;
  }
  const int i = T.pairs[2 * P], j = T.pairs[2 * P + 1];
  const int64_t off = T.dom_off[P];
  const int d = (int)(T.dom_off[P + 1] - off);
  const int64_t m2 = (int64_t)d * (d - 1) / 2;
  const int nsp = (i == j) ? 1 : 2;
  const int64_t cnt = d + nsp * m2;
  const double ei = T.eps[i], ej = T.eps[j];
  const double iiii = o4(i, i, i, i), iijj = o4(i, i, j, j), ijij = o4(i, j, i, j);
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < cnt; q += (int64_t)gridDim.x * blockDim.x) {
    double v;
    if (q < d) {
      const int a = T.dom_idx[off + q];
      const double ea = T.eps[no + a];
      const double aaaa = vvd[(int64_t)a * nv + a];
      if (i == j)
        v = 2 * ea - 2 * ei + aaaa + iiii + 2 * Kd(i, a) - 4 * Jd(i, a);
      else
        v = 2 * ea - ei - ej + aaaa + iijj + ijij + Kd(i, a) + Kd(j, a) - 2 * Jd(i, a) - 2 * Jd(j, a);
    } else {
      int64_t idx = q - d;
      const bool trip = idx >= m2;
      if (trip) idx -= m2;
      int hi, lo;
      tri_decode(idx, hi, lo);
      const int a = T.dom_idx[off + hi], b = T.dom_idx[off + lo];
      const double de = T.eps[no + a] + T.eps[no + b] - ei - ej;
      const double aabb = vvd[(int64_t)a * nv + b], abab = vvx[(int64_t)a * nv + b];
      if (i == j) {
        v = de + aabb + abab + iiii + Kd(i, a) + Kd(i, b) - 2 * Jd(i, a) - 2 * Jd(i, b);
      } else {
        const double ks = Kd(i, a) + Kd(i, b) + Kd(j, a) + Kd(j, b);
        const double js = Jd(i, a) + Jd(i, b) + Jd(j, a) + Jd(j, b);
        v = trip ? de + aabb - abab + iijj - ijij + 1.5 * ks - js
                 : de + aabb + abab + iijj + ijij + 0.5 * ks - js;
      }
    }
    diag[T.aa_off[P] + q] = v;
  }
}

// Lpm[(p(p+1)/2+q)][J] -> Lvv[J][a][b] (both halves)
template <class OT>
__global__ void gather_lvv_kernel(const double* __restrict__ Lpm, int64_t n_chol, int n_int, int n_v,
                                  OT* __restrict__ Lvv) {
  // block: one virtual row a, a chunk of J; threads over (b <= a)
  const int a = blockIdx.x;
  const int64_t J0 = (int64_t)blockIdx.y * 32;
  const int64_t pa = (int64_t)(n_int + a);
  for (int t = threadIdx.x; t < (a + 1) * 32; t += blockDim.x) {
    const int jj = t & 31, b = t >> 5;
    const int64_t J = J0 + jj;
    if (J >= n_chol) continue;
    const int64_t rec = pa * (pa + 1) / 2 + n_int + b;
    const double v = Lpm[rec * n_chol + J];
    const int64_t base = J * n_v * n_v;
    Lvv[base + (int64_t)a * n_v + b] = (OT)v;
    Lvv[base + (int64_t)b * n_v + a] = (OT)v;
  }
}

__global__ void gather_lov_kernel(const double* __restrict__ Lpm, int64_t n_chol, int n_int, int n_v,
                                  double* __restrict__ Lov) {
  const int a = blockIdx.x;
  const int64_t J0 = (int64_t)blockIdx.y * 32;
  const int64_t pa = (int64_t)(n_int + a);
  for (int t = threadIdx.x; t < n_int * 32; t += blockDim.x) {
    const int jj = t & 31, i = t >> 5;
    const int64_t J = J0 + jj;
    if (J >= n_chol) continue;
    const int64_t rec = pa * (pa + 1) / 2 + i;
    Lov[(J * n_int + i) * n_v + a] = Lpm[rec * n_chol + J];
  }
}

template <class LT>
__global__ void vv_exch_kernel(const LT* __restrict__ Lvv, int64_t n_chol, int n_v, double* __restrict__ out) {
  const int64_t n = (int64_t)n_v * n_v;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int64_t J = 0; J < n_chol; ++J) {
      const double v = (double)Lvv[J * n + e];
      acc += v * v;
    }
    out[e] = acc;
  }
}

__global__ void build_mss_kernel(const double* __restrict__ Kov, const double* __restrict__ Jvv, int n_int,
                                 int n_v, double* __restrict__ Mss) {
  const int64_t niv = (int64_t)n_int * n_v;
  const int64_t n = niv * niv;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e / niv, col = e - row * niv;
    const int i = (int)(row / n_v), a = (int)(row - (int64_t)i * n_v);
    const int j = (int)(col / n_v), b = (int)(col - (int64_t)j * n_v);
    const int hi = i > j ? i : j, lo = i > j ? j : i;
    const int64_t pij = (int64_t)hi * (hi + 1) / 2 + lo;
    Mss[e] = 2.0 * Kov[e] - Jvv[pij * n_v * n_v + (int64_t)a * n_v + b];
  }
}

// block (local pair q, jj): copy the d x d gathered block of L^{J0+jj}
template <class LT>
__global__ void pack_L_kernel(PairTab T, PackTab K, const LT* __restrict__ Lvv, int64_t J0, int64_t len,
                              LT* __restrict__ Lp) {
  const int q = blockIdx.x;
  const int64_t jj = blockIdx.y;
  const int P = K.pairs_local[q];
  const int64_t off = T.dom_off[P];
  const int d = (int)(T.dom_off[P + 1] - off);
  const int64_t nv = T.n_v;
  const LT* L = Lvv + (J0 + jj) * nv * nv;
  LT* out = Lp + K.lp_off[q] + jj * (int64_t)d * d;
  for (int r = threadIdx.y; r < d; r += blockDim.y) {
    const int64_t row = (int64_t)T.dom_idx[off + r] * nv;
    for (int c = threadIdx.x; c < d; c += blockDim.x) out[(int64_t)r * d + c] = L[row + T.dom_idx[off + c]];
  }
}

// Cw[k][P d + c] = C_P[k][c]: the C blocks of equal size d side by side
__global__ void concat_cols_kernel(const double* __restrict__ C, int n_pairs, int d, double* __restrict__ Cw) {
  const int64_t d2 = (int64_t)d * d, n = (int64_t)n_pairs * d2, ld = (int64_t)n_pairs * d;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t P = e / d2, rem = e - P * d2;
    const int64_t k = rem / d, c = rem - k * d;
    Cw[k * ld + P * d + c] = C[e];
  }
}

// y = A x for a square row-major n x n matrix: one wave per row
__global__ void gemv_kernel(const double* __restrict__ A, int64_t n, const double* __restrict__ x,
                            double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < n;
       r += (int64_t)gridDim.x * (blockDim.x >> 6)) {
    const double* a = A + r * n;
    double acc = 0.0;
    for (int64_t c = lane; c < n; c += 64) acc += a[c] * x[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) y[r] = acc;
  }
}

__global__ void fill_kernel(double* p, int64_t n, double v) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = v;
}

// ---- explicit one-electron integrals ---------------------------------------
// F_pq = h_pq + sum_k [2 (pq|kk) - (pk|kq)] (closed-shell Fock matrix), n_orb x
// n_orb, from the packed h (element (p,q), p >= q, at p(p+1)/2 + q) and the
// assembled blocks: O4[a][b][c][d] = (ab|cd) (internal), O3[k][p][q][b] =
// (kp|bq), Kov[(i,a)][(j,b)] = (ia|jb), Jvv[ij pair][a][b] = (ij|ab).
__global__ void fock_kernel(const double* __restrict__ hp, const double* __restrict__ O4,
                            const double* __restrict__ O3, const double* __restrict__ Kov,
                            const double* __restrict__ Jvv, int n_int, int n_v, double* __restrict__ F) {
  const int64_t no = n_int, nv = n_v, n_orb = no + nv, niv = no * nv;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_orb * n_orb;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = e / n_orb, q = e - p * n_orb;
    if (p < q) continue;  // lower triangle, mirrored below
    double g = 0.0;
    if (p < no) {  // (internal, internal)
      for (int64_t k = 0; k < no; ++k)
        g += 2.0 * O4[((p * no + q) * no + k) * no + k] - O4[((p * no + k) * no + k) * no + q];
    } else if (q < no) {  // (virtual a = p - no, internal i = q): 2 (kk|ai) - (ik|ak)
      const int64_t a = p - no, i = q;
      for (int64_t k = 0; k < no; ++k)
        g += 2.0 * O3[((k * no + k) * no + i) * nv + a] - O3[((i * no + k) * no + k) * nv + a];
    } else {  // (virtual, virtual): 2 (kk|ab) - (ka|kb)
      const int64_t a = p - no, b = q - no;
      for (int64_t k = 0; k < no; ++k)
        g += 2.0 * Jvv[((k * (k + 1) / 2 + k) * nv + a) * nv + b] - Kov[(k * nv + a) * niv + k * nv + b];
    }
    const double f = hp[p * (p + 1) / 2 + q] + g;
    F[p * n_orb + q] = f;
    F[q * n_orb + p] = f;
  }
}

// O3[k][k][q][b] -= f_{q,b}: the occupied-virtual Fock block joins the
// 3-internal singles coupling (ki|bj) - delta_ki f_jb in both directions
__global__ void fold_o3_kernel(double* __restrict__ O3, const double* __restrict__ F, int n_int, int n_v) {
  const int64_t no = n_int, nv = n_v, n_orb = no + nv;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < no * no * nv;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e / (no * nv), r = e - k * no * nv, q = r / nv, b = r - q * nv;
    O3[((k * no + k) * no + q) * nv + b] -= F[q * n_orb + no + b];
  }
}

// Mss[(i,a)][(j,b)] += delta_ij f'_ab - delta_ab f'_ij (off-diagonal Fock
// elements; the diagonal ones act through the orbital energies)
__global__ void fold_mss_kernel(double* __restrict__ Mss, const double* __restrict__ F, int n_int, int n_v) {
  const int64_t no = n_int, nv = n_v, n_orb = no + nv, niv = no * nv;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < no * nv * nv;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / (nv * nv), r = e - i * nv * nv, a = r / nv, b = r - a * nv;
    if (a != b) Mss[(i * nv + a) * niv + i * nv + b] += F[(no + a) * n_orb + no + b];
  }
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nv * no * no;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = e / (no * no), r = e - a * no * no, i = r / no, j = r - i * no;
    if (i != j) Mss[(i * nv + a) * niv + j * nv + a] -= F[i * n_orb + j];
  }
}

// Reference <-> singles through f_ia: g[i][a] += c0 f_ia and
// sigma_0 += 2 sum_ia f_ia s_i^a (one workgroup; after the singles reduction
// and the reference dot of the same build)
__global__ void fov_terms_kernel(const double* x, const IoSlot* io, const double* __restrict__ s,
                                 const double* __restrict__ Fov, int64_t niv, double* __restrict__ g,
                                 double* __restrict__ sig0) {
  if (io) x = io->x;
  const double c0 = x[0];
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < niv; e += blockDim.x) {
    const double f = Fov[e];
    g[e] += c0 * f;
    acc += f * s[e];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *sig0 += 2.0 * red[0];
}

inline unsigned gx(int64_t work) { return (unsigned)std::min<int64_t>(std::max<int64_t>(cdiv(work, 256), 1), 4096); }

// out[c][r] = in[r][c] (rows x cols, small matrices)
__global__ void transpose_small_kernel(const double* __restrict__ in, int rows, int cols, double* __restrict__ out) {
  const int64_t n = (int64_t)rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / cols), c = (int)(e - (int64_t)r * cols);
    out[(int64_t)c * rows + r] = in[e];
  }
}

}  // namespace

void launch_transpose_small(const double* in, int rows, int cols, double* out, hipStream_t st) {
  if ((int64_t)rows * cols <= 0) return;
  hipLaunchKernelGGL(transpose_small_kernel, dim3(gx((int64_t)rows * cols)), dim3(256), 0, st, in, rows, cols, out);
  TCI_HIP(hipGetLastError());
}

void launch_unpack(const PairTab& T, int dmax, const double* x, double* C, double* Ct, double* s,
                   hipStream_t st, const IoSlot* io) {
  const int64_t w = std::max<int64_t>((int64_t)dmax * dmax, (int64_t)T.n_int * T.n_v);
  hipLaunchKernelGGL(unpack_kernel, dim3(std::min<unsigned>(gx(w), 256), T.n_pairs + 1), dim3(256), 0, st, T, x, C,
                     Ct, s, io);
  TCI_HIP(hipGetLastError());
}

void launch_rterms(const PairTab& T, int dmax, const double* x, const double* s, const double* Kov,
                   const double* O3, double* R, hipStream_t st, const IoSlot* io) {
  if (T.n_pairs == 0) return;
  hipLaunchKernelGGL(rterms_kernel, dim3((unsigned)cdiv(dmax, RT_ROWS), T.n_pairs), dim3(RT_THREADS), 0, st, T, x, s,
                     Kov, O3, R, io);
  TCI_HIP(hipGetLastError());
}

void launch_hole(const PairTab& T, const HoleTab& H, int dmax, const double* C, const double* O4, double* R,
                 hipStream_t st) {
  if (T.n_pairs == 0) return;
  hipLaunchKernelGGL(hole_kernel, dim3((unsigned)cdiv(dmax, 16), T.n_pairs), dim3(256), 0, st, T, H, C, O4, R);
  TCI_HIP(hipGetLastError());
}

int ref_dot_slices() { return kRefSlices; }

void launch_ref_dot(const PairTab& T, const double* Kov, const double* Ct, double* pair_part, double* sig0,
                    hipStream_t st) {
  if (T.n_pairs > 0) {
    hipLaunchKernelGGL(ref_dot_kernel, dim3(T.n_pairs, kRefSlices), dim3(256), 0, st, T, Kov, Ct, pair_part);
    TCI_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, st, pair_part, T.n_pairs * kRefSlices, sig0);
  TCI_HIP(hipGetLastError());
}

void launch_s3int_singles(const PairTab& T, int dmax, const int64_t* gpart_off, const double* Ct, const double* O3,
                          double* gpart, hipStream_t st) {
  if (T.n_pairs == 0) return;
  hipLaunchKernelGGL(s3int_singles_kernel, dim3(T.n_pairs, 2, (unsigned)cdiv(dmax, S3R)), dim3(256), 0, st, T,
                     gpart_off, Ct, O3, gpart);
  TCI_HIP(hipGetLastError());
}

void launch_singles_reduce(const PairTab& T, const InvTab& I, const int64_t* gpart_off, const double* gpart,
                           const double* gss, const double* gW, int n_split, double* g, hipStream_t st) {
  hipLaunchKernelGGL(singles_reduce_kernel, dim3((unsigned)T.n_v), dim3(64 * kRedSlices), 0, st, T, I, gpart_off,
                     gpart, gss, gW, n_split, g);
  TCI_HIP(hipGetLastError());
}

void launch_pack(const PairTab& T, int dmax, const double* x, const double* R, const double* g, const double* sig0,
                 double e_ref, int flags, double* sigma, hipStream_t st, const IoSlot* io) {
  const int64_t w = std::max<int64_t>((int64_t)dmax * dmax, 1 + (int64_t)T.n_int * T.n_v);
  hipLaunchKernelGGL(pack_kernel, dim3(std::min<unsigned>(gx(w), 256), T.n_pairs + 1), dim3(256), 0, st, T, x, R, g,
                     sig0, e_ref, flags, sigma, io);
  TCI_HIP(hipGetLastError());
}

void launch_set_io(IoSlot* io, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(set_io_kernel, dim3(1), dim3(64), 0, st, io, x, y);
  TCI_HIP(hipGetLastError());
}

void launch_diag(const PairTab& T, int dmax, const double* Kov, const double* Jvv, const double* O4,
                 const double* vv_diag, const double* vv_exch, double* diag, hipStream_t st) {
  const int64_t w = std::max<int64_t>((int64_t)dmax * dmax, 1 + (int64_t)T.n_int * T.n_v);
  hipLaunchKernelGGL(diag_kernel, dim3(std::min<unsigned>(gx(w), 256), T.n_pairs + 1), dim3(256), 0, st, T, Kov, Jvv,
                     O4, vv_diag, vv_exch, diag);
  TCI_HIP(hipGetLastError());
}

void launch_gather_Lvv(const double* Lpm, int64_t n_chol, int n_int, int n_v, void* Lvv, bool f32, hipStream_t st) {
  dim3 grid(n_v, (unsigned)cdiv(n_chol, 32));
  if (f32) hipLaunchKernelGGL(gather_lvv_kernel<float>, grid, dim3(256), 0, st, Lpm, n_chol, n_int, n_v, (float*)Lvv);
  else hipLaunchKernelGGL(gather_lvv_kernel<double>, grid, dim3(256), 0, st, Lpm, n_chol, n_int, n_v, (double*)Lvv);
  TCI_HIP(hipGetLastError());
}

void launch_gather_Lov(const double* Lpm, int64_t n_chol, int n_int, int n_v, double* Lov, hipStream_t st) {
  dim3 grid(n_v, (unsigned)cdiv(n_chol, 32));
  hipLaunchKernelGGL(gather_lov_kernel, grid, dim3(256), 0, st, Lpm, n_chol, n_int, n_v, Lov);
  TCI_HIP(hipGetLastError());
}

void launch_vv_exch(const void* Lvv, bool f32, int64_t n_chol, int n_v, double* out, hipStream_t st) {
  const unsigned g = gx((int64_t)n_v * n_v);
  if (f32) hipLaunchKernelGGL(vv_exch_kernel<float>, dim3(g), dim3(256), 0, st, (const float*)Lvv, n_chol, n_v, out);
  else hipLaunchKernelGGL(vv_exch_kernel<double>, dim3(g), dim3(256), 0, st, (const double*)Lvv, n_chol, n_v, out);
  TCI_HIP(hipGetLastError());
}

void launch_build_mss(const double* Kov, const double* Jvv, int n_int, int n_v, double* Mss, hipStream_t st) {
  const int64_t n = (int64_t)n_int * n_v * n_int * n_v;
  hipLaunchKernelGGL(build_mss_kernel, dim3(gx(n)), dim3(256), 0, st, Kov, Jvv, n_int, n_v, Mss);
  TCI_HIP(hipGetLastError());
}

void launch_pack_L(const PairTab& T, const PackTab& K, const void* Lvv, bool f32, int64_t J0, int64_t len, void* Lp,
                   hipStream_t st) {
  if (K.n_local == 0 || len == 0) return;
  dim3 grid(K.n_local, (unsigned)len), block(64, 4);
  if (f32) hipLaunchKernelGGL(pack_L_kernel<float>, grid, block, 0, st, T, K, (const float*)Lvv, J0, len, (float*)Lp);
  else hipLaunchKernelGGL(pack_L_kernel<double>, grid, block, 0, st, T, K, (const double*)Lvv, J0, len, (double*)Lp);
  TCI_HIP(hipGetLastError());
}

void launch_concat_cols(const double* C, int n_pairs, int d, double* Cw, hipStream_t st) {
  const int64_t n = (int64_t)n_pairs * d * d;
  if (n <= 0) return;
  hipLaunchKernelGGL(concat_cols_kernel, dim3(gx(n)), dim3(256), 0, st, C, n_pairs, d, Cw);
  TCI_HIP(hipGetLastError());
}

void launch_gemv(const double* A, int64_t n, const double* x, double* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gemv_kernel, dim3((unsigned)std::min<int64_t>(cdiv(n, 4), 8192)), dim3(256), 0, st, A, n, x, y);
  TCI_HIP(hipGetLastError());
}

void launch_fill(double* p, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(gx(n)), dim3(256), 0, st, p, n, v);
  TCI_HIP(hipGetLastError());
}

void launch_fock(const double* hp, const double* O4, const double* O3, const double* Kov, const double* Jvv,
                 int n_int, int n_v, double* F, hipStream_t st) {
  const int64_t n = (int64_t)(n_int + n_v) * (n_int + n_v);
  hipLaunchKernelGGL(fock_kernel, dim3(gx(n)), dim3(256), 0, st, hp, O4, O3, Kov, Jvv, n_int, n_v, F);
  TCI_HIP(hipGetLastError());
}

void launch_fold_fock(const double* F, int n_int, int n_v, double* O3, double* Mss, hipStream_t st) {
  const int64_t n = (int64_t)n_int * n_int * n_v + (int64_t)n_int * n_v * n_v;
  hipLaunchKernelGGL(fold_o3_kernel, dim3(gx(n)), dim3(256), 0, st, O3, F, n_int, n_v);
  TCI_HIP(hipGetLastError());
  hipLaunchKernelGGL(fold_mss_kernel, dim3(gx(n)), dim3(256), 0, st, Mss, F, n_int, n_v);
  TCI_HIP(hipGetLastError());
}

void launch_fov_terms(const double* x, const double* s, const double* Fov, int64_t niv, double* g, double* sig0,
                      hipStream_t st, const IoSlot* io) {
  hipLaunchKernelGGL(fov_terms_kernel, dim3(1), dim3(256), 0, st, x, io, s, Fov, niv, g, sig0);
  TCI_HIP(hipGetLastError());
}

}  // namespace tci
