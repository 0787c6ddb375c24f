// Pair-pair terms as dense products over ordered internal pairs, with every
// pair block embedded in the full virtual space.
//
// N = n_int n_v.  An ordered internal pair (p, q) names an n_v x n_v block;
// X^{pq} is the walk Q = {p, q}'s dense block X_Q (p >= q) or X_Q^T (p < q),
// scattered to the virtuals of its domain D_Q (zero elsewhere, and zero for
// pairs without doubles).  Per sigma:
//   CR [(p,r)][(q,c)] = C^{pq}[r][c],   CtR likewise from Ct = 2C - C^T,
//   CC [(p,r)][(q,c)] = C^{qp}[r][c]    (CR, CtR and CC are symmetric),
//   Yt [(i,a)][J] = sum_c L^J_ac s_i^c,
// and once per plan
//   Kov[(k,c)][(j,b)] = (kc|jb),  Jbig[(k,c)][(j,b)] = (kj|cb)  (both symmetric),
//   Lov[J][(j,b)] = L^J_jb.
// Then
//   Z  = CtR Kov - CR Jbig + Yt Lov,   Z3 = Jbig CC,   W[i][J][b] = (Lov CtR)[J][(i,b)]
// and since Kov = Lov^T Lov, Z = U Lov - CR Jbig with U = Yt + CtR Lov^T
// (N x n_chol): 2 N^2 n_chol flops for U and as many for U Lov, in place of
// 2 N^3 for CtR Kov (FP64 and bf16x3; fp16x3 keeps CtR Kov).
// hold for every ordered (i, j): Z(i,j) = sum_k [Ct^{ik} K^{kj} - C^{ik} J^{kj}] + X^{ij},
// Z3(j,i) = sum_k J^{kj} C^{ik}, so the ring and 3-external terms of the walk
// P = (i, j) are, at positions x, y of D_P (a = D_P[x], b = D_P[y]),
//   R_P[x][y] += Z(i,j)[a][b] - Z3(j,i)[a][b] + Z(j,i)[b][a] - Z3(i,j)[b][a].
// A restriction of an embedded product equals the product over the domain
// intersections, which is what the per-walk gather programs compute.  The
// 4-internal term sum_{Q=(k,l)} [(ki|lj) C_Q + (li|kj) C_Q^T] is one product of
// an n_pairs x 2 n_pairs coefficient matrix with the embedded [C; C^T]
// (2 n_pairs x n_v^2 rows), gathered back the same way.  All products run on
// the dense FP64 kernel, or for the split precisions on the split-row kernel
// (Z, Z3 and W with k padded per n_v block to a multiple of 8; the 4-internal
// product over k = walk, its B operand the embedded blocks written directly
// as split rows over k).  The plan
// takes this path when the products cost at most a few times the gathered
// per-walk work (full or wide domains).
#include "plan.hpp"

#include <cstdlib>
#include <initializer_list>
#include <string>

namespace tci {

namespace {

struct EmbedTab {
  const int32_t* pairid;  // [p n_int + q] -> strong walk of the unordered pair {p, q}, or -1
  const int32_t* pos;     // [Q n_v + a]   -> position of virtual a in D_Q, or -1
};

__device__ __forceinline__ int64_t rec_of(int p, int q) {
  const int hi = max(p, q), lo = min(p, q);
  return (int64_t)hi * (hi + 1) / 2 + lo;
}

// Jbig[(p,r)][(q,s)] = (pq|rs) from Jvv[rec(p,q)][r][s]
__global__ void jbig_kernel(const double* __restrict__ Jvv, int n_int, int n_v, double* __restrict__ Jbig) {
  const int64_t N = (int64_t)n_int * n_v;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * N; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e / N, col = e - row * N;
    const int p = (int)(row / n_v), r = (int)(row - (int64_t)p * n_v);
    const int q = (int)(col / n_v), s = (int)(col - (int64_t)q * n_v);
    Jbig[e] = Jvv[(rec_of(p, q) * n_v + r) * n_v + s];
  }
}

// Output block (p, q) = sign X^{pq} (cc = false) or sign X^{qp} = sign (X^{pq})^T
// (cc = true), embedded.  32 x 32 tiles through LDS so both orientations read
// along rows of the source block.  FP64: out[(p n_v + r) ldo + q n_v + c];
// split rows: out16[(p n_v + r) ldo + xidx(q vp + c)], vp = n_v rounded up to 8.
template <bool SPLIT>
__global__ void relayout_kernel(PairTab T, EmbedTab E, const double* __restrict__ src, bool cc, double sign,
                                void* __restrict__ out, int64_t ldo, const float* scl, int f16) {
  __shared__ double tile[32][33];
  const int nv = T.n_v, nt = (nv + 31) >> 5;
  const int q = blockIdx.x / nt, ct = blockIdx.x - q * nt;
  const int p = blockIdx.y / nt, rt = blockIdx.y - p * nt;
  const int Q = E.pairid[p * T.n_int + q];
  const bool direct = (p >= q) != cc;
  const int r0 = rt * 32, c0 = ct * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (Q >= 0) {
    const double* blk = src + T.cd_off[Q];
    const int d = (int)(T.dom_off[Q + 1] - T.dom_off[Q]);
    const int32_t* ps = E.pos + (int64_t)Q * nv;
    for (int yy = ty; yy < 32; yy += 8) {
      if (direct) {  // element (r, c) = blk[pos r][pos c]
        const int r = r0 + yy, c = c0 + tx;
        const int pr = r < nv ? ps[r] : -1, pc = c < nv ? ps[c] : -1;
        tile[yy][tx] = (pr >= 0 && pc >= 0) ? blk[(int64_t)pr * d + pc] : 0.0;
      } else {  // element (r, c) = blk[pos c][pos r]: walk along row pos c
        const int c = c0 + yy, r = r0 + tx;
        const int pr = r < nv ? ps[r] : -1, pc = c < nv ? ps[c] : -1;
        tile[tx][yy] = (pr >= 0 && pc >= 0) ? blk[(int64_t)pc * d + pr] : 0.0;
      }
    }
  } else {
    for (int yy = ty; yy < 32; yy += 8) tile[yy][tx] = 0.0;
  }
  __syncthreads();
  if (!SPLIT) {
    double* o = static_cast<double*>(out);
    for (int yy = ty; yy < 32; yy += 8) {
      const int r = r0 + yy, c = c0 + tx;
      if (r < nv && c < nv) o[((int64_t)p * nv + r) * ldo + (int64_t)q * nv + c] = sign * tile[yy][tx];
    }
  } else if (threadIdx.x < 128) {
    const int vp = (nv + 7) & ~7;
    const int yy = threadIdx.x >> 2, c = c0 + 8 * (threadIdx.x & 3);
    const int r = r0 + yy;
    if (r < nv && c < vp) {
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = (c + t < nv) ? (float)(sign * tile[yy][c - c0 + t]) : 0.f;
      put_split8(static_cast<uint16_t*>(out) + ((int64_t)p * nv + r) * ldo + xidx((int64_t)q * vp + c), v,
                 scale_at(scl), f16);
    }
  }
}

// split rows of an FP64 matrix whose columns are nb blocks of d, each block's
// k padded to dp: out16[row ldo + xidx(q dp + c)] = src[row ld + q d + c]
__global__ void split_blocked_kernel(const double* __restrict__ src, int64_t rows, int64_t ld, int nb, int d,
                                     uint16_t* __restrict__ out, int64_t ldo, const float* scl, int f16) {
  const float sc = scale_at(scl);
  const int dp = (d + 7) & ~7, nch = dp >> 3;
  const int64_t per_row = (int64_t)nb * nch;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < rows * per_row;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e / per_row;
    const int rem = (int)(e - row * per_row);
    const int q = rem / nch, c = 8 * (rem - q * nch);
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = (c + t < d) ? (float)src[row * ld + (int64_t)q * d + c + t] : 0.f;
    put_split8(out + row * ldo + xidx((int64_t)q * dp + c), v, sc, f16);
  }
}

// split rows of a transposed FP64 matrix: out16[c ldo + xidx(r)] = src[r ld + c],
// r < nr, c < nc; k (= r) padded with zeros to a multiple of 8
__global__ void transpose_split_kernel(const double* __restrict__ src, int64_t nr, int64_t nc, int64_t ld,
                                       uint16_t* __restrict__ out, int64_t ldo, const float* scl, int f16) {
  __shared__ float tile[32][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t r = r0 + yy, c = c0 + tx;
    tile[yy][tx] = (r < nr && c < nc) ? (float)src[r * ld + c] : 0.f;
  }
  __syncthreads();
  if (threadIdx.x < 128) {
    const int cy = threadIdx.x >> 2, ch = threadIdx.x & 3;
    const int64_t c = c0 + cy, k = r0 + 8 * ch;
    if (c < nc && k < ((nr + 7) & ~(int64_t)7)) {
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = tile[8 * ch + t][cy];
      put_split8(out + c * ldo + xidx(k), v, scale_at(scl), f16);
    }
  }
}

// dst[c][r] = src[r][c] (rows x cols FP64, 32 x 32 tiles through LDS)
__global__ void transpose_kernel(const double* __restrict__ src, int64_t rows, int64_t cols, double* __restrict__ dst) {
  __shared__ double tile[32][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t r = r0 + yy, c = c0 + tx;
    tile[yy][tx] = (r < rows && c < cols) ? src[r * cols + c] : 0.0;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t c = c0 + yy, r = r0 + tx;
    if (r < rows && c < cols) dst[c * rows + r] = tile[tx][yy];
  }
}

// Yt[(i,a)][J] = Y[(J,a)][i]  (32 x 32 tiles of (J, a) per i)
__global__ void y_transpose_kernel(const double* __restrict__ Y, int n_int, int n_v, int64_t n_chol,
                                   double* __restrict__ Yt) {
  __shared__ double tile[32][33];
  const int i = blockIdx.z;
  const int64_t J0 = (int64_t)blockIdx.x * 32;
  const int a0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t J = J0 + yy;
    const int a = a0 + tx;
    tile[yy][tx] = (J < n_chol && a < n_v) ? Y[(J * n_v + a) * n_int + i] : 0.0;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int a = a0 + yy;
    const int64_t J = J0 + tx;
    if (J < n_chol && a < n_v) Yt[((int64_t)i * n_v + a) * n_chol + J] = tile[tx][yy];
  }
}

// Cemb[Q][a][b] = C_Q[pos a][pos b], CTemb[Q][b][a] = the same value (0 outside D_Q)
__global__ void embed_blocks_kernel(PairTab T, EmbedTab E, const double* __restrict__ C, double* __restrict__ Cemb,
                                    double* __restrict__ CTemb) {
  __shared__ double tile[32][33];
  const int nv = T.n_v, nt = (nv + 31) >> 5;
  const int Q = blockIdx.y;
  const int a0 = (blockIdx.x / nt) * 32, b0 = (blockIdx.x % nt) * 32;
  const int d = (int)(T.dom_off[Q + 1] - T.dom_off[Q]);
  const double* blk = C + T.cd_off[Q];
  const int32_t* ps = E.pos + (int64_t)Q * nv;
  const int64_t base = (int64_t)Q * nv * nv;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int a = a0 + yy, b = b0 + tx;
    const int pa = a < nv ? ps[a] : -1, pb = b < nv ? ps[b] : -1;
    const double v = (pa >= 0 && pb >= 0) ? blk[(int64_t)pa * d + pb] : 0.0;
    tile[yy][tx] = v;
    if (a < nv && b < nv) Cemb[base + (int64_t)a * nv + b] = v;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int b = b0 + yy, a = a0 + tx;
    if (a < nv && b < nv) CTemb[base + (int64_t)b * nv + a] = tile[tx][yy];
  }
}

// R_P[x][y] += Remb[P][a][b] + Z(i,j)[a][b] - Z3(j,i)[a][b] + Z(j,i)[b][a] - Z3(i,j)[b][a],
// a = D_P[x], b = D_P[y]
__global__ void ring_combine_kernel(PairTab T, const double* __restrict__ Z, const double* __restrict__ Z3,
                                    const double* __restrict__ Remb, double* __restrict__ R) {
  __shared__ double tile[32][33];
  const int nv = T.n_v;
  const int64_t N = (int64_t)T.n_int * nv;
  const int P = blockIdx.y;
  const int d = (int)(T.dom_off[P + 1] - T.dom_off[P]);
  const int nt = (d + 31) >> 5;
  if ((int)blockIdx.x >= nt * nt) return;
  const int i = T.pairs[2 * P], j = T.pairs[2 * P + 1];
  const int32_t* D = T.dom_idx + T.dom_off[P];
  const int x0 = (blockIdx.x / nt) * 32, y0 = (blockIdx.x % nt) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {  // transposed part: rows b, columns a
    const int y = y0 + yy, x = x0 + tx;
    double v = 0.0;
    if (x < d && y < d) {
      const int a = D[x], b = D[y];
      v = Z[((int64_t)j * nv + b) * N + (int64_t)i * nv + a] - Z3[((int64_t)i * nv + b) * N + (int64_t)j * nv + a];
    }
    tile[yy][tx] = v;
  }
  __syncthreads();
  double* Rb = R + T.cd_off[P];
  const double* Hp = Remb + (int64_t)P * nv * nv;
  for (int yy = ty; yy < 32; yy += 8) {
    const int x = x0 + yy, y = y0 + tx;
    if (x < d && y < d) {
      const int a = D[x], b = D[y];
      Rb[(int64_t)x * d + y] += Hp[(int64_t)a * nv + b] + Z[((int64_t)i * nv + a) * N + (int64_t)j * nv + b] -
                                Z3[((int64_t)j * nv + a) * N + (int64_t)i * nv + b] + tile[tx][yy];
    }
  }
}

inline unsigned grid_for(int64_t work) {
  return (unsigned)std::min<int64_t>(std::max<int64_t>(cdiv(work, 256), 1), 16384);
}

Op dop(const void* base, int64_t rs) {
  Op o;
  o.base = base;
  o.rs = rs;
  o.cs = 1;
  o.rmap = nullptr;
  o.cmap = nullptr;
  o.jstride = 0;
  o.f32 = 0;
  return o;
}

Seg dseg(const Op& A, const Op& B, int K, double alpha) {
  Seg s;
  s.A = A;
  s.B = B;
  s.K = K;
  s.alpha = alpha;
  return s;
}

EmbedTab etab(const tci_plan& P) { return EmbedTab{P.pairid_d.as<int32_t>(), P.pos.as<int32_t>()}; }

}  // namespace

// Embedded products when they cost at most 8x the per-walk gathered ring
// flops (f_gather) and their FP64 work matrices fit in 48 GB; TCI_DENSE_PAIRS=0
// turns them off, =1 forces them within the memory bound.
bool dense_pairs_choose(const tci_plan& P, double f_gather) {
  const double N = (double)P.n_int * P.n_v;
  if (P.n_pairs == 0 || N >= 46340.0) return false;  // N^2 must index with int32 tiles
  const double bytes = 8.0 * (6.0 * N * N + 3.0 * P.n_pairs * (double)P.n_v * P.n_v);
  if (bytes > 48.0e9) return false;
  const char* env = std::getenv("TCI_DENSE_PAIRS");
  if (env && env[0] == '0') return false;
  if (env && env[0] == '1') return true;
  const double f_dense = 6.0 * N * N * N + 4.0 * N * N * (double)P.n_chol;
  return f_dense <= 8.0 * f_gather;
}

// FP64 tables: Jbig, work matrices, coefficient matrix and the dense programs;
// o4 is the host copy of (ki|lj) = o4[((k n + i) n + l) n + j]
// Split rows over k = walk of the embedded blocks, for the 4-internal product
// in split precision: row (a, b) of n_v^2 holds C_Q[pos a][pos b] at k = Q
// and C_Q[pos b][pos a] at k = n_pairs + Q (the Cemb and CTemb entries; 0
// outside D_Q), zero-padded to a multiple of 8.  One thread per row and
// 8-wide k chunk (blockIdx.y), so that a block's reads of each C_Q row are
// contiguous.
__global__ void embed_split_kernel(PairTab T, EmbedTab E, const double* __restrict__ C, uint16_t* __restrict__ out,
                                   int64_t ldo, const float* scl, int f16) {
  const int nv = T.n_v, np = T.n_pairs;
  const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= (int64_t)nv * nv) return;
  const int a = (int)(row / nv), b = (int)(row - (int64_t)a * nv);
  const int k0 = 8 * (int)blockIdx.y;
  float v[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int k = k0 + t;
    double x = 0.0;
    if (k < 2 * np) {
      const int Q = k < np ? k : k - np;
      const int32_t* ps = E.pos + (int64_t)Q * nv;
      const int pa = ps[a], pb = ps[b];
      if (pa >= 0 && pb >= 0) {
        const int d = (int)(T.dom_off[Q + 1] - T.dom_off[Q]);
        const double* blk = C + T.cd_off[Q];
        x = k < np ? blk[(int64_t)pa * d + pb] : blk[(int64_t)pb * d + pa];
      }
    }
    v[t] = (float)x;
  }
  put_split8(out + row * ldo + xidx(k0), v, scale_at(scl), f16);
}

void dense_pairs_build(tci_plan& P, const std::vector<double>& o4, hipStream_t st) {
  const int n = P.n_int, nv = P.n_v, np = P.n_pairs;
  const int64_t N = (int64_t)n * nv, v2 = (int64_t)nv * nv;
  std::vector<int32_t> pid((size_t)n * n);
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) pid[(size_t)p * n + q] = P.pairid[(size_t)p * n + q];
  upload(P.pairid_d, pid, st);
  P.Jbig.alloc((size_t)N * N * 8);
  hipLaunchKernelGGL(jbig_kernel, dim3(grid_for(N * N)), dim3(256), 0, st, P.Jvv.as<double>(), n, nv,
                     P.Jbig.as<double>());
  TCI_HIP(hipGetLastError());
  for (DevBuf* b : {&P.CRb, &P.CtRb, &P.CCb, &P.Zb, &P.Z3b}) b->alloc((size_t)N * N * 8);
  for (DevBuf* b : {&P.Cemb, &P.CTemb, &P.Remb}) b->alloc((size_t)np * v2 * 8);
  P.Yt.alloc((size_t)N * P.n_chol * 8);
  std::vector<double> hA((size_t)np * 2 * np, 0.0);
  for (int X = 0; X < np; ++X) {
    const int i = P.h_pairs[2 * X], j = P.h_pairs[2 * X + 1];
    for (int Q = 0; Q < np; ++Q) {
      const int k = P.h_pairs[2 * Q], l = P.h_pairs[2 * Q + 1];
      hA[(size_t)X * 2 * np + Q] = o4[(((size_t)k * n + i) * n + l) * n + j];                          // (ki|lj)
      hA[(size_t)X * 2 * np + np + Q] = k != l ? o4[(((size_t)l * n + i) * n + k) * n + j] : 0.0;  // (li|kj)
    }
  }
  upload(P.holeA, hA, st);
  const int iN = (int)N, K = (int)P.n_chol;
  {
    const char* e = std::getenv("TCI_Z_LOWRANK");
    P.z_lowrank = !(e && e[0] == '0');
  }
  P.ringZ.clear();
  P.ringZ.dense = true;
  P.ringZ.tile = 128;
  P.progU.clear();
  if (P.z_lowrank) {
    // Kov = Lov^T Lov, so C~R Kov + Yt Lov = U Lov with U = Yt + C~R Lov^T
    // (accumulated into Yt after its transpose in every sigma)
    P.LovT.alloc((size_t)N * P.n_chol * 8);
    hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)cdiv(N, 32), (unsigned)cdiv(P.n_chol, 32)), dim3(256), 0, st,
                       P.Lov.as<double>(), P.n_chol, N, P.LovT.as<double>());
    TCI_HIP(hipGetLastError());
    P.progU.dense = true;
    P.progU.tile = 128;
    P.progU.add(P.Yt.as<double>(), P.n_chol, iN, K, {dseg(dop(P.CtRb.p, N), dop(P.LovT.p, P.n_chol), iN, 1.0)}, true);
    P.progU.upload_tables(st);
    P.ringZ.add(P.Zb.as<double>(), N, iN, iN,
                {dseg(dop(P.Yt.p, P.n_chol), dop(P.Lov.p, N), K, 1.0), dseg(dop(P.CRb.p, N), dop(P.Jbig.p, N), iN, -1.0)},
                false);
  } else {
    P.ringZ.add(P.Zb.as<double>(), N, iN, iN,
                {dseg(dop(P.CtRb.p, N), dop(P.Kov.p, N), iN, 1.0), dseg(dop(P.CRb.p, N), dop(P.Jbig.p, N), iN, -1.0),
                 dseg(dop(P.Yt.p, P.n_chol), dop(P.Lov.p, N), K, 1.0)},
                false);
  }
  P.ringZ.upload_tables(st);
  P.ringZ3.clear();
  P.ringZ3.dense = true;
  P.ringZ3.tile = 128;
  P.ringZ3.add(P.Z3b.as<double>(), N, iN, iN, {dseg(dop(P.Jbig.p, N), dop(P.CCb.p, N), iN, 1.0)}, false);
  P.ringZ3.upload_tables(st);
  // W[i][J][b] = (Lov CtR)[J][(i,b)]: column block i of n_v goes to W + i n_chol n_v
  Blocking bw;
  bw.cdiv = nv;
  bw.cstride = P.n_chol * nv;
  P.progWd.clear();
  P.progWd.dense = true;
  P.progWd.tile = 128;
  P.progWd.add(P.W.as<double>(), nv, K, iN, {dseg(dop(P.Lov.p, N), dop(P.CtRb.p, N), iN, 1.0)}, false, false, 0, bw);
  P.progWd.upload_tables(st);
  P.holeG.clear();
  P.holeG.dense = true;
  P.holeG.tile = 128;
  P.holeG.add(P.Remb.as<double>(), v2, np, (int)v2,
              {dseg(dop(P.holeA.p, 2 * np), dop(P.Cemb.p, v2), np, 1.0),
               dseg(dop(P.holeA.as<double>() + np, 2 * np), dop(P.CTemb.p, v2), np, 1.0)},
              false);
  P.holeG.upload_tables(st);
}

// split copies of Kov, Jbig, Lov, Lov^T and the split programs for Z, Z3 and W
void dense_pairs_build_split(tci_plan& P, hipStream_t st, int cfg_z, int cfg_uw) {
  tci_plan::MixedLadder& M = P.mx;
  const int n = P.n_int, nv = P.n_v;
  const int64_t N = (int64_t)n * nv, ldx = 2 * (int64_t)n * xpad(nv), ldj = 2 * xpad(P.n_chol);
  // the low-rank Z in bf16x3 only: in fp16x3 the scale of U would have to be
  // known before U exists (C~R is split with the scale of C, Yt and U together)
  M.z_lowrank = P.z_lowrank && M.prec == TCI_SIGMA_BF16X3;
  for (DevBuf* b : {&M.Kx, &M.Jx, &M.CRx, &M.CtRx, &M.CCx}) {
    if (b == &M.Kx && M.z_lowrank) {
      b->reset();
      continue;
    }
    b->alloc((size_t)N * ldx * 2);
    M.bytes += (int64_t)b->bytes;
  }
  M.Lovx.alloc((size_t)P.n_chol * ldx * 2);
  M.LovTx.alloc((size_t)N * ldj * 2);
  M.Ytx.alloc((size_t)N * ldj * 2);
  M.bytes += (int64_t)(M.Lovx.bytes + M.LovTx.bytes + M.Ytx.bytes);
  const int64_t work = N * n * (xpad(nv) / 8);
  const float* s0 = M.scales.as<float>();  // constants' scale
  const int f16 = M.prec == TCI_SIGMA_FP16X3 ? 1 : 0;
  if (!M.z_lowrank)
    hipLaunchKernelGGL(split_blocked_kernel, dim3(grid_for(work)), dim3(256), 0, st, P.Kov.as<double>(), N, N, n, nv,
                       M.Kx.as<uint16_t>(), ldx, s0, f16);
  hipLaunchKernelGGL(split_blocked_kernel, dim3(grid_for(work)), dim3(256), 0, st, P.Jbig.as<double>(), N, N, n, nv,
                     M.Jx.as<uint16_t>(), ldx, s0, f16);
  hipLaunchKernelGGL(split_blocked_kernel, dim3(grid_for(P.n_chol * n * (xpad(nv) / 8))), dim3(256), 0, st,
                     P.Lov.as<double>(), P.n_chol, N, n, nv, M.Lovx.as<uint16_t>(), ldx, s0, f16);
  hipLaunchKernelGGL(transpose_split_kernel, dim3((unsigned)cdiv(N, 32), (unsigned)cdiv(P.n_chol, 32)), dim3(256), 0, st,
                     P.Lov.as<double>(), P.n_chol, N, N, M.LovTx.as<uint16_t>(), ldj, s0, f16);
  TCI_HIP(hipGetLastError());
  const int iN = (int)N, K = (int)(n * xpad(nv)), KJ = (int)xpad(P.n_chol);
  auto xop = [&](const DevBuf& b, int64_t ld) { return dop(b.p, ld); };
  // B operands as [n][k] rows: Kov, Jbig and CC are symmetric; Lov enters transposed.
  // Z and Z3 are square N x N outputs: 128 x 128 tiles, or 192 x 192 with
  // TCI_PAIRS_TILE=192 (each operand panel fetched a third less often), fed by
  // direct-load stages with TCI_PAIRS_TILE=192d
  // cfg_z (Z, Z3 and the 4-internal product) and cfg_uw (U and W): 0 for
  // 128 x 128 tiles, 1, 2 and 3 for 256 x 256 tiles of the 32 x 32-MFMA
  // kernel with 4, 8 or 16 waves, 4 and 5 for 128 x 192 and 128 x 128 tiles
  // with two direct-load stages; -1 for TCI_PAIRS_TILE (128, 192 or 192d: Z
  // and Z3 alone on 192 x 192 tiles, 256, 256w8 or 256w16: every product on
  // 256 x 256, 128x192d2 or 128d2: every product on those direct-load tiles)
  const char* te = std::getenv("TCI_PAIRS_TILE");
  const std::string tes = te ? te : "128";
  if (cfg_z < 0)
    cfg_z = tes == "256" ? 1
            : tes == "256w8" ? 2
            : tes == "256w16" ? 3
            : tes == "128x192d2" ? 4
            : tes == "128d2" ? 5
                             : 0;
  if (cfg_uw < 0) cfg_uw = cfg_z;
  static const int waves_of[6] = {4, 4, 8, 16, 4, 4};
  auto tile_of = [](int c) { return c == 1 || c == 2 || c == 3 ? 256 : 128; };
  const int ztile = cfg_z ? tile_of(cfg_z) : (tes == "192" || tes == "192d" ? 192 : 128);
  const int ztn = cfg_z == 4 ? 192 : 0;                            // tile_n of the 128 x 192 form
  const int zdma = cfg_z >= 4 ? 2 : (ztile == 192 && tes == "192d" ? 3 : 0);
  const int htile = tile_of(cfg_z), pw = waves_of[cfg_z];          // 4-internal product, waves of Z's 256 tiles
  const int utile = tile_of(cfg_uw), uw = waves_of[cfg_uw];        // U and W
  const int utn = cfg_uw == 4 ? 192 : 0, udma = cfg_uw >= 4 ? 2 : 0;
  M.ringZx.clear();
  M.ringZx.split = true;
  M.ringZx.waves = pw;
  M.ringZx.tile = ztile;
  M.ringZx.tile_n = ztn;  // tasks are cut by tile x tile_n when added
  if (M.z_lowrank)  // Ytx holds the split rows of U = Yt + C~R Lov^T
    M.ringZx.add(P.Zb.as<double>(), N, iN, iN,
                 {dseg(xop(M.CRx, ldx), xop(M.Jx, ldx), K, 1.0), dseg(xop(M.Ytx, ldj), xop(M.LovTx, ldj), KJ, 1.0)},
                 false, false, 0, Blocking(), 5);
  else
    M.ringZx.add(P.Zb.as<double>(), N, iN, iN,
                 {dseg(xop(M.CtRx, ldx), xop(M.Kx, ldx), K, 1.0), dseg(xop(M.CRx, ldx), xop(M.Jx, ldx), K, 1.0),
                  dseg(xop(M.Ytx, ldj), xop(M.LovTx, ldj), KJ, 1.0)},
                 false, false, 0, Blocking(), 5);
  M.ringZx.dma = zdma;
  M.ringZx.upload_tables(st);
  M.ringZ3x.clear();
  M.ringZ3x.split = true;
  M.ringZ3x.waves = pw;
  M.ringZ3x.tile = ztile;
  M.ringZ3x.tile_n = ztn;
  M.ringZ3x.add(P.Z3b.as<double>(), N, iN, iN, {dseg(xop(M.Jx, ldx), xop(M.CCx, ldx), K, 1.0)}, false, false, 0,
                Blocking(), 5);
  M.ringZ3x.dma = zdma;
  M.ringZ3x.upload_tables(st);
  // 4-internal term: Remb = holeA [Cemb; CTemb] with both operands as split
  // rows over k = walk (2 n_pairs, padded); holeA carries the constants' scale
  // and the embedded C that of the pair products' C, so the output scale is
  // the pair products' one
  {
    const int np = P.n_pairs;
    const int64_t v2 = (int64_t)nv * nv, kh = xpad(2 * (int64_t)np), ldh = 2 * kh;
    M.holeAx.alloc((size_t)np * ldh * 2);
    M.embTx.alloc((size_t)v2 * ldh * 2);
    M.bytes += (int64_t)(M.holeAx.bytes + M.embTx.bytes);
    hipLaunchKernelGGL(split_blocked_kernel, dim3(grid_for(np * (kh / 8))), dim3(256), 0, st, P.holeA.as<double>(),
                       (int64_t)np, 2 * (int64_t)np, 1, 2 * np, M.holeAx.as<uint16_t>(), ldh, s0, f16);
    TCI_HIP(hipGetLastError());
    M.holeGx.clear();
    M.holeGx.split = true;
    M.holeGx.waves = pw;
    M.holeGx.tile = htile;
    M.holeGx.tile_n = ztn;
    M.holeGx.add(P.Remb.as<double>(), v2, np, (int)v2, {dseg(xop(M.holeAx, ldh), xop(M.embTx, ldh), (int)kh, 1.0)},
                 false, false, 0, Blocking(), 5);
    M.holeGx.dma = cfg_z >= 4 ? 2 : 0;
    M.holeGx.upload_tables(st);
  }
  Blocking bw;
  bw.cdiv = nv;
  bw.cstride = P.n_chol * nv;
  M.progWx.clear();
  M.progWx.split = true;
  M.progWx.waves = uw;
  M.progWx.tile = utile;
  M.progWx.tile_n = utn;
  M.progWx.add(P.W.as<double>(), nv, (int)P.n_chol, iN, {dseg(xop(M.Lovx, ldx), xop(M.CtRx, ldx), K, 1.0)}, false,
               false, 0, bw, 5);
  M.progWx.dma = udma;
  M.progWx.upload_tables(st);
  // U[(i,a)][J] = Yt + sum_(j,b) C~R[(i,a)][(j,b)] Lov[J][(j,b)], accumulated into Yt
  M.progUx.clear();
  if (M.z_lowrank) {
    M.progUx.split = true;
    M.progUx.waves = uw;
    M.progUx.tile = utile;
    M.progUx.tile_n = utn;
    M.progUx.add(P.Yt.as<double>(), P.n_chol, iN, (int)P.n_chol, {dseg(xop(M.CtRx, ldx), xop(M.Lovx, ldx), K, 1.0)},
                 true, false, 0, Blocking(), 5);
    M.progUx.dma = udma;
    M.progUx.upload_tables(st);
  }
  // With 128 x 128 tiles the products go out as one launch, tasks with the
  // longest k first (Z, then Z3 and W, then the 4-internal product):
  // workgroups take tasks in order as slots free up, so the short tasks fill
  // the last round instead of each launch ending in a partly empty one.  With
  // the low-rank Z, U and W go out first as a launch of their own (Z reads U).
  auto merge = [](GemmProgram& out, std::initializer_list<const GemmProgram*> gs, hipStream_t s) {
    out.clear();
    out.split = true;
    out.tile = (*gs.begin())->tile;
    out.tile_n = (*gs.begin())->tile_n;
    out.dma = (*gs.begin())->dma;
    out.waves = (*gs.begin())->waves;
    for (const GemmProgram* g : gs) {
      const int s0 = (int)out.segs.size();
      for (const Seg& x : g->segs) out.segs.push_back(x);
      for (Task t : g->tasks) {
        t.seg0 += s0;
        t.seg1 += s0;
        out.tasks.push_back(t);
      }
    }
    out.upload_tables(s);
  };
  M.pairsx.clear();
  M.pairsUWx.clear();
  // one launch holds tasks of one tile shape only
  // (programs merged into one launch share tile, tile_n and stages)
  const bool zmerge = ((ztile == 128 || ztile == 256) && !zdma) || cfg_z >= 4;
  M.pairs_merged = zmerge && (M.z_lowrank || (utile == ztile && utn == ztn && udma == zdma));
  if (M.pairs_merged) {
    if (M.z_lowrank) {
      merge(M.pairsUWx, {&M.progUx, &M.progWx}, st);
      merge(M.pairsx, {&M.ringZx, &M.ringZ3x, &M.holeGx}, st);
    } else {
      merge(M.pairsx, {&M.ringZx, &M.ringZ3x, &M.progWx, &M.holeGx}, st);
    }
  }
}

// ring, 3-external (X and W) and 4-internal terms into R and W; needs Y
void dense_pairs_terms(tci_plan& P, int precision, hipStream_t st) {
  const int n = P.n_int, nv = P.n_v;
  const int64_t N = (int64_t)n * nv;
  const int nt = (nv + 31) / 32;
  const dim3 g(n * nt, n * nt), b(256);
  const EmbedTab E = etab(P);
  const std::string pn = precision == TCI_SIGMA_FP64 ? "fp64" : (precision == TCI_SIGMA_FP16X3 ? "fp16x3" : "bf16x3");
  auto run = [&](const std::string& role, const GemmProgram& gp) {
    P.kt_run(pn + " " + role, gp, precision, st, [&] {
      if (precision == TCI_SIGMA_FP64) gp.launch(st);
      else gp.launch_split(precision, st, 0, P.mx.scales.as<float>());
    });
  };
  int e0 = P.kt_begin(st);
  hipLaunchKernelGGL(y_transpose_kernel, dim3((unsigned)cdiv(P.n_chol, 32), nt, n), b, 0, st, P.Y.as<double>(), n, nv,
                     P.n_chol, P.Yt.as<double>());
  if (precision == TCI_SIGMA_FP64) {
    hipLaunchKernelGGL(relayout_kernel<false>, g, b, 0, st, P.tab, E, P.Ct.as<double>(), false, 1.0, P.CtRb.p, N,
                       nullptr, 0);
    hipLaunchKernelGGL(relayout_kernel<false>, g, b, 0, st, P.tab, E, P.C.as<double>(), false, 1.0, P.CRb.p, N,
                       nullptr, 0);
    hipLaunchKernelGGL(relayout_kernel<false>, g, b, 0, st, P.tab, E, P.C.as<double>(), true, 1.0, P.CCb.p, N,
                       nullptr, 0);
    TCI_HIP(hipGetLastError());
    P.kt_end(e0, pn + " pair relayouts", 0.0, precision, st);
    if (P.z_lowrank) run("pair product U", P.progU);
    run("pair product Z", P.ringZ);
    run("pair product Z3", P.ringZ3);
    run("pair product W", P.progWd);
  } else {
    tci_plan::MixedLadder& M = P.mx;
    const int64_t ldx = 2 * (int64_t)n * xpad(nv), ldj = 2 * xpad(P.n_chol);
    float* sc = M.scales.as<float>();
    const int f16 = precision == TCI_SIGMA_FP16X3 ? 1 : 0;
    // fp16x3: one scale for C, Ct and Yt (|Ct| <= 3 max|C|; max|C| from this sigma's ladder)
    if (f16) launch_split_scales(2, P.Yt.as<double>(), N * P.n_chol, M.maxbits.as<unsigned>(), 1, sc, (float)P.dmax, st);
    const float* sv = sc + 6;
    hipLaunchKernelGGL(relayout_kernel<true>, g, b, 0, st, P.tab, E, P.Ct.as<double>(), false, 1.0, M.CtRx.p, ldx, sv,
                       f16);
    hipLaunchKernelGGL(relayout_kernel<true>, g, b, 0, st, P.tab, E, P.C.as<double>(), false, -1.0, M.CRx.p, ldx, sv,
                       f16);
    hipLaunchKernelGGL(relayout_kernel<true>, g, b, 0, st, P.tab, E, P.C.as<double>(), true, 1.0, M.CCx.p, ldx, sv, f16);
    if (!M.z_lowrank)
      hipLaunchKernelGGL(split_blocked_kernel, dim3(grid_for(N * (xpad(P.n_chol) / 8))), b, 0, st, P.Yt.as<double>(), N,
                         P.n_chol, 1, (int)P.n_chol, M.Ytx.as<uint16_t>(), ldj, sv, f16);
    TCI_HIP(hipGetLastError());
    const int64_t v2 = (int64_t)nv * nv, kh = xpad(2 * (int64_t)P.n_pairs);
    hipLaunchKernelGGL(embed_split_kernel, dim3((unsigned)cdiv(v2, 256), (unsigned)(kh / 8)), b, 0, st, P.tab, E,
                       P.C.as<double>(), M.embTx.as<uint16_t>(), 2 * kh, sv, f16);
    TCI_HIP(hipGetLastError());
    P.kt_end(e0, pn + " pair relayouts", 0.0, precision, st);
    if (M.z_lowrank) {  // U (into Yt) and W, then the split rows of U, then Z, Z3 and the 4-internal product
      if (M.pairs_merged) {
        run("pair products U, W", M.pairsUWx);
      } else {
        run("pair product U", M.progUx);
        run("pair product W", M.progWx);
      }
      e0 = P.kt_begin(st);
      hipLaunchKernelGGL(split_blocked_kernel, dim3(grid_for(N * (xpad(P.n_chol) / 8))), b, 0, st, P.Yt.as<double>(),
                         N, P.n_chol, 1, (int)P.n_chol, M.Ytx.as<uint16_t>(), ldj, sv, f16);
      TCI_HIP(hipGetLastError());
      P.kt_end(e0, pn + " pair relayouts", 0.0, precision, st);
      if (M.pairs_merged) {
        run("pair products Z, Z3, 4-internal", M.pairsx);
      } else {
        run("pair product Z", M.ringZx);
        run("pair product Z3", M.ringZ3x);
        run("4-internal product", M.holeGx);
      }
    } else if (M.pairs_merged) {
      run("pair products Z, Z3, W, 4-internal", M.pairsx);
    } else {
      run("pair product Z", M.ringZx);
      run("pair product Z3", M.ringZ3x);
      run("pair product W", M.progWx);
      run("4-internal product", M.holeGx);
    }
  }
  if (precision == TCI_SIGMA_FP64) {
    e0 = P.kt_begin(st);
    hipLaunchKernelGGL(embed_blocks_kernel, dim3(nt * nt, P.n_pairs), b, 0, st, P.tab, E, P.C.as<double>(),
                       P.Cemb.as<double>(), P.CTemb.as<double>());
    TCI_HIP(hipGetLastError());
    P.kt_end(e0, pn + " pair relayouts", 0.0, precision, st);
    run("4-internal product", P.holeG);
  }
  const int ntd = (P.dmax + 31) / 32;
  e0 = P.kt_begin(st);
  hipLaunchKernelGGL(ring_combine_kernel, dim3(ntd * ntd, P.n_pairs), b, 0, st, P.tab, P.Zb.as<double>(),
                     P.Z3b.as<double>(), P.Remb.as<double>(), P.R.as<double>());
  TCI_HIP(hipGetLastError());
  P.kt_end(e0, pn + " ring_combine_kernel", 0.0, precision, st);
}

}  // namespace tci
