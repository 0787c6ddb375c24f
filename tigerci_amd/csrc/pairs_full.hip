// Pair-pair terms as dense products when every internal pair i >= j is a
// strong walk whose domain is the whole virtual space (the untruncated
// synthetic problems).  N = n_int n_v; an ordered internal pair (p, q) names an
// n_v x n_v block, X^{pq} = X_Q for p >= q and X_Q^T for p < q, with
// Q = p(p+1)/2 + q the walk of the unordered pair.  Per sigma:
//   CR [(p,r)][(q,c)] = C^{pq}[r][c],   CtR likewise from Ct = 2C - C^T,
//   CC [(p,r)][(q,c)] = C^{qp}[r][c]    (CR and CC are symmetric matrices)
// and once per plan
//   Kov[(k,c)][(j,b)] = (kc|jb),  Jbig[(k,c)][(j,b)] = (kj|cb)  (both symmetric).
// Then Z = CtR Kov - CR Jbig and Z3 = Jbig CC hold, for every ordered (i, j),
//   Z (i,j) = sum_k [Ct^{ik} K^{kj} - C^{ik} J^{kj}],  Z3(j,i) = sum_k J^{kj} C^{ik},
// so the ring terms Z^{ij} + Z^{ji T} of R^{ij} are
//   R^{ij}[a][b] += Z(i,j)[a][b] - Z3(j,i)[a][b] + Z(j,i)[b][a] - Z3(i,j)[b][a].
// The 4-internal term sum_{Q=(k,l)} [(ki|lj) C_Q + (li|kj) C_Q^T] is one
// product of an n_pairs x 2 n_pairs coefficient matrix with [C; C^T] seen as
// 2 n_pairs x n_v^2.  All products run on the dense FP64 kernel, or for the
// split precisions Z and Z3 on the split-row kernel (k padded per block to a
// multiple of 8).
#include "plan.hpp"

namespace tci {

namespace {

__device__ __forceinline__ int64_t walk_of(int p, int q) {
  const int hi = max(p, q), lo = min(p, q);
  return (int64_t)hi * (hi + 1) / 2 + lo;
}

// Jbig[(p,r)][(q,s)] = (pq|rs) from Jvv[walk(p,q)][r][s]
__global__ void jbig_kernel(const double* __restrict__ Jvv, int n_int, int n_v, double* __restrict__ Jbig) {
  const int64_t N = (int64_t)n_int * n_v;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * N; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e / N, col = e - row * N;
    const int p = (int)(row / n_v), r = (int)(row - (int64_t)p * n_v);
    const int q = (int)(col / n_v), s = (int)(col - (int64_t)q * n_v);
    Jbig[e] = Jvv[(walk_of(p, q) * n_v + r) * n_v + s];
  }
}

// Output block (p, q) = sign * X^{pq} (cc = false) or sign * X^{qp} = sign * (X^{pq})^T
// (cc = true), from the dense d x d blocks of src (walk order, d^2 apart).
// 32 x 32 tiles through LDS so both orientations read and write coalesced.
// FP64: out[(p d + r) ldo + q d + c];  split rows: out16[(p d + r) ldo + xidx(q dp + c)],
// dp = d rounded up to 8, padding columns zero.
template <bool SPLIT>
__global__ void relayout_kernel(const double* __restrict__ src, int d, bool cc, double sign, void* __restrict__ out,
                                int64_t ldo) {
  __shared__ double tile[32][33];
  const int nt = (d + 31) >> 5;
  const int q = blockIdx.x / nt, ct = blockIdx.x - q * nt;
  const int p = blockIdx.y / nt, rt = blockIdx.y - p * nt;
  const double* blk = src + walk_of(p, q) * d * d;
  const bool direct = (p >= q) != cc;
  const int r0 = rt * 32, c0 = ct * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    if (direct) {
      const int r = r0 + yy, c = c0 + tx;
      tile[yy][tx] = (r < d && c < d) ? blk[(int64_t)r * d + c] : 0.0;
    } else {  // element (r, c) = blk[c][r]: read row c, column r
      const int c = c0 + yy, r = r0 + tx;
      tile[tx][yy] = (r < d && c < d) ? blk[(int64_t)c * d + r] : 0.0;
    }
  }
  __syncthreads();
  if (!SPLIT) {
    double* o = static_cast<double*>(out);
    for (int yy = ty; yy < 32; yy += 8) {
      const int r = r0 + yy, c = c0 + tx;
      if (r < d && c < d) o[((int64_t)p * d + r) * ldo + (int64_t)q * d + c] = sign * tile[yy][tx];
    }
  } else if (threadIdx.x < 128) {
    const int dp = (d + 7) & ~7;
    const int yy = threadIdx.x >> 2, c = c0 + 8 * (threadIdx.x & 3);
    const int r = r0 + yy;
    if (r < d && c < dp) {
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = (c + t < d) ? (float)(sign * tile[yy][c - c0 + t]) : 0.f;
      put_split8(static_cast<uint16_t*>(out) + ((int64_t)p * d + r) * ldo + xidx((int64_t)q * dp + c), v);
    }
  }
}

// split rows of an FP64 matrix whose columns are nb blocks of d, each block's
// k padded to dp: out16[row ldo + xidx(q dp + c)] = src[row ld + q d + c]
__global__ void split_blocked_kernel(const double* __restrict__ src, int64_t rows, int64_t ld, int nb, int d,
                                     uint16_t* __restrict__ out, int64_t ldo) {
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
    put_split8(out + row * ldo + xidx((int64_t)q * dp + c), v);
  }
}

// R_P[a][b] += Z(i,j)[a][b] - Z3(j,i)[a][b] + Z(j,i)[b][a] - Z3(i,j)[b][a]
__global__ void ring_combine_kernel(PairTab T, const double* __restrict__ Z, const double* __restrict__ Z3,
                                    double* __restrict__ R) {
  __shared__ double tile[32][33];
  const int d = T.n_v;
  const int64_t N = (int64_t)T.n_int * d;
  const int nt = (d + 31) >> 5;
  const int P = blockIdx.y;
  const int i = T.pairs[2 * P], j = T.pairs[2 * P + 1];
  const int a0 = (blockIdx.x / nt) * 32, b0 = (blockIdx.x % nt) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {  // transposed part, read along a
    const int b = b0 + yy, a = a0 + tx;
    tile[yy][tx] = (a < d && b < d) ? Z[((int64_t)j * d + b) * N + (int64_t)i * d + a] -
                                          Z3[((int64_t)i * d + b) * N + (int64_t)j * d + a]
                                    : 0.0;
  }
  __syncthreads();
  double* Rb = R + T.cd_off[P];
  for (int yy = ty; yy < 32; yy += 8) {
    const int a = a0 + yy, b = b0 + tx;
    if (a < d && b < d)
      Rb[(int64_t)a * d + b] += Z[((int64_t)i * d + a) * N + (int64_t)j * d + b] -
                                Z3[((int64_t)j * d + a) * N + (int64_t)i * d + b] + tile[tx][yy];
  }
}

// CT_Q = C_Q^T for every walk (d_Q x d_Q blocks at cd_off)
__global__ void transpose_blocks_kernel(PairTab T, const double* __restrict__ C, double* __restrict__ CT) {
  __shared__ double tile[32][33];
  const int P = blockIdx.y;
  const int d = (int)(T.dom_off[P + 1] - T.dom_off[P]);
  const int nt = (d + 31) >> 5;
  if ((int)blockIdx.x >= nt * nt) return;
  const int r0 = (blockIdx.x / nt) * 32, c0 = (blockIdx.x % nt) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const double* src = C + T.cd_off[P];
  double* dst = CT + T.cd_off[P];
  for (int yy = ty; yy < 32; yy += 8) {
    const int r = r0 + yy, c = c0 + tx;
    tile[yy][tx] = (r < d && c < d) ? src[(int64_t)r * d + c] : 0.0;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int c = c0 + yy, r = r0 + tx;
    if (r < d && c < d) dst[(int64_t)c * d + r] = tile[tx][yy];
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

}  // namespace

bool dense_pairs_eligible(const tci_plan& P) {
  const int n = P.n_int;
  if (P.n_pairs != n * (n + 1) / 2) return false;
  if ((int64_t)n * P.n_v >= ((int64_t)1 << 31) / 2) return false;
  for (int Q = 0; Q < P.n_pairs; ++Q) {
    const int i = P.h_pairs[2 * Q], j = P.h_pairs[2 * Q + 1];
    if ((int64_t)i * (i + 1) / 2 + j != Q) return false;
    if (P.h_dom_off[Q + 1] - P.h_dom_off[Q] != P.n_v) return false;
  }
  return true;
}

// FP64 tables: Jbig, the work matrices and the three dense programs; o4 is the
// host copy of (ki|lj) = o4[((k n + i) n + l) n + j]
void dense_pairs_build(tci_plan& P, const std::vector<double>& o4, hipStream_t st) {
  const int n = P.n_int, d = P.n_v, np = P.n_pairs;
  const int64_t N = (int64_t)n * d, d2 = (int64_t)d * d;
  P.Jbig.alloc((size_t)N * N * 8);
  hipLaunchKernelGGL(jbig_kernel, dim3(grid_for(N * N)), dim3(256), 0, st, P.Jvv.as<double>(), n, d,
                     P.Jbig.as<double>());
  TCI_HIP(hipGetLastError());
  for (DevBuf* b : {&P.CRb, &P.CtRb, &P.CCb, &P.Zb, &P.Z3b}) b->alloc((size_t)N * N * 8);
  P.CTb.alloc((size_t)P.sum_d2 * 8);
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
  const int iN = (int)N;
  P.ringZ.clear();
  P.ringZ.dense = true;
  P.ringZ.tile = 128;
  P.ringZ.add(P.Zb.as<double>(), N, iN, iN,
              {dseg(dop(P.CtRb.p, N), dop(P.Kov.p, N), iN, 1.0), dseg(dop(P.CRb.p, N), dop(P.Jbig.p, N), iN, -1.0)},
              false);
  P.ringZ.upload_tables(st);
  P.ringZ3.clear();
  P.ringZ3.dense = true;
  P.ringZ3.tile = 128;
  P.ringZ3.add(P.Z3b.as<double>(), N, iN, iN, {dseg(dop(P.Jbig.p, N), dop(P.CCb.p, N), iN, 1.0)}, false);
  P.ringZ3.upload_tables(st);
  P.holeG.clear();
  P.holeG.dense = true;
  P.holeG.tile = 128;
  P.holeG.add(P.R.as<double>(), d2, np, (int)d2,
              {dseg(dop(P.holeA.p, 2 * np), dop(P.C.p, d2), np, 1.0),
               dseg(dop(P.holeA.as<double>() + np, 2 * np), dop(P.CTb.p, d2), np, 1.0)},
              true);
  P.holeG.upload_tables(st);
}

// split copies of Kov and Jbig and the split programs for Z and Z3
void dense_pairs_build_split(tci_plan& P, hipStream_t st) {
  tci_plan::MixedLadder& M = P.mx;
  const int n = P.n_int, d = P.n_v;
  const int64_t N = (int64_t)n * d, ldx = 2 * (int64_t)n * xpad(d);
  for (DevBuf* b : {&M.Kx, &M.Jx, &M.CRx, &M.CtRx, &M.CCx}) {
    b->alloc((size_t)N * ldx * 2);
    M.bytes += (int64_t)b->bytes;
  }
  const int64_t work = N * n * (xpad(d) / 8);
  hipLaunchKernelGGL(split_blocked_kernel, dim3(grid_for(work)), dim3(256), 0, st, P.Kov.as<double>(), N, N, n, d,
                     M.Kx.as<uint16_t>(), ldx);
  hipLaunchKernelGGL(split_blocked_kernel, dim3(grid_for(work)), dim3(256), 0, st, P.Jbig.as<double>(), N, N, n, d,
                     M.Jx.as<uint16_t>(), ldx);
  TCI_HIP(hipGetLastError());
  const int iN = (int)N, K = (int)(n * xpad(d));
  auto xop = [&](const DevBuf& b) { return dop(b.p, ldx); };
  M.ringZx.clear();
  M.ringZx.split = true;
  M.ringZx.tile = 128;
  // B operands as [n][k] rows: Kov and Jbig are symmetric, CC is symmetric
  M.ringZx.add(P.Zb.as<double>(), N, iN, iN, {dseg(xop(M.CtRx), xop(M.Kx), K, 1.0), dseg(xop(M.CRx), xop(M.Jx), K, 1.0)},
               false);
  M.ringZx.upload_tables(st);
  M.ringZ3x.clear();
  M.ringZ3x.split = true;
  M.ringZ3x.tile = 128;
  M.ringZ3x.add(P.Z3b.as<double>(), N, iN, iN, {dseg(xop(M.Jx), xop(M.CCx), K, 1.0)}, false);
  M.ringZ3x.upload_tables(st);
}

// ring (Z, Z3 and their combination) and 4-internal terms into R
void dense_pairs_terms(tci_plan& P, int precision, hipStream_t st) {
  const int n = P.n_int, d = P.n_v;
  const int64_t N = (int64_t)n * d;
  const int nt = (d + 31) / 32;
  const dim3 g(n * nt, n * nt), b(256);
  if (precision == TCI_SIGMA_FP64) {
    hipLaunchKernelGGL(relayout_kernel<false>, g, b, 0, st, P.Ct.as<double>(), d, false, 1.0, P.CtRb.p, N);
    hipLaunchKernelGGL(relayout_kernel<false>, g, b, 0, st, P.C.as<double>(), d, false, 1.0, P.CRb.p, N);
    hipLaunchKernelGGL(relayout_kernel<false>, g, b, 0, st, P.C.as<double>(), d, true, 1.0, P.CCb.p, N);
    TCI_HIP(hipGetLastError());
    P.ringZ.launch(st);
    P.ringZ3.launch(st);
  } else {
    tci_plan::MixedLadder& M = P.mx;
    const int64_t ldx = 2 * (int64_t)n * xpad(d);
    hipLaunchKernelGGL(relayout_kernel<true>, g, b, 0, st, P.Ct.as<double>(), d, false, 1.0, M.CtRx.p, ldx);
    hipLaunchKernelGGL(relayout_kernel<true>, g, b, 0, st, P.C.as<double>(), d, false, -1.0, M.CRx.p, ldx);
    hipLaunchKernelGGL(relayout_kernel<true>, g, b, 0, st, P.C.as<double>(), d, true, 1.0, M.CCx.p, ldx);
    TCI_HIP(hipGetLastError());
    M.ringZx.launch_split(precision, st);
    M.ringZ3x.launch_split(precision, st);
  }
  hipLaunchKernelGGL(ring_combine_kernel, dim3(nt * nt, P.n_pairs), b, 0, st, P.tab, P.Zb.as<double>(),
                     P.Z3b.as<double>(), P.R.as<double>());
  hipLaunchKernelGGL(transpose_blocks_kernel, dim3(nt * nt, P.n_pairs), b, 0, st, P.tab, P.C.as<double>(),
                     P.CTb.as<double>());
  TCI_HIP(hipGetLastError());
  P.holeG.launch(st);
}

}  // namespace tci
