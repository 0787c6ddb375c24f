// Split-precision (bf16x3 / scaled fp16x3) two-phase ladder
//   R_P = sum_J L^J[D,D] C_P L^J[D,D].
//
// Both phases are C = A B^T products with k-contiguous split-row operands
// (dense_split.hip).  With T' = L C^T per Cholesky vector,
//   phase 1  T'_P[(jj,a)][c] = sum_k L^J[a][k] C_P[c][k]      (A = L rows, B = C rows)
//   phase 2  R_P[n][a]      += sum_jj sum_c L^J[n][c] T'_P[jj][a][c]
//                                                             (A = L rows, B = T' rows)
// and L^J symmetric, so R = L T'^T = L C L.  Phase 1 writes T' directly in
// split-row form; phase 2 accumulates FP32 tile sums into FP64 R (split into
// partial R buffers along K when the outputs alone cannot fill the chip).
// Split inputs: L^J[v][v] of all J once per plan (full-domain pairs), packed
// L^J[D,D] of local-domain pairs per chunk, C_P per sigma call.
#include "plan.hpp"

#include <algorithm>
#include <cstdlib>

namespace tci {

namespace {

// Cx[P] rows r of C_P (d x d, FP64) in split-row form, row stride 2 xpad(d)
// Cx[P] rows c = split row c of C_P; with nvec > 1 the walks hold nvec such
// blocks ([P][v][c][k] at nvec cx_off[P]) and this writes slot vec
__global__ void split_blocks_kernel(PairTab T, const int64_t* __restrict__ cx_off, const double* __restrict__ C,
                                    uint16_t* __restrict__ Cx, const float* scl, int f16, int nvec, int vec) {
  const float sc = scale_at(scl);
  const int P = blockIdx.y;
  const int d = (int)(T.dom_off[P + 1] - T.dom_off[P]);
  const int nch = (d + 7) >> 3;
  const double* src = C + T.cd_off[P];
  uint16_t* dst = Cx + (int64_t)nvec * cx_off[P] + (int64_t)vec * d * 16 * nch;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < d * nch; e += gridDim.x * blockDim.x) {
    const int r = e / nch, ch = e - r * nch;
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = 8 * ch + t;
      v[t] = k < d ? (float)src[(int64_t)r * d + k] : 0.f;
    }
    put_split8(dst + (int64_t)r * 16 * nch + 16 * ch, v, sc, f16);
  }
}

// Cx[P] rows c = split column c of C_P (row c of C_P^T), 32 x 32 tiles via LDS
__global__ void split_blocks_T_kernel(PairTab T, const int64_t* __restrict__ cx_off, const double* __restrict__ C,
                                      uint16_t* __restrict__ Cx, const float* scl, int f16) {
  __shared__ float tile[32][33];
  const int P = blockIdx.y;
  const int d = (int)(T.dom_off[P + 1] - T.dom_off[P]);
  const int nt = (d + 31) >> 5;
  if ((int)blockIdx.x >= nt * nt) return;
  const int k0 = (blockIdx.x / nt) * 32, c0 = (blockIdx.x % nt) * 32;
  const double* src = C + T.cd_off[P];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int k = k0 + yy, c = c0 + tx;
    tile[yy][tx] = (k < d && c < d) ? (float)src[(int64_t)k * d + c] : 0.f;
  }
  __syncthreads();
  if (threadIdx.x < 128) {
    const int nch = (d + 7) >> 3;
    const int cy = threadIdx.x >> 2, ch = threadIdx.x & 3;
    const int c = c0 + cy, k = k0 + 8 * ch;
    if (c < d && k < 8 * nch) {
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = tile[8 * ch + t][cy];
      put_split8(Cx + cx_off[P] + (int64_t)c * 16 * nch + 2 * k, v, scale_at(scl), f16);
    }
  }
}

// Lx[(J, a)] = split row a of L^J[v][v] for all J, row stride 2 xpad(n_v)
template <class LT>
__global__ void split_lvv_kernel(const LT* __restrict__ Lvv, int64_t rows, int n_v, uint16_t* __restrict__ Lx,
                                 const float* scl, int f16) {
  const float sc = scale_at(scl);
  const int nch = (n_v + 7) >> 3;
  const int64_t n = rows * nch;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / nch;
    const int ch = (int)(e - r * nch);
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = 8 * ch + t;
      v[t] = k < n_v ? (float)Lvv[r * n_v + k] : 0.f;
    }
    put_split8(Lx + r * 16 * nch + 16 * ch, v, sc, f16);
  }
}

// Lpx[q] rows (jj, r) = split row r of L^{J0+jj}[D_P, D_P] for local pair q.
// One wave per row: the lanes first gather up to 512 domain columns of the row
// in order (lane l takes columns l, l + 64, ..., so that one load instruction
// reads increasing addresses), the row passes through LDS, and lane l then
// splits columns 8l .. 8l + 7 into one chunk.  256 threads.
template <class LT>
__global__ __launch_bounds__(256) void pack_L_split_kernel(PairTab T, const int32_t* __restrict__ pairs_local,
                                                           const int64_t* __restrict__ lp_off,
                                                           const LT* __restrict__ Lvv, int64_t J0, int len,
                                                           uint16_t* __restrict__ Lpx, const float* scl, int f16) {
  const float sc = scale_at(scl);
  const int q = blockIdx.y;
  const int P = pairs_local[q];
  const int64_t off = T.dom_off[P];
  const int d = (int)(T.dom_off[P + 1] - off);
  const int nch = (d + 7) >> 3;
  const int64_t nv = T.n_v;
  const int64_t rows = (int64_t)len * d;
  const int32_t* dix = T.dom_idx + off;
  uint16_t* out = Lpx + lp_off[q];
  __shared__ float rowbuf[4][512];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // the workgroup steps through rows together, so every barrier is reached
  // by all of its threads
  for (int64_t base = (int64_t)blockIdx.x * 4; base < rows; base += (int64_t)gridDim.x * 4) {
    const int64_t row = base + w;  // jj d + r
    const bool ok = row < rows;
    const LT* L = Lvv;
    if (ok) {
      const int64_t jj = row / d;
      const int r = (int)(row - jj * d);
      L = Lvv + (J0 + jj) * nv * nv + (int64_t)dix[r] * nv;
    }
    for (int c0 = 0; c0 < d; c0 += 512) {
      const int cn = min(512, d - c0);
      for (int t = lane; t < 512; t += 64) rowbuf[w][t] = (ok && t < cn) ? (float)L[dix[c0 + t]] : 0.f;
      __syncthreads();
      if (ok)
        for (int ch = lane; 8 * ch < cn; ch += 64) {
          float v[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) v[t] = rowbuf[w][8 * ch + t];
          put_split8(out + row * 16 * nch + 16 * ((c0 >> 3) + ch), v, sc, f16);
        }
      __syncthreads();
    }
  }
}

// out = max(out, max |x|) as float bits (non-negative floats order like their bits)
template <class T>
__global__ void absmax_kernel(const T* __restrict__ x, int64_t n, unsigned* __restrict__ out) {
  __shared__ float part[4];
  float m = 0.f;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf((float)x[e]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
    atomicMax(out, __float_as_uint(m));
  }
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// power of two s with m s in [2^13, 2^14): FP16 parts keep 22 bits with room
__device__ __forceinline__ float p2scale(float m) { return m > 0.f ? ldexpf(1.f, 13 - ilogbf(m)) : 1.f; }

// Scales of the scaled-FP16 split (indices documented at MixedLadder::scales).
//   mode 0: constants from mx[2]     mode 1: ladder from mx[0] = max|C|
//   mode 2: pair products from mx[0], mx[1] = max|Yt|
__global__ void set_scales_kernel(int mode, const unsigned* __restrict__ mx, float* __restrict__ sc, float dmax) {
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    const float mi = __uint_as_float(mx[2]);
    sc[0] = p2scale(mi);
    sc[8] = mi;
  } else if (mode == 1) {
    const float mc = __uint_as_float(mx[0]);
    sc[1] = p2scale(mc);
    sc[2] = p2scale(dmax * sc[8] * mc);  // |T| <= d max|L| max|C|
    sc[3] = sc[2] / (sc[0] * sc[1]);
    sc[4] = 1.f / (sc[2] * sc[0]);
  } else {
    const float mv = fmaxf(3.f * __uint_as_float(mx[0]), __uint_as_float(mx[1]));  // |Ct| <= 3 max|C|
    sc[6] = p2scale(mv);
    sc[5] = 1.f / (sc[0] * sc[6]);
  }
}

inline unsigned blocks_for(int64_t work, unsigned cap) {
  return (unsigned)std::min<int64_t>(std::max<int64_t>(cdiv(work, 256), 1), cap);
}

Op xop(const void* base, int64_t rs, int64_t jstride = 0) {
  Op o;
  o.base = base;
  o.rs = rs;
  o.cs = 1;
  o.rmap = nullptr;
  o.cmap = nullptr;
  o.jstride = jstride;
  o.f32 = 0;
  return o;
}

Seg xseg(const Op& A, const Op& B, int K) {
  Seg s;
  s.A = A;
  s.B = B;
  s.K = K;
  s.alpha = 1.0;
  return s;
}

}  // namespace

void launch_split_scales(int mode, const double* x, int64_t n, unsigned* mx, int slot, float* sc, float dmax,
                         hipStream_t st) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, st, mx + slot, 1);
  if (n > 0) hipLaunchKernelGGL(absmax_kernel<double>, dim3(blocks_for(n, 8192)), dim3(256), 0, st, x, n, mx + slot);
  hipLaunchKernelGGL(set_scales_kernel, dim3(1), dim3(64), 0, st, mode, mx, sc, dmax);
  TCI_HIP(hipGetLastError());
}

// One chunk of Cholesky vectors of the split ladder: pack L^J[D,D] of the
// local-domain walks unless prepacked, then phase 1 and phase 2.
void mixed_ladder_chunk(tci_plan& P, int precision, const std::pair<int64_t, int>& ch, hipStream_t st) {
  tci_plan::MixedLadder& M = P.mx;
  const int f16 = precision == TCI_SIGMA_FP16X3 ? 1 : 0;
  const float* sc = M.scales.as<float>();
  const tci_plan::LadderPhase& ph = M.phase[ch.second];
  const std::string pn = precision == TCI_SIGMA_FP16X3 ? "fp16x3" : "bf16x3";
  if (M.n_local && !M.prepacked) {
    const unsigned gxb = 1024;
    const int e0 = P.kt_begin(st);
    if (P.lvv_f32)
      hipLaunchKernelGGL(pack_L_split_kernel<float>, dim3(gxb, M.n_local), dim3(256), 0, st, P.tab,
                         M.pairs_local.as<int32_t>(), ph.lp_off.as<int64_t>(), static_cast<const float*>(P.Lvv.p),
                         ch.first, (int)ph.Lc, M.Lpx.as<uint16_t>(), sc, f16);
    else
      hipLaunchKernelGGL(pack_L_split_kernel<double>, dim3(gxb, M.n_local), dim3(256), 0, st, P.tab,
                         M.pairs_local.as<int32_t>(), ph.lp_off.as<int64_t>(), static_cast<const double*>(P.Lvv.p),
                         ch.first, (int)ph.Lc, M.Lpx.as<uint16_t>(), sc, f16);
    TCI_HIP(hipGetLastError());
    P.kt_end(e0, pn + " ladder pack_L_split_kernel", 0.0, precision, st);
  }
  auto run = [&](const std::string& role, const GemmProgram& g) {
    P.kt_run(pn + " ladder " + role, g, precision, st, [&] { g.launch_split(precision, st, ch.first, sc); });
  };
  if (M.merged) {
    run("phase 1", ph.p1);
    run("phase 2", ph.p2);
  } else {
    for (const auto& g : ph.p1x) run("phase 1", g);
    for (const auto& g : ph.pp1) run("phase 1 (one walk)", g);
    for (const auto& g : ph.p2x) run("phase 2", g);
    for (const auto& g : ph.pp2) run("phase 2 (one walk)", g);
  }
}

void mixed_prepare(tci_plan& P, int precision, hipStream_t st) {
  tci_plan::MixedLadder& M = P.mx;
  TCI_REQUIRE(precision == TCI_SIGMA_BF16X3 || precision == TCI_SIGMA_FP16X3, "not a split precision");
  if (M.ready && M.prec == precision) return;
  TCI_REQUIRE(!P.capturing, "split operands are prepared outside graph capture");
  // captured graphs point into the copies rebuilt below
  P.drop_graphs();
  M.ready = false;
  M.bat.nvec = 0;  // batched programs point into the copies rebuilt below
  M.prec = precision;
  const int f16 = precision == TCI_SIGMA_FP16X3 ? 1 : 0;
  // scales: all 1 for bf16x3 (FP32 exponent range); for fp16x3 the constants'
  // scale now, the per-sigma ones inside every sigma
  {
    std::vector<float> one(16, 1.f);
    upload(M.scales, one, st);
    std::vector<unsigned> z(4, 0u);
    upload(M.maxbits, z, st);
  }
  float* sc = M.scales.as<float>();
  if (f16) {
    unsigned* mx = M.maxbits.as<unsigned>();
    const int64_t nvv = P.n_chol * (int64_t)P.n_v * P.n_v, nov = P.n_chol * (int64_t)P.n_int * P.n_v;
    const int64_t niv = (int64_t)P.n_int * P.n_v;
    if (P.lvv_f32)
      hipLaunchKernelGGL(absmax_kernel<float>, dim3(blocks_for(nvv, 8192)), dim3(256), 0, st, P.Lvv.as<float>(), nvv,
                         mx + 2);
    else
      hipLaunchKernelGGL(absmax_kernel<double>, dim3(blocks_for(nvv, 8192)), dim3(256), 0, st, P.Lvv.as<double>(), nvv,
                         mx + 2);
    if (P.dense_pairs) {
      hipLaunchKernelGGL(absmax_kernel<double>, dim3(blocks_for(niv * niv, 8192)), dim3(256), 0, st,
                         P.Kov.as<double>(), niv * niv, mx + 2);
      hipLaunchKernelGGL(absmax_kernel<double>, dim3(blocks_for(niv * niv, 8192)), dim3(256), 0, st,
                         P.Jbig.as<double>(), niv * niv, mx + 2);
      hipLaunchKernelGGL(absmax_kernel<double>, dim3(blocks_for(nov, 8192)), dim3(256), 0, st, P.Lov.as<double>(),
                         nov, mx + 2);
      const int64_t nh = 2 * (int64_t)P.n_pairs * P.n_pairs;  // 4-internal coefficients
      hipLaunchKernelGGL(absmax_kernel<double>, dim3(blocks_for(nh, 8192)), dim3(256), 0, st, P.holeA.as<double>(),
                         nh, mx + 2);
    }
    hipLaunchKernelGGL(set_scales_kernel, dim3(1), dim3(64), 0, st, 0, mx, sc, (float)P.dmax);
    TCI_HIP(hipGetLastError());
  }
  const int n_pairs = P.n_pairs, n_v = P.n_v;
  const int64_t nvp = xpad(n_v), n_chol = P.n_chol;
  std::vector<int64_t> cxo(n_pairs + 1, 0);
  std::vector<int32_t> loc;
  int64_t sum_x = 0, sum_local_x = 0;  // sum over pairs of d * 2 xpad(d) (16-bit elements)
  bool any_full = false;
  for (int Q = 0; Q < n_pairs; ++Q) {
    const int64_t d = P.h_dom_off[Q + 1] - P.h_dom_off[Q];
    const int64_t x = d * 2 * xpad(d);
    cxo[Q + 1] = cxo[Q] + x;
    sum_x += x;
    if (d == n_v) {
      any_full = true;
    } else {
      loc.push_back(Q);
      sum_local_x += x;
    }
  }
  M.bytes = 0;
  upload(M.cx_off, cxo, st);
  M.Cx.alloc((size_t)std::max<int64_t>(sum_x, 16) * 2);
  M.n_local = (int)loc.size();
  upload(M.pairs_local, loc, st);
  M.bytes += (int64_t)M.Cx.bytes;
  if (any_full) {
    const int64_t rows = n_chol * n_v;
    M.Lx.alloc((size_t)rows * 2 * nvp * 2);
    if (P.lvv_f32)
      hipLaunchKernelGGL(split_lvv_kernel<float>, dim3(blocks_for(rows * (nvp / 8), 65536)), dim3(256), 0, st,
                         static_cast<const float*>(P.Lvv.p), rows, n_v, M.Lx.as<uint16_t>(), sc, f16);
    else
      hipLaunchKernelGGL(split_lvv_kernel<double>, dim3(blocks_for(rows * (nvp / 8), 65536)), dim3(256), 0, st,
                         static_cast<const double*>(P.Lvv.p), rows, n_v, M.Lx.as<uint16_t>(), sc, f16);
    TCI_HIP(hipGetLastError());
    M.bytes += (int64_t)M.Lx.bytes;
  }
  // local-domain L^J[D,D] for all J packed once when it fits 64 GB; otherwise
  // packed per chunk of Cholesky vectors inside every sigma
  M.prepacked = sum_local_x > 0 && 2.0 * (double)sum_local_x * (double)n_chol <= 64.0e9;
  std::vector<int64_t> lpo_all(loc.size() + 1, 0);
  if (M.prepacked) {
    for (size_t q = 0; q < loc.size(); ++q) {
      const int64_t d = P.h_dom_off[loc[q] + 1] - P.h_dom_off[loc[q]];
      lpo_all[q + 1] = lpo_all[q] + n_chol * d * 2 * xpad(d);
    }
    upload(M.lp_off_all, lpo_all, st);
    // 16 KiB tail: the fused kernel for 129-192 reads rows past a block's end
    M.Lpx.alloc((size_t)lpo_all.back() * 2 + 16384);
    M.bytes += (int64_t)M.Lpx.bytes;
    if (P.lvv_f32)
      hipLaunchKernelGGL(pack_L_split_kernel<float>, dim3(4096, (unsigned)loc.size()), dim3(256), 0, st, P.tab,
                         M.pairs_local.as<int32_t>(), M.lp_off_all.as<int64_t>(), static_cast<const float*>(P.Lvv.p),
                         (int64_t)0, (int)n_chol, M.Lpx.as<uint16_t>(), sc, f16);
    else
      hipLaunchKernelGGL(pack_L_split_kernel<double>, dim3(4096, (unsigned)loc.size()), dim3(256), 0, st, P.tab,
                         M.pairs_local.as<int32_t>(), M.lp_off_all.as<int64_t>(), static_cast<const double*>(P.Lvv.p),
                         (int64_t)0, (int)n_chol, M.Lpx.as<uint16_t>(), sc, f16);
    TCI_HIP(hipGetLastError());
  }
  // chunk length: T' (all pairs) + packed L (local pairs, unless prepacked), 2 bytes per element
  const double budget = 8.0e9;
  const double per_J = 2.0 * (double)(sum_x + (M.prepacked ? 0 : sum_local_x));
  int64_t Lc = std::max<int64_t>(1, std::min<int64_t>(n_chol, (int64_t)(budget / std::max(per_J, 1.0))));
  for (int64_t q = Lc; q >= std::max<int64_t>(1, Lc / 2); --q)
    if (n_chol % q == 0) {
      Lc = q;
      break;
    }
  if (const char* e = std::getenv("TCI_SPLIT_CHUNK")) {  // Cholesky vectors per chunk, for sweeps
    const long long v = std::atoll(e);
    if (v > 0) Lc = std::min<int64_t>(n_chol, (int64_t)std::min<long long>(v, (long long)Lc));
  }
  M.Tx.alloc((size_t)Lc * sum_x * 2);
  // padding columns of T' are never written by phase 1 and must read as zero
  launch_fill(M.Tx.as<double>(), (int64_t)(M.Tx.bytes / 8), 0.0, st);
  M.bytes += (int64_t)M.Tx.bytes;
  if (sum_local_x && !M.prepacked) {
    M.Lpx.alloc((size_t)Lc * sum_local_x * 2);
    M.bytes += (int64_t)M.Lpx.bytes;
  }
  {
    const char* env = std::getenv("TCI_LADDER_MERGED");
    M.merged = loc.empty() && n_pairs > 0 && !(env && env[0] == '0');
    const char* pp = std::getenv("TCI_LADDER_PER_PAIR");
    M.per_pair = pp && pp[0] == '1';
    if (M.per_pair) M.merged = false;
  }
  // programs for a K-split fill target: phase 2 of each tile shape is split
  // along K until it has about `target` workgroups (at most 16 splits);
  // big: phase 2 of domains of 129-192 on single 192 x 192 tiles
  // fused: walks of 129-192 (prepacked) run the fused kernel over all J
  // instead of the two phases, with about fused_target workgroups
  // fused2_target > 0: walks of 65-128 (prepacked) run the whole-walk fused
  // kernel, split along J into about fused2_target workgroups
  auto fusable = [&](int64_t d) { return M.prepacked && d > 128 && d <= 192 && d != n_v; };
  auto fusable2 = [&](int64_t d) { return M.prepacked && d > 64 && d <= 128 && d != n_v; };
  // fused3_target > 0: walks of 129-192 (prepacked) run the whole-walk kernel
  // with C in LDS (before the C-panel kernel, which then gets none of them)
  auto fusable3 = [&](int64_t d) { return M.prepacked && d > 128 && d <= 192 && d != n_v; };
  auto build = [&](int64_t target, bool big, bool fused, int64_t fused_target, int64_t fused2_target,
                   int64_t fused3_target) {
    const int tile2 = 128;
    const int64_t tiles2m = cdiv((int64_t)n_pairs * n_v, tile2) * cdiv(n_v, tile2);
    M.rpart_n = (int)std::max<int64_t>(1, std::min<int64_t>(16, cdiv(target, std::max<int64_t>(tiles2m, 1))));
    // per-walk tile shapes (local domains): phase 2 of walk Q runs on tiles
    // t2m[v2] x t2n[v2]; every shape gets enough K-splits to fill the chip in its
    // own launch
    static const int t2m[4] = {64, 128, 128, 192}, t2n[4] = {64, 128, 192, 192};
    std::vector<int> v2of(n_pairs, 1);
    int nsplit_v[4] = {1, 1, 1, 1};
    int nsf = 1, nsf2 = 1, nsf3 = 1;
    if (!M.merged) {
      int64_t tv[4] = {0, 0, 0, 0}, tf = 0, tf2 = 0, tf3 = 0;
      for (int Q = 0; Q < n_pairs; ++Q) {
        const int64_t d = P.h_dom_off[Q + 1] - P.h_dom_off[Q];
        if (fused3_target > 0 && fusable3(d)) {
          ++tf3;
          continue;
        }
        if (fused && fusable(d)) {
          tf += cdiv(d, 64);
          continue;
        }
        if (fused2_target > 0 && fusable2(d)) {
          ++tf2;
          continue;
        }
        int v = d <= 64 ? 0 : (d <= 128 ? 1 : ((double)d / (192.0 * cdiv(d, 192)) >= (double)d / (128.0 * cdiv(d, 128)) ? 2 : 1));
        if (big && d > 128 && d <= 192) v = 3;
        v2of[Q] = v;
        tv[v] += cdiv(d, t2m[v]) * cdiv(d, t2n[v]);
      }
      M.rpart_n = 1;
      for (int v = 0; v < 4; ++v) {
        nsplit_v[v] = tv[v] ? (int)std::max<int64_t>(1, std::min<int64_t>(16, cdiv(target, tv[v]))) : 1;
        M.rpart_n = std::max(M.rpart_n, nsplit_v[v]);
      }
      if (tf) {  // each (panel, J range) of a fused walk writes its own partial-R slice
        nsf = (int)std::max<int64_t>(1, std::min<int64_t>(5, cdiv(fused_target, tf)));
        M.rpart_n = std::max(M.rpart_n, 3 * nsf);
      }
      if (tf2) {  // each J range of a whole-walk fused task writes its own slice
        nsf2 = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(16, n_chol), cdiv(fused2_target, tf2)));
        M.rpart_n = std::max(M.rpart_n, nsf2);
      }
      if (tf3) {
        nsf3 = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(16, n_chol), cdiv(fused3_target, tf3)));
        M.rpart_n = std::max(M.rpart_n, nsf3);
      }
    }
    if (M.rpart_n > 1 && M.Rpart.bytes < (size_t)M.rpart_n * P.sum_d2 * 8) {
      M.bytes -= (int64_t)M.Rpart.bytes;
      M.Rpart.alloc((size_t)M.rpart_n * P.sum_d2 * 8);
      M.bytes += (int64_t)M.Rpart.bytes;
    }
    M.phase.clear();
    M.chunks.clear();
    M.fused_h.clear();
    M.fused2_h.clear();
    M.fused3_h.clear();
    std::vector<int64_t> lens;
    for (int64_t J0 = 0; J0 < n_chol; J0 += Lc) {
      const int64_t len = std::min(Lc, n_chol - J0);
      int idx = -1;
      for (size_t q = 0; q < lens.size(); ++q)
        if (lens[q] == len) idx = (int)q;
      if (idx < 0) {
        idx = (int)lens.size();
        lens.push_back(len);
      }
      M.chunks.push_back({J0, idx});
    }
    const uint16_t* Lx = M.Lx.as<uint16_t>();
    const uint16_t* Cx = M.Cx.as<uint16_t>();
    uint16_t* Tx = M.Tx.as<uint16_t>();
    const int64_t jstride_full = (int64_t)n_v * 2 * nvp;
    for (int64_t len : lens) {
      M.phase.emplace_back();
      tci_plan::LadderPhase& ph = M.phase.back();
      ph.Lc = len;
      ph.nsplit = (int)std::min<int64_t>(M.rpart_n, len);
      std::vector<int64_t> lpo(loc.size() + 1, 0);
      for (size_t q = 0; q < loc.size(); ++q) {
        const int64_t d = P.h_dom_off[loc[q] + 1] - P.h_dom_off[loc[q]];
        lpo[q + 1] = lpo[q] + len * d * 2 * xpad(d);
      }
      upload(ph.lp_off, lpo, st);
      ph.p1.clear();
      ph.p2.clear();
      ph.p1.split = ph.p2.split = true;
      ph.p1.tile = 128;
      ph.p2.tile = tile2;
      if (M.merged) {
        // every walk has the full domain: one product per phase over all walks,
        //   phase 1 T''[(jj,a)][(Q,c)] = sum_k L^J[a][k] C_Q[k][c]   (B = rows of C_Q^T),
        //            stored as T''[jj][(Q,a)][c]
        //   phase 2 R[(Q,a)][n] += sum_jj sum_c T''[jj][(Q,a)][c] L^J[n][c]
        const int d = n_v, NP = n_pairs * n_v;
        const int64_t ldx = 2 * nvp, Sjj = (int64_t)NP * ldx;
        Blocking bk;
        bk.rdiv = d;
        bk.rstride = Sjj;
        bk.cdiv = d;
        bk.cstride = (int64_t)d * ldx;
        ph.p1.add(reinterpret_cast<double*>(Tx), ldx, (int)(len * d), NP,
                  {xseg(xop(Lx, ldx, jstride_full), xop(Cx, ldx), d)}, false, true, 0, bk, 3);
        for (int sp = 0; sp < ph.nsplit; ++sp) {
          const int64_t j0 = len * sp / ph.nsplit, j1 = len * (sp + 1) / ph.nsplit;
          std::vector<Seg> sg;
          for (int64_t jj = j0; jj < j1; ++jj)
            sg.push_back(xseg(xop(Tx + jj * Sjj, ldx), xop(Lx + jj * jstride_full, ldx, jstride_full), d));
          double* Rout = M.rpart_n > 1 ? M.Rpart.as<double>() + (int64_t)sp * P.sum_d2 : P.R.as<double>();
          ph.p2.add(Rout, d, NP, d, sg, true, false, 0, Blocking(), 4);
        }
        ph.p1.upload_tables(st);
        ph.p2.upload_tables(st);
        continue;
      }
      // per-walk tile shapes: phase 1 (M = len d rows) 128 x {64,128,192} and
      // phase 2 (d x d) {64x64, 128x128, 128x192}, chosen so one tile spans the
      // short side of the domain where possible (each operand panel read once)
      ph.p1x.clear();
      ph.p2x.clear();
      ph.pp1.clear();
      ph.pp2.clear();
      ph.p1x.resize(3);
      ph.p2x.resize(4);
      static const int tn1[3] = {64, 128, 192};
      for (int v = 0; v < 3; ++v) {
        ph.p1x[v].split = true;
        ph.p1x[v].tile = 128;
        ph.p1x[v].tile_n = tn1[v];
      }
      for (int v = 0; v < 4; ++v) {
        ph.p2x[v].split = true;
        ph.p2x[v].tile = t2m[v];
        ph.p2x[v].tile_n = t2n[v];
      }
      auto util = [](int64_t d, int t) { return (double)d / (double)(t * cdiv(d, t)); };
      int64_t toff = 0;
      int qloc = 0;
      for (int Q = 0; Q < n_pairs; ++Q) {
        const int d = (int)(P.h_dom_off[Q + 1] - P.h_dom_off[Q]);
        const int64_t ldx = 2 * xpad(d);
        const bool full = d == n_v;
        // L operand: rows (jj, a), one jj-block every jstep elements
        const uint16_t* Lb;
        int64_t jstep, jstride;
        if (full) {
          Lb = Lx;
          jstep = jstride_full;
          jstride = jstride_full;
        } else if (M.prepacked) {  // [q][J][r][k]: chunk start through jstride
          Lb = M.Lpx.as<uint16_t>() + lpo_all[qloc++];
          jstep = (int64_t)d * ldx;
          jstride = jstep;
        } else {
          Lb = M.Lpx.as<uint16_t>() + lpo[qloc++];
          jstep = (int64_t)d * ldx;
          jstride = 0;
        }
        if (fused3_target > 0 && fusable3(d)) {
          if (&ph == &M.phase.front())
            for (int sp = 0; sp < nsf3; ++sp) {
              FusedTask f{};
              f.L = Lb;
              f.C = Cx + cxo[Q];
              f.R = (M.rpart_n > 1 ? M.Rpart.as<double>() + (int64_t)sp * P.sum_d2 : P.R.as<double>()) +
                    P.h_cd_off[Q];
              f.jstep = jstep;
              f.d = d;
              f.ldx = (int32_t)ldx;
              f.j0 = (int32_t)(n_chol * sp / nsf3);
              f.j1 = (int32_t)(n_chol * (sp + 1) / nsf3);
              M.fused3_h.push_back(f);
            }
          toff += len * d * ldx;
          continue;
        }
        if (fused && fusable(d)) {
          // every J at once (prepacked), so only the first chunk length's
          // phase builds the tasks
          if (&ph == &M.phase.front())
            for (int sp = 0; sp < nsf; ++sp)
              for (int a0 = 0; a0 < d; a0 += 64) {  // the panels of one J range next to each other
                FusedTask f{};
                f.L = Lb;
                f.C = Cx + cxo[Q];
                f.R = M.Rpart.as<double>() + (int64_t)((a0 / 64) * nsf + sp) * P.sum_d2 + P.h_cd_off[Q];
                f.jstep = jstep;
                f.d = d;
                f.ldx = (int32_t)ldx;
                f.a0 = a0;
                f.j0 = (int32_t)(n_chol * sp / nsf);
                f.j1 = (int32_t)(n_chol * (sp + 1) / nsf);
                M.fused_h.push_back(f);
              }
          toff += len * d * ldx;
          continue;
        }
        if (fused2_target > 0 && fusable2(d)) {
          if (&ph == &M.phase.front())
            for (int sp = 0; sp < nsf2; ++sp) {
              FusedTask f{};
              f.L = Lb;
              f.C = Cx + cxo[Q];
              f.R = (M.rpart_n > 1 ? M.Rpart.as<double>() + (int64_t)sp * P.sum_d2 : P.R.as<double>()) +
                    P.h_cd_off[Q];
              f.jstep = jstep;
              f.d = d;
              f.ldx = (int32_t)ldx;
              f.j0 = (int32_t)(n_chol * sp / nsf2);
              f.j1 = (int32_t)(n_chol * (sp + 1) / nsf2);
              M.fused2_h.push_back(f);
            }
          toff += len * d * ldx;
          continue;
        }
        uint16_t* Tq = Tx + toff;
        int v1 = 0;
        for (int v = 1; v < 3; ++v)
          if (util(d, tn1[v]) >= util(d, tn1[v1])) v1 = v;
        const int v2 = v2of[Q];
        const int nsp = (int)std::min<int64_t>(nsplit_v[v2], len);
        // one program per walk and phase instead (a launch per walk) with
        // TCI_LADDER_PER_PAIR=1, to compare against the grouped launches
        GemmProgram* g1 = &ph.p1x[v1];
        GemmProgram* g2 = &ph.p2x[v2];
        if (M.per_pair) {
          ph.pp1.emplace_back();
          ph.pp2.emplace_back();
          g1 = &ph.pp1.back();
          g2 = &ph.pp2.back();
          g1->split = g2->split = true;
          g1->tile = 128;
          g1->tile_n = tn1[v1];
          g2->tile = ph.p2x[v2].tile;
          g2->tile_n = ph.p2x[v2].tile_n;
        }
        g1->add(reinterpret_cast<double*>(Tq), ldx, (int)(len * d), d,
                {xseg(xop(Lb, ldx, jstride), xop(Cx + cxo[Q], ldx), d)}, false, true, 0, Blocking(), 3);
        for (int sp = 0; sp < nsp; ++sp) {
          const int64_t j0 = len * sp / nsp, j1 = len * (sp + 1) / nsp;
          std::vector<Seg> sg;
          for (int64_t jj = j0; jj < j1; ++jj)
            sg.push_back(xseg(xop(Lb + jj * jstep, ldx, jstride), xop(Tq + jj * d * ldx, ldx), d));
          double* Rout = (M.rpart_n > 1 ? M.Rpart.as<double>() + (int64_t)sp * P.sum_d2 : P.R.as<double>()) +
                         P.h_cd_off[Q];
          g2->add(Rout, d, d, d, sg, true, false, 0, Blocking(), 4);
        }
        toff += len * d * ldx;
      }
      for (auto& g : ph.p1x) g.upload_tables(st);
      for (auto& g : ph.p2x) g.upload_tables(st);
      for (auto& g : ph.pp1) g.upload_tables(st);
      for (auto& g : ph.pp2) g.upload_tables(st);
    }
    M.fused = !M.fused_h.empty();
    if (M.fused) upload(M.fused_d, M.fused_h, st);
    // largest domains first, so the smallest tasks fill the tail of the launch
    std::stable_sort(M.fused2_h.begin(), M.fused2_h.end(),
                     [](const FusedTask& x, const FusedTask& y) { return x.d > y.d; });
    M.fused2 = !M.fused2_h.empty();
    if (M.fused2) upload(M.fused2_d, M.fused2_h, st);
    std::stable_sort(M.fused3_h.begin(), M.fused3_h.end(),
                     [](const FusedTask& x, const FusedTask& y) { return x.d > y.d; });
    M.fused3 = !M.fused3_h.empty();
    if (M.fused3) upload(M.fused3_d, M.fused3_h, st);
  };
  const int64_t def_target = 512;
  auto progs = [&](tci_plan::LadderPhase& ph) {
    std::vector<GemmProgram*> v;
    if (M.merged) {
      v.push_back(&ph.p1);
      v.push_back(&ph.p2);
    }
    for (auto& g : ph.p1x) v.push_back(&g);
    for (auto& g : ph.p2x) v.push_back(&g);
    return v;
  };
  // phase 2 of the domains of 129-192 on 192 x 192 tiles: register-staged or
  // fed by direct global-to-LDS loads (three stages)
  auto set_dma = [&](bool on) {
    for (auto& ph : M.phase)
      if (ph.p2x.size() > 3) ph.p2x[3].dma = on && ph.p2x[3].tile == 192 ? 3 : 0;
  };
  // phase 1 of the domains of 129-192 (128 x 192 tiles): direct-load stages (0: off, 2 or 3)
  auto set_dma1 = [&](int nst) {
    for (auto& ph : M.phase)
      if (ph.p1x.size() > 2 && ph.p1x[2].tile == 128 && ph.p1x[2].tile_n == 192) ph.p1x[2].dma = nst;
  };
  if (M.per_pair || !autotune_enabled()) {
    // untuned: 128-row phase-2 tiles unless TCI_SPLIT_P2_192=1
    const char* e = std::getenv("TCI_SPLIT_P2_192");
    const char* f = std::getenv("TCI_SPLIT_FUSED");
    const char* f2 = std::getenv("TCI_SPLIT_FUSED128");
    const char* f3 = std::getenv("TCI_SPLIT_FUSED192");
    build(def_target, e && e[0] == '1', f && f[0] == '1', 1024, f2 && (f2[0] == '1' || f2[0] == '8') ? 1024 : 0,
          f3 && f3[0] == '1' ? 1024 : 0);
    M.fused2_waves = f2 && f2[0] == '8' ? 8 : 4;  // TCI_SPLIT_FUSED128=8: eight-wave workgroups
    const char* fd = std::getenv("TCI_SPLIT_DMA");
    set_dma(fd && fd[0] == '1');
    const char* fd1 = std::getenv("TCI_SPLIT_DMA1");
    set_dma1(fd1 && (fd1[0] == '2' || fd1[0] == '3') ? fd1[0] - '0' : 0);
  } else if (M.tuned_target > 0) {
    // tuned before for the other split precision: same schedule, so one plan
    // gives the same sigma for a precision every time it is used
    build(M.tuned_target, M.tuned_big, M.tuned_fused, M.tuned_fused_target, M.tuned_fused2_target,
          M.tuned_fused3_target);
    M.fused_remap = M.tuned_fused_order;
    M.fused2_remap = M.tuned_fused2_order;
    M.fused2_waves = M.tuned_fused2_waves;
    M.fused3_remap = M.tuned_fused3_order;
    set_dma(M.tuned_dma);
    set_dma1(M.tuned_dma1);
    for (auto& ph : M.phase) {
      std::vector<GemmProgram*> g = progs(ph);
      for (size_t k = 0; k < g.size() && k < M.tuned_orders.size(); ++k) g[k]->remap = M.tuned_orders[k];
    }
  } else {
    // autotuning: time the first chunk of the ladder for several fill
    // targets and phase-2 tile shapes, keep the fastest, then pick each
    // program's workgroup order
    bool any_big = false;
    for (int Q = 0; Q < n_pairs; ++Q) {
      const int64_t d = P.h_dom_off[Q + 1] - P.h_dom_off[Q];
      any_big = any_big || (d > 128 && d <= 192 && d != n_v);
    }
    const int64_t cand[4] = {512, 1024, 2048, 4096};
    double best = 1e300;
    int64_t bt = def_target;
    bool bb = false;
    std::string log = "split ladder: (K-split target, 192-tiles)";
    for (int big = 0; big <= (any_big && !M.merged ? 1 : 0); ++big)
      for (int64_t t : cand) {
        build(t, big != 0, false, 0, 0, 0);
        TCI_HIP(hipStreamSynchronize(st));
        const double ms = time_ms(st, 2, [&] { mixed_ladder_chunk(P, precision, M.chunks[0], st); });
        log += " " + std::to_string(t) + (big ? "b" : "") + ":" + fmt_ms(ms);
        if (ms < best * 0.98) {
          best = ms;
          bt = t;
          bb = big != 0;
        }
      }
    build(bt, bb, false, 0, 0, 0);
    TCI_HIP(hipStreamSynchronize(st));
    log += " -> " + std::to_string(bt) + (bb ? "b" : "");
    bool bdma = false;
    if (bb) {  // the 192 x 192 phase-2 tiles with direct-load stages
      set_dma(true);
      const double md = time_ms(st, 2, [&] { mixed_ladder_chunk(P, precision, M.chunks[0], st); });
      log += "; direct-load stages:" + fmt_ms(md);
      bdma = md < best * 0.98;
      if (bdma) best = md;
      set_dma(bdma);
      log += bdma ? " -> on" : " -> off";
    }
    int bdma1 = 0;  // then direct-load stages for the 128 x 192 phase-1 tiles
    for (int nst : {2, 3}) {
      set_dma1(nst);
      const double m1 = time_ms(st, 2, [&] { mixed_ladder_chunk(P, precision, M.chunks[0], st); });
      log += "; phase-1 direct loads " + std::to_string(nst) + ":" + fmt_ms(m1);
      if (m1 < best * 0.98) {
        best = m1;
        bdma1 = nst;
      }
    }
    set_dma1(bdma1);
    // then the fused kernel for the walks of 129-192 (whole ladder timed: it
    // runs over every Cholesky vector in one launch)
    bool bf = false;
    int64_t bft = 0;
    bool any_fusable = false;
    for (int Q = 0; Q < n_pairs && !M.merged; ++Q) any_fusable = any_fusable || fusable(P.h_dom_off[Q + 1] - P.h_dom_off[Q]);
    if (any_fusable) {
      const double t2 = time_ms(st, 2, [&] { mixed_ladder_run(P, precision, st); });
      log += "; whole two-phase:" + fmt_ms(t2);
      double tb = t2;
      int bfo = 1;
      for (int64_t ft : {512, 1024, 2048})
        for (int fo : {1, 0}) {
          build(bt, bb, true, ft, 0, 0);
          set_dma(bdma);
          set_dma1(bdma1);
          M.fused_remap = fo;
          TCI_HIP(hipStreamSynchronize(st));
          const double tf = time_ms(st, 2, [&] { mixed_ladder_run(P, precision, st); });
          log += " fused" + std::to_string(ft) + (fo ? "x" : "") + ":" + fmt_ms(tf);
          if (tf < tb * 0.98) {
            tb = tf;
            bf = true;
            bft = ft;
            bfo = fo;
          }
        }
      M.tuned_fused_order = bfo;
      build(bt, bb, bf, bft, 0, 0);
      set_dma(bdma);
      set_dma1(bdma1);
      M.fused_remap = bfo;
      TCI_HIP(hipStreamSynchronize(st));
      log += bf ? " -> fused" + std::to_string(bft) + (bfo ? "x" : "") : " -> two-phase";
    }
    // then the whole-walk fused kernel for the walks of 65-128
    int64_t bft2 = 0;
    int bfo2 = 0, bnw2 = 4;
    bool any_fusable2 = false;
    for (int Q = 0; Q < n_pairs && !M.merged; ++Q) any_fusable2 = any_fusable2 || fusable2(P.h_dom_off[Q + 1] - P.h_dom_off[Q]);
    if (any_fusable2) {
      double tb = time_ms(st, 2, [&] { mixed_ladder_run(P, precision, st); });
      log += "; 65-128 two-phase:" + fmt_ms(tb);
      // eight-wave workgroups (TCI_SPLIT_FUSED128=8) measured 2% slower at
      // 34/192/1200 and are not a candidate
      for (int64_t ft : {512, 1024, 2048})
        for (int fo : {0, 1})
          for (int nw : {4}) {
            build(bt, bb, bf, bft, ft, 0);
            set_dma(bdma);
            set_dma1(bdma1);
            M.fused2_remap = fo;
            M.fused2_waves = nw;
            TCI_HIP(hipStreamSynchronize(st));
            const double tf = time_ms(st, 2, [&] { mixed_ladder_run(P, precision, st); });
            log += " fused128-" + std::to_string(ft) + (fo ? "x" : "") + (nw == 8 ? "w8" : "") + ":" + fmt_ms(tf);
            if (tf < tb * 0.98) {
              tb = tf;
              bft2 = ft;
              bfo2 = fo;
              bnw2 = nw;
            }
          }
      build(bt, bb, bf, bft, bft2, 0);
      set_dma(bdma);
      set_dma1(bdma1);
      M.fused2_remap = bfo2;
      M.fused2_waves = bnw2;
      TCI_HIP(hipStreamSynchronize(st));
      log += bft2 ? " -> fused128-" + std::to_string(bft2) + (bfo2 ? "x" : "") + (bnw2 == 8 ? "w8" : "") : " -> two-phase";
    }
    M.tuned_fused2_target = bft2;
    M.tuned_fused2_order = bfo2;
    M.tuned_fused2_waves = bnw2;
    // the whole-walk kernel with C in LDS for the walks of 129-192
    // (TCI_SPLIT_FUSED192) is not a candidate: at 34/192/1200 it measured 15%
    // slower than the two phases for those walks
    M.tuned_fused3_target = 0;
    M.tuned_fused3_order = 0;
    log += "; order";
    const auto& ch = M.chunks[0];
    std::vector<GemmProgram*> g0 = progs(M.phase[ch.second]);
    const float* scl = M.scales.as<float>();
    M.tuned_orders.assign(g0.size(), -1);
    for (size_t k = 0; k < g0.size(); ++k) {
      if (g0[k]->tasks.empty()) continue;
      const int o = tune_order(*g0[k], st, [&] { g0[k]->launch_split(precision, st, ch.first, scl); });
      log += " " + std::to_string(o);
      M.tuned_orders[k] = o;
      for (auto& ph : M.phase) progs(ph)[k]->remap = o;  // same role in the other chunk lengths
    }
    M.tuned_target = bt;
    M.tuned_big = bb;
    M.tuned_dma = bdma;
    M.tuned_dma1 = bdma1;
    M.tuned_fused = bf;
    M.tuned_fused_target = bft;
    P.tuning += log + "\n";
  }
  if (P.dense_pairs) {
    const float* scl = M.scales.as<float>();
    auto pair_programs = [&] {
      std::vector<GemmProgram*> gs;
      if (M.pairs_merged) gs = {&M.pairsUWx, &M.pairsx};
      else gs = {&M.progUx, &M.ringZx, &M.ringZ3x, &M.progWx, &M.holeGx};
      return gs;
    };
    auto run_pairs = [&] {
      for (GemmProgram* g : pair_programs())
        if (!g->tasks.empty()) g->launch_split(precision, st, 0, scl);
    };
    const char* te = std::getenv("TCI_PAIRS_TILE");
    if (autotune_enabled() && !te) {
      // tiles of the split pair products (128 x 128, or 256 x 256 on 32 x 32
      // MFMAs with 4, 8 or 16 waves): Z, Z3 and the 4-internal product first,
      // then U and W; then their tile order
      std::string log = (precision == TCI_SIGMA_FP16X3 ? "fp16x3" : "bf16x3") + std::string(" pair products: tile");
      static const char* cname[6] = {"128", "256w4", "256w8", "256w16", "128x192d2", "128d2"};
      const int64_t bytes0 = M.bytes;  // each build counts its split copies again
      auto trial = [&](int cz, int cu) {
        M.bytes = bytes0;
        dense_pairs_build_split(P, st, cz, cu);
        return time_ms(st, 2, run_pairs);
      };
      double best = 1e300;
      int bz = 0, bu = 0;
      if (M.tuned_pairs_z >= 0) {  // tuned for the other split precision: same tiles, same sums
        bz = M.tuned_pairs_z;
        bu = M.tuned_pairs_uw;
        log += " kept";
        trial(bz, bu);
      } else {
        for (int cz = 0; cz < 6; ++cz) {
          const double ms = trial(cz, 0);
          log += std::string(" z") + cname[cz] + ":" + fmt_ms(ms);
          if (ms < best * 0.98) {
            best = ms;
            bz = cz;
          }
        }
        if (M.z_lowrank) {  // U and W go out as a launch of their own
          for (int cu : {2, 4, 5}) {
            const double ms = trial(bz, cu);
            log += std::string(" uw") + cname[cu] + ":" + fmt_ms(ms);
            if (ms < best * 0.98) {
              best = ms;
              bu = cu;
            }
          }
        } else {
          bu = bz;
        }
        if (!(M.z_lowrank && bu == 5)) trial(bz, bu);  // unless the last trial built the chosen tiles
        M.tuned_pairs_z = bz;
        M.tuned_pairs_uw = bu;
      }
      log += std::string(" -> z") + cname[bz] + " uw" + cname[bu] + "; order";
      for (GemmProgram* g : pair_programs())
        if (!g->tasks.empty()) log += " " + tune_block_order(*g, st, [&] { g->launch_split(precision, st, 0, scl); });
      P.tuning += log + "\n";
    } else {
      dense_pairs_build_split(P, st);
    }
  }
  // ladder from the resident (ab|cd) block when the plan holds it: timed
  // against the split Cholesky-direct ladder above (kept when faster) unless
  // autotuning is off or TCI_LADDER_VV=1; the loser's operand copies are freed
  P.vv.split = false;
  if (P.vv.built) {
    const char* env = std::getenv("TCI_LADDER_VV");
    const bool force = env && env[0] == '1';
    const std::string pn = precision == TCI_SIGMA_FP16X3 ? "fp16x3" : "bf16x3";
    vv_prepare_split(P, precision, st);
    std::string log = pn + " ladder (ab|cd) resident:";
    bool use = true;
    if (P.vv.split_use >= 0) {  // decided for the other split precision: same schedule, same sums
      use = P.vv.split_use == 1;
      if (use && P.vv.split_tile > 0) {
        vv_set_schedule(P, precision, P.vv.split_tile, P.vv.split_tile_n, P.vv.split_dma, P.vv.split_ks, st,
                        P.vv.split_trans);
        P.vv.gx.remap = P.vv.split_order;
      }
      log += " kept";
    } else if (autotune_enabled()) {
      const double tv = vv_tune(P, precision, st, log);
      if (!force) {
        const double tc = time_ms(st, 1, [&] { mixed_ladder_run(P, precision, st); });
        log += "; resident " + fmt_ms(tv) + ", Cholesky-direct " + fmt_ms(tc);
        use = tv < tc;
      }
      P.vv.split_use = use ? 1 : 0;
    }
    P.vv.split = use;
    P.tuning += log + (use ? " -> resident\n" : " -> Cholesky-direct\n");
    TCI_HIP(hipStreamSynchronize(st));
    if (use) {
      for (DevBuf* b : {&M.Lx, &M.Lpx, &M.Tx, &M.Rpart}) {
        M.bytes -= (int64_t)b->bytes;
        b->reset();
      }
    } else {
      for (DevBuf* b : {&P.vv.Vpx, &P.vv.Vmx, &P.vv.Xpx, &P.vv.Xmx}) b->reset();
      P.vv.prec = -1;
    }
  }
  TCI_HIP(hipStreamSynchronize(st));
  M.ready = true;
}

void mixed_ladder_run(tci_plan& P, int precision, hipStream_t st) {
  tci_plan::MixedLadder& M = P.mx;
  const int f16 = precision == TCI_SIGMA_FP16X3 ? 1 : 0;
  float* sc = M.scales.as<float>();
  if (f16) {  // scale of C and of the intermediate from max|C|
    unsigned* mx = M.maxbits.as<unsigned>();
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, st, mx, 2);
    hipLaunchKernelGGL(absmax_kernel<double>, dim3(blocks_for(P.sum_d2, 8192)), dim3(256), 0, st, P.C.as<double>(),
                       P.sum_d2, mx);
    hipLaunchKernelGGL(set_scales_kernel, dim3(1), dim3(64), 0, st, 1, mx, sc, (float)P.dmax);
    TCI_HIP(hipGetLastError());
  }
  if (M.merged) {
    const int nt = (P.dmax + 31) / 32;
    hipLaunchKernelGGL(split_blocks_T_kernel, dim3(nt * nt, P.n_pairs), dim3(256), 0, st, P.tab,
                       M.cx_off.as<int64_t>(), P.C.as<double>(), M.Cx.as<uint16_t>(), sc + 1, f16);
  } else {
    hipLaunchKernelGGL(split_blocks_kernel, dim3(blocks_for((int64_t)P.dmax * ((P.dmax + 7) / 8), 256), P.n_pairs),
                       dim3(256), 0, st, P.tab, M.cx_off.as<int64_t>(), P.C.as<double>(), M.Cx.as<uint16_t>(), sc + 1,
                       f16, 1, 0);
  }
  TCI_HIP(hipGetLastError());
  double* Rz = M.rpart_n > 1 ? M.Rpart.as<double>() : P.R.as<double>();
  launch_fill(Rz, (int64_t)M.rpart_n * P.sum_d2, 0.0, st);
  // algorithmic flops of a fused task list: 4 d^2 (panel width) per Cholesky vector
  auto fflops = [](const std::vector<FusedTask>& v, bool panels) {
    double f = 0.0;
    for (const FusedTask& t : v)
      f += 4.0 * t.d * t.d * (panels ? std::min(64, t.d - t.a0) : t.d) * (double)(t.j1 - t.j0);
    return f;
  };
  const std::string pn = precision == TCI_SIGMA_FP16X3 ? "fp16x3" : "bf16x3";
  if (M.fused) {
    const int e0 = P.kt_begin(st);
    launch_xladder_fused(precision, M.fused_d.as<FusedTask>(), (int64_t)M.fused_h.size(), sc, st,
                         M.fused_remap < 0 || xcd_remap_forced() ? xcd_remap_default() : M.fused_remap);
    if (e0 >= 0) P.kt_end(e0, pn + " ladder xladder_fused_kernel (129-192 panels)", fflops(M.fused_h, true), precision, st);
  }
  if (M.fused2) {
    const int e0 = P.kt_begin(st);
    launch_xladder_fused128(precision, M.fused2_d.as<FusedTask>(), (int64_t)M.fused2_h.size(), sc, st,
                            xcd_remap_forced() ? xcd_remap_default() : M.fused2_remap, M.fused2_waves);
    if (e0 >= 0) P.kt_end(e0, pn + " ladder xladder_fused128_kernel", fflops(M.fused2_h, false), precision, st);
  }
  if (M.fused3) {
    const int e0 = P.kt_begin(st);
    launch_xladder_fused192(precision, M.fused3_d.as<FusedTask>(), (int64_t)M.fused3_h.size(), sc, st,
                            xcd_remap_forced() ? xcd_remap_default() : M.fused3_remap);
    if (e0 >= 0) P.kt_end(e0, pn + " ladder xladder_fused192_kernel", fflops(M.fused3_h, false), precision, st);
  }
  for (const auto& ch : M.chunks) mixed_ladder_chunk(P, precision, ch, st);
  if (P.timing()) TCI_HIP(hipEventRecord(P.ev[3], st));
  if (M.rpart_n > 1) ladder_reduce(M.Rpart.as<double>(), P.sum_d2, M.rpart_n, P.R.as<double>(), P.sum_d2, st);
}

void mixed_ladder(tci_plan& P, int precision, hipStream_t st) {
  mixed_prepare(P, precision, st);
  if (P.vv.split) {
    vv_ladder(P, precision, st);
    if (P.timing()) TCI_HIP(hipEventRecord(P.ev[3], st));
    return;
  }
  mixed_ladder_run(P, precision, st);
}

// ---------------------------------------------------------------------------
// Batched split ladder: nvec vectors in one two-phase pass, so every L^J block
// is read once per batch instead of once per vector.
//   phase 1 T'[(jj,a)][(v,c)] = sum_k L^J[a][k] C_v[c][k], stored as
//           T'[jj][v][a][c] (blocked output: rows by jj, columns by v)
//   phase 2 R_v[n][a] += sum_jj sum_c L^J[n][c] T'[jj][v][a][c] (columns (v,a)
//           blocked into the R_v arrays)
// Needs per-walk programs and every L^J[D,D] resident (prepacked or the full
// domain); the schedule reuses the single-vector tuning's fill target.
bool mixed_batch_supported(const tci_plan& P) {
  const tci_plan::MixedLadder& M = P.mx;
  return M.ready && !P.vv.split && !M.merged && (M.prepacked || M.n_local == 0) && P.n_pairs > 0;
}

void mixed_batch_prepare(tci_plan& P, int precision, int nvec, hipStream_t st) {
  mixed_prepare(P, precision, st);
  tci_plan::MixedLadder& M = P.mx;
  tci_plan::MixedLadder::Batch& B = M.bat;
  if (B.nvec == nvec) return;
  TCI_REQUIRE(mixed_batch_supported(P), "batched split ladder needs per-walk programs with resident L blocks");
  TCI_REQUIRE(nvec >= 2 && nvec <= 64, "batch of 2..64 vectors");
  B.nvec = 0;
  const int n_pairs = P.n_pairs, n_v = P.n_v;
  const int64_t n_chol = P.n_chol, nvp = xpad(n_v), jstride_full = (int64_t)n_v * 2 * nvp;
  std::vector<int64_t> cxo(n_pairs + 1, 0), lpo_all;
  std::vector<int32_t> locq;
  for (int Q = 0; Q < n_pairs; ++Q) {
    const int64_t d = P.h_dom_off[Q + 1] - P.h_dom_off[Q];
    cxo[Q + 1] = cxo[Q] + d * 2 * xpad(d);
    if (d != n_v) locq.push_back(Q);
  }
  if (M.prepacked) {  // offsets of the local walks in the prepacked Lpx, as in mixed_prepare
    lpo_all.assign(locq.size() + 1, 0);
    for (size_t q = 0; q < locq.size(); ++q) {
      const int64_t d = P.h_dom_off[locq[q] + 1] - P.h_dom_off[locq[q]];
      lpo_all[q + 1] = lpo_all[q] + n_chol * d * 2 * xpad(d);
    }
  }
  const int64_t sum_x = cxo.back();
  B.Cxb.alloc((size_t)nvec * sum_x * 2);
  B.Rb.alloc((size_t)nvec * P.sum_d2 * 8);
  // T' of one chunk: 8 GB per vector of the batch (the single-vector chunks)
  int64_t Lc = std::max<int64_t>(1, std::min<int64_t>(n_chol, (int64_t)(8.0e9 / (2.0 * (double)sum_x))));
  for (int64_t q = Lc; q >= std::max<int64_t>(1, Lc / 2); --q)
    if (n_chol % q == 0) {
      Lc = q;
      break;
    }
  B.Txb.alloc((size_t)Lc * nvec * sum_x * 2);
  // phase-2 tiles by domain: 64 x 64 up to 64, 128 x 128 up to 128, 192 x 192 above
  static const int tt[3] = {64, 128, 192};
  auto cls = [](int64_t d) { return d <= 64 ? 0 : (d <= 128 ? 1 : 2); };
  const int64_t target = M.tuned_target > 0 ? M.tuned_target : 512;
  int64_t tv[3] = {0, 0, 0};
  for (int Q = 0; Q < n_pairs; ++Q) {
    const int64_t d = P.h_dom_off[Q + 1] - P.h_dom_off[Q];
    const int c = cls(d);
    tv[c] += cdiv(d, tt[c]) * cdiv(nvec * d, tt[c]);
  }
  int nsp_v[3];
  B.rpart_n = 1;
  for (int c = 0; c < 3; ++c) {
    nsp_v[c] = tv[c] ? (int)std::max<int64_t>(1, std::min<int64_t>(16, cdiv(target, tv[c]))) : 1;
    B.rpart_n = std::max(B.rpart_n, nsp_v[c]);
  }
  if (B.rpart_n > 1) B.Rpartb.alloc((size_t)B.rpart_n * nvec * P.sum_d2 * 8);
  B.phase.clear();
  B.chunks.clear();
  std::vector<int64_t> lens;
  for (int64_t J0 = 0; J0 < n_chol; J0 += Lc) {
    const int64_t len = std::min(Lc, n_chol - J0);
    int idx = -1;
    for (size_t q = 0; q < lens.size(); ++q)
      if (lens[q] == len) idx = (int)q;
    if (idx < 0) {
      idx = (int)lens.size();
      lens.push_back(len);
    }
    B.chunks.push_back({J0, idx});
  }
  const uint16_t* Cxb = B.Cxb.as<uint16_t>();
  uint16_t* Tx = B.Txb.as<uint16_t>();
  for (int64_t len : lens) {
    B.phase.emplace_back();
    tci_plan::LadderPhase& ph = B.phase.back();
    ph.Lc = len;
    ph.p1x.resize(1);
    ph.p2x.resize(3);
    ph.p1x[0].split = true;
    ph.p1x[0].tile = 128;
    ph.p1x[0].tile_n = 192;
    for (int c = 0; c < 3; ++c) {
      ph.p2x[c].split = true;
      ph.p2x[c].tile = tt[c];
      ph.p2x[c].tile_n = tt[c];
    }
    int64_t toff = 0;
    int qloc = 0;
    for (int Q = 0; Q < n_pairs; ++Q) {
      const int d = (int)(P.h_dom_off[Q + 1] - P.h_dom_off[Q]);
      const int64_t ldx = 2 * xpad(d);
      const uint16_t* Lb;
      int64_t jstep, jstride;
      if (d == n_v) {
        Lb = M.Lx.as<uint16_t>();
        jstep = jstride = jstride_full;
      } else {
        Lb = M.Lpx.as<uint16_t>() + lpo_all[qloc++];
        jstep = jstride = (int64_t)d * ldx;
      }
      uint16_t* Tq = Tx + toff;
      const int64_t Sv = (int64_t)d * ldx;  // one vector's block of T' for one jj
      Blocking b1;
      b1.rdiv = d;
      b1.rstride = nvec * Sv;
      b1.cdiv = d;
      b1.cstride = Sv;
      ph.p1x[0].add(reinterpret_cast<double*>(Tq), ldx, (int)(len * d), nvec * d,
                    {xseg(xop(Lb, ldx, jstride), xop(Cxb + nvec * cxo[Q], ldx), d)}, false, true, 0, b1, 3);
      const int c = cls(d);
      const int nsp = (int)std::min<int64_t>(nsp_v[c], len);
      Blocking b2;
      b2.cdiv = d;
      b2.cstride = P.sum_d2;
      for (int sp = 0; sp < nsp; ++sp) {
        const int64_t j0 = len * sp / nsp, j1 = len * (sp + 1) / nsp;
        std::vector<Seg> sg;
        for (int64_t jj = j0; jj < j1; ++jj)
          sg.push_back(xseg(xop(Lb + jj * jstep, ldx, jstride), xop(Tq + jj * nvec * Sv, ldx), d));
        double* Rout = (B.rpart_n > 1 ? B.Rpartb.as<double>() + (int64_t)sp * nvec * P.sum_d2 : B.Rb.as<double>()) +
                       P.h_cd_off[Q];
        ph.p2x[c].add(Rout, d, d, nvec * d, sg, true, false, 0, b2, 4);
      }
      toff += len * nvec * Sv;
    }
    for (auto& g : ph.p1x) g.upload_tables(st);
    for (auto& g : ph.p2x) g.upload_tables(st);
  }
  TCI_HIP(hipStreamSynchronize(st));
  B.nvec = nvec;
}

void mixed_batch_absmax(tci_plan& P, bool first, hipStream_t st) {
  unsigned* mx = P.mx.maxbits.as<unsigned>();
  if (first) hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, st, mx, 2);
  hipLaunchKernelGGL(absmax_kernel<double>, dim3(blocks_for(P.sum_d2, 8192)), dim3(256), 0, st, P.C.as<double>(),
                     P.sum_d2, mx);
  TCI_HIP(hipGetLastError());
}

void mixed_batch_set_scales(tci_plan& P, hipStream_t st) {
  hipLaunchKernelGGL(set_scales_kernel, dim3(1), dim3(64), 0, st, 1, P.mx.maxbits.as<unsigned>(),
                     P.mx.scales.as<float>(), (float)P.dmax);
  TCI_HIP(hipGetLastError());
}

void mixed_batch_split_c(tci_plan& P, int precision, int v, hipStream_t st) {
  tci_plan::MixedLadder& M = P.mx;
  const int f16 = precision == TCI_SIGMA_FP16X3 ? 1 : 0;
  hipLaunchKernelGGL(split_blocks_kernel, dim3(blocks_for((int64_t)P.dmax * ((P.dmax + 7) / 8), 256), P.n_pairs),
                     dim3(256), 0, st, P.tab, M.cx_off.as<int64_t>(), P.C.as<double>(), M.bat.Cxb.as<uint16_t>(),
                     M.scales.as<float>() + 1, f16, M.bat.nvec, v);
  TCI_HIP(hipGetLastError());
}

void mixed_batch_ladder(tci_plan& P, int precision, hipStream_t st) {
  tci_plan::MixedLadder& M = P.mx;
  tci_plan::MixedLadder::Batch& B = M.bat;
  const int64_t n = (int64_t)B.nvec * P.sum_d2;
  launch_fill(B.rpart_n > 1 ? B.Rpartb.as<double>() : B.Rb.as<double>(), (int64_t)B.rpart_n * n, 0.0, st);
  const float* sc = M.scales.as<float>();
  for (const auto& ch : B.chunks) {
    const tci_plan::LadderPhase& ph = B.phase[ch.second];
    for (const auto& g : ph.p1x) g.launch_split(precision, st, ch.first, sc);
    for (const auto& g : ph.p2x) g.launch_split(precision, st, ch.first, sc);
  }
  if (P.timing()) TCI_HIP(hipEventRecord(P.ev[3], st));
  if (B.rpart_n > 1) ladder_reduce(B.Rpartb.as<double>(), n, B.rpart_n, B.Rb.as<double>(), n, st);
}

}  // namespace tci
