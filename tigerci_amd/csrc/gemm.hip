// Grouped gather-GEMM on FP64 MFMA (v_mfma_f64_16x16x4_f64), gfx950.
//
// One workgroup (4 waves, 256 lanes) owns one 64x64 output tile of one task
// and loops over all of that task's K-segments, so tiles never race and no
// atomics are needed.  Operands are read through optional gather maps (local
// pair domains, pair-index maps into the Cholesky records), which lets one
// launch cover every pair block of a term regardless of its domain size.
//
// MFMA 16x16x4 f64 lane layout (checked by tci_selftest_mfma):
//   A[i][k] at lane i + 16k, B[k][j] at lane j + 16k,
//   D[i][j] at lane j + 16*(i%4), register i/4.
#include "common.hpp"

namespace tci {

namespace {

constexpr int TM = 64, TN = 64, TK = 16;

__device__ __forceinline__ double fetch(const Op& o, int r, int k, int R, int K) {
  if (r >= R || k >= K || o.base == nullptr) return 0.0;
  int64_t rr = o.rmap ? (int64_t)o.rmap[r] : (int64_t)r;
  int64_t kk = o.cmap ? (int64_t)o.cmap[k] : (int64_t)k;
  if (rr < 0 || kk < 0) return 0.0;
  const int64_t e = rr * o.rs + kk * o.cs;
  return o.f32 ? (double)static_cast<const float*>(o.base)[e] : static_cast<const double*>(o.base)[e];
}

__global__ __launch_bounds__(256) void ggemm_f64_kernel(const Task* __restrict__ tasks,
                                                        const Seg* __restrict__ segs) {
  const Task t = tasks[blockIdx.x];
  __shared__ double As[TK][TM + 1];
  __shared__ double Bs[TK][TN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  double ra[4], rb[4];
  int seg = t.seg0, k0 = 0;
  auto load_regs = [&](int sg, int kb) {
    const Seg S = segs[sg];
    const bool a_k_contig = (S.A.cs == 1);
    const bool b_n_contig = (S.B.cs == 1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int e = tid + 256 * q;
      int m, k;
      if (a_k_contig) { m = e >> 4; k = e & 15; } else { m = e & 63; k = e >> 6; }
      ra[q] = S.alpha * fetch(S.A, t.m0 + m, kb + k, t.M, S.K);
      int n, kk;
      if (b_n_contig) { n = e & 63; kk = e >> 6; } else { n = e >> 4; kk = e & 15; }
      rb[q] = fetch(S.B, kb + kk, t.n0 + n, S.K, t.N);
    }
  };
  auto store_regs = [&](int sg) {
    const Seg S = segs[sg];
    const bool a_k_contig = (S.A.cs == 1);
    const bool b_n_contig = (S.B.cs == 1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int e = tid + 256 * q;
      int m, k;
      if (a_k_contig) { m = e >> 4; k = e & 15; } else { m = e & 63; k = e >> 6; }
      As[k][m] = ra[q];
      int n, kk;
      if (b_n_contig) { n = e & 63; kk = e >> 6; } else { n = e >> 4; kk = e & 15; }
      Bs[kk][n] = rb[q];
    }
  };

  // skip empty segments
  while (seg < t.seg1 && segs[seg].K <= 0) ++seg;
  if (seg < t.seg1) load_regs(seg, 0);
  while (seg < t.seg1) {
    const int cur_seg = seg;
    store_regs(cur_seg);
    __syncthreads();
    // advance to the next (segment, k-block)
    k0 += TK;
    if (k0 >= segs[seg].K) {
      k0 = 0;
      ++seg;
      while (seg < t.seg1 && segs[seg].K <= 0) ++seg;
    }
    if (seg < t.seg1) load_regs(seg, k0);
#pragma unroll
    for (int kk = 0; kk < TK / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      double a0 = As[kr][wm + (lane & 15)];
      double a1 = As[kr][wm + 16 + (lane & 15)];
      double b0 = Bs[kr][wn + (lane & 15)];
      double b1 = Bs[kr][wn + 16 + (lane & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  // write back
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = t.m0 + wm + i * 16 + (lane >> 4) + 4 * r;
        int col = t.n0 + wn + j * 16 + (lane & 15);
        if (row < t.M && col < t.N) {
          double* p = t.C + (int64_t)row * t.ldc + col;
          *p = t.accumulate ? *p + acc[i][j][r] : acc[i][j][r];
        }
      }
}

}  // namespace

void launch_ggemm(const Task* tasks, const Seg* segs, int64_t ntasks, hipStream_t st) {
  if (ntasks <= 0) return;
  const int64_t maxgrid = 1 << 30;
  TCI_REQUIRE(ntasks < maxgrid, "too many gemm tasks");
  hipLaunchKernelGGL(ggemm_f64_kernel, dim3((unsigned)ntasks), dim3(256), 0, st, tasks, segs);
  TCI_HIP(hipGetLastError());
}

void GemmProgram::add(double* C, int64_t ldc, int M, int N, const std::vector<Seg>& s,
                      bool accumulate) {
  if (M <= 0 || N <= 0) return;
  int seg0 = (int)segs.size();
  for (const Seg& x : s) segs.push_back(x);
  int seg1 = (int)segs.size();
  for (int m0 = 0; m0 < M; m0 += TM)
    for (int n0 = 0; n0 < N; n0 += TN) {
      Task t;
      t.C = C;
      t.ldc = ldc;
      t.M = M;
      t.N = N;
      t.m0 = m0;
      t.n0 = n0;
      t.seg0 = seg0;
      t.seg1 = seg1;
      t.accumulate = accumulate ? 1 : 0;
      tasks.push_back(t);
    }
}

void GemmProgram::upload_tables(hipStream_t st) {
  upload(d_tasks, tasks, st);
  upload(d_segs, segs, st);
}

void GemmProgram::launch(hipStream_t st) const {
  launch_ggemm(d_tasks.as<Task>(), d_segs.as<Seg>(), (int64_t)tasks.size(), st);
}

}  // namespace tci
