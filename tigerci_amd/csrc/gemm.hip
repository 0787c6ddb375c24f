// Grouped gather-GEMM on FP64 MFMA (v_mfma_f64_16x16x4_f64), gfx950.
//
// One workgroup owns one output tile of one task and loops over all of that
// task's K-segments, so tiles never race and no atomics are needed.  Operands
// are read through optional gather maps (local pair domains, pair-index maps
// into the Cholesky records, (J,row) maps of the ladder), so one launch covers
// every pair block of a term regardless of its domain size.
//
// Two tile shapes:
//   64x64   4 waves of 32x32, single LDS buffer + register prefetch; used for
//           skinny outputs (a GEMV, n_int-wide panels).
//   128x128 4 waves of 64x64 (4x4 MFMA blocks, 16 accumulators), two LDS
//           stages so global loads of stage s+1 overlap the 64 MFMAs per wave
//           of stage s; one workgroup per CU with the full 512-register budget
//           per lane (no spills).  16x16 blocks that lie outside the output
//           are skipped (wave-uniform), so a 300x300 pair block costs 304x304,
//           not 384x384.
//
// MFMA 16x16x4 f64 lane layout (checked by tci_selftest_mfma):
//   A[i][k] at lane i + 16k, B[k][j] at lane j + 16k,
//   D[i][j] at lane j + 16*(i%4), register i/4.
#include <cstdlib>

#include <algorithm>

#include "common.hpp"

namespace tci {

namespace {

__device__ __forceinline__ double ldop(const Op& o, int64_t e) {
  return o.f32 ? (double)static_cast<const float*>(o.base)[e] : static_cast<const double*>(o.base)[e];
}

__device__ __forceinline__ double fetch(const Op& o, int64_t J0, int r, int k, int R, int K) {
  if (r >= R || k >= K || o.base == nullptr) return 0.0;
  int64_t rr = o.rmap ? (int64_t)o.rmap[r] : (int64_t)r;
  int64_t kk = o.cmap ? (int64_t)o.cmap[k] : (int64_t)k;
  if (rr < 0 || kk < 0) return 0.0;
  return ldop(o, J0 * o.jstride + rr * o.rs + kk * o.cs);
}

// ------------------------------------------------------------------ 64 x 64
constexpr int TM = 64, TN = 64, TK = 16;

__global__ __launch_bounds__(256) void ggemm64_kernel(const Task* __restrict__ tasks, const Seg* __restrict__ segs,
                                                      int64_t J0, int remap) {
  const Task t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
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
      ra[q] = S.alpha * fetch(S.A, J0, t.m0 + m, kb + k, t.M, S.K);
      int n, kk;
      if (b_n_contig) { n = e & 63; kk = e >> 6; } else { n = e >> 4; kk = e & 15; }
      rb[q] = fetch(S.B, J0, kb + kk, t.n0 + n, S.K, t.N);
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

  while (seg < t.seg1 && segs[seg].K <= 0) ++seg;
  if (seg < t.seg1) load_regs(seg, 0);
  while (seg < t.seg1) {
    const int cur_seg = seg;
    store_regs(cur_seg);
    __syncthreads();
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

// ---------------------------------------------------------------- 128 x 128
constexpr int BM = 128, BK = 16, LP = BM + 4;  // 128 x 128 tiles

// Each thread owns 8 elements of the A tile and 8 of the B tile: one row (A)
// or column (B) and 8 consecutive k.  Consecutive threads walk the operand's
// contiguous axis.
__device__ __forceinline__ void a_slot(bool kc, int tid, int& m, int& kbase) {
  if (kc) { m = tid >> 1; kbase = (tid & 1) * 8; }
  else { m = tid & 127; kbase = (tid >> 7) * 8; }
}
__device__ __forceinline__ void b_slot(bool nc, int tid, int& n, int& kbase) {
  if (nc) { n = tid & 127; kbase = (tid >> 7) * 8; }
  else { n = tid >> 1; kbase = (tid & 1) * 8; }
}

// Per-thread state of one stage: which (segment, k-block), the fixed row
// offset of its A row and column offset of its B column (-1: reads zero), and
// the gathered k indices of its 8 A and 8 B elements (-1: zero).
struct StageIdx {
  int seg, kb;
  int64_t aoff, boff;
  int32_t ka[8], kbr[8];
};

__device__ __forceinline__ bool advance(const Seg* segs, int seg1, int& seg, int& kb) {
  kb += BK;
  if (kb >= segs[seg].K) {
    kb = 0;
    ++seg;
    while (seg < seg1 && segs[seg].K <= 0) ++seg;
  }
  return seg < seg1;
}

__device__ __forceinline__ void fixed_offsets(const Seg& S, const Task& t, int64_t J0, int tid, int64_t& aoff,
                                              int64_t& boff) {
  int m, ka, n, kb;
  a_slot(S.A.cs == 1, tid, m, ka);
  b_slot(S.B.cs == 1, tid, n, kb);
  const int r = t.m0 + m, c = t.n0 + n;
  const int rc = min(r, t.M - 1), cc = min(c, t.N - 1);
  const int64_t rr = S.A.rmap ? (int64_t)gld(S.A.rmap + rc) : (int64_t)rc;
  const int64_t ccol = S.B.cmap ? (int64_t)gld(S.B.cmap + cc) : (int64_t)cc;
  aoff = (r < t.M && rr >= 0) ? J0 * S.A.jstride + rr * S.A.rs : -1;
  boff = (c < t.N && ccol >= 0) ? J0 * S.B.jstride + ccol * S.B.cs : -1;
}

__device__ __forceinline__ void kmaps(const Seg& S, int kb, int tid, int32_t* ka, int32_t* kbr) {
  // branch-free: clamp the index so the map load is always in bounds, select after
  int m, kaa, n, kbb;
  a_slot(S.A.cs == 1, tid, m, kaa);
  b_slot(S.B.cs == 1, tid, n, kbb);
  const int kmax = S.K - 1;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = kb + kaa + q;
    const int kc = min(k, kmax);
    const int32_t va = S.A.cmap ? gld(S.A.cmap + kc) : kc;
    ka[q] = k <= kmax ? va : -1;
    const int k2 = kb + kbb + q;
    const int kc2 = min(k2, kmax);
    const int32_t vb = S.B.rmap ? gld(S.B.rmap + kc2) : kc2;
    kbr[q] = k2 <= kmax ? vb : -1;
  }
}

__device__ __forceinline__ void load_data(const Seg& S, const StageIdx& x, double* a, double* b) {
  // every load reads a valid address (offset 0 when the element is outside the
  // operand) and is masked afterwards, so all 16 loads are in flight together
  if (S.A.f32) {
    const float* pa = static_cast<const float*>(S.A.base);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bool ok = x.aoff >= 0 && x.ka[q] >= 0;
      const float v = gld(pa + (ok ? x.aoff + (int64_t)x.ka[q] * S.A.cs : 0));
      a[q] = ok ? S.alpha * (double)v : 0.0;
    }
  } else {
    const double* pa = static_cast<const double*>(S.A.base);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bool ok = x.aoff >= 0 && x.ka[q] >= 0;
      const double v = gld(pa + (ok ? x.aoff + (int64_t)x.ka[q] * S.A.cs : 0));
      a[q] = ok ? S.alpha * v : 0.0;
    }
  }
  if (S.B.f32) {
    const float* pb = static_cast<const float*>(S.B.base);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bool ok = x.boff >= 0 && x.kbr[q] >= 0;
      const float v = gld(pb + (ok ? x.boff + (int64_t)x.kbr[q] * S.B.rs : 0));
      b[q] = ok ? (double)v : 0.0;
    }
  } else {
    const double* pb = static_cast<const double*>(S.B.base);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bool ok = x.boff >= 0 && x.kbr[q] >= 0;
      const double v = gld(pb + (ok ? x.boff + (int64_t)x.kbr[q] * S.B.rs : 0));
      b[q] = ok ? v : 0.0;
    }
  }
}

__device__ __forceinline__ void store_data(const Seg& S, int tid, const double* a, const double* b, double* As,
                                           double* Bs) {
  int m, ka, n, kb;
  a_slot(S.A.cs == 1, tid, m, ka);
  b_slot(S.B.cs == 1, tid, n, kb);
#pragma unroll
  for (int q = 0; q < 8; ++q) As[(ka + q) * LP + m] = a[q];
#pragma unroll
  for (int q = 0; q < 8; ++q) Bs[(kb + q) * LP + n] = b[q];
}

// Software pipeline: while the MFMAs of stage s run, the data of stage s+1
// (whose k indices were fetched during stage s-1) and the k indices of stage
// s+2 are in flight, so no load waits on another load.  The loop is
// specialised on the number of 16x16 blocks (NBM x NBN) this wave owns inside
// the output, so it is straight-line MFMA code without predicates; all
// variants execute the same loads and barriers.
template <int NBM, int NBN>
__device__ __forceinline__ void ggemm128_body(const Task& t, const Seg* __restrict__ segs, int64_t J0, int seg0,
                                              double* smem, int tid, int wm, int wn, int li, int lg) {
  double* As0 = smem;
  double* Bs0 = smem + BK * LP;
  double* As1 = smem + 2 * BK * LP;
  double* Bs1 = smem + 3 * BK * LP;
  d4 acc[NBM > 0 ? NBM : 1][NBN > 0 ? NBN : 1];
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  double ra[8], rb[8];
  StageIdx nxt;
  nxt.seg = seg0;
  nxt.kb = 0;
  fixed_offsets(segs[nxt.seg], t, J0, tid, nxt.aoff, nxt.boff);
  kmaps(segs[nxt.seg], nxt.kb, tid, nxt.ka, nxt.kbr);
  load_data(segs[nxt.seg], nxt, ra, rb);
  store_data(segs[nxt.seg], tid, ra, rb, As0, Bs0);
  const int seg_prev = nxt.seg;
  bool more = advance(segs, t.seg1, nxt.seg, nxt.kb);
  if (more) {
    if (nxt.seg != seg_prev) fixed_offsets(segs[nxt.seg], t, J0, tid, nxt.aoff, nxt.boff);
    kmaps(segs[nxt.seg], nxt.kb, tid, nxt.ka, nxt.kbr);
  }
  __syncthreads();
  int buf = 0;
  while (true) {
    StageIdx nn;
    bool more2 = false;
    if (more) {
      load_data(segs[nxt.seg], nxt, ra, rb);
      nn = nxt;
      more2 = advance(segs, t.seg1, nn.seg, nn.kb);
      if (more2) {
        if (nn.seg != nxt.seg) fixed_offsets(segs[nn.seg], t, J0, tid, nn.aoff, nn.boff);
        kmaps(segs[nn.seg], nn.kb, tid, nn.ka, nn.kbr);
      }
    }
    if (NBM > 0 && NBN > 0) {
      const double* As = buf ? As1 : As0;
      const double* Bs = buf ? Bs1 : Bs0;
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        const int kr = kk * 4 + lg;
        double a[NBM > 0 ? NBM : 1], b[NBN > 0 ? NBN : 1];
#pragma unroll
        for (int i = 0; i < NBM; ++i) a[i] = As[kr * LP + wm + i * 16 + li];
#pragma unroll
        for (int j = 0; j < NBN; ++j) b[j] = Bs[kr * LP + wn + j * 16 + li];
#pragma unroll
        for (int i = 0; i < NBM; ++i)
#pragma unroll
          for (int j = 0; j < NBN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    }
    if (!more) break;
    store_data(segs[nxt.seg], tid, ra, rb, buf ? As0 : As1, buf ? Bs0 : Bs1);
    __syncthreads();
    buf ^= 1;
    nxt = nn;
    more = more2;
  }
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = t.m0 + wm + i * 16 + lg + 4 * r;
        const int col = t.n0 + wn + j * 16 + li;
        if (row < t.M && col < t.N) {
          __attribute__((address_space(1))) double* p =
              (__attribute__((address_space(1))) double*)(t.C + (int64_t)row * t.ldc + col);
          *p = t.accumulate ? *p + acc[i][j][r] : acc[i][j][r];
        }
      }
}

__global__ __launch_bounds__(256, 1) void ggemm128_kernel(const Task* __restrict__ tasks,
                                                          const Seg* __restrict__ segs, int64_t J0, int remap) {
  const Task t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
  __shared__ double smem[4 * BK * LP];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
  // wave-uniform counts of the 16x16 blocks inside the output
  const int nbm = __builtin_amdgcn_readfirstlane(max(0, min(4, (t.M - (t.m0 + wm) + 15) >> 4)));
  const int nbn = __builtin_amdgcn_readfirstlane(max(0, min(4, (t.N - (t.n0 + wn) + 15) >> 4)));
  int seg0 = t.seg0;
  while (seg0 < t.seg1 && segs[seg0].K <= 0) ++seg0;
  if (seg0 >= t.seg1) {
    // no K: write zeros (or leave an accumulated output untouched)
    if (!t.accumulate)
      for (int e = tid; e < 128 * 128; e += 256) {
        const int row = t.m0 + (e >> 7), col = t.n0 + (e & 127);
        if (row < t.M && col < t.N) t.C[(int64_t)row * t.ldc + col] = 0.0;
      }
    return;
  }
#define TCI_G128(M_, N_) \
  case M_ * 5 + N_:      \
    ggemm128_body<M_, N_>(t, segs, J0, seg0, smem, tid, wm, wn, li, lg); \
    break;
  switch (nbm * 5 + nbn) {
    TCI_G128(0, 0) TCI_G128(0, 1) TCI_G128(0, 2) TCI_G128(0, 3) TCI_G128(0, 4)
    TCI_G128(1, 0) TCI_G128(1, 1) TCI_G128(1, 2) TCI_G128(1, 3) TCI_G128(1, 4)
    TCI_G128(2, 0) TCI_G128(2, 1) TCI_G128(2, 2) TCI_G128(2, 3) TCI_G128(2, 4)
    TCI_G128(3, 0) TCI_G128(3, 1) TCI_G128(3, 2) TCI_G128(3, 3) TCI_G128(3, 4)
    TCI_G128(4, 0) TCI_G128(4, 1) TCI_G128(4, 2) TCI_G128(4, 3) TCI_G128(4, 4)
    default: break;
  }
#undef TCI_G128
}

}  // namespace

namespace {
int remap_env() {
  static const int v = [] {
    const char* e = std::getenv("TCI_XCD_REMAP");
    return e && (e[0] == '0' || e[0] == '1') ? e[0] - '0' : -1;
  }();
  return v;
}
}  // namespace

int xcd_remap_default() { return remap_env() < 0 ? 0 : remap_env(); }
bool xcd_remap_forced() { return remap_env() >= 0; }

void launch_ggemm(int tile, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0, hipStream_t st,
                  int remap) {
  if (ntasks <= 0) return;
  TCI_REQUIRE(ntasks < ((int64_t)1 << 31), "too many gemm tasks");
  if (tile == 128)
    hipLaunchKernelGGL(ggemm128_kernel, dim3((unsigned)ntasks), dim3(256), 0, st, tasks, segs, J0, remap);
  else
    hipLaunchKernelGGL(ggemm64_kernel, dim3((unsigned)ntasks), dim3(256), 0, st, tasks, segs, J0, remap);
  TCI_HIP(hipGetLastError());
}

void GemmProgram::add(double* C, int64_t ldc, int M, int N, const std::vector<Seg>& s, bool accumulate,
                      bool out_split, int orow_div, Blocking blk, int oscale) {
  if (M <= 0 || N <= 0) return;
  TCI_REQUIRE((blk.rdiv == 0 && blk.cdiv == 0) || ((dense || split) && orow_div == 0 && blk.rdiv >= 0 && blk.cdiv >= 0),
              "blocked output needs a dense or split program");
  TCI_REQUIRE(oscale < 0 || split, "output scales belong to split programs");
  TCI_REQUIRE(!out_split || (split && !accumulate), "split-row output needs a split program and no accumulation");
  TCI_REQUIRE(!(split && tile == 256) || (!out_split && orow_div == 0 && (tile_n == 0 || tile_n == 256)),
              "256 x 256 split tiles write FP64 output only");
  if (split)
    for (const Seg& x : s)
      TCI_REQUIRE(x.alpha == 1.0 && orow_div == 0 && x.A.base && x.B.base && !x.A.rmap && !x.A.cmap && !x.B.rmap &&
                      !x.B.cmap && x.A.rs % 16 == 0 && x.B.rs % 16 == 0 && x.A.jstride % 16 == 0 &&
                      x.B.jstride % 16 == 0 && (reinterpret_cast<uintptr_t>(x.A.base) & 31) == 0 &&
                      (reinterpret_cast<uintptr_t>(x.B.base) & 31) == 0,
                  "split GEMM segments need 32-byte aligned split-row operands and alpha = 1");
  TCI_REQUIRE(orow_div == 0 || (dense && M % orow_div == 0), "row permutation needs a dense program and M % div == 0");
  if (dense && !split)
    for (const Seg& x : s)
      TCI_REQUIRE(x.A.cs == 1 && x.B.cs == 1 && !x.A.rmap && !x.A.cmap && !x.B.rmap && !x.B.cmap && x.A.base &&
                      x.B.base,
                  "dense GEMM segments need contiguous row-major operands");
  int seg0 = (int)segs.size();
  for (const Seg& x : s) segs.push_back(x);
  int seg1 = (int)segs.size();
  const int tn = (split && tile_n > 0) ? tile_n : tile;
  for (int m0 = 0; m0 < M; m0 += tile)
    for (int n0 = 0; n0 < N; n0 += tn) {
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
      t.out_split = out_split ? 1 : 0;
      t.orow_div = orow_div;
      t.rdiv = blk.rdiv;
      t.cdiv = blk.cdiv;
      t.rstride = blk.rstride;
      t.cstride = blk.cstride;
      t.oscale = oscale;
      t.pad_ = 0;
      tasks.push_back(t);
    }
}

void GemmProgram::upload_tables(hipStream_t st) {
  scaled = false;
  for (const Task& t : tasks) scaled = scaled || t.oscale >= 0;
  alg_flops = 0.0;
  const int tm = tile, tn = (split && tile_n > 0) ? tile_n : tile;
  for (const Task& t : tasks) {
    const double mv = std::min(tm, t.M - t.m0), nv = std::min(tn, t.N - t.n0);
    for (int q = t.seg0; q < t.seg1; ++q) alg_flops += 2.0 * mv * nv * segs[q].K;
  }
  upload(d_tasks, tasks, st);
  upload(d_segs, segs, st);
}

void GemmProgram::block_order(int G, const std::vector<Task>& base, hipStream_t st) {
  tasks = base;
  if (G > 0) {
    const int tm = tile, tn = (split && tile_n > 0) ? tile_n : tile;
    // grids in order of first appearance
    std::vector<std::pair<const double*, int32_t>> keys;
    std::vector<int> grid(tasks.size());
    for (size_t k = 0; k < tasks.size(); ++k) {
      const std::pair<const double*, int32_t> key(tasks[k].C, tasks[k].seg0);
      size_t g = 0;
      while (g < keys.size() && keys[g] != key) ++g;
      if (g == keys.size()) keys.push_back(key);
      grid[k] = (int)g;
    }
    std::vector<size_t> idx(tasks.size());
    for (size_t k = 0; k < idx.size(); ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
      const Task &x = base[a], &y = base[b];
      if (grid[a] != grid[b]) return grid[a] < grid[b];
      const int xb = x.m0 / (G * tm), yb = y.m0 / (G * tm);
      if (xb != yb) return xb < yb;
      const int xc = x.n0 / (G * tn), yc = y.n0 / (G * tn);
      if (xc != yc) return xc < yc;
      if (x.m0 != y.m0) return x.m0 < y.m0;
      return x.n0 < y.n0;
    });
    for (size_t k = 0; k < idx.size(); ++k) tasks[k] = base[idx[k]];
  }
  upload_tables(st);
}

std::string GemmProgram::kernel_name() const {
  const std::string t = std::to_string(tile), tn = std::to_string(tile_n > 0 ? tile_n : tile);
  if (split && tile == 256) return "xgemm256_kernel<" + std::to_string(waves) + ">";
  if (split) return dma ? "xgemmd_kernel<" + t + "," + tn + "," + std::to_string(dma) + ">"
                        : "xgemm_kernel<" + t + "," + tn + ">";
  if (dense) return "dgemm_kernel<" + t + ">";
  return "ggemm" + t + "_kernel";
}

void GemmProgram::launch(hipStream_t st, int64_t J0) const {
  TCI_REQUIRE(!split, "split program launched without a precision");
  const int rm = remap < 0 || xcd_remap_forced() ? xcd_remap_default() : remap;
  if (dense) launch_dgemm(tile, d_tasks.as<Task>(), d_segs.as<Seg>(), (int64_t)tasks.size(), J0, st, rm);
  else launch_ggemm(tile, d_tasks.as<Task>(), d_segs.as<Seg>(), (int64_t)tasks.size(), J0, st, rm);
}

}  // namespace tci

