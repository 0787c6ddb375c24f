// Dense grouped GEMM on FP64 MFMA for row-major operands (the ladder phases).
//
// Same Task/Seg tables as the gather kernels, restricted to contiguous
// operands: A is M x K row-major (A.cs == 1, leading dimension A.rs), B is
// K x N row-major (B.cs == 1, leading dimension B.rs), no gather maps.  That
// restriction is what makes the staging cheap: every global load is a
// coalesced 8-byte access addressed by a 32-bit offset from a per-segment
// base, and out-of-range rows/columns are clamped and masked without branches.
//
// Tiles of TS x TS (TS = 128: 4 waves of 64 x 64, 66 KiB of LDS, two
// workgroups per CU; TS = 64: 4 waves of 32 x 32, 17 KiB, for small outputs
// such as 300 x 300 pair blocks where 128-wide tiles would leave a third of
// the waves idle), BK = 16, two LDS stages, global loads of stage s+1 in
// flight during the MFMAs of stage s.  Outputs may be blocked (Task rdiv /
// cdiv): one launch then scatters row and column blocks to separate places.  Each wave's loop is specialised on its
// count of 16 x 16 blocks inside the output (no predicated MFMAs, no
// accumulator copies).  An optional output-row permutation lets one launch
// write its rows in a different order (row r to (r % div)(M/div) + r/div).
// MFMA 16x16x4 f64: A[i][k] lane i+16k, B[k][j] lane j+16k, D[i][j] lane
// j+16(i%4), register i/4.
#include "common.hpp"

namespace tci {

namespace {

constexpr int DK = 16;

struct DStage {
  int seg, kb;
};

__device__ __forceinline__ bool dadvance(const Seg* segs, int seg1, int& seg, int& kb) {
  kb += DK;
  if (kb >= segs[seg].K) {
    kb = 0;
    ++seg;
    while (seg < seg1 && segs[seg].K <= 0) ++seg;
  }
  return seg < seg1;
}

// A: thread owns k = tid & 15 of rows (tid >> 4) + 16q; B: column tid % TS of
// rows tid / TS + (256 / TS) q; q < TS / 16 (a TS x 16 and a 16 x TS tile).
template <int TS, class TA, class TB>
__device__ __forceinline__ void dload_t(const Seg& S, const Task& t, int64_t J0, int kb, int tid, double* a, double* b) {
  constexpr int NQ = TS / 16, BR = 256 / TS;
  const TA* pa = static_cast<const TA*>(S.A.base) + J0 * S.A.jstride;
  const TB* pb = static_cast<const TB*>(S.B.base) + J0 * S.B.jstride;
  const int ka = kb + (tid & 15);
  const bool kaok = ka < S.K;
  const int kac = min(ka, S.K - 1);
  const int lda = (int)S.A.rs, ldb = (int)S.B.rs;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int r = t.m0 + (tid >> 4) + 16 * q;
    const bool ok = kaok && r < t.M;
    const int rc = min(r, t.M - 1);
    const double v = (double)gld(pa + ((int64_t)rc * lda + kac));
    a[q] = ok ? S.alpha * v : 0.0;
  }
  const int c = t.n0 + (tid % TS);
  const bool cok = c < t.N;
  const int cc = min(c, t.N - 1);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int k = kb + tid / TS + BR * q;
    const bool ok = cok && k < S.K;
    const int kc = min(k, S.K - 1);
    const double v = (double)gld(pb + ((int64_t)kc * ldb + cc));
    b[q] = ok ? v : 0.0;
  }
}

template <int TS>
__device__ __forceinline__ void dload(const Seg& S, const Task& t, int64_t J0, int kb, int tid, double* a, double* b) {
  if (S.A.f32) {
    if (S.B.f32) dload_t<TS, float, float>(S, t, J0, kb, tid, a, b);
    else dload_t<TS, float, double>(S, t, J0, kb, tid, a, b);
  } else {
    if (S.B.f32) dload_t<TS, double, float>(S, t, J0, kb, tid, a, b);
    else dload_t<TS, double, double>(S, t, J0, kb, tid, a, b);
  }
}

template <int TS>
__device__ __forceinline__ void dstore(int tid, const double* a, const double* b, double* As, double* Bs) {
  constexpr int NQ = TS / 16, BR = 256 / TS, DLP = TS + 2;
#pragma unroll
  for (int q = 0; q < NQ; ++q) As[(tid & 15) * DLP + (tid >> 4) + 16 * q] = a[q];
#pragma unroll
  for (int q = 0; q < NQ; ++q) Bs[(tid / TS + BR * q) * DLP + (tid % TS)] = b[q];
}

// Epilogue shared by the tile shapes: lane (li, lg) of block (i, j) holds rows
// 16i + lg + 4r, column 16j + li; plain, row-permuted or blocked output.
template <int NBM, int NBN, int BSM = 16, int BSN = 16>
__device__ __forceinline__ void dstore_out(const Task& t, const d4 (&acc)[NBM > 0 ? NBM : 1][NBN > 0 ? NBN : 1],
                                           int wm, int wn, int li, int lg) {
  // column offsets of this lane (one per j), then one row offset per (i, r)
  int64_t coff[NBN > 0 ? NBN : 1];
#pragma unroll
  for (int j = 0; j < NBN; ++j) {
    const int col = t.n0 + wn + j * BSN + li;
    const int cb = t.cdiv ? col / t.cdiv : 0;
    coff[j] = (int64_t)cb * t.cstride + (col - cb * t.cdiv);
  }
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = t.m0 + wm + i * BSM + lg + 4 * r;
      if (row >= t.M) continue;
      const int prow = t.orow_div > 0 ? (row % t.orow_div) * (t.M / t.orow_div) + row / t.orow_div : row;
      const int64_t ro = t.rdiv ? (int64_t)(row / t.rdiv) * t.rstride + (int64_t)(row % t.rdiv) * t.ldc
                                : (int64_t)prow * t.ldc;
#pragma unroll
      for (int j = 0; j < NBN; ++j) {
        const int col = t.n0 + wn + j * BSN + li;
        if (col < t.N) {
          __attribute__((address_space(1))) double* p = (__attribute__((address_space(1))) double*)(t.C + ro + coff[j]);
          *p = t.accumulate ? *p + acc[i][j][r] : acc[i][j][r];
        }
      }
    }
}

// Output tile of a task without K-segments: zeros unless accumulating.
template <int TS, int NT>
__device__ __forceinline__ void zero_tile(const Task& t, int tid) {
  for (int e = tid; e < TS * TS; e += NT) {
    const int row = t.m0 + e / TS, col = t.n0 + e % TS;
    if (row < t.M && col < t.N) {
      const int prow = t.orow_div > 0 ? (row % t.orow_div) * (t.M / t.orow_div) + row / t.orow_div : row;
      const int64_t ro = t.rdiv ? (int64_t)(row / t.rdiv) * t.rstride + (int64_t)(row % t.rdiv) * t.ldc
                                : (int64_t)prow * t.ldc;
      const int cb = t.cdiv ? col / t.cdiv : 0;
      t.C[ro + (int64_t)cb * t.cstride + (col - cb * t.cdiv)] = 0.0;
    }
  }
}

template <int TS, int NBM, int NBN>
__device__ __forceinline__ void dgemm_body(const Task& t, const Seg* __restrict__ segs, int64_t J0, int seg0,
                                           double* smem, int tid, int wm, int wn, int li, int lg) {
  constexpr int DLP = TS + 2, NQ = TS / 16;
  double* As0 = smem;
  double* Bs0 = smem + DK * DLP;
  double* As1 = smem + 2 * DK * DLP;
  double* Bs1 = smem + 3 * DK * DLP;
  d4 acc[NBM > 0 ? NBM : 1][NBN > 0 ? NBN : 1];
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  double ra[NQ], rb[NQ];
  DStage cur{seg0, 0};
  dload<TS>(segs[cur.seg], t, J0, cur.kb, tid, ra, rb);
  dstore<TS>(tid, ra, rb, As0, Bs0);
  DStage nxt = cur;
  bool more = dadvance(segs, t.seg1, nxt.seg, nxt.kb);
  __syncthreads();
  int buf = 0;
  while (true) {
    if (more) dload<TS>(segs[nxt.seg], t, J0, nxt.kb, tid, ra, rb);
    if (NBM > 0 && NBN > 0) {
      const double* As = buf ? As1 : As0;
      const double* Bs = buf ? Bs1 : Bs0;
#pragma unroll
      for (int kk = 0; kk < DK / 4; ++kk) {
        const int kr = kk * 4 + lg;
        double a[NBM > 0 ? NBM : 1], b[NBN > 0 ? NBN : 1];
#pragma unroll
        for (int i = 0; i < NBM; ++i) a[i] = As[kr * DLP + wm + i * 16 + li];
#pragma unroll
        for (int j = 0; j < NBN; ++j) b[j] = Bs[kr * DLP + wn + j * 16 + li];
#pragma unroll
        for (int i = 0; i < NBM; ++i)
#pragma unroll
          for (int j = 0; j < NBN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    }
    if (!more) break;
    dstore<TS>(tid, ra, rb, buf ? As0 : As1, buf ? Bs0 : Bs1);
    __syncthreads();
    buf ^= 1;
    more = dadvance(segs, t.seg1, nxt.seg, nxt.kb);
  }
  dstore_out<NBM, NBN>(t, acc, wm, wn, li, lg);
}

template <int TS>
__global__ __launch_bounds__(256, TS == 128 ? 2 : 4) void dgemm_kernel(const Task* __restrict__ tasks,
                                                                       const Seg* __restrict__ segs, int64_t J0,
                                                                       int remap) {
  constexpr int WB = TS / 32;  // 16x16 blocks per wave side
  const Task t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
  __shared__ double smem[4 * DK * (TS + 2)];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int wm = (w >> 1) * (TS / 2), wn = (w & 1) * (TS / 2);
  const int nbm = __builtin_amdgcn_readfirstlane(max(0, min(WB, (t.M - (t.m0 + wm) + 15) >> 4)));
  const int nbn = __builtin_amdgcn_readfirstlane(max(0, min(WB, (t.N - (t.n0 + wn) + 15) >> 4)));
  int seg0 = t.seg0;
  while (seg0 < t.seg1 && segs[seg0].K <= 0) ++seg0;
  if (seg0 >= t.seg1) {
    if (!t.accumulate) zero_tile<TS, 256>(t, tid);
    return;
  }
#define TCI_DB(M_, N_) \
  case M_ * 5 + N_:    \
    if constexpr (M_ <= WB && N_ <= WB) dgemm_body<TS, M_, N_>(t, segs, J0, seg0, smem, tid, wm, wn, li, lg); \
    break;
  switch (nbm * 5 + nbn) {
    TCI_DB(0, 0) TCI_DB(0, 1) TCI_DB(0, 2) TCI_DB(0, 3) TCI_DB(0, 4)
    TCI_DB(1, 0) TCI_DB(1, 1) TCI_DB(1, 2) TCI_DB(1, 3) TCI_DB(1, 4)
    TCI_DB(2, 0) TCI_DB(2, 1) TCI_DB(2, 2) TCI_DB(2, 3) TCI_DB(2, 4)
    TCI_DB(3, 0) TCI_DB(3, 1) TCI_DB(3, 2) TCI_DB(3, 3) TCI_DB(3, 4)
    TCI_DB(4, 0) TCI_DB(4, 1) TCI_DB(4, 2) TCI_DB(4, 3) TCI_DB(4, 4)
    default: break;
  }
#undef TCI_DB
}


// ---------------------------------------------------------------- 192 x 192
// Eight waves (2 x 4) on one 192 x 192 tile: a whole d x d pair
// block of a domain of up to 192 per workgroup, so each operand panel is read
// once (24 flops per byte against 8 for 64-wide tiles).  99 KiB of LDS, one
// workgroup per CU; up to 18 accumulator blocks per wave.  A stage is 192 x 16 of A
// (thread: k = tid & 15, rows tid >> 4 + 32 q) and 16 x 192 of B (thread:
// columns tid % 64 + 64 c, rows tid / 64 + 8 r), loads clamped and masked as
// in dload_t.
constexpr int T192 = 192, L192 = T192 + 2;

template <class TA, class TB>
__device__ __forceinline__ void dload192_t(const Seg& S, const Task& t, int64_t J0, int kb, int tid, double* a,
                                           double* b) {
  const TA* pa = static_cast<const TA*>(S.A.base) + J0 * S.A.jstride;
  const TB* pb = static_cast<const TB*>(S.B.base) + J0 * S.B.jstride;
  const int ka = kb + (tid & 15);
  const bool kaok = ka < S.K;
  const int kac = min(ka, S.K - 1);
  const int lda = (int)S.A.rs, ldb = (int)S.B.rs;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int r = t.m0 + (tid >> 4) + 32 * q;
    const bool ok = kaok && r < t.M;
    const int rc = min(r, t.M - 1);
    const double v = (double)gld(pa + ((int64_t)rc * lda + kac));
    a[q] = ok ? S.alpha * v : 0.0;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int col = t.n0 + (tid & 63) + 64 * c;
    const bool cok = col < t.N;
    const int cc = min(col, t.N - 1);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int k = kb + (tid >> 6) + 8 * r;
      const bool ok = cok && k < S.K;
      const int kc = min(k, S.K - 1);
      const double v = (double)gld(pb + ((int64_t)kc * ldb + cc));
      b[c * 2 + r] = ok ? v : 0.0;
    }
  }
}

__device__ __forceinline__ void dload192(const Seg& S, const Task& t, int64_t J0, int kb, int tid, double* a,
                                         double* b) {
  if (S.A.f32) {
    if (S.B.f32) dload192_t<float, float>(S, t, J0, kb, tid, a, b);
    else dload192_t<float, double>(S, t, J0, kb, tid, a, b);
  } else {
    if (S.B.f32) dload192_t<double, float>(S, t, J0, kb, tid, a, b);
    else dload192_t<double, double>(S, t, J0, kb, tid, a, b);
  }
}

__device__ __forceinline__ void dstore192(int tid, const double* a, const double* b, double* As, double* Bs) {
#pragma unroll
  for (int q = 0; q < 6; ++q) As[(tid & 15) * L192 + (tid >> 4) + 32 * q] = a[q];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 2; ++r) Bs[((tid >> 6) + 8 * r) * L192 + (tid & 63) + 64 * c] = b[c * 2 + r];
}

template <int NBM, int NBN>
__device__ __forceinline__ void dgemm192_body(const Task& t, const Seg* __restrict__ segs, int64_t J0, int seg0,
                                              double* smem, int tid, int wm, int wn, int li, int lg) {
  double* As0 = smem;
  double* Bs0 = smem + DK * L192;
  double* As1 = smem + 2 * DK * L192;
  double* Bs1 = smem + 3 * DK * L192;
  d4 acc[NBM > 0 ? NBM : 1][NBN > 0 ? NBN : 1];
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  double ra[6], rb[6];
  DStage cur{seg0, 0};
  dload192(segs[cur.seg], t, J0, cur.kb, tid, ra, rb);
  dstore192(tid, ra, rb, As0, Bs0);
  DStage nxt = cur;
  bool more = dadvance(segs, t.seg1, nxt.seg, nxt.kb);
  __syncthreads();
  int buf = 0;
  while (true) {
    if (more) dload192(segs[nxt.seg], t, J0, nxt.kb, tid, ra, rb);
    if (NBM > 0 && NBN > 0) {
      const double* As = buf ? As1 : As0;
      const double* Bs = buf ? Bs1 : Bs0;
#pragma unroll
      for (int kk = 0; kk < DK / 4; ++kk) {
        const int kr = kk * 4 + lg;
        double a[NBM > 0 ? NBM : 1], b[NBN > 0 ? NBN : 1];
#pragma unroll
        for (int i = 0; i < NBM; ++i) a[i] = As[kr * L192 + wm + i * 32 + li];
#pragma unroll
        for (int j = 0; j < NBN; ++j) b[j] = Bs[kr * L192 + wn + j * 64 + li];
#pragma unroll
        for (int i = 0; i < NBM; ++i)
#pragma unroll
          for (int j = 0; j < NBN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    }
    if (!more) break;
    dstore192(tid, ra, rb, buf ? As0 : As1, buf ? Bs0 : Bs1);
    __syncthreads();
    buf ^= 1;
    more = dadvance(segs, t.seg1, nxt.seg, nxt.kb);
  }
  dstore_out<NBM, NBN, 32, 64>(t, acc, wm, wn, li, lg);
}

__global__ __launch_bounds__(512, 1) void dgemm192_kernel(const Task* __restrict__ tasks, const Seg* __restrict__ segs,
                                                         int64_t J0, int remap) {
  const Task t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
  __shared__ double smem[4 * DK * L192];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  // 16 x 16 blocks dealt round-robin to the waves (row block 2i + w/4, column
  // block 4j + w%4), so a domain smaller than 192 leaves all waves about
  // equally loaded
  const int wm = (w >> 2) * 16, wn = (w & 3) * 16;
  const int nbm = __builtin_amdgcn_readfirstlane(max(0, min(6, (t.M - (t.m0 + wm) + 31) / 32)));
  const int nbn = __builtin_amdgcn_readfirstlane(max(0, min(3, (t.N - (t.n0 + wn) + 63) / 64)));
  int seg0 = t.seg0;
  while (seg0 < t.seg1 && segs[seg0].K <= 0) ++seg0;
  if (seg0 >= t.seg1) {
    if (!t.accumulate) zero_tile<T192, 512>(t, tid);
    return;
  }
#define TCI_DB(M_, N_) \
  case M_ * 4 + N_:    \
    dgemm192_body<M_, N_>(t, segs, J0, seg0, smem, tid, wm, wn, li, lg); \
    break;
  switch (nbm * 4 + nbn) {
    TCI_DB(0, 0) TCI_DB(0, 1) TCI_DB(0, 2) TCI_DB(0, 3)
    TCI_DB(1, 0) TCI_DB(1, 1) TCI_DB(1, 2) TCI_DB(1, 3)
    TCI_DB(2, 0) TCI_DB(2, 1) TCI_DB(2, 2) TCI_DB(2, 3)
    TCI_DB(3, 0) TCI_DB(3, 1) TCI_DB(3, 2) TCI_DB(3, 3)
    TCI_DB(4, 0) TCI_DB(4, 1) TCI_DB(4, 2) TCI_DB(4, 3)
    TCI_DB(5, 0) TCI_DB(5, 1) TCI_DB(5, 2) TCI_DB(5, 3)
    TCI_DB(6, 0) TCI_DB(6, 1) TCI_DB(6, 2) TCI_DB(6, 3)
    default: break;
  }
#undef TCI_DB
}

}  // namespace

void launch_dgemm(int tile, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0, hipStream_t st,
                  int remap) {
  if (ntasks <= 0) return;
  TCI_REQUIRE(ntasks < ((int64_t)1 << 31), "too many gemm tasks");
  if (tile == 192)
    hipLaunchKernelGGL(dgemm192_kernel, dim3((unsigned)ntasks), dim3(512), 0, st, tasks, segs, J0, remap);
  else if (tile == 128)
    hipLaunchKernelGGL(dgemm_kernel<128>, dim3((unsigned)ntasks), dim3(256), 0, st, tasks, segs, J0, remap);
  else if (tile == 64)
    hipLaunchKernelGGL(dgemm_kernel<64>, dim3((unsigned)ntasks), dim3(256), 0, st, tasks, segs, J0, remap);
  else
    throw Error(TCI_ERR_ARG, "unsupported FP64 GEMM tile " + std::to_string(tile));
  TCI_HIP(hipGetLastError());
}

}  // namespace tci
