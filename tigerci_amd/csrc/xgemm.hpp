// Dense grouped GEMM in split precision on 16-bit MFMA:
//   C[tile] (+)= s sum_seg A_seg B_seg^T,   A_seg: M x K,  B_seg: N x K,
// both operands in split-row storage (hi/lo chunks, common.hpp), so
// A B^T ~= Ahi Bhi^T + Ahi Blo^T + Alo Bhi^T with FP32 accumulation.
// bf16x3 (v_mfma_f32_16x16x32_bf16): hi + lo keeps ~17 significant bits of
// each operand and the dropped lo lo^T term is of the same relative size.
// Scaled fp16x3 (v_mfma_f32_16x16x32_f16): ~22 bits; operands carry
// power-of-two scales chosen by their producers, and the task's output scale s
// (scales[oscale]) undoes them exactly.
//
// Because both operands are k-contiguous and pre-split, staging is a pure
// copy: per 32-wide k-step a thread moves 32-byte chunks (hi 8 + lo 8) from
// global memory into LDS, no conversion in the loop.  LDS keeps each operand
// as a hi plane and a lo plane of [row][32] bf16 (64 B per row) with the
// 16-byte chunk index XOR-swizzled by (row >> 1) & 3: any 8 consecutive lanes
// of a fragment read or of a staging write then touch 8 distinct 16-byte bank
// groups.  A lane's fragment (row = lane % 16, k-chunk = lane / 16) is one
// 16-byte LDS read.
//
// Tiles TM x TN (4 waves of TM/2 x TN/2): 128 x 128 (64 KiB LDS, 2 workgroups
// per CU), 128 x 192 and 128 x 64 (for outputs whose short side is a domain of
// up to 192 or 64, so that one tile spans it and each operand panel is read
// once), 192 x 192 (96 KiB, one workgroup per CU: a whole d x d pair block of
// up to 192 per tile, both operand panels read once), 64 x 64 (32 KiB, 4
// workgroups per CU); two LDS
// stages with the global loads of stage s+1 in flight during the MFMAs of
// stage s; each wave's loop specialised on its count of 16 x 16 blocks inside
// the output.  Output: FP64 (optionally accumulated) or, with out_split, a
// split-row matrix (the operand of a following split GEMM).
// MFMA 16x16x32 bf16: A[i][k] lane i + 16(k/8) element k%8, B[k][j] lane
// j + 16(k/8) element k%8, D[i][j] lane j + 16(i/4) register i%4.  The
// kernel feeds the B fragment as the MFMA's A operand, so a lane's four
// result registers are four consecutive output columns of one row: FP64
// results leave as 32-byte vectors and split results as 8-byte hi and lo
// groups.
//
// This header holds the kernel templates; dense_split.hip instantiates the
// 64/128-row tiles and the launch dispatch, xgemm192_*.hip the 192 x 192 tile
// (one translation unit per precision, so the builds run in parallel).
#pragma once

#include "common.hpp"

namespace tci {

namespace {

constexpr int XK = 32;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct XT;
template <>
struct XT<TCI_SIGMA_BF16X3> {
  typedef __bf16 T;
  typedef bf16x8 V;
  static __device__ __forceinline__ f4 mfma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <>
struct XT<TCI_SIGMA_FP16X3> {
  typedef _Float16 T;
  typedef h8 V;
  static __device__ __forceinline__ f4 mfma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

// element offset of k-chunk ch (8 values) of LDS row `row`
__device__ __forceinline__ int xsw(int row, int ch) { return row * XK + (((ch ^ (row >> 1)) & 3) << 3); }

// The same for the 32 x 32 MFMA fragments of the 256-tile kernel.  The LDS
// serves a 128-bit read 16 lanes per pass, lanes 8q..8q+7 with the eight lanes
// 16 further (the grouping the bank-conflict counters of both fragment
// shapes fit), over 64 banks: rows 4 apart share banks.  The 16 x 16 fragments
// read chunks 0 and 1 (or 2 and 3) of rows r..r+7 in a pass, which xsw keeps
// apart; the 32 x 32 fragments read one chunk of rows r..r+7 and r+16..r+23,
// so the chunk position takes row bits 2 and 4.
__device__ __forceinline__ int xsw32f(int row) { return ((row >> 1) & 2) | ((row >> 4) & 1); }
__device__ __forceinline__ int xsw32(int row, int ch) { return row * XK + (((ch ^ xsw32f(row)) & 3) << 3); }

__device__ __forceinline__ bool xadvance(const Seg* segs, int seg1, int& seg, int& kb) {
  kb += XK;
  if (kb >= segs[seg].K) {
    kb = 0;
    ++seg;
    while (seg < seg1 && segs[seg].K <= 0) ++seg;
  }
  return seg < seg1;
}

// Rows r0 + (tid >> 2) + 64 q (q < R / 64), k-chunk tid & 3 of one operand.
// Rows beyond `rows` and chunks beyond K load from a zero page instead, so no
// select waits on the loaded data.
__device__ __attribute__((aligned(32))) u32x4 xzero[2];

template <int R>
__device__ __forceinline__ void xload(const Op& O, int64_t J0, int rows, int r0, int kb, int K, int tid,
                                      u32x4 (*v)[2]) {
  constexpr int NL = R / 64;
  const uint16_t* base = static_cast<const uint16_t*>(O.base) + J0 * O.jstride;
  const int k = kb + 8 * (tid & 3);
  const bool kok = k < K;
  const int64_t koff = (int64_t)(k >> 3) << 4;
#pragma unroll
  for (int q = 0; q < NL; ++q) {
    const int r = r0 + (tid >> 2) + 64 * q;
    const u32x4* p = (kok && r < rows) ? reinterpret_cast<const u32x4*>(base + (int64_t)r * O.rs + koff) : xzero;
    v[q][0] = gld(p);
    v[q][1] = gld(p + 1);
  }
}

template <int R>
__device__ __forceinline__ void xstore(int tid, const u32x4 (*v)[2], uint16_t* hi, uint16_t* lo) {
  constexpr int NL = R / 64;
#pragma unroll
  for (int q = 0; q < NL; ++q) {
    const int e = xsw((tid >> 2) + 64 * q, tid & 3);
    *reinterpret_cast<u32x4*>(hi + e) = v[q][0];
    *reinterpret_cast<u32x4*>(lo + e) = v[q][1];
  }
}

template <int TM, int TN>
__device__ __forceinline__ void xstage(const Seg& S, const Task& t, int64_t J0, int kb, int tid, u32x4 (*va)[2],
                                       u32x4 (*vb)[2]) {
  xload<TM>(S.A, J0, t.M, t.m0, kb, S.K, tid, va);
  xload<TN>(S.B, J0, t.N, t.n0, kb, S.K, tid, vb);
}

// Direct loads of one stage (k-step kb of segment S) into stage buffer sb
// (A hi, A lo planes of TM rows, then B hi, B lo of TN rows): a plane of R rows
// is R / 16 instructions of 64 lanes; wave w of NW issues instructions w, w + NW, ...
// Chunks are placed by xsw, or by xsw32 when SW32 is set.
template <int TM, int TN, int NW = 4, bool SW32 = false>
__device__ __forceinline__ void dstage(const Seg& S, const Task& t, int64_t J0, int kb, int w, int lane,
                                       uint16_t* sb) {
  constexpr int PA = TM * XK, PB = TN * XK;
#pragma unroll
  for (int op = 0; op < 2; ++op) {
    const Op& O = op ? S.B : S.A;
    const int rows = op ? t.N : t.M, r0 = op ? t.n0 : t.m0;
    const uint16_t* base = static_cast<const uint16_t*>(O.base) + J0 * O.jstride;
    uint16_t* plane = sb + (op ? 2 * PA : 0);
    const int pl = op ? PB : PA;
#pragma unroll
    for (int q = 0; q < (op ? TN : TM) / (16 * NW); ++q) {
      const int ins = w + NW * q, s = 64 * ins + lane;
      const int row = s >> 2, ch = (s & 3) ^ (SW32 ? xsw32f(row) : ((row >> 1) & 3));
      const int k = kb + 8 * ch, r = r0 + row;
      const u32x4* p = (k < S.K && r < rows)
                           ? reinterpret_cast<const u32x4*>(base + (int64_t)r * O.rs + ((int64_t)(k >> 3) << 4))
                           : xzero;
      uint16_t* hi = plane + ins * 512;
      uint16_t* lo = plane + pl + ins * 512;
      __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)p,
                                       (void __attribute__((address_space(3)))*)hi, 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)(p + 1),
                                       (void __attribute__((address_space(3)))*)lo, 16, 0, 0);
    }
  }
}

// Next k-step of a task's segment list with the segment index kept in
// scalar registers, so that the segment reads are scalar loads: a vector load
// here would make the compiler wait for every outstanding load, the direct
// loads of the stages in flight included.
__device__ __forceinline__ bool dadvance(const Seg* segs, int seg1, int& seg, int& kb) {
  kb += XK;
  if (kb >= segs[__builtin_amdgcn_readfirstlane(seg)].K) {
    kb = 0;
    seg = __builtin_amdgcn_readfirstlane(seg + 1);
    while (seg < seg1 && segs[seg].K <= 0) seg = __builtin_amdgcn_readfirstlane(seg + 1);
  }
  seg = __builtin_amdgcn_readfirstlane(seg);
  kb = __builtin_amdgcn_readfirstlane(kb);
  return seg < seg1;
}

// distance between a wave's consecutive 16 x 16 blocks: 16 (quadrants) or 32
// (round-robin over two waves, 192 x 192 tiles)
template <int TM, int TN>
struct XInterleave {
  static constexpr int BS = (TM == 192 && TN == 192) ? 32 : 16;
};

// One 32-wide k-step of a wave's blocks from an LDS stage (A hi, A lo, B hi,
// B lo planes of PA / PB elements).
template <int TM, int TN, int PREC, int NBM, int NBN>
__device__ __forceinline__ void xgemm_step(const uint16_t* base, f4 (&acc)[NBM > 0 ? NBM : 1][NBN > 0 ? NBN : 1],
                                           int wm, int wn, int li, int lg) {
  typedef typename XT<PREC>::V V;
  constexpr int BS = XInterleave<TM, TN>::BS;
  constexpr int PA = TM * XK, PB = TN * XK;
  if (NBM > 0 && NBN > 0) {
    V ah[NBM > 0 ? NBM : 1], al[NBM > 0 ? NBM : 1], bh[NBN > 0 ? NBN : 1], bl[NBN > 0 ? NBN : 1];
#pragma unroll
    for (int i = 0; i < NBM; ++i) {
      const int e = xsw(wm + i * BS + li, lg);
      ah[i] = *reinterpret_cast<const V*>(base + e);
      al[i] = *reinterpret_cast<const V*>(base + PA + e);
    }
#pragma unroll
    for (int j = 0; j < NBN; ++j) {
      const int e = xsw(wn + j * BS + li, lg);
      bh[j] = *reinterpret_cast<const V*>(base + 2 * PA + e);
      bl[j] = *reinterpret_cast<const V*>(base + 2 * PA + PB + e);
    }
    // B fragment as the MFMA's first operand: acc[i][j] holds the transposed
    // 16 x 16 block, so each lane ends with 4 consecutive output columns.
    // Small terms first; the three passes keep 16 independent chains.
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int j = 0; j < NBN; ++j) acc[i][j] = XT<PREC>::mfma(bh[j], al[i], acc[i][j]);
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int j = 0; j < NBN; ++j) acc[i][j] = XT<PREC>::mfma(bl[j], ah[i], acc[i][j]);
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int j = 0; j < NBN; ++j) acc[i][j] = XT<PREC>::mfma(bh[j], ah[i], acc[i][j]);
  }
}

// Epilogue: FP64 (optionally accumulated) or split-row output of a wave's blocks.
template <int TM, int TN, int PREC, int NBM, int NBN>
__device__ __forceinline__ void xgemm_store(const Task& t, const float* __restrict__ scales,
                                            const f4 (&acc)[NBM > 0 ? NBM : 1][NBN > 0 ? NBN : 1], int wm, int wn,
                                            int li, int lg) {
  typedef typename XT<PREC>::T T16;
  constexpr int BS = XInterleave<TM, TN>::BS;
  // lane (li, lg) of block (i, j) holds output row 16i + li, columns 16j + 4lg .. +3
  const float osc = t.oscale >= 0 ? scales[t.oscale] : 1.f;
  int64_t roff[NBM > 0 ? NBM : 1];
#pragma unroll
  for (int i = 0; i < NBM; ++i) {
    const int row = t.m0 + wm + i * BS + li;
    roff[i] = t.rdiv ? (int64_t)(row / t.rdiv) * t.rstride + (int64_t)(row % t.rdiv) * t.ldc : (int64_t)row * t.ldc;
  }
#pragma unroll
  for (int j = 0; j < NBN; ++j) {
    const int col = t.n0 + wn + j * BS + 4 * lg;
    if (col >= t.N) continue;
    const int cb = t.cdiv ? col / t.cdiv : 0, cc = col - cb * t.cdiv;
    const int64_t coff = (int64_t)cb * t.cstride;
    // one vector store when the 4 columns stay in one column block (and, for a
    // split row, in one 8-wide chunk); split rows without blocking may spill
    // into their zero padding, which receives the exact zeros of masked B rows
    const bool whole = t.cdiv ? ((cc & 3) == 0 && cc + 3 < t.cdiv) : (t.out_split || col + 3 < t.N);
#pragma unroll
    for (int i = 0; i < NBM; ++i) {
      const int row = t.m0 + wm + i * BS + li;
      if (row >= t.M) continue;
      const f4 v = acc[i][j] * osc;
      if (whole) {
        if (t.out_split) {
          typedef T16 b4 __attribute__((ext_vector_type(4)));
          b4 h, l;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            h[r] = (T16)v[r];
            l[r] = (T16)(v[r] - (float)h[r]);
          }
          __attribute__((address_space(1))) b4* p = (__attribute__((address_space(1))) b4*)(
              reinterpret_cast<uint16_t*>(t.C) + roff[i] + coff + xidx(cc));
          p[0] = h;
          p[2] = l;
        } else {
          typedef double d4u __attribute__((ext_vector_type(4), aligned(8)));  // ldc is arbitrary
          __attribute__((address_space(1))) d4u* p = (__attribute__((address_space(1))) d4u*)(t.C + roff[i] + coff + cc);
          const d4u o = {(double)v[0], (double)v[1], (double)v[2], (double)v[3]};
          if (t.accumulate) *p = *p + o;
          else *p = o;
        }
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int cr = col + r;
          if (cr >= t.N) break;
          const int cbr = t.cdiv ? cr / t.cdiv : 0, ccr = cr - cbr * t.cdiv;
          const int64_t o = roff[i] + (int64_t)cbr * t.cstride;
          if (t.out_split) {
            __attribute__((address_space(1))) T16* p =
                (__attribute__((address_space(1))) T16*)(reinterpret_cast<uint16_t*>(t.C) + o + xidx(ccr));
            const T16 h = (T16)v[r];
            p[0] = h;
            p[8] = (T16)(v[r] - (float)h);
          } else {
            __attribute__((address_space(1))) double* p = (__attribute__((address_space(1))) double*)(t.C + o + ccr);
            *p = t.accumulate ? *p + (double)v[r] : (double)v[r];
          }
        }
      }
    }
  }
}

template <int TM, int TN, int PREC, int NBM, int NBN>
__device__ __forceinline__ void xgemm_body(const Task& t, const Seg* __restrict__ segs, int64_t J0, int seg0,
                                           const float* __restrict__ scales, uint16_t* sm, int tid, int wm, int wn,
                                           int li, int lg) {
  // stage layout: A hi, A lo (TM rows each), B hi, B lo (TN rows each)
  constexpr int PA = TM * XK, PB = TN * XK, STG = 2 * PA + 2 * PB;
  f4 acc[NBM > 0 ? NBM : 1][NBN > 0 ? NBN : 1];
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  u32x4 va[TM / 64][2], vb[TN / 64][2];
  int seg = seg0, kb = 0;
  xstage<TM, TN>(segs[seg], t, J0, kb, tid, va, vb);
  xstore<TM>(tid, va, sm, sm + PA);
  xstore<TN>(tid, vb, sm + 2 * PA, sm + 2 * PA + PB);
  bool more = xadvance(segs, t.seg1, seg, kb);
  __syncthreads();
  int buf = 0;
  while (true) {
    if (more) xstage<TM, TN>(segs[seg], t, J0, kb, tid, va, vb);
    xgemm_step<TM, TN, PREC, NBM, NBN>(sm + buf * STG, acc, wm, wn, li, lg);
    if (!more) break;
    uint16_t* nb = sm + (buf ^ 1) * STG;
    xstore<TM>(tid, va, nb, nb + PA);
    xstore<TN>(tid, vb, nb + 2 * PA, nb + 2 * PA + PB);
    __syncthreads();
    buf ^= 1;
    more = xadvance(segs, t.seg1, seg, kb);
  }
  xgemm_store<TM, TN, PREC, NBM, NBN>(t, scales, acc, wm, wn, li, lg);
}

// Output tile of a task without K-segments: zeros unless accumulating.
template <int TM, int TN>
__device__ __forceinline__ void xgemm_zero_tile(const Task& t, int tid) {
  if (!t.accumulate)
    for (int e = tid; e < TM * TN; e += 256) {
      const int row = t.m0 + e / TN, col = t.n0 + e % TN;
      if (row < t.M && col < t.N) {
        const int64_t ro = t.rdiv ? (int64_t)(row / t.rdiv) * t.rstride + (int64_t)(row % t.rdiv) * t.ldc
                                  : (int64_t)row * t.ldc;
        const int cb = t.cdiv ? col / t.cdiv : 0, cc = col - cb * t.cdiv;
        if (t.out_split) {
          uint16_t* p = reinterpret_cast<uint16_t*>(t.C) + ro + (int64_t)cb * t.cstride + xidx(cc);
          p[0] = 0;
          p[8] = 0;
        } else {
          t.C[ro + (int64_t)cb * t.cstride + cc] = 0.0;
        }
      }
    }
}

template <int TM, int TN, int PREC>
__global__ __launch_bounds__(256, TM == 64 ? 4 : (TM + TN > 320 ? 1 : 2)) void xgemm_kernel(const Task* __restrict__ tasks,
                                                                      const Seg* __restrict__ segs, int64_t J0,
                                                                      const float* __restrict__ scales, int remap) {
  constexpr int WBM = TM / 32, WBN = TN / 32;  // 16x16 blocks per wave side
  const Task t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
  __shared__ __attribute__((aligned(16))) uint16_t sm[2 * 2 * (TM + TN) * XK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  // 192 x 192 tiles deal their 16 x 16 blocks round-robin to the waves (row
  // block 2i + w/2, column block 2j + w%2), so a domain smaller than 192 leaves
  // the four waves about equally loaded; other tiles give each wave a quadrant
  constexpr int BS = XInterleave<TM, TN>::BS;
  const int wm = BS == 32 ? (w >> 1) * 16 : (w >> 1) * (TM / 2);
  const int wn = BS == 32 ? (w & 1) * 16 : (w & 1) * (TN / 2);
  const int nbm = __builtin_amdgcn_readfirstlane(max(0, min(WBM, (t.M - (t.m0 + wm) + BS - 1) / BS)));
  const int nbn = __builtin_amdgcn_readfirstlane(max(0, min(WBN, (t.N - (t.n0 + wn) + BS - 1) / BS)));
  int seg0 = t.seg0;
  while (seg0 < t.seg1 && segs[seg0].K <= 0) ++seg0;
  if (seg0 >= t.seg1) {
    xgemm_zero_tile<TM, TN>(t, tid);
    return;
  }
#define TCI_XB(M_, N_) \
  case M_ * 8 + N_:    \
    if constexpr (M_ <= WBM && N_ <= WBN)                                                              \
      xgemm_body<TM, TN, PREC, M_, N_>(t, segs, J0, seg0, scales, sm, tid, wm, wn, li, lg);            \
    break;
#define TCI_XROW(M_) TCI_XB(M_, 0) TCI_XB(M_, 1) TCI_XB(M_, 2) TCI_XB(M_, 3) TCI_XB(M_, 4) TCI_XB(M_, 5) TCI_XB(M_, 6)
  switch (nbm * 8 + nbn) {
    TCI_XROW(0) TCI_XROW(1) TCI_XROW(2) TCI_XROW(3) TCI_XROW(4) TCI_XROW(5) TCI_XROW(6)
    default: break;
  }
#undef TCI_XROW
#undef TCI_XB
}

}  // namespace

template <int TM, int TN, int PREC>
inline void launch_xgemm_shape(const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0, const float* scales,
                               hipStream_t st, int remap) {
  hipLaunchKernelGGL((xgemm_kernel<TM, TN, PREC>), dim3((unsigned)ntasks), dim3(256), 0, st, tasks, segs, J0, scales,
                     remap);
}

// 192 x 192 tiles, one translation unit per precision
void launch_xgemm192_bf16(const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0, const float* scales,
                          hipStream_t st, int remap);
void launch_xgemm192_f16(const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0, const float* scales,
                         hipStream_t st, int remap);
// 256 x 256 tiles on 32 x 32 MFMAs with direct-load stages (xgemm256.hip);
// FP64 output only
void launch_xgemm256(int precision, int waves, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0,
                     const float* scales, hipStream_t st, int remap);
// LDS stages filled by direct loads (xgemm192d.hip): 192 x 192 with 3 stages,
// 128 x 192 with 2 or 3
void launch_xgemmd(int precision, int tm, int tn, int nst, const Task* tasks, const Seg* segs, int64_t ntasks,
                   int64_t J0, const float* scales, hipStream_t st, int remap);

}  // namespace tci
