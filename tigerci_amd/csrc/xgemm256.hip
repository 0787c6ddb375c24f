// Split GEMM on 256 x 256 tiles with 32 x 32 MFMAs.
//
// The 128 x 128 split kernel gives each wave a 64 x 64 quadrant: per 32-wide
// k-step a wave reads 16 KiB of fragments from LDS for 48 16x16x32 MFMAs,
// and with two workgroups per CU the LDS carries about 125 bytes per clock
// at full MFMA rate, its whole bandwidth.  Here each of the four waves owns a
// 128 x 128 quadrant as 4 x 4 blocks of 32 x 32 (v_mfma_f32_32x32x16_{bf16,f16},
// 256 accumulator registers): per k-step 32 KiB of fragments feed 96 MFMAs of
// twice the work, and one workgroup per CU writes 64 KiB of operands into LDS,
// so the LDS runs at about half its bandwidth.  Both stages (2 x 64 KiB) are
// filled by direct global-to-LDS loads (dstage, xgemm.hpp, chunks placed by
// xsw32 so that the fragment reads do not conflict); the loads of the
// next stage are in flight during the 3 x 32 MFMAs of the current one.
//
// An eight-wave form gives each wave 128 x 64 (4 x 2 blocks, 128 accumulator
// registers) and puts two waves on every SIMD.
//
// MFMA 32x32x16: src0[i][k] and src1[k][j] sit in lane i (j) + 32 (k / 8),
// element k % 8; D[i][j] in lane j + 32 ((i / 4) % 2), register
// 4 (i / 8) + i % 4.  The B fragment goes in as src0, so lane l holds output
// row l % 32 of its block and, in registers 4g .. 4g + 3, the four
// consecutive columns 8g + 4 (l / 32) .. +3.
#include "xgemm.hpp"

namespace tci {

namespace {

typedef float f16v __attribute__((ext_vector_type(16)));

template <int PREC>
struct X32;
template <>
struct X32<TCI_SIGMA_BF16X3> {
  typedef bf16x8 V;
  static __device__ __forceinline__ f16v mfma(V a, V b, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct X32<TCI_SIGMA_FP16X3> {
  typedef h8 V;
  static __device__ __forceinline__ f16v mfma(V a, V b, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};

constexpr int T256 = 256;

// One 32-wide k-step of a wave's NBM x NBN blocks (two 16-wide halves).
template <int PREC, int NBM, int NBN>
__device__ __forceinline__ void x256_step(const uint16_t* base, f16v (&acc)[NBM][NBN], int wm, int wn, int lane) {
  typedef typename X32<PREC>::V V;
  constexpr int PA = T256 * XK;
  const int r = lane & 31, hsel = lane >> 5;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    V ah[NBM], al[NBM], bh[NBN], bl[NBN];
#pragma unroll
    for (int i = 0; i < NBM; ++i) {
      const int e = xsw32(wm + 32 * i + r, 2 * h + hsel);
      ah[i] = *reinterpret_cast<const V*>(base + e);
      al[i] = *reinterpret_cast<const V*>(base + PA + e);
    }
#pragma unroll
    for (int j = 0; j < NBN; ++j) {
      const int e = xsw32(wn + 32 * j + r, 2 * h + hsel);
      bh[j] = *reinterpret_cast<const V*>(base + 2 * PA + e);
      bl[j] = *reinterpret_cast<const V*>(base + 3 * PA + e);
    }
    // small terms first (B fragment as src0: acc holds the transposed block)
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int j = 0; j < NBN; ++j) acc[i][j] = X32<PREC>::mfma(bh[j], al[i], acc[i][j]);
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int j = 0; j < NBN; ++j) acc[i][j] = X32<PREC>::mfma(bl[j], ah[i], acc[i][j]);
#pragma unroll
    for (int i = 0; i < NBM; ++i)
#pragma unroll
      for (int j = 0; j < NBN; ++j) acc[i][j] = X32<PREC>::mfma(bh[j], ah[i], acc[i][j]);
  }
}

// FP64 output (optionally accumulated, scaled, row- or column-blocked) of a
// wave's blocks: lane l writes row l % 32 of each block, four columns per
// register group.
template <int NBM, int NBN>
__device__ __forceinline__ void x256_store(const Task& t, const float* __restrict__ scales,
                                           const f16v (&acc)[NBM][NBN], int wm, int wn, int lane) {
  const float osc = t.oscale >= 0 ? scales[t.oscale] : 1.f;
  const int r = lane & 31, hsel = lane >> 5;
#pragma unroll
  for (int i = 0; i < NBM; ++i) {
    const int row = t.m0 + wm + 32 * i + r;
    if (row >= t.M) continue;
    const int64_t ro =
        t.rdiv ? (int64_t)(row / t.rdiv) * t.rstride + (int64_t)(row % t.rdiv) * t.ldc : (int64_t)row * t.ldc;
#pragma unroll
    for (int j = 0; j < NBN; ++j)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int col = t.n0 + wn + 32 * j + 8 * g + 4 * hsel;
        if (col >= t.N) continue;
        const int cb = t.cdiv ? col / t.cdiv : 0, cc = col - cb * t.cdiv;
        const bool whole = t.cdiv ? ((cc & 3) == 0 && cc + 3 < t.cdiv) : col + 3 < t.N;
        const float v0 = acc[i][j][4 * g] * osc, v1 = acc[i][j][4 * g + 1] * osc, v2 = acc[i][j][4 * g + 2] * osc,
                    v3 = acc[i][j][4 * g + 3] * osc;
        if (whole) {
          typedef double d4u __attribute__((ext_vector_type(4), aligned(8)));  // ldc is arbitrary
          __attribute__((address_space(1))) d4u* p =
              (__attribute__((address_space(1))) d4u*)(t.C + ro + (int64_t)cb * t.cstride + cc);
          const d4u o = {(double)v0, (double)v1, (double)v2, (double)v3};
          if (t.accumulate) *p = *p + o;
          else *p = o;
        } else {
          const float v[4] = {v0, v1, v2, v3};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int c = col + q;
            if (c >= t.N) break;
            const int cbq = t.cdiv ? c / t.cdiv : 0, ccq = c - cbq * t.cdiv;
            __attribute__((address_space(1))) double* p =
                (__attribute__((address_space(1))) double*)(t.C + ro + (int64_t)cbq * t.cstride + ccq);
            *p = t.accumulate ? *p + (double)v[q] : (double)v[q];
          }
        }
      }
  }
}

template <int PREC, int NW, int NBM, int NBN>
__device__ __forceinline__ void x256_body(const Task& t, const Seg* __restrict__ segs, int64_t J0, int seg0,
                                          const float* __restrict__ scales, uint16_t* sm, int w, int lane, int wm,
                                          int wn) {
  constexpr int STG = 4 * T256 * XK;
  f16v acc[NBM > 0 ? NBM : 1][NBN > 0 ? NBN : 1];
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;
  int lseg = seg0, lkb = 0;
  dstage<T256, T256, NW, true>(segs[lseg], t, J0, lkb, w, lane, sm);
  bool lmore = dadvance(segs, t.seg1, lseg, lkb);
  int cb = 0;
  while (true) {
    // s_waitcnt immediate: vmcnt(0) (expcnt and lgkmcnt fields left at their maximum)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __builtin_amdgcn_s_barrier();  // every wave's loads of this stage are in LDS; the other buffer is free
    bool issued = false;
    if (lmore) {
      dstage<T256, T256, NW, true>(segs[lseg], t, J0, lkb, w, lane, sm + (cb ^ 1) * STG);
      lmore = dadvance(segs, t.seg1, lseg, lkb);
      issued = true;
    }
    if constexpr (NBM > 0 && NBN > 0) x256_step<PREC, NBM, NBN>(sm + cb * STG, acc, wm, wn, lane);
    if (!issued) break;
    cb ^= 1;
  }
  if constexpr (NBM > 0 && NBN > 0) x256_store<NBM, NBN>(t, scales, acc, wm, wn, lane);
}

// NW = 4: four waves of 128 x 128 (4 x 4 blocks, one wave per SIMD);
// NW = 8: eight waves of 128 x 64 (4 x 2 blocks, two waves per SIMD to hide
// each other's LDS and barrier waits); NW = 16: sixteen waves of 64 x 64
// (2 x 2 blocks, four per SIMD)
template <int PREC, int NW>
__global__ __launch_bounds__(64 * NW, 1) void xgemm256_kernel(const Task* __restrict__ tasks,
                                                              const Seg* __restrict__ segs, int64_t J0,
                                                              const float* __restrict__ scales, int remap) {
  constexpr int WM = NW == 16 ? 64 : 128, WN = NW == 4 ? 128 : 64;  // wave tile
  constexpr int BM = WM / 32, BN = WN / 32, WR = T256 / WM;         // blocks; waves along M
  const Task t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
  __shared__ __attribute__((aligned(16))) uint16_t sm[2 * 4 * T256 * XK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (w % WR) * WM, wn = (w / WR) * WN;
  const int nbm = __builtin_amdgcn_readfirstlane(max(0, min(BM, (t.M - (t.m0 + wm) + 31) / 32)));
  const int nbn = __builtin_amdgcn_readfirstlane(max(0, min(BN, (t.N - (t.n0 + wn) + 31) / 32)));
  int seg0 = t.seg0;
  while (seg0 < t.seg1 && segs[seg0].K <= 0) ++seg0;
  if (seg0 >= t.seg1) {
    for (int e = tid; e < T256 * T256; e += 64 * NW) {
      const int row = t.m0 + e / T256, col = t.n0 + e % T256;
      if (row < t.M && col < t.N && !t.accumulate) {
        const int64_t ro = t.rdiv ? (int64_t)(row / t.rdiv) * t.rstride + (int64_t)(row % t.rdiv) * t.ldc
                                  : (int64_t)row * t.ldc;
        const int cb = t.cdiv ? col / t.cdiv : 0, cc = col - cb * t.cdiv;
        t.C[ro + (int64_t)cb * t.cstride + cc] = 0.0;
      }
    }
    return;
  }
  // every wave takes part in the loads and barriers; a wave whose part lies
  // outside the output (nbm or nbn 0) computes nothing, and a wave at the edge
  // computes all its blocks (rows and columns beyond M and N are loaded as
  // zeros and not stored), which keeps one instantiation per kernel
  if (nbm > 0 && nbn > 0) x256_body<PREC, NW, BM, BN>(t, segs, J0, seg0, scales, sm, w, lane, wm, wn);
  else x256_body<PREC, NW, 0, 0>(t, segs, J0, seg0, scales, sm, w, lane, wm, wn);
}

template <int NW>
void launch256(int precision, dim3 g, const Task* tasks, const Seg* segs, int64_t J0, const float* scales,
               hipStream_t st, int remap) {
  if (precision == TCI_SIGMA_BF16X3)
    hipLaunchKernelGGL((xgemm256_kernel<TCI_SIGMA_BF16X3, NW>), g, dim3(64 * NW), 0, st, tasks, segs, J0, scales,
                       remap);
  else
    hipLaunchKernelGGL((xgemm256_kernel<TCI_SIGMA_FP16X3, NW>), g, dim3(64 * NW), 0, st, tasks, segs, J0, scales,
                       remap);
}

}  // namespace

void launch_xgemm256(int precision, int waves, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0,
                     const float* scales, hipStream_t st, int remap) {
  if (ntasks <= 0) return;
  TCI_REQUIRE(ntasks < ((int64_t)1 << 31), "too many gemm tasks");
  TCI_REQUIRE(precision == TCI_SIGMA_BF16X3 || precision == TCI_SIGMA_FP16X3, "split GEMM needs a split precision");
  const dim3 g((unsigned)ntasks);
  if (waves == 16) launch256<16>(precision, g, tasks, segs, J0, scales, st, remap);
  else if (waves == 8) launch256<8>(precision, g, tasks, segs, J0, scales, st, remap);
  else launch256<4>(precision, g, tasks, segs, J0, scales, st, remap);
  TCI_HIP(hipGetLastError());
}

}  // namespace tci
