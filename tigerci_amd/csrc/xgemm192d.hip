// Split GEMM tiles (192 x 192, 128 x 192, 128 x 128) with LDS stages filled by direct
// global-to-LDS loads (global_load_lds_dwordx4).
//
// The register-staged kernel (xgemm.hpp) has one stage in flight while it
// computes the other, and its staging registers bound how many bytes a
// workgroup keeps in flight; at one workgroup per CU (96 KiB of LDS) a
// long-K product such as phase 2 of the split ladder then waits on memory
// latency.  Here the loads go straight to LDS without staging registers: with
// three stages, two are in flight while the third is computed (192 x 192:
// 144 KiB, one workgroup per CU); with two stages a 128 x 192 tile takes
// 80 KiB, two workgroups per CU.  Each stage holds the same four planes (A hi,
// A lo, B hi, B lo of [rows][32 k], XOR-swizzled 16-byte chunks) as the
// register-staged kernel, so the MFMA step and the epilogue are shared.  A direct load writes lane l's 16 bytes to
// LDS at base + 16 l; the swizzle is therefore applied on the source side:
// slot s of a plane (row s / 4, LDS chunk s % 4) loads global chunk
// (s % 4) ^ ((row >> 1) & 3).  Rows beyond M or N and chunks at or beyond K
// load from the zero page, as in the register-staged kernel.
//
// Per k-step: wait until only the later stages' loads are outstanding
// (s_waitcnt vmcnt(2 (TM + TN) / 64) with one stage ahead: that many direct
// loads per wave per stage), barrier, issue stage + NST - 1 into the buffer
// every wave has just finished with, then the MFMAs.  The barrier is a bare s_barrier: a __syncthreads would also wait
// for the loads still in flight.
#include "xgemm.hpp"

namespace tci {

namespace {

template <int TM, int TN, int NST, int PREC, int NBM, int NBN>
__device__ __forceinline__ void xgemmd_body(const Task& t, const Seg* __restrict__ segs, int64_t J0, int seg0,
                                            const float* __restrict__ scales, uint16_t* sm, int w, int lane,
                                            int wm, int wn, int li, int lg) {
  constexpr int STG = 2 * (TM + TN) * XK;
  constexpr int PER = 2 * (TM + TN) / 64;  // direct loads per wave per stage
  static_assert(NST == 2 || NST == 3, "two or three stages");
  static_assert(PER <= 15, "vmcnt field");
  f4 acc[NBM > 0 ? NBM : 1][NBN > 0 ? NBN : 1];
#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // load cursor: the next stage to issue; stages 0 .. NST - 2 go out first
  int lseg = seg0, lkb = 0;
  dstage<TM, TN>(segs[lseg], t, J0, lkb, w, lane, sm);
  bool lmore = dadvance(segs, t.seg1, lseg, lkb);
  bool has1 = false;  // with three stages: the stage after the current one is in flight
  if (NST == 3 && lmore) {
    dstage<TM, TN>(segs[lseg], t, J0, lkb, w, lane, sm + STG);
    lmore = dadvance(segs, t.seg1, lseg, lkb);
    has1 = true;
  }
  int cb = 0;
  while (true) {
    // s_waitcnt immediate: vmcnt in bits 3:0 (and 15:14), expcnt 6:4, lgkmcnt 11:8
    if (has1) __builtin_amdgcn_s_waitcnt(0x0F70 | PER);  // vmcnt(PER): this stage has landed
    else __builtin_amdgcn_s_waitcnt(0x0F70);             // vmcnt(0)
    __builtin_amdgcn_s_barrier();  // every wave's loads of this stage are in LDS; the oldest buffer is free
    bool issued = false;
    if (lmore) {
      dstage<TM, TN>(segs[lseg], t, J0, lkb, w, lane, sm + ((cb + NST - 1) % NST) * STG);
      lmore = dadvance(segs, t.seg1, lseg, lkb);
      issued = true;
    }
    xgemm_step<TM, TN, PREC, NBM, NBN>(sm + cb * STG, acc, wm, wn, li, lg);
    if (!has1 && !issued) break;
    cb = (cb + 1) % NST;
    has1 = NST == 3 ? issued : false;
  }
  xgemm_store<TM, TN, PREC, NBM, NBN>(t, scales, acc, wm, wn, li, lg);
}

template <int TM, int TN, int NST, int PREC>
__global__ __launch_bounds__(256, 2 * (TM + TN) * XK * 2 * NST > 81920 ? 1 : 2) void xgemmd_kernel(
    const Task* __restrict__ tasks, const Seg* __restrict__ segs, int64_t J0, const float* __restrict__ scales,
    int remap) {
  constexpr int WBM = TM / 32, WBN = TN / 32;  // 16x16 blocks per wave side
  constexpr int BS = XInterleave<TM, TN>::BS;
  const Task t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
  __shared__ __attribute__((aligned(16))) uint16_t sm[NST * 2 * (TM + TN) * XK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  // blocks of a wave as in the register-staged kernel: round-robin (192 x 192)
  // or quadrants
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
#define TCI_DB(M_, N_) \
  case M_ * 8 + N_:    \
    if constexpr (M_ <= WBM && N_ <= WBN)                                                           \
      xgemmd_body<TM, TN, NST, PREC, M_, N_>(t, segs, J0, seg0, scales, sm, w, lane, wm, wn, li, lg); \
    break;
#define TCI_DROW(M_) TCI_DB(M_, 0) TCI_DB(M_, 1) TCI_DB(M_, 2) TCI_DB(M_, 3) TCI_DB(M_, 4) TCI_DB(M_, 5) TCI_DB(M_, 6)
  switch (nbm * 8 + nbn) {
    TCI_DROW(0) TCI_DROW(1) TCI_DROW(2) TCI_DROW(3) TCI_DROW(4) TCI_DROW(5) TCI_DROW(6)
    default: break;
  }
#undef TCI_DROW
#undef TCI_DB
}

template <int TM, int TN, int NST>
void launch_xgemmd_t(int precision, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0,
                     const float* scales, hipStream_t st, int remap) {
  if (precision == TCI_SIGMA_BF16X3)
    hipLaunchKernelGGL((xgemmd_kernel<TM, TN, NST, TCI_SIGMA_BF16X3>), dim3((unsigned)ntasks), dim3(256), 0, st, tasks,
                       segs, J0, scales, remap);
  else if (precision == TCI_SIGMA_FP16X3)
    hipLaunchKernelGGL((xgemmd_kernel<TM, TN, NST, TCI_SIGMA_FP16X3>), dim3((unsigned)ntasks), dim3(256), 0, st, tasks,
                       segs, J0, scales, remap);
  else
    throw Error(TCI_ERR_ARG, "split GEMM needs a split precision");
}

}  // namespace

void launch_xgemmd(int precision, int tm, int tn, int nst, const Task* tasks, const Seg* segs, int64_t ntasks,
                   int64_t J0, const float* scales, hipStream_t st, int remap) {
  if (ntasks <= 0) return;
  TCI_REQUIRE(ntasks < ((int64_t)1 << 31), "too many gemm tasks");
  if (tm == 192 && tn == 192 && nst == 3)
    launch_xgemmd_t<192, 192, 3>(precision, tasks, segs, ntasks, J0, scales, st, remap);
  else if (tm == 128 && tn == 192 && nst == 2)
    launch_xgemmd_t<128, 192, 2>(precision, tasks, segs, ntasks, J0, scales, st, remap);
  else if (tm == 128 && tn == 192 && nst == 3)
    launch_xgemmd_t<128, 192, 3>(precision, tasks, segs, ntasks, J0, scales, st, remap);
  else if (tm == 128 && tn == 128 && nst == 2)
    launch_xgemmd_t<128, 128, 2>(precision, tasks, segs, ntasks, J0, scales, st, remap);
  else
    throw Error(TCI_ERR_ARG, "no direct-load split GEMM for " + std::to_string(tm) + "x" + std::to_string(tn) + ", " +
                                 std::to_string(nst) + " stages");
  TCI_HIP(hipGetLastError());
}

}  // namespace tci
