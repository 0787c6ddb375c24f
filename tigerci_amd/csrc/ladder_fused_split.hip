// Fused split-precision ladder for pair domains of 129-192 virtuals.
//
// The two-phase ladder writes T'_J = L_J C^T (d x d per walk and Cholesky
// vector J) to HBM in phase 1 and reads it back in phase 2.  Here one
// workgroup owns a 64-wide panel c0..c0+63 of the c index for one walk and a
// range of J; the panel C[c0.., :] stays in LDS for all of its J, and for
// every J
//   phase 1  T'[a][c] = sum_k L_J[a][k] C[c][k]    (192 x 64, K = d)
//   phase 2  R[n][a] += sum_c L_J[n][c] T'[a][c]   (192 x 192, K = 64)
// with T' kept in LDS as split rows, so HBM carries L_J only; the panels of
// one walk and J read the same L_J block, and run next to each other (one XCD
// when the launch deals contiguous task ranges to the XCDs) so that they
// share it in L2.  R of a panel is a partial sum over c: every (panel, J
// range) writes its own K-split slice.  Both phases are NT products on
// pre-split k-contiguous operands with the xgemm_kernel LDS swizzle and
// fragment layouts, three MFMAs per product, FP32 accumulation; T' is scaled
// by scales[3] before it is split and R by scales[4] (fp16x3; all 1 for
// bf16x3), as in the two phases.
//
// Eight waves: phase 1 gives each a 48 x 32 piece of T' (3 x 2 blocks of
// 16 x 16), phase 2 a 48 x 96 piece of R (3 x 6 blocks).  Rows and columns
// beyond d read the zero page, so a domain below 192 costs 192-wide work.
#include "ladder_fused.hpp"
#include "xgemm.hpp"

namespace tci {

namespace {

constexpr int FP = 64, FW = 192, FKS = FW / XK;  // c panel, full width, k-steps of a row

// xload / xstore for 512 threads: local rows (tid >> 2) + 128 q below R
template <int R>
__device__ __forceinline__ void fload(const Op& O, int rows, int r0, int kb, int K, int tid, u32x4 (*v)[2]) {
  const uint16_t* base = static_cast<const uint16_t*>(O.base);
  const int k = kb + 8 * (tid & 3);
  const bool kok = k < K;
  const int64_t koff = (int64_t)(k >> 3) << 4;
#pragma unroll
  for (int q = 0; q < (R + 127) / 128; ++q) {
    const int lr = (tid >> 2) + 128 * q, r = r0 + lr;
    if (lr < R) {
      const u32x4* p = (kok && r < rows) ? reinterpret_cast<const u32x4*>(base + (int64_t)r * O.rs + koff) : xzero;
      v[q][0] = gld(p);
      v[q][1] = gld(p + 1);
    }
  }
}

template <int R>
__device__ __forceinline__ void fstore(int tid, const u32x4 (*v)[2], uint16_t* hi, uint16_t* lo) {
#pragma unroll
  for (int q = 0; q < (R + 127) / 128; ++q) {
    const int lr = (tid >> 2) + 128 * q;
    if (lr < R) {
      const int e = xsw(lr, tid & 3);
      *reinterpret_cast<u32x4*>(hi + e) = v[q][0];
      *reinterpret_cast<u32x4*>(lo + e) = v[q][1];
    }
  }
}

template <int PREC>
__global__ __launch_bounds__(512, 1) void xladder_fused_kernel(const FusedTask* __restrict__ tasks,
                                                               const float* __restrict__ scales, int remap) {
  typedef typename XT<PREC>::T T16;
  typedef typename XT<PREC>::V V;
  const FusedTask t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
  constexpr int STG = 2 * FW * XK;                         // one stage: 192 rows, hi and lo
  __shared__ __attribute__((aligned(16))) uint16_t sm[2 * STG];
  __shared__ __attribute__((aligned(16))) uint16_t cpn[2 * FKS * FP * XK];  // C panel: hi planes, then lo
  __shared__ __attribute__((aligned(16))) uint16_t tp[2 * 2 * FW * XK];     // T' panel: 2 hi planes, 2 lo
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int d = t.d, nks = (d + XK - 1) / XK, c0 = t.a0;
  const int nk2 = c0 + XK < d ? 2 : 1;  // k-steps of phase 2 inside the domain
  const float s_t = scales ? scales[3] : 1.f, s_r = scales ? scales[4] : 1.f;
  // phase 1 waves: T' rows 48 (w / 2) + 16 i, panel columns 32 (w % 2) + 16 j;
  // phase 2 waves: R rows 48 (w / 2) + 16 i, columns 96 (w % 2) + 16 j
  const int wr = 48 * (w >> 1), p1c = 32 * (w & 1), p2c = 96 * (w & 1);
  u32x4 va[2][2];
  {  // the C panel, every k-step
    Op oc{};
    oc.base = t.C;
    oc.rs = t.ldx;
    for (int ks = 0; ks < nks; ++ks) {
      fload<FP>(oc, d, c0, ks * XK, d, tid, va);
      fstore<FP>(tid, va, cpn + ks * FP * XK, cpn + FKS * FP * XK + ks * FP * XK);
    }
  }
  f4 acc2[3][6];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) acc2[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int J = t.j0; J < t.j1; ++J) {
    Op ol{};
    ol.base = t.L + (int64_t)J * t.jstep;
    ol.rs = t.ldx;
    // ---- phase 1: T'[0..191][c0 + 0..63] over k = 0..d-1
    f4 acc1[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc1[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    fload<FW>(ol, d, 0, 0, d, tid, va);
    __syncthreads();  // the previous J's phase 2 is done with the stages and T'
    fstore<FW>(tid, va, sm, sm + FW * XK);
    __syncthreads();
    for (int ks = 0; ks < nks; ++ks) {
      const bool more = ks + 1 < nks;
      if (more) fload<FW>(ol, d, 0, (ks + 1) * XK, d, tid, va);
      const uint16_t* base = sm + (ks & 1) * STG;
      const uint16_t* ch_ = cpn + ks * FP * XK;
      const uint16_t* cl_ = cpn + FKS * FP * XK + ks * FP * XK;
      V ah[3], al[3], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int e = xsw(wr + i * 16 + li, lg);
        ah[i] = *reinterpret_cast<const V*>(base + e);
        al[i] = *reinterpret_cast<const V*>(base + FW * XK + e);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int e = xsw(p1c + j * 16 + li, lg);
        bh[j] = *reinterpret_cast<const V*>(ch_ + e);
        bl[j] = *reinterpret_cast<const V*>(cl_ + e);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc1[i][j] = XT<PREC>::mfma(bh[j], al[i], acc1[i][j]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc1[i][j] = XT<PREC>::mfma(bl[j], ah[i], acc1[i][j]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc1[i][j] = XT<PREC>::mfma(bh[j], ah[i], acc1[i][j]);
      if (more) {
        uint16_t* nb = sm + ((ks + 1) & 1) * STG;
        fstore<FW>(tid, va, nb, nb + FW * XK);
      }
      __syncthreads();
    }
    // ---- T' to LDS: lane holds T'[wr + 16i + li][p1c + 16j + 4lg .. +3]
    // (panel-local columns, 2 k-steps of 32)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int c = p1c + 16 * j + 4 * lg;
        const int ks = c >> 5, ch = (c & 31) >> 3;
        const int e = ks * FW * XK + xsw(wr + 16 * i + li, ch) + (c & 7);
        typedef T16 b4 __attribute__((ext_vector_type(4)));
        b4 h, l;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = acc1[i][j][r] * s_t;
          h[r] = (T16)v;
          l[r] = (T16)(v - (float)h[r]);
        }
        *reinterpret_cast<b4*>(tp + e) = h;
        *reinterpret_cast<b4*>(tp + 2 * FW * XK + e) = l;
      }
    // ---- phase 2: R[0..191][0..191] += L_J[n][c0 + c] T'[a][c] over the panel
    fload<FW>(ol, d, 0, c0, d, tid, va);
    fstore<FW>(tid, va, sm, sm + FW * XK);
    __syncthreads();
    for (int ks = 0; ks < nk2; ++ks) {
      const bool more = ks + 1 < nk2;
      if (more) fload<FW>(ol, d, 0, c0 + (ks + 1) * XK, d, tid, va);
      const uint16_t* base = sm + (ks & 1) * STG;
      const uint16_t* th = tp + ks * FW * XK;
      const uint16_t* tl = tp + 2 * FW * XK + ks * FW * XK;
      V ah[3], al[3], bh[6], bl[6];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int e = xsw(wr + i * 16 + li, lg);
        ah[i] = *reinterpret_cast<const V*>(base + e);
        al[i] = *reinterpret_cast<const V*>(base + FW * XK + e);
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const int e = xsw(p2c + j * 16 + li, lg);
        bh[j] = *reinterpret_cast<const V*>(th + e);
        bl[j] = *reinterpret_cast<const V*>(tl + e);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) acc2[i][j] = XT<PREC>::mfma(bh[j], al[i], acc2[i][j]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) acc2[i][j] = XT<PREC>::mfma(bl[j], ah[i], acc2[i][j]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) acc2[i][j] = XT<PREC>::mfma(bh[j], ah[i], acc2[i][j]);
      if (more) {
        uint16_t* nb = sm + ((ks + 1) & 1) * STG;
        fstore<FW>(tid, va, nb, nb + FW * XK);
      }
      __syncthreads();
    }
  }
  // ---- R: lane holds R[wr + 16i + li][p2c + 16j + 4lg .. +3], accumulated
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int n = wr + 16 * i + li;
    if (n >= d) continue;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int a = p2c + 16 * j + 4 * lg;
      if (a >= d) continue;
      const f4 v = acc2[i][j] * s_r;
      double* p = t.R + (int64_t)n * d + a;
      if (a + 3 < d) {
        typedef double d4u __attribute__((ext_vector_type(4), aligned(8)));
        __attribute__((address_space(1))) d4u* q = (__attribute__((address_space(1))) d4u*)p;
        const d4u o = {(double)v[0], (double)v[1], (double)v[2], (double)v[3]};
        *q = *q + o;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (a + r < d) {
            __attribute__((address_space(1))) double* q = (__attribute__((address_space(1))) double*)(p + r);
            *q = *q + (double)v[r];
          }
      }
    }
  }
}

}  // namespace

void launch_xladder_fused(int precision, const FusedTask* tasks, int64_t ntasks, const float* scales, hipStream_t st,
                          int remap) {
  if (ntasks <= 0) return;
  TCI_REQUIRE(ntasks < ((int64_t)1 << 31), "too many fused ladder tasks");
  if (precision == TCI_SIGMA_BF16X3)
    hipLaunchKernelGGL(xladder_fused_kernel<TCI_SIGMA_BF16X3>, dim3((unsigned)ntasks), dim3(512), 0, st, tasks, scales,
                       remap);
  else if (precision == TCI_SIGMA_FP16X3)
    hipLaunchKernelGGL(xladder_fused_kernel<TCI_SIGMA_FP16X3>, dim3((unsigned)ntasks), dim3(512), 0, st, tasks, scales,
                       remap);
  else
    throw Error(TCI_ERR_ARG, "fused ladder needs a split precision");
  TCI_HIP(hipGetLastError());
}

}  // namespace tci
