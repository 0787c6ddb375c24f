// Fused split-precision ladder for pair domains of 129-192 virtuals.
//
// The two-phase ladder writes T'_J = L_J C^T (d x d per walk and Cholesky
// vector J) to HBM in phase 1 and reads it back in phase 2.  Here one
// workgroup owns the output columns a0..a0+63 of R = sum_J L_J T'_J^T for one
// walk and a range of J, and for every J
//   phase 1  T'[a][c] = sum_k L_J[a][k] C[c][k]    (64 x 192, K = d)
//   phase 2  R[n][a] += sum_c L_J[n][c] T'[a][c]   (192 x 64, K = d)
// with T' kept in LDS as split rows (hi and lo planes of six 32-wide k-steps),
// so HBM carries L_J and C only.  Both phases are NT products on pre-split
// k-contiguous operands, staged through LDS like xgemm_kernel (same swizzle
// and fragment layouts), three MFMAs per product, FP32 accumulation; T' is
// scaled by scales[3] before it is split and R by scales[4] (fp16x3; all 1 for
// bf16x3), exactly as the two phases do.
//
// Eight waves: phase 1 gives each a 32 x 48 piece of T' (2 x 3 blocks of
// 16 x 16), phase 2 a 48 x 32 piece of the R columns (3 x 2 blocks).  Rows and
// columns beyond d read the zero page, so a domain below 192 costs 192-wide
// work.
#include "ladder_fused.hpp"
#include "xgemm.hpp"

namespace tci {

namespace {

constexpr int FA = 64, FW = 192, FKS = FW / XK;  // a-block, full width, k-steps of T'

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
  // stage: phase 1 A hi/lo (64 rows) + B hi/lo (192 rows); phase 2 uses the
  // first 2 x 192 rows of a stage for its A operand
  constexpr int STG = 2 * (FA + FW) * XK;
  __shared__ __attribute__((aligned(16))) uint16_t sm[2 * STG];
  __shared__ __attribute__((aligned(16))) uint16_t tp[2 * FKS * FA * XK];  // T' hi planes, then lo planes
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int d = t.d, nks = (d + XK - 1) / XK;
  const float s_t = scales ? scales[3] : 1.f, s_r = scales ? scales[4] : 1.f;
  Op oc{};
  oc.base = t.C;
  oc.rs = t.ldx;
  f4 acc2[3][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc2[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  u32x4 va[2][2], vb[2][2];
  // phase 1 waves: rows 32 (w / 4) + 16 i of the a-block, columns 48 (w % 4) + 16 j;
  // phase 2 waves: rows 48 (w / 2) + 16 i, columns 32 (w % 2) + 16 j of the a-block
  const int p1r = 32 * (w >> 2), p1c = 48 * (w & 3), p2r = 48 * (w >> 1), p2c = 32 * (w & 1);
  for (int J = t.j0; J < t.j1; ++J) {
    Op ol{};
    ol.base = t.L + (int64_t)J * t.jstep;
    ol.rs = t.ldx;
    // ---- phase 1: T'[a0 + 0..63][0..191] over k = 0..d-1
    f4 acc1[2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) acc1[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    fload<FA>(ol, d, t.a0, 0, d, tid, va);
    fload<FW>(oc, d, 0, 0, d, tid, vb);
    __syncthreads();  // previous J's phase 2 is done with the stage buffers
    fstore<FA>(tid, va, sm, sm + FA * XK);
    fstore<FW>(tid, vb, sm + 2 * FA * XK, sm + 2 * FA * XK + FW * XK);
    __syncthreads();
    for (int ks = 0; ks < nks; ++ks) {
      const bool more = ks + 1 < nks;
      if (more) {
        fload<FA>(ol, d, t.a0, (ks + 1) * XK, d, tid, va);
        fload<FW>(oc, d, 0, (ks + 1) * XK, d, tid, vb);
      }
      const uint16_t* base = sm + (ks & 1) * STG;
      V ah[2], al[2], bh[3], bl[3];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int e = xsw(p1r + i * 16 + li, lg);
        ah[i] = *reinterpret_cast<const V*>(base + e);
        al[i] = *reinterpret_cast<const V*>(base + FA * XK + e);
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int e = xsw(p1c + j * 16 + li, lg);
        bh[j] = *reinterpret_cast<const V*>(base + 2 * FA * XK + e);
        bl[j] = *reinterpret_cast<const V*>(base + 2 * FA * XK + FW * XK + e);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc1[i][j] = XT<PREC>::mfma(bh[j], al[i], acc1[i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc1[i][j] = XT<PREC>::mfma(bl[j], ah[i], acc1[i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc1[i][j] = XT<PREC>::mfma(bh[j], ah[i], acc1[i][j]);
      if (more) {
        uint16_t* nb = sm + ((ks + 1) & 1) * STG;
        fstore<FA>(tid, va, nb, nb + FA * XK);
        fstore<FW>(tid, vb, nb + 2 * FA * XK, nb + 2 * FA * XK + FW * XK);
      }
      __syncthreads();
    }
    // ---- T' to LDS: lane holds T'[p1r + 16i + li][p1c + 16j + 4lg .. +3]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int c = p1c + 16 * j + 4 * lg;
        const int ks = c >> 5, ch = (c & 31) >> 3;
        const int e = ks * FA * XK + xsw(p1r + 16 * i + li, ch) + (c & 7);
        typedef T16 b4 __attribute__((ext_vector_type(4)));
        b4 h, l;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = acc1[i][j][r] * s_t;
          h[r] = (T16)v;
          l[r] = (T16)(v - (float)h[r]);
        }
        *reinterpret_cast<b4*>(tp + e) = h;
        *reinterpret_cast<b4*>(tp + FKS * FA * XK + e) = l;
      }
    // ---- phase 2: R[0..191][a0 + 0..63] += L_J[n][c] T'[a][c] over c = 0..d-1
    fload<FW>(ol, d, 0, 0, d, tid, va);
    fstore<FW>(tid, va, sm, sm + FW * XK);
    __syncthreads();
    for (int ks = 0; ks < nks; ++ks) {
      const bool more = ks + 1 < nks;
      if (more) fload<FW>(ol, d, 0, (ks + 1) * XK, d, tid, va);
      const uint16_t* base = sm + (ks & 1) * STG;
      const uint16_t* th = tp + ks * FA * XK;
      const uint16_t* tl = tp + FKS * FA * XK + ks * FA * XK;
      V ah[3], al[3], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int e = xsw(p2r + i * 16 + li, lg);
        ah[i] = *reinterpret_cast<const V*>(base + e);
        al[i] = *reinterpret_cast<const V*>(base + FW * XK + e);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int e = xsw(p2c + j * 16 + li, lg);
        bh[j] = *reinterpret_cast<const V*>(th + e);
        bl[j] = *reinterpret_cast<const V*>(tl + e);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc2[i][j] = XT<PREC>::mfma(bh[j], al[i], acc2[i][j]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc2[i][j] = XT<PREC>::mfma(bl[j], ah[i], acc2[i][j]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc2[i][j] = XT<PREC>::mfma(bh[j], ah[i], acc2[i][j]);
      if (more) {
        uint16_t* nb = sm + ((ks + 1) & 1) * STG;
        fstore<FW>(tid, va, nb, nb + FW * XK);
      }
      __syncthreads();
    }
  }
  // ---- R block: lane holds R[p2r + 16i + li][a0 + p2c + 16j + 4lg .. +3], accumulated
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int n = p2r + 16 * i + li;
    if (n >= d) continue;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int a = t.a0 + p2c + 16 * j + 4 * lg;
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
