// Fused split-precision ladder for pair domains of 129-192 virtuals.
//
// The two-phase ladder writes T'_J = L_J C^T (d x d per walk and Cholesky
// vector J) to HBM in phase 1 and reads it back in phase 2.  For domains
// above 128 the two split operands of one J (2 x 147 KiB) do not fit in LDS
// together, so here LDS holds only C_Q, for the whole J range of the task,
// and the waves read L_J from global memory:
//   phase 1  T'[a][c] = sum_k L_J[a][k] C[c][k]   (MFMA A = C rows from LDS,
//            B = the wave's own L_J rows a, from HBM)
//   phase 2  R[n][a] += sum_c L_J[n][c] T'[a][c]  (MFMA A = T' in registers,
//            B = L_J rows n, all of them, from L2: the workgroup has just read
//            the block)
// HBM carries each L_J block once and T' never leaves the registers.  As in
// the kernel for 65-128, phase 1 computes T'^T so that each lane's
// accumulators are an MFMA A operand of phase 2 once the 32 c of a k-step are
// taken in the order {4g..4g+3, 16+4g..16+4g+3} for lane group g, and phase 2
// reads L_J in the same order (8-byte loads).  To bound the registers, the c
// range is done in three parts of up to 64: phase 1 of a part (its c blocks
// over all k, the wave's L_J rows read again from L2 after the first part),
// then the phase-2 k-steps of that part.
//
// Four waves; wave w owns the a blocks w, w + 4 and w + 8 (those inside the
// domain), with T'^T of one c part (up to 4 x 3 blocks) and R[0..191][its a]
// (12 x 3 blocks) in registers.  The body is specialised on the domain's
// block count (9..12) and the wave's block count (2 or 3): unrolled,
// branch-free loops; global fragments are loaded one step ahead of their
// MFMAs.  The J loop has no barrier: LDS is read-only after C is
// staged.  Rows beyond d and chunks starting at or beyond d read the zero
// page; the split rows keep their padding to a multiple of 8 at zero.  T' is
// scaled by scales[3] before it is split and R by scales[4] (fp16x3; all 1
// for bf16x3), three MFMAs per product, FP32 accumulation.  Each task
// accumulates its J range into its own zero-filled K-split slice of R.
#include "ladder_fused.hpp"
#include "xgemm.hpp"

namespace tci {

namespace {

constexpr int HW = 192;          // padded domain width
constexpr int HCH = HW / 8;      // 8-element chunks per row
constexpr int HRS = 2 * HW + 8;  // LDS row stride (16-bit elements): 784 B, 196 dwords = 4 mod 64 banks
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// Global fragments come from one lane base pointer per phase plus
// wave-uniform offsets, so no per-load address stays live across the J loop.
// Chunks starting at or beyond d are zeroed after the load (their products
// would meet zeros of T'); rows at or beyond d are read as they are, since
// they only reach rows of T' and R outside the domain, which are never
// stored; the buffer of L blocks carries a 16 KiB tail for those reads.

// phase-1 operand: hi and lo of the lane's 8 k values at base + off
template <class V>
__device__ __forceinline__ void lfrag(const uint16_t* base, int64_t off, bool ok, V& h, V& l) {
  const u32x4* p = reinterpret_cast<const u32x4*>(base + off);
  const u32x4 z = {0u, 0u, 0u, 0u};
  const u32x4 a = gld(p), b = gld(p + 1);
  h = __builtin_bit_cast(V, ok ? a : z);
  l = __builtin_bit_cast(V, ok ? b : z);
}

// phase-2 operand: c = 32s + 4g + r and c + 16 of one row: hi at base + off
// and base + off + 32, lo 8 elements later
template <class V>
__device__ __forceinline__ void nfrag(const uint16_t* base, int64_t off, bool ok0, bool ok1, V& h, V& l) {
  const u32x2* p0 = reinterpret_cast<const u32x2*>(base + off);
  const u32x2* p1 = reinterpret_cast<const u32x2*>(base + off + 32);
  const u32x2 z = {0u, 0u};
  u32x2 h0 = gld(p0), h1 = gld(p1), l0 = gld(p0 + 2), l1 = gld(p1 + 2);
  h0 = ok0 ? h0 : z;
  l0 = ok0 ? l0 : z;
  h1 = ok1 ? h1 : z;
  l1 = ok1 ? l1 : z;
  const u32x4 hq = {h0[0], h0[1], h1[0], h1[1]};
  const u32x4 lq = {l0[0], l0[1], l1[0], l1[1]};
  h = __builtin_bit_cast(V, hq);
  l = __builtin_bit_cast(V, lq);
}

// One c part (HF) of one Cholesky vector: phase 1 for the c blocks of the
// part over all k, then the phase-2 k-steps of the part.
template <int PREC, int NB, int NAB, int HF>
__device__ __forceinline__ void fused192_half(const uint16_t* Lj, int64_t ldx, int d, const uint16_t* cs, int w,
                                              int li, int lg, float s_t, f4 (&acc2)[NAB][NB]) {
  typedef typename XT<PREC>::T T16;
  typedef typename XT<PREC>::V V;
  constexpr int NKS = (NB + 1) / 2;  // 32-wide k-steps (and c-block pairs, n-block pairs)
  constexpr int H1 = 2;              // phase-2 k-steps (c-block pairs) per c part
  constexpr int P1 = 2, P2 = 2;      // global fragment slots of phases 1 and 2
  constexpr int sp0 = HF * H1, sp1 = sp0 + H1 < NKS ? sp0 + H1 : NKS;  // phase-2 k-steps of this part
  // ---- phase 1: acc1[q][ab] lane holds T'[16(w + 4ab) + li][16(2 sp0 + q) + 4lg + r]
  f4 acc1[2 * H1][NAB];
#pragma unroll
  for (int q = 0; q < 2 * H1; ++q)
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab) acc1[q][ab] = f4{0.f, 0.f, 0.f, 0.f};
  // lane bases: phase 1 row 16w + li, chunk lg; phase 2 row li, element
  // 16(lg >> 1) + 4(lg & 1) (c = 4lg in its chunk)
  const uint16_t* b1 = Lj + (16 * w + li) * ldx + 16 * lg;
  const uint16_t* b2 = Lj + li * ldx + 16 * (lg >> 1) + 4 * (lg & 1);
  V lh[P1][NAB], ll[P1][NAB];
  auto ldl = [&](int k, int slot) {
    const bool ok = 8 * (4 * k + lg) < d;
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab) lfrag(b1, 64 * ab * ldx + 64 * k, ok, lh[slot][ab], ll[slot][ab]);
  };
#pragma unroll
  for (int k = 0; k < P1; ++k)
    if (k < NKS) ldl(k, k);
#pragma unroll
  for (int k = 0; k < NKS; ++k) {
    const int sl = k % P1;
    const int ko = 16 * (4 * k + lg);
#pragma unroll
    for (int q = 0; q < 2 * (sp1 - sp0); q += 2) {
      const int cb = 2 * sp0 + q;
      V ch[2], cl[2];
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (cb + u < NB) {
          const uint16_t* p = cs + (16 * (cb + u) + li) * HRS + ko;
          ch[u] = *reinterpret_cast<const V*>(p);
          cl[u] = *reinterpret_cast<const V*>(p + 8);
        }
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (cb + u < NB) acc1[q + u][ab] = XT<PREC>::mfma(ch[u], ll[sl][ab], acc1[q + u][ab]);
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (cb + u < NB) acc1[q + u][ab] = XT<PREC>::mfma(cl[u], lh[sl][ab], acc1[q + u][ab]);
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (cb + u < NB) acc1[q + u][ab] = XT<PREC>::mfma(ch[u], lh[sl][ab], acc1[q + u][ab]);
    }
    if (k + P1 < NKS) ldl(k + P1, sl);
    __builtin_amdgcn_sched_barrier(0);  // keep the loads of later steps out of this one
  }
  // ---- phase 2 over the k-steps s of this half: T' of step s split in
  // registers when the step starts (th[ab] lane row a = 16(w + 4ab) + li,
  // k positions 0..3 = c 32s + 4lg + r, 4..7 = c 32s + 16 + 4lg + r); steps
  // i = (s, n-block pair np) with global fragments P2 - 1 steps ahead
  constexpr int NS2MAX = H1 * NKS;
  constexpr int ns2 = (sp1 - sp0) * NKS;
  V nh[P2][2], nl[P2][2];
  auto ldn = [&](int i, int slot) {
    const int s = sp0 + i / NKS, np = 2 * (i % NKS);
    const bool ok0 = 8 * (4 * s + (lg >> 1)) < d, ok1 = 8 * (4 * s + 2 + (lg >> 1)) < d;
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (np + u < NB) nfrag(b2, 16 * (np + u) * ldx + 64 * s, ok0, ok1, nh[slot][u], nl[slot][u]);
  };
#pragma unroll
  for (int k = 0; k < P2; ++k)
    if (k < ns2) ldn(k, k);
  V th[NAB], tl[NAB];
#pragma unroll
  for (int i = 0; i < NS2MAX; ++i) {
    if (i >= ns2) break;
    const int sl = i % P2;
    const int q0 = 2 * (i / NKS), np = 2 * (i % NKS);
    if (i % NKS == 0) {
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float v = acc1[q0 + h][ab][r] * s_t;
            const T16 hv = (T16)v;
            th[ab][4 * h + r] = hv;
            tl[ab][4 * h + r] = (T16)(v - (float)hv);
          }
    }
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (np + u < NB) acc2[ab][np + u] = XT<PREC>::mfma(th[ab], nl[sl][u], acc2[ab][np + u]);
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (np + u < NB) acc2[ab][np + u] = XT<PREC>::mfma(tl[ab], nh[sl][u], acc2[ab][np + u]);
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (np + u < NB) acc2[ab][np + u] = XT<PREC>::mfma(th[ab], nh[sl][u], acc2[ab][np + u]);
    if (i + P2 < ns2) ldn(i + P2, sl);
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int PREC, int NB, int NAB>
__device__ __forceinline__ void fused192_body(const FusedTask& t, int d, int j0, int j1, const uint16_t* cs,
                                              int w, int li, int lg, float s_t, float s_r) {
  const int64_t ldx = t.ldx;
  f4 acc2[NAB][NB];
#pragma unroll
  for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
    for (int nb2 = 0; nb2 < NB; ++nb2) acc2[ab][nb2] = f4{0.f, 0.f, 0.f, 0.f};
  for (int J = j0; J < j1; ++J) {
    const uint16_t* Lj = t.L + (int64_t)J * t.jstep;
    fused192_half<PREC, NB, NAB, 0>(Lj, ldx, d, cs, w, li, lg, s_t, acc2);
    fused192_half<PREC, NB, NAB, 1>(Lj, ldx, d, cs, w, li, lg, s_t, acc2);
    fused192_half<PREC, NB, NAB, 2>(Lj, ldx, d, cs, w, li, lg, s_t, acc2);
  }
  // ---- R: lane holds R[16 nb2 + li][16(w + 4ab) + 4lg .. +3]
#pragma unroll
  for (int ab = 0; ab < NAB; ++ab) {
    const int a = 16 * (w + 4 * ab) + 4 * lg;
    if (a >= d) continue;
#pragma unroll
    for (int nb2 = 0; nb2 < NB; ++nb2) {
      const int n = 16 * nb2 + li;
      if (n >= d) continue;
      const f4 v = acc2[ab][nb2] * s_r;
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

template <int PREC>
__global__ __launch_bounds__(256, 1) void xladder_fused192_kernel(const FusedTask* __restrict__ tasks,
                                                                  const float* __restrict__ scales, int remap) {
  const FusedTask t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
  __shared__ __attribute__((aligned(16))) uint16_t cs[HW * HRS];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const float s_t = scales ? scales[3] : 1.f, s_r = scales ? scales[4] : 1.f;
  const int d = __builtin_amdgcn_readfirstlane(t.d);
  const int j0 = __builtin_amdgcn_readfirstlane(t.j0), j1 = __builtin_amdgcn_readfirstlane(t.j1);
  const int nb = (d + 15) >> 4;  // 9..12 for domains of 129..192
  // C_Q rows 0..16 nb - 1 (zero beyond d), chunk pairs e = tid + 256 q
  for (int e = tid; e < 16 * nb * HCH; e += 256) {
    const int r = e / HCH, ch = e % HCH;
    const u32x4* p = (r < d && 8 * ch < d) ? reinterpret_cast<const u32x4*>(t.C + (int64_t)r * t.ldx + 16 * ch) : xzero;
    const u32x4 h = gld(p), l = gld(p + 1);
    u32x4* q = reinterpret_cast<u32x4*>(cs + r * HRS + 16 * ch);
    q[0] = h;
    q[1] = l;
  }
  __syncthreads();
#define TCI_F192(NB_, NAB_) \
  case NB_ * 4 + NAB_:      \
    fused192_body<PREC, NB_, NAB_>(t, d, j0, j1, cs, w, li, lg, s_t, s_r); \
    break;
  switch (nb * 4 + (w + 8 < nb ? 3 : 2)) {
    TCI_F192(9, 2) TCI_F192(9, 3) TCI_F192(10, 2) TCI_F192(10, 3)
    TCI_F192(11, 2) TCI_F192(11, 3) TCI_F192(12, 2) TCI_F192(12, 3)
    default: break;
  }
#undef TCI_F192
}

}  // namespace

void launch_xladder_fused192(int precision, const FusedTask* tasks, int64_t ntasks, const float* scales,
                             hipStream_t st, int remap) {
  if (ntasks <= 0) return;
  TCI_REQUIRE(ntasks < ((int64_t)1 << 31), "too many fused ladder tasks");
  if (precision == TCI_SIGMA_BF16X3)
    hipLaunchKernelGGL(xladder_fused192_kernel<TCI_SIGMA_BF16X3>, dim3((unsigned)ntasks), dim3(256), 0, st, tasks,
                       scales, remap);
  else if (precision == TCI_SIGMA_FP16X3)
    hipLaunchKernelGGL(xladder_fused192_kernel<TCI_SIGMA_FP16X3>, dim3((unsigned)ntasks), dim3(256), 0, st, tasks,
                       scales, remap);
  else
    throw Error(TCI_ERR_ARG, "fused ladder needs a split precision");
  TCI_HIP(hipGetLastError());
}

}  // namespace tci
