// Fused split-precision ladder for pair domains of 65-128 virtuals.
//
// The two-phase ladder writes T'_J = L_J C^T (d x d per walk and Cholesky
// vector J) to HBM in phase 1 and reads it back in phase 2.  For a domain of
// at most 128 both split operands of one J fit in LDS next to each other
// (2 x 66 KiB), so here one workgroup owns a whole walk for a range of J:
//   C_Q (split rows) stays in LDS for the whole range,
//   L_J (split rows) is staged in LDS once per J, and the next L_J is loaded
//   into registers while the current one is used,
//   phase 1  T'[a][c] = sum_k L_J[a][k] C[c][k]   (MFMA A = C rows, B = L_J rows)
//   phase 2  R[n][a] += sum_c L_J[n][c] T'[a][c]  (MFMA A = T', B = L_J rows)
// HBM then carries each L_J block once (4 bytes per element) and T' never
// leaves the registers: phase 1 computes the transpose T'^T, whose 16 x 16
// accumulator blocks hold in each lane four consecutive c of one a, which
// is the layout of an MFMA A operand once the 32 c of a k-step are taken in
// the order {4g..4g+3, 16+4g..16+4g+3} for lane group g; phase 2 reads the
// L_J operand in the same order (two 8-byte LDS reads per plane).  The k
// order inside one MFMA does not change the sum.
//
// Four waves, wave w owning the 16-column blocks w and w + 4 of a (T' and R),
// or eight waves owning one block each (two waves per SIMD hide each other's
// LDS and barrier waits, for twice the LDS reads); a wave keeps T'[a][0..127]
// and R[0..127][a] of its blocks in registers.
// The body is specialised on the domain's block count (5..8) and on the
// wave's block count (1 or 2), so every loop is unrolled and branch-free; a
// domain of d costs ceil(d/16) blocks per side and ceil(d/32) k-steps.  Rows
// beyond d and chunks starting at or beyond d stage zeros; the split rows keep
// their padding to a multiple of 8 at zero (as the two-phase kernels assume).
// T' is scaled by scales[3] before it is split and R by scales[4] (fp16x3;
// all 1 for bf16x3), as in the two phases; three MFMAs per product, FP32
// accumulation.  Each task accumulates its J range into its own zero-filled
// K-split slice of R.
#include "ladder_fused.hpp"
#include "xgemm.hpp"

namespace tci {

namespace {

constexpr int GW = 128;               // padded domain width
constexpr int GCH = GW / 8;           // 8-element chunks per row
constexpr int GRS = 2 * GW + 8;       // LDS row stride (16-bit elements): hi/lo chunk pairs + 16 B

// Stage a split-row matrix (rows 0..16 NB - 1): thread tid moves the chunk
// pairs e = tid + 256 q, row e / GCH, chunk e % GCH (hi 16 B, lo 16 B); rows
// >= d and chunks starting at or beyond d come from the zero page.  Rows from
// 16 NB on are never read.
// (NT threads: passes of NT chunk pairs; the 256 NB chunk pairs of rows
// 0..16 NB - 1 take (256 NB + NT - 1) / NT of them.)
template <int NB, int NT>
__device__ __forceinline__ void gload(const uint16_t* base, int64_t rs, int d, int tid, u32x4 (*v)[2]) {
  constexpr int Q = (256 * NB + NT - 1) / NT;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = tid + NT * q, r = e / GCH, ch = e % GCH;
    const u32x4* p = (e < 256 * NB && r < d && 8 * ch < d)
                         ? reinterpret_cast<const u32x4*>(base + (int64_t)r * rs + 16 * ch)
                         : xzero;
    v[q][0] = gld(p);
    v[q][1] = gld(p + 1);
  }
}

template <int NB, int NT>
__device__ __forceinline__ void gstore(int tid, const u32x4 (*v)[2], uint16_t* s) {
  constexpr int Q = (256 * NB + NT - 1) / NT;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = tid + NT * q, r = e / GCH, ch = e % GCH;
    if (e < 256 * NB) {
      u32x4* p = reinterpret_cast<u32x4*>(s + r * GRS + 16 * ch);
      p[0] = v[q][0];
      p[1] = v[q][1];
    }
  }
}

// One wave's whole task for a domain of NB 16-blocks per side, in a
// workgroup of NW waves, owning the a blocks w + NW ab for ab < NAB (NAB = 0:
// the wave only helps stage L_J); every loop is unrolled and branch-free.
template <int PREC, int NB, int NAB, int NW>
__device__ __forceinline__ void fused128_body(const FusedTask& t, int d, int j0, int j1, uint16_t* cs, uint16_t* ls,
                                              int tid, int w, int li, int lg, float s_t, float s_r) {
  typedef typename XT<PREC>::T T16;
  typedef typename XT<PREC>::V V;
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  constexpr int NKS = (NB + 1) / 2;  // 32-wide k-steps
  constexpr int NT = 64 * NW, QS = (256 * NB + NT - 1) / NT;
  constexpr int NA = NAB > 0 ? NAB : 1;
  u32x4 pf[QS][2];
  gload<NB, NT>(t.C, t.ldx, d, tid, pf);
  gstore<NB, NT>(tid, pf, cs);
  gload<NB, NT>(t.L + (int64_t)j0 * t.jstep, t.ldx, d, tid, pf);
  gstore<NB, NT>(tid, pf, ls);
  __syncthreads();
  f4 acc2[NA][NB];
#pragma unroll
  for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
    for (int nb2 = 0; nb2 < NB; ++nb2) acc2[ab][nb2] = f4{0.f, 0.f, 0.f, 0.f};
  const int c0 = 16 * (lg >> 1) + 4 * (lg & 1);  // element of c = 32s + 4lg in its row
  for (int J = j0; J < j1; ++J) {
    const bool more = J + 1 < j1;
    if (more) gload<NB, NT>(t.L + (int64_t)(J + 1) * t.jstep, t.ldx, d, tid, pf);
    if constexpr (NAB > 0) {
    // ---- phase 1: acc1[cb][ab] lane holds T'[16(w + 4ab) + li][16cb + 4lg + r].
    // Steps i = (k-step s, block pair cp); the fragments of step i + 1 are read
    // from LDS before the MFMAs of step i.
    constexpr int NPC = (NB + 1) / 2, NI1 = NKS * NPC;
    f4 acc1[2 * NKS][NAB];
#pragma unroll
    for (int cb = 0; cb < 2 * NKS; ++cb)
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab) acc1[cb][ab] = f4{0.f, 0.f, 0.f, 0.f};
    auto ldc = [&](int i, V* h, V* l) {
      const int ko = 16 * (4 * (i / NPC) + lg), cp = 2 * (i % NPC);
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (cp + u < NB) {
          const uint16_t* p = cs + (16 * (cp + u) + li) * GRS + ko;
          h[u] = *reinterpret_cast<const V*>(p);
          l[u] = *reinterpret_cast<const V*>(p + 8);
        }
    };
    auto ldl = [&](int sk, V* h, V* l) {
      const int ko = 16 * (4 * sk + lg);
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab) {
        const uint16_t* p = ls + (16 * (w + NW * ab) + li) * GRS + ko;
        h[ab] = *reinterpret_cast<const V*>(p);
        l[ab] = *reinterpret_cast<const V*>(p + 8);
      }
    };
    {
      V ch[2], cl[2], lh[NAB], ll[NAB];
      ldc(0, ch, cl);
      ldl(0, lh, ll);
#pragma unroll
      for (int i = 0; i < NI1; ++i) {
        const int cp = 2 * (i % NPC);
        const bool nexts = i + 1 < NI1 && (i + 1) % NPC == 0;
        V nch[2], ncl[2], nlh[NAB], nll[NAB];
        if (i + 1 < NI1) ldc(i + 1, nch, ncl);
        if (nexts) ldl((i + 1) / NPC, nlh, nll);
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (cp + u < NB) acc1[cp + u][ab] = XT<PREC>::mfma(ch[u], ll[ab], acc1[cp + u][ab]);
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (cp + u < NB) acc1[cp + u][ab] = XT<PREC>::mfma(cl[u], lh[ab], acc1[cp + u][ab]);
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (cp + u < NB) acc1[cp + u][ab] = XT<PREC>::mfma(ch[u], lh[ab], acc1[cp + u][ab]);
        if (i + 1 < NI1) {
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            ch[u] = nch[u];
            cl[u] = ncl[u];
          }
        }
        if (nexts) {
#pragma unroll
          for (int ab = 0; ab < NAB; ++ab) {
            lh[ab] = nlh[ab];
            ll[ab] = nll[ab];
          }
        }
      }
    }
    // ---- phase 2: acc2[ab][nb2] lane holds R[16 nb2 + li][16(w + 4ab) + 4lg + r].
    // T' of k-step s is split in registers when the step starts: th[ab] lane
    // row a = 16(w + 4ab) + li, k positions 0..3 = c 32s + 4lg + r, 4..7 =
    // c 32s + 16 + 4lg + r.  Steps i = (s, block pair np), fragments one ahead.
    constexpr int NI2 = NKS * NPC;
    auto ldn = [&](int i, V* h, V* l) {
      const int sk = i / NPC, np = 2 * (i % NPC);
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (np + u < NB) {
          // c = 32s + 4lg + r at chunk 4s + lg/2, c + 16 two chunks later
          const uint16_t* p = ls + (16 * (np + u) + li) * GRS + 64 * sk + c0;
          const u32x2 h0 = *reinterpret_cast<const u32x2*>(p);
          const u32x2 h1 = *reinterpret_cast<const u32x2*>(p + 32);
          const u32x2 l0 = *reinterpret_cast<const u32x2*>(p + 8);
          const u32x2 l1 = *reinterpret_cast<const u32x2*>(p + 40);
          const u32x4 hq = {h0[0], h0[1], h1[0], h1[1]};
          const u32x4 lq = {l0[0], l0[1], l1[0], l1[1]};
          h[u] = __builtin_bit_cast(V, hq);
          l[u] = __builtin_bit_cast(V, lq);
        }
    };
    {
      V nh[2], nl[2], th[NAB], tl[NAB];
      ldn(0, nh, nl);
#pragma unroll
      for (int i = 0; i < NI2; ++i) {
        const int sk = i / NPC, np = 2 * (i % NPC);
        if (i % NPC == 0) {
#pragma unroll
          for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const float v = acc1[2 * sk + h][ab][r] * s_t;
                const T16 hv = (T16)v;
                th[ab][4 * h + r] = hv;
                tl[ab][4 * h + r] = (T16)(v - (float)hv);
              }
        }
        V xh[2], xl[2];
        if (i + 1 < NI2) ldn(i + 1, xh, xl);
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (np + u < NB) acc2[ab][np + u] = XT<PREC>::mfma(th[ab], nl[u], acc2[ab][np + u]);
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (np + u < NB) acc2[ab][np + u] = XT<PREC>::mfma(tl[ab], nh[u], acc2[ab][np + u]);
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (np + u < NB) acc2[ab][np + u] = XT<PREC>::mfma(th[ab], nh[u], acc2[ab][np + u]);
        if (i + 1 < NI2) {
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            nh[u] = xh[u];
            nl[u] = xl[u];
          }
        }
      }
    }
    }
    if (more) {
      __syncthreads();  // every wave is done with L_J
      gstore<NB, NT>(tid, pf, ls);
      __syncthreads();
    }
  }
  // ---- R: lane holds R[16 nb2 + li][16(w + NW ab) + 4lg .. +3]
#pragma unroll
  for (int ab = 0; ab < NAB; ++ab) {
    const int a = 16 * (w + NW * ab) + 4 * lg;
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

template <int PREC, int NW>
__global__ __launch_bounds__(64 * NW, 1) void xladder_fused128_kernel(const FusedTask* __restrict__ tasks,
                                                                      const float* __restrict__ scales, int remap) {
  const FusedTask t = tasks[remap ? xcd_task(blockIdx.x, gridDim.x) : blockIdx.x];
  __shared__ __attribute__((aligned(16))) uint16_t cs[GW * GRS];
  __shared__ __attribute__((aligned(16))) uint16_t ls[GW * GRS];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int d = __builtin_amdgcn_readfirstlane(t.d);
  const int j0 = __builtin_amdgcn_readfirstlane(t.j0), j1 = __builtin_amdgcn_readfirstlane(t.j1);
  const int nb = (d + 15) >> 4;  // 5..8 for domains of 65..128
  const float s_t = scales ? scales[3] : 1.f, s_r = scales ? scales[4] : 1.f;
  // a wave runs the body for its count of blocks w + NW ab inside the domain
  // (four waves: 1 or 2; eight waves: 0 or 1); all bodies meet the same
  // barriers (one after staging, two per Cholesky vector after the first)
  const int nab = (w < nb ? 1 : 0) + (w + NW < nb ? 1 : 0);
#define TCI_F128(NB_, NAB_) \
  case NB_ * 4 + NAB_:      \
    if constexpr (NAB_ * NW <= 8)                                                     \
      fused128_body<PREC, NB_, NAB_, NW>(t, d, j0, j1, cs, ls, tid, w, li, lg, s_t, s_r); \
    break;
  switch (nb * 4 + nab) {
    TCI_F128(5, 0) TCI_F128(5, 1) TCI_F128(5, 2) TCI_F128(6, 0) TCI_F128(6, 1) TCI_F128(6, 2)
    TCI_F128(7, 0) TCI_F128(7, 1) TCI_F128(7, 2) TCI_F128(8, 0) TCI_F128(8, 1) TCI_F128(8, 2)
    default: break;
  }
#undef TCI_F128
}

}  // namespace

void launch_xladder_fused128(int precision, const FusedTask* tasks, int64_t ntasks, const float* scales,
                             hipStream_t st, int remap, int waves) {
  if (ntasks <= 0) return;
  TCI_REQUIRE(ntasks < ((int64_t)1 << 31), "too many fused ladder tasks");
  TCI_REQUIRE(waves == 4 || waves == 8, "fused ladder workgroups have 4 or 8 waves");
  TCI_REQUIRE(precision == TCI_SIGMA_BF16X3 || precision == TCI_SIGMA_FP16X3, "fused ladder needs a split precision");
  const dim3 g((unsigned)ntasks);
  if (precision == TCI_SIGMA_BF16X3 && waves == 4)
    hipLaunchKernelGGL((xladder_fused128_kernel<TCI_SIGMA_BF16X3, 4>), g, dim3(256), 0, st, tasks, scales, remap);
  else if (precision == TCI_SIGMA_BF16X3)
    hipLaunchKernelGGL((xladder_fused128_kernel<TCI_SIGMA_BF16X3, 8>), g, dim3(512), 0, st, tasks, scales, remap);
  else if (waves == 4)
    hipLaunchKernelGGL((xladder_fused128_kernel<TCI_SIGMA_FP16X3, 4>), g, dim3(256), 0, st, tasks, scales, remap);
  else
    hipLaunchKernelGGL((xladder_fused128_kernel<TCI_SIGMA_FP16X3, 8>), g, dim3(512), 0, st, tasks, scales, remap);
  TCI_HIP(hipGetLastError());
}

}  // namespace tci
