// Split-precision grouped gather-GEMM: the same Task/Seg tables as the FP64
// kernels, computed as  A B ~= Ahi Bhi + Ahi Blo + Alo Bhi  on 16-bit MFMA
// (v_mfma_f32_16x16x32_bf16 / _f16) with FP32 accumulation.
//
// Operands are read from FP64 or FP32 arrays through the same gather maps and
// split while they are staged: hi = round(x), lo = round(x - hi), so hi + lo
// carries ~16 significant bits (bf16) or ~22 (fp16, values scaled by 2^s with
// the per-launch exponent shift chosen by the caller so that no value leaves
// the fp16 range).  The neglected lo*lo term is below the FP32 accumulator's
// rounding.
//
// Tiles are 128 x 128 with BK = 32 (one MFMA k-step); LDS holds each 16-bit
// operand k-contiguous ([row][32 + 8 pad]) so a lane's 8-element fragment
// (A[i][8g..8g+7], lane i + 16g) is one 16-byte read.  Two stages of
// {Ahi, Alo, Bhi, Blo} = 80 KiB, one workgroup per CU.
// Result layout (f32 16x16): D[i][j] at lane j + 16*(i/4), register i%4.
#include "common.hpp"

namespace tci {

namespace {

constexpr int SM = 128, SK = 32, SKP = SK + 8;  // rows x k, padded row length in 16-bit elements

template <int PREC>
struct Split;
template <>
struct Split<TCI_SIGMA_BF16X3> {
  typedef __bf16 T;
  typedef bf16x8 V;
  static __device__ __forceinline__ void split(float x, T& hi, T& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
  }
  static __device__ __forceinline__ f4 mfma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <>
struct Split<TCI_SIGMA_FP16X3> {
  typedef _Float16 T;
  typedef h8 V;
  static __device__ __forceinline__ void split(float x, T& hi, T& lo) {
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
  }
  static __device__ __forceinline__ f4 mfma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

// thread -> (row, first k) of the 16 elements it stages per operand
__device__ __forceinline__ void slot(bool kc, int tid, int& r, int& kbase, int& kstride) {
  if (kc) { r = tid >> 1; kbase = (tid & 1) * 16; kstride = 1; }  // 16 consecutive k of one row
  else { r = tid & 127; kbase = (tid >> 7); kstride = 2; }        // k = kbase, kbase+2, ... (rows walk fastest)
}

struct SIdx {
  int seg, kb;
  int64_t aoff, boff;
  int32_t ka[16], kbr[16];
};

__device__ __forceinline__ bool sadvance(const Seg* segs, int seg1, int& seg, int& kb) {
  kb += SK;
  if (kb >= segs[seg].K) {
    kb = 0;
    ++seg;
    while (seg < seg1 && segs[seg].K <= 0) ++seg;
  }
  return seg < seg1;
}

__device__ __forceinline__ void sfixed(const Seg& S, const Task& t, int64_t J0, int tid, int64_t& aoff, int64_t& boff) {
  int m, ka, ks, n, kb, ks2;
  slot(S.A.cs == 1, tid, m, ka, ks);
  slot(S.B.rs == 1, tid, n, kb, ks2);
  const int r = t.m0 + m, c = t.n0 + n;
  const int rc = min(r, t.M - 1), cc = min(c, t.N - 1);
  const int64_t rr = S.A.rmap ? (int64_t)gld(S.A.rmap + rc) : (int64_t)rc;
  const int64_t ccol = S.B.cmap ? (int64_t)gld(S.B.cmap + cc) : (int64_t)cc;
  aoff = (r < t.M && rr >= 0) ? J0 * S.A.jstride + rr * S.A.rs : -1;
  boff = (c < t.N && ccol >= 0) ? J0 * S.B.jstride + ccol * S.B.cs : -1;
}

__device__ __forceinline__ void skmaps(const Seg& S, int kb, int tid, int32_t* ka, int32_t* kbr) {
  int m, k0a, ksa, n, k0b, ksb;
  slot(S.A.cs == 1, tid, m, k0a, ksa);
  slot(S.B.rs == 1, tid, n, k0b, ksb);
  const int kmax = S.K - 1;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int k = kb + k0a + q * ksa;
    const int kc = min(k, kmax);
    const int32_t va = S.A.cmap ? gld(S.A.cmap + kc) : kc;
    ka[q] = k <= kmax ? va : -1;
    const int k2 = kb + k0b + q * ksb;
    const int kc2 = min(k2, kmax);
    const int32_t vb = S.B.rmap ? gld(S.B.rmap + kc2) : kc2;
    kbr[q] = k2 <= kmax ? vb : -1;
  }
}

template <class TA, class TB>
__device__ __forceinline__ void sload_t(const Seg& S, const SIdx& x, float sa, float sb, float* a, float* b) {
  const TA* pa = static_cast<const TA*>(S.A.base);
  const TB* pb = static_cast<const TB*>(S.B.base);
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const bool oka = x.aoff >= 0 && x.ka[q] >= 0;
    const float va = (float)gld(pa + (oka ? x.aoff + (int64_t)x.ka[q] * S.A.cs : 0));
    a[q] = oka ? sa * va : 0.f;
    const bool okb = x.boff >= 0 && x.kbr[q] >= 0;
    const float vb = (float)gld(pb + (okb ? x.boff + (int64_t)x.kbr[q] * S.B.rs : 0));
    b[q] = okb ? sb * vb : 0.f;
  }
}

__device__ __forceinline__ void sload(const Seg& S, const SIdx& x, float scale_a, float scale_b, float* a, float* b) {
  const float sa = scale_a * (float)S.alpha;
  if (S.A.f32) {
    if (S.B.f32) sload_t<float, float>(S, x, sa, scale_b, a, b);
    else sload_t<float, double>(S, x, sa, scale_b, a, b);
  } else {
    if (S.B.f32) sload_t<double, float>(S, x, sa, scale_b, a, b);
    else sload_t<double, double>(S, x, sa, scale_b, a, b);
  }
}

template <int PREC>
__device__ __forceinline__ void sstore(const Seg& S, int tid, const float* a, const float* b,
                                       typename Split<PREC>::T* Ah, typename Split<PREC>::T* Al,
                                       typename Split<PREC>::T* Bh, typename Split<PREC>::T* Bl) {
  typedef typename Split<PREC>::T T;
  int m, k0, ks;
  slot(S.A.cs == 1, tid, m, k0, ks);
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    T h, l;
    Split<PREC>::split(a[q], h, l);
    const int k = k0 + q * ks;
    Ah[m * SKP + k] = h;
    Al[m * SKP + k] = l;
  }
  slot(S.B.rs == 1, tid, m, k0, ks);
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    T h, l;
    Split<PREC>::split(b[q], h, l);
    const int k = k0 + q * ks;
    Bh[m * SKP + k] = h;
    Bl[m * SKP + k] = l;
  }
}

template <int PREC>
__global__ __launch_bounds__(256, 1) void ggemm_split_kernel(const Task* __restrict__ tasks,
                                                             const Seg* __restrict__ segs, int64_t J0,
                                                             float scale_a, float scale_b) {
  typedef typename Split<PREC>::T T;
  typedef typename Split<PREC>::V V;
  const Task t = tasks[blockIdx.x];
  extern __shared__ char sm_raw[];
  T* sm = reinterpret_cast<T*>(sm_raw);
  const int STG = 4 * SM * SKP;  // elements per stage
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
  const int nbm = __builtin_amdgcn_readfirstlane(max(0, min(4, (t.M - (t.m0 + wm) + 15) >> 4)));
  const int nbn = __builtin_amdgcn_readfirstlane(max(0, min(4, (t.N - (t.n0 + wn) + 15) >> 4)));
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  int seg0 = t.seg0;
  while (seg0 < t.seg1 && segs[seg0].K <= 0) ++seg0;
  if (seg0 < t.seg1) {
    float ra[16], rb[16];
    SIdx nxt;
    nxt.seg = seg0;
    nxt.kb = 0;
    sfixed(segs[nxt.seg], t, J0, tid, nxt.aoff, nxt.boff);
    skmaps(segs[nxt.seg], nxt.kb, tid, nxt.ka, nxt.kbr);
    sload(segs[nxt.seg], nxt, scale_a, scale_b, ra, rb);
    sstore<PREC>(segs[nxt.seg], tid, ra, rb, sm, sm + SM * SKP, sm + 2 * SM * SKP, sm + 3 * SM * SKP);
    SIdx cur = nxt;
    bool more = sadvance(segs, t.seg1, nxt.seg, nxt.kb);
    if (more) {
      if (nxt.seg != cur.seg) sfixed(segs[nxt.seg], t, J0, tid, nxt.aoff, nxt.boff);
      skmaps(segs[nxt.seg], nxt.kb, tid, nxt.ka, nxt.kbr);
    }
    __syncthreads();
    int buf = 0;
    while (true) {
      SIdx nn;
      bool more2 = false;
      if (more) {
        sload(segs[nxt.seg], nxt, scale_a, scale_b, ra, rb);
        nn = nxt;
        more2 = sadvance(segs, t.seg1, nn.seg, nn.kb);
        if (more2) {
          if (nn.seg != nxt.seg) sfixed(segs[nn.seg], t, J0, tid, nn.aoff, nn.boff);
          skmaps(segs[nn.seg], nn.kb, tid, nn.ka, nn.kbr);
        }
      }
      const T* Ah = sm + buf * STG;
      const T* Al = Ah + SM * SKP;
      const T* Bh = Ah + 2 * SM * SKP;
      const T* Bl = Ah + 3 * SM * SKP;
      V ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ah[i] = *reinterpret_cast<const V*>(Ah + (wm + i * 16 + li) * SKP + 8 * lg);
        al[i] = *reinterpret_cast<const V*>(Al + (wm + i * 16 + li) * SKP + 8 * lg);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        bh[j] = *reinterpret_cast<const V*>(Bh + (wn + j * 16 + li) * SKP + 8 * lg);
        bl[j] = *reinterpret_cast<const V*>(Bl + (wn + j * 16 + li) * SKP + 8 * lg);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i < nbm && j < nbn) acc[i][j] = Split<PREC>::mfma(al[i], bh[j], acc[i][j]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i < nbm && j < nbn) acc[i][j] = Split<PREC>::mfma(ah[i], bl[j], acc[i][j]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i < nbm && j < nbn) acc[i][j] = Split<PREC>::mfma(ah[i], bh[j], acc[i][j]);
      if (!more) break;
      T* Wh = sm + (buf ^ 1) * STG;
      sstore<PREC>(segs[nxt.seg], tid, ra, rb, Wh, Wh + SM * SKP, Wh + 2 * SM * SKP, Wh + 3 * SM * SKP);
      __syncthreads();
      buf ^= 1;
      nxt = nn;
      more = more2;
    }
  }
  const double unscale = 1.0 / ((double)scale_a * (double)scale_b);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (i < nbm && j < nbn) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = t.m0 + wm + i * 16 + 4 * lg + r;
          const int col = t.n0 + wn + j * 16 + li;
          if (row < t.M && col < t.N) {
            const double v = (double)acc[i][j][r] * unscale;
            if (t.out_f32) {
              float* p = reinterpret_cast<float*>(t.C) + (int64_t)row * t.ldc + col;
              *p = t.accumulate ? (float)(*p + v) : (float)v;
            } else {
              double* p = t.C + (int64_t)row * t.ldc + col;
              *p = t.accumulate ? *p + v : v;
            }
          }
        }
      }
}

}  // namespace

void launch_ggemm_split(int precision, const Task* tasks, const Seg* segs, int64_t ntasks, int64_t J0,
                        float scale_a, float scale_b, hipStream_t st) {
  if (ntasks <= 0) return;
  TCI_REQUIRE(ntasks < ((int64_t)1 << 31), "too many gemm tasks");
  const size_t shm = 2 * 4 * SM * SKP * 2;
  if (precision == TCI_SIGMA_BF16X3) {
    auto k = ggemm_split_kernel<TCI_SIGMA_BF16X3>;
    TCI_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(k, dim3((unsigned)ntasks), dim3(256), shm, st, tasks, segs, J0, scale_a, scale_b);
  } else if (precision == TCI_SIGMA_FP16X3) {
    auto k = ggemm_split_kernel<TCI_SIGMA_FP16X3>;
    TCI_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(k, dim3((unsigned)ntasks), dim3(256), shm, st, tasks, segs, J0, scale_a, scale_b);
  } else {
    throw Error(TCI_ERR_ARG, "split GEMM needs a split precision");
  }
  TCI_HIP(hipGetLastError());
}

void GemmProgram::launch_split(int precision, hipStream_t st, int64_t J0, float scale_a, float scale_b) const {
  launch_ggemm_split(precision, d_tasks.as<Task>(), d_segs.as<Seg>(), (int64_t)tasks.size(), J0, scale_a, scale_b, st);
}

}  // namespace tci
