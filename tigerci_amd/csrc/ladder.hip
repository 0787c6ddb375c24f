// 4-external ladder, Cholesky-direct, grouped over all pair walks (FP64 MFMA).
//
//   R^P += sum_J  L^J[D,D] . C^P . L^J[D,D]        (L^J symmetric)
//
// = sum_cd (ac|bd) C^P_cd for every pair block, the term the reference builds
// per virtual pair (a,c) as (ab|cd) by DGEMM and contracts walk by walk with
// DGEMV (three_and_four_seg_mod_1.f90:472-632, 3240-3627).  Here each work
// item is (pair P, 64-row panel, J-chunk); a wave owns 16 rows a of R^P and
// for every 16-column block cb of C computes
//   phase 1:  T^T[cb][a] = sum_k C[k][cb] L[k][a]       (MFMA, 4 doubles/lane)
//   phase 2:  R[a][:]   += sum_{c in cb} T[a][c] L[c][:]
// The f64 MFMA result layout puts D[i][j] at lane j+16(i%4), register i/4, so
// after phase 1 lane l holds T[a=l%16][c=l/16+4r].  That is exactly the
// phase-2 A-operand layout (A[i][k] at lane i+16k) for k-step s = r once the
// summation index inside the block is permuted as c = l/16 + 4s, so T never
// leaves registers.
// C panels and L row panels are staged in LDS and shared by the 4 waves.
// Items are laid out so that consecutive workgroup ids on one XCD (ids
// congruent mod 8) walk the same pair: C^P and L^J[D,D] stay in that XCD's L2.
#include "ladder.hpp"

namespace tci {

namespace {

template <class LT>
__device__ __forceinline__ double ldL(const LT* p) {
  return (double)(*p);
}

template <int NBMAX, bool CACHE_B, class LT>
__global__ __launch_bounds__(256) void ladder_f64_kernel(const LadderItem* __restrict__ items,
                                                         int64_t n_items_per_xcd,
                                                         const int64_t* __restrict__ dom_off,
                                                         const int32_t* __restrict__ dom_idx,
                                                         const int64_t* __restrict__ cd_off,
                                                         const double* __restrict__ Cd,
                                                         const LT* __restrict__ Lvv, int64_t n_v,
                                                         double* __restrict__ partial,
                                                         int64_t chunk_stride) {
  // XCD-aware mapping: hardware dispatches workgroup b to XCD b % 8.
  const int64_t b = blockIdx.x;
  const int64_t item_id = (b % 8) * n_items_per_xcd + (b / 8);
  const LadderItem it = items[item_id];
  if (it.P < 0) return;  // padding item
  const int P = it.P;
  const int64_t off = dom_off[P];
  const int d = (int)(dom_off[P + 1] - off);
  const int NB = (d + 15) >> 4;
  const int dpad = NB * 16;

  extern __shared__ double smem[];
  double* Cp = smem;                   // [dpad][16]
  double* Lr = smem + dpad * 16;       // [16][dpad]
  int* Dl = reinterpret_cast<int*>(Lr + 16 * dpad);  // [dpad]

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  for (int e = tid; e < dpad; e += 256) Dl[e] = (e < d) ? dom_idx[off + e] : 0;
  __syncthreads();

  const double* Cb = Cd + cd_off[P];
  const int a0 = it.r0 + 16 * w;       // this wave's 16 rows
  const int acol = a0 + li;             // phase-1 B column of this lane
  const bool a_ok = acol < d;
  const int64_t a_phys = a_ok ? (int64_t)Dl[acol] : 0;

  d4 acc[NBMAX];
#pragma unroll
  for (int nb = 0; nb < NBMAX; ++nb) acc[nb] = d4{0.0, 0.0, 0.0, 0.0};

  double bcache[CACHE_B ? NBMAX * 4 : 1];

  for (int J = it.J0; J < it.J1; ++J) {
    const LT* LJ = Lvv + (int64_t)J * n_v * n_v;
    if constexpr (CACHE_B) {
#pragma unroll
      for (int ks = 0; ks < NBMAX * 4; ++ks) {
        const int k = ks * 4 + lg;
        bcache[ks] = (ks < NB * 4 && k < d && a_ok) ? ldL(LJ + (int64_t)Dl[k] * n_v + a_phys) : 0.0;
      }
    }
    for (int cb = 0; cb < NB; ++cb) {
      __syncthreads();
      for (int e = tid; e < dpad * 16; e += 256) {
        const int k = e >> 4, c = e & 15, cc = cb * 16 + c;
        Cp[e] = (k < d && cc < d) ? Cb[(int64_t)k * d + cc] : 0.0;
        const int cr = e / dpad, n = e - cr * dpad, crr = cb * 16 + cr;
        Lr[e] = (crr < d && n < d) ? ldL(LJ + (int64_t)Dl[crr] * n_v + Dl[n]) : 0.0;
      }
      __syncthreads();
      // phase 1: T^T block for (cb, this wave's rows)
      d4 t0 = d4{0.0, 0.0, 0.0, 0.0}, t1 = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < NBMAX * 4; ks += 2) {
        if (ks < NB * 4) {
          const int k0 = ks * 4 + lg, k1 = k0 + 4;
          double bf0, bf1;
          if constexpr (CACHE_B) {
            bf0 = bcache[ks];
            bf1 = bcache[ks + 1];
          } else {
            bf0 = (k0 < d && a_ok) ? ldL(LJ + (int64_t)Dl[k0] * n_v + a_phys) : 0.0;
            bf1 = (k1 < d && a_ok) ? ldL(LJ + (int64_t)Dl[k1] * n_v + a_phys) : 0.0;
          }
          t0 = __builtin_amdgcn_mfma_f64_16x16x4f64(Cp[k0 * 16 + li], bf0, t0, 0, 0, 0);
          t1 = __builtin_amdgcn_mfma_f64_16x16x4f64(Cp[k1 * 16 + li], bf1, t1, 0, 0, 0);
        }
      }
      const d4 T = t0 + t1;
      // phase 2: acc[nb] += T(a, c-block) . L(c-block, nb)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double af = T[s];
        const double* lrow = Lr + (lg + 4 * s) * dpad + li;
#pragma unroll
        for (int nb = 0; nb < NBMAX; ++nb)
          if (nb < NB) acc[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, lrow[nb * 16], acc[nb], 0, 0, 0);
      }
    }
  }
  // write this chunk's partial: lane l holds R[a0 + lg + 4r][nb*16 + li]
  double* out = partial + (int64_t)it.chunk * chunk_stride + cd_off[P];
#pragma unroll
  for (int nb = 0; nb < NBMAX; ++nb) {
    if (nb < NB) {
      const int n = nb * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = a0 + lg + 4 * r;
        if (a < d && n < d) out[(int64_t)a * d + n] = acc[nb][r];
      }
    }
  }
}

__global__ void ladder_reduce_kernel(const double* __restrict__ partial, int64_t chunk_stride,
                                     int n_chunks, double* __restrict__ R, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += partial[(int64_t)c * chunk_stride + e];
    R[e] = s;
  }
}

template <int NBMAX, bool CACHE_B, class LT>
void launch_one(const LadderLaunch& L, const LadderArgs& a, hipStream_t st) {
  const size_t shm = (size_t)L.dpad_max * 16 * 2 * sizeof(double) + (size_t)L.dpad_max * sizeof(int);
  auto kern = ladder_f64_kernel<NBMAX, CACHE_B, LT>;
  TCI_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  hipLaunchKernelGGL(kern, dim3((unsigned)(L.n_per_xcd * 8)), dim3(256), shm, st,
                     L.d_items.as<LadderItem>(), L.n_per_xcd, a.dom_off, a.dom_idx, a.cd_off, a.Cd,
                     static_cast<const LT*>(a.Lvv), a.n_v, a.partial, a.chunk_stride);
  TCI_HIP(hipGetLastError());
}

}  // namespace

void LadderLaunch::build(const std::vector<LadderItem>& all, const std::vector<double>& cost, int dmax,
                         hipStream_t st) {
  // greedy longest-processing-time assignment of pairs' items to the 8 XCDs,
  // each XCD's list ordered by decreasing cost
  std::vector<size_t> order(all.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
  std::vector<std::vector<LadderItem>> lists(8);
  std::vector<double> load(8, 0.0);
  // keep all items of one pair on one XCD: group by pair first
  std::vector<int> pair_xcd;
  for (size_t oi : order) {
    const LadderItem& it = all[oi];
    if ((int)pair_xcd.size() <= it.P) pair_xcd.resize(it.P + 1, -1);
    int x = pair_xcd[it.P];
    if (x < 0) {
      x = 0;
      for (int q = 1; q < 8; ++q)
        if (load[q] < load[x]) x = q;
      pair_xcd[it.P] = x;
    }
    lists[x].push_back(it);
    load[x] += cost[oi];
  }
  size_t m = 0;
  for (auto& l : lists) m = std::max(m, l.size());
  n_per_xcd = (int64_t)m;
  std::vector<LadderItem> flat(8 * m);
  for (int x = 0; x < 8; ++x)
    for (size_t i = 0; i < m; ++i) {
      LadderItem pad{};
      pad.P = -1;
      flat[x * m + i] = i < lists[x].size() ? lists[x][i] : pad;
    }
  n_items = (int64_t)all.size();
  dpad_max = ((dmax + 15) / 16) * 16;
  upload(d_items, flat, st);
}

void LadderLaunch::launch(const LadderArgs& a, hipStream_t st) const {
  if (n_items == 0) return;
  const int nb = dpad_max / 16;
  TCI_REQUIRE(nb <= 32, "pair domain larger than 512 virtuals is not supported by the ladder kernel");
  if (a.lvv_f32) {
    if (nb <= 8) launch_one<8, true, float>(*this, a, st);
    else if (nb <= 16) launch_one<16, true, float>(*this, a, st);
    else launch_one<32, false, float>(*this, a, st);
  } else {
    if (nb <= 8) launch_one<8, true, double>(*this, a, st);
    else if (nb <= 16) launch_one<16, true, double>(*this, a, st);
    else launch_one<32, false, double>(*this, a, st);
  }
}

void ladder_reduce(const double* partial, int64_t chunk_stride, int n_chunks, double* R, int64_t n,
                   hipStream_t st) {
  if (n <= 0) return;
  int64_t blocks = std::min<int64_t>(cdiv(n, 256), 8192);
  hipLaunchKernelGGL(ladder_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, st, partial, chunk_stride,
                     n_chunks, R, n);
  TCI_HIP(hipGetLastError());
}

}  // namespace tci
