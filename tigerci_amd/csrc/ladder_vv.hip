// Ladder term from the resident virtual four-index block.
//
// The ladder R_P[a][b] = sum_cd (ac|bd) C_P[c][d] is what the reference builds
// per outer virtual a from the Cholesky vectors (DGEMM of L^J blocks into
// (ac|bd) for that a, then a DGEMV per walk over its allowed virtuals;
// F/three_and_four_seg_mod_1.f90, ladder records of the 4-external loops).
// With 288 GB of HBM the whole virtual block can stay resident instead: it is
// assembled once per plan from L^J, and every sigma is then one GEMM over all
// walks, 2 n_p^2 n_walks flops, against 4 d^3 n_chol per walk for the
// Cholesky-direct two-phase products.
//
// Symmetric and antisymmetric parts.  C_P = C+ + C-, C+ symmetric, C-
// antisymmetric, over global virtual indices.  With packed pair indices
// p(a,b) = a(a+1)/2 + b (a >= b, n_p of them) and q(a,b) = a(a-1)/2 + b
// (a > b, n_m of them):
//   V+[p(a,b)][p(c,d)] = (ac|bd) + (ad|bc)      X+[P][p(c,d)] = C+_cd (c > d), C_cc / 2
//   V-[q(a,b)][q(c,d)] = (ac|bd) - (ad|bc)      X-[P][q(c,d)] = C-_cd
//   S+ = X+ V+,  S- = X- V-   (V+- symmetric)
//   R_P[a][b] = S+[p(a,b)] + S-[q(a,b)],  R_P[b][a] = S+[p(a,b)] - S-[q(a,b)]  (a > b),
//   R_P[a][a] = S+[p(a,a)].
// Walks with i == j have symmetric C_P and take no row of X-.  Entries of X
// outside a walk's domain are zero once per plan and never written again, so a
// sigma writes only the sum_P d_P (d_P + 1) / 2 entries its walks own.
//
// Layouts (row-major): V+ [n_p][n_p], V- [n_m][n_m] (FP64, plan lifetime);
// X+ [P][n_p], X- [P'][n_m]; S+, S- the same shapes, one copy per K-split.
// Split precisions keep V+- and X+- as split rows (hi/lo 16-bit parts,
// common.hpp) and compute S = X V^T on the split GEMM; FP64 uses the dense
// FP64 GEMM with V as the K x N operand.  The split product can also run as
// S^T = V X^T: with a few hundred walks, the walks then lie along the tiles'
// N side, where 192-wide tiles cover 518 walks in three tiles (V read three
// times) against five 128-row tiles the other way round; S^T is transposed
// back before the gather.
#include <algorithm>
#include <cmath>

#include "plan.hpp"

namespace tci {

namespace {

__host__ __device__ __forceinline__ int64_t tri(int64_t a) { return a * (a + 1) / 2; }

// (c, d), c >= d, of the packed index p = c(c+1)/2 + d
__device__ __forceinline__ void untri(int64_t p, int& c, int& d) {
  int64_t cc = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (tri(cc + 1) <= p) ++cc;
  while (tri(cc) > p) --cc;
  c = (int)cc;
  d = (int)(p - tri(cc));
}

// LT[r][J] = L^J[r], r < nv2 (r = a n_v + c)
template <class T>
__global__ void transpose_lvv_kernel(const T* __restrict__ L, int64_t n_chol, int64_t nv2, double* __restrict__ LT) {
  __shared__ double tile[32][33];
  const int64_t J0 = (int64_t)blockIdx.y * 32, r0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t J = J0 + yy, r = r0 + tx;
    tile[yy][tx] = (J < n_chol && r < nv2) ? (double)L[J * nv2 + r] : 0.0;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int64_t r = r0 + yy, J = J0 + tx;
    if (r < nv2 && J < n_chol) LT[r * n_chol + J] = tile[tx][yy];
  }
}

// Rows p(a,b) of V+ and q(a,b) of V- for a in [a0, a0 + na), from the slab
// W[(a - a0) n_v + c][b n_v + d] = (ac|bd).  blockIdx.y = (a - a0) n_v + b.
__global__ void vv_assemble_kernel(const double* __restrict__ W, int n_v, int a0, double* __restrict__ Vp,
                                   double* __restrict__ Vm) {
  const int al = (int)blockIdx.y / n_v, b = (int)blockIdx.y % n_v;
  const int a = a0 + al;
  if (b > a) return;
  const int64_t nv2 = (int64_t)n_v * n_v, np = tri(n_v), nm = tri(n_v - 1);
  const double* Wa = W + (int64_t)al * n_v * nv2 + (int64_t)b * n_v;
  const int64_t rp = tri(a) + b, rm = a > b ? tri(a - 1) + b : -1;
  for (int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; col < np; col += (int64_t)gridDim.x * blockDim.x) {
    int c, d;
    untri(col, c, d);
    const double x = Wa[(int64_t)c * nv2 + d];  // (ac|bd)
    const double y = Wa[(int64_t)d * nv2 + c];  // (ad|bc)
    Vp[rp * np + col] = x + y;
    if (rm >= 0 && c > d) Vm[rm * nm + tri(c - 1) + d] = x - y;
  }
}

__device__ __forceinline__ void put_split1(uint16_t* p, double v, float s, int f16) {
  const float x = (float)v * s;
  if (f16) {
    const _Float16 h = (_Float16)x;
    const _Float16 l = (_Float16)(x - (float)h);
    p[0] = __builtin_bit_cast(uint16_t, h);
    p[8] = __builtin_bit_cast(uint16_t, l);
  } else {
    const __bf16 h = (__bf16)x;
    const __bf16 l = (__bf16)(x - (float)h);
    p[0] = __builtin_bit_cast(uint16_t, h);
    p[8] = __builtin_bit_cast(uint16_t, l);
  }
}

// split rows of a row-major FP64 n x n matrix: out[r ldo + xidx(k)] (k padded to 8)
__global__ void split_square_kernel(const double* __restrict__ src, int64_t n, uint16_t* __restrict__ out,
                                    int64_t ldo, const float* scl, int f16) {
  const float s = scale_at(scl);
  const int64_t nch = ((n + 7) & ~(int64_t)7) >> 3;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * nch; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / nch, c = 8 * (e - r * nch);
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = c + t < n ? (float)src[r * n + c + t] : 0.f;
    put_split8(out + r * ldo + xidx(c), v, s, f16 != 0);
  }
}

// max |x| over n doubles into *out (float bits, atomic max of non-negative floats)
__global__ void vv_absmax_kernel(const double* __restrict__ x, int64_t n, unsigned* out) {
  float m = 0.f;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf((float)x[e]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

__global__ void vv_zero_kernel(unsigned* p) {
  if (threadIdx.x == 0) *p = 0u;
}

// scale of V (power of two with max|V| s in [2^13, 2^14)) into sc[kScV]
__global__ void vv_set_vscale_kernel(const unsigned* mx, float* sc, int slot) {
  if (threadIdx.x != 0) return;
  const float m = __uint_as_float(*mx);
  sc[slot] = m > 0.f ? ldexpf(1.f, 13 - ilogbf(m)) : 1.f;
}

// output scale of the split product: 1 / (scale of V * scale of C)
__global__ void vv_set_oscale_kernel(float* sc, int vslot, int cslot, int oslot) {
  if (threadIdx.x == 0) sc[oslot] = 1.f / (sc[vslot] * sc[cslot]);
}

// X+ / X- rows of every walk from its dense C block.  One thread per unordered
// local pair (r >= s); the global indices need not be sorted within a domain.
// FP64: Xp[P ldp + p], Xm[colm[P] ldm + q].  Split: the same rows as split
// rows (ldp, ldm in 16-bit elements), values times *scl.
template <bool SPLIT>
__global__ void vv_scatter_kernel(PairTab T, const double* __restrict__ C, const int32_t* __restrict__ colm,
                                  void* __restrict__ Xp, int64_t ldp, void* __restrict__ Xm, int64_t ldm,
                                  const float* scl, int f16) {
  const int P = blockIdx.y;
  const int64_t o = T.dom_off[P];
  const int d = (int)(T.dom_off[P + 1] - o);
  const double* Cp = C + T.cd_off[P];
  const int cm = colm[P];
  const float s = SPLIT ? scale_at(scl) : 1.f;
  const int64_t nt = (int64_t)d * (d + 1) / 2;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
    int r, c;
    untri(t, r, c);
    int gr = T.dom_idx[o + r], gc = T.dom_idx[o + c];
    double hr = Cp[(int64_t)r * d + c], rh = Cp[(int64_t)c * d + r];
    if (gr < gc) {  // orient as (higher, lower) global index
      const int g = gr;
      gr = gc;
      gc = g;
      const double v = hr;
      hr = rh;
      rh = v;
    }
    const double vp = r == c ? 0.5 * hr : 0.5 * (hr + rh);
    const int64_t kp = tri(gr) + gc;
    if (SPLIT) put_split1(static_cast<uint16_t*>(Xp) + P * ldp + xidx(kp), vp, s, f16);
    else static_cast<double*>(Xp)[P * ldp + kp] = vp;
    if (cm >= 0 && gr != gc) {
      const double vm = 0.5 * (hr - rh);
      const int64_t km = tri(gr - 1) + gc;
      if (SPLIT) put_split1(static_cast<uint16_t*>(Xm) + cm * ldm + xidx(km), vm, s, f16);
      else static_cast<double*>(Xm)[cm * ldm + km] = vm;
    }
  }
}

// R_P[r][s] (overwritten) from the K-split partials of S+ [P][n_p] and S- [P'][n_m]
__global__ void vv_gather_kernel(PairTab T, const double* __restrict__ Sp, int64_t sp_part, int64_t ldp,
                                 const double* __restrict__ Sm, int64_t sm_part, int64_t ldm, int nparts,
                                 const int32_t* __restrict__ colm, double* __restrict__ R) {
  const int P = blockIdx.y;
  const int64_t o = T.dom_off[P];
  const int d = (int)(T.dom_off[P + 1] - o);
  const int cm = colm[P];
  double* Rp = R + T.cd_off[P];
  const int64_t n2 = (int64_t)d * d;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n2; t += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(t / d), c = (int)(t - (int64_t)r * d);
    const int gr = T.dom_idx[o + r], gc = T.dom_idx[o + c];
    const int hi = max(gr, gc), lo = min(gr, gc);
    const double* sp = Sp + P * ldp + tri(hi) + lo;
    double v = 0.0;
    for (int k = 0; k < nparts; ++k) v += sp[k * sp_part];
    if (cm >= 0 && hi != lo) {
      const double* sm = Sm + cm * ldm + tri(hi - 1) + lo;
      double w = 0.0;
      for (int k = 0; k < nparts; ++k) w += sm[k * sm_part];
      v += gr > gc ? w : -w;
    }
    Rp[t] = v;
  }
}

// S[r][c] = sum over the nparts K-split partials of ST[c][r] (ST: [cols][rows]
// per partial, S: [rows][cols]); 32 x 32 tiles through LDS, 256 threads
__global__ void vv_untranspose_kernel(const double* __restrict__ ST, int64_t part, int64_t rows, int64_t cols,
                                      int nparts, double* __restrict__ S) {
  __shared__ double t[32][33];
  const int64_t c0 = (int64_t)blockIdx.x * 32, r0 = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int i = ty; i < 32; i += 8) {
    const int64_t c = c0 + i, r = r0 + tx;
    double v = 0.0;
    if (c < cols && r < rows)
      for (int k = 0; k < nparts; ++k) v += ST[k * part + c * rows + r];
    t[i][tx] = v;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int64_t r = r0 + i, c = c0 + tx;
    if (r < rows && c < cols) S[r * cols + c] = t[tx][i];
  }
}

inline unsigned blocks(int64_t work, unsigned cap) {
  return (unsigned)std::min<int64_t>(std::max<int64_t>(cdiv(work, 256), 1), cap);
}

Op vop(const void* base, int64_t rs) {
  Op o;
  o.base = base;
  o.rs = rs;
  o.cs = 1;
  o.rmap = nullptr;
  o.cmap = nullptr;
  o.jstride = 0;
  o.f32 = 0;
  return o;
}

Seg vseg(const Op& A, const Op& B, int K) {
  Seg s;
  s.A = A;
  s.B = B;
  s.K = K;
  s.alpha = 1.0;
  return s;
}

// Scale slots of the split form: V's scale and the product's output scale
// (slot 1 holds the scale of C, set per sigma by the split ladder's helper).
constexpr int kScV = 9, kScOut = 10, kScC = 1;

// S = X V (FP64: V as K x N; split: X V^T over split rows) for both parts as
// one program: tasks ordered by column block of V, then K-split, then row
// block of X, so the row tiles that read one V panel are neighbours.
void build_products(tci_plan& P, GemmProgram& g, bool split, int tile, int tile_n, int dma, int ks, bool trans,
                    hipStream_t st) {
  tci_plan::VVLadder& V = P.vv;
  g.clear();
  g.dense = !split;
  g.split = split;
  g.tile = tile;
  g.tile_n = split ? tile_n : 0;
  // 256 x 256 split tiles take `dma` as their wave count (4, 8 or 16); others as direct-load stages
  g.dma = split && tile == 256 ? 0 : dma;
  g.waves = split && tile == 256 && (dma == 8 || dma == 16) ? dma : 4;
  const int tn = split ? tile_n : tile;
  struct Part {
    const void* X;
    const void* Vm;
    double* S;
    int64_t n, rows, ldx;
  };
  const int64_t lx_p = split ? 2 * xpad(V.np) : V.np, lx_m = split ? 2 * xpad(V.nm) : V.nm;
  std::vector<Part> parts = {{split ? V.Xpx.p : V.Xp.p, split ? V.Vpx.p : V.Vp.p,
                              (trans ? V.STp : V.Sp).as<double>(), V.np, (int64_t)P.n_pairs, lx_p}};
  if (V.ncm > 0)
    parts.push_back({split ? V.Xmx.p : V.Xm.p, split ? V.Vmx.p : V.Vm.p, (trans ? V.STm : V.Sm).as<double>(), V.nm,
                     (int64_t)V.ncm, lx_m});
  if (trans) {
    TCI_REQUIRE(split, "the transposed product is a split-precision schedule");
    // S^T = V X^T: tasks by row block of V (tile rows of pairs), then K-split;
    // the program tiles the walks along N
    for (const Part& pt : parts) {
      const int64_t K = xpad(pt.n);
      const int64_t spart = pt.rows * pt.n;
      for (int64_t m0 = 0; m0 < pt.n; m0 += tile) {
        const int mm = (int)std::min<int64_t>(tile, pt.n - m0);
        for (int k = 0; k < ks; ++k) {
          const int64_t k0 = (K * k / ks) & ~(int64_t)15, k1 = k + 1 == ks ? K : (K * (k + 1) / ks) & ~(int64_t)15;
          TCI_REQUIRE(k1 > k0, "empty K-split range");
          const uint16_t* Vb = static_cast<const uint16_t*>(pt.Vm) + m0 * pt.ldx + 2 * k0;
          const uint16_t* Xb = static_cast<const uint16_t*>(pt.X) + 2 * k0;
          g.add(pt.S + k * spart + m0 * pt.rows, pt.rows, mm, (int)pt.rows,
                {vseg(vop(Vb, pt.ldx), vop(Xb, pt.ldx), (int)(k1 - k0))}, false, false, 0, Blocking(), kScOut);
        }
      }
    }
    g.upload_tables(st);
    return;
  }
  for (const Part& pt : parts) {
    const int64_t K = split ? xpad(pt.n) : pt.n;
    const int64_t spart = pt.rows * pt.n;  // elements of one partial S
    for (int64_t n0 = 0; n0 < pt.n; n0 += tn) {
      const int nn = (int)std::min<int64_t>(tn, pt.n - n0);
      for (int k = 0; k < ks; ++k) {
        // K ranges on multiples of 16 (8 for split rows suffices; 16 keeps both aligned)
        const int64_t k0 = (K * k / ks) & ~(int64_t)15, k1 = k + 1 == ks ? K : (K * (k + 1) / ks) & ~(int64_t)15;
        TCI_REQUIRE(k1 > k0, "empty K-split range");
        Seg s;
        if (split) {
          const uint16_t* Xb = static_cast<const uint16_t*>(pt.X) + 2 * k0;
          const uint16_t* Vb = static_cast<const uint16_t*>(pt.Vm) + n0 * pt.ldx + 2 * k0;
          s = vseg(vop(Xb, pt.ldx), vop(Vb, pt.ldx), (int)(k1 - k0));
        } else {
          const double* Xb = static_cast<const double*>(pt.X) + k0;
          const double* Vb = static_cast<const double*>(pt.Vm) + k0 * pt.n + n0;
          s = vseg(vop(Xb, pt.ldx), vop(Vb, pt.n), (int)(k1 - k0));
        }
        g.add(pt.S + k * spart + n0, pt.n, (int)pt.rows, nn, {s}, false, false, 0, Blocking(), split ? kScOut : -1);
      }
    }
  }
  g.upload_tables(st);
}

void alloc_partials(tci_plan& P, int ks, bool trans) {
  tci_plan::VVLadder& V = P.vv;
  const size_t need_p = (size_t)ks * P.n_pairs * V.np * 8, need_m = (size_t)ks * std::max(V.ncm, 1) * V.nm * 8;
  if (V.Sp.bytes < need_p) V.Sp.alloc(need_p);
  if (V.Sm.bytes < need_m) V.Sm.alloc(need_m);
  if (trans) {
    if (V.STp.bytes < need_p) V.STp.alloc(need_p);
    if (V.STm.bytes < need_m) V.STm.alloc(need_m);
  }
}

}  // namespace

bool vv_fits(const tci_plan& P, double* bytes_out) {
  const int64_t n_v = P.n_v, np = tri(n_v), nm = tri(n_v - 1);
  const double v64 = 8.0 * ((double)np * np + (double)nm * nm);
  // resident V (FP64 and one split copy), X and two K-split partials of S
  const double need = 1.5 * v64 + 8.0 * 3.0 * (double)P.n_pairs * (np + nm);
  if (bytes_out) *bytes_out = need;
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return false;
  // every Task and GEMM index of the products stays within 32-bit extents
  return n_v >= 2 && np < (int64_t)1 << 30 && need <= 64.0e9 && need + 48.0e9 <= (double)freeb;
}

double vv_flops(const tci_plan& P) {
  const double np = (double)tri(P.n_v), nm = (double)tri(P.n_v - 1);
  int ncm = 0;
  for (int Q = 0; Q < P.n_pairs; ++Q) ncm += P.h_pairs[2 * Q] != P.h_pairs[2 * Q + 1];
  return 2.0 * np * np * P.n_pairs + 2.0 * nm * nm * ncm;
}

// Assemble V+- (FP64) from L^J, in slabs of outer virtuals: W = LT[(a,c)][J]
// L^J[(b,d)] on the dense FP64 GEMM, then the symmetric and antisymmetric
// combinations.  Sets up X, the column map of X- and FP64 product tables.
void vv_build(tci_plan& P, hipStream_t st) {
  tci_plan::VVLadder& V = P.vv;
  const int n_v = P.n_v;
  const int64_t nv2 = (int64_t)n_v * n_v;
  V.np = tri(n_v);
  V.nm = tri(n_v - 1);
  std::vector<int32_t> colm(P.n_pairs, -1);
  V.ncm = 0;
  for (int Q = 0; Q < P.n_pairs; ++Q)
    if (P.h_pairs[2 * Q] != P.h_pairs[2 * Q + 1]) colm[Q] = V.ncm++;
  upload(V.colm, colm, st);
  V.Vp.alloc((size_t)V.np * V.np * 8);
  V.Vm.alloc((size_t)std::max<int64_t>(V.nm, 1) * std::max<int64_t>(V.nm, 1) * 8);
  {
    DevBuf LT;
    LT.alloc((size_t)nv2 * P.n_chol * 8);
    const dim3 gt((unsigned)cdiv(nv2, 32), (unsigned)cdiv(P.n_chol, 32));
    if (P.lvv_f32)
      hipLaunchKernelGGL(transpose_lvv_kernel<float>, gt, dim3(256), 0, st, P.Lvv.as<float>(), P.n_chol, nv2, LT.as<double>());
    else
      hipLaunchKernelGGL(transpose_lvv_kernel<double>, gt, dim3(256), 0, st, P.Lvv.as<double>(), P.n_chol, nv2,
                         LT.as<double>());
    TCI_HIP(hipGetLastError());
    // slab of na outer virtuals: na n_v^3 doubles, at most 4 GB and 65535 grid rows
    const int64_t per_a = (int64_t)n_v * nv2 * 8;
    const int na = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)n_v, (int64_t)4.0e9 / per_a, 65535 / n_v}));
    DevBuf W;
    W.alloc((size_t)na * per_a);
    for (int a0 = 0; a0 < n_v; a0 += na) {
      const int nb = std::min(na, n_v - a0);
      GemmProgram g;
      g.dense = true;
      g.tile = 128;
      Op A = vop(LT.as<double>() + (int64_t)a0 * n_v * P.n_chol, P.n_chol);
      Op B = vop(P.Lvv.p, nv2);
      B.f32 = P.lvv_f32 ? 1 : 0;
      g.add(W.as<double>(), nv2, nb * n_v, (int)nv2, {vseg(A, B, (int)P.n_chol)}, false);
      g.upload_tables(st);
      g.launch(st);
      hipLaunchKernelGGL(vv_assemble_kernel, dim3(blocks(V.np, 64), (unsigned)(nb * n_v)), dim3(256), 0, st,
                         W.as<double>(), n_v, a0, V.Vp.as<double>(), V.Vm.as<double>());
      TCI_HIP(hipGetLastError());
      TCI_HIP(hipStreamSynchronize(st));  // the tables and W are reused by the next slab
    }
  }
  V.Xp.alloc((size_t)P.n_pairs * V.np * 8);
  V.Xm.alloc((size_t)std::max(V.ncm, 1) * V.nm * 8);
  launch_fill(V.Xp.as<double>(), (int64_t)(V.Xp.bytes / 8), 0.0, st);
  launch_fill(V.Xm.as<double>(), (int64_t)(V.Xm.bytes / 8), 0.0, st);
  V.built = true;
  vv_set_schedule(P, TCI_SIGMA_FP64, 128, 0, 0, 1, st);
  TCI_HIP(hipStreamSynchronize(st));
}

void vv_set_schedule(tci_plan& P, int precision, int tile, int tile_n, int dma, int ks, hipStream_t st, bool trans) {
  tci_plan::VVLadder& V = P.vv;
  // every K-split range non-empty (at least 64 long), so every partial of S is written
  const int64_t kmin = V.ncm > 0 ? std::min(V.np, V.nm) : V.np;
  ks = (int)std::max<int64_t>(1, std::min<int64_t>(ks, kmin / 64));
  if (precision == TCI_SIGMA_FP64) {
    alloc_partials(P, ks, false);
    V.ks = ks;
    build_products(P, V.g64, false, tile, 0, 0, ks, false, st);
  } else {
    alloc_partials(P, ks, trans);
    V.ksx = ks;
    V.xtrans = trans;
    build_products(P, V.gx, true, tile, tile_n, dma, ks, trans, st);
  }
}

// Split copies of V for a split precision (fp16x3: scaled by a power of two
// from max|V| into scales[9]) and the split X rows, zeroed.
void vv_prepare_split(tci_plan& P, int precision, hipStream_t st) {
  tci_plan::VVLadder& V = P.vv;
  TCI_REQUIRE(V.built, "resident (ab|cd) block not built");
  const int f16 = precision == TCI_SIGMA_FP16X3 ? 1 : 0;
  float* sc = P.mx.scales.as<float>();
  if (f16) {
    unsigned* mx = P.mx.maxbits.as<unsigned>() + 3;
    hipLaunchKernelGGL(vv_zero_kernel, dim3(1), dim3(64), 0, st, mx);
    hipLaunchKernelGGL(vv_absmax_kernel, dim3(blocks(V.np * V.np, 8192)), dim3(256), 0, st, V.Vp.as<double>(),
                       V.np * V.np, mx);
    if (V.nm > 0)
      hipLaunchKernelGGL(vv_absmax_kernel, dim3(blocks(V.nm * V.nm, 8192)), dim3(256), 0, st, V.Vm.as<double>(),
                         V.nm * V.nm, mx);
    hipLaunchKernelGGL(vv_set_vscale_kernel, dim3(1), dim3(64), 0, st, mx, sc, kScV);
    TCI_HIP(hipGetLastError());
  }
  const int64_t lp = 2 * xpad(V.np), lm = 2 * xpad(V.nm);
  V.Vpx.alloc((size_t)V.np * lp * 2);
  V.Vmx.alloc((size_t)std::max<int64_t>(V.nm, 1) * lm * 2);
  hipLaunchKernelGGL(split_square_kernel, dim3(blocks(V.np * (lp / 16), 16384)), dim3(256), 0, st, V.Vp.as<double>(),
                     V.np, V.Vpx.as<uint16_t>(), lp, sc + kScV, f16);
  if (V.nm > 0)
    hipLaunchKernelGGL(split_square_kernel, dim3(blocks(V.nm * (lm / 16), 16384)), dim3(256), 0, st,
                       V.Vm.as<double>(), V.nm, V.Vmx.as<uint16_t>(), lm, sc + kScV, f16);
  TCI_HIP(hipGetLastError());
  V.Xpx.alloc((size_t)P.n_pairs * lp * 2);
  V.Xmx.alloc((size_t)std::max(V.ncm, 1) * lm * 2);
  launch_fill(static_cast<double*>(V.Xpx.p), (int64_t)(V.Xpx.bytes / 8), 0.0, st);
  launch_fill(static_cast<double*>(V.Xmx.p), (int64_t)(V.Xmx.bytes / 8), 0.0, st);
  V.prec = precision;
  // untuned schedule: 128 x 192 tiles, TCI_VV_TILE=256 (256w8, 256w16) for the
  // 32 x 32-MFMA kernel, TCI_VV_TILE=128x192T (128x192T<k>: k K-splits) for
  // the transposed product, TCI_VV_TILE=128x192d2 or 128x192d3 for two or
  // three direct-load stages
  const char* te = std::getenv("TCI_VV_TILE");
  const std::string ts = te ? te : "";
  if (te && std::atoi(te) == 256)
    vv_set_schedule(P, precision, 256, 256, ts == "256w8" ? 8 : (ts == "256w16" ? 16 : 0), 2, st);
  else if (ts.rfind("128x192T", 0) == 0)
    vv_set_schedule(P, precision, 128, 192, 0, ts.size() > 8 ? std::max(1, std::atoi(te + 8)) : 1, st, true);
  else if (ts == "128x192d2" || ts == "128x192d3")
    vv_set_schedule(P, precision, 128, 192, ts == "128x192d2" ? 2 : 3, 1, st);
  else if (ts == "128x128d2")
    vv_set_schedule(P, precision, 128, 128, 2, 1, st);
  else vv_set_schedule(P, precision, 128, 192, 0, 1, st);
  TCI_HIP(hipStreamSynchronize(st));
}

// R = ladder of the unpacked C (overwrites R), FP64 or the prepared split precision.
void vv_ladder(tci_plan& P, int precision, hipStream_t st) {
  tci_plan::VVLadder& V = P.vv;
  const bool split = precision != TCI_SIGMA_FP64;
  TCI_REQUIRE(V.built && (!split || V.prec == precision), "resident (ab|cd) block not prepared for this precision");
  const std::string pn = precision == TCI_SIGMA_FP64 ? "fp64" : (precision == TCI_SIGMA_FP16X3 ? "fp16x3" : "bf16x3");
  const int f16 = precision == TCI_SIGMA_FP16X3 ? 1 : 0;
  float* sc = P.mx.scales.as<float>();
  const dim3 gs(blocks((int64_t)P.dmax * (P.dmax + 1) / 2, 64), (unsigned)P.n_pairs);
  int e0 = P.kt_begin(st);
  if (split) {
    if (f16) {  // scale of C from max|C|, then the output scale 1 / (s_V s_C)
      launch_split_scales(1, P.C.as<double>(), P.sum_d2, P.mx.maxbits.as<unsigned>(), 0, sc, (float)P.dmax, st);
      hipLaunchKernelGGL(vv_set_oscale_kernel, dim3(1), dim3(64), 0, st, sc, kScV, kScC, kScOut);
    }
    hipLaunchKernelGGL(vv_scatter_kernel<true>, gs, dim3(256), 0, st, P.tab, P.C.as<double>(), V.colm.as<int32_t>(),
                       V.Xpx.p, 2 * xpad(V.np), V.Xmx.p, 2 * xpad(V.nm), f16 ? sc + kScC : nullptr, f16);
  } else {
    hipLaunchKernelGGL(vv_scatter_kernel<false>, gs, dim3(256), 0, st, P.tab, P.C.as<double>(), V.colm.as<int32_t>(),
                       V.Xp.p, V.np, V.Xm.p, V.nm, nullptr, 0);
  }
  TCI_HIP(hipGetLastError());
  P.kt_end(e0, pn + " ladder vv_scatter_kernel", 0.0, precision, st);
  const GemmProgram& g = split ? V.gx : V.g64;
  int ks = split ? V.ksx : V.ks;  // K-splits change the launch (and its traffic), so they name it
  P.kt_run(pn + " ladder (ab|cd) product" + (ks > 1 ? " K-split " + std::to_string(ks) : std::string()), g,
           precision, st, [&] {
    if (split) g.launch_split(precision, st, 0, sc);
    else g.launch(st);
  });
  if (split && V.xtrans) {  // S^T partials -> S (summed)
    e0 = P.kt_begin(st);
    const int64_t rp = P.n_pairs, rm = std::max(V.ncm, 1);
    hipLaunchKernelGGL(vv_untranspose_kernel, dim3((unsigned)cdiv(V.np, 32), (unsigned)cdiv(rp, 32)), dim3(256), 0,
                       st, V.STp.as<double>(), rp * V.np, rp, V.np, ks, V.Sp.as<double>());
    if (V.ncm > 0)
      hipLaunchKernelGGL(vv_untranspose_kernel, dim3((unsigned)cdiv(V.nm, 32), (unsigned)cdiv(rm, 32)), dim3(256), 0,
                         st, V.STm.as<double>(), rm * V.nm, rm, V.nm, ks, V.Sm.as<double>());
    TCI_HIP(hipGetLastError());
    P.kt_end(e0, pn + " ladder vv_untranspose_kernel", 0.0, precision, st);
    ks = 1;
  }
  e0 = P.kt_begin(st);
  hipLaunchKernelGGL(vv_gather_kernel, dim3(blocks((int64_t)P.dmax * P.dmax, 64), (unsigned)P.n_pairs), dim3(256), 0,
                     st, P.tab, V.Sp.as<double>(), (int64_t)P.n_pairs * V.np, V.np, V.Sm.as<double>(),
                     (int64_t)std::max(V.ncm, 1) * V.nm, V.nm, ks, V.colm.as<int32_t>(), P.R.as<double>());
  TCI_HIP(hipGetLastError());
  P.kt_end(e0, pn + " ladder vv_gather_kernel", 0.0, precision, st);
}

// Autotuned product schedule: tile shape and K-splits, then workgroup order.
// Returns the time of the whole resident ladder with the chosen schedule.
double vv_tune(tci_plan& P, int precision, hipStream_t st, std::string& log) {
  tci_plan::VVLadder& V = P.vv;
  struct Cand {
    int tile, tile_n, dma, ks;
    bool trans = false;
  };
  std::vector<Cand> cands;
  if (precision == TCI_SIGMA_FP64) {
    for (int t : {128, 192})
      for (int ks : {1, 2, 4}) cands.push_back({t, 0, 0, ks});
  } else {
    for (int ks : {1, 2, 4}) {
      cands.push_back({128, 128, 0, ks});
      cands.push_back({128, 192, 0, ks});
      cands.push_back({192, 192, 0, ks});
      cands.push_back({192, 192, 3, ks});
      cands.push_back({256, 256, 0, ks});
      cands.push_back({256, 256, 8, ks});  // dma field: eight waves
      cands.push_back({256, 256, 16, ks});  // sixteen waves
      cands.push_back({128, 192, 2, ks});  // 128 x 192 with two or three direct-load stages
      cands.push_back({128, 192, 3, ks});
      cands.push_back({128, 128, 2, ks});
      cands.push_back({128, 192, 0, ks, true});  // S^T = V X^T
      cands.push_back({128, 128, 0, ks, true});
      cands.push_back({192, 192, 3, ks, true});
      cands.push_back({128, 192, 3, ks, true});
    }
  }
  auto label = [](const Cand& c) {
    return std::to_string(c.tile) + (c.tile_n ? "x" + std::to_string(c.tile_n) : "") +
           (c.dma ? (c.tile == 256 ? "w" + std::to_string(c.dma) : "d" + (c.tile == 192 ? "" : std::to_string(c.dma)))
                  : "") +
           (c.trans ? "T" : "") + "/" + std::to_string(c.ks);
  };
  double best = 1e300;
  Cand bc = cands[0];
  for (const Cand& c : cands) {
    vv_set_schedule(P, precision, c.tile, c.tile_n, c.dma, c.ks, st, c.trans);
    const double ms = time_ms(st, 2, [&] { vv_ladder(P, precision, st); });
    log += " " + label(c) + ":" + fmt_ms(ms);
    if (ms < best * 0.98) {
      best = ms;
      bc = c;
    }
  }
  vv_set_schedule(P, precision, bc.tile, bc.tile_n, bc.dma, bc.ks, st, bc.trans);
  GemmProgram& g = precision == TCI_SIGMA_FP64 ? V.g64 : V.gx;
  const float* sc = P.mx.scales.as<float>();
  const int o = tune_order(g, st, [&] {
    if (precision == TCI_SIGMA_FP64) g.launch(st);
    else g.launch_split(precision, st, 0, sc);
  });
  if (precision != TCI_SIGMA_FP64) {
    V.split_tile = bc.tile;
    V.split_tile_n = bc.tile_n;
    V.split_dma = bc.dma;
    V.split_ks = bc.ks;
    V.split_order = o;
    V.split_trans = bc.trans;
  }
  log += " -> " + label(bc) + " order " + std::to_string(o);
  TCI_HIP(hipStreamSynchronize(st));
  return time_ms(st, 2, [&] { vv_ladder(P, precision, st); });
}

int64_t vv_device_bytes(const tci_plan& P) {
  const tci_plan::VVLadder& V = P.vv;
  int64_t b = 0;
  for (const DevBuf* x : {&V.Vp, &V.Vm, &V.Vpx, &V.Vmx, &V.Xp, &V.Xm, &V.Xpx, &V.Xmx, &V.Sp, &V.Sm, &V.STp, &V.STm,
                          &V.colm})
    b += (int64_t)x->bytes;
  return b;
}

}  // namespace tci
