/*
 * CPU restatement of the reference's 4-external ladder, in its own loop
 * structure, with OpenMP over the outer virtual index: the CPU baseline that
 * is timed beside the GPU, and a CPU check of the GPU ladder.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ and the CPU-baseline leg of
 * bench.py, never by the tigerci_amd product path.
 *
 * Loop structure (three_and_four_seg_mod_1.f90):
 *   all_external_3_4_loops_cho_2 (:318-469)   !$omp do over virtual a
 *   ext34_parallel_load (:472-632)            per a: gather the Cholesky
 *       vectors of the pairs (a,x) and of every virtual pair (c,d), d <= c,
 *       then one DGEMM (:556)  I(b, cd) = sum_J L^J_ab L^J_cd = (ab|cd)
 *   loop c = first virtual .. a, and every N-2 walk whose allowed-virtual
 *       domain holds a and c (the abcd records of build_ext_34seg_dryrun,
 *       :3142-3238)
 *   sort_ints_34_vec_lmo (:3520-3627), direct branch: over domain positions
 *       i > j (packed order (2,1),(3,1),(3,2),...)
 *         mod2 = (a D_j|c D_i) + (a D_i|c D_j), mod3 = the difference,
 *         mod4[i] = (a D_i|c D_i)
 *   ext_3_4_seg_loops_vec_lmo_res_2 (:3240-3518): DGEMV products of the CSF
 *       blocks with mod2/mod3/mod4 added to the (a,c) CSF of the walk
 *         a != c: singlet  += C_ab . mod2 + sqrt2 C_aa . mod4
 *                 triplet  += -(C_ab^T . mod3)
 *         a == c: aa[a]    += C_aa . mod4 + (C_ab . mod2) / sqrt2
 *                 (the triplet term there is (1/sqrt2) C_ab^T . mod3, and
 *                  mod3 vanishes identically for a == c by (ab|cd) = (cd|ab))
 * CSF layout and walk tables as in include/tigersigma.h.  BLAS enters through
 * function pointers (Fortran interface, 32-bit integers), so the caller
 * chooses the library; the reference links a system BLAS or MKL.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef void (*dgemm_fn)(const char* ta, const char* tb, const int* m, const int* n, const int* k, const double* alpha,
                         const double* a, const int* lda, const double* b, const int* ldb, const double* beta,
                         double* c, const int* ldc);
typedef void (*dgemv_fn)(const char* t, const int* m, const int* n, const double* alpha, const double* a,
                         const int* lda, const double* x, const int* incx, const double* beta, double* y,
                         const int* incy);

static int64_t rec(int64_t p, int64_t q) { /* packed orbital-pair record, 0-based */
  if (p < q) {
    int64_t t = p;
    p = q;
    q = t;
  }
  return p * (p + 1) / 2 + q;
}

static int64_t vcol(int64_t x, int64_t y) { /* packed virtual pair (max, min), 0-based */
  return x >= y ? x * (x + 1) / 2 + y : y * (y + 1) / 2 + x;
}

/* dot through DGEMV with one column, as the reference's dgemv('t', m, 1, ...) */
static double gemv_dot(dgemv_fn gemv, int m, double alpha, const double* a, const double* v) {
  if (m <= 0) return 0.0;
  const int one = 1;
  const double zero = 0.0;
  double out = 0.0;
  gemv("T", &m, &one, &alpha, a, &m, v, &one, &zero, &out, &one);
  return out;
}

/*
 * sigma[k] += (ladder part of H c)[k] for the outer virtuals a_list[0..n_a)
 * (every virtual when a_list is NULL).  L: records x n_chol (record
 * p(p+1)/2+q for orbitals p >= q, internal first).  Returns 0, or -1 when
 * scratch memory cannot be had.
 */
int tci_cpu_ladder(void* dgemm_ptr, void* dgemv_ptr, int64_t n_int, int64_t n_v, int64_t n_chol, const double* L,
                   int64_t n_pairs, const int32_t* pairs, const int64_t* dom_off, const int32_t* dom_idx,
                   const int64_t* aa_off, const int64_t* ab_off, const double* x, double* sigma,
                   const int64_t* a_list, int64_t n_a, int nthreads) {
  dgemm_fn gemm = (dgemm_fn)dgemm_ptr;
  dgemv_fn gemv = (dgemv_fn)dgemv_ptr;
  const int64_t nonred = n_v * (n_v + 1) / 2;
  const double sqrt2 = sqrt(2.0);
  int64_t dmax = 0;
  for (int64_t P = 0; P < n_pairs; ++P)
    if (dom_off[P + 1] - dom_off[P] > dmax) dmax = dom_off[P + 1] - dom_off[P];
  /* position of every virtual in every domain (-1: absent), walks per virtual */
  int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_pairs > 0 ? n_pairs : 1) * n_v);
  int64_t* wo = (int64_t*)calloc((size_t)n_v + 1, sizeof(int64_t));
  int32_t* wl = (int32_t*)malloc(sizeof(int32_t) * (size_t)(dom_off[n_pairs] > 0 ? dom_off[n_pairs] : 1));
  /* |cd) vectors of every virtual pair, n_chol x nonred (shared, read-only) */
  double* cd = (double*)malloc(sizeof(double) * (size_t)n_chol * nonred);
  if (!pos || !wo || !wl || !cd) {
    free(pos);
    free(wo);
    free(wl);
    free(cd);
    return -1;
  }
  for (int64_t e = 0; e < n_pairs * n_v; ++e) pos[e] = -1;
  for (int64_t P = 0; P < n_pairs; ++P)
    for (int64_t q = dom_off[P]; q < dom_off[P + 1]; ++q) {
      pos[P * n_v + dom_idx[q]] = (int32_t)(q - dom_off[P]);
      wo[dom_idx[q] + 1]++;
    }
  for (int64_t a = 0; a < n_v; ++a) wo[a + 1] += wo[a];
  {
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_v > 0 ? n_v : 1));
    if (!fill) {
      free(pos);
      free(wo);
      free(wl);
      free(cd);
      return -1;
    }
    memcpy(fill, wo, sizeof(int64_t) * (size_t)n_v);
    for (int64_t P = 0; P < n_pairs; ++P)
      for (int64_t q = dom_off[P]; q < dom_off[P + 1]; ++q) wl[fill[dom_idx[q]]++] = (int32_t)P;
    free(fill);
  }
  for (int64_t c = 0; c < n_v; ++c)
    for (int64_t d = 0; d <= c; ++d)
      memcpy(cd + (size_t)n_chol * vcol(c, d), L + (size_t)n_chol * rec(n_int + c, n_int + d),
             sizeof(double) * (size_t)n_chol);

  int failed = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : failed)
#endif
  {
    const int64_t m2max = dmax * (dmax - 1) / 2;
    double* ab = (double*)malloc(sizeof(double) * (size_t)n_chol * n_v);
    double* I = (double*)malloc(sizeof(double) * (size_t)n_v * nonred);
    double* mod2 = (double*)malloc(sizeof(double) * (size_t)(m2max > 0 ? m2max : 1));
    double* mod3 = (double*)malloc(sizeof(double) * (size_t)(m2max > 0 ? m2max : 1));
    double* mod4 = (double*)malloc(sizeof(double) * (size_t)(dmax > 0 ? dmax : 1));
    if (!ab || !I || !mod2 || !mod3 || !mod4) failed = 1;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int64_t ia = 0; ia < (a_list ? n_a : n_v); ++ia) {
      const int64_t a = a_list ? a_list[ia] : ia;
      if (failed || a < 0 || a >= n_v) continue;
      if (wo[a + 1] == wo[a]) continue; /* no walk allows a */
      /* (ab| for every virtual b, then I(b, cd) = (ab|cd): DGEMM (:556) */
      for (int64_t b = 0; b < n_v; ++b)
        memcpy(ab + (size_t)n_chol * b, L + (size_t)n_chol * rec(n_int + a, n_int + b), sizeof(double) * (size_t)n_chol);
      {
        const int m = (int)n_v, n = (int)nonred, k = (int)n_chol;
        const double one = 1.0, zero = 0.0;
        gemm("T", "N", &m, &n, &k, &one, ab, &k, cd, &k, &zero, I, &m);
      }
      for (int64_t c = 0; c <= a; ++c) {
        for (int64_t w = wo[a]; w < wo[a + 1]; ++w) {
          const int64_t P = wl[w];
          const int32_t pa = pos[P * n_v + a], pc = pos[P * n_v + c];
          if (pc < 0) continue;
          const int32_t* D = dom_idx + dom_off[P];
          const int d = (int)(dom_off[P + 1] - dom_off[P]);
          const int m2 = d * (d - 1) / 2;
          /* sort_ints_34_vec_lmo, direct branch */
          int cnt = 0;
          for (int i = 0; i < d; ++i) {
            const int64_t xi = D[i], ci = vcol(xi, c);
            for (int j = 0; j < i; ++j) {
              const int64_t yj = D[j];
              const double t1 = I[yj + n_v * ci];           /* (a D_j|c D_i) */
              const double t2 = I[xi + n_v * vcol(yj, c)];  /* (a D_i|c D_j) */
              mod2[cnt] = t1 + t2;
              mod3[cnt] = t1 - t2;
              ++cnt;
            }
            mod4[i] = I[xi + n_v * ci];
          }
          const int nsp = pairs[2 * P] == pairs[2 * P + 1] ? 1 : 2;
          const double* c_aa = x + aa_off[P];
          const double* c_ab = x + ab_off[P];
          if (a != c) {
            const int64_t tgt = (int64_t)pa * (pa - 1) / 2 + pc;
            const double vs = gemv_dot(gemv, d, sqrt2, c_aa, mod4) + gemv_dot(gemv, m2, 1.0, c_ab, mod2);
#ifdef _OPENMP
#pragma omp atomic
#endif
            sigma[ab_off[P] + tgt] += vs;
            if (nsp == 2) {
              const double vt = gemv_dot(gemv, m2, -1.0, c_ab + m2, mod3);
#ifdef _OPENMP
#pragma omp atomic
#endif
              sigma[ab_off[P] + m2 + tgt] += vt;
            }
          } else {
            const double va = gemv_dot(gemv, d, 1.0, c_aa, mod4) + gemv_dot(gemv, m2, 1.0 / sqrt2, c_ab, mod2);
#ifdef _OPENMP
#pragma omp atomic
#endif
            sigma[aa_off[P] + pa] += va;
          }
        }
      }
    }
    free(ab);
    free(I);
    free(mod2);
    free(mod3);
    free(mod4);
  }
  free(pos);
  free(wo);
  free(wl);
  free(cd);
  return failed ? -1 : 0;
}
