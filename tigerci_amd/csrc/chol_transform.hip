// AO -> MO transformation of Cholesky vectors on the GPU.
//
// The reference transforms each vector in two halves (first_half_transform,
// F/cholesky_transform.f90:305-364: a loop over the kept AO pairs (mu, nu)
// that adds C[nu][i] L^J(mu, nu) to H[J][mu][i], and C[mu][i] L^J(mu, nu) to
// H[J][nu][i] when mu != nu; second_half_transform :371-380: a DGEMM with C)
// and writes record ij = i(i+1)/2 + j of the MO vectors with its norm
// (write_MO_cho_vec :387-411).  Here the pair list is scattered into dense
// symmetric L^J (n_ao x n_ao), and both halves are GEMMs on the FP64 MFMA
// kernel: H[(J, mu)][i] = sum_nu L^J[mu][nu] C[nu][i] as one product over all
// vectors, then MO^J = C^T H^J as one grouped launch of n_chol products; a
// gather writes the records [ij][J] and a reduction their norms.
#include <unordered_set>

#include "common.hpp"
#include "terms.hpp"

namespace tci {

namespace {

// L[J][mu][nu] = L[J][nu][mu] = vec[J][k] for pair k = (mu, nu)
__global__ void scatter_ao_pairs_kernel(const int32_t* __restrict__ pairs, const double* __restrict__ vec,
                                        int64_t n_pairs, int n_ao, double* __restrict__ L) {
  const int64_t J = blockIdx.y;
  const double* v = vec + J * n_pairs;
  double* LJ = L + J * (int64_t)n_ao * n_ao;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n_pairs; k += (int64_t)gridDim.x * blockDim.x) {
    const int mu = pairs[2 * k], nu = pairs[2 * k + 1];
    const double x = v[k];
    LJ[(int64_t)mu * n_ao + nu] = x;
    LJ[(int64_t)nu * n_ao + mu] = x;
  }
}

// rec[ij][J] = MO[J][j][i] for ij = i(i+1)/2 + j, i >= j (J fastest: coalesced writes)
__global__ void pack_mo_records_kernel(const double* __restrict__ MO, int n_mo, int64_t n_chol, double* __restrict__ rec) {
  const int64_t n_rec = (int64_t)n_mo * (n_mo + 1) / 2;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n_rec * n_chol;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ij = t / n_chol, J = t - ij * n_chol;
    int64_t i = (int64_t)((sqrt(8.0 * (double)ij + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= ij) ++i;
    while (i * (i + 1) / 2 > ij) --i;
    const int64_t j = ij - i * (i + 1) / 2;
    rec[t] = MO[(J * n_mo + j) * n_mo + i];
  }
}

// norms[r] = |rec[r][:]|_2, one wavefront per record
__global__ void record_norms_kernel(const double* __restrict__ rec, int64_t n_rec, int64_t n_chol,
                                    double* __restrict__ norms) {
  const int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= n_rec) return;
  double s = 0.0;
  for (int64_t J = lane; J < n_chol; J += 64) {
    const double x = rec[r * n_chol + J];
    s += x * x;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) norms[r] = sqrt(fabs(s));
}

Op cop(const void* base, int64_t rs) {
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

Seg cseg(const Op& A, const Op& B, int K) {
  Seg s;
  s.A = A;
  s.B = B;
  s.K = K;
  s.alpha = 1.0;
  return s;
}

}  // namespace

}  // namespace tci

using namespace tci;

extern "C" int tci_cholesky_ao_to_mo(int32_t n_ao, int32_t n_mo, int64_t n_chol, int64_t n_ao_pairs,
                                     const int32_t* ao_pairs, const double* ao_vec, const double* mo_coef,
                                     double* mo_records, double* norms) {
  try {
    TCI_REQUIRE(ao_pairs && ao_vec && mo_coef && mo_records, "null argument");
    TCI_REQUIRE(n_ao >= 1 && n_mo >= 1 && n_chol >= 1 && n_ao_pairs >= 1, "invalid sizes");
    TCI_REQUIRE((int64_t)n_chol * n_ao <= ((int64_t)1 << 31) - 1, "n_chol x n_ao exceeds the GEMM row range");
    std::unordered_set<int64_t> seen;
    for (int64_t k = 0; k < n_ao_pairs; ++k) {
      const int a = ao_pairs[2 * k], b = ao_pairs[2 * k + 1];
      TCI_REQUIRE(a >= 0 && a < n_ao && b >= 0 && b < n_ao, "AO pair index out of range");
      const int64_t key = (int64_t)std::max(a, b) * n_ao + std::min(a, b);
      TCI_REQUIRE(seen.insert(key).second, "AO pair listed twice");
    }
    // C as [mu][p] (K x N operand of the first half) and as [p][mu] (A operand of the second)
    std::vector<double> crm((size_t)n_ao * n_mo), ct((size_t)n_mo * n_ao);
    for (int p = 0; p < n_mo; ++p)
      for (int mu = 0; mu < n_ao; ++mu) {
        const double c = mo_coef[(size_t)p * n_ao + mu];
        crm[(size_t)mu * n_mo + p] = c;
        ct[(size_t)p * n_ao + mu] = c;
      }
    std::vector<int32_t> pr(ao_pairs, ao_pairs + 2 * n_ao_pairs);
    std::vector<double> av(ao_vec, ao_vec + (size_t)n_chol * n_ao_pairs);
    DevBuf dpairs, dvec, dcrm, dct, dL, dH, dMO, drec, dnorm;
    upload(dpairs, pr);
    upload(dvec, av);
    upload(dcrm, crm);
    upload(dct, ct);
    const int64_t nao2 = (int64_t)n_ao * n_ao, nmo2 = (int64_t)n_mo * n_mo;
    dL.alloc((size_t)n_chol * nao2 * 8);
    dH.alloc((size_t)n_chol * n_ao * n_mo * 8);
    dMO.alloc((size_t)n_chol * nmo2 * 8);
    launch_fill(dL.as<double>(), n_chol * nao2, 0.0, nullptr);
    TCI_REQUIRE(n_chol <= 65535, "at most 65535 Cholesky vectors per call");
    hipLaunchKernelGGL(scatter_ao_pairs_kernel,
                       dim3((unsigned)std::min<int64_t>(cdiv(n_ao_pairs, 256), 1024), (unsigned)n_chol), dim3(256), 0,
                       nullptr, dpairs.as<int32_t>(), dvec.as<double>(), n_ao_pairs, n_ao, dL.as<double>());
    TCI_HIP(hipGetLastError());
    // first half: H[(J, mu)][i] = sum_nu L[(J, mu)][nu] C[nu][i]
    GemmProgram h1;
    h1.dense = true;
    h1.tile = 128;
    h1.add(dH.as<double>(), n_mo, (int)(n_chol * n_ao), n_mo, {cseg(cop(dL.p, n_ao), cop(dcrm.p, n_mo), n_ao)}, false);
    h1.upload_tables(nullptr);
    h1.launch(nullptr);
    // second half: MO^J[j][i] = sum_mu C^T[j][mu] H^J[mu][i], one product per vector
    GemmProgram h2;
    h2.dense = true;
    h2.tile = n_mo > 96 ? 128 : 64;
    for (int64_t J = 0; J < n_chol; ++J)
      h2.add(dMO.as<double>() + J * nmo2, n_mo, n_mo, n_mo,
             {cseg(cop(dct.p, n_ao), cop(dH.as<double>() + J * n_ao * n_mo, n_mo), n_ao)}, false);
    h2.upload_tables(nullptr);
    h2.launch(nullptr);
    const int64_t n_rec = (int64_t)n_mo * (n_mo + 1) / 2;
    drec.alloc((size_t)n_rec * n_chol * 8);
    hipLaunchKernelGGL(pack_mo_records_kernel, dim3((unsigned)std::min<int64_t>(cdiv(n_rec * n_chol, 256), 65536)),
                       dim3(256), 0, nullptr, dMO.as<double>(), n_mo, n_chol, drec.as<double>());
    TCI_HIP(hipGetLastError());
    if (norms) {
      dnorm.alloc((size_t)n_rec * 8);
      hipLaunchKernelGGL(record_norms_kernel, dim3((unsigned)cdiv(n_rec, 4)), dim3(256), 0, nullptr, drec.as<double>(),
                         n_rec, n_chol, dnorm.as<double>());
      TCI_HIP(hipGetLastError());
      TCI_HIP(hipMemcpy(norms, dnorm.p, (size_t)n_rec * 8, hipMemcpyDeviceToHost));
    }
    TCI_HIP(hipMemcpy(mo_records, drec.p, (size_t)n_rec * n_chol * 8, hipMemcpyDeviceToHost));
    return TCI_OK;
  } catch (const tci::Error& ex) {
    tci::set_last_error(ex.what());
    return ex.code;
  } catch (const std::exception& ex) {
    tci::set_last_error(ex.what());
    return TCI_ERR_ARG;
  } catch (...) {
    tci::set_last_error("unknown error");
    return TCI_ERR_ARG;
  }
}
