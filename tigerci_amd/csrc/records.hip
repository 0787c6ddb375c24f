// Ladder loop records of the iteration-1 dry run on the GPU.
//
// The reference's first sigma walks every virtual pair a >= c (a outer) and,
// for each, every kept N-2 walk, and writes a record for each walk whose
// allowed-virtual list holds both (all_external_3_4_loops_cho_2 with
// build_ext_34seg_dryrun, F/three_and_four_seg_mod_1.f90:318-469,3142-3238).
// Here one thread owns one pair: it counts the walks that hold both
// virtuals, a prefix sum over the pairs (in the reference's pair order) gives
// each pair its first record, and the same thread writes its records in walk
// order.  Membership is a table of 1-based list positions per walk and
// virtual, so a test is one load.
#include "common.hpp"

namespace tci {

namespace {

// pos[w][a] = 1-based position of virtual a in walk w's list, 0 if absent
__global__ void record_positions_kernel(int64_t n_walks, int n_v, const int64_t* __restrict__ list_off,
                                        const int32_t* __restrict__ list_virt, int32_t* __restrict__ pos) {
  const int64_t w = blockIdx.x;
  if (w >= n_walks) return;
  const int64_t q0 = list_off[w], q1 = list_off[w + 1];
  for (int64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) pos[w * n_v + list_virt[q]] = (int32_t)(q - q0 + 1);
}

__device__ inline void pair_of(int64_t p, int& a, int& c) {
  int64_t x = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while ((x + 1) * (x + 2) / 2 <= p) ++x;
  while (x * (x + 1) / 2 > p) --x;
  a = (int)x;
  c = (int)(p - x * (x + 1) / 2);
}

// pass 0: count[p] = records of pair p; pass 1: write them from first[p]
template <bool WRITE>
__global__ void ladder_records_kernel(int n_int, int n_v, int64_t n_pairs, int64_t n_walks,
                                      const int32_t* __restrict__ pos, const uint8_t* __restrict__ ignorable,
                                      const int64_t* __restrict__ head_weight, const int32_t* __restrict__ spins,
                                      const int64_t* __restrict__ first, int64_t* __restrict__ count,
                                      int32_t* __restrict__ rec) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  int a, c;
  pair_of(p, a, c);  // p = a(a+1)/2 + c, c <= a
  if (ignorable && ignorable[(int64_t)c * n_v + a]) {
    if (!WRITE) count[p] = 0;
    return;
  }
  const int64_t ag = n_int + 1 + a, cg = n_int + 1 + c;
  int64_t k = WRITE ? first[p] : 0;
  for (int64_t w = 0; w < n_walks; ++w) {
    const int32_t pa = pos[w * n_v + a], pc = pos[w * n_v + c];
    if (pa == 0 || pc == 0) continue;
    if (WRITE) {
      int32_t* r = rec + 7 * k;
      r[0] = (int32_t)(ag * (ag - 1) / 2 + cg);
      r[1] = (int32_t)head_weight[w];
      r[2] = spins[3 * w];
      r[3] = spins[3 * w + 1];
      r[4] = spins[3 * w + 2];
      r[5] = pa;
      r[6] = pc;
    }
    ++k;
  }
  if (!WRITE) count[p] = k;
}

int64_t spin_count(int n, int mult) {
  if (n < 0 || (n + mult - 1) % 2 || mult - 1 > n) return 0;
  const int s2 = mult - 1;
  double v = mult;
  for (int k = 2; k <= n; ++k) v *= k;
  for (int k = 2; k <= (n + s2) / 2 + 1; ++k) v /= k;
  for (int k = 2; k <= (n - s2) / 2; ++k) v /= k;
  return std::llround(v);
}

}  // namespace

}  // namespace tci

using namespace tci;

extern "C" int tci_ladder_records(int32_t n_int, int32_t n_v, int32_t spin_mult, int64_t n_walks,
                                  const int64_t* head_weight, const int32_t* num_singles, const int64_t* list_off,
                                  const int32_t* list_virt, const uint8_t* ignorable, int64_t capacity,
                                  int32_t* records, int64_t* n_records) {
  try {
    TCI_REQUIRE(head_weight && num_singles && list_off && list_virt && n_records && (capacity == 0 || records),
                "null argument");
    TCI_REQUIRE(n_int >= 1 && n_v >= 1 && n_walks >= 1 && spin_mult >= 1, "invalid sizes");
    const int64_t orb = (int64_t)n_int + n_v;
    TCI_REQUIRE(orb * (orb + 1) / 2 < ((int64_t)1 << 31), "pair index beyond 32 bits");
    TCI_REQUIRE(list_off[0] == 0, "list_off must start at 0");
    std::vector<int32_t> spins((size_t)n_walks * 3);
    for (int64_t w = 0; w < n_walks; ++w) {
      TCI_REQUIRE(list_off[w + 1] >= list_off[w] && list_off[w + 1] - list_off[w] <= n_v, "bad list offsets");
      TCI_REQUIRE(head_weight[w] >= 0 && head_weight[w] < ((int64_t)1 << 31), "head weight beyond 32 bits");
      const int s = num_singles[w];
      TCI_REQUIRE(s >= 0, "negative open-shell count");
      spins[(size_t)3 * w] = (int32_t)spin_count(s, spin_mult);
      spins[(size_t)3 * w + 1] = s >= 2 ? (int32_t)spin_count(s - 2, spin_mult) : 0;
      spins[(size_t)3 * w + 2] = (int32_t)spin_count(s + 2, spin_mult);
      for (int64_t q = list_off[w]; q < list_off[w + 1]; ++q) {
        TCI_REQUIRE(list_virt[q] >= 0 && list_virt[q] < n_v, "virtual index out of range");
        TCI_REQUIRE(q == list_off[w] || list_virt[q] > list_virt[q - 1], "lists must be sorted and unique");
      }
    }
    const int64_t nl = list_off[n_walks];
    std::vector<int64_t> hw(head_weight, head_weight + n_walks), off(list_off, list_off + n_walks + 1);
    std::vector<int32_t> lv(list_virt, list_virt + nl);
    DevBuf dhw, doff, dlv, dsp, dpos, dign, dcnt;
    upload(dhw, hw);
    upload(doff, off);
    upload(dlv, lv);
    upload(dsp, spins);
    if (ignorable) {
      std::vector<uint8_t> ig(ignorable, ignorable + (size_t)n_v * n_v);
      upload(dign, ig);
    }
    dpos.alloc((size_t)n_walks * n_v * 4);
    TCI_HIP(hipMemset(dpos.p, 0, (size_t)n_walks * n_v * 4));
    TCI_REQUIRE(n_walks < ((int64_t)1 << 31), "too many walks");
    hipLaunchKernelGGL(record_positions_kernel, dim3((unsigned)n_walks), dim3(64), 0, nullptr, n_walks, n_v,
                       doff.as<int64_t>(), dlv.as<int32_t>(), dpos.as<int32_t>());
    TCI_HIP(hipGetLastError());
    const int64_t n_pairs = (int64_t)n_v * (n_v + 1) / 2;
    dcnt.alloc((size_t)n_pairs * 8);
    const unsigned g = (unsigned)cdiv(n_pairs, 256);
    const uint8_t* ig = ignorable ? dign.as<uint8_t>() : nullptr;
    hipLaunchKernelGGL(ladder_records_kernel<false>, dim3(g), dim3(256), 0, nullptr, (int)n_int, (int)n_v, n_pairs,
                       n_walks, dpos.as<int32_t>(), ig, dhw.as<int64_t>(), dsp.as<int32_t>(), nullptr,
                       dcnt.as<int64_t>(), nullptr);
    TCI_HIP(hipGetLastError());
    std::vector<int64_t> cnt((size_t)n_pairs);
    TCI_HIP(hipMemcpy(cnt.data(), dcnt.p, (size_t)n_pairs * 8, hipMemcpyDeviceToHost));
    std::vector<int64_t> first((size_t)n_pairs);
    int64_t tot = 0;
    for (int64_t p = 0; p < n_pairs; ++p) {
      first[(size_t)p] = tot;
      tot += cnt[(size_t)p];
    }
    *n_records = tot;
    if (tot == 0 || capacity == 0) return TCI_OK;  // capacity 0: size query
    TCI_REQUIRE(capacity >= tot, "record capacity " + std::to_string(capacity) + " < " + std::to_string(tot));
    DevBuf dfirst, drec;
    upload(dfirst, first);
    drec.alloc((size_t)tot * 7 * 4);
    hipLaunchKernelGGL(ladder_records_kernel<true>, dim3(g), dim3(256), 0, nullptr, (int)n_int, (int)n_v, n_pairs,
                       n_walks, dpos.as<int32_t>(), ig, dhw.as<int64_t>(), dsp.as<int32_t>(), dfirst.as<int64_t>(),
                       nullptr, drec.as<int32_t>());
    TCI_HIP(hipGetLastError());
    TCI_HIP(hipMemcpy(records, drec.p, (size_t)tot * 7 * 4, hipMemcpyDeviceToHost));
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
