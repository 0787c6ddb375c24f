// Internal walk table and allowed-virtual lists on the GPU.
//
// tci_internal_walks: the walks of the internal distinct row table (DRT) that
// end with n_elec - n_ext electrons, in the order of their lexical index
// (index = 1 + sum of the arc weights along the walk, F/locist_mod.f90:394),
// with the keep flag of the reference's tree_search_internal.  The graph is
// built on the host (graph_border F/graph_mod.f90:101-262: cumulative
// occupation borders of every level, narrowed to the SD band of the
// references; drt_maker :263-380: vertex weights y0 and arc weights y1, y2).
// Every walk is then produced on the GPU by unranking its index from the
// bottom vertex: at vertex (i, J) the walks arriving by a 0-, 1- or 2-step
// from (i-1, J-d) hold the consecutive index ranges [w_d, w_d + |y0(i-1, J-d)|),
// w_0 = 0, w_1 = y1(i-1, J-1), w_2 = y2(i-1, J-2).  A walk is kept when it is at
// most doubly excited from some reference (check_excitation, :963-1002) and
// keeps its frozen levels doubly occupied (check_frozen, :1004-1026).
//
// tci_allowed_virtuals: per walk, the hole pairs against every reference it is
// a double excitation of (find_hole_pairs, F/locist_mod.f90:533-607), and the
// allowed virtuals: every virtual for a reference or single excitation, else
// the union over the hole pairs that interact under WP of the virtuals that
// interact under TOV (find_allowed_virtuals, :453-518).  One wavefront per
// walk: the lanes compute the hole pairs, then each lane ORs 32-virtual words
// of the TOV rows.  The host numbers the distinct lists in order of first
// appearance (add_new_virtual_list, F/allowed_virtuals_mod.f90:163-219).
#include <unordered_map>

#include "common.hpp"

namespace tci {

namespace {

// ---------------------------------------------------------------------------
// DRT tables (host)
struct Drt {
  int n = 0;                        // levels 1..n (orbitals)
  int E = 0;                        // electron dimension of the tables
  std::vector<int> left, right;     // borders per level
  std::vector<int64_t> y0, y1, y2;  // [(n + 1) * E]; absent entries: y0 = 0, y1 = y2 = -1
  bool belong(int i, int j) const { return i >= 0 && j >= 0 && i <= n && j < E && left[i] <= j && j <= right[i]; }
  int64_t& Y0(int i, int j) { return y0[(size_t)i * E + j]; }
  int64_t& Y1(int i, int j) { return y1[(size_t)i * E + j]; }
  int64_t& Y2(int i, int j) { return y2[(size_t)i * E + j]; }
  int64_t g0(int i, int j) const { return belong(i, j) ? y0[(size_t)i * E + j] : 0; }
  int64_t g1(int i, int j) const { return belong(i, j) ? y1[(size_t)i * E + j] : -1; }
  int64_t g2(int i, int j) const { return belong(i, j) ? y2[(size_t)i * E + j] : -1; }
  bool add0(int i, int j) const { return belong(i, j) && g0(i, j) > 0; }
  bool add1(int i, int j) const { return belong(i, j) && g1(i, j) >= 0; }
  bool add2(int i, int j) const { return belong(i, j) && g2(i, j) >= 0; }
};

// graph_border + drt_maker over n_orb levels; refs are occupations over all
// orbitals (n_ref x n_orb)
Drt build_drt(int n_orb, int n_elec, const std::vector<std::vector<int>>& refs, int spin_mult, int exilevel) {
  Drt g;
  g.n = n_orb;
  g.E = 2 * n_orb + 3;
  const int min_singles = spin_mult - 1;
  const int right_singles = min_singles, right_doubly = (n_elec - right_singles) / 2;
  const int left_singles = std::max(0, min_singles), left_doubly = (n_elec - left_singles) / 2;
  const int left_unocc = n_orb - left_singles - left_doubly;
  g.left.assign(n_orb + 1, 0);
  g.right.assign(n_orb + 1, 0);
  for (int i = 1; i <= n_orb; ++i) {
    int add = i <= left_unocc ? 0 : (i <= left_unocc + left_singles ? 1 : 2);
    g.left[i] = i == 1 ? add : g.left[i - 1] + add;
    add = i <= right_doubly ? 2 : (i <= right_doubly + right_singles ? 1 : 0);
    g.right[i] = i == 1 ? add : g.right[i - 1] + add;
  }
  std::vector<int> cum(refs.size());
  for (size_t r = 0; r < refs.size(); ++r) cum[r] = refs[r][0];
  for (int i = 2; i <= n_orb; ++i) {
    int lo = 1 << 30, hi = -(1 << 30);
    for (size_t r = 0; r < refs.size(); ++r) {
      cum[r] += refs[r][i - 1];
      lo = std::min(lo, cum[r]);
      hi = std::max(hi, cum[r]);
    }
    g.left[i] = std::max(lo - exilevel, g.left[i]);
    g.right[i] = std::min(hi + exilevel, g.right[i]);
  }
  for (int i = 0; i <= n_orb; ++i) {
    g.left[i] = std::max(g.left[i], 0);
    g.right[i] = std::min(g.right[i], g.E - 1);
  }
  const size_t sz = (size_t)(n_orb + 1) * g.E;
  g.y0.assign(sz, 0);
  g.y1.assign(sz, -1);
  g.y2.assign(sz, -1);
  g.Y0(0, 0) = g.belong(1, 0) ? 1 : -1;
  g.Y1(0, 0) = g.belong(1, 1) ? 0 : -1;
  g.Y2(0, 0) = g.belong(1, 2) ? 0 : -1;
  for (int i = 1; i <= n_orb; ++i)
    for (int j = g.right[i]; j >= g.left[i]; --j) {
      int64_t w = 0;
      if (g.add0(i - 1, j)) w += std::llabs(g.g0(i - 1, j));
      if (g.add1(i - 1, j - 1)) w += std::llabs(g.g0(i - 1, j - 1));
      if (g.add2(i - 1, j - 2)) w += std::llabs(g.g0(i - 1, j - 2));
      g.Y0(i, j) = g.belong(i + 1, j) ? w : -w;
      g.Y1(i, j) = !g.belong(i + 1, j + 1) ? -1 : (g.add0(i, j + 1) ? std::llabs(g.g0(i, j + 1)) : 0);
      int64_t w2 = -1;
      if (g.belong(i + 1, j + 2)) {
        w2 = 0;
        if (g.add1(i, j + 1)) w2 += std::llabs(g.g0(i, j + 1));
        if (g.add0(i, j + 2)) w2 += std::llabs(g.g0(i, j + 2));
      }
      g.Y2(i, j) = w2;
    }
  return g;
}

// ---------------------------------------------------------------------------
// unranking (one thread per walk)
struct DrtView {
  const int64_t *y0, *y1, *y2;
  const int32_t *left, *right;
  int32_t E, n;
};

__device__ __forceinline__ bool d_belong(const DrtView& g, int i, int j) {
  return i >= 0 && j >= 0 && i <= g.n && j < g.E && g.left[i] <= j && j <= g.right[i];
}

// walk k (0-based lexical index - 1) ending at (n_int, e); refs: n_ref x n_int
__global__ void unrank_walks_kernel(DrtView g, int n_int, int e, int64_t n_walks, const int8_t* __restrict__ refs,
                                    int n_ref, int n_elec, int n_frozen, int8_t* __restrict__ occ,
                                    uint8_t* __restrict__ kept, int* __restrict__ bad) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= n_walks) return;
  int8_t* o = occ + w * n_int;
  int64_t k = w;
  int J = e;
  bool ok = true;
  for (int i = n_int; i >= 1 && ok; --i) {
    bool found = false;
    for (int d = 0; d <= 2 && !found; ++d) {
      const int p = J - d;
      if (p < 0 || !d_belong(g, i - 1, p)) continue;
      const size_t at = (size_t)(i - 1) * g.E + p;
      const int64_t y0 = g.y0[at];
      const bool valid = d == 0 ? y0 > 0 : (d == 1 ? g.y1[at] >= 0 : g.y2[at] >= 0);
      if (!valid) continue;
      const int64_t lo = d == 0 ? 0 : (d == 1 ? g.y1[at] : g.y2[at]);
      const int64_t cnt = y0 < 0 ? -y0 : y0;
      if (k >= lo && k < lo + cnt) {
        o[i - 1] = (int8_t)d;
        k -= lo;
        J = p;
        found = true;
      }
    }
    ok = found;
  }
  if (!ok || J != 0 || k != 0) {
    atomicAdd(bad, 1);
    kept[w] = 0;
    return;
  }
  // check_excitation and check_frozen
  int tot = 0;
  for (int i = 0; i < n_int; ++i) tot += o[i];
  bool exc = false;
  for (int r = 0; r < n_ref && !exc; ++r) {
    int dsum = 0;
    for (int i = 0; i < n_int; ++i) {
      const int x = (int)o[i] - (int)refs[(size_t)r * n_int + i];
      dsum += x < 0 ? -x : x;
    }
    exc = dsum + (n_elec - tot) <= 4;
  }
  bool frz = true;
  for (int i = 0; i < n_frozen; ++i) frz = frz && o[i] == 2;
  kept[w] = (exc && frz) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// allowed virtuals (one wavefront per walk).  status: 0 reference, 1 single
// excitation (every virtual), 2 double excitation, 3 double excitation whose
// hole pairs all fail WP (empty list), 4 neither (an error for the caller)
constexpr int kMaxRef = 16;

__global__ void allowed_virtuals_kernel(int n_int, int nw, int n_v, int n_ref, const int8_t* __restrict__ refs,
                                        int64_t n_walks, const int8_t* __restrict__ occ,
                                        const uint8_t* __restrict__ wp, const uint32_t* __restrict__ tov,
                                        int32_t* __restrict__ status, uint32_t* __restrict__ mask) {
  const int lane = threadIdx.x & 63;
  const int64_t w = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= n_walks) return;
  const int8_t* o = occ + w * n_int;
  // hole pairs, the same in every lane
  int pi[kMaxRef], pj[kMaxRef];
  int np = 0, st = 2;
  bool is_ref = false;
  for (int r = 0; r < n_ref && !is_ref; ++r) {
    bool same = true;
    for (int i = 0; i < n_int; ++i) same = same && o[i] == refs[(size_t)r * n_int + i];
    is_ref = same;
  }
  if (is_ref) st = 0;
  for (int r = 0; r < n_ref && st == 2; ++r) {
    int holes = 0, parts = 0, first = -1, second = -1;
    for (int i = 0; i < n_int; ++i) {
      const int x = (int)refs[(size_t)r * n_int + i] - (int)o[i];
      if (x > 0) holes += x;
      else parts -= x;
      if (x == 2) first = second = i;
      if (x == 1) {
        if (first < 0) first = i;
        else second = i;
      }
    }
    if (holes > 2 || parts > 2) continue;
    if (holes == 1) {
      st = 1;
      break;
    }
    pi[np] = first;
    pj[np] = second;
    ++np;
  }
  if (st == 2 && np == 0) st = 4;
  bool any_wp = false;
  if (st == 2) {
    for (int q = 0; q < np; ++q) any_wp = any_wp || wp[(size_t)pi[q] * n_int + pj[q]] != 0;
    if (!any_wp) st = 3;
  }
  if (lane == 0) status[w] = st;
  for (int q = lane; q < nw; q += 64) {
    uint32_t m = 0;
    if (st == 0 || st == 1) {
      const int rem = n_v - 32 * q;
      m = rem >= 32 ? 0xffffffffu : ((1u << rem) - 1u);
    } else if (st == 2) {
      for (int t = 0; t < np; ++t)
        if (wp[(size_t)pi[t] * n_int + pj[t]]) m |= tov[((size_t)pi[t] * n_int + pj[t]) * nw + q];
    }
    mask[w * nw + q] = m;
  }
}

struct VecHash {
  size_t operator()(const std::vector<uint32_t>& v) const {
    uint64_t h = 1469598103934665603ull;
    for (uint32_t x : v) h = (h ^ x) * 1099511628211ull;
    return (size_t)h;
  }
};

}  // namespace

}  // namespace tci

using namespace tci;

extern "C" int tci_internal_walks(int32_t n_int, int32_t n_v, int32_t n_elec, int32_t spin_mult, int32_t n_ref,
                                  const int8_t* references, int32_t n_frozen, int32_t n_ext, int64_t capacity,
                                  int8_t* occ, uint8_t* kept, int64_t* n_walks) {
  try {
    TCI_REQUIRE(references && n_walks && (capacity == 0 || (occ && kept)), "null argument");
    TCI_REQUIRE(n_int >= 1 && n_v >= 0 && n_ref >= 1 && n_ext >= 0 && n_ext <= 2 && spin_mult >= 1,
                "invalid walk-table sizes");
    TCI_REQUIRE(n_frozen >= 0 && n_frozen <= n_int, "n_frozen must be in 0..n_int");
    TCI_REQUIRE(n_elec >= n_ext && n_elec - n_ext <= 2 * n_int, "electron count outside the internal space");
    std::vector<std::vector<int>> refs(n_ref, std::vector<int>(n_int + n_v, 0));
    for (int r = 0; r < n_ref; ++r)
      for (int i = 0; i < n_int; ++i) {
        const int x = references[(size_t)r * n_int + i];
        TCI_REQUIRE(x >= 0 && x <= 2, "reference occupations must be 0, 1 or 2");
        refs[r][i] = x;
      }
    const Drt g = build_drt(n_int + n_v, n_elec, refs, spin_mult, 2);
    const int e = n_elec - n_ext;
    const int64_t nw = g.belong(n_int, e) ? std::llabs(g.g0(n_int, e)) : 0;
    *n_walks = nw;
    if (nw == 0 || capacity == 0) return TCI_OK;  // capacity 0: size query
    TCI_REQUIRE(capacity >= nw, "walk capacity " + std::to_string(capacity) + " < " + std::to_string(nw) + " walks");
    // only the internal levels 0..n_int are read on the device
    const int L = n_int + 1;
    std::vector<int64_t> y0(g.y0.begin(), g.y0.begin() + (size_t)L * g.E), y1(g.y1.begin(), g.y1.begin() + (size_t)L * g.E),
        y2(g.y2.begin(), g.y2.begin() + (size_t)L * g.E);
    std::vector<int32_t> lf(g.left.begin(), g.left.begin() + L), rt(g.right.begin(), g.right.begin() + L);
    std::vector<int8_t> rr((size_t)n_ref * n_int);
    for (int r = 0; r < n_ref; ++r)
      for (int i = 0; i < n_int; ++i) rr[(size_t)r * n_int + i] = (int8_t)refs[r][i];
    DevBuf d0, d1, d2, dl, dr, dref, docc, dkept, dbad;
    upload(d0, y0);
    upload(d1, y1);
    upload(d2, y2);
    upload(dl, lf);
    upload(dr, rt);
    upload(dref, rr);
    docc.alloc((size_t)nw * n_int);
    dkept.alloc((size_t)nw);
    std::vector<int> zero(1, 0);
    upload(dbad, zero);
    DrtView v{d0.as<int64_t>(), d1.as<int64_t>(), d2.as<int64_t>(), dl.as<int32_t>(), dr.as<int32_t>(), g.E, n_int};
    const unsigned blocks = (unsigned)cdiv(nw, 256);
    TCI_REQUIRE(cdiv(nw, 256) < ((int64_t)1 << 31), "too many walks");
    hipLaunchKernelGGL(unrank_walks_kernel, dim3(blocks), dim3(256), 0, nullptr, v, n_int, e, nw,
                       dref.as<int8_t>(), n_ref, n_elec, n_frozen, docc.as<int8_t>(), dkept.as<uint8_t>(),
                       dbad.as<int>());
    TCI_HIP(hipGetLastError());
    int bad = 0;
    TCI_HIP(hipMemcpy(&bad, dbad.p, sizeof(int), hipMemcpyDeviceToHost));
    TCI_HIP(hipMemcpy(occ, docc.p, (size_t)nw * n_int, hipMemcpyDeviceToHost));
    TCI_HIP(hipMemcpy(kept, dkept.p, (size_t)nw, hipMemcpyDeviceToHost));
    TCI_REQUIRE(bad == 0, std::to_string(bad) + " walk indices did not unrank (inconsistent graph)");
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

extern "C" int tci_allowed_virtuals(int32_t n_int, int32_t n_v, int32_t n_ref, const int8_t* references,
                                    int64_t n_walks, const int8_t* walk_occ, const uint8_t* wp, const uint8_t* tov,
                                    int32_t* status, int64_t* list_of_walk, int64_t* list_off, int32_t* list_virt,
                                    int64_t list_capacity, int64_t* n_lists) {
  try {
    TCI_REQUIRE(references && walk_occ && wp && tov && status && list_of_walk && list_off && n_lists &&
                    (list_capacity == 0 || list_virt),
                "null argument");
    TCI_REQUIRE(n_int >= 1 && n_v >= 1 && n_ref >= 1 && n_walks >= 1, "invalid sizes");
    TCI_REQUIRE(n_ref <= kMaxRef, "at most 16 references");
    const int nw = (int)cdiv(n_v, 32);
    // TOV rows as 32-virtual bit words per internal pair
    std::vector<uint32_t> bits((size_t)n_int * n_int * nw, 0u);
    for (int64_t p = 0; p < (int64_t)n_int * n_int; ++p)
      for (int a = 0; a < n_v; ++a)
        if (tov[p * n_v + a]) bits[p * nw + a / 32] |= 1u << (a % 32);
    std::vector<int8_t> rr(references, references + (size_t)n_ref * n_int);
    std::vector<int8_t> ww(walk_occ, walk_occ + (size_t)n_walks * n_int);
    std::vector<uint8_t> wpv(wp, wp + (size_t)n_int * n_int);
    DevBuf dref, docc, dwp, dtov, dst, dmask;
    upload(dref, rr);
    upload(docc, ww);
    upload(dwp, wpv);
    upload(dtov, bits);
    dst.alloc((size_t)n_walks * 4);
    dmask.alloc((size_t)n_walks * nw * 4);
    const int64_t blocks = cdiv(n_walks, 4);  // four wavefronts (walks) per workgroup
    TCI_REQUIRE(blocks < ((int64_t)1 << 31), "too many walks");
    hipLaunchKernelGGL(allowed_virtuals_kernel, dim3((unsigned)blocks), dim3(256), 0, nullptr, n_int, nw, n_v, n_ref,
                       dref.as<int8_t>(), n_walks, docc.as<int8_t>(), dwp.as<uint8_t>(), dtov.as<uint32_t>(),
                       dst.as<int32_t>(), dmask.as<uint32_t>());
    TCI_HIP(hipGetLastError());
    std::vector<uint32_t> mask((size_t)n_walks * nw);
    TCI_HIP(hipMemcpy(status, dst.p, (size_t)n_walks * 4, hipMemcpyDeviceToHost));
    TCI_HIP(hipMemcpy(mask.data(), dmask.p, mask.size() * 4, hipMemcpyDeviceToHost));
    // distinct lists in order of first appearance
    std::unordered_map<std::vector<uint32_t>, int64_t, VecHash> seen;
    int64_t nl = 0, used = 0;
    list_off[0] = 0;
    for (int64_t w = 0; w < n_walks; ++w) {
      TCI_REQUIRE(status[w] != 4, "walk " + std::to_string(w) +
                                      " is not a reference, single or double excitation of the references");
      if (status[w] == 0) {
        list_of_walk[w] = -1;
        continue;
      }
      std::vector<uint32_t> key(mask.begin() + w * nw, mask.begin() + (w + 1) * nw);
      auto it = seen.find(key);
      if (it != seen.end()) {
        list_of_walk[w] = it->second;
        continue;
      }
      int64_t cnt = 0;
      for (uint32_t x : key) cnt += __builtin_popcount(x);
      TCI_REQUIRE(used + cnt <= list_capacity,
                  "list capacity " + std::to_string(list_capacity) + " exceeded at list " + std::to_string(nl));
      for (int q = 0; q < nw; ++q)
        for (int b = 0; b < 32; ++b)
          if (key[q] >> b & 1u) list_virt[used++] = 32 * q + b;
      seen.emplace(std::move(key), nl);
      list_of_walk[w] = nl;
      list_off[++nl] = used;
    }
    *n_lists = nl;
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
