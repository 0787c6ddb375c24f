// Spin functions and symmetric-group representation matrices of the SGGA
// (symmetric group graphical approach) on the GPU, and the contracted cycles
// that make the loop coupling coefficients.
//
// Spin functions of n open shells and multiplicity 2S+1 are the paths of the
// branching diagram from vertex (a, b) = (1, 1) to (up, down + 1): a step
// "up" adds one to a, a step "down" one to b, and a down step is allowed only
// while b < down + 1 and b < a + 1.  The reference enumerates them depth
// first, down before up, the shell index running from n - 1 to 1
// (branching_diagram, sft_graph, F/spin_mod.f90:99-521), so the rank of a
// path is the sum, over its up steps taken where a down step was allowed, of
// the number of completions of the down step.  Here every spin function is
// unranked from its index by one thread and writes its column of the
// elementary-transposition table sft and of the intermediate spins; then
// every (k, spin function) column of the cycle matrices (l..k), l < k, runs
// the reference's recurrence (make_cycles, :523-698), the transpositions are
// products of two cycles (make_transpositions, :700-860), and each contracted
// cycle (contracted_cycle_stateless, :861-1312) is one workgroup whose
// threads own its output columns.
#include "common.hpp"
#include "terms.hpp"

namespace tci {

namespace {

constexpr int kMaxOpen = 12;    // fsn(12, 1) = 132 functions; tables of 66 x 132 x 132
constexpr int kMaxFsn = 924;    // fsn(12, 13): bound on any multiplicity's count
constexpr int kMaxCcFsn = 132;  // contracted cycles: singlet tables of up to 12 open shells

int fsn_of(int n, int mult) {
  if (n < 0 || (n + mult - 1) % 2 || mult - 1 > n) return 0;
  const int s2 = mult - 1;
  // (2S+1) n! / ((n/2 + S + 1)! (n/2 - S)!)
  double v = mult;
  for (int k = 2; k <= n; ++k) v *= k;
  for (int k = 2; k <= (n + s2) / 2 + 1; ++k) v /= k;
  for (int k = 2; k <= (n - s2) / 2; ++k) v /= k;
  return (int)std::llround(v);
}

struct Branch {
  int up = 0, down = 0, ld = 0;  // br[a * ld + b], a = 0..up + 1, b = 0..down + 2
  std::vector<int32_t> br;
  int at(int a, int b) const { return br[(size_t)a * ld + b]; }
};

// completions of every vertex (a, b) of the branching diagram, a >= 1 and
// 1 <= b <= a + 1 (zero elsewhere, as in the reference's table)
Branch branching_diagram(int n, int mult) {
  Branch B;
  B.up = (n + mult - 1) / 2;
  B.down = n - B.up;
  B.ld = B.down + 3;
  B.br.assign((size_t)(B.up + 2) * B.ld, 0);
  for (int a = B.up; a >= 1; --a)
    for (int b = std::min(B.down + 1, a + 1); b >= 1; --b) {
      int64_t c = 0;
      if (a == B.up && b == B.down + 1) {
        c = 1;
      } else {
        const bool forced_up = b == B.down + 1 || b == a + 1;
        if (a < B.up) c += B.br[(size_t)(a + 1) * B.ld + b];
        if (!forced_up && b < B.down + 1) c += B.br[(size_t)a * B.ld + b + 1];
      }
      B.br[(size_t)a * B.ld + b] = (int32_t)c;
    }
  return B;
}

// Thread w (0-based) unranks spin function w + 1: its path, intermediate
// spins spin[v][w] (v = 0..n) and elementary transpositions sft[i-1][w] for
// i = 1..n-1 (the reference's signed weights).
__global__ void spin_functions_kernel(int n, int mult, int up, int down, const int32_t* __restrict__ br, int ld,
                                      int nf, int32_t* __restrict__ sft, int32_t* __restrict__ spin) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nf) return;
  bool path[kMaxOpen + 1];
  int pa[kMaxOpen + 1], pb[kMaxOpen + 1];
  path[n] = true;
  int r = w, a = 1, b = 1;
  for (int s = n - 1; s >= 1; --s) {
    pa[s] = a;
    pb[s] = b;
    const bool forced_up = b == down + 1 || b == a + 1;
    bool u = true;
    if (!forced_up) {
      const int f = br[a * ld + b + 1];
      if (r < f) u = false;
      else r -= f;
    }
    path[s] = u;
    if (u) ++a;
    else ++b;
  }
  const int weight = w + 1;
  int sp = 1;
  spin[(int64_t)n * nf + w] = 1;
  spin[w] = mult;
  for (int i = n - 1; i >= 1; --i) {
    const bool prev = path[i + 1], cur = path[i];
    const int ai = pa[i], bi = pb[i];
    sp = prev ? sp + 1 : sp - 1;
    spin[(int64_t)i * nf + w] = sp;
    int v = 0;
    if (prev && !cur && ai == bi) {
      v = weight;
    } else if (prev && !cur) {
      v = weight - br[(ai - 1) * ld + bi + 1];
      if (bi + 2 <= down + 1) v += br[(ai - 1) * ld + bi + 2];
    } else if (!prev && cur) {
      int t = weight + br[ai * ld + bi];
      if (bi <= ai && bi < down + 1) t -= br[ai * ld + bi + 1];
      v = -t;
    }
    sft[(int64_t)(i - 1) * nf + w] = v;
  }
}

__host__ __device__ inline int cyc_label(int l, int k) { return (k - 1) * (k - 2) / 2 + l - 1; }  // 0-based, l < k

// Column i (0-based) of every cycle (l..k), l = k-1 .. 1, for one k: the
// reference's recurrence over elementary transpositions, with its 1e-9 floor.
// The sparse column (spin functions and values, at most nf + 1 entries) lives
// in a workspace slice of the thread's own.
__global__ void cycles_kernel(int n, int nf, const int32_t* __restrict__ sft, const int32_t* __restrict__ spin,
                              double* __restrict__ cyc, int32_t* __restrict__ ws_reorder,
                              double* __restrict__ ws_column) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y + 2;
  if (i >= nf) return;
  const int64_t slot = ((int64_t)blockIdx.y * nf + i) * (nf + 1);
  int32_t* reorder = ws_reorder + slot;
  double* column = ws_column + slot;
  reorder[0] = i + 1;
  column[0] = 1.0;
  int ns = 1;
  for (int l = k - 1; l >= 1; --l) {
    const int ns0 = ns;
    for (int j = 0; j < ns0; ++j) {
      const int f = reorder[j];  // 1-based spin function
      const double akl = 1.0 / (double)spin[(int64_t)(l + 1) * nf + f - 1];
      const double net = column[j];
      const int sv = sft[(int64_t)(l - 1) * nf + f - 1];
      if (sv == 0) {
        column[j] = -net;
      } else if (sv > 0) {
        column[j] = akl * net;
        if (sv != f && ns <= nf) {
          reorder[ns] = sv;
          column[ns] = -sqrt(1.0 - akl * akl) * net;
          ++ns;
        }
      } else if (ns <= nf) {
        column[j] = -akl * net;
        reorder[ns] = -sv;
        column[ns] = -sqrt(1.0 - akl * akl) * net;
        ++ns;
      }
    }
    double* C = cyc + (int64_t)cyc_label(l, k) * nf * nf;
    for (int jj = 0; jj < ns; ++jj)
      if (reorder[jj] >= 1 && reorder[jj] <= nf)
        C[(int64_t)(reorder[jj] - 1) * nf + i] = fabs(column[jj]) >= 1e-9 ? column[jj] : 0.0;
  }
}

// transposition (p, q), p > q: (q+1 .. p)^T (q .. p), or the cycle (q..p) when p = q + 1
__global__ void transpositions_kernel(int n, int nf, const double* __restrict__ cyc, double* __restrict__ tr) {
  const int lab = blockIdx.y, row = blockIdx.x;
  int p = 2;
  while ((p - 1) * p / 2 <= lab) ++p;  // lab in [(p-1)(p-2)/2, (p-1)p/2)
  const int q = lab - (p - 1) * (p - 2) / 2 + 1;
  const int64_t nn = (int64_t)nf * nf;
  double* T = tr + (int64_t)lab * nn;
  for (int c = threadIdx.x; c < nf; c += blockDim.x) {
    if (p == q + 1) {
      T[(int64_t)row * nf + c] = cyc[(int64_t)cyc_label(q, p) * nn + (int64_t)row * nf + c];
    } else {
      const double* A = cyc + (int64_t)cyc_label(q + 1, p) * nn;  // used transposed
      const double* B = cyc + (int64_t)cyc_label(q, p) * nn;
      double s = 0.0;
      for (int t = 0; t < nf; ++t) s += A[(int64_t)t * nf + row] * B[(int64_t)t * nf + c];
      T[(int64_t)row * nf + c] = s;
    }
  }
}

struct CcRec {
  int i, j, k, l, d1, d2, left, right;
};

// element (row, col) (1-based) of the first matrix of the product: the cycle
// (j..i) for j < i, its transpose for j > i (columns beyond tcols zero), the
// identity for i = j
__device__ inline double con_at(const double* __restrict__ cyc, int nf, int i, int j, int row, int col, int tcols) {
  if (col < 1 || col > nf) return 0.0;
  if (i == j) return row == col ? 1.0 : 0.0;
  const double* C = cyc + (int64_t)cyc_label(min(i, j), max(i, j)) * nf * nf;
  if (j < i) return C[(int64_t)(row - 1) * nf + col - 1];
  return col <= tcols ? C[(int64_t)(col - 1) * nf + row - 1] : 0.0;
}

// One workgroup per record; thread c builds column c of the compressed
// product (the columns of spin functions whose shells d1, d2 pair up) and
// multiplies it by the cycle (l..k) (transposed for l > k, identity for l = k).
__global__ void contracted_cycles_kernel(int nf, const int32_t* __restrict__ sft, const int32_t* __restrict__ spin,
                                         const double* __restrict__ cyc, const int32_t* __restrict__ fsn_tab,
                                         const CcRec* __restrict__ recs, const int64_t* __restrict__ off,
                                         double* __restrict__ out) {
  const CcRec R = recs[blockIdx.x];
  const int d1a = max(R.d1, R.d2), d2b = min(R.d1, R.d2);
  const int nsfl = fsn_tab[R.left], nsfr = fsn_tab[R.right];
  int temp = R.right;
  if (d1a > 0 && d1a < R.left - 1) temp += 2;
  if (d2b > 0 && d2b < R.left - 3) temp += 2;
  const int nsfri = fsn_tab[temp];
  auto S = [&](int d, int f) { return sft[(int64_t)(d - 1) * nf + f - 1]; };
  double v[kMaxCcFsn + 1];
  for (int c = threadIdx.x; c < nsfr; c += blockDim.x) {
    // column c + 1 of the compressed matrix, rows 1..nsfl, into v
    if (d1a == 0) {
      for (int row = 1; row <= nsfl; ++row) v[row] = con_at(cyc, nf, R.i, R.j, row, c + 1, nsfl);
    } else {
      // the (c+1)-th spin function f <= nsfri whose shells qualify
      int f = 0, seen = 0;
      for (int g = 1; g <= nsfri; ++g) {
        const bool ok = S(d1a, g) > 0 && (d2b == 0 || S(d2b, g) > 0);
        if (ok && seen++ == c) {
          f = g;
          break;
        }
      }
      if (f == 0) {
        for (int row = 1; row <= nsfl; ++row) v[row] = 0.0;
      } else if (d2b == 0) {
        const int sv = S(d1a, f);
        if (sv == f) {
          for (int row = 1; row <= nsfl; ++row) v[row] = sqrt(2.0) * con_at(cyc, nf, R.i, R.j, row, f, nsfl);
        } else {
          const double ak = 1.0 / (double)spin[(int64_t)(d1a + 1) * nf + f - 1];
          const double m = -1.0;  // sv > 0 here
          const double cf = -m * sqrt(1.0 - m * ak), cs = m * sqrt(1.0 + m * ak);
          for (int row = 1; row <= nsfl; ++row)
            v[row] = cf * con_at(cyc, nf, R.i, R.j, row, f, nsfl) + cs * con_at(cyc, nf, R.i, R.j, row, sv, nsfl);
        }
      } else {
        const int s1 = S(d1a, f), s2 = S(d2b, f);
        if (s1 == f && s2 == f) {
          for (int row = 1; row <= nsfl; ++row) v[row] = 2.0 * con_at(cyc, nf, R.i, R.j, row, f, nsfl);
        } else {
          const double a1 = 1.0 / (double)spin[(int64_t)(d1a + 1) * nf + f - 1];
          const double a2 = 1.0 / (double)spin[(int64_t)(d2b + 1) * nf + f - 1];
          const double m1 = -1.0, m2 = -1.0;  // s1, s2 > 0 here
          const double c_f = (-m1 * sqrt(1 - m1 * a1)) * (-m2 * sqrt(1 - m2 * a2));
          const double c_2 = (-m1 * sqrt(1 - m1 * a1)) * (m2 * sqrt(1 + m2 * a2));
          const double c_1 = (m1 * sqrt(1 + m1 * a1)) * (-m2 * sqrt(1 - m2 * a2));
          const double c_12 = (m1 * sqrt(1 + m1 * a1)) * (m2 * sqrt(1 + m2 * a2));
          const int s12 = S(d1a, s2);
          for (int row = 1; row <= nsfl; ++row)
            v[row] = c_f * con_at(cyc, nf, R.i, R.j, row, f, nsfl) + c_1 * con_at(cyc, nf, R.i, R.j, row, s1, nsfl) +
                     c_2 * con_at(cyc, nf, R.i, R.j, row, s2, nsfl) +
                     (s12 >= 1 ? c_12 * con_at(cyc, nf, R.i, R.j, row, s12, nsfl) : 0.0);
        }
      }
    }
    double* o = out + off[blockIdx.x];
    for (int row = 1; row <= nsfl; ++row) {
      double s;
      if (R.l == R.k) {
        s = v[row];
      } else {
        s = 0.0;
        for (int t = 1; t <= nsfl; ++t) s += con_at(cyc, nf, R.k, R.l, row, t, nsfl) * v[t];
      }
      o[(int64_t)(row - 1) * nsfr + c] = s;
    }
  }
}

// device tables of n open shells (sft, spin, cycles)
struct SpinDev {
  int n = 0, mult = 1, nf = 0, ncyc = 0;
  DevBuf sft, spin, cyc;
};

void build_spin_dev(int n, int mult, SpinDev& D, hipStream_t st) {
  TCI_REQUIRE(n >= 2 && n <= kMaxOpen, "open shells must be in 2.." + std::to_string(kMaxOpen));
  TCI_REQUIRE(mult >= 1 && (n + mult - 1) % 2 == 0 && mult - 1 <= n, "multiplicity does not fit the open shells");
  D.n = n;
  D.mult = mult;
  D.nf = fsn_of(n, mult);
  TCI_REQUIRE(D.nf >= 1 && D.nf <= kMaxFsn, "spin-function count out of range");
  D.ncyc = n * (n - 1) / 2;
  const Branch B = branching_diagram(n, mult);
  TCI_REQUIRE(B.at(1, 1) == D.nf, "branching diagram disagrees with the spin-function count");
  DevBuf dbr;
  upload(dbr, B.br, st);
  D.sft.alloc((size_t)(n - 1) * D.nf * 4);
  D.spin.alloc((size_t)(n + 1) * D.nf * 4);
  hipLaunchKernelGGL(spin_functions_kernel, dim3((unsigned)cdiv(D.nf, 64)), dim3(64), 0, st, n, mult, B.up, B.down,
                     dbr.as<int32_t>(), B.ld, D.nf, D.sft.as<int32_t>(), D.spin.as<int32_t>());
  TCI_HIP(hipGetLastError());
  const int64_t nn = (int64_t)D.nf * D.nf;
  D.cyc.alloc((size_t)D.ncyc * nn * 8);
  launch_fill(D.cyc.as<double>(), D.ncyc * nn, 0.0, st);
  DevBuf wr, wc;
  const size_t nslot = (size_t)(n - 1) * D.nf * (D.nf + 1);
  wr.alloc(nslot * 4);
  wc.alloc(nslot * 8);
  hipLaunchKernelGGL(cycles_kernel, dim3((unsigned)cdiv(D.nf, 64), (unsigned)(n - 1)), dim3(64), 0, st, n, D.nf,
                     D.sft.as<int32_t>(), D.spin.as<int32_t>(), D.cyc.as<double>(), wr.as<int32_t>(), wc.as<double>());
  TCI_HIP(hipGetLastError());
  TCI_HIP(hipStreamSynchronize(st));  // dbr and the workspaces live until here
}

}  // namespace

}  // namespace tci

using namespace tci;

extern "C" int tci_spin_tables(int32_t n_open, int32_t spin_mult, int32_t* sft, int32_t* spin_matrix, double* cycles,
                               double* transpositions) {
  try {
    SpinDev D;
    build_spin_dev(n_open, spin_mult, D, nullptr);
    const int64_t nn = (int64_t)D.nf * D.nf;
    if (sft) TCI_HIP(hipMemcpy(sft, D.sft.p, (size_t)(n_open - 1) * D.nf * 4, hipMemcpyDeviceToHost));
    if (spin_matrix) TCI_HIP(hipMemcpy(spin_matrix, D.spin.p, (size_t)(n_open + 1) * D.nf * 4, hipMemcpyDeviceToHost));
    if (cycles) TCI_HIP(hipMemcpy(cycles, D.cyc.p, (size_t)D.ncyc * nn * 8, hipMemcpyDeviceToHost));
    if (transpositions) {
      DevBuf dt;
      dt.alloc((size_t)D.ncyc * nn * 8);
      hipLaunchKernelGGL(transpositions_kernel, dim3((unsigned)D.nf, (unsigned)D.ncyc), dim3(64), 0, nullptr, n_open,
                         D.nf, D.cyc.as<double>(), dt.as<double>());
      TCI_HIP(hipGetLastError());
      TCI_HIP(hipMemcpy(transpositions, dt.p, (size_t)D.ncyc * nn * 8, hipMemcpyDeviceToHost));
    }
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

extern "C" int tci_contracted_cycles(int32_t n_open, int32_t spin_mult, int64_t n_rec, const int32_t* records,
                                     const int64_t* out_off, double* out) {
  try {
    TCI_REQUIRE(records && out_off && out, "null argument");
    TCI_REQUIRE(spin_mult == 1, "contracted cycles are evaluated for singlet spin functions");
    TCI_REQUIRE(n_open >= 2 && n_open <= kMaxOpen, "open shells must be in 2.." + std::to_string(kMaxOpen));
    TCI_REQUIRE(n_rec >= 1 && n_rec < ((int64_t)1 << 31), "invalid record count");
    std::vector<int32_t> fs(n_open + 1);
    for (int m = 0; m <= n_open; ++m) fs[m] = fsn_of(m, spin_mult);
    std::vector<CcRec> rc((size_t)n_rec);
    int64_t total = 0;
    for (int64_t r = 0; r < n_rec; ++r) {
      const int32_t* x = records + 8 * r;
      CcRec c{x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7]};
      TCI_REQUIRE(c.left >= 2 && c.left <= n_open && c.right >= 0 && c.right <= c.left && (c.left - c.right) % 2 == 0,
                  "record " + std::to_string(r) + ": open-shell counts out of range");
      TCI_REQUIRE(c.i >= 1 && c.i <= c.left && c.j >= 1 && c.j <= c.left && c.k >= 1 && c.k <= c.left && c.l >= 1 &&
                      c.l <= c.left,
                  "record " + std::to_string(r) + ": cycle indices out of range");
      TCI_REQUIRE(c.d1 >= 0 && c.d1 < c.left && c.d2 >= 0 && c.d2 < c.left,
                  "record " + std::to_string(r) + ": shell indices out of range");
      int temp = c.right;
      const int d1a = std::max(c.d1, c.d2), d2b = std::min(c.d1, c.d2);
      if (d1a > 0 && d1a < c.left - 1) temp += 2;
      if (d2b > 0 && d2b < c.left - 3) temp += 2;
      TCI_REQUIRE(temp <= n_open, "record " + std::to_string(r) + ": intermediate open shells exceed the tables");
      rc[(size_t)r] = c;
      TCI_REQUIRE(out_off[r] >= total, "output blocks must not overlap and must come in record order");
      total = out_off[r] + (int64_t)fs[c.left] * fs[c.right];
    }
    SpinDev D;
    build_spin_dev(n_open, spin_mult, D, nullptr);
    TCI_REQUIRE(D.nf <= kMaxCcFsn, "too many spin functions for contracted cycles");
    DevBuf drec, doff, dfs, dout;
    upload(drec, rc);
    std::vector<int64_t> offs(out_off, out_off + n_rec);
    upload(doff, offs);
    upload(dfs, fs);
    dout.alloc((size_t)total * 8);
    hipLaunchKernelGGL(contracted_cycles_kernel, dim3((unsigned)n_rec), dim3(64), 0, nullptr, D.nf,
                       D.sft.as<int32_t>(), D.spin.as<int32_t>(), D.cyc.as<double>(), dfs.as<int32_t>(),
                       drec.as<CcRec>(), doff.as<int64_t>(), dout.as<double>());
    TCI_HIP(hipGetLastError());
    TCI_HIP(hipMemcpy(out, dout.p, (size_t)total * 8, hipMemcpyDeviceToHost));
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
