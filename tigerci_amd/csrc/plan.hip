// Plan construction, sigma orchestration and the C ABI.
//
// Per sigma build (all device work on the caller's stream):
//   unpack        CSF vector -> dense C^{ij}, Ct^{ij} = 2C - C^T, singles s
//   ladder        R^P  = sum_J L^J C^P L^J                 (grouped MFMA launch)
//   rterms        R   += c0 K^{ij} - (ki|bj) s_k^a - (kj|ai) s_k^b
//   hole          R   += sum_kl (ki|lj) C^{kl}
//   Y             Y[J][a][i] = sum_c L^J_ac s_i^c
//   ring          R   += Z^{ij} + Z^{ji T} + (ac|bj) s_i^c + (bc|ai) s_j^c   (grouped gather-GEMM)
//   singles       g = (2K - J) s + sum (ab|jc) Ct^{ij}_bc - sum (ji|kb) Ct^{jk}_ab
//   reference     sigma_0 = sum_ij <K^{ij}, Ct^{ij}>
//   pack          sigma = S^T R (CSF factors) + Fock diagonal
// The formulas are those of the closed-shell MRSDCI sigma of htimesc
// (htimesc.f90:46-147) written in pair-block form.
#include "plan.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>

namespace tci {

static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }

namespace {

inline int64_t rec(int64_t p, int64_t q) {  // 0-based packed orbital-pair record
  if (p < q) std::swap(p, q);
  return p * (p + 1) / 2 + q;
}

Op op(const void* base, int64_t rs, int64_t cs, const int32_t* rmap = nullptr, const int32_t* cmap = nullptr,
      bool f32 = false, int64_t jstride = 0) {
  Op o;
  o.base = base;
  o.rs = rs;
  o.cs = cs;
  o.rmap = rmap;
  o.cmap = cmap;
  o.jstride = jstride;
  o.f32 = f32 ? 1 : 0;
  return o;
}

Seg seg(const Op& A, const Op& B, int K, double alpha) {
  Seg s;
  s.A = A;
  s.B = B;
  s.K = K;
  s.alpha = alpha;
  return s;
}

template <class T>
void upload_now(DevBuf& b, const std::vector<T>& h) {
  upload(b, h, nullptr);
}

}  // namespace
}  // namespace tci

using namespace tci;

void tci_plan::build(const tci_problem& p) {
  TCI_REQUIRE(p.n_int >= 1 && p.n_v >= 1 && p.n_chol >= 1, "n_int, n_v and n_chol must be positive");
  TCI_REQUIRE(p.n_int <= 256 && p.n_v <= 65536, "orbital counts out of range");
  TCI_REQUIRE((p.eps || p.h) && p.chol, "chol and one of eps or h are required");
  TCI_REQUIRE(p.chol_dtype == TCI_F64 || p.chol_dtype == TCI_F32, "chol_dtype must be TCI_F64 or TCI_F32");
  TCI_REQUIRE(p.ladder_dtype == TCI_F64 || p.ladder_dtype == TCI_F32, "ladder_dtype must be TCI_F64 or TCI_F32");
  TCI_REQUIRE(p.n_pairs >= 0 && (p.n_pairs == 0 || (p.pairs && p.dom_off && p.dom_idx)), "pair tables missing");
  n_int = (int)p.n_int;
  n_v = (int)p.n_v;
  n_chol = p.n_chol;
  n_orb = n_int + n_v;
  npair_all = n_orb * (n_orb + 1) / 2;
  n_pairs = (int)p.n_pairs;
  lvv_f32 = p.ladder_dtype == TCI_F32;
  explicit_h = p.h != nullptr;
  if (explicit_h) h_eps.assign(n_orb, 0.0);  // diag(F), set once F is built
  else h_eps.assign(p.eps, p.eps + n_orb);
  h_pairs.assign(p.pairs, p.pairs + 2 * n_pairs);
  h_dom_off.assign(p.dom_off, p.dom_off + n_pairs + 1);
  TCI_REQUIRE(h_dom_off[0] == 0, "dom_off[0] must be 0");
  h_dom_idx.assign(p.dom_idx, p.dom_idx + h_dom_off[n_pairs]);
  pairid.assign((size_t)n_int * n_int, -1);
  for (int P = 0; P < n_pairs; ++P) {
    const int i = h_pairs[2 * P], j = h_pairs[2 * P + 1];
    TCI_REQUIRE(0 <= j && j <= i && i < n_int, "pair walks must satisfy 0 <= j <= i < n_int");
    TCI_REQUIRE(pairid[(size_t)i * n_int + j] < 0, "duplicate pair walk");
    pairid[(size_t)i * n_int + j] = pairid[(size_t)j * n_int + i] = P;
    const int64_t d = h_dom_off[P + 1] - h_dom_off[P];
    TCI_REQUIRE(d >= 1 && d <= 512, "pair domain size must be within [1, 512]");
    for (int64_t q = h_dom_off[P]; q < h_dom_off[P + 1]; ++q) {
      TCI_REQUIRE(h_dom_idx[q] >= 0 && h_dom_idx[q] < n_v, "domain index out of range");
      if (q > h_dom_off[P]) TCI_REQUIRE(h_dom_idx[q] > h_dom_idx[q - 1], "domains must be sorted and unique");
    }
    dmax = std::max<int>(dmax, (int)d);
  }
  // ---- CSF layout ----------------------------------------------------------
  h_aa_off.resize(n_pairs);
  h_ab_off.resize(n_pairs);
  h_cd_off.resize(n_pairs + 1);
  int64_t off = 1 + (int64_t)n_int * n_v;
  sum_d2 = 0;
  for (int P = 0; P < n_pairs; ++P) {
    const int64_t d = h_dom_off[P + 1] - h_dom_off[P];
    const int nsp = h_pairs[2 * P] == h_pairs[2 * P + 1] ? 1 : 2;
    h_aa_off[P] = off;
    off += d;
    h_ab_off[P] = off;
    off += nsp * (d * (d - 1) / 2);
    h_cd_off[P] = sum_d2;
    sum_d2 += d * d;
  }
  h_cd_off[n_pairs] = sum_d2;
  n_csf = off;

  hipStream_t st = nullptr;
  upload_now(pairs, h_pairs);
  upload_now(dom_off, h_dom_off);
  upload_now(dom_idx, h_dom_idx);
  upload_now(cd_off, h_cd_off);
  upload_now(aa_off, h_aa_off);
  upload_now(ab_off, h_ab_off);
  upload_now(eps, h_eps);
  tab.n_pairs = n_pairs;
  tab.n_int = n_int;
  tab.n_v = n_v;
  tab.pairs = pairs.as<int32_t>();
  tab.dom_off = dom_off.as<int64_t>();
  tab.dom_idx = dom_idx.as<int32_t>();
  tab.cd_off = cd_off.as<int64_t>();
  tab.aa_off = aa_off.as<int64_t>();
  tab.ab_off = ab_off.as<int64_t>();
  tab.eps = eps.as<double>();

  // ---- Cholesky records -> device layouts ----------------------------------
  const int64_t niv = (int64_t)n_int * n_v;
  DevBuf Lpm;
  Lpm.alloc((size_t)npair_all * n_chol * sizeof(double));
  {
    const size_t esz = p.chol_dtype == TCI_F64 ? 8 : 4;
    const size_t total = (size_t)npair_all * n_chol;
    if (esz == 8) {
      TCI_HIP(hipMemcpy(Lpm.p, p.chol, total * 8, hipMemcpyHostToDevice));
    } else {
      // convert FP32 records on the host in bounded chunks
      const float* src = static_cast<const float*>(p.chol);
      const size_t chunk = (size_t)1 << 24;
      std::vector<double> buf(std::min(chunk, total));
      for (size_t o = 0; o < total; o += chunk) {
        const size_t m = std::min(chunk, total - o);
        for (size_t q = 0; q < m; ++q) buf[q] = (double)src[o + q];
        TCI_HIP(hipMemcpy(Lpm.as<double>() + o, buf.data(), m * 8, hipMemcpyHostToDevice));
      }
    }
  }
  Lvv.alloc((size_t)n_chol * n_v * n_v * (lvv_f32 ? 4 : 8));
  launch_gather_Lvv(Lpm.as<double>(), n_chol, n_int, n_v, Lvv.p, lvv_f32, st);
  Lov.alloc((size_t)n_chol * niv * 8);
  launch_gather_Lov(Lpm.as<double>(), n_chol, n_int, n_v, Lov.as<double>(), st);

  // ---- one-off integral assembly on MFMA -------------------------------------
  std::vector<double> o4((size_t)n_int * n_int * n_int * n_int);
  {
    std::vector<int32_t> rec_ov(niv), rec_oo((size_t)n_int * n_int), rec_aa(n_v),
        rec_pairs((size_t)n_int * (n_int + 1) / 2), rec_vv((size_t)n_v * n_v);
    for (int i = 0; i < n_int; ++i)
      for (int a = 0; a < n_v; ++a) {
        rec_ov[(size_t)i * n_v + a] = (int32_t)rec(n_int + a, i);
      }
    for (int k = 0; k < n_int; ++k)
      for (int i = 0; i < n_int; ++i) rec_oo[(size_t)k * n_int + i] = (int32_t)rec(k, i);
    for (int k = 0; k < n_int; ++k)
      for (int l = 0; l <= k; ++l) rec_pairs[(size_t)k * (k + 1) / 2 + l] = (int32_t)rec(k, l);
    for (int a = 0; a < n_v; ++a) {
      rec_aa[a] = (int32_t)rec(n_int + a, n_int + a);
      for (int b = 0; b < n_v; ++b) rec_vv[(size_t)a * n_v + b] = (int32_t)rec(n_int + a, n_int + b);
    }
    TCI_REQUIRE(npair_all < ((int64_t)1 << 31), "too many orbital pairs");
    DevBuf d_ov, d_oo, d_aa, d_pr, d_vv;
    upload_now(d_ov, rec_ov);
    upload_now(d_oo, rec_oo);
    upload_now(d_aa, rec_aa);
    upload_now(d_pr, rec_pairs);
    upload_now(d_vv, rec_vv);
    const double* L = Lpm.as<double>();
    const int npr = n_int * (n_int + 1) / 2;
    Kov.alloc((size_t)niv * niv * 8);
    Jvv.alloc((size_t)npr * n_v * n_v * 8);
    O4.alloc((size_t)n_int * n_int * n_int * n_int * 8);
    O3.alloc((size_t)n_int * n_int * niv * 8);
    vvd.alloc((size_t)n_v * n_v * 8);
    vvx.alloc((size_t)n_v * n_v * 8);
    GemmProgram asmb;
    asmb.tile = 128;
    const int K = (int)n_chol;
    // rows: records (k contiguous, cs = 1); cols: records transposed (rs = 1, cs = n_chol)
    asmb.add(Kov.as<double>(), niv, (int)niv, (int)niv,
             {seg(op(L, n_chol, 1, d_ov.as<int32_t>()), op(L, 1, n_chol, nullptr, d_ov.as<int32_t>()), K, 1.0)},
             false);
    asmb.add(Jvv.as<double>(), (int64_t)n_v * n_v, npr, n_v * n_v,
             {seg(op(L, n_chol, 1, d_pr.as<int32_t>()), op(L, 1, n_chol, nullptr, d_vv.as<int32_t>()), K, 1.0)},
             false);
    asmb.add(O4.as<double>(), (int64_t)n_int * n_int, n_int * n_int, n_int * n_int,
             {seg(op(L, n_chol, 1, d_oo.as<int32_t>()), op(L, 1, n_chol, nullptr, d_oo.as<int32_t>()), K, 1.0)},
             false);
    // O3[k][p][q][b] = (kp|bq): rows (k,p), columns (q,b) with b fastest
    asmb.add(O3.as<double>(), niv, n_int * n_int, (int)niv,
             {seg(op(L, n_chol, 1, d_oo.as<int32_t>()), op(L, 1, n_chol, nullptr, d_ov.as<int32_t>()), K, 1.0)},
             false);
    asmb.add(vvd.as<double>(), n_v, n_v, n_v,
             {seg(op(L, n_chol, 1, d_aa.as<int32_t>()), op(L, 1, n_chol, nullptr, d_aa.as<int32_t>()), K, 1.0)},
             false);
    asmb.upload_tables(st);
    asmb.launch(st);
    launch_vv_exch(Lvv.p, lvv_f32, n_chol, n_v, vvx.as<double>(), st);
    Mss.alloc((size_t)niv * niv * 8);
    launch_build_mss(Kov.as<double>(), Jvv.as<double>(), n_int, n_v, Mss.as<double>(), st);
    TCI_HIP(hipMemcpy(o4.data(), O4.p, o4.size() * 8, hipMemcpyDeviceToHost));
    if (!explicit_h) {
      // E_ref = 2 sum_i eps_i - sum_ik [2 (ii|kk) - (ik|ik)]  (closed shell, Fock = diag(eps))
      auto O = [&](int a, int b, int c, int d) { return o4[(((size_t)a * n_int + b) * n_int + c) * n_int + d]; };
      double e = 0.0;
      for (int i = 0; i < n_int; ++i) {
        e += 2.0 * h_eps[i];
        for (int k = 0; k < n_int; ++k) e -= 2.0 * O(i, i, k, k) - O(i, k, i, k);
      }
      e_ref = e;
    }
    TCI_HIP(hipDeviceSynchronize());
  }
  Lpm.reset();

  // ---- explicit one-electron integrals: Fock matrix, orbital energies, E_ref --
  // The pp array of the reference (getpp, get_integrals_mod.f90) in its packed
  // lower-triangle order.  F = h + G; the orbital energies are diag(F) and
  // E_ref = sum_i (h_ii + f_ii).  Off-diagonal elements couple the CSFs:
  // f_ij (i != j) through O4 and Mss, f_ab (a != b) through Mss and the
  // per-walk products F'[D,D] C_P + C_P F'[D,D], f_ia through O3 (singles <->
  // doubles) and the reference <-> singles kernel.
  std::vector<double> hF;
  if (explicit_h) {
    const int64_t npk = n_orb * (n_orb + 1) / 2;
    std::vector<double> hp(p.h, p.h + npk);
    DevBuf d_h;
    upload_now(d_h, hp);
    Fm.alloc((size_t)n_orb * n_orb * 8);
    launch_fock(d_h.as<double>(), O4.as<double>(), O3.as<double>(), Kov.as<double>(), Jvv.as<double>(), n_int, n_v,
                Fm.as<double>(), st);
    hF.resize((size_t)n_orb * n_orb);
    TCI_HIP(hipMemcpy(hF.data(), Fm.p, hF.size() * 8, hipMemcpyDeviceToHost));
    double e = 0.0;
    for (int64_t q = 0; q < n_orb; ++q) {
      h_eps[q] = hF[(size_t)q * n_orb + q];
      TCI_REQUIRE(std::isfinite(h_eps[q]), "non-finite Fock matrix element");
    }
    for (int i = 0; i < n_int; ++i) e += hp[(size_t)i * (i + 1) / 2 + i] + h_eps[i];
    e_ref = e;
    TCI_HIP(hipMemcpy(eps.p, h_eps.data(), (size_t)n_orb * 8, hipMemcpyHostToDevice));
  }

  diag.alloc((size_t)n_csf * 8);
  launch_diag(tab, dmax, Kov.as<double>(), Jvv.as<double>(), O4.as<double>(), vvd.as<double>(), vvx.as<double>(),
              diag.as<double>(), st);

  if (explicit_h) {
    // fold the off-diagonal Fock elements into the coupling tables (after the
    // diagonal, which reads only O4 elements the fold leaves unchanged):
    //   O4[k][i][l][j] -= f'_ki d_lj + d_ki f'_lj   (hole term over C^{kl})
    //   O3[k][k][q][b] -= f_qb,  Mss += d_ij f'_ab - d_ab f'_ij
    auto Fh = [&](int64_t a, int64_t b) { return hF[(size_t)a * n_orb + b]; };
    const size_t n = n_int;
    for (size_t k = 0; k < n; ++k)
      for (size_t i = 0; i < n; ++i)
        for (size_t l = 0; l < n; ++l)
          for (size_t j = 0; j < n; ++j) {
            double v = 0.0;
            if (l == j && k != i) v += Fh(k, i);
            if (k == i && l != j) v += Fh(l, j);
            if (v != 0.0) o4[((k * n + i) * n + l) * n + j] -= v;
          }
    TCI_HIP(hipMemcpy(O4.p, o4.data(), o4.size() * 8, hipMemcpyHostToDevice));
    launch_fold_fock(Fm.as<double>(), n_int, n_v, O3.as<double>(), Mss.as<double>(), st);
    std::vector<double> fvv((size_t)n_v * n_v, 0.0), fov((size_t)niv, 0.0);
    for (int a = 0; a < n_v; ++a)
      for (int b = 0; b < n_v; ++b)
        if (a != b) {
          fvv[(size_t)a * n_v + b] = Fh(n_int + a, n_int + b);
          has_fvv = has_fvv || fvv[(size_t)a * n_v + b] != 0.0;
        }
    for (int i = 0; i < n_int; ++i)
      for (int a = 0; a < n_v; ++a) {
        fov[(size_t)i * n_v + a] = Fh(i, n_int + a);
        has_fov = has_fov || fov[(size_t)i * n_v + a] != 0.0;
      }
    if (has_fvv) upload_now(Fvvp, fvv);
    if (has_fov) upload_now(Fov, fov);
    TCI_HIP(hipDeviceSynchronize());
  }

  // ---- domain maps ---------------------------------------------------------
  std::vector<int32_t> h_pos((size_t)std::max(n_pairs, 1) * n_v, -1);
  for (int Q = 0; Q < n_pairs; ++Q)
    for (int64_t q = h_dom_off[Q]; q < h_dom_off[Q + 1]; ++q)
      h_pos[(size_t)Q * n_v + h_dom_idx[q]] = (int32_t)(q - h_dom_off[Q]);
  upload_now(pos, h_pos);
  // composite maps D_P -> position in D_Q for every (P,Q) sharing an internal index
  std::map<std::pair<int, int>, int64_t> xoff;
  std::vector<int32_t> h_xmap;
  auto need = [&](int P, int Q) {
    auto key = std::make_pair(P, Q);
    if (xoff.count(key)) return;
    xoff[key] = (int64_t)h_xmap.size();
    for (int64_t q = h_dom_off[P]; q < h_dom_off[P + 1]; ++q) h_xmap.push_back(h_pos[(size_t)Q * n_v + h_dom_idx[q]]);
  };
  for (int P = 0; P < n_pairs; ++P) {
    const int i = h_pairs[2 * P], j = h_pairs[2 * P + 1];
    for (int k = 0; k < n_int; ++k) {
      int Q1 = pairid[(size_t)i * n_int + k], Q2 = pairid[(size_t)j * n_int + k];
      if (Q1 >= 0) need(P, Q1);
      if (Q2 >= 0) need(P, Q2);
    }
  }
  upload_now(xmap, h_xmap);
  auto XM = [&](int P, int Q) { return xmap.as<int32_t>() + xoff.at(std::make_pair(P, Q)); };
  auto DL = [&](int P) { return dom_idx.as<int32_t>() + h_dom_off[P]; };
  auto dsz = [&](int P) { return (int)(h_dom_off[P + 1] - h_dom_off[P]); };
  // intersections for the 4-internal term, sorted by position in P
  {
    std::vector<int64_t> ioff((size_t)n_pairs * n_pairs + 1, 0);
    std::vector<int32_t> is;
    for (int P = 0; P < n_pairs; ++P)
      for (int Q = 0; Q < n_pairs; ++Q) {
        ioff[(size_t)P * n_pairs + Q] = (int64_t)is.size() / 2;
        int64_t a = h_dom_off[P], b = h_dom_off[Q];
        while (a < h_dom_off[P + 1] && b < h_dom_off[Q + 1]) {
          if (h_dom_idx[a] < h_dom_idx[b]) ++a;
          else if (h_dom_idx[a] > h_dom_idx[b]) ++b;
          else {
            is.push_back((int32_t)(a - h_dom_off[P]));
            is.push_back((int32_t)(b - h_dom_off[Q]));
            ++a;
            ++b;
          }
        }
      }
    ioff[(size_t)n_pairs * n_pairs] = (int64_t)is.size() / 2;
    upload_now(isect_off, ioff);
    upload_now(isect, is);
    // inverse lists: virtual a -> (Q, position)
    std::vector<int64_t> ivo(n_v + 1, 0);
    for (int64_t q = 0; q < (int64_t)h_dom_idx.size(); ++q) ivo[h_dom_idx[q] + 1]++;
    for (int a = 0; a < n_v; ++a) ivo[a + 1] += ivo[a];
    std::vector<int32_t> iq(h_dom_idx.size()), ip(h_dom_idx.size());
    std::vector<int64_t> fill(ivo.begin(), ivo.end() - 1);
    for (int Q = 0; Q < n_pairs; ++Q)
      for (int64_t q = h_dom_off[Q]; q < h_dom_off[Q + 1]; ++q) {
        const int a = h_dom_idx[q];
        iq[fill[a]] = Q;
        ip[fill[a]] = (int32_t)(q - h_dom_off[Q]);
        fill[a]++;
      }
    upload_now(inv_off, ivo);
    upload_now(inv_q, iq);
    upload_now(inv_pos, ip);
    std::vector<int64_t> go(2 * (size_t)n_pairs + 1, 0);
    for (int Q = 0; Q < n_pairs; ++Q) {
      go[2 * Q + 1] = go[2 * Q] + (int64_t)n_int * dsz(Q);
      go[2 * Q + 2] = go[2 * Q + 1] + (int64_t)n_int * dsz(Q);
    }
    upload_now(gpart_off, go);
    gpart.alloc((size_t)go[2 * n_pairs] * 8);
  }

  // ---- work buffers ----------------------------------------------------------
  C.alloc((size_t)sum_d2 * 8);
  Ct.alloc((size_t)sum_d2 * 8);
  R.alloc((size_t)sum_d2 * 8);
  s.alloc((size_t)niv * 8);
  sT.alloc((size_t)niv * 8);
  g.alloc((size_t)niv * 8);
  gss.alloc((size_t)niv * 8);
  Y.alloc((size_t)n_chol * niv * 8);
  W.alloc((size_t)n_chol * niv * 8);
  part.alloc((size_t)std::max(n_pairs, 1) * ref_dot_slices() * 8);
  sig0.alloc(8);
  {
    // K-splits of gW = W L (a product with only n_int rows that streams all of
    // L): about 2048 64 x 64 tiles, at least 4 Cholesky vectors per split
    const int64_t tiles = cdiv(n_int, 64) * cdiv(n_v, 64);
    n_split = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(n_chol / 4, 512), cdiv(2048, tiles)));
  }
  gW.alloc((size_t)n_split * niv * 8);

  // ---- grouped task tables -----------------------------------------------------
  const double* Kp = Kov.as<double>();
  const double* Jp = Jvv.as<double>();
  const double* Cp = C.as<double>();
  const double* Ctp = Ct.as<double>();
  auto Kop = [&](int k, int l, const int32_t* rm, const int32_t* cm) {  // K^{kl}[rows, cols] = (ka|lb)
    return op(Kp + (int64_t)k * n_v * niv + (int64_t)l * n_v, niv, 1, rm, cm);
  };
  auto Jop = [&](int k, int l, const int32_t* rm, const int32_t* cm) {  // J^{kl}[rows, cols] = (kl|ab)
    return op(Jp + rec(k, l) * n_v * n_v, n_v, 1, rm, cm);
  };
  // dense block of strong walk Q seen as the ordered pair (u,v): element [r][c]
  // of X^{uv} = (u >= v) ? X_Q[r][c] : X_Q[c][r]
  auto blk = [&](const double* X, int Q, bool transposed, const int32_t* rm, const int32_t* cm) {
    const int d = dsz(Q);
    const double* b = X + h_cd_off[Q];
    return transposed ? op(b, 1, d, rm, cm) : op(b, d, 1, rm, cm);
  };
  {
    // flops of the per-walk gathered ring and 3-external products
    double fg = 0.0;
    for (int P = 0; P < n_pairs; ++P) {
      const double d = dsz(P);
      const int i = h_pairs[2 * P], j = h_pairs[2 * P + 1];
      for (int k = 0; k < n_int; ++k) {
        const int Q1 = pairid[(size_t)i * n_int + k], Q2 = pairid[(size_t)j * n_int + k];
        if (Q1 >= 0) fg += 3.0 * 2.0 * d * d * dsz(Q1);
        if (Q2 >= 0) fg += 3.0 * 2.0 * d * d * dsz(Q2);
      }
      fg += 2.0 * 2.0 * d * d * (double)n_chol + 2.0 * d * n_v * (double)n_chol;
    }
    dense_pairs = dense_pairs_choose(*this, fg);
  }
  ring.clear();
  ring.tile = 128;
  for (int P = 0; P < n_pairs; ++P) {
    const int i = h_pairs[2 * P], j = h_pairs[2 * P + 1];
    const int d = dsz(P);
    const int32_t* DP = DL(P);
    std::vector<Seg> sg;
    for (int k = 0; k < n_int && !dense_pairs; ++k) {
      const int Q1 = pairid[(size_t)i * n_int + k];
      if (Q1 >= 0) {  // Z^{ij}: Ct^{ik} K^{kj} - C^{ik} J^{kj} - J^{kj} C^{ik}
        const int dq = dsz(Q1);
        const bool tr = i < k;
        const int32_t* m = XM(P, Q1);
        sg.push_back(seg(blk(Ctp, Q1, tr, m, nullptr), Kop(k, j, DL(Q1), DP), dq, 1.0));
        sg.push_back(seg(blk(Cp, Q1, tr, m, nullptr), Jop(k, j, DL(Q1), DP), dq, -1.0));
        sg.push_back(seg(Jop(k, j, DP, DL(Q1)), blk(Cp, Q1, tr, nullptr, m), dq, -1.0));
      }
      const int Q2 = pairid[(size_t)j * n_int + k];
      if (Q2 >= 0) {  // (Z^{ji})^T: K^{ik} Ct^{kj} - J^{ki} C^{kj} - C^{kj} J^{ki}
        const int dq = dsz(Q2);
        // X^{kj}[c][b] = X^{jk}[b][c]: element [c][b] of the ordered pair (k,j)
        const bool tr = k < j;
        const int32_t* m = XM(P, Q2);
        sg.push_back(seg(Kop(i, k, DP, DL(Q2)), blk(Ctp, Q2, tr, nullptr, m), dq, 1.0));
        sg.push_back(seg(Jop(k, i, DP, DL(Q2)), blk(Cp, Q2, tr, nullptr, m), dq, -1.0));
        sg.push_back(seg(blk(Cp, Q2, tr, m, nullptr), Jop(k, i, DL(Q2), DP), dq, -1.0));
      }
    }
    if (dense_pairs) continue;
    // 3-external singles: X^{ij}_ab + X^{ji}_ba, X^{ij}_ab = sum_J Y[J][a][i] L^J_{jb}
    const double* Yp = Y.as<double>();
    const double* Lo = Lov.as<double>();
    sg.push_back(seg(op(Yp + i, n_int, niv, DP, nullptr), op(Lo + (int64_t)j * n_v, niv, 1, nullptr, DP), (int)n_chol, 1.0));
    sg.push_back(seg(op(Lo + (int64_t)i * n_v, 1, niv, DP, nullptr), op(Yp + j, niv, n_int, nullptr, DP), (int)n_chol, 1.0));
    ring.add(R.as<double>() + h_cd_off[P], d, d, d, sg, true);
  }
  ring.upload_tables(st);
  // off-diagonal virtual Fock elements on every pair block:
  // R_P += F'[D,D] C_P + C_P F'[D,D]
  fockvv.clear();
  fockvv.tile = 128;
  if (has_fvv) {
    const double* Fv = Fvvp.as<double>();
    for (int P = 0; P < n_pairs; ++P) {
      const int d = dsz(P);
      const int32_t* DP = DL(P);
      fockvv.add(R.as<double>() + h_cd_off[P], d, d, d,
                 {seg(op(Fv, n_v, 1, DP, DP), blk(Cp, P, false, nullptr, nullptr), d, 1.0),
                  seg(blk(Cp, P, false, nullptr, nullptr), op(Fv, n_v, 1, DP, DP), d, 1.0)},
                 true);
    }
    fockvv.upload_tables(st);
  }
  if (dense_pairs) dense_pairs_build(*this, o4, st);
  // Y[(J,a)][i] = sum_c L^J_ac s_i^c
  // (a dense product: L^J rows are contiguous, and s is transposed into sT per sigma)
  progY.clear();
  progY.dense = true;
  progY.tile = 64;
  progY.add(Y.as<double>(), n_int, (int)(n_chol * n_v), n_int,
            {seg(op(Lvv.p, n_v, 1, nullptr, nullptr, lvv_f32), op(sT.as<double>(), n_int, 1), n_v, 1.0)}, false);
  progY.upload_tables(st);
  // W[i][J][b] = sum_j sum_c Ct^{ij}[pos(b)][c] L^J_{j D[c]}
  progW.clear();
  progW.tile = 128;
  for (int i = 0; i < n_int && !dense_pairs; ++i) {
    std::vector<Seg> sg;
    for (int j = 0; j < n_int; ++j) {
      const int Q = pairid[(size_t)i * n_int + j];
      if (Q < 0) continue;
      const int dq = dsz(Q);
      const bool tr = i < j;  // Ct^{ij} = Ct_Q^T when i < j
      const double* b = Ctp + h_cd_off[Q];
      // B(c, b) = Ct^{ij}[pos(b)][c]
      Op B = tr ? op(b, dq, 1, nullptr, pos.as<int32_t>() + (size_t)Q * n_v)
                : op(b, 1, dq, nullptr, pos.as<int32_t>() + (size_t)Q * n_v);
      sg.push_back(seg(op(Lov.as<double>() + (int64_t)j * n_v, niv, 1, nullptr, DL(Q)), B, dq, 1.0));
    }
    progW.add(W.as<double>() + (int64_t)i * n_chol * n_v, n_v, (int)n_chol, n_v, sg, false);
  }
  progW.upload_tables(st);
  // gW[s][i][a] = sum_{J in split s} sum_b W[i][J][b] L^J_ba (a dense product)
  progGW.clear();
  progGW.dense = true;
  progGW.tile = 64;
  for (int sp = 0; sp < n_split; ++sp) {
    const int64_t J0 = n_chol * sp / n_split, J1 = n_chol * (sp + 1) / n_split;
    const void* Lb = lvv_f32 ? (const void*)(Lvv.as<float>() + J0 * n_v * n_v) : (const void*)(Lvv.as<double>() + J0 * n_v * n_v);
    progGW.add(gW.as<double>() + (int64_t)sp * niv, n_v, n_int, n_v,
               {seg(op(W.as<double>() + J0 * n_v, n_chol * n_v, 1), op(Lb, n_v, 1, nullptr, nullptr, lvv_f32),
                    (int)((J1 - J0) * n_v), 1.0)},
               false);
  }
  progGW.upload_tables(st);
  // gss = (2K - J) s runs as a GEMV (launch_gemv) inside every sigma

  // ---- ladder work items -------------------------------------------------------
  {
    int64_t panels = 0;
    for (int P = 0; P < n_pairs; ++P) panels += cdiv(dsz(P), 64);
    const int64_t target = 2048;
    lad_chunks = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(16, n_chol), cdiv(target, std::max<int64_t>(panels, 1))));
    std::vector<LadderItem> items;
    std::vector<double> cost;
    for (int P = 0; P < n_pairs; ++P) {
      const int d = dsz(P);
      const double dp = 16.0 * cdiv(d, 16);
      for (int r0 = 0; r0 < d; r0 += 64)
        for (int c = 0; c < lad_chunks; ++c) {
          LadderItem it{};
          it.P = P;
          it.r0 = r0;
          it.J0 = (int)(n_chol * c / lad_chunks);
          it.J1 = (int)(n_chol * (c + 1) / lad_chunks);
          it.chunk = c;
          items.push_back(it);
          cost.push_back(std::min(64, d - r0) * dp * dp * (it.J1 - it.J0));
        }
    }
    const char* env = std::getenv("TCI_LADDER");
    ladder_fused = env && std::string(env) == "fused";
    if (ladder_fused) {
      ladder.build(items, cost, dmax, st);
      lad_part.alloc((size_t)lad_chunks * sum_d2 * 8);
    }
  }
  if (!ladder_fused) build_two_phase_ladder(st);
  // ladder from the resident (ab|cd) block: with autotuning, built when it
  // fits and its product costs under twice the Cholesky-direct flops, then
  // timed against the Cholesky-direct ladder (the faster is kept);
  // TCI_LADDER_VV=1 forces it whenever it fits (with or without autotuning),
  // TCI_LADDER_VV=0 switches it off
  if (!ladder_fused && n_pairs > 0) {
    const char* env = std::getenv("TCI_LADDER_VV");
    const int force = env && env[0] == '1' ? 1 : (env && env[0] == '0' ? 0 : -1);
    double f_chol = 0.0;
    for (int P = 0; P < n_pairs; ++P) f_chol += 4.0 * std::pow((double)dsz(P), 3) * (double)n_chol;
    const double f_vv = vv_flops(*this);
    double need = 0.0;
    const bool fits = force != 0 && vv_fits(*this, &need);
    const bool tune = autotune_enabled();
    const bool try_vv = fits && (force == 1 || (tune && f_vv < 2.0 * f_chol));
    if (force == 1 && !fits) tuning += "fp64 ladder (ab|cd) resident: does not fit (" + std::to_string(need / 1e9) + " GB)\n";
    if (try_vv) {
      vv_build(*this, st);
      char head[160];
      std::snprintf(head, sizeof head, "fp64 ladder (ab|cd) resident: %.3g flops against %.3g Cholesky-direct;",
                    f_vv, f_chol);
      std::string log = head;
      bool use = true;
      if (tune) {
        const double tv = vv_tune(*this, TCI_SIGMA_FP64, st, log);
        if (force != 1) {
          const double tc = time_ms(st, 1, [&] { two_phase_ladder(st); });
          log += "; resident " + fmt_ms(tv) + ", Cholesky-direct " + fmt_ms(tc);
          use = tv < tc;
        }
      }
      vv.fp64 = use;
      tuning += log + (use ? " -> resident\n" : " -> Cholesky-direct\n");
      if (use) {  // the two-phase ladder's buffers are not used again
        Tbuf.reset();
        Lp.reset();
        Rpart.reset();
        Cw.reset();
        lchunks.clear();
        lphase.clear();
      }
    }
  }

  // order of the dense pair products' tiles (plain, or XCD-contiguous in
  // blocks of tiles that share operand panels in one XCD's L2)
  if (dense_pairs && autotune_enabled()) {
    std::string log = "fp64 pair products: order";
    for (GemmProgram* g : {&ringZ, &ringZ3, &progU, &progWd, &holeG})
      if (!g->tasks.empty()) log += " " + tune_block_order(*g, st, [&] { g->launch(st); });
    tuning += log + "\n";
  }

  // ---- accounting ---------------------------------------------------------------
  flops_ladder = 0.0;
  double f_ring = 0.0, bytes = 0.0;
  for (int P = 0; P < n_pairs; ++P) {
    const double d = dsz(P);
    flops_ladder += 4.0 * d * d * d * (double)n_chol;
    bytes += (double)n_chol * d * d * (lvv_f32 ? 4 : 8) + 2.0 * 8.0 * d * d;
    const int i = h_pairs[2 * P], j = h_pairs[2 * P + 1];
    for (int k = 0; k < n_int; ++k) {
      const int Q1 = pairid[(size_t)i * n_int + k], Q2 = pairid[(size_t)j * n_int + k];
      if (Q1 >= 0) f_ring += 3.0 * 2.0 * d * d * dsz(Q1);
      if (Q2 >= 0) f_ring += 3.0 * 2.0 * d * d * dsz(Q2);
    }
    f_ring += 2.0 * 2.0 * d * d * (double)n_chol;
  }
  bytes_ladder = bytes;
  if (dense_pairs) {  // executed products: Z (two or three terms), U, Z3, W, the 4-internal GEMM
    const double N = (double)niv, d2 = (double)n_v * n_v;
    f_ring = (z_lowrank ? 4.0 * N * N * N + 6.0 * N * N * (double)n_chol : 6.0 * N * N * N + 4.0 * N * N * (double)n_chol) +
             2.0 * n_pairs * 2.0 * n_pairs * d2;
  }
  flops_sigma = flops_ladder + f_ring + 2.0 * (double)n_chol * n_v * niv * 3.0 + 2.0 * (double)niv * niv;
  TCI_HIP(hipDeviceSynchronize());
  device_bytes = 0;
  for (DevBuf* b : {&pairs, &dom_off, &dom_idx, &cd_off, &aa_off, &ab_off, &eps, &pos, &xmap, &isect_off, &isect,
                    &inv_off, &inv_q, &inv_pos, &gpart_off, &Lvv, &Lov, &Kov, &Jvv, &Mss, &O4, &O3, &vvd, &vvx,
                    &diag, &C, &Ct, &R, &s, &sT, &g, &gss, &gW, &Y, &W, &gpart, &part, &sig0, &lad_part, &Jbig, &CRb,
                    &CtRb, &CCb, &Zb, &Z3b, &Cemb, &CTemb, &Remb, &Yt, &holeA, &LovT, &Tbuf, &Rpart, &Lp, &Cw, &lp_off_all, &Fm,
                    &Fvvp, &Fov})
    device_bytes += (int64_t)b->bytes;
  device_bytes += vv_device_bytes(*this);
}

void tci_plan::drop_graphs() {
  if (graphs.empty() && iter_graphs.empty()) return;
  // a replay may still be in flight on any stream
  TCI_HIP(hipDeviceSynchronize());
  for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
  graphs.clear();
  for (auto& kv : iter_graphs) (void)hipGraphExecDestroy(kv.second);
  iter_graphs.clear();
}

tci_plan::~tci_plan() {
  for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : iter_graphs) (void)hipGraphExecDestroy(kv.second);
  if (iter_host) (void)hipHostFree(iter_host);
  if (gstream) (void)hipStreamDestroy(gstream);
  for (hipEvent_t& e : ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t& e : kev) (void)hipEventDestroy(e);
}

int tci_plan::kt_begin(hipStream_t st) {
  if (!timing()) return -1;
  if (kev_used + 2 > kev.size()) {
    if (kev.size() >= 8192) return -1;  // pool exhausted: this launch goes untimed
    for (int q = 0; q < 64; ++q) {
      hipEvent_t e = nullptr;
      TCI_HIP(hipEventCreate(&e));
      kev.push_back(e);
    }
  }
  TCI_HIP(hipEventRecord(kev[kev_used], st));
  return (int)kev_used++;
}

void tci_plan::kt_end(int e0, const std::string& name, double flops, int prec, hipStream_t st) {
  if (e0 < 0) return;
  TCI_HIP(hipEventRecord(kev[kev_used], st));
  kpend.push_back(KRec{name, flops, e0, (int)kev_used, prec});
  ++kev_used;
}

void tci_plan::kt_collect() {
  for (const KRec& r : kpend) {
    TCI_HIP(hipEventSynchronize(kev[r.e1]));
    float ms = 0.f;
    TCI_HIP(hipEventElapsedTime(&ms, kev[r.e0], kev[r.e1]));
    KStat& k = kstats[r.prec][r.name];
    ++k.n;
    k.ms += ms;
    k.flops += r.flops;
  }
  kpend.clear();
  kev_used = 0;
}

// Two-phase ladder tables (dense FP64 MFMA programs).  Per chunk of Lc
// Cholesky vectors starting at J0:
//   pack    Lp_P[jj] = L^{J0+jj}[D_P, D_P]  (local-domain pairs only; full-domain
//           pairs read L^J directly)
//   phase 1 T_P[(jj,a)][c] = sum_k Lp_P[jj][a][k] C_P[k][c]       (128-wide tiles)
//   phase 2 R_P[a][n]    += sum_jj sum_c T_P[jj][a][c] Lp_P[jj][c][n]
//           (64-wide tiles, one K-segment per jj, K split into partial R
//           buffers when the outputs alone cannot fill the chip)
// Lc bounds T (8 bytes) plus Lp (4 or 8 bytes) per element of sum_P d_P^2.
void tci_plan::build_two_phase_ladder(hipStream_t st) {
  lphase.clear();
  lchunks.clear();
  const size_t lsz = lvv_f32 ? 4 : 8;
  int64_t sum_local = 0;
  std::vector<int32_t> loc;
  for (int P = 0; P < n_pairs; ++P) {
    const int64_t d = h_dom_off[P + 1] - h_dom_off[P];
    if (d != n_v) {
      loc.push_back(P);
      sum_local += d * d;
    }
  }
  n_local = (int)loc.size();
  upload(pairs_local, loc, st);
  {
    const char* env = std::getenv("TCI_LADDER_MERGED");
    ladder_merged = n_local == 0 && n_pairs > 0 && !(env && env[0] == '0');
  }
  // L^J[D_P, D_P] of every local walk packed once per plan when that takes at
  // most 96 GB and leaves 64 GB free (the split ladder's copies and the
  // Davidson vectors); otherwise packed per chunk inside every sigma
  {
    const char* env = std::getenv("TCI_FP64_PREPACK");
    const double need = (double)sum_local * (double)n_chol * (double)lsz;
    size_t freeb = 0, totalb = 0;
    TCI_HIP(hipMemGetInfo(&freeb, &totalb));
    lp_all = sum_local > 0 && !(env && env[0] == '0') && need <= 96.0e9 && need + 64.0e9 <= (double)freeb;
  }
  std::vector<int64_t> lpo_all(loc.size() + 1, 0);
  if (lp_all) {
    for (size_t q = 0; q < loc.size(); ++q) {
      const int64_t d = h_dom_off[loc[q] + 1] - h_dom_off[loc[q]];
      lpo_all[q + 1] = lpo_all[q] + n_chol * d * d;
    }
    upload(lp_off_all, lpo_all, st);
    Lp.alloc((size_t)lpo_all.back() * lsz);
    PackTab K{pairs_local.as<int32_t>(), lp_off_all.as<int64_t>(), n_local};
    launch_pack_L(tab, K, Lvv.p, lvv_f32, 0, n_chol, Lp.p, st);
  }
  const double budget = 8.0e9;  // bytes of T + Lp per chunk
  const double per_J = 8.0 * (double)std::max<int64_t>(sum_d2, 1) + (lp_all ? 0.0 : (double)lsz * sum_local);
  int64_t Lc = std::max<int64_t>(1, std::min<int64_t>(n_chol, (int64_t)(budget / per_J)));
  for (int64_t q = Lc; q >= std::max<int64_t>(1, Lc / 2); --q)
    if (n_chol % q == 0) { Lc = q; break; }
  std::vector<int64_t> lens;
  for (int64_t J0 = 0; J0 < n_chol; J0 += Lc) {
    const int64_t len = std::min(Lc, n_chol - J0);
    int idx = -1;
    for (size_t q = 0; q < lens.size(); ++q)
      if (lens[q] == len) idx = (int)q;
    if (idx < 0) {
      idx = (int)lens.size();
      lens.push_back(len);
    }
    lchunks.push_back({J0, idx});
  }
  if (ladder_merged) Cw.alloc((size_t)sum_d2 * 8);
  Tbuf.alloc((size_t)Lc * sum_d2 * 8);
  if (sum_local && !lp_all) Lp.alloc((size_t)Lc * sum_local * lsz);
  // programs for a K-split fill target: phase 2 is split along K until it has
  // about `target` workgroups (at most 16 splits)
  // tile2: phase-2 tile of the per-walk programs (64 or 128)
  // tile1: phase-1 tile of the per-walk programs for domains above 64 (64 or
  //        128; 192 puts the domains of 129-192 on 192 x 192 tiles, 65-128 on 128)
  // big: both phases of local domains of 129-192 on 192 x 192 tiles (one
  //      workgroup per CU, so their K-split target is a quarter)
  auto is_big = [&](int64_t d) { return d > 128 && d <= 192 && d != n_v; };
  auto build = [&](int64_t target, int tile2, int tile1, bool big) {
    int64_t tiles2 = 0, tiles_big = 0;
    for (int P = 0; P < n_pairs; ++P) {
      const int64_t d = h_dom_off[P + 1] - h_dom_off[P];
      if (big && is_big(d)) ++tiles_big;
      else tiles2 += cdiv(d, tile2) * cdiv(d, tile2);
    }
    const int nsp_small = tiles2 ? (int)std::max<int64_t>(1, std::min<int64_t>(16, cdiv(target, tiles2))) : 1;
    const int nsp_big =
        tiles_big ? (int)std::max<int64_t>(1, std::min<int64_t>(16, cdiv(std::max<int64_t>(target / 4, 1), tiles_big)))
                  : 1;
    rpart_n = std::max(nsp_small, nsp_big);
    if (ladder_merged) {
      const int64_t tm = cdiv((int64_t)n_pairs * n_v, 128) * cdiv(n_v, 128);
      rpart_n = (int)std::max<int64_t>(1, std::min<int64_t>(16, cdiv(target, tm)));
    }
    if (rpart_n > 1 && Rpart.bytes < (size_t)rpart_n * sum_d2 * 8) Rpart.alloc((size_t)rpart_n * sum_d2 * 8);
    lphase.clear();
    const int64_t nv2 = (int64_t)n_v * n_v;
    for (int64_t len : lens) {
      lphase.emplace_back();
      LadderPhase& ph = lphase.back();
      ph.Lc = len;
      ph.nsplit = (int)std::min<int64_t>(rpart_n, len);
      std::vector<int64_t> lpo(n_local + 1, 0);
      for (int q = 0; q < n_local; ++q) {
        const int64_t d = h_dom_off[loc[q] + 1] - h_dom_off[loc[q]];
        lpo[q + 1] = lpo[q] + len * d * d;
      }
      upload(ph.lp_off, lpo, st);
      ph.p1.clear();
      ph.p2.clear();
      ph.p1.dense = ph.p2.dense = true;
      ph.p1.tile = 128;
      ph.p2.tile = 64;
      if (ladder_merged) {
        // phase 1: T[(jj,a)][(P,c)] = sum_k L^J[a][k] Cw[k][(P,c)], stored as T[jj][P][a][c]
        // phase 2: R[(P,a)][n] += sum_jj sum_c T[jj][(P,a)][c] L^J[c][n]  (R blocks back to back)
        const int d = n_v, NP = n_pairs * n_v;
        const int64_t Sjj = (int64_t)n_pairs * d * d;
        Blocking bk;
        bk.rdiv = d;
        bk.rstride = Sjj;
        bk.cdiv = d;
        bk.cstride = (int64_t)d * d;
        ph.p1.add(Tbuf.as<double>(), d, (int)(len * d), NP,
                  {seg(op(Lvv.p, n_v, 1, nullptr, nullptr, lvv_f32, nv2), op(Cw.p, NP, 1), d, 1.0)}, false, false, 0, bk);
        ph.p2.tile = 128;
        for (int sp = 0; sp < ph.nsplit; ++sp) {
          const int64_t j0 = len * sp / ph.nsplit, j1 = len * (sp + 1) / ph.nsplit;
          std::vector<Seg> sg;
          for (int64_t jj = j0; jj < j1; ++jj)
            sg.push_back(seg(op(Tbuf.as<double>() + jj * Sjj, d, 1),
                             op(static_cast<const char*>(Lvv.p) + (size_t)(jj * nv2) * lsz, n_v, 1, nullptr, nullptr,
                                lvv_f32, nv2),
                             d, 1.0));
          double* Rout = rpart_n > 1 ? Rpart.as<double>() + (int64_t)sp * sum_d2 : R.as<double>();
          ph.p2.add(Rout, d, NP, d, sg, true);
        }
        ph.p1.upload_tables(st);
        ph.p2.upload_tables(st);
        continue;
      }
      // per-walk programs: phase 1 (M = len d) on 64- or 128-wide tiles by domain
      // size; phase 2 on 64-wide tiles (4 workgroups per CU measure faster than
      // 128-wide tiles for the long-K d x d outputs)
      ph.p1x.clear();
      ph.p2x.clear();
      ph.p1x.resize(3);
      ph.p2x.resize(2);
      for (int v = 0; v < 3; ++v) {
        ph.p1x[v].dense = true;
        ph.p1x[v].tile = v == 2 ? 192 : (v ? (tile1 == 192 ? 128 : tile1) : 64);
      }
      ph.p2x[0].dense = ph.p2x[1].dense = true;
      ph.p2x[0].tile = tile2;
      ph.p2x[1].tile = 192;
      int64_t toff = 0;
      int qloc = 0;
      for (int P = 0; P < n_pairs; ++P) {
        const int d = (int)(h_dom_off[P + 1] - h_dom_off[P]);
        const bool full = d == n_v;
        const bool b192 = big && is_big(d);
        // tile1 == 192: phase 1 of the domains of 129-192 on 192 x 192 tiles, phase 2 on tile2
        const int v = (b192 || (tile1 == 192 && is_big(d))) ? 2 : (d <= 64 ? 0 : 1);
        const int nsp = (int)std::min<int64_t>(b192 ? nsp_big : nsp_small, len);
        double* T = Tbuf.as<double>() + toff;
        // L operand of this pair: rows (jj, a) with leading dimension ld, one
        // jj-block every jstep elements; full domains read L^J itself (J0 via jstride)
        const void* Lb;
        int64_t ld, jstep, jstride;
        if (full) {
          Lb = Lvv.p;
          ld = n_v;
          jstep = nv2;
          jstride = nv2;
        } else if (lp_all) {  // [q][J][r][c]: chunk start through jstride
          Lb = static_cast<const char*>(Lp.p) + (size_t)lpo_all[qloc++] * lsz;
          ld = d;
          jstep = (int64_t)d * d;
          jstride = jstep;
        } else {
          Lb = static_cast<const char*>(Lp.p) + (size_t)lpo[qloc++] * lsz;
          ld = d;
          jstep = (int64_t)d * d;
          jstride = 0;
        }
        // phase 1: rows (jj, a) of L (M = len*d); rows jj*ld.. are contiguous when
        // ld == d, and for full domains ld == n_v == d as well
        ph.p1x[v].add(T, d, (int)(len * d), d, {seg(op(Lb, ld, 1, nullptr, nullptr, lvv_f32, jstride),
                                                    op(C.as<double>() + h_cd_off[P], d, 1), d, 1.0)},
                      false);
        for (int sp = 0; sp < nsp; ++sp) {
          const int64_t j0 = len * sp / nsp, j1 = len * (sp + 1) / nsp;
          std::vector<Seg> sg;
          for (int64_t jj = j0; jj < j1; ++jj) {
            const void* Lj = static_cast<const char*>(Lb) + (size_t)(jj * jstep) * lsz;
            sg.push_back(seg(op(T + jj * d * d, d, 1), op(Lj, ld, 1, nullptr, nullptr, lvv_f32, jstride), d, 1.0));
          }
          double* Rout = (rpart_n > 1 ? Rpart.as<double>() + (int64_t)sp * sum_d2 : R.as<double>()) + h_cd_off[P];
          ph.p2x[b192 ? 1 : 0].add(Rout, d, d, d, sg, true);
        }
        toff += len * (int64_t)d * d;
      }
      for (auto& g : ph.p1x) g.upload_tables(st);
      for (auto& g : ph.p2x) g.upload_tables(st);
    }
  };
  // 64-wide phase-2 tiles run 4 workgroups per CU, 128-wide ones 2
  const int64_t def_target = ladder_merged ? 512 : 2048;
  if (!autotune_enabled()) {
    // untuned: no 192 x 192 tiles unless TCI_FP64_P192=1 (both phases of
    // 129-192) or TCI_FP64_P1_192=1 (their phase 1)
    const char* e = std::getenv("TCI_FP64_P192");
    const char* e1 = std::getenv("TCI_FP64_P1_192");
    build(def_target, 64, e1 && e1[0] == '1' ? 192 : 128, e && e[0] == '1');
    return;
  }
  // autotuning on the first chunk (C holds no vector yet; the timing does
  // not depend on the values) over fill targets and, for per-walk programs,
  // the phase-2 tile and the 192 x 192 tiles; then the phase-1 tile and each
  // program's workgroup order
  bool any_big = false;
  for (int P = 0; P < n_pairs; ++P) any_big = any_big || is_big(h_dom_off[P + 1] - h_dom_off[P]);
  double best = 1e300;
  int64_t bt = def_target;
  int btile = 64;
  bool bbig = false;
  std::string log = "fp64 ladder: (K-split target, phase-2 tile, 192-tiles)";
  for (int big = 0; big <= (any_big && !ladder_merged ? 1 : 0); ++big)
    for (int tile2 : {64, 128}) {
      if ((ladder_merged || big) && tile2 == 128) break;  // merged programs have their own tiles
      const int64_t base = tile2 == 64 ? def_target : def_target / 4;
      for (int64_t t : {base / 2, base, 2 * base, 4 * base}) {
        build(t, tile2, 128, big != 0);
        TCI_HIP(hipStreamSynchronize(st));
        const double ms = time_ms(st, 2, [&] { ladder_chunk(lchunks[0], st); });
        log += " " + std::to_string(t) + "/" + std::to_string(tile2) + (big ? "b" : "") + ":" + fmt_ms(ms);
        if (ms < best * 0.98) {
          best = ms;
          bt = t;
          btile = tile2;
          bbig = big != 0;
        }
      }
    }
  int btile1 = 128;
  if (!ladder_merged && n_local > 0) {  // then the phase-1 tile of domains above 64
    build(bt, btile, 64, bbig);
    TCI_HIP(hipStreamSynchronize(st));
    const double ms = time_ms(st, 2, [&] { ladder_chunk(lchunks[0], st); });
    log += "; phase-1 tile 64:" + fmt_ms(ms);
    if (ms < best * 0.98) {
      best = ms;
      btile1 = 64;
    }
    if (any_big && !bbig) {  // phase 1 of 129-192 alone on 192 x 192 tiles
      build(bt, btile, 192, bbig);
      TCI_HIP(hipStreamSynchronize(st));
      const double m2 = time_ms(st, 2, [&] { ladder_chunk(lchunks[0], st); });
      log += " 192:" + fmt_ms(m2);
      if (m2 < best * 0.98) {
        best = m2;
        btile1 = 192;
      }
    }
  }
  build(bt, btile, btile1, bbig);
  TCI_HIP(hipStreamSynchronize(st));
  log += " -> " + std::to_string(bt) + "/" + std::to_string(btile) + (bbig ? "b" : "") + "/p1 " +
         std::to_string(btile1) + "; order";
  auto progs = [&](LadderPhase& ph) {
    std::vector<GemmProgram*> v;
    if (ladder_merged) {
      v.push_back(&ph.p1);
      v.push_back(&ph.p2);
    }
    for (auto& g : ph.p1x) v.push_back(&g);
    for (auto& g : ph.p2x) v.push_back(&g);
    return v;
  };
  const int64_t J0 = lchunks[0].first;
  std::vector<GemmProgram*> g0 = progs(lphase[lchunks[0].second]);
  for (size_t k = 0; k < g0.size(); ++k) {
    if (g0[k]->tasks.empty()) continue;
    const int o = tune_order(*g0[k], st, [&] { g0[k]->launch(st, J0); });
    log += " " + std::to_string(o);
    for (auto& ph : lphase) progs(ph)[k]->remap = o;
  }
  tuning += log + "\n";
}

// One chunk of Cholesky vectors of the FP64 two-phase ladder: pack L^J[D,D]
// of the local-domain walks, then phase 1 and phase 2.
void tci_plan::ladder_chunk(const std::pair<int64_t, int>& ch, hipStream_t st) {
  const LadderPhase& ph = lphase[ch.second];
  if (n_local && !lp_all) {
    PackTab K{pairs_local.as<int32_t>(), ph.lp_off.as<int64_t>(), n_local};
    const int e0 = kt_begin(st);
    launch_pack_L(tab, K, Lvv.p, lvv_f32, ch.first, ph.Lc, Lp.p, st);
    kt_end(e0, "fp64 ladder pack_L_kernel", 0.0, TCI_SIGMA_FP64, st);
  }
  auto run = [&](const char* role, const GemmProgram& g) {
    kt_run(role, g, TCI_SIGMA_FP64, st, [&] { g.launch(st, ch.first); });
  };
  if (ladder_merged) {
    run("fp64 ladder phase 1", ph.p1);
    run("fp64 ladder phase 2", ph.p2);
  } else {
    for (const auto& g : ph.p1x) run("fp64 ladder phase 1", g);
    for (const auto& g : ph.p2x) run("fp64 ladder phase 2", g);
  }
}

// The FP64 two-phase ladder of one sigma: R (zeroed here) += sum over the
// chunks of Cholesky vectors, then the K-split partials reduced.
void tci_plan::two_phase_ladder(hipStream_t st) {
  double* Rz = rpart_n > 1 ? Rpart.as<double>() : R.as<double>();
  launch_fill(Rz, (int64_t)(rpart_n > 1 ? rpart_n : 1) * sum_d2, 0.0, st);
  if (ladder_merged) launch_concat_cols(C.as<double>(), n_pairs, n_v, Cw.as<double>(), st);
  for (const auto& ch : lchunks) ladder_chunk(ch, st);
  if (timing()) TCI_HIP(hipEventRecord(ev[3], st));
  if (rpart_n > 1) ladder_reduce(Rpart.as<double>(), sum_d2, rpart_n, R.as<double>(), sum_d2, st);
}

void tci_plan::unpack_and_ladder(const double* x, const IoSlot* io, int precision, hipStream_t st) {
  launch_unpack(tab, dmax, x, C.as<double>(), Ct.as<double>(), s.as<double>(), st, io);
  if (timing()) TCI_HIP(hipEventRecord(ev[2], st));
  if (precision != TCI_SIGMA_FP64) {
    mixed_ladder(*this, precision, st);
    return;
  }
  if (vv.fp64) {
    vv_ladder(*this, TCI_SIGMA_FP64, st);
    if (timing()) TCI_HIP(hipEventRecord(ev[3], st));
    return;
  }
  if (!ladder_fused) {
    two_phase_ladder(st);
    return;
  }
  LadderArgs a;
  a.dom_off = dom_off.as<int64_t>();
  a.dom_idx = dom_idx.as<int32_t>();
  a.cd_off = cd_off.as<int64_t>();
  a.Cd = C.as<double>();
  a.Lvv = Lvv.p;
  a.lvv_f32 = lvv_f32;
  a.n_v = n_v;
  a.partial = lad_part.as<double>();
  a.chunk_stride = sum_d2;
  const int e0 = kt_begin(st);
  ladder.launch(a, st);
  kt_end(e0, "fp64 fused ladder", flops_ladder, TCI_SIGMA_FP64, st);
  if (timing()) TCI_HIP(hipEventRecord(ev[3], st));
  ladder_reduce(lad_part.as<double>(), sum_d2, lad_chunks, R.as<double>(), sum_d2, st);
}

void tci_plan::sigma_body(const double* x, double* y, const IoSlot* io, int flags, hipStream_t st) {
  unpack_and_ladder(x, io, flags & 0xff, st);
  sigma_rest(x, y, io, flags, st);
}

// every term after the ladder: needs C, Ct and s of x (unpacked) and the
// ladder's R
void tci_plan::sigma_rest(const double* x, double* y, const IoSlot* io, int flags, hipStream_t st) {
  const int precision = flags & 0xff;
  int e0 = kt_begin(st);
  launch_rterms(tab, dmax, x, s.as<double>(), Kov.as<double>(), O3.as<double>(), R.as<double>(), st, io);
  kt_end(e0, "rterms_kernel", 0.0, precision, st);
  launch_transpose_small(s.as<double>(), n_int, n_v, sT.as<double>(), st);
  kt_run("singles Y", progY, precision, st, [&] { progY.launch(st); });
  if (dense_pairs) {
    dense_pairs_terms(*this, precision, st);
  } else {
    HoleTab H{isect_off.as<int64_t>(), isect.as<int32_t>()};
    e0 = kt_begin(st);
    launch_hole(tab, H, dmax, C.as<double>(), O4.as<double>(), R.as<double>(), st);
    kt_end(e0, "hole_kernel", 0.0, precision, st);
    kt_run("ring and 3-external", ring, precision, st, [&] { ring.launch(st); });
    kt_run("singles W", progW, precision, st, [&] { progW.launch(st); });
  }
  launch_gemv(Mss.as<double>(), (int64_t)n_int * n_v, s.as<double>(), gss.as<double>(), st);
  progGW.launch(st);
  launch_s3int_singles(tab, dmax, gpart_off.as<int64_t>(), Ct.as<double>(), O3.as<double>(), gpart.as<double>(), st);
  InvTab I{inv_off.as<int64_t>(), inv_q.as<int32_t>(), inv_pos.as<int32_t>()};
  launch_singles_reduce(tab, I, gpart_off.as<int64_t>(), gpart.as<double>(), gss.as<double>(), gW.as<double>(), n_split,
                        g.as<double>(), st);
  launch_ref_dot(tab, Kov.as<double>(), Ct.as<double>(), part.as<double>(), sig0.as<double>(), st);
  if (has_fvv) kt_run("virtual Fock", fockvv, precision, st, [&] { fockvv.launch(st); });
  if (has_fov) launch_fov_terms(x, s.as<double>(), Fov.as<double>(), (int64_t)n_int * n_v, g.as<double>(),
                                sig0.as<double>(), st, io);
  launch_pack(tab, dmax, x, R.as<double>(), g.as<double>(), sig0.as<double>(), e_ref, flags, y, st, io);
}

// The whole build as one graph launch: the first call with a flag set captures
// the launch sequence of sigma_body on a private stream (x and y read from the
// io slot), later calls set the slot and replay.  Split precisions prepare
// their operand copies eagerly first: a precision switch rebuilds those copies
// and drops every captured graph.  Every buffer a graph touches belongs to the
// plan and keeps its address for the plan's lifetime otherwise.
void tci_plan::sigma_graph(const double* x, double* y, int flags, hipStream_t st) {
  const int precision = flags & 0xff;
  if (precision != TCI_SIGMA_FP64) mixed_prepare(*this, precision, st);
  auto it = graphs.find(flags);
  if (it == graphs.end()) {
    if (!io.p) io.alloc(sizeof(IoSlot));
    if (!gstream) TCI_HIP(hipStreamCreateWithFlags(&gstream, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    TCI_HIP(hipStreamBeginCapture(gstream, hipStreamCaptureModeThreadLocal));
    capturing = true;
    try {
      sigma_body(nullptr, nullptr, io.as<IoSlot>(), flags, gstream);
    } catch (...) {
      capturing = false;
      (void)hipStreamEndCapture(gstream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    capturing = false;
    TCI_HIP(hipStreamEndCapture(gstream, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TCI_HIP(e);
    it = graphs.emplace(flags, ex).first;
  }
  if (profiling) TCI_HIP(hipEventRecord(ev[0], st));
  launch_set_io(io.as<IoSlot>(), x, y, st);
  TCI_HIP(hipGraphLaunch(it->second, st));
  if (profiling) {
    TCI_HIP(hipEventRecord(ev[1], st));
    TCI_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    TCI_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    prof_sigma_ms[precision] += ms;
    ++prof_sigma_n[precision];
  }
}

void tci_plan::sigma_batch(const double* const* x, double* const* y, int nvec, int flags, hipStream_t st) {
  const int precision = flags & 0xff;
  TCI_REQUIRE(precision == TCI_SIGMA_FP64 || precision == TCI_SIGMA_BF16X3 || precision == TCI_SIGMA_FP16X3,
              "unknown precision");
  TCI_REQUIRE((flags & ~(0xff | TCI_ACCUMULATE | TCI_ADD_EREF | TCI_SIGMA_GRAPH)) == 0, "unknown sigma flag bits");
  bool batched = nvec > 1 && (precision == TCI_SIGMA_BF16X3 || precision == TCI_SIGMA_FP16X3);
  if (batched) {
    mixed_prepare(*this, precision, st);
    batched = mixed_batch_supported(*this);
  }
  if (!batched) {
    for (int v = 0; v < nvec; ++v) sigma(x[v], y[v], flags, st);
    return;
  }
  const int fl = flags & ~TCI_SIGMA_GRAPH;
  mixed_batch_prepare(*this, precision, nvec, st);
  if (profiling) TCI_HIP(hipEventRecord(ev[0], st));
  if (precision == TCI_SIGMA_FP16X3) {  // one scale set for the whole batch
    for (int v = 0; v < nvec; ++v) {
      launch_unpack(tab, dmax, x[v], C.as<double>(), Ct.as<double>(), s.as<double>(), st);
      mixed_batch_absmax(*this, v == 0, st);
    }
    mixed_batch_set_scales(*this, st);
  }
  for (int v = 0; v < nvec; ++v) {
    launch_unpack(tab, dmax, x[v], C.as<double>(), Ct.as<double>(), s.as<double>(), st);
    mixed_batch_split_c(*this, precision, v, st);
  }
  if (profiling) TCI_HIP(hipEventRecord(ev[2], st));
  mixed_batch_ladder(*this, precision, st);
  for (int v = 0; v < nvec; ++v) {
    launch_unpack(tab, dmax, x[v], C.as<double>(), Ct.as<double>(), s.as<double>(), st);
    TCI_HIP(hipMemcpyAsync(R.p, mx.bat.Rb.as<double>() + (int64_t)v * sum_d2, (size_t)sum_d2 * 8,
                           hipMemcpyDeviceToDevice, st));
    sigma_rest(x[v], y[v], nullptr, fl, st);
  }
  if (profiling) {
    TCI_HIP(hipEventRecord(ev[1], st));
    TCI_HIP(hipEventSynchronize(ev[1]));
    kt_collect();
    float ms = 0.f;
    TCI_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    prof_sigma_ms[precision] += ms;
    prof_sigma_n[precision] += nvec;
    TCI_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
    prof_ladder_ms[precision] += ms;
    prof_ladder_n[precision] += nvec;
  }
}

void tci_plan::sigma(const double* x, double* y, int flags, hipStream_t st) {
  const int precision = flags & 0xff;
  TCI_REQUIRE(precision == TCI_SIGMA_FP64 || precision == TCI_SIGMA_BF16X3 || precision == TCI_SIGMA_FP16X3,
              "unknown precision");
  TCI_REQUIRE((flags & ~(0xff | TCI_ACCUMULATE | TCI_ADD_EREF | TCI_SIGMA_GRAPH)) == 0, "unknown sigma flag bits");
  if (flags & TCI_SIGMA_GRAPH) {
    sigma_graph(x, y, flags & ~TCI_SIGMA_GRAPH, st);
    return;
  }
  if (profiling) TCI_HIP(hipEventRecord(ev[0], st));
  sigma_body(x, y, nullptr, flags, st);
  if (profiling) {
    TCI_HIP(hipEventRecord(ev[1], st));
    TCI_HIP(hipEventSynchronize(ev[1]));
    kt_collect();
    float ms = 0.f;
    TCI_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    prof_sigma_ms[precision] += ms;
    ++prof_sigma_n[precision];
    TCI_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
    prof_ladder_ms[precision] += ms;
    ++prof_ladder_n[precision];
  }
}

// ============================================================================
// C ABI
// ============================================================================
#define TCI_GUARD_BEGIN try {
#define TCI_GUARD_END                                   \
  }                                                     \
  catch (const tci::Error& e) {                         \
    tci::set_last_error(e.what());                      \
    return e.code;                                      \
  }                                                     \
  catch (const std::exception& e) {                     \
    tci::set_last_error(e.what());                      \
    return TCI_ERR_ARG;                                 \
  }                                                     \
  catch (...) {                                         \
    tci::set_last_error("unknown error");               \
    return TCI_ERR_ARG;                                 \
  }

extern "C" {

const char* tci_last_error(void) { return tci::g_last_error.c_str(); }

int tci_plan_create(const tci_problem* prob, tci_plan** out) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(prob && out, "null argument");
  *out = nullptr;
  tci_plan* p = new tci_plan();
  try {
    p->build(*prob);
  } catch (...) {
    delete p;
    throw;
  }
  *out = p;
  return TCI_OK;
  TCI_GUARD_END
}

void tci_plan_destroy(tci_plan* plan) {
  if (plan) {
    (void)hipDeviceSynchronize();
    delete plan;
  }
}

int tci_sigma_apply_batch(tci_plan* plan, const double* c, double* sigma, int64_t n, int32_t nvec, int32_t flags,
                          void* stream) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(plan && c && sigma, "null argument");
  TCI_REQUIRE(n == plan->n_csf, "vector length does not match the plan");
  TCI_REQUIRE(nvec >= 1 && nvec <= 64, "nvec must be 1..64");
  TCI_REQUIRE(c + (int64_t)nvec * n <= sigma || sigma + (int64_t)nvec * n <= c, "c and sigma must not overlap");
  std::vector<const double*> xs(nvec);
  std::vector<double*> ys(nvec);
  for (int v = 0; v < nvec; ++v) {
    xs[v] = c + (int64_t)v * n;
    ys[v] = sigma + (int64_t)v * n;
  }
  plan->sigma_batch(xs.data(), ys.data(), nvec, flags, static_cast<hipStream_t>(stream));
  return TCI_OK;
  TCI_GUARD_END
}

int tci_plan_tuning(const tci_plan* plan, char* buf, int64_t len) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(plan && buf && len > 0, "null argument");
  const std::string& t = plan->tuning;
  const size_t n = std::min<size_t>(t.size(), (size_t)len - 1);
  std::memcpy(buf, t.data(), n);
  buf[n] = 0;
  return TCI_OK;
  TCI_GUARD_END
}

int tci_plan_get_info(const tci_plan* plan, tci_plan_info* info) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(plan && info, "null argument");
  info->n_csf = plan->n_csf;
  info->n_ladder_items = plan->ladder.n_items;
  info->e_ref = plan->e_ref;
  info->flops_ladder = plan->flops_ladder;
  info->flops_sigma = plan->flops_sigma;
  info->bytes_ladder = plan->bytes_ladder;
  info->device_bytes = plan->device_bytes;
  return TCI_OK;
  TCI_GUARD_END
}

int tci_sigma_apply(tci_plan* plan, const double* c, double* sigma, int64_t n, int32_t flags, void* stream) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(plan && c && sigma, "null argument");
  TCI_REQUIRE(n == plan->n_csf, "vector length does not match the plan's CSF count");
  TCI_REQUIRE(c + n <= sigma || sigma + n <= c, "c and sigma must not overlap");
  plan->sigma(c, sigma, flags, static_cast<hipStream_t>(stream));
  return TCI_OK;
  TCI_GUARD_END
}

int tci_htimesc(tci_plan* plan, const double* civec, double* sigmavec, int64_t n) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(plan && civec && sigmavec, "null argument");
  TCI_REQUIRE(n == plan->n_csf, "vector length does not match the plan's CSF count");
  DevBuf dc, ds;
  dc.alloc((size_t)n * 8);
  ds.alloc((size_t)n * 8);
  TCI_HIP(hipMemcpy(dc.p, civec, (size_t)n * 8, hipMemcpyHostToDevice));
  TCI_HIP(hipMemcpy(ds.p, sigmavec, (size_t)n * 8, hipMemcpyHostToDevice));
  plan->sigma(dc.as<double>(), ds.as<double>(), TCI_SIGMA_FP64 | TCI_ACCUMULATE | TCI_ADD_EREF, nullptr);
  TCI_HIP(hipMemcpy(sigmavec, ds.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return TCI_OK;
  TCI_GUARD_END
}

int tci_diagonal(tci_plan* plan, double* diag, int64_t n, void* stream) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(plan && diag, "null argument");
  TCI_REQUIRE(n == plan->n_csf, "vector length does not match the plan's CSF count");
  TCI_HIP(hipMemcpyAsync(diag, plan->diag.p, (size_t)n * 8, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
  return TCI_OK;
  TCI_GUARD_END
}

int tci_profile_enable(tci_plan* plan, int32_t on) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(plan, "null argument");
  for (hipEvent_t& e : plan->ev)
    if (!e) TCI_HIP(hipEventCreate(&e));
  plan->profiling = on != 0;
  for (int q = 0; q < 3; ++q) {
    plan->prof_ladder_ms[q] = plan->prof_sigma_ms[q] = 0.0;
    plan->prof_ladder_n[q] = plan->prof_sigma_n[q] = 0;
    plan->kstats[q].clear();
  }
  plan->kpend.clear();
  plan->kev_used = 0;
  return TCI_OK;
  TCI_GUARD_END
}

int tci_profile_read(const tci_plan* plan, int32_t precision, double* ladder_ms, int64_t* ladder_launches,
                     double* sigma_ms, int64_t* sigma_calls) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(plan, "null argument");
  TCI_REQUIRE(precision >= 0 && precision < 3, "unknown precision");
  if (ladder_ms) *ladder_ms = plan->prof_ladder_ms[precision];
  if (ladder_launches) *ladder_launches = plan->prof_ladder_n[precision];
  if (sigma_ms) *sigma_ms = plan->prof_sigma_ms[precision];
  if (sigma_calls) *sigma_calls = plan->prof_sigma_n[precision];
  return TCI_OK;
  TCI_GUARD_END
}

int tci_profile_kernels(const tci_plan* plan, int32_t precision, char* buf, int64_t len) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(plan && buf && len > 0, "null argument");
  TCI_REQUIRE(precision >= 0 && precision < 3, "unknown precision");
  std::string out;
  char line[512];
  for (const auto& kv : plan->kstats[precision]) {
    std::snprintf(line, sizeof line, "%s|%lld|%.6f|%.6e\n", kv.first.c_str(), (long long)kv.second.n, kv.second.ms,
                  kv.second.n ? kv.second.flops / (double)kv.second.n : 0.0);
    out += line;
  }
  const size_t m = std::min<size_t>(out.size(), (size_t)len - 1);
  std::memcpy(buf, out.data(), m);
  buf[m] = 0;
  return TCI_OK;
  TCI_GUARD_END
}

int tci_device_alloc(int64_t bytes, void** ptr) {
  TCI_GUARD_BEGIN
  TCI_REQUIRE(ptr && bytes >= 0, "bad argument");
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 8));
  if (e != hipSuccess) throw tci::Error(TCI_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return TCI_OK;
  TCI_GUARD_END
}

int tci_device_free(void* ptr) {
  TCI_GUARD_BEGIN
  if (ptr) TCI_HIP(hipFree(ptr));
  return TCI_OK;
  TCI_GUARD_END
}

int tci_memcpy_h2d(void* dst, const void* src, int64_t bytes) {
  TCI_GUARD_BEGIN
  TCI_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return TCI_OK;
  TCI_GUARD_END
}

int tci_memcpy_d2h(void* dst, const void* src, int64_t bytes) {
  TCI_GUARD_BEGIN
  TCI_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return TCI_OK;
  TCI_GUARD_END
}

int tci_memcpy_d2d(void* dst, const void* src, int64_t bytes) {
  TCI_GUARD_BEGIN
  TCI_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice));
  return TCI_OK;
  TCI_GUARD_END
}

int tci_device_synchronize(void) {
  TCI_GUARD_BEGIN
  TCI_HIP(hipDeviceSynchronize());
  return TCI_OK;
  TCI_GUARD_END
}

}  // extern "C"
