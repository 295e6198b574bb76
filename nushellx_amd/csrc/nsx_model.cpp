#include "nsx_model.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace nsx {

// ---------------------------------------------------------------- recoupling
// 6j symbol from the Racah single-sum formula (doubled arguments), evaluated in
// long double through log-factorials.
static long double lfac(int n) { return lgammal((long double)n + 1.0L); }
static bool tri(int a, int b, int c) { return ((a + b + c) & 1) == 0 && std::abs(a - b) <= c && c <= a + b; }
static long double ldelta(int a, int b, int c) {
  return 0.5L * (lfac((a + b - c) / 2) + lfac((a - b + c) / 2) + lfac((-a + b + c) / 2) -
                 lfac((a + b + c) / 2 + 1));
}

double sixj(int a, int b, int c, int d, int e, int f) {
  if (!tri(a, b, c) || !tri(a, e, f) || !tri(d, b, f) || !tri(d, e, c)) return 0.0;
  const int t1 = (a + b + c) / 2, t2 = (a + e + f) / 2, t3 = (d + b + f) / 2, t4 = (d + e + c) / 2;
  const int p1 = (a + b + d + e) / 2, p2 = (b + c + e + f) / 2, p3 = (c + a + f + d) / 2;
  const long double pre = ldelta(a, b, c) + ldelta(a, e, f) + ldelta(d, b, f) + ldelta(d, e, c);
  long double s = 0.0L;
  const int tmin = std::max(std::max(t1, t2), std::max(t3, t4));
  const int tmax = std::min(p1, std::min(p2, p3));
  for (int t = tmin; t <= tmax; ++t) {
    const long double lg = pre + lfac(t + 1) - lfac(t - t1) - lfac(t - t2) - lfac(t - t3) -
                           lfac(t - t4) - lfac(p1 - t) - lfac(p2 - t) - lfac(p3 - t);
    s += ((t & 1) ? -1.0L : 1.0L) * expl(lg);
  }
  return (double)s;
}

// W(j1 j2 j3 j4; J K) = (-1)^{(j1+j2+j3+j4)/2} {j1 j2 J; j4 j3 K}  (src/NuCommon/NuCommon.f90:3255-3265)
double racah_w(int j1, int j2, int j3, int j4, int J, int K) {
  const double ph = (((j1 + j2 + j3 + j4) / 2) & 1) ? -1.0 : 1.0;
  return ph * sixj(j1, j2, J, j4, j3, K);
}

int tfind_ref(const std::vector<std::array<uint8_t, 8>>& keys, const uint8_t* key) {
  const int n = (int)keys.size();
  if (n == 0) return 0;
  if (std::memcmp(keys[n - 1].data(), key, 8) == 0) return n;
  float rn = (float)n / 2.0f, drn = rn;
  for (;;) {
    const int in = (int)lroundf(rn);
    if (in < 1 || in > n) return 0;
    const int c = std::memcmp(key, keys[in - 1].data(), 8);
    if (c == 0) return in;
    drn = drn / 2.0f;
    if (std::fabs(drn) < 0.5f) return 0;
    rn = c > 0 ? rn + drn : rn - drn;
  }
}

// ---------------------------------------------------------------- helpers
static int find_key(const std::vector<Density>& d, const uint8_t* key) {
  int lo = 0, hi = (int)d.size() - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) / 2;
    const int c = std::memcmp(d[mid].key, key, 8);
    if (c == 0) return mid;
    if (c < 0) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

static inline bool bit(const std::vector<uint32_t>& mask, int words, int col, int row) {
  return (mask[size_t(col) * words + (row >> 5)] >> (row & 31)) & 1u;
}

// Dense d(Jl) x d(Jr) matrix of one density: partition blocks in next_ij order
// (column partition outer, row partition inner), each block column-major.
std::vector<double> densify(const SideFiles& s, const Density& dd, int J2l, int J2r) {
  const int dl = s.d(J2l), dr = s.d(J2r);
  std::vector<double> D(size_t(dl) * dr, 0.0);
  int64_t pt = dd.pos;
  for (int j = 0; j < s.nsp; ++j)
    for (int i = 0; i < s.nsp; ++i) {
      if (!bit(dd.mask, s.words, j, i)) continue;
      const int oi = s.offset(J2l, i), oj = s.offset(J2r, j);
      for (int k = 0; k < s.n(J2r, j); ++k)
        for (int l = 0; l < s.n(J2l, i); ++l) D[size_t(oj + k) * dl + (oi + l)] = s.vals[pt++];
    }
  return D;
}

// Union-mask sum of densities: the dense form of NuTRLZ's coefor merge
// (src/NuXpct/NuCoreXE.f90:70-138 is the same routine), which adds the values of
// partition blocks present in both masks and copies the others.
struct Merged {
  std::vector<double> D;         // d(Jl) x d(Jr), column-major
  std::vector<uint32_t> mask;    // union of the contributing masks
  int64_t count = 0;             // stored values in the union layout
};

static void merge_into(Merged& g, const SideFiles& s, const Density& dd, int J2l, int J2r, double scale) {
  const std::vector<double> D = densify(s, dd, J2l, J2r);
  if (g.D.empty()) {
    g.D.assign(D.size(), 0.0);
    g.mask.assign(dd.mask.size(), 0u);
    for (size_t q = 0; q < D.size(); ++q) g.D[q] = scale * D[q];
  } else {
    for (size_t q = 0; q < D.size(); ++q) g.D[q] = g.D[q] + scale * D[q];
  }
  for (size_t w = 0; w < dd.mask.size(); ++w) g.mask[w] |= dd.mask[w];
  g.count = 0;
  for (int j = 0; j < s.nsp; ++j)
    for (int i = 0; i < s.nsp; ++i)
      if (bit(g.mask, s.words, j, i)) g.count += int64_t(s.n(J2r, j)) * s.n(J2l, i);
}

// pn part of NuTRLZ: prd() merges the densities along the .dop eigen-terms
// (src/NuTRLZ/NuCoreXZ.f90:1613-1887) and drdbh() builds the term tables over
// (lam, parity, n) with the Racah factor alone (:1536-1595).
static std::string build_separable(Model& m, const std::string& dopfile, char typea,
                                   const std::map<std::array<int, 2>, int>& bidx) {
  const SideFiles& A = m.A;
  const SideFiles& B = m.B;
  int max_lmo = 0;
  std::vector<DopSection> secs;
  std::string err = read_dop(dopfile, max_lmo, secs);
  if (!err.empty()) return err;
  const int totlms = max_lmo / 2 + 1;
  int max_elm = 0;
  for (auto& sc : secs) max_elm = std::max(max_elm, (int)sc.terms.size());
  // lmel(lam index, n, parity) = 2lam of every eigen-term present (:1778)
  std::vector<int> lmel(size_t(totlms) * (max_elm + 1) * 2, -1);
  auto lmel_at = [&](int t, int n, int pp) -> int& { return lmel[(size_t(t) * (max_elm + 1) + n) * 2 + pp]; };
  for (auto& sc : secs)
    for (int n = 1; n <= (int)sc.terms.size(); ++n) lmel_at(sc.lm / 2, n, (sc.par + 3) / 2 - 1) = sc.lm;

  using Key = std::array<int, 5>;  // (2Jl, 2Jr, n, parity slot, 2lam)
  std::map<Key, Merged> amerg, bmerg;
  for (int pass = 0; pass < 2; ++pass) {
    const SideFiles& S = pass ? B : A;
    auto& out = pass ? bmerg : amerg;
    for (int jl = S.minJ2; jl <= S.maxJ2; jl += S.delJ2)
      for (int jr = S.minJ2; jr <= S.maxJ2; jr += S.delJ2)
        for (auto& sc : secs) {
          const int lm = sc.lm;
          if (!(jr >= std::abs(jl - lm) && jr <= jl + lm)) continue;
          const int pp = (sc.par + 3) / 2 - 1;
          for (int n = 1; n <= (int)sc.terms.size(); ++n) {
            const DopTerm& t = sc.terms[n - 1];
            for (const DopLine& l : t.lines) {
              // proton pair first on a line (:1645-1646)
              const bool a_first = typea == 'p';
              const int s1 = (a_first != (pass == 1)) ? l.s[0] : l.s[2];
              const int s2 = (a_first != (pass == 1)) ? l.s[1] : l.s[3];
              const uint8_t k[8] = {(uint8_t)jl, (uint8_t)jr, (uint8_t)s1, (uint8_t)s2, (uint8_t)lm, 0, 0, 0};
              const int ds = find_key(S.dens, k);
              if (ds < 0) continue;
              // a side: e*renorm*me*trdav (:1678); b side: me*trdbv (:1800)
              const double scale = pass ? l.me : (t.e * m.renorm) * l.me;
              merge_into(out[{jl, jr, n, pp, lm}], S, S.dens[ds], jl, jr, scale);
            }
          }
        }
  }
  for (auto& kv : amerg) m.nnz_a += kv.second.count;
  for (auto& kv : bmerg) m.nnz_b += kv.second.count;

  int max_sj2u = 0;
  for (int s : m.nus.sj2) max_sj2u = std::max(max_sj2u, s);
  std::map<Key, int> adense_of, bhat_of;
  std::vector<uint32_t> lbr(B.words);
  for (int L = 0; L < m.basd; ++L) {
    if (!m.bas_active[L]) continue;  // truncated basis: no output for this block
    const int ajl = m.bas_a[L], bjl = m.bas_b[L];
    for (int ajr = A.minJ2; ajr <= A.maxJ2; ajr += A.delJ2) {
      Item it;
      it.L = L;
      it.ajr = ajr;
      for (int t = 0; t < totlms; ++t)
        for (int pp = 0; pp < 2; ++pp)
          for (int n = 1; n <= max_elm; ++n) {
            const int lm = lmel_at(t, n, pp);
            if (lm < 0) continue;
            if (!(ajr >= std::abs(ajl - lm) && ajr <= ajl + lm)) continue;
            auto fa = amerg.find({ajl, ajr, n, pp, lm});
            if (fa == amerg.end()) continue;
            ATerm at;
            at.adens = -1;
            std::fill(lbr.begin(), lbr.end(), 0u);
            int bjrn = std::max(std::abs(bjl - lm), std::abs(ajr - m.J2));
            const int bjrx = std::min(std::min(bjl + lm, ajr + m.J2), B.maxJ2);
            bjrn += (bjrn - B.minJ2) % B.delJ2;
            for (int bjr = bjrn; bjr <= bjrx; bjr += B.delJ2) {
              // the Racah table holds |W| > 1e-6 for 2lam <= 2 max(2j) and Jb' on the b grid (:1115-1133)
              if (bjr < B.minJ2 || lm > 2 * max_sj2u) continue;
              const double w = racah_w(ajl, bjl, ajr, bjr, m.J2, lm);
              if (!(std::fabs(w) > 1e-6)) continue;
              const double ph = (((ajl + bjr - m.J2 - lm) / 2) & 1) ? -1.0 : 1.0;
              const double r = ph * w * std::sqrt(double(lm + 1));
              auto bi = bidx.find({ajr, bjr});
              if (bi == bidx.end()) return "find error bh tables";
              if (!(bjr >= std::abs(bjl - lm) && bjr <= bjl + lm)) continue;
              const Key bk = {bjl, bjr, n, pp, lm};
              auto fb = bmerg.find(bk);
              if (fb == bmerg.end()) continue;
              int slot;
              auto f = bhat_of.find(bk);
              if (f == bhat_of.end()) {
                slot = (int)m.bhat.size();
                m.bhat.push_back(fb->second.D);
                m.bhat_J.push_back({bjl, bjr});
                m.bhat_key.push_back({bjl, bjr, lm, n, pp});
                bhat_of[bk] = slot;
              } else {
                slot = f->second;
              }
              at.rt.push_back({bi->second, slot, r});
              const Merged& mb = fb->second;
              for (int c = 0; c < B.nsp; ++c)
                for (int w2 = 0; w2 < B.words; ++w2) lbr[w2] |= mb.mask[size_t(c) * B.words + w2];
              int dar = 0;
              const Merged& ma = fa->second;
              for (int c = 0; c < A.nsp; ++c) {
                bool nz = false;
                for (int w2 = 0; w2 < A.words; ++w2) nz |= ma.mask[size_t(c) * A.words + w2] != 0u;
                if (nz) dar += A.n(ajr, c);
              }
              m.flops_ref += 2.0 * dar * mb.count;
            }
            if (at.rt.empty()) continue;
            {
              int dbl = 0;
              for (int i = 0; i < B.nsp; ++i)
                if ((lbr[i >> 5] >> (i & 31)) & 1u) dbl += B.n(bjl, i);
              m.flops_ref += 2.0 * dbl * fa->second.count;
            }
            const Key ak = {ajl, ajr, n, pp, lm};
            auto fd = adense_of.find(ak);
            if (fd == adense_of.end()) {
              at.adens = (int)m.adense.size();
              m.adense.push_back(fa->second.D);
              m.adense_J.push_back({ajl, ajr});
              adense_of[ak] = at.adens;
            } else {
              at.adens = fd->second;
            }
            m.n_aterms++;
            m.n_rterms += (int64_t)at.rt.size();
            m.n_bterms += (int64_t)at.rt.size();
            it.at.push_back(std::move(at));
          }
      if (!it.at.empty()) m.items.push_back(std::move(it));
    }
  }
  return "";
}

std::string build_model(const std::string& nus_prefix, const std::string& opfile_a,
                        const std::string& opfile_b, const std::string& intfile, Model& m,
                        const std::string& dopfile, bool keep_zero_ops) {
  std::string err;
  m.keep_zero_ops = keep_zero_ops;
  if (!(err = read_nus(nus_prefix, m.nus)).empty()) return err;
  m.intfile = intfile;
  m.opfile_a = opfile_a;
  m.opfile_b = opfile_b;
  const NusInput& nus = m.nus;
  if (!(err = read_side(opfile_a, nus, m.A)).empty()) return err;
  if (!(err = read_side(opfile_b, nus, m.B)).empty()) return err;
  // load-time zero limits of the reference (src/NuTRL/NuCoreX.f90:1163,1195;
  // the separable build uses 1e-6 on both sides, src/NuTRLZ/NuCoreXZ.f90:1205,1237)
  m.separable = !dopfile.empty();
  const double lim_a = m.separable ? 1e-6 : 1e-11, lim_b = m.separable ? 1e-6 : 1e-10;
  for (double& v : m.A.vals) if (std::fabs(v) < lim_a) v = 0.0;
  for (double& v : m.B.vals) if (std::fabs(v) < lim_b) v = 0.0;
  char typea = m.A.proton ? 'p' : 'n', typeb = m.B.proton ? 'p' : 'n';
  if (typea == typeb) { typea = 'n'; typeb = 'p'; }
  int sa0, sa1, sb0, sb1;
  if (typea == 'p') { sa0 = 1; sa1 = nus.np_shells; sb0 = nus.np_shells + 1; sb1 = nus.no_shells; }
  else { sa0 = nus.np_shells + 1; sa1 = nus.no_shells; sb0 = 1; sb1 = nus.np_shells; }
  int no_cA, no_cB;
  if (typea == 'p') { no_cB = (nus.no_cN + nus.nmp) / 2; no_cA = (nus.no_cN - nus.nmp) / 2; }
  else { no_cB = (nus.no_cN - nus.nmp) / 2; no_cA = (nus.no_cN + nus.nmp) / 2; }
  m.azero = no_cA == 0;
  m.bzero = no_cB == 0;
  m.J2 = nus.min_cJ2;
  const SideFiles& A = m.A;
  const SideFiles& B = m.B;

  // block basis (src/NuTRL/NuCoreX.f90:932-988)
  std::vector<std::array<int, 2>> blk;
  for (int ajl = A.minJ2; ajl <= A.maxJ2; ajl += A.delJ2) {
    int bjln = std::max(std::abs(ajl - m.J2), B.minJ2);
    const int bjlx = std::min(ajl + m.J2, B.maxJ2);
    bjln += (bjln - B.minJ2) % B.delJ2;
    for (int bjl = bjln; bjl <= bjlx; bjl += B.delJ2) blk.push_back({ajl, bjl});
  }
  std::sort(blk.begin(), blk.end());
  m.basd = (int)blk.size();
  std::map<std::array<int, 2>, int> bidx;
  m.N = 0;
  for (int b = 0; b < m.basd; ++b) {
    m.bas_a.push_back(blk[b][0]);
    m.bas_b.push_back(blk[b][1]);
    m.da.push_back(A.d(blk[b][0]));
    m.db.push_back(B.d(blk[b][1]));
    m.posnv.push_back(m.N);
    m.N += int64_t(m.da.back()) * m.db.back();
    bidx[blk[b]] = b;
  }
  // parity mask, state order b-partition, b-state, a-partition, a-state (:998-1030);
  // with an N-hbar-omega window also the window mask and the active blocks
  // (:990-1037: trunc = lcoef inside [min_MSon, max_MSon]; a block is active
  // when any partition pair of it lies inside, whatever its states' parity)
  m.truncate = nus.min_MSon != 0 || nus.max_MSon != 999;  // :813
  m.lcoef.assign(size_t(m.N), 0);
  m.trunc.assign(m.truncate ? size_t(m.N) : 0, 0);
  m.bas_active.assign(m.basd, 1);
  {
    int64_t k = 0;
    for (int b = 0; b < m.basd; ++b) {
      bool any = false;
      for (int bp = 0; bp < B.nsp; ++bp)
        for (int bs = 0; bs < B.n(m.bas_b[b], bp); ++bs)
          for (int ap = 0; ap < A.nsp; ++ap) {
            const int ntot = A.nq[ap] + B.nq[bp];
            const bool in = ntot >= nus.min_MSon && ntot <= nus.max_MSon;
            any = any || in;
            for (int as = 0; as < A.n(m.bas_a[b], ap); ++as) {
              const uint8_t lc = (nus.tparity == 0 || A.parity[ap] * B.parity[bp] == nus.tparity) ? 1 : 0;
              if (m.truncate) m.trunc[k] = in ? lc : 0;
              m.lcoef[k++] = lc;
            }
          }
      if (m.truncate) m.bas_active[b] = any ? 1 : 0;
    }
  }
  const char xtra = intfile.empty() ? 'x' : intfile.back();
  if (!(err = read_mats(opfile_a, nus, typea == 'n' ? no_cA : -no_cA, xtra, m.A)).empty()) return err;
  if (!(err = read_mats(opfile_b, nus, typeb == 'n' ? no_cB : -no_cB, xtra, m.B)).empty()) return err;
  if (!(err = read_spe(intfile, m.renorm, m.no_level)).empty()) return err;
  if (m.azero || m.bzero) return "";

  std::vector<OphEntry> oph;
  int max_lmo = 0;
  if (!(err = read_oph(intfile, oph, max_lmo)).empty()) return err;
  if (m.separable) return build_separable(m, dopfile, typea, bidx);
  for (const auto& d : A.dens) m.nnz_a += d.count;
  for (const auto& d : B.dens) m.nnz_b += d.count;
  std::map<std::array<uint8_t, 8>, std::pair<double, int>> opn;  // key -> (renorm * me, record index)
  m.oph_me.clear();
  for (auto& e : oph) {
    std::array<uint8_t, 8> k;
    std::memcpy(k.data(), e.key, 8);
    double v = std::fabs(e.me) < 1e-10 ? 0.0 : e.me;
    opn.emplace(k, std::make_pair(m.renorm * v, (int)m.oph_me.size()));
    m.oph_me.push_back(e.me);
  }
  int max_sj2u = 0;
  for (int s : nus.sj2) max_sj2u = std::max(max_sj2u, s);
  int max_lambda = nus.del_cJ2;
  if (max_lambda <= 2 || (max_lambda & 1)) max_lambda = 2 * max_sj2u;
  int del_lm = nus.del_cT2;
  if (del_lm < 2 || (del_lm & 1)) del_lm = 2;
  const int max_lm = std::min(std::min(A.max_lm, max_lmo), std::min(2 * max_sj2u, max_lambda));

  std::map<int, int> adense_of;               // a-density index -> adense slot
  std::map<std::array<int, 5>, int> bhat_of;  // (bjl, bjr, lm, ashl, ashr) -> bhat slot
  std::vector<uint32_t> lbr(B.words);

  // drdbh loop order (:1486-1565), regrouped per (L, Ja')
  for (int L = 0; L < m.basd; ++L) {
    if (!m.bas_active[L]) continue;  // truncated basis: no output for this block (:155-157)
    const int ajl = m.bas_a[L], bjl = m.bas_b[L];
    for (int ajr = A.minJ2; ajr <= A.maxJ2; ajr += A.delJ2) {
      Item it;
      it.L = L;
      it.ajr = ajr;
      for (int ashl = sa0; ashl <= sa1; ++ashl)
        for (int ashr = sa0; ashr <= sa1; ++ashr) {
          const int j1 = nus.sj2[ashl - 1], j2 = nus.sj2[ashr - 1];
          int lmn = std::abs(j1 - j2);
          const int lmx = std::min(j1 + j2, max_lm);
          lmn += lmn % del_lm;
          for (int lm = lmn; lm <= lmx; lm += del_lm) {
            if (!(ajr >= std::abs(ajl - lm) && ajr <= ajl + lm)) continue;
            const uint8_t ta[8] = {(uint8_t)ajl, (uint8_t)ajr, (uint8_t)ashl, (uint8_t)ashr, (uint8_t)lm, 0, 0, 0};
            const int atsc = find_key(A.dens, ta);
            if (atsc < 0) continue;
            ATerm at;
            at.adens = -1;
            int nb_this = 0;
            std::fill(lbr.begin(), lbr.end(), 0u);
            int bjrn = std::max(std::abs(bjl - lm), std::abs(ajr - m.J2));
            const int bjrx = std::min(std::min(bjl + lm, ajr + m.J2), B.maxJ2);
            bjrn += (bjrn - B.minJ2) % B.delJ2;
            for (int bjr = bjrn; bjr <= bjrx; bjr += B.delJ2) {
              const double w = racah_w(ajl, bjl, ajr, bjr, m.J2, lm);
              if (std::fabs(w) == 0.0) continue;
              const double ph = (((ajl + bjr - m.J2 - lm) / 2) & 1) ? -1.0 : 1.0;
              const double r = ph * w * std::sqrt(double(lm + 1));
              auto bi = bidx.find({ajr, bjr});
              if (bi == bidx.end()) return "find error bh tables";
              const int R = bi->second;
              // b-densities of this (Jb, Jb', lam) coupled to (sa, sa')
              std::vector<std::pair<int, double>> bl;
              std::vector<std::pair<int, int>> src;  // (b-density, .oph record)
              for (int bshl = sb0; bshl <= sb1; ++bshl)
                for (int bshr = sb0; bshr <= sb1; ++bshr) {
                  const int k1 = nus.sj2[bshl - 1], k2 = nus.sj2[bshr - 1];
                  if (lm < std::abs(k1 - k2) || lm > std::min(k1 + k2, max_lm)) continue;
                  if (!(bjr >= std::abs(bjl - lm) && bjr <= bjl + lm)) continue;
                  const uint8_t tb[8] = {(uint8_t)bjl, (uint8_t)bjr, (uint8_t)bshl, (uint8_t)bshr, (uint8_t)lm, 0, 0, 0};
                  const int btsc = find_key(B.dens, tb);
                  if (btsc < 0) continue;
                  std::array<uint8_t, 8> op = {(uint8_t)ashl, (uint8_t)ashr, (uint8_t)bshl, (uint8_t)bshr, (uint8_t)lm, 0, 0, 0};
                  auto oi = opn.find(op);
                  if (oi == opn.end() || (oi->second.first == 0.0 && !keep_zero_ops)) continue;
                  bl.push_back({btsc, oi->second.first});
                  src.push_back({btsc, oi->second.second});
                }
              if (bl.empty()) continue;
              const std::array<int, 5> bk = {bjl, bjr, lm, ashl, ashr};
              auto f = bhat_of.find(bk);
              int slot;
              if (f == bhat_of.end()) {
                std::vector<double> acc(size_t(B.d(bjl)) * B.d(bjr), 0.0);
                for (auto& pr : bl) {
                  const std::vector<double> D = densify(B, B.dens[pr.first], bjl, bjr);
                  for (size_t q = 0; q < acc.size(); ++q) acc[q] += pr.second * D[q];
                }
                slot = (int)m.bhat.size();
                m.bhat.push_back(std::move(acc));
                m.bhat_J.push_back({bjl, bjr});
                m.bhat_key.push_back(bk);
                m.bhat_src.push_back(std::move(src));
                bhat_of[bk] = slot;
              } else {
                slot = f->second;
              }
              at.rt.push_back({R, slot, r});
              // algorithmic work (reference loop nest): 2*dar*nnz per b-density
              for (auto& pr : bl) {
                const Density& bd = B.dens[pr.first];
                for (int c = 0; c < B.nsp; ++c)
                  for (int w2 = 0; w2 < B.words; ++w2) lbr[w2] |= bd.mask[size_t(c) * B.words + w2];
                nb_this++;
                int dar = 0;
                const Density& ad = A.dens[atsc];
                for (int c = 0; c < A.nsp; ++c) {
                  bool nz = false;
                  for (int w2 = 0; w2 < A.words; ++w2) nz |= ad.mask[size_t(c) * A.words + w2] != 0u;
                  if (nz) dar += A.n(ajr, c);
                }
                m.flops_ref += 2.0 * dar * bd.count;
              }
            }
            if (at.rt.empty()) continue;
            {
              int dbl = 0;
              for (int i = 0; i < B.nsp; ++i)
                if ((lbr[i >> 5] >> (i & 31)) & 1u) dbl += B.n(bjl, i);
              m.flops_ref += 2.0 * dbl * A.dens[atsc].count;
            }
            auto fa = adense_of.find(atsc);
            if (fa == adense_of.end()) {
              at.adens = (int)m.adense.size();
              m.adense.push_back(densify(A, A.dens[atsc], ajl, ajr));
              m.adense_J.push_back({ajl, ajr});
              adense_of[atsc] = at.adens;
            } else {
              at.adens = fa->second;
            }
            m.n_aterms++;
            m.n_rterms += (int64_t)at.rt.size();
            m.n_bterms += nb_this;
            it.at.push_back(std::move(at));
          }
        }
      if (!it.at.empty()) m.items.push_back(std::move(it));
    }
  }
  // like-particle work: every stored element touches a column of the block
  for (int L = 0; L < m.basd; ++L) {
    if (!m.bas_active[L]) continue;
    const auto& Ha = m.A.H[m.bas_a[L] / 2];
    const auto& Hb = m.B.H[m.bas_b[L] / 2];
    for (double v : Ha) if (v != 0.0) m.flops_like += 2.0 * m.db[L];
    for (double v : Hb) if (v != 0.0) m.flops_like += 2.0 * m.da[L];
  }
  m.flops_ref += m.flops_like;
  return "";
}

}  // namespace nsx
