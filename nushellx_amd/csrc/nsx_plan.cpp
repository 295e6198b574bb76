#include "nsx_plan.hpp"

#include <stdexcept>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <array>
#include <map>
#include <unordered_map>
#include <set>
#include <string>

namespace nsx {

// Hash of a matrix's bit patterns, one 64-bit word per step (multiply-xorshift
// mixing); equality is still decided by an exact comparison.
static uint64_t content_hash(const double* v, size_t n) {
  uint64_t h = 0x9e3779b97f4a7c15ull ^ (uint64_t)n;
  for (size_t i = 0; i < n; ++i) {
    uint64_t w;
    std::memcpy(&w, v + i, 8);
    h = (h ^ w) * 0xff51afd7ed558ccdull;
    h ^= h >> 32;
  }
  h *= 0xc4ceb9fe1a85ec53ull;
  return h ^ (h >> 29);
}

// Append a matrix to `store` unless an identical one is already there; returns its offset.
static int64_t intern(std::vector<double>& store, std::unordered_map<uint64_t, std::vector<int64_t>>& idx,
                      const double* v, size_t n, std::vector<std::pair<int64_t, int64_t>>& blocks) {
  const uint64_t h = content_hash(v, n);
  auto& cands = idx[h];
  for (int64_t off : cands)
    if (off + (int64_t)n <= (int64_t)store.size() && std::equal(v, v + n, store.begin() + off)) return off;
  const int64_t off = (int64_t)store.size();
  store.insert(store.end(), v, v + n);
  cands.push_back(off);
  blocks.push_back({off, (int64_t)n});
  return off;
}

// Linear-ensemble interning: a Bhat^T run is identified by the files it is
// folded from (b side, interaction) and its group key, Jb' range and slots; a
// new run records, per slot, its fold recipe over the b densities.
static int64_t ens_intern_run(HostPlan& p, const Model& m, const std::array<int, 4>& grp,
                              const std::map<int, int>& slot_of, std::map<int, int64_t>& within,
                              const std::vector<double>& run) {
  std::string key = "B|" + m.opfile_b + "|" + m.intfile;
  for (int x : grp) key += "," + std::to_string(x);
  for (auto& sl : slot_of) key += ";" + std::to_string(sl.first);
  key += "|" + std::to_string(run.size());
  auto f = p.ens_key.find(key);
  if (f != p.ens_key.end()) return f->second;
  const int64_t off = (int64_t)p.Bh.size();
  p.Bh.insert(p.Bh.end(), run.begin(), run.end());
  p.b_blocks.push_back({off, (int64_t)run.size()});
  p.ens_key[key] = off;
  const int bjl = grp[0], db = m.B.d(bjl);
  for (auto& sl : slot_of) {
    const int bjr = sl.first, dbr = m.B.d(bjr);
    EnsFold fd;
    fd.dst = off + within[bjr];
    fd.n = dbr * db;
    fd.p_begin = (int32_t)p.ens_pair.size();
    for (const auto& pr : m.bhat_src[sl.second]) {
      // the density in the Bhat^T layout, stored once per (b side, density, Jb, Jb')
      const std::string dk = "D|" + m.opfile_b + "|" + std::to_string(pr.first) + "," + std::to_string(bjl) +
                             "," + std::to_string(bjr);
      auto g = p.ens_key.find(dk);
      int64_t src;
      if (g == p.ens_key.end()) {
        const std::vector<double> D = densify(m.B, m.B.dens[pr.first], bjl, bjr);  // db x dbr
        src = (int64_t)p.ens_bsrc.size();
        for (int c = 0; c < db; ++c)
          for (int i = 0; i < dbr; ++i) p.ens_bsrc.push_back(D[size_t(i) * db + c]);
        p.ens_key[dk] = src;
      } else {
        src = g->second;
      }
      EnsPair ep;
      ep.src = src;
      ep.c = m.renorm;
      ep.val = p.ens_opcol + pr.second;
      ep.pad = 0;
      p.ens_pair.push_back(ep);
    }
    fd.p_end = (int32_t)p.ens_pair.size();
    fd.pad = 0;
    p.ens_fold.push_back(fd);
  }
  return off;
}

// Linear-ensemble interning of a like-particle H(J) block: identified by the
// side's file name, the interaction suffix and J.
static int64_t ens_intern_h(HostPlan& p, const Model& m, int side, int j) {
  const SideFiles& sf = side ? m.B : m.A;
  const std::string key = "H|" + (side ? m.opfile_b : m.opfile_a) + "|" +
                          (m.intfile.empty() ? std::string("x") : m.intfile.substr(m.intfile.size() - 1)) + "|" +
                          std::to_string(j);
  auto f = p.ens_key.find(key);
  if (f != p.ens_key.end()) {
    p.ens_h_alias.push_back({f->second, (int64_t)sf.H[j].size(), p.ens_t, side, j, 0});
    return f->second;
  }
  const int64_t off = (int64_t)p.H.size();
  p.H.insert(p.H.end(), sf.H[j].begin(), sf.H[j].end());
  p.h_blocks.push_back({off, (int64_t)sf.H[j].size()});
  p.ens_key[key] = off;
  p.ens_h.push_back({off, (int64_t)sf.H[j].size(), p.ens_t, side, j, 0});
  return off;
}

static int col_tile(int rows) {
  // widest column tile whose da x nc outputs fit one workgroup's registers
  int nc = kTileElems / std::max(rows, 1);
  return std::max(nc, 1);
}

void plan_add(HostPlan& p, const Model& m) {
  const int64_t vb = p.N;
  p.vbase.push_back(vb);
  if (m.truncate && p.vmask.empty()) p.vmask.assign(size_t(vb), 1);
  if (m.truncate)
    p.vmask.insert(p.vmask.end(), m.trunc.begin(), m.trunc.end());
  else if (!p.vmask.empty())
    p.vmask.insert(p.vmask.end(), size_t(m.N), 1);
  // operator values of this instance
  std::vector<int64_t> a_off(m.adense.size()), b_off(m.bhat.size());
  for (size_t k = 0; k < m.adense.size(); ++k)
    a_off[k] = intern(p.Ad, p.dedup_a, m.adense[k].data(), m.adense[k].size(), p.a_blocks);
  // Bhat^T blocks grouped by (Jb, lam, sa, sa') with every Jb' of the group's
  // range adjacent (zeros where a Jb' has no operator), so that the blocks one
  // a-term needs form a single contiguous run.
  {
    std::map<std::array<int, 4>, std::vector<int>> groups;
    for (size_t k = 0; k < m.bhat.size(); ++k) {
      const auto& key = m.bhat_key[k];
      groups[{key[0], key[2], key[3], key[4]}].push_back((int)k);
    }
    std::vector<double> run;
    for (auto& g : groups) {
      const int bjl = g.first[0];
      const int db = m.B.d(bjl);
      int jmin = 1 << 30, jmax = -1;
      std::map<int, int> slot_of;
      for (int k : g.second) {
        jmin = std::min(jmin, m.bhat_key[k][1]);
        jmax = std::max(jmax, m.bhat_key[k][1]);
        slot_of[m.bhat_key[k][1]] = k;
      }
      run.clear();
      std::map<int, int64_t> within;
      for (int bjr = jmin; bjr <= jmax; bjr += m.B.delJ2) {
        const int dbr = m.B.d(bjr);
        within[bjr] = (int64_t)run.size();
        auto it = slot_of.find(bjr);
        // Bhat^T (dbr x db, column-major): element (i, c) = Bhat(c, i)
        for (int c = 0; c < db; ++c)
          for (int i = 0; i < dbr; ++i)
            run.push_back(it == slot_of.end() ? 0.0 : m.bhat[it->second][size_t(i) * db + c]);
      }
      const int64_t off = p.ens ? ens_intern_run(p, m, g.first, slot_of, within, run)
                                : intern(p.Bh, p.dedup_b, run.data(), run.size(), p.b_blocks);
      for (int k : g.second) b_off[k] = off + within[m.bhat_key[k][1]];
    }
  }
  p.inst_a_off.push_back(a_off);
  p.inst_b_off.push_back(b_off);
  std::vector<int64_t> ha_off(m.A.nJ(), -1), hb_off(m.B.nJ(), -1);
  int64_t nnz_like = 0;
  for (int j = 0; j < m.A.nJ(); ++j)
    if (!m.A.H.empty() && !m.A.H[j].empty()) {
      ha_off[j] = p.ens ? ens_intern_h(p, m, 0, j)
                        : intern(p.H, p.dedup_h, m.A.H[j].data(), m.A.H[j].size(), p.h_blocks);
      for (double v : m.A.H[j]) nnz_like += v != 0.0;
    }
  for (int j = 0; j < m.B.nJ(); ++j)
    if (!m.B.H.empty() && !m.B.H[j].empty()) {
      hb_off[j] = p.ens ? ens_intern_h(p, m, 1, j)
                        : intern(p.H, p.dedup_h, m.B.H[j].data(), m.B.H[j].size(), p.h_blocks);
      for (double v : m.B.H[j]) nnz_like += v != 0.0;
    }
  // items: partial tiles
  std::vector<std::vector<int64_t>> parts_of(m.basd);
  for (const Item& it : m.items) {
    const int da = m.da[it.L], db = m.db[it.L], dar = m.A.d(it.ajr);
    const int64_t base = p.part_size;
    p.part_size += int64_t(da) * db;
    parts_of[it.L].push_back(base);
    const int a_begin = (int)p.aterms.size();
    for (const ATerm& at : it.at) {
      DATerm t;
      t.a_off = a_off[at.adens];
      t.r_begin = (int)p.rterms.size();
      for (const RTerm& rt : at.rt) {
        DRTerm r;
        r.v_off = vb + m.posnv[rt.R];
        r.b_off = b_off[rt.bhat];
        r.r = rt.r;
        r.dbr = m.db[rt.R];
        r.pad = 0;
        p.rterms.push_back(r);
      }
      t.r_end = (int)p.rterms.size();
      p.aterms.push_back(t);
    }
    const int a_end = (int)p.aterms.size();
    int maxdbr = 0;
    for (const ATerm& at : it.at)
      for (const RTerm& rt : at.rt) maxdbr = std::max(maxdbr, m.db[rt.R]);
    if (std::max(std::max(da, dar), std::max(db, maxdbr)) <= p.cfg.big_min) {
      // phases: one per a-term (split when its runs exceed kFatCap); items are
      // cut into chunks of <= kMaxPhases phases, each chunk owning a partial
      std::vector<DPhase> ph_item;
      for (const ATerm& at : it.at) {
        const DATerm& t = p.aterms[a_begin + (int)(&at - &it.at[0])];
        const int na = da * dar;
        int u0 = t.r_begin;
        while (u0 < t.r_end) {
          int u1 = u0;
          int64_t v0 = p.rterms[u0].v_off, b0 = p.rterms[u0].b_off;
          int nv = 0, nb = 0;
          while (u1 < t.r_end) {
            const DRTerm& r = p.rterms[u1];
            if (r.v_off < v0 || r.b_off < b0) break;   // runs must grow monotonically
            const int nv2 = (int)(r.v_off - v0) + dar * r.dbr;
            const int nb2 = (int)(r.b_off - b0) + r.dbr * db;
            const bool last = (u1 + 1 == t.r_end);
            if (u1 > u0 && nv2 + nb2 + (last ? na : 0) > p.cfg.fat_cap) break;
            nv = nv2; nb = nb2;
            ++u1;
          }
          DPhase ph;
          ph.v_off = v0; ph.b_off = b0;
          ph.nv = nv; ph.nb = nb;
          const bool has_a = (u1 == t.r_end);
          ph.a_off = has_a ? t.a_off : -1;
          ph.na = has_a ? na : 0;
          ph.pair_begin = (int)p.pairs.size();
          for (int u = u0; u < u1; ++u) {
            const DRTerm& r = p.rterms[u];
            DPair pr;
            pr.r = r.r;
            pr.v_lds = (int)(r.v_off - v0);
            pr.b_lds = nv + (int)(r.b_off - b0);
            pr.dbr = r.dbr;
            pr.pad = 0;
            p.pairs.push_back(pr);
          }
          ph.pair_end = (int)p.pairs.size();
          ph.pad = 0;
          p.lds_buf = std::max(p.lds_buf, ph.nv + ph.nb + ph.na);
          ph_item.push_back(ph);
          u0 = u1;
        }
      }
      parts_of[it.L].pop_back();
      p.part_size = base;
      size_t q = 0;
      while (q < ph_item.size()) {
        DItem2 d2;
        d2.w_off = vb + m.posnv[it.L];
        d2.part_off = p.part_size;
        p.part_size += int64_t(da) * db;
        parts_of[it.L].push_back(d2.part_off);
        d2.da = da; d2.dar = dar; d2.db = db; d2.pad = 0;
        d2.ph_begin = (int)p.phases.size();
        int n = 0;
        // cut only after a phase carrying A (an a-term never spans two chunks)
        while (q < ph_item.size()) {
          p.phases.push_back(ph_item[q]);
          ++n;
          const bool a_done = ph_item[q].na > 0;
          ++q;
          if (a_done && n >= p.cfg.max_phases) break;
        }
        d2.ph_end = (int)p.phases.size();
        p.items2.push_back(d2);
      }
      p.lds_y = std::max(p.lds_y, dar * db);
      p.max_units = std::max(p.max_units, std::max(da, dar) * ((db + kColBlk - 1) / kColBlk));
      continue;
    }
    int nc = std::min(col_tile(std::max(da, dar)), db);
    nc = std::min(nc, std::max(1, kMaxLdsDoubles / std::max(dar, 1)));
    for (int j0 = 0; j0 < db; j0 += nc) {
      DItem d;
      d.part_off = base;
      d.w_off = vb + m.posnv[it.L];
      d.da = da; d.dar = dar; d.db = db;
      d.j0 = j0; d.nc = std::min(nc, db - j0);
      d.a_begin = a_begin; d.a_end = a_end; d.pad = 0;
      p.items.push_back(d);
      p.max_lds = std::max(p.max_lds, dar * d.nc);
    }
  }
  for (int L = 0; L < m.basd; ++L) {
    const int da = m.da[L], db = m.db[L];
    const int p_begin = (int)p.part_list.size();
    p.part_list.insert(p.part_list.end(), parts_of[L].begin(), parts_of[L].end());
    const int p_end = (int)p.part_list.size();
    const int nc = std::min(col_tile(da), db);
    for (int j0 = 0; j0 < db; j0 += nc) {
      DBlock b;
      b.v_off = vb + m.posnv[L];
      // an inactive block of a truncated basis gets W_L = 0 (multiplya/b skip it, :86-88)
      b.ha_off = (m.azero || !m.bas_active[L]) ? -1 : ha_off[m.bas_a[L] / 2];
      b.hb_off = (m.bzero || !m.bas_active[L]) ? -1 : hb_off[m.bas_b[L] / 2];
      b.da = da; b.db = db; b.j0 = j0; b.nc = std::min(nc, db - j0);
      b.p_begin = p_begin; b.p_end = p_end;
      p.blocks.push_back(b);
    }
  }
  p.N += m.N;
  p.flops_ref += m.flops_ref;
  p.flops_like += m.flops_like;
  // algorithmic bytes (FP64): read v, write w, read every operator value once
  p.bytes_alg += 16 * m.N + 8 * (m.nnz_a + m.nnz_b + nnz_like + m.n_bterms);
}

namespace {
// The two stacked stages' work, grouped: distinct (V block, Bhat^T block)
// products by Bhat^T block, and a-terms (with their X references) by left block.
struct StkU1 { int64_t v_off; int dar, dbr, db; int64_t x_off; };
struct StkT2 { int64_t a_off; int dar; std::vector<std::pair<int64_t, double>> x; };
struct StkGroups {
  std::map<std::array<int64_t, 3>, std::vector<StkU1>> by_b;  // (b_off, dbr, db) -> products
  // (instance, 2Jb, 2Ja, w_off, da, db) -> a-terms: left blocks of one instance
  // with the same Jb read the same X products, so their chunks are kept adjacent
  std::map<std::array<int64_t, 6>, std::vector<StkT2>> by_w;
};

// Fills g with the small-block items of `models` and sets p.x_size (X buffer
// elements; each product is dar x db).
void stacked_groups(HostPlan& p, const std::vector<const Model*>& models, StkGroups& g) {
  p.x_size = 0;
  std::map<std::pair<int64_t, int64_t>, int64_t> x_of;        // (b_off, v_off) -> x_off
  for (size_t mi = 0; mi < models.size(); ++mi) {
    const Model& m = *models[mi];
    const int64_t vb = p.vbase[mi];
    const auto& a_off = p.inst_a_off[mi];
    const auto& b_off = p.inst_b_off[mi];
    for (const Item& it : m.items) {
      const int da = m.da[it.L], db = m.db[it.L], dar = m.A.d(it.ajr);
      int maxdbr = 0;
      for (const ATerm& at : it.at)
        for (const RTerm& rt : at.rt) maxdbr = std::max(maxdbr, m.db[rt.R]);
      if (std::max(std::max(da, dar), std::max(db, maxdbr)) > p.cfg.big_min) continue;  // direct kernel
      for (const ATerm& at : it.at) {
        StkT2 t;
        t.a_off = a_off[at.adens];
        t.dar = dar;
        for (const RTerm& rt : at.rt) {
          const int64_t v = vb + m.posnv[rt.R], b = b_off[rt.bhat];
          auto f = x_of.find({b, v});
          int64_t x;
          if (f == x_of.end()) {
            x = p.x_size;
            p.x_size += int64_t(dar) * db;
            x_of[{b, v}] = x;
            g.by_b[{b, m.db[rt.R], db}].push_back({v, dar, m.db[rt.R], db, x});
          } else {
            x = f->second;
          }
          t.x.push_back({x, rt.r});
        }
        g.by_w[{(int64_t)mi, m.bas_b[it.L], m.bas_a[it.L], vb + m.posnv[it.L], da, db}].push_back(std::move(t));
      }
    }
  }
}
}  // namespace

void build_stacked(HostPlan& p, const std::vector<const Model*>& models) {
  p.s_chunk1.clear(); p.s_unit1.clear(); p.s_chunk2.clear(); p.s_term2.clear(); p.s_xref.clear();
  p.s_lds1 = p.s_lds2 = 1;
  using U1 = StkU1;
  using T2 = StkT2;
  StkGroups grp;
  stacked_groups(p, models, grp);
  auto& by_b = grp.by_b;
  auto& by_w = grp.by_w;
  // an a-term's references must fit one chunk's table; otherwise keep the staged path
  for (auto& g : by_w)
    for (const T2& t : g.second)
      if ((int)t.x.size() > kStkXref) {
        p.cfg.stacked = 0;
        p.x_size = 0;
        return;
      }
  // a chunk holds at least one pair (term); the cap grows to fit the largest single one
  int cap = p.cfg.stk_cap;
  for (auto& g : by_b)
    for (const U1& u : g.second) cap = std::max(cap, u.dbr * u.db + u.dar * u.dbr);
  for (auto& g : by_w)
    for (const T2& t : g.second) cap = std::max(cap, (int)(g.first[4] * t.dar + t.dar * g.first[5]));
  const int wcap = std::max(p.cfg.stk_work, 1);
  for (auto& g : by_b) {
    const int dbr = (int)g.first[1], db = (int)g.first[2];
    const int ncb = (db + kColBlk - 1) / kColBlk;
    SChunk1 c{};
    auto open = [&]() {
      c = SChunk1{};
      c.b_off = g.first[0]; c.dbr = dbr; c.db = db;
      c.u_begin = c.u_end = (int)p.s_unit1.size();
      c.lds = dbr * db;
    };
    open();
    for (const U1& u : g.second) {
      const int need = u.dar * dbr, w = u.dar * ncb;
      if (c.u_end > c.u_begin && (c.lds + need > cap || c.work + w > wcap || c.u_end - c.u_begin >= kStkMeta)) {
        p.s_chunk1.push_back(c);
        p.s_lds1 = std::max(p.s_lds1, c.lds);
        open();
      }
      SUnit1 s{};
      s.v_off = u.v_off; s.x_off = u.x_off; s.dar = u.dar; s.lds = c.lds; s.work0 = c.work;
      p.s_unit1.push_back(s);
      c.u_end++;
      c.lds += need;
      c.work += w;
    }
    p.s_chunk1.push_back(c);
    p.s_lds1 = std::max(p.s_lds1, c.lds);
  }
  for (auto& g : by_w) {
    const int da = (int)g.first[4], db = (int)g.first[5];
    SChunk2 c{};
    auto open = [&]() {
      c = SChunk2{};
      c.w_off = g.first[3]; c.da = da; c.db = db;
      c.t_begin = c.t_end = (int)p.s_term2.size();
    };
    auto close = [&]() {
      if (c.t_end == c.t_begin) return;
      c.x_begin = p.s_term2[c.t_begin].x_begin;
      c.x_end = p.s_term2[c.t_end - 1].x_end;
      p.s_chunk2.push_back(c);
      p.s_lds2 = std::max(p.s_lds2, c.lds);
    };
    open();
    for (const T2& t : g.second) {
      const int need = da * t.dar + t.dar * db;
      const int nx = (int)p.s_xref.size() - (c.t_end > c.t_begin ? p.s_term2[c.t_begin].x_begin : 0);
      if (c.t_end > c.t_begin && (c.lds + need > cap || c.t_end - c.t_begin >= kStkTerms ||
                                  nx + (int)t.x.size() > kStkXref)) {
        close();
        open();
      }
      STerm2 s{};
      s.a_off = t.a_off; s.dar = t.dar; s.lds = c.lds;
      s.x_begin = (int)p.s_xref.size();
      for (auto& x : t.x) p.s_xref.push_back({x.first, x.second});
      s.x_end = (int)p.s_xref.size();
      p.s_term2.push_back(s);
      c.t_end++;
      c.lds += need;
    }
    close();
  }
}

void build_mstk(HostPlan& p, const std::vector<const Model*>& models) {
  p.s_chunk1.clear(); p.s_unit1.clear(); p.s_chunk2.clear(); p.s_term2.clear(); p.s_xref.clear();
  p.m_task1.clear(); p.m_unit1.clear(); p.m_task2.clear();
  p.s_lds1 = p.s_lds2 = 1;
  StkGroups grp;
  stacked_groups(p, models, grp);
  const int u1 = std::max(p.cfg.mx_units1, 1), t2 = std::max(p.cfg.mx_terms2, 1);
  p.mx_ks = 1;
  // the stage-1 launch carries the like-particle blocks: LDS for the largest
  // small one (V_L, H_a, H_b), at most 16 KiB; larger ones read global memory
  int need = 1;
  for (const DBlock& b : p.blocks)
    if (b.j0 == 0 && std::max(b.da, b.db) <= kSmallMax) need = std::max(need, b.da * b.db + b.da * b.da + b.db * b.db);
  p.s_lds1 = std::min(need, 2048);
  for (auto& g : grp.by_b) {
    const int dbr = (int)g.first[1], db = (int)g.first[2];
    p.mx_ks = std::max(p.mx_ks, (dbr + 3) / 4);
    const int n = (int)g.second.size();
    for (int u0 = 0; u0 < n; u0 += u1) {
      const int ub = (int)p.m_unit1.size();
      for (int u = u0; u < std::min(n, u0 + u1); ++u) {
        const StkU1& x = g.second[u];
        for (int i0 = 0; i0 < x.dar; i0 += kMxTile) {
          MUnit1 s{};
          s.v_off = x.v_off + i0; s.x_off = x.x_off + int64_t(db) * i0;
          s.ld = x.dar; s.rows = std::min(kMxTile, x.dar - i0);
          p.m_unit1.push_back(s);
        }
      }
      for (int j0 = 0; j0 < db; j0 += kMxTile) {
        MTask1 t{};
        t.b_off = g.first[0]; t.dbr = dbr; t.db = db; t.j0 = j0;
        t.u_begin = ub; t.u_end = (int)p.m_unit1.size();
        p.m_task1.push_back(t);
      }
    }
  }
  for (auto& g : grp.by_w) {
    const int da = (int)g.first[4], db = (int)g.first[5];
    const int n = (int)g.second.size();
    for (int q0 = 0; q0 < n; q0 += t2) {
      const int tb = (int)p.s_term2.size();
      for (int q = q0; q < std::min(n, q0 + t2); ++q) {
        const StkT2& t = g.second[q];
        p.mx_ks = std::max(p.mx_ks, (t.dar + 3) / 4);
        STerm2 s{};
        s.a_off = t.a_off; s.dar = t.dar;
        s.x_begin = (int)p.s_xref.size();
        for (auto& x : t.x) p.s_xref.push_back({x.first, x.second});
        s.x_end = (int)p.s_xref.size();
        p.s_term2.push_back(s);
      }
      for (int j0 = 0; j0 < db; j0 += kMxTile)
        for (int i0 = 0; i0 < da; i0 += kMxTile) {
          MTask2 t{};
          t.w_off = g.first[3]; t.da = da; t.db = db; t.i0 = i0; t.j0 = j0;
          t.t_begin = tb; t.t_end = (int)p.s_term2.size();
          p.m_task2.push_back(t);
        }
    }
  }
}

static void build_mxf_impl(HostPlan& p, const std::vector<const Model*>& models) {
  p.f_mtask.clear(); p.f_mstep.clear(); p.f_mu.clear(); p.Bn.clear(); p.bn_src.clear();
  p.f_dtile.clear(); p.f_dpart.clear(); p.mxf_det = false;
  p.mx_ks = 1;
  // Bhat^T block of the Bh store -> its untransposed copy in Bn
  std::unordered_map<int64_t, int64_t> bn_of;
  auto bn = [&](int64_t bo, int dbr, int db) {
    auto f = bn_of.find(bo);
    if (f != bn_of.end()) return f->second;
    const int64_t o = (int64_t)p.Bn.size();
    p.Bn.resize(p.Bn.size() + size_t(dbr) * db);
    for (int j = 0; j < db; ++j)
      for (int k = 0; k < dbr; ++k) p.Bn[o + j + int64_t(db) * k] = p.Bh[bo + k + int64_t(dbr) * j];
    bn_of[bo] = o;
    p.bn_src.push_back({bo, o, dbr, db});
    return o;
  };
  p.mxf_slots = p.cfg.mxf_slots == 2 ? 2 : kMfSlots;
  // the like-particle part of the small blocks folded into the path (then W
  // is cleared first instead of stored by k_reduce): not on truncated plans,
  // whose pn part reads the masked vector, nor on linear ensembles
  const bool fold_like = p.cfg.mxf_like && p.vmask.empty() && !p.ens;
  p.mxf_like = fold_like;
  std::unordered_map<int64_t, const DBlock*> blk_of;  // small block by its vector offset
  if (fold_like)
    for (const DBlock& b : p.blocks)
      if (b.j0 == 0 && b.nc == b.db && std::max(b.da, b.db) <= kSmallMax) blk_of[b.v_off] = &b;
  std::unordered_map<int, int64_t> id_bn, id_ad;
  std::unordered_map<int64_t, int64_t> hb_bn, ha_ad;
  auto ident = [](std::vector<double>& st, int d) {
    const int64_t o = (int64_t)st.size();
    st.resize(st.size() + size_t(d) * d, 0.0);
    for (int i = 0; i < d; ++i) st[o + i + int64_t(d) * i] = 1.0;
    return o;
  };
  auto copy_h = [&](std::vector<double>& st, int64_t h, int d) {
    const int64_t o = (int64_t)st.size();
    st.insert(st.end(), p.H.begin() + h, p.H.begin() + h + int64_t(d) * d);
    return o;
  };
  auto ident_bn = [&](int d) { auto f = id_bn.find(d); return f != id_bn.end() ? f->second : id_bn[d] = ident(p.Bn, d); };
  auto ident_ad = [&](int d) { auto f = id_ad.find(d); return f != id_ad.end() ? f->second : id_ad[d] = ident(p.Ad, d); };
  auto h_in_bn = [&](int64_t h, int d) { auto f = hb_bn.find(h); return f != hb_bn.end() ? f->second : hb_bn[h] = copy_h(p.Bn, h, d); };
  auto h_in_ad = [&](int64_t h, int d) { auto f = ha_ad.find(h); return f != ha_ad.end() ? f->second : ha_ad[h] = copy_h(p.Ad, h, d); };
  struct Cons { int64_t a = -1; std::map<int, double> r; };
  struct Slab {
    std::map<int, std::pair<int64_t, int>> prod;  // R -> (Bhat^T offset, d(Jb'))
    std::map<int, Cons> cons;                     // 2Ja -> a-term
  };
  const size_t nslot = p.cfg.mxf_slots == 2 ? 2 : kMfSlots;
  for (size_t mi = 0; mi < models.size(); ++mi) {
    const Model& m = *models[mi];
    const int64_t vb = p.vbase[mi];
    const auto& a_off = p.inst_a_off[mi];
    const auto& b_off = p.inst_b_off[mi];
    // (Ja', 2Jb) -> (lam, sa, sa') -> slab; per group the left blocks by 2Ja
    std::map<std::array<int, 2>, std::map<std::array<int, 3>, Slab>> groups;
    std::map<std::array<int, 2>, std::map<int, int>> lblk;  // group -> 2Ja -> L
    for (const Item& it : m.items) {
      const int L = it.L, da = m.da[L], db = m.db[L], dar = m.A.d(it.ajr);
      int maxdbr = 0;
      for (const ATerm& at : it.at)
        for (const RTerm& rt : at.rt) maxdbr = std::max(maxdbr, m.db[rt.R]);
      if (std::max(std::max(da, dar), std::max(db, maxdbr)) > p.cfg.big_min) continue;  // large-block paths
      const std::array<int, 2> g{it.ajr, m.bas_b[L]};
      lblk[g][m.bas_a[L]] = L;
      for (const ATerm& at : it.at) {
        if (at.rt.empty()) continue;
        const auto& bk = m.bhat_key[at.rt[0].bhat];
        Slab& sl = groups[g][{bk[2], bk[3], bk[4]}];
        Cons& c = sl.cons[m.bas_a[L]];
        if (c.a >= 0) throw std::runtime_error("build_mxf: two a-terms of one left block in one slab");
        c.a = a_off[at.adens];
        for (const RTerm& rt : at.rt) {
          const int64_t bo = b_off[rt.bhat];
          auto f = sl.prod.find(rt.R);
          if (f == sl.prod.end())
            sl.prod[rt.R] = {bo, m.db[rt.R]};
          else if (f->second.first != bo)
            throw std::runtime_error("build_mxf: one right block with two Bhat blocks in one slab");
          c.r[rt.R] += rt.r;
        }
      }
    }
    if (fold_like) {
      // the like-particle part as two more slabs of the group (Ja, Jb) of each
      // small block L: H_a V_L (product V_L I, consumer A = H_a) and V_L H_b^T
      // (product V_L H_b^T, consumer A = I), each with r = 1
      for (int L = 0; L < m.basd; ++L) {
        auto f = blk_of.find(vb + m.posnv[L]);
        if (f == blk_of.end()) continue;  // a large block (its like-particle tiles are k_big0's)
        const DBlock& b = *f->second;
        if (b.ha_off < 0 && b.hb_off < 0) continue;
        const std::array<int, 2> g{m.bas_a[L], m.bas_b[L]};
        lblk[g][m.bas_a[L]] = L;
        if (b.ha_off >= 0) {
          Slab& sl = groups[g][{-1, 0, 0}];
          sl.prod[L] = {-(ident_bn(b.db) + 1), b.db};
          Cons& c = sl.cons[m.bas_a[L]];
          c.a = h_in_ad(b.ha_off, b.da);
          c.r[L] = 1.0;
        }
        if (b.hb_off >= 0) {
          Slab& sl = groups[g][{-2, 0, 0}];
          sl.prod[L] = {-(h_in_bn(b.hb_off, b.db) + 1), b.db};
          Cons& c = sl.cons[m.bas_a[L]];
          c.a = ident_ad(b.da);
          c.r[L] = 1.0;
        }
      }
    }
    for (auto& g : groups) {
      const int dar = m.A.d(g.first[0]);
      const auto& lb = lblk[g.first];
      const int db = m.db[lb.begin()->second];
      struct Slot { int ja, L, i0; };
      std::vector<Slot> slots;
      for (auto& kv : lb)
        for (int i0 = 0; i0 < m.da[kv.second]; i0 += kMxTile) slots.push_back({kv.first, kv.second, i0});
      for (int a0 = 0; a0 < dar; a0 += kMxTile)
        for (int j0 = 0; j0 < db; j0 += kMxTile)
          for (size_t s0 = 0; s0 < slots.size(); s0 += nslot) {
            const int ns = (int)std::min<size_t>(nslot, slots.size() - s0);
            MFTask t{};
            for (int q = 0; q < kMfSlots; ++q) {
              t.w_off[q] = -1;
              t.da[q] = 1;
            }
            for (int q = 0; q < ns; ++q) {
              const Slot& sl = slots[s0 + q];
              const int da = m.da[sl.L];
              t.w_off[q] = vb + m.posnv[sl.L] + sl.i0 + int64_t(da) * j0;
              t.da[q] = da;
              t.rows_i[q] = std::min(kMxTile, da - sl.i0);
            }
            t.ld_v = dar;
            t.rows_a = std::min(kMxTile, dar - a0);
            t.db = db;
            t.cols_j = std::min(kMxTile, db - j0);
            // steps: the group's slabs that feed any slot of the task
            std::vector<int> step_ids;
            std::vector<double> cost;
            for (auto& kv : g.second) {
              const Slab& sl = kv.second;
              bool any = false;
              for (int q = 0; q < ns; ++q) any |= sl.cons.count(slots[s0 + q].ja) > 0;
              if (!any) continue;
              MFStep st{};
              double c = 0.0;
              for (int q = 0; q < kMfSlots; ++q) st.a_off[q] = -1;
              for (int q = 0; q < ns; ++q) {
                const Slot& so = slots[s0 + q];
                auto f = sl.cons.find(so.ja);
                if (f == sl.cons.end()) continue;
                st.a_off[q] = f->second.a + so.i0 + int64_t(m.da[so.L]) * a0;
                c += 4.0;
              }
              st.u_begin = (int)p.f_mu.size();
              for (auto& pr : sl.prod) {
                MFU u{};
                u.v_off = vb + m.posnv[pr.first] + a0;
                u.dbr = pr.second.second;
                const int64_t bo = pr.second.first;  // < 0: already in the Bn store, at -bo - 1
                u.b_off = (bo >= 0 ? bn(bo, u.dbr, db) : -bo - 1) + j0;
                for (int q = 0; q < ns; ++q) {
                  auto f = sl.cons.find(slots[s0 + q].ja);
                  if (f == sl.cons.end()) continue;
                  auto fr = f->second.r.find(pr.first);
                  if (fr != f->second.r.end()) u.r[q] = fr->second;
                }
                p.mx_ks = std::max(p.mx_ks, (u.dbr + 3) / 4);
                c += (u.dbr + 3) / 4 + 1;
                p.f_mu.push_back(u);
              }
              st.u_end = (int)p.f_mu.size();
              step_ids.push_back((int)p.f_mstep.size());
              cost.push_back(c);
              p.f_mstep.push_back(st);
            }
            if (step_ids.empty()) continue;
            // workgroups of at most 4 * mxf_steps steps (each a copy of the
            // task's slots), four contiguous step ranges of about equal cost
            const int per = p.cfg.mxf_steps > 0 ? 4 * p.cfg.mxf_steps : (int)step_ids.size();
            for (size_t c0 = 0; c0 < step_ids.size(); c0 += per) {
              const int n = (int)std::min<size_t>(per, step_ids.size() - c0);
              const int first = step_ids[c0];
              double total = 0.0;
              for (int q = 0; q < n; ++q) total += cost[c0 + q];
              t.p[0] = first;
              double acc = 0.0;
              int k = 0;
              for (int w = 1; w < 4; ++w) {
                while (k < n && acc + 0.5 * cost[c0 + k] < total * w / 4.0) acc += cost[c0 + k++];
                t.p[w] = first + k;
              }
              t.p[4] = first + n;
              p.f_mtask.push_back(t);
            }
          }
    }
  }
  // XCD ranges: a task's time is taken as its longest wavefront's cost (the
  // cost of the step split above: K steps + 1 per product, 4 per slot's A;
  // counting dependent load rounds instead balanced worse); contiguous ranges
  // keep neighbouring tasks (shared operators) in one XCD's L2, cut at equal
  // shares of the summed costs, and the longest tasks of a range go first, so
  // that the range does not end on one long task
  const int nt = (int)p.f_mtask.size();
  std::vector<double> crit(size_t(nt), 0.0);
  double total = 0.0;
  for (int i = 0; i < nt; ++i) {
    const MFTask& t = p.f_mtask[size_t(i)];
    for (int w = 0; w < 4; ++w) {
      double c = 0.0;
      for (int st = t.p[w]; st < t.p[w + 1]; ++st) {
        const MFStep& S = p.f_mstep[size_t(st)];
        for (int u = S.u_begin; u < S.u_end; ++u) c += (p.f_mu[size_t(u)].dbr + 3) / 4 + 1;
        for (int q = 0; q < kMfSlots; ++q) c += S.a_off[q] >= 0 ? 4.0 : 0.0;
      }
      crit[size_t(i)] = std::max(crit[size_t(i)], c);
    }
    total += crit[size_t(i)];
  }
  p.mxf_xcd[0] = 0;
  double acc = 0.0;
  for (int i = 0, x = 1; x < 8; ++x) {
    while (i < nt && acc + 0.5 * crit[size_t(i)] < total * x / 8.0) acc += crit[size_t(i++)];
    p.mxf_xcd[size_t(x)] = i;
  }
  p.mxf_xcd[8] = nt;
  std::vector<MFTask> sorted;
  sorted.reserve(size_t(nt));
  for (int x = 0; x < 8; ++x) {
    std::vector<int> ord;
    for (int i = p.mxf_xcd[size_t(x)]; i < p.mxf_xcd[size_t(x) + 1]; ++i) ord.push_back(i);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return crit[size_t(a)] > crit[size_t(b)]; });
    for (int i : ord) sorted.push_back(p.f_mtask[size_t(i)]);
  }
  p.f_mtask.swap(sorted);
  // deterministic mode: each W tile's partials, in task order
  p.mxf_det = p.cfg.deterministic != 0;
  if (p.mxf_det) {
    const int sl = p.mxf_slots;
    std::map<int64_t, MFDTile> tiles;
    std::map<int64_t, std::vector<int32_t>> parts;
    for (int i = 0; i < nt; ++i) {
      const MFTask& t = p.f_mtask[size_t(i)];
      for (int q = 0; q < sl; ++q) {
        if (t.w_off[q] < 0) continue;
        tiles[t.w_off[q]] = MFDTile{t.w_off[q], t.da[q], t.rows_i[q], t.cols_j, 0, 0, 0};
        parts[t.w_off[q]].push_back(i * sl + q);
      }
    }
    for (auto& kv : tiles) {
      MFDTile d = kv.second;
      d.c_begin = (int32_t)p.f_dpart.size();
      const auto& pl = parts[kv.first];
      p.f_dpart.insert(p.f_dpart.end(), pl.begin(), pl.end());
      d.c_end = (int32_t)p.f_dpart.size();
      p.f_dtile.push_back(d);
    }
  }
}

bool build_mxf(HostPlan& p, const std::vector<const Model*>& models) {
  const size_t ad0 = p.Ad.size();
  try {
    build_mxf_impl(p, models);
  } catch (const std::exception&) {
    p.f_mtask.clear(); p.f_mstep.clear(); p.f_mu.clear(); p.Bn.clear(); p.bn_src.clear();
    p.f_dtile.clear(); p.f_dpart.clear(); p.mxf_det = false;
    p.Ad.resize(ad0);
    p.mxf_like = false;
    return false;
  }
  return true;
}

bool build_fused(HostPlan& p, const std::vector<const Model*>& models, int threads) {
  p.f_task.clear(); p.f_blk.clear(); p.f_round.clear(); p.f_seg.clear(); p.f_x.clear();
  p.f_cons.clear(); p.f_ref.clear(); p.f_blocks.clear(); p.f_part_list.clear();
  p.f_part_size = 0;
  p.f_lds = p.f_units = 1;
  auto fail = [&]() {
    p.f_task.clear(); p.f_blk.clear(); p.f_round.clear(); p.f_seg.clear(); p.f_x.clear();
    p.f_cons.clear(); p.f_ref.clear(); p.f_blocks.clear(); p.f_part_list.clear();
    p.f_part_size = 0;
    return false;
  };
  // items that only the direct kernel takes would add partials of their own
  if (!p.items.empty() && p.g_tile2.empty()) return fail();
  const int cap = std::max(p.cfg.stk_cap, 256);
  const double target = std::max(p.cfg.stk_work, 1000);
  struct Cons { int bl, da; int64_t a_off; std::vector<std::pair<int, double>> refs; };  // refs: (slab X, r)
  struct XP { int64_t v_off, b_off; int dbr; };
  struct Slab { int dar = 0; std::vector<XP> x; std::vector<Cons> cons; double work = 0.0; };
  std::map<int64_t, std::vector<int64_t>> bases_of;  // block (by its v_off) -> partial bases
  for (size_t mi = 0; mi < models.size(); ++mi) {
    const Model& m = *models[mi];
    const int64_t vb = p.vbase[mi];
    const auto& a_off = p.inst_a_off[mi];
    const auto& b_off = p.inst_b_off[mi];
    std::map<int, std::vector<int>> col_blocks;  // Jb -> blocks, Ja ascending (basis order)
    for (int L = 0; L < m.basd; ++L) col_blocks[m.bas_b[L]].push_back(L);
    std::map<int, std::map<std::array<int, 4>, Slab>> cols;
    for (const Item& it : m.items) {
      const int da = m.da[it.L], db = m.db[it.L], dar = m.A.d(it.ajr);
      int maxdbr = 0;
      for (const ATerm& at : it.at)
        for (const RTerm& rt : at.rt) maxdbr = std::max(maxdbr, m.db[rt.R]);
      if (std::max(std::max(da, dar), std::max(db, maxdbr)) > p.cfg.big_min) continue;  // large-block tiles
      const int Jb = m.bas_b[it.L];
      const std::vector<int>& cb = col_blocks[Jb];
      const int bl = (int)(std::find(cb.begin(), cb.end(), it.L) - cb.begin());
      for (const ATerm& at : it.at) {
        const auto& bk = m.bhat_key[at.rt[0].bhat];
        Slab& sl = cols[Jb][{it.ajr, bk[2], bk[3], bk[4]}];
        sl.dar = dar;
        Cons c;
        c.bl = bl;
        c.da = da;
        c.a_off = a_off[at.adens];
        for (const RTerm& rt : at.rt) {
          const int64_t v = vb + m.posnv[rt.R], b = b_off[rt.bhat];
          int xi = 0;
          while (xi < (int)sl.x.size() && !(sl.x[xi].v_off == v && sl.x[xi].b_off == b)) ++xi;
          if (xi == (int)sl.x.size()) {
            sl.x.push_back({v, b, m.db[rt.R]});
            sl.work += 2.0 * dar * m.db[rt.R] * db;
          }
          c.refs.push_back({xi, rt.r});
        }
        sl.work += 2.0 * da * dar * db + double(c.refs.size()) * dar * db;
        sl.cons.push_back(std::move(c));
      }
    }
    for (auto& col : cols) {
      const std::vector<int>& blocks = col_blocks[col.first];
      const int db = m.B.d(col.first);
      const int ncb = (db + kColBlk - 1) / kColBlk;
      if ((int)blocks.size() > kFMaxBlk) return fail();
      const int b_begin = (int)p.f_blk.size();
      int wcol = 0, units = 0;
      for (int L : blocks) {
        p.f_blk.push_back(FBlk{wcol, m.da[L], units, 0});
        wcol += m.da[L] * db;
        units += m.da[L] * ncb;
      }
      const int b_end = (int)p.f_blk.size();
      if (units > kFMaxU * threads) return fail();
      p.f_units = std::max(p.f_units, units);
      // a slab whose round alone would exceed the LDS cap is split into groups of
      // its consumers, each with the products it reads (computed once per group)
      std::vector<Slab> split;
      auto lds_of = [&](const Slab& sl) {
        std::set<int64_t> v, b, a;
        int na = 0, np = 0;
        for (const XP& x : sl.x) {
          if (v.insert(x.v_off).second) np += sl.dar * x.dbr;
          if (b.insert(x.b_off).second) np += x.dbr * db;
        }
        for (const Cons& c : sl.cons)
          if (a.insert(c.a_off).second) na += c.da * sl.dar;
        const int nx = (int)sl.x.size() * sl.dar * db, ny = (int)sl.cons.size() * sl.dar * db;
        return na + nx + std::max(np, ny);
      };
      for (auto& sp : col.second) {
        const Slab& sl = sp.second;
        if (lds_of(sl) <= cap || sl.cons.size() < 2) {
          split.push_back(sl);
          continue;
        }
        size_t c0 = 0;
        while (c0 < sl.cons.size()) {
          Slab g;
          g.dar = sl.dar;
          std::map<int, int> xmap;  // slab X -> group X
          size_t c1 = c0;
          while (c1 < sl.cons.size()) {
            Slab t = g;
            std::map<int, int> tm = xmap;
            Cons c = sl.cons[c1];
            for (auto& f : c.refs) {
              auto it = tm.find(f.first);
              if (it == tm.end()) {
                it = tm.emplace(f.first, (int)t.x.size()).first;
                t.x.push_back(sl.x[f.first]);
              }
              f.first = it->second;
            }
            t.cons.push_back(std::move(c));
            if (c1 > c0 && lds_of(t) > cap) break;
            g = std::move(t);
            xmap = std::move(tm);
            ++c1;
          }
          const int db_ = db;
          for (const XP& x : g.x) g.work += 2.0 * g.dar * x.dbr * db_;
          for (const Cons& c : g.cons) g.work += 2.0 * c.da * g.dar * db_ + double(c.refs.size()) * g.dar * db_;
          split.push_back(std::move(g));
          c0 = c1;
        }
      }
      std::vector<const Slab*> slabs;
      double total = 0.0;
      for (const Slab& sl : split) {
        slabs.push_back(&sl);
        total += sl.work;
      }
      const int ntask = std::max(1, (int)std::ceil(total / target));
      const double per = total / ntask;
      size_t si = 0;
      FTask tk{};
      auto open_task = [&]() {
        tk = FTask{};
        tk.part_off = p.f_part_size;
        p.f_part_size += wcol;
        tk.b_begin = b_begin;
        tk.b_end = b_end;
        tk.db = db;
        tk.units = units;
        tk.r_begin = (int)p.f_round.size();
        for (size_t k = 0; k < blocks.size(); ++k)
          bases_of[vb + m.posnv[blocks[k]]].push_back(tk.part_off + p.f_blk[b_begin + k].wc_off);
      };
      for (int t = 0; t < ntask && si < slabs.size(); ++t) {
        open_task();
        // this task's slabs: up to its share of the column's work (the last takes the rest)
        double acc = 0.0;
        const size_t s0 = si;
        while (si < slabs.size() && (si == s0 || t == ntask - 1 || acc + 0.5 * slabs[si]->work <= per)) {
          acc += slabs[si]->work;
          ++si;
        }
        // rounds of slabs that fit the LDS buffer and the round tables.  A round's
        // LDS holds [A blocks][X products][V_R and Bhat^T blocks, later Y]: the
        // staged V_R and Bhat^T are dead once X is formed, so Y reuses them.
        size_t ri = s0;
        while (ri < si) {
          if ((int)p.f_round.size() - tk.r_begin == std::min(std::max(p.cfg.f_rounds, 1), kFMaxRounds)) {
            // continue in a task of its own
            tk.r_end = (int)p.f_round.size();
            p.f_task.push_back(tk);
            open_task();
          }
          FRound R{};
          R.s_begin = (int)p.f_seg.size();
          R.x_begin = (int)p.f_x.size();
          std::map<int64_t, int> v_lds, b_lds, a_lds;  // region-relative offsets
          int na_tot = 0, nx_tot = 0, np_tot = 0, ny_tot = 0, nw_tot = 0, nref = 0;
          struct RC { int bl, da, dar, a_lds, y_lds; std::vector<std::pair<int, double>> refs; };  // refs: (x_lds, r)
          std::vector<RC> rc;
          int last_seg[3] = {-1, -1, -1};
          // lengths and region-relative offsets of this round's segments (FSeg
          // stores the LDS offset and the prefix in 16 bits once the round is laid out)
          std::vector<int> seg_len, seg_off;
          auto seg = [&](int64_t src, int len, int kind, int off) {
            const int ls = last_seg[kind == 2 ? 2 : 0];
            if (ls >= 0) {
              const int k = ls - R.s_begin;
              if (p.f_seg[ls].kind == kind && p.f_seg[ls].src + seg_len[k] == src && seg_off[k] + seg_len[k] == off) {
                seg_len[k] += len;
                return;
              }
            }
            last_seg[kind == 2 ? 2 : 0] = (int)p.f_seg.size();
            FSeg g{};
            g.src = src;
            g.kind = (int16_t)kind;
            p.f_seg.push_back(g);
            seg_len.push_back(len);
            seg_off.push_back(off);
          };
          while (ri < si) {
            const Slab& sl = *slabs[ri];
            const int dar = sl.dar;
            int da_ = 0, dx = 0, dp = 0, dy = 0;
            std::set<int64_t> nv, nb, na;
            for (const XP& x : sl.x) {
              if (!v_lds.count(x.v_off) && nv.insert(x.v_off).second) dp += dar * x.dbr;
              if (!b_lds.count(x.b_off) && nb.insert(x.b_off).second) dp += x.dbr * db;
              dx += dar * db;
            }
            int nr = 0;
            for (const Cons& c : sl.cons) {
              if (!a_lds.count(c.a_off) && na.insert(c.a_off).second) da_ += c.da * dar;
              dy += dar * db;
              nr += (int)c.refs.size();
            }
            int dw = 0;  // W-item partial sums (4 per item), in the X space once Y is formed
            for (const Cons& c : sl.cons) dw += kColBlk * c.da * ncb;
            const int after = na_tot + da_ + std::max(nx_tot + dx, nw_tot + dw) + std::max(np_tot + dp, ny_tot + dy);
            const int nx = (int)p.f_x.size() - R.x_begin;
            const int ns = (int)p.f_seg.size() - R.s_begin;
            const int new_segs = (int)(nv.size() + nb.size() + na.size());
            if (!rc.empty() && (after > cap || nx + (int)sl.x.size() > kFMaxX ||
                                (int)(rc.size() + sl.cons.size()) > kFMaxC || nref + nr > kFMaxRef ||
                                ns + new_segs > kFMaxSeg))
              break;
            if ((int)sl.x.size() > kFMaxX || (int)sl.cons.size() > kFMaxC || nr > kFMaxRef || new_segs > kFMaxSeg)
              return fail();
            std::vector<int> x_lds(sl.x.size());
            for (size_t k = 0; k < sl.x.size(); ++k) {
              const XP& x = sl.x[k];
              auto fv = v_lds.find(x.v_off);
              if (fv == v_lds.end()) {
                fv = v_lds.emplace(x.v_off, np_tot).first;
                seg(x.v_off, dar * x.dbr, 0, np_tot);
                np_tot += dar * x.dbr;
              }
              auto fb = b_lds.find(x.b_off);
              if (fb == b_lds.end()) {
                fb = b_lds.emplace(x.b_off, np_tot).first;
                seg(x.b_off, x.dbr * db, 1, np_tot);
                np_tot += x.dbr * db;
              }
              FX fx{};
              fx.v_lds = fv->second;
              fx.b_lds = fb->second;
              fx.x_lds = nx_tot;
              fx.dar = dar;
              fx.dbr = x.dbr;
              x_lds[k] = nx_tot;
              nx_tot += dar * db;
              p.f_x.push_back(fx);
            }
            for (const Cons& c : sl.cons) {
              auto fa = a_lds.find(c.a_off);
              if (fa == a_lds.end()) {
                fa = a_lds.emplace(c.a_off, na_tot).first;
                seg(c.a_off, c.da * dar, 2, na_tot);
                na_tot += c.da * dar;
              }
              RC r{c.bl, c.da, dar, fa->second, ny_tot, {}};
              ny_tot += dar * db;
              nw_tot += kColBlk * c.da * ncb;
              for (auto& f : c.refs) r.refs.push_back({x_lds[f.first], f.second});
              rc.push_back(std::move(r));
            }
            nref += nr;
            ++ri;
          }
          // region bases: A at 0, X (later the W partials) after A, V_R / Bhat^T
          // (later Y) after that
          const int xbase = na_tot, pbase = na_tot + std::max(nx_tot, nw_tot);
          const int lds = pbase + std::max(np_tot, ny_tot);
          R.wp_lds = xbase;
          if (lds > kFMaxLds) return fail();
          R.s_end = (int)p.f_seg.size();
          int pre = 0;
          for (int k = R.s_begin; k < R.s_end; ++k) {
            p.f_seg[k].lds = (int16_t)(seg_off[k - R.s_begin] + (p.f_seg[k].kind == 2 ? 0 : pbase));
            p.f_seg[k].pre = (int16_t)pre;
            pre += seg_len[k - R.s_begin];
          }
          R.stage = pre;
          R.x_end = (int)p.f_x.size();
          int w = 0;
          for (int k = R.x_begin; k < R.x_end; ++k) {
            FX& fx = p.f_x[k];
            fx.v_lds += pbase;
            fx.b_lds += pbase;
            fx.x_lds += xbase;
            fx.work0 = w;
            w += fx.dar * ncb;
          }
          R.xwork = w;
          std::stable_sort(rc.begin(), rc.end(), [](const RC& a, const RC& b) { return a.bl < b.bl; });
          R.c_begin = (int)p.f_cons.size();
          R.f_begin = (int)p.f_ref.size();
          int y0 = 0, w0 = 0;
          for (const RC& r : rc) {
            FCons c{};
            c.w0 = w0;
            w0 += r.da * ncb;
            c.a_lds = r.a_lds;
            c.y_lds = pbase + r.y_lds;
            c.da = r.da;
            c.dar = r.dar;
            c.f_begin = (int)p.f_ref.size() - R.f_begin;
            for (auto& f : r.refs) p.f_ref.push_back(FRef{f.second, xbase + f.first, 0});
            c.f_end = (int)p.f_ref.size() - R.f_begin;
            c.y0 = y0;
            y0 += r.dar * db;
            p.f_cons.push_back(c);
          }
          R.c_end = (int)p.f_cons.size();
          R.f_end = (int)p.f_ref.size();
          R.ywork = y0;
          R.wwork = w0;
          for (int b = 0, c = 0; b <= kFMaxBlk; ++b) {
            while (c < (int)rc.size() && rc[c].bl < b) ++c;
            R.cb[b] = c;
          }
          p.f_lds = std::max(p.f_lds, lds);
          p.f_round.push_back(R);
        }
        tk.r_end = (int)p.f_round.size();
        p.f_task.push_back(tk);
      }
    }
  }
  for (const DBlock& b : p.blocks) {
    DBlock f = b;
    f.p_begin = (int)p.f_part_list.size();
    auto it = bases_of.find(b.v_off);
    if (it != bases_of.end()) p.f_part_list.insert(p.f_part_list.end(), it->second.begin(), it->second.end());
    f.p_end = (int)p.f_part_list.size();
    p.f_blocks.push_back(f);
  }
  return true;
}

// K-step mask of the columns [c0, c0 + w) x rows [r0, r0 + h) of a column-major
// matrix M (ld rows), over steps of kBigStep along the columns (col_steps) or
// along the rows (!col_steps)
static void step_mask(const double* M, int ld, int r0, int h, int c0, int w, bool col_steps, uint64_t (&mask)[2]) {
  mask[0] = mask[1] = 0;
  for (int c = c0; c < c0 + w; ++c)
    for (int r = r0; r < r0 + h; ++r)
      if (M[r + (int64_t)ld * c] != 0.0) {
        const int s = (col_steps ? c : r) / kBigStep;
        mask[s >> 6] |= uint64_t(1) << (s & 63);
      }
}

void build_big(HostPlan& p, const std::vector<const Model*>& models) {
  p.g_task1.clear(); p.g_tile1.clear(); p.g_tile2.clear(); p.g_tile3.clear(); p.g_term2.clear();
  p.g_xref.clear();
  p.xg_size = 0;
  // the K-step masks cover inner dimensions up to kBigMaxK; larger blocks keep the direct kernel
  for (size_t mi = 0; mi < models.size(); ++mi)
    for (int L = 0; L < models[mi]->basd; ++L)
      if (models[mi]->da[L] > kBigMaxK || models[mi]->db[L] > kBigMaxK) return;
  // like-particle tiles of every large block (the reduce kernels skip these blocks)
  for (const DBlock& b : p.blocks) {
    if (b.j0 != 0 || std::max(b.da, b.db) <= kSmallMax) continue;
    for (int j0 = 0; j0 < b.db; j0 += kBigTile)
      for (int i0 = 0; i0 < b.da; i0 += kBigTile)
        p.g_tile3.push_back(GTile3{b.v_off, b.ha_off, b.hb_off, b.da, b.db, i0, j0});
  }
  struct T { int64_t a_off; int dar; std::vector<std::pair<int, double>> u; };  // refs: (task, r)
  std::map<std::pair<int64_t, int64_t>, int> task_of;      // (b_off, v_off) -> task
  std::vector<std::vector<char>> rows_needed;               // per task: i-tiles of X some term reads
  std::map<std::pair<size_t, int>, std::vector<T>> terms_of;
  std::map<std::pair<size_t, int>, std::array<int64_t, 3>> block_of;  // -> (w_off, da, db)
  for (size_t mi = 0; mi < models.size(); ++mi) {
    const Model& m = *models[mi];
    const int64_t vb = p.vbase[mi];
    const auto& a_off = p.inst_a_off[mi];
    const auto& b_off = p.inst_b_off[mi];
    for (const Item& it : m.items) {
      const int da = m.da[it.L], db = m.db[it.L], dar = m.A.d(it.ajr);
      int maxdbr = 0;
      for (const ATerm& at : it.at)
        for (const RTerm& rt : at.rt) maxdbr = std::max(maxdbr, m.db[rt.R]);
      if (std::max(std::max(da, dar), std::max(db, maxdbr)) <= p.cfg.big_min) continue;  // small-block paths
      block_of[{mi, it.L}] = {vb + m.posnv[it.L], da, db};
      auto& tl = terms_of[{mi, it.L}];
      for (const ATerm& at : it.at) {
        T t;
        t.a_off = a_off[at.adens];
        t.dar = dar;
        // X rows this term reads: the columns of A that are not all zero
        const double* A = p.Ad.data() + t.a_off;
        std::vector<char> used((dar + kBigTile - 1) / kBigTile, 0);
        for (int c = 0; c < dar; ++c)
          for (int r = 0; r < da; ++r)
            if (A[r + (int64_t)da * c] != 0.0) { used[c / kBigTile] = 1; break; }
        for (const RTerm& rt : at.rt) {
          const int64_t v = vb + m.posnv[rt.R], b = b_off[rt.bhat];
          auto f = task_of.find({b, v});
          int task;
          if (f == task_of.end()) {
            task = (int)p.g_task1.size();
            task_of[{b, v}] = task;
            p.g_task1.push_back(GTask1{v, b, p.xg_size, dar, m.db[rt.R], db, 0});
            p.xg_size += int64_t(big_ldx(dar)) * db;  // stays a multiple of kBigStep
            rows_needed.emplace_back(used.size(), 0);
          } else {
            task = f->second;
          }
          for (size_t q = 0; q < used.size(); ++q) rows_needed[task][q] |= used[q];
          t.u.push_back({task, rt.r});
        }
        tl.push_back(std::move(t));
      }
    }
  }
  // stage 1: the X tiles some term reads, with the K steps of their Bhat^T columns
  for (size_t task = 0; task < p.g_task1.size(); ++task) {
    const GTask1& g = p.g_task1[task];
    const double* Bt = p.Bh.data() + g.b_off;
    for (int j0 = 0; j0 < g.n; j0 += kBigTile) {
      GTile1 t{};
      step_mask(Bt, g.k, 0, g.k, j0, std::min(kBigTile, g.n - j0), false, t.kmask);
      if (!(t.kmask[0] | t.kmask[1])) continue;  // X columns identically zero
      for (int i0 = 0; i0 < g.m; i0 += p.cfg.big_rows1) {
        bool need = false;  // rows_needed is kept per kBigTile rows
        for (int q = i0 / kBigTile; q < (int)rows_needed[task].size() && q * kBigTile < i0 + p.cfg.big_rows1; ++q)
          need = need || rows_needed[task][q];
        if (!need) continue;
        t.task = (int)task;
        t.i0 = i0;
        t.j0 = j0;
        p.g_tile1.push_back(t);
      }
    }
  }
  // stage 2: per left block and row tile, the terms whose A rows there are not
  // all zero, with their K steps, in chunks of about kBigKChunk active steps
  for (auto& kv : terms_of) {
    const auto& blk = block_of[kv.first];
    const int da = (int)blk[1], db = (int)blk[2];
    for (int i0 = 0; i0 < da; i0 += kBigTile) {
      std::vector<GTerm2> row;
      for (const T& t : kv.second) {
        GTerm2 g{};
        step_mask(p.Ad.data() + t.a_off, da, i0, std::min(kBigTile, da - i0), 0, t.dar, true, g.kmask);
        if (!(g.kmask[0] | g.kmask[1])) continue;
        g.a_off = t.a_off;
        g.dar = t.dar;
        g.x_begin = (int)p.g_xref.size();
        for (auto& u : t.u) p.g_xref.push_back({p.g_task1[u.first].x_off, u.second});
        g.x_end = (int)p.g_xref.size();
        row.push_back(g);
      }
      size_t q = 0;
      while (q < row.size()) {
        const int t_begin = (int)p.g_term2.size();
        int steps = 0;
        while (q < row.size()) {
          const int s = __builtin_popcountll(row[q].kmask[0]) + __builtin_popcountll(row[q].kmask[1]);
          if (steps > 0 && steps + s > p.cfg.big_kchunk) break;
          steps += s;
          p.g_term2.push_back(row[q++]);
        }
        const int t_end = (int)p.g_term2.size();
        for (int j0 = 0; j0 < db; j0 += kBigTile)
          p.g_tile2.push_back(GTile2{blk[0], da, db, i0, j0, t_begin, t_end});
      }
    }
  }
}

// Maps offsets of src's operator store into dst's, interning src's blocks.
struct StoreMap {
  std::vector<int64_t> start, shift;  // sorted block starts in src, and the offset change
  void add(const std::vector<double>& src_store, const std::vector<std::pair<int64_t, int64_t>>& blocks,
           std::vector<double>& dst_store, std::unordered_map<uint64_t, std::vector<int64_t>>& idx,
           std::vector<std::pair<int64_t, int64_t>>& dst_blocks) {
    auto sorted = blocks;
    std::sort(sorted.begin(), sorted.end());
    for (auto& b : sorted) {
      const int64_t off = intern(dst_store, idx, src_store.data() + b.first, (size_t)b.second, dst_blocks);
      start.push_back(b.first);
      shift.push_back(off - b.first);
    }
  }
  // any offset inside a block moves with it
  int64_t operator()(int64_t off) const {
    if (off < 0) return off;
    const auto it = std::upper_bound(start.begin(), start.end(), off);
    return off + shift[size_t(it - start.begin()) - 1];
  }
};

void plan_merge(HostPlan& p, const HostPlan& q) {
  const int64_t sv = p.N, sp = p.part_size;
  StoreMap ma, mb, mh;
  ma.add(q.Ad, q.a_blocks, p.Ad, p.dedup_a, p.a_blocks);
  mb.add(q.Bh, q.b_blocks, p.Bh, p.dedup_b, p.b_blocks);
  mh.add(q.H, q.h_blocks, p.H, p.dedup_h, p.h_blocks);
  const int i_pairs = (int)p.pairs.size(), i_phases = (int)p.phases.size();
  const int i_aterms = (int)p.aterms.size(), i_rterms = (int)p.rterms.size();
  const int i_plist = (int)p.part_list.size();
  for (int64_t v : q.vbase) p.vbase.push_back(v + sv);
  if (!q.vmask.empty() || !p.vmask.empty()) {
    p.vmask.resize(size_t(sv), 1);
    if (q.vmask.empty())
      p.vmask.insert(p.vmask.end(), size_t(q.N), 1);
    else
      p.vmask.insert(p.vmask.end(), q.vmask.begin(), q.vmask.end());
  }
  for (auto a : q.inst_a_off) {
    for (auto& x : a) x = ma(x);
    p.inst_a_off.push_back(std::move(a));
  }
  for (auto b : q.inst_b_off) {
    for (auto& x : b) x = mb(x);
    p.inst_b_off.push_back(std::move(b));
  }
  p.pairs.insert(p.pairs.end(), q.pairs.begin(), q.pairs.end());
  for (DPhase ph : q.phases) {
    ph.v_off += sv;
    ph.b_off = mb(ph.b_off);
    ph.a_off = ma(ph.a_off);
    ph.pair_begin += i_pairs;
    ph.pair_end += i_pairs;
    p.phases.push_back(ph);
  }
  for (DItem2 it : q.items2) {
    it.part_off += sp;
    it.w_off += sv;
    it.ph_begin += i_phases;
    it.ph_end += i_phases;
    p.items2.push_back(it);
  }
  for (DRTerm r : q.rterms) {
    r.v_off += sv;
    r.b_off = mb(r.b_off);
    p.rterms.push_back(r);
  }
  for (DATerm t : q.aterms) {
    t.a_off = ma(t.a_off);
    t.r_begin += i_rterms;
    t.r_end += i_rterms;
    p.aterms.push_back(t);
  }
  for (DItem it : q.items) {
    it.part_off += sp;
    it.w_off += sv;
    it.a_begin += i_aterms;
    it.a_end += i_aterms;
    p.items.push_back(it);
  }
  for (int64_t o : q.part_list) p.part_list.push_back(o + sp);
  for (DBlock b : q.blocks) {
    b.v_off += sv;
    b.ha_off = mh(b.ha_off);
    b.hb_off = mh(b.hb_off);
    b.p_begin += i_plist;
    b.p_end += i_plist;
    p.blocks.push_back(b);
  }
  p.max_units = std::max(p.max_units, q.max_units);
  p.lds_y = std::max(p.lds_y, q.lds_y);
  p.lds_buf = std::max(p.lds_buf, q.lds_buf);
  p.max_lds = std::max(p.max_lds, q.max_lds);
  p.N += q.N;
  p.part_size += q.part_size;
  p.flops_ref += q.flops_ref;
  p.flops_like += q.flops_like;
  p.bytes_alg += q.bytes_alg;
}

}  // namespace nsx
