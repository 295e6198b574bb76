#include "nsx_plan.hpp"

#include <algorithm>

namespace nsx {

static int col_tile(int rows) {
  // widest column tile whose da x nc outputs fit one workgroup's registers
  int nc = kTileElems / std::max(rows, 1);
  return std::max(nc, 1);
}

void plan_add(HostPlan& p, const Model& m) {
  const int64_t vb = p.N;
  p.vbase.push_back(vb);
  // operator values of this instance
  std::vector<int64_t> a_off(m.adense.size()), b_off(m.bhat.size());
  for (size_t k = 0; k < m.adense.size(); ++k) {
    a_off[k] = (int64_t)p.Ad.size();
    p.Ad.insert(p.Ad.end(), m.adense[k].begin(), m.adense[k].end());
  }
  for (size_t k = 0; k < m.bhat.size(); ++k) {
    b_off[k] = (int64_t)p.Bh.size();
    p.Bh.insert(p.Bh.end(), m.bhat[k].begin(), m.bhat[k].end());
  }
  std::vector<int64_t> ha_off(m.A.nJ(), -1), hb_off(m.B.nJ(), -1);
  int64_t nnz_like = 0;
  for (int j = 0; j < m.A.nJ(); ++j)
    if (!m.A.H.empty() && !m.A.H[j].empty()) {
      ha_off[j] = (int64_t)p.H.size();
      p.H.insert(p.H.end(), m.A.H[j].begin(), m.A.H[j].end());
      for (double v : m.A.H[j]) nnz_like += v != 0.0;
    }
  for (int j = 0; j < m.B.nJ(); ++j)
    if (!m.B.H.empty() && !m.B.H[j].empty()) {
      hb_off[j] = (int64_t)p.H.size();
      p.H.insert(p.H.end(), m.B.H[j].begin(), m.B.H[j].end());
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
    int nc = std::min(col_tile(std::max(da, dar)), db);
    nc = std::min(nc, std::max(1, kMaxLdsDoubles / std::max(dar, 1)));
    for (int j0 = 0; j0 < db; j0 += nc) {
      DItem d;
      d.part_off = base;
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
      b.ha_off = m.azero ? -1 : ha_off[m.bas_a[L] / 2];
      b.hb_off = m.bzero ? -1 : hb_off[m.bas_b[L] / 2];
      b.da = da; b.db = db; b.j0 = j0; b.nc = std::min(nc, db - j0);
      b.p_begin = p_begin; b.p_end = p_end;
      p.blocks.push_back(b);
    }
  }
  p.N += m.N;
  p.flops_ref += m.flops_ref;
  p.flops_like += m.flops_like;
  // algorithmic bytes (FP64): read v, write w, read every operator value once
  int64_t nnzA = 0, nnzB = 0;
  for (const auto& d : m.A.dens) nnzA += d.count;
  for (const auto& d : m.B.dens) nnzB += d.count;
  p.bytes_alg += 16 * m.N + 8 * (nnzA + nnzB + nnz_like + m.n_bterms);
}

}  // namespace nsx
