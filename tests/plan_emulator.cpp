// Test helper (CPU only): executes the device work plan of libnsx.so with
// plain loops, so the plan construction (block lists, phase runs, Bhat^T
// layout, partial bookkeeping) is checked against the oracle without a GPU.
// It mirrors the arithmetic of the HIP kernels, not their scheduling.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../nushellx_amd/csrc/nsx_eig.hpp"
#include "../nushellx_amd/csrc/nsx_model.hpp"
#include "../nushellx_amd/csrc/nsx_plan.hpp"

using namespace nsx;

static thread_local std::string g_err;

extern "C" const char* emu_last_error() { return g_err.c_str(); }

// H*v of every instance of a plan, computed on the CPU from the plan arrays
// The pn part's input: v masked to the window states of truncated instances.
static const double* pn_input(const HostPlan& p, const double* vin, std::vector<double>& store) {
  if (p.vmask.empty()) return vin;
  store.assign(vin, vin + p.N);
  for (int64_t i = 0; i < p.N; ++i)
    if (!p.vmask[i]) store[i] = 0.0;
  return store.data();
}

static void emulate(const HostPlan& p, const double* vin_all, double* vout) {
  std::vector<double> vm;
  const double* vin = vin_all;            // like-particle input
  const double* vp = pn_input(p, vin_all, vm);
  std::vector<double> part(p.part_size, 0.0), buf;
  // staged items
  for (const DItem2& it : p.items2) {
    std::vector<double> W(size_t(it.da) * it.db, 0.0), Y(size_t(it.dar) * it.db, 0.0);
    for (int ph = it.ph_begin; ph < it.ph_end; ++ph) {
      const DPhase& P = p.phases[ph];
      buf.assign(P.nv + P.nb + P.na, 0.0);
      for (int i = 0; i < P.nv; ++i) buf[i] = vp[P.v_off + i];
      for (int i = 0; i < P.nb; ++i) buf[P.nv + i] = p.Bh[P.b_off + i];
      for (int i = 0; i < P.na; ++i) buf[P.nv + P.nb + i] = p.Ad[P.a_off + i];
      for (int q = P.pair_begin; q < P.pair_end; ++q) {
        const DPair& Q = p.pairs[q];
        for (int j = 0; j < it.db; ++j)
          for (int i = 0; i < it.dar; ++i) {
            double s = 0.0;
            for (int k = 0; k < Q.dbr; ++k) s += buf[Q.v_lds + i + it.dar * k] * buf[Q.b_lds + k + Q.dbr * j];
            Y[i + size_t(it.dar) * j] += Q.r * s;
          }
      }
      if (P.na > 0) {
        const double* A = buf.data() + P.nv + P.nb;
        for (int j = 0; j < it.db; ++j)
          for (int i = 0; i < it.da; ++i) {
            double s = 0.0;
            for (int k = 0; k < it.dar; ++k) s += A[i + it.da * k] * Y[k + size_t(it.dar) * j];
            W[i + size_t(it.da) * j] += s;
          }
        std::fill(Y.begin(), Y.end(), 0.0);
      }
    }
    for (size_t q = 0; q < W.size(); ++q) part[it.part_off + q] = W[q];
  }
  // direct items (large blocks), unless the large-block GEMM path takes them
  const bool big = !p.g_tile2.empty();
  for (const DItem& it : p.items) {
    if (big) break;
    for (int jj = 0; jj < it.nc; ++jj) {
      const int j = it.j0 + jj;
      std::vector<double> w(it.da, 0.0);
      for (int a = it.a_begin; a < it.a_end; ++a) {
        const DATerm& T = p.aterms[a];
        std::vector<double> y(it.dar, 0.0);
        for (int u = T.r_begin; u < T.r_end; ++u) {
          const DRTerm& R = p.rterms[u];
          for (int i = 0; i < it.dar; ++i) {
            double s = 0.0;
            for (int k = 0; k < R.dbr; ++k) s += vp[R.v_off + i + int64_t(it.dar) * k] * p.Bh[R.b_off + int64_t(R.dbr) * j + k];
            y[i] += R.r * s;
          }
        }
        for (int i = 0; i < it.da; ++i) {
          double s = 0.0;
          for (int k = 0; k < it.dar; ++k) s += p.Ad[T.a_off + i + int64_t(it.da) * k] * y[k];
          w[i] += s;
        }
      }
      for (int i = 0; i < it.da; ++i) part[it.part_off + i + int64_t(it.da) * j] = w[i];
    }
  }
  // reduce
  for (const DBlock& b : p.blocks) {
    for (int jj = 0; jj < b.nc; ++jj) {
      const int j = b.j0 + jj;
      for (int i = 0; i < b.da; ++i) {
        const int64_t ij = i + int64_t(b.da) * j;
        double w = 0.0;
        for (int q = b.p_begin; q < b.p_end; ++q) w += part[p.part_list[q] + ij];
        if (b.ha_off >= 0) {
          double s = 0.0;
          for (int k = 0; k < b.da; ++k) s += p.H[b.ha_off + i + int64_t(b.da) * k] * vin[b.v_off + k + int64_t(b.da) * j];
          w += s;
        }
        if (b.hb_off >= 0) {
          double s = 0.0;
          for (int k = 0; k < b.db; ++k) s += vin[b.v_off + i + int64_t(b.da) * k] * p.H[b.hb_off + j + int64_t(b.db) * k];
          w += s;
        }
        vout[b.v_off + ij] = w;
      }
    }
  }
  if (!big) return;
  // large-block path: X_u = V_R Bhat^T on the listed tiles and their active K
  // steps (X starts at zero), then W tiles += sum over the tile's terms and
  // their active K steps of A_k sum_u r_u X_u
  auto active = [](const uint64_t (&mk)[2], int k) { const int s = k / kBigStep; return (mk[s >> 6] >> (s & 63)) & 1u; };
  std::vector<double> XG(size_t(p.xg_size), 0.0);
  for (const GTile1& t : p.g_tile1) {
    const GTask1& g = p.g_task1[t.task];
    for (int j = t.j0; j < std::min(t.j0 + kBigTile, g.n); ++j)
      for (int i = t.i0; i < std::min(t.i0 + p.cfg.big_rows1, g.m); ++i) {
        double s = 0.0;
        for (int k = 0; k < g.k; ++k)
          if (active(t.kmask, k)) s += vp[g.v_off + i + int64_t(g.m) * k] * p.Bh[g.b_off + k + int64_t(g.k) * j];
        XG[g.x_off + i + int64_t(big_ldx(g.m)) * j] = s;
      }
  }
  for (const GTile2& t : p.g_tile2)
    for (int j = t.j0; j < std::min(t.j0 + kBigTile, t.db); ++j)
      for (int i = t.i0; i < std::min(t.i0 + kBigTile, t.da); ++i) {
        double s = 0.0;
        for (int q = t.t_begin; q < t.t_end; ++q) {
          const GTerm2& tm = p.g_term2[q];
          for (int k = 0; k < tm.dar; ++k) {
            if (!active(tm.kmask, k)) continue;
            double y = 0.0;
            for (int x = tm.x_begin; x < tm.x_end; ++x)
              y += p.g_xref[x].r * XG[p.g_xref[x].x_off + k + int64_t(big_ldx(tm.dar)) * j];
            s += p.Ad[tm.a_off + i + int64_t(t.da) * k] * y;
          }
        }
        vout[t.w_off + i + int64_t(t.da) * j] += s;
      }
}

extern "C" int emu_hv(const char* nus, const char* opa, const char* opb, const char* intf,
                      const double* vin, double* vout, int64_t n) {
  Model m;
  std::string e = build_model(nus, opa, opb, intf, m);
  if (!e.empty()) { g_err = e; return -1; }
  if (m.N != n) { g_err = "dimension mismatch"; return -2; }
  HostPlan p;
  plan_add(p, m);
  build_big(p, {&m});
  emulate(p, vin, vout);
  return 0;
}

// The plan with items whose side exceeds big_min on the large-block GEMM path
// (dop: optional .dop prefix of a separable model, or empty).
extern "C" int emu_hv_big(const char* nus, const char* opa, const char* opb, const char* intf, const char* dop,
                          int big_min, const double* vin, double* vout, int64_t n, int64_t* tiles) {
  Model m;
  std::string e = build_model(nus, opa, opb, intf, m, dop ? dop : "");
  if (!e.empty()) { g_err = e; return -1; }
  if (m.N != n) { g_err = "dimension mismatch"; return -2; }
  HostPlan p;
  p.cfg.big_min = big_min;
  plan_add(p, m);
  build_big(p, {&m});
  emulate(p, vin, vout);
  if (tiles) {
    tiles[0] = (int64_t)p.g_tile1.size();
    tiles[1] = (int64_t)p.g_tile2.size();
    tiles[2] = (int64_t)p.g_xref.size();
    tiles[3] = (int64_t)p.g_term2.size();
  }
  return 0;
}

// The separable (NuTRLZ) model of <dop>.dop through the same plan.
extern "C" int emu_hv_sep(const char* nus, const char* opa, const char* opb, const char* intf, const char* dop,
                          const double* vin, double* vout, int64_t n, double* flops) {
  Model m;
  std::string e = build_model(nus, opa, opb, intf, m, dop);
  if (!e.empty()) { g_err = e; return -1; }
  if (m.N != n) { g_err = "dimension mismatch"; return -2; }
  HostPlan p;
  plan_add(p, m);
  emulate(p, vin, vout);
  if (flops) *flops = m.flops_ref;
  return 0;
}

// A batch of instances planned into one plan (merge = 0) or planned one by one
// and merged (merge = 1); args holds 4 file names per instance.
extern "C" int emu_hv_batch(int n, const char* const* args, int merge, const double* vin, double* vout,
                            int64_t ntot, int64_t* sizes) {
  std::vector<Model> ms(n);
  for (int i = 0; i < n; ++i) {
    std::string e = build_model(args[4 * i], args[4 * i + 1], args[4 * i + 2], args[4 * i + 3], ms[i]);
    if (!e.empty()) { g_err = e; return -1; }
  }
  HostPlan p;
  for (int i = 0; i < n; ++i) {
    if (merge) {
      HostPlan q;
      q.cfg = p.cfg;
      plan_add(q, ms[i]);
      plan_merge(p, q);
    } else {
      plan_add(p, ms[i]);
    }
  }
  if (p.N != ntot) { g_err = "dimension mismatch"; return -2; }
  emulate(p, vin, vout);
  sizes[0] = (int64_t)p.Ad.size();
  sizes[1] = (int64_t)p.Bh.size();
  sizes[2] = (int64_t)p.H.size();
  return 0;
}

// plan statistics for tests and tuning: items2, phases, pairs, direct items, blocks, partial doubles
extern "C" int emu_plan_stats(const char* nus, const char* opa, const char* opb, const char* intf,
                              int64_t* out8) {
  Model m;
  std::string e = build_model(nus, opa, opb, intf, m);
  if (!e.empty()) { g_err = e; return -1; }
  HostPlan p;
  plan_add(p, m);
  out8[0] = (int64_t)p.items2.size();
  out8[1] = (int64_t)p.phases.size();
  out8[2] = (int64_t)p.pairs.size();
  out8[3] = (int64_t)p.items.size();
  out8[4] = (int64_t)p.blocks.size();
  out8[5] = p.part_size;
  out8[6] = p.lds_buf;
  out8[7] = (int64_t)(p.Ad.size() + p.Bh.size() + p.H.size());
  return 0;
}

// the product's recoupling coefficients, for the golden-value tests
extern "C" double emu_sixj(int a, int b, int c, int d, int e, int f) { return nsx::sixj(a, b, c, d, e, f); }

extern "C" int emu_sym_eig(const double* A, int n, int full, double* w, double* Z) {
  return nsx::sym_eig(A, n, full != 0, w, Z) ? 0 : 1;
}

// Matrix-core stacked path, emulated lane by lane as the k_mx1 / k_mx2
// kernels run it (16 x 16 x 4 FP64 steps with their operand and result lane
// layouts, X transposed per product), with every memory index checked against
// its array.  Returns -4 on an out-of-range access.
namespace {
struct Mfma16 {  // one 16 x 16 x 4 FP64 step from the 64 lanes' operands
  static void step(const double (&a)[64], const double (&b)[64], double (&acc)[64][4]) {
    double A[16][4], B[4][16];
    for (int l = 0; l < 64; ++l) { A[l % 16][l / 16] = a[l]; B[l / 16][l % 16] = b[l]; }
    for (int l = 0; l < 64; ++l)
      for (int v = 0; v < 4; ++v) {
        const int row = l / 16 + 4 * v, col = l % 16;
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += A[row][k] * B[k][col];
        acc[l][v] += s;
      }
  }
  static int row(int l, int v) { return l / 16 + 4 * v; }
};
}  // namespace

extern "C" int emu_hv_mstk(const char* nus, const char* opa, const char* opb, const char* intf,
                           const double* vin, double* vout, int64_t n, int units1, int terms2, int big_min,
                           int64_t* stats4) {
  Model m;
  std::string e = build_model(nus, opa, opb, intf, m);
  if (!e.empty()) { g_err = e; return -1; }
  if (m.N != n) { g_err = "dimension mismatch"; return -2; }
  HostPlan p;
  p.cfg.stacked = 3;
  p.cfg.mx_units1 = units1;
  p.cfg.mx_terms2 = terms2;
  if (big_min > 0) p.cfg.big_min = big_min;
  plan_add(p, m);
  std::vector<const Model*> ms{&m};
  build_mstk(p, ms);
  build_big(p, ms);
  if (!p.g_tile2.empty() || (!p.items.empty() && p.g_task1.empty() && p.cfg.big_min < kSmallMax)) {
    // large blocks are not emulated here
  }
  std::vector<double> vm;
  const double* vp = pn_input(p, vin, vm);
  std::vector<double> X(p.x_size, std::nan(""));
  bool oob = false;
  auto chk = [&](int64_t i, int64_t size) { if (i < 0 || i >= size) oob = true; return i >= 0 && i < size; };
  const int KS = p.mx_ks <= 6 ? 6 : 10;
  if (p.mx_ks > 10) { g_err = "inner size above 40"; return -3; }
  for (const MTask1& t : p.m_task1) {
    const int nks = (t.dbr + 3) >> 2;
    double b[10][64];
    for (int l = 0; l < 64; ++l) {
      const int kq = l >> 4, j = t.j0 + (l & 15);
      for (int ks = 0; ks < KS; ++ks) {
        const int k = 4 * ks + kq;
        b[ks][l] = 0.0;
        const bool ok = k < t.dbr && j < t.db;
        const int64_t ix = t.b_off + (ok ? k + int64_t(t.dbr) * j : 0);
        if (ks < nks && chk(ix, (int64_t)p.Bh.size()) && ok) b[ks][l] = p.Bh[ix];
      }
    }
    for (int u = t.u_begin; u < t.u_end; ++u) {
      const MUnit1& s = p.m_unit1[u];
      if (s.rows < 1 || s.rows > kMxTile) { g_err = "row tile out of range"; return -3; }
      double a[10][64], acc[64][4] = {};
      for (int l = 0; l < 64; ++l)
        for (int ks = 0; ks < KS; ++ks) {
          const int k = 4 * ks + (l >> 4), i = l & 15;
          const bool ok = k < t.dbr && i < s.rows;
          // the kernel reads the tile's first element for lanes outside it
          const int64_t ix = s.v_off + (ok ? i + int64_t(s.ld) * k : 0);
          a[ks][l] = 0.0;
          if (ks < nks && chk(ix, n) && ok) a[ks][l] = vp[ix];
        }
      for (int ks = 0; ks < nks; ++ks) Mfma16::step(a[ks], b[ks], acc);
      for (int l = 0; l < 64; ++l)
        for (int v = 0; v < 4; ++v) {
          const int r = Mfma16::row(l, v), j = t.j0 + (l & 15);
          if (r < s.rows && j < t.db) {
            const int64_t ix = s.x_off + j + int64_t(t.db) * r;
            if (chk(ix, p.x_size)) X[ix] = acc[l][v];
          }
        }
    }
  }
  std::fill(vout, vout + n, 0.0);
  for (const DBlock& b : p.blocks)
    for (int jj = 0; jj < b.nc; ++jj) {
      const int j = b.j0 + jj;
      for (int i = 0; i < b.da; ++i) {
        double w = 0.0;
        if (b.ha_off >= 0)
          for (int k = 0; k < b.da; ++k) w += p.H[b.ha_off + i + int64_t(b.da) * k] * vin[b.v_off + k + int64_t(b.da) * j];
        if (b.hb_off >= 0)
          for (int k = 0; k < b.db; ++k) w += vin[b.v_off + i + int64_t(b.da) * k] * p.H[b.hb_off + j + int64_t(b.db) * k];
        vout[b.v_off + i + int64_t(b.da) * j] = w;
      }
    }
  int64_t writes = 0;
  for (const MTask2& t : p.m_task2) {
    double accw[4][64][4] = {};
    for (int wave = 0; wave < 4; ++wave)
      for (int tt = t.t_begin + wave; tt < t.t_end; tt += 4) {
        const STerm2& s = p.s_term2[tt];
        const int nks = (s.dar + 3) >> 2;
        double a[10][64], y[10][64];
        for (int l = 0; l < 64; ++l) {
          const int kq = l >> 4, j = t.j0 + (l & 15), i = t.i0 + (l & 15);
          for (int ks = 0; ks < KS; ++ks) {
            const int k = 4 * ks + kq;
            a[ks][l] = 0.0;
            y[ks][l] = 0.0;
            const bool aok = k < s.dar && i < t.da;
            const int64_t ia = s.a_off + (aok ? i + int64_t(t.da) * k : 0);
            if (ks < nks && chk(ia, (int64_t)p.Ad.size()) && aok) a[ks][l] = p.Ad[ia];
            const bool xok = k < s.dar && j < t.db;
            for (int x = s.x_begin; x < s.x_end; ++x) {
              const int64_t ix = p.s_xref[x].x_off + (xok ? j + int64_t(t.db) * k : 0);
              if (ks < nks && chk(ix, p.x_size) && xok) y[ks][l] += p.s_xref[x].r * X[ix];
            }
          }
        }
        for (int ks = 0; ks < nks; ++ks) Mfma16::step(y[ks], a[ks], accw[wave]);
      }
    for (int l = 0; l < 64; ++l)
      for (int v = 0; v < 4; ++v) {
        const int jr = t.j0 + Mfma16::row(l, v), i = t.i0 + (l & 15);
        if (jr < t.db && i < t.da) {
          const int64_t ix = t.w_off + i + int64_t(t.da) * jr;
          if (chk(ix, n)) vout[ix] += (accw[0][l][v] + accw[1][l][v]) + (accw[2][l][v] + accw[3][l][v]);
          ++writes;
        }
      }
  }
  if (stats4) {
    stats4[0] = (int64_t)p.m_task1.size();
    stats4[1] = (int64_t)p.m_task2.size();
    stats4[2] = p.x_size;
    stats4[3] = writes;
  }
  if (oob) { g_err = "out-of-range access in the matrix-core stacked path"; return -4; }
  if (!p.items.empty()) { g_err = "large-block items are not emulated by emu_hv_mstk"; return -5; }
  return 0;
}

// Fused matrix-core path, emulated lane by lane as k_mxf runs it: per step
// the X_u tiles (16 x 16 x 4 FP64 steps), Y_s += r_su X_u in the result
// layout, W_s^T += Y_s^T A_s^T with Y's result registers as the first operand's
// K steps, the four wavefronts' tiles summed; every memory index checked.
extern "C" int emu_hv_mxf(const char* nus, const char* opa, const char* opb, const char* intf,
                          const double* vin, double* vout, int64_t n, int steps_per_wave, int slots,
                          int like, int det, int64_t* stats4) {
  Model m;
  std::string e = build_model(nus, opa, opb, intf, m);
  if (!e.empty()) { g_err = e; return -1; }
  if (m.N != n) { g_err = "dimension mismatch"; return -2; }
  HostPlan p;
  p.cfg.stacked = 4;
  p.cfg.mxf_steps = steps_per_wave;
  p.cfg.mxf_slots = slots;
  p.cfg.mxf_like = like;
  p.cfg.deterministic = det;
  plan_add(p, m);
  std::vector<const Model*> ms{&m};
  if (!build_mxf(p, ms)) { g_err = "build_mxf refused the model"; return -6; }
  if (p.mxf_det != (det != 0)) { g_err = "deterministic tables not built"; return -8; }
  const int sl = p.mxf_slots;
  // deterministic mode: the tasks' partial tiles (element row + 16 col)
  std::vector<double> P(p.mxf_det ? p.f_mtask.size() * size_t(sl) * 256 : 0, 0.0);
  if (p.mx_ks > 10) { g_err = "inner size above 40"; return -3; }
  std::vector<double> vm;
  const double* vp = pn_input(p, vin, vm);
  bool oob = false;
  auto chk = [&](int64_t i, int64_t size) { if (i < 0 || i >= size) oob = true; return i >= 0 && i < size; };
  std::fill(vout, vout + n, 0.0);
  // the like-particle part: stored first by k_reduce, or folded into the tasks
  if (!p.mxf_like)
  for (const DBlock& b : p.blocks)
    for (int jj = 0; jj < b.nc; ++jj) {
      const int j = b.j0 + jj;
      for (int i = 0; i < b.da; ++i) {
        double w = 0.0;
        if (b.ha_off >= 0)
          for (int k = 0; k < b.da; ++k) w += p.H[b.ha_off + i + int64_t(b.da) * k] * vin[b.v_off + k + int64_t(b.da) * j];
        if (b.hb_off >= 0)
          for (int k = 0; k < b.db; ++k) w += vin[b.v_off + i + int64_t(b.da) * k] * p.H[b.hb_off + j + int64_t(b.db) * k];
        vout[b.v_off + i + int64_t(b.da) * j] = w;
      }
    }
  int64_t steps = 0, prods = 0;
  // the kernel's workgroup -> task map: workgroup b on XCD b % 8 takes the
  // (b / 8)-th task of that XCD's range; the ranges must partition the tasks
  const int ntask = (int)p.f_mtask.size();
  int most = 0;
  for (int x = 0; x < 8; ++x) {
    if (p.mxf_xcd[x + 1] < p.mxf_xcd[x]) { g_err = "XCD ranges out of order"; return -7; }
    most = std::max(most, p.mxf_xcd[x + 1] - p.mxf_xcd[x]);
  }
  if (p.mxf_xcd[0] != 0 || p.mxf_xcd[8] != ntask) { g_err = "XCD ranges do not cover the tasks"; return -7; }
  std::vector<int> visits(size_t(ntask), 0);
  for (int wg = 0; wg < 8 * most; ++wg) {
    const int x = wg % 8, ti = p.mxf_xcd[x] + wg / 8;
    if (ti >= p.mxf_xcd[x + 1]) continue;
    ++visits[size_t(ti)];
    const MFTask& t = p.f_mtask[size_t(ti)];
    double Wt[4][kMfSlots][64][4] = {};
    for (int wave = 0; wave < 4; ++wave)
      for (int st = t.p[wave]; st < t.p[wave + 1]; ++st) {
        const MFStep& S = p.f_mstep[st];
        ++steps;
        double Y[kMfSlots][64][4] = {};
        for (int u = S.u_begin; u < S.u_end; ++u) {
          const MFU& U = p.f_mu[u];
          ++prods;
          const int nk = (U.dbr + 3) >> 2;
          double X[64][4] = {};
          for (int ks = 0; ks < nk; ++ks) {
            double a[64], b[64];
            for (int l = 0; l < 64; ++l) {
              const int k = 4 * ks + (l >> 4), l16 = l & 15;
              const bool okv = k < U.dbr && l16 < t.rows_a;
              const int64_t iv = U.v_off + (okv ? l16 + int64_t(t.ld_v) * k : 0);
              a[l] = (chk(iv, n) && okv) ? vp[iv] : 0.0;
              const bool okb = k < U.dbr && l16 < t.cols_j;
              const int64_t ib = U.b_off + (okb ? l16 + int64_t(t.db) * k : 0);
              b[l] = (chk(ib, (int64_t)p.Bn.size()) && okb) ? p.Bn[ib] : 0.0;
            }
            Mfma16::step(a, b, X);
          }
          for (int q = 0; q < kMfSlots; ++q)
            for (int l = 0; l < 64; ++l)
              for (int v = 0; v < 4; ++v) Y[q][l][v] += U.r[q] * X[l][v];
        }
        for (int q = 0; q < kMfSlots; ++q) {
          if (S.a_off[q] < 0) continue;
          for (int ks = 0; ks < 4; ++ks) {
            double a[64], b[64];
            for (int l = 0; l < 64; ++l) {
              const int kp = Mfma16::row(l, ks), l16 = l & 15;
              a[l] = Y[q][l][ks];
              const bool ok = kp < t.rows_a && l16 < t.rows_i[q];
              const int64_t ia = S.a_off[q] + (ok ? l16 + int64_t(t.da[q]) * kp : 0);
              b[l] = (chk(ia, (int64_t)p.Ad.size()) && ok) ? p.Ad[ia] : 0.0;
            }
            Mfma16::step(a, b, Wt[wave][q]);
          }
        }
      }
    for (int q = 0; q < kMfSlots; ++q) {
      if (t.w_off[q] < 0) continue;
      for (int l = 0; l < 64; ++l)
        for (int v = 0; v < 4; ++v) {
          const int jr = Mfma16::row(l, v), i = l & 15;
          if (p.mxf_det) {
            if (q >= sl) { oob = true; continue; }
            P[(size_t(ti) * sl + q) * 256 + i + 16 * jr] = (Wt[0][q][l][v] + Wt[1][q][l][v]) + (Wt[2][q][l][v] + Wt[3][q][l][v]);
          } else if (i < t.rows_i[q] && jr < t.cols_j) {
            const int64_t iw = t.w_off[q] + i + int64_t(t.da[q]) * jr;
            if (chk(iw, n))
              vout[iw] += (Wt[0][q][l][v] + Wt[1][q][l][v]) + (Wt[2][q][l][v] + Wt[3][q][l][v]);
          }
        }
    }
  }
  if (p.mxf_det) {  // k_mxf_sum: each W tile's partials in its fixed order, each partial of a used slot once
    std::vector<int> used(P.size() / 256, 0);
    for (const MFDTile& d : p.f_dtile) {
      for (int e = 0; e < 256; ++e) {
        const int r = e & 15, c = e >> 4;
        if (r >= d.rows || c >= d.cols) continue;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};  // partial n of the list into accumulator n % 4
        for (int k = d.c_begin; k < d.c_end; ++k)
          if (chk(p.f_dpart[size_t(k)], (int64_t)used.size()))
            acc[(k - d.c_begin) & 3] += P[size_t(p.f_dpart[size_t(k)]) * 256 + e];
        const int64_t iw = d.w_off + r + int64_t(d.da) * c;
        if (chk(iw, n)) vout[iw] += (acc[0] + acc[1]) + (acc[2] + acc[3]);
      }
      for (int k = d.c_begin; k < d.c_end; ++k)
        if (chk(p.f_dpart[size_t(k)], (int64_t)used.size())) ++used[size_t(p.f_dpart[size_t(k)])];
    }
    for (size_t i = 0; i < p.f_mtask.size(); ++i)
      for (int q = 0; q < sl; ++q)
        if ((p.f_mtask[i].w_off[q] >= 0) != (used[i * sl + q] == 1)) { g_err = "a partial tile not summed exactly once"; return -8; }
  }
  if (stats4) {
    stats4[0] = (int64_t)p.f_mtask.size();
    stats4[1] = steps;
    stats4[2] = prods;
    stats4[3] = p.mx_ks + (p.mxf_like ? 100 : 0);
  }
  for (int v : visits)
    if (v != 1) { g_err = "a task not run exactly once by the workgroup map"; return -7; }
  if (oob) { g_err = "out-of-range access in the fused matrix-core path"; return -4; }
  if (!p.items.empty()) { g_err = "large-block items are not emulated by emu_hv_mxf"; return -5; }
  return 0;
}

// Stacked two-stage path: X = V Bhat^T per distinct pair, then W += A sum r X,
// plus the like-particle part and the direct (large-block) items, all from the
// plan arrays the stacked kernels read.
extern "C" int emu_hv_stacked(const char* nus, const char* opa, const char* opb, const char* intf,
                              const double* vin, double* vout, int64_t n, int cap, int work) {
  Model m;
  std::string e = build_model(nus, opa, opb, intf, m);
  if (!e.empty()) { g_err = e; return -1; }
  if (m.N != n) { g_err = "dimension mismatch"; return -2; }
  HostPlan p;
  p.cfg.stacked = 1;
  p.cfg.stk_cap = cap;
  p.cfg.stk_work = work;
  plan_add(p, m);
  std::vector<const Model*> ms{&m};
  build_stacked(p, ms);
  std::vector<double> vm;
  const double* vp = pn_input(p, vin, vm);
  std::vector<double> X(p.x_size, 0.0);
  for (const SChunk1& c : p.s_chunk1) {
    if (c.lds > p.s_lds1) { g_err = "chunk LDS above the plan maximum"; return -3; }
    for (int u = c.u_begin; u < c.u_end; ++u) {
      const SUnit1& U = p.s_unit1[u];
      for (int j = 0; j < c.db; ++j)
        for (int i = 0; i < U.dar; ++i) {
          double s = 0.0;
          for (int k = 0; k < c.dbr; ++k) s += vp[U.v_off + i + int64_t(U.dar) * k] * p.Bh[c.b_off + k + int64_t(c.dbr) * j];
          X[U.x_off + i + int64_t(U.dar) * j] = s;
        }
    }
  }
  std::fill(vout, vout + n, 0.0);
  // like-particle part (reduce kernel without partials)
  for (const DBlock& b : p.blocks)
    for (int jj = 0; jj < b.nc; ++jj) {
      const int j = b.j0 + jj;
      for (int i = 0; i < b.da; ++i) {
        double w = 0.0;
        if (b.ha_off >= 0)
          for (int k = 0; k < b.da; ++k) w += p.H[b.ha_off + i + int64_t(b.da) * k] * vin[b.v_off + k + int64_t(b.da) * j];
        if (b.hb_off >= 0)
          for (int k = 0; k < b.db; ++k) w += vin[b.v_off + i + int64_t(b.da) * k] * p.H[b.hb_off + j + int64_t(b.db) * k];
        vout[b.v_off + i + int64_t(b.da) * j] = w;
      }
    }
  for (const SChunk2& c : p.s_chunk2) {
    if (c.lds > p.s_lds2) { g_err = "chunk LDS above the plan maximum"; return -3; }
    std::vector<double> Wc(size_t(c.da) * c.db, 0.0);
    for (int t = c.t_begin; t < c.t_end; ++t) {
      const STerm2& T = p.s_term2[t];
      std::vector<double> Y(size_t(T.dar) * c.db, 0.0);
      for (int x = T.x_begin; x < T.x_end; ++x)
        for (size_t q = 0; q < Y.size(); ++q) Y[q] += p.s_xref[x].r * X[p.s_xref[x].x_off + q];
      for (int j = 0; j < c.db; ++j)
        for (int i = 0; i < c.da; ++i) {
          double s = 0.0;
          for (int k = 0; k < T.dar; ++k) s += p.Ad[T.a_off + i + int64_t(c.da) * k] * Y[k + size_t(T.dar) * j];
          Wc[i + size_t(c.da) * j] += s;
        }
    }
    for (size_t q = 0; q < Wc.size(); ++q) vout[c.w_off + q] += Wc[q];
  }
  // direct items (large blocks) add into W
  for (const DItem& it : p.items)
    for (int jj = 0; jj < it.nc; ++jj) {
      const int j = it.j0 + jj;
      std::vector<double> w(it.da, 0.0);
      for (int a = it.a_begin; a < it.a_end; ++a) {
        const DATerm& T = p.aterms[a];
        std::vector<double> y(it.dar, 0.0);
        for (int u = T.r_begin; u < T.r_end; ++u) {
          const DRTerm& R = p.rterms[u];
          for (int i = 0; i < it.dar; ++i) {
            double s = 0.0;
            for (int k = 0; k < R.dbr; ++k) s += vp[R.v_off + i + int64_t(it.dar) * k] * p.Bh[R.b_off + int64_t(R.dbr) * j + k];
            y[i] += R.r * s;
          }
        }
        for (int i = 0; i < it.da; ++i) {
          double s = 0.0;
          for (int k = 0; k < it.dar; ++k) s += p.Ad[T.a_off + i + int64_t(it.da) * k] * y[k];
          w[i] += s;
        }
      }
      for (int i = 0; i < it.da; ++i) vout[it.w_off + i + int64_t(it.da) * j] += w[i];
    }
  return 0;
}

// Fused path: every task of the plan replayed in the kernel's order (rounds of
// staging, X = V Bhat^T, Y = sum r X, W += A Y into the task's column outputs),
// through an LDS image of the round, then the reduce over the column partials
// with the like-particle part, and the large-block items added on top.
// stats (optional, 5 values): tasks, rounds, largest round LDS, partial doubles, largest task units.
extern "C" int emu_hv_fused(const char* nus, const char* opa, const char* opb, const char* intf,
                            const double* vin, double* vout, int64_t n, int cap, int work, int threads,
                            int64_t* stats) {
  Model m;
  std::string e = build_model(nus, opa, opb, intf, m);
  if (!e.empty()) { g_err = e; return -1; }
  if (m.N != n) { g_err = "dimension mismatch"; return -2; }
  HostPlan p;
  p.cfg.stacked = 2;
  p.cfg.stk_cap = cap;
  p.cfg.stk_work = work;
  if (const char* v = std::getenv("NSX_BIG_MIN")) p.cfg.big_min = std::min(kSmallMax, std::max(1, std::atoi(v)));
  plan_add(p, m);
  std::vector<const Model*> ms{&m};
  build_big(p, ms);
  if (!build_fused(p, ms, threads)) { g_err = "fused plan not possible"; return -4; }
  std::vector<double> vm;
  const double* vp = pn_input(p, vin, vm);
  std::vector<double> part(p.f_part_size, 0.0);
  std::vector<double> lds(p.f_lds + 1);
  for (const FTask& t : p.f_task) {
    const int ncb = (t.db + kColBlk - 1) / kColBlk;
    if (t.units > kFMaxU * threads) { g_err = "task units above the kernel's limit"; return -3; }
    std::vector<double> acc(size_t(t.units) * kColBlk, 0.0);
    for (int r = t.r_begin; r < t.r_end; ++r) {
      const FRound& R = p.f_round[r];
      std::fill(lds.begin(), lds.end(), std::nan(""));
      if (R.x_end - R.x_begin > kFMaxX || R.c_end - R.c_begin > kFMaxC || R.f_end - R.f_begin > kFMaxRef ||
          R.s_end - R.s_begin > kFMaxSeg || t.r_end - t.r_begin > kFMaxRounds) {
        g_err = "round tables above the kernel's limits";
        return -3;
      }
      for (int s = R.s_begin; s < R.s_end; ++s) {
        const FSeg& g = p.f_seg[s];
        const int len = (s + 1 < R.s_end ? p.f_seg[s + 1].pre : R.stage) - g.pre;
        if (g.lds < 0 || len <= 0 || g.lds + len > p.f_lds) { g_err = "segment outside the round's LDS"; return -3; }
        const double* src = g.kind == 0 ? vp : g.kind == 1 ? p.Bh.data() : p.Ad.data();
        for (int k = 0; k < len; ++k) lds[g.lds + k] = src[g.src + k];
      }
      for (int x = R.x_begin; x < R.x_end; ++x) {
        const FX& X = p.f_x[x];
        for (int w = 0; w < X.dar * ncb; ++w) {
          const int i = w % X.dar, j0 = (w / X.dar) * kColBlk;
          for (int q = 0; q < kColBlk && j0 + q < t.db; ++q) {
            double sum = 0.0;
            for (int k = 0; k < X.dbr; ++k) sum += lds[X.v_lds + i + X.dar * k] * lds[X.b_lds + k + X.dbr * (j0 + q)];
            lds[X.x_lds + i + X.dar * (j0 + q)] = sum;
          }
        }
      }
      // Y overwrites the staged V / Bhat^T: form every Y from X first, as the kernel's barrier does
      std::vector<std::vector<double>> Y;
      for (int c = R.c_begin; c < R.c_end; ++c) {
        const FCons& C = p.f_cons[c];
        std::vector<double> y(size_t(C.dar) * t.db, 0.0);
        for (int f = C.f_begin; f < C.f_end; ++f) {
          const FRef& F = p.f_ref[R.f_begin + f];
          for (size_t q = 0; q < y.size(); ++q) y[q] += F.r * lds[F.x_lds + q];
        }
        Y.push_back(std::move(y));
      }
      for (int c = R.c_begin; c < R.c_end; ++c) {
        const FCons& C = p.f_cons[c];
        for (size_t q = 0; q < Y[c - R.c_begin].size(); ++q) lds[C.y_lds + q] = Y[c - R.c_begin][q];
      }
      // W items (consumer, row, 4 columns) write their partial sums over the dead X space
      for (int c = R.c_begin; c < R.c_end; ++c) {
        const FCons& C = p.f_cons[c];
        for (int lu = 0; lu < C.da * ncb; ++lu) {
          const int it = C.w0 + lu, i = lu % C.da, j0 = (lu / C.da) * kColBlk;
          if (R.wp_lds + kColBlk * (it + 1) > p.f_lds) { g_err = "W partials outside the round's LDS"; return -3; }
          for (int q = 0; q < kColBlk; ++q) {
            double sum = 0.0;
            if (j0 + q < t.db)
              for (int k = 0; k < C.dar; ++k) sum += lds[C.a_lds + i + C.da * k] * lds[C.y_lds + k + C.dar * (j0 + q)];
            lds[R.wp_lds + kColBlk * it + q] = sum;
          }
        }
      }
      for (int u = 0; u < t.units; ++u) {
        int b = t.b_begin;
        while (b + 1 < t.b_end && p.f_blk[b + 1].unit0 <= u) ++b;
        const FBlk& B = p.f_blk[b];
        const int lu = u - B.unit0, j0 = (lu / B.da) * kColBlk;
        for (int c = R.cb[b - t.b_begin]; c < R.cb[b - t.b_begin + 1]; ++c) {
          const FCons& C = p.f_cons[R.c_begin + c];
          if (C.da != B.da) { g_err = "consumer rows differ from its block's"; return -3; }
          for (int q = 0; q < kColBlk && j0 + q < t.db; ++q) acc[size_t(u) * kColBlk + q] += lds[R.wp_lds + kColBlk * (C.w0 + lu) + q];
        }
      }
    }
    for (int u = 0; u < t.units; ++u) {
      int b = t.b_begin;
      while (b + 1 < t.b_end && p.f_blk[b + 1].unit0 <= u) ++b;
      const FBlk& B = p.f_blk[b];
      const int lu = u - B.unit0, i = lu % B.da, j0 = (lu / B.da) * kColBlk;
      for (int q = 0; q < kColBlk && j0 + q < t.db; ++q)
        part[t.part_off + B.wc_off + i + B.da * (j0 + q)] = acc[size_t(u) * kColBlk + q];
    }
  }
  std::fill(vout, vout + n, 0.0);
  for (const DBlock& b : p.f_blocks)
    for (int jj = 0; jj < b.nc; ++jj) {
      const int j = b.j0 + jj;
      for (int i = 0; i < b.da; ++i) {
        const int64_t ij = i + int64_t(b.da) * j;
        double w = 0.0;
        for (int q = b.p_begin; q < b.p_end; ++q) w += part[p.f_part_list[q] + ij];
        if (b.ha_off >= 0)
          for (int k = 0; k < b.da; ++k) w += p.H[b.ha_off + i + int64_t(b.da) * k] * vin[b.v_off + k + int64_t(b.da) * j];
        if (b.hb_off >= 0)
          for (int k = 0; k < b.db; ++k) w += vin[b.v_off + i + int64_t(b.da) * k] * p.H[b.hb_off + j + int64_t(b.db) * k];
        vout[b.v_off + ij] = w;
      }
    }
  // large-block items (the GEMM tiles on the GPU) add into W
  for (const DItem& it : p.items)
    for (int jj = 0; jj < it.nc; ++jj) {
      const int j = it.j0 + jj;
      std::vector<double> w(it.da, 0.0);
      for (int a = it.a_begin; a < it.a_end; ++a) {
        const DATerm& T = p.aterms[a];
        std::vector<double> y(it.dar, 0.0);
        for (int u = T.r_begin; u < T.r_end; ++u) {
          const DRTerm& R = p.rterms[u];
          for (int i = 0; i < it.dar; ++i) {
            double s = 0.0;
            for (int k = 0; k < R.dbr; ++k) s += vp[R.v_off + i + int64_t(it.dar) * k] * p.Bh[R.b_off + int64_t(R.dbr) * j + k];
            y[i] += R.r * s;
          }
        }
        for (int i = 0; i < it.da; ++i) {
          double s = 0.0;
          for (int k = 0; k < it.dar; ++k) s += p.Ad[T.a_off + i + int64_t(it.da) * k] * y[k];
          w[i] += s;
        }
      }
      for (int i = 0; i < it.da; ++i) vout[it.w_off + i + int64_t(it.da) * j] += w[i];
    }
  if (const char* v = std::getenv("NSX_EMU_ROUND_STATS")) {
    int mx = 0, mc = 0, mf = 0, ms = 0, mr = 0;
    for (const FRound& R : p.f_round) {
      mx = std::max(mx, R.x_end - R.x_begin); mc = std::max(mc, R.c_end - R.c_begin);
      mf = std::max(mf, R.f_end - R.f_begin); ms = std::max(ms, R.s_end - R.s_begin);
    }
    for (const FTask& t : p.f_task) mr = std::max(mr, t.r_end - t.r_begin);
    std::fprintf(stderr, "round max: X %d cons %d refs %d segs %d; rounds per task %d\n", mx, mc, mf, ms, mr);
  }
  if (stats) {
    stats[0] = (int64_t)p.f_task.size();
    stats[1] = (int64_t)p.f_round.size();
    stats[2] = p.f_lds;
    stats[3] = p.f_part_size;
    stats[4] = p.f_units;
  }
  return 0;
}
