// Reduced-precision operand copies (FP32 or BF16 hi/lo) of a host plan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <tuple>
#include <vector>

#include "nsx_kernels.hpp"

namespace nsx {

template <typename T>
static std::string upload_vec(const std::vector<T>& v, T*& dst) {
  dst = nullptr;
  if (v.empty()) return "";
  if (hipMalloc(&dst, sizeof(T) * v.size()) != hipSuccess) return "hipMalloc failed (panel records)";
  if (hipMemcpy(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
    return "hipMemcpy failed (panel records)";
  return "";
}

// 256-unit chunks of panel blocks
static std::vector<PanelChunk> panel_chunks(const std::vector<PanelBlk>& blks) {
  std::vector<PanelChunk> c;
  for (size_t b = 0; b < blks.size(); ++b) {
    const int64_t units = int64_t(blks[b].kcols ? blks[b].rows : blks[b].cols) * blks[b].ldq;
    for (int64_t u = 0; u < units; u += 256) c.push_back(PanelChunk{(int32_t)b, (int32_t)u});
  }
  return c;
}

// The large-block tiles' operands as BF16 panels (see PanelBlk): one panel
// per distinct Bhat^T block and V_R block of the stage-1 tasks and per
// distinct A block of the stage-2 terms; the Bhat^T and A panels are made
// here from the Bf16x2 stores, the V_R panels by every H*v.
static std::string build_panels(const HostPlan& hp, LowpPlan& lp) {
  auto ldq_of = [](int K) { return (K + kBigStep - 1) / kBigStep * 4; };
  std::vector<PanelBlk> bblk, vblk, ablk;
  std::map<std::tuple<int64_t, int, int>, int64_t> bmap, vmap, amap;  // (offset, rows, cols) -> unit offset
  int64_t bu = 0, vu = 0, au = 0;
  auto intern = [&](std::map<std::tuple<int64_t, int, int>, int64_t>& mp, std::vector<PanelBlk>& blks, int64_t& used,
                    int64_t off, int rows, int cols, int kcols) {
    const auto key = std::make_tuple(off, rows, cols);
    auto it = mp.find(key);
    if (it != mp.end()) return it->second;
    const int ldq = ldq_of(kcols ? cols : rows);
    blks.push_back(PanelBlk{off, used, rows, cols, kcols, ldq});
    mp[key] = used;
    const int64_t at = used;
    used += int64_t(kcols ? rows : cols) * ldq;
    return at;
  };
  std::vector<GTask1> t1 = hp.g_task1;
  for (GTask1& g : t1) {
    g.b_off = intern(bmap, bblk, bu, g.b_off, g.k, g.n, 0);  // Bhat^T: k x n, K along its rows
    g.v_off = intern(vmap, vblk, vu, g.v_off, g.m, g.k, 1);  // V_R: m x k, K along its columns
  }
  std::vector<int> term_da(hp.g_term2.size(), -1);
  for (const GTile2& t : hp.g_tile2)
    for (int q = t.t_begin; q < t.t_end; ++q) term_da[size_t(q)] = t.da;
  std::vector<GTerm2> t2 = hp.g_term2;
  for (size_t q = 0; q < t2.size(); ++q) {
    if (term_da[q] < 0) continue;  // read by no tile
    t2[q].a_off = intern(amap, ablk, au, t2[q].a_off, term_da[q], t2[q].dar, 1);  // A: da x dar
  }
  std::string e;
  if (hipMalloc(&lp.Bp, sizeof(uint4) * (size_t)std::max<int64_t>(bu, 1)) != hipSuccess ||
      hipMalloc(&lp.Ap, sizeof(uint4) * (size_t)std::max<int64_t>(au, 1)) != hipSuccess ||
      hipMalloc(&lp.Vt, sizeof(uint4) * (size_t)std::max<int64_t>(vu, 1)) != hipSuccess)
    return "hipMalloc failed (panels)";
  if (!(e = upload_vec(t1, lp.g_task1p)).empty()) return e;
  if (!(e = upload_vec(t2, lp.g_term2p)).empty()) return e;
  // the operator panels, once
  for (int w = 0; w < 2; ++w) {
    const std::vector<PanelBlk>& blks = w == 0 ? bblk : ablk;
    const std::vector<PanelChunk> ch = panel_chunks(blks);
    PanelBlk* dblk = nullptr;
    PanelChunk* dch = nullptr;
    e = upload_vec(blks, dblk);
    if (e.empty()) e = upload_vec(ch, dch);
    if (e.empty() && launch_panels(dch, (int)ch.size(), dblk, w == 0 ? lp.Bh : lp.Ad, w == 0 ? lp.Bp : lp.Ap,
                                   nullptr) != hipSuccess)
      e = "panel kernel launch failed";
    if (e.empty() && hipDeviceSynchronize() != hipSuccess) e = "panel kernel failed";
    (void)hipFree(dblk);
    (void)hipFree(dch);
    if (!e.empty()) return e;
  }
  const std::vector<PanelChunk> vch = panel_chunks(vblk);
  if (!(e = upload_vec(vblk, lp.vt_blk)).empty()) return e;
  if (!(e = upload_vec(vch, lp.vt_chunk)).empty()) return e;
  lp.n_vt_chunk = (int)vch.size();
  return "";
}

template <typename T, typename F>
static std::string upload_as(const std::vector<double>& src, void*& dst, F conv) {
  dst = nullptr;
  std::vector<T> tmp(src.size());
  for (size_t i = 0; i < src.size(); ++i) tmp[i] = conv(src[i]);
  if (tmp.empty()) return "";
  if (hipMalloc(&dst, sizeof(T) * tmp.size()) != hipSuccess) return "hipMalloc failed (reduced-precision operands)";
  if (hipMemcpy(dst, tmp.data(), sizeof(T) * tmp.size(), hipMemcpyHostToDevice) != hipSuccess)
    return "hipMemcpy failed (reduced-precision operands)";
  return "";
}

std::string build_lowp(const HostPlan& hp, int precision, LowpPlan& lp, bool panels) {
  lp = LowpPlan();
  lp.precision = precision;
  std::string e;
  size_t esz;
  if (precision == 1) {
    auto cf = [](double x) { return (float)x; };
    if (!(e = upload_as<float>(hp.Ad, lp.Ad, cf)).empty()) return e;
    if (!(e = upload_as<float>(hp.Bh, lp.Bh, cf)).empty()) return e;
    if (!(e = upload_as<float>(hp.H, lp.H, cf)).empty()) return e;
    if (!(e = upload_as<float>(hp.Bn, lp.Bn, cf)).empty()) return e;
    esz = sizeof(float);
  } else if (precision == 2) {
    auto cb = [](double x) { return bf16x2_split((float)x); };
    if (!(e = upload_as<Bf16x2>(hp.Ad, lp.Ad, cb)).empty()) return e;
    if (!(e = upload_as<Bf16x2>(hp.Bh, lp.Bh, cb)).empty()) return e;
    if (!(e = upload_as<Bf16x2>(hp.H, lp.H, cb)).empty()) return e;
    if (!(e = upload_as<Bf16x2>(hp.Bn, lp.Bn, cb)).empty()) return e;
    esz = sizeof(Bf16x2);
  } else {
    return "";
  }
  if (hipMalloc(&lp.vin, esz * (size_t)(hp.N > 0 ? hp.N : 1)) != hipSuccess) return "hipMalloc failed (vector)";
  if (!hp.vmask.empty() && hipMalloc(&lp.vin_m, esz * (size_t)(hp.N > 0 ? hp.N : 1)) != hipSuccess)
    return "hipMalloc failed (masked vector)";
  // NSX_BF_PANELS=0 (environment) keeps the tiles on the Bf16x2 stores
  const char* env = getenv("NSX_BF_PANELS");
  if (env && atoi(env) == 0) panels = false;
  // NSX_BIG_XCD_GROUP=g1[,g2] (environment): the panel tiles' XCD runs
  if (const char* g = getenv("NSX_BIG_XCD_GROUP")) {
    int a = 0, b = -1;
    if (sscanf(g, "%d,%d", &a, &b) >= 1) {
      lp.xcd_grp1 = std::max(a, 0);
      lp.xcd_grp2 = b < 0 ? lp.xcd_grp1 : b;
    }
  }
  if (precision == 2 && panels && !hp.g_tile2.empty()) return build_panels(hp, lp);
  return "";
}

void free_lowp(LowpPlan& lp) {
  (void)hipFree(lp.Ad);
  (void)hipFree(lp.Bh);
  (void)hipFree(lp.H);
  (void)hipFree(lp.vin);
  (void)hipFree(lp.vin_m);
  (void)hipFree(lp.Bn);
  (void)hipFree(lp.Vt);
  (void)hipFree(lp.Bp);
  (void)hipFree(lp.Ap);
  (void)hipFree(lp.g_task1p);
  (void)hipFree(lp.g_term2p);
  (void)hipFree(lp.vt_blk);
  (void)hipFree(lp.vt_chunk);
  lp = LowpPlan();
}

}  // namespace nsx
