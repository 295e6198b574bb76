// Linear ensembles: many interaction samples of the same nuclei and J values
// in one batch, with the operators of every sample produced on the device.
//
// NuShellX's operator files are linear in the interaction parameters theta
// (SPEs and TBMEs): the like-particle matrices H(J) (.mat, built by NuOrth as a
// sum over the interaction's matrix elements, src/NuOrth/NuOrth.f90:206-218)
// and the pn multipoles (.oph, NuOp's particle-hole recoupling of the TBMEs,
// src/NuOp/NuOp.f90:352-363).  The densities (.tri/.trd) do not depend on the
// interaction at all.  So a template instance per (nucleus, J), read once from
// files, fixes the structure of H*v (blocks, Racah terms, b-density fold
// recipes), and a sample only changes operator values:
//   values[s] = theta[s] . basis                       (one FP64 GEMM)
//   Bhat^T[s] = sum_p renorm val[s, p] Bsrc_p           (fold, per recipe)
//   H[s]      = the H columns of values[s]
// after which the stacked kernels run all samples at fixed strides, every
// sample's workgroups on one XCD, and the large-block tiles (blocks with a side
// above 40) run every sample as the second grid dimension.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "nsx_api_internal.hpp"

using namespace nsx;

namespace {

constexpr int kGemmS = 16;   // samples per GEMM workgroup
constexpr int kFoldS = 8;    // samples per fold workgroup
constexpr int kMaxSamples = 65528;  // grid y of the stacked kernels (samples rounded up to 8)

// values[s][k] = sum_i theta[s][i] basis[i][k]: a workgroup owns 256 columns and
// kGemmS samples; theta rows pass through LDS 256 parameters at a time.
__global__ __launch_bounds__(256) void k_ens_gemm(const double* __restrict__ theta,
                                                  const double* __restrict__ basis, double* __restrict__ values,
                                                  int S, int P, int64_t K) {
  __shared__ double th[kGemmS][256];
  const int tid = threadIdx.x;
  const int64_t k = (int64_t)blockIdx.x * 256 + tid;
  const int s0 = blockIdx.y * kGemmS;
  double acc[kGemmS];
#pragma unroll
  for (int q = 0; q < kGemmS; ++q) acc[q] = 0.0;
  for (int i0 = 0; i0 < P; i0 += 256) {
    const int ni = min(256, P - i0);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kGemmS; ++q) {
      const int s = s0 + q;
      th[q][tid] = (s < S && tid < ni) ? theta[(int64_t)s * P + i0 + tid] : 0.0;
    }
    __syncthreads();
    if (k < K)
      for (int i = 0; i < ni; ++i) {
        const double bv = basis[(int64_t)(i0 + i) * K + k];
#pragma unroll
        for (int q = 0; q < kGemmS; ++q) acc[q] = fma(th[q][i], bv, acc[q]);
      }
  }
  if (k < K)
#pragma unroll
    for (int q = 0; q < kGemmS; ++q)
      if (s0 + q < S) values[(int64_t)(s0 + q) * K + k] = acc[q];
}

// One Bhat^T sub-block for kFoldS samples: element e sums renorm * me * Bsrc
// over the recipe's b densities, with me zeroed below 1e-10 as the .oph loader
// does.  Padding between sub-blocks of a run is never written (zero from set-up).
__global__ __launch_bounds__(256) void k_ens_fold(const EnsFold* __restrict__ folds,
                                                  const EnsPair* __restrict__ pairs,
                                                  const double* __restrict__ bsrc,
                                                  const double* __restrict__ values, int64_t K, int S,
                                                  double* __restrict__ Bh, int64_t b_t) {
  const EnsFold f = folds[blockIdx.x];
  const int s0 = blockIdx.y * kFoldS;
  for (int e = threadIdx.x; e < f.n; e += 256) {
    double acc[kFoldS];
#pragma unroll
    for (int q = 0; q < kFoldS; ++q) acc[q] = 0.0;
    for (int p = f.p_begin; p < f.p_end; ++p) {
      const EnsPair pr = pairs[p];
      const double bv = bsrc[pr.src + e];
#pragma unroll
      for (int q = 0; q < kFoldS; ++q) {
        const int s = min(s0 + q, S - 1);
        double me = values[(int64_t)s * K + pr.val];
        me = fabs(me) < 1e-10 ? 0.0 : me;
        acc[q] = fma(pr.c * me, bv, acc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < kFoldS; ++q)
      if (s0 + q < S) Bh[(int64_t)(s0 + q) * b_t + f.dst + e] = acc[q];
  }
}

// The fused matrix-core path's untransposed copies of Bhat^T blocks, per
// sample: Bn[s][bn + j + db k] = Bh[s][bo + k + dbr j]; one workgroup per
// (block, sample)
__global__ __launch_bounds__(256) void k_ens_bn(const int64_t* __restrict__ src, const double* __restrict__ Bh,
                                                int64_t b_t, double* __restrict__ Bn, int64_t bn_t) {
  const int64_t* q = src + 4 * (int64_t)blockIdx.x;
  const int64_t bo = q[0], bn = q[1];
  const int dbr = (int)q[2], db = (int)q[3];
  const int64_t s = blockIdx.y;
  for (int e = threadIdx.x; e < dbr * db; e += 256) {
    const int k = e % dbr, j = e / dbr;
    Bn[s * bn_t + bn + j + (int64_t)db * k] = Bh[s * b_t + bo + e];
  }
}

// H[s][k] = values[s][n_o + k]
__global__ __launch_bounds__(256) void k_ens_hcopy(const double* __restrict__ values, int64_t K, int64_t n_o,
                                                   int64_t h_t, int64_t total, double* __restrict__ H) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const int64_t s = i / h_t, k = i - s * h_t;
  H[i] = values[s * K + n_o + k];
}

#define ENS_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) {                                                  \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e));         \
      return NSX_EGPU;                                                       \
    }                                                                        \
  } while (0)

// The H blocks of a template in its value layout: side a then side b, J ascending.
struct HEntry { int side, j, d; };
std::vector<HEntry> h_layout(const Model& m) {
  std::vector<HEntry> out;
  for (int side = 0; side < 2; ++side) {
    const SideFiles& sf = side ? m.B : m.A;
    for (int j = 0; j < (int)sf.H.size(); ++j)
      if (!sf.H[j].empty()) {
        const int d = (int)std::lround(std::sqrt((double)sf.H[j].size()));
        out.push_back({side, j, d});
      }
  }
  return out;
}

int64_t layout_len(const Model& m) {
  int64_t k = (int64_t)m.oph_me.size();
  for (const HEntry& h : h_layout(m)) k += int64_t(h.d) * h.d;
  return k;
}

// Produces every sample's Bhat^T and H stores (and their reduced-precision
// copies) from theta.
int produce(nsx_batch* b) {
  EnsData& E = b->ens;
  DevicePlan& d = b->dplan;
  hipStream_t s = b->stream;
  if (E.K > 0) {
    const dim3 g((unsigned)((E.K + 255) / 256), (unsigned)((E.S + kGemmS - 1) / kGemmS));
    hipLaunchKernelGGL(k_ens_gemm, g, dim3(256), 0, s, E.theta, E.basis, E.values, E.S, E.P, E.K);
    ENS_HIP(hipGetLastError());
  }
  if (E.n_fold > 0) {
    const dim3 g((unsigned)E.n_fold, (unsigned)((E.S + kFoldS - 1) / kFoldS));
    hipLaunchKernelGGL(k_ens_fold, g, dim3(256), 0, s, E.fold, E.pair, E.bsrc, E.values, E.K, E.S, d.Bh, E.b_t);
    ENS_HIP(hipGetLastError());
  }
  if (E.n_bn > 0) {
    hipLaunchKernelGGL(k_ens_bn, dim3((unsigned)E.n_bn, (unsigned)E.S), dim3(256), 0, s, E.bn_src, d.Bh, E.b_t, d.Bn,
                       E.bn_t);
    ENS_HIP(hipGetLastError());
  }
  const int64_t nh = (int64_t)E.S * E.h_t;
  if (nh > 0) {
    hipLaunchKernelGGL(k_ens_hcopy, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, E.values, E.K, E.n_o,
                       E.h_t, nh, d.H);
    ENS_HIP(hipGetLastError());
  }
  if (b->precision != NSX_FP64) {
    ENS_HIP(convert_to(b->precision, d.Bh, b->lowp.Bh, (int64_t)E.S * E.b_t, s));
    ENS_HIP(convert_to(b->precision, d.H, b->lowp.H, nh, s));
    if (E.n_bn > 0) ENS_HIP(convert_to(b->precision, d.Bn, b->lowp.Bn, (int64_t)E.S * E.bn_t, s));
  }
  ENS_HIP(hipStreamSynchronize(s));
  return NSX_OK;
}

template <typename T>
std::string up(T*& dst, const T* src, size_t n) {
  dst = nullptr;
  if (n == 0) return "";
  if (hipMalloc(&dst, sizeof(T) * n) != hipSuccess) return "hipMalloc failed (ensemble)";
  if (hipMemcpy(dst, src, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) return "hipMemcpy failed (ensemble)";
  return "";
}

}  // namespace

namespace nsx {
void free_ens(EnsData& e) {
  (void)hipFree(e.theta);
  (void)hipFree(e.basis);
  (void)hipFree(e.values);
  (void)hipFree(e.bsrc);
  (void)hipFree(e.fold);
  (void)hipFree(e.pair);
  (void)hipFree(e.bn_src);
  e = EnsData();
}
}  // namespace nsx

extern "C" {

int nsx_ens_layout(const nsx_hv* h, int64_t* n_oph, int32_t* n_h, int32_t* h_desc, int32_t cap) {
  if (!h || !n_oph || !n_h) { set_error("null argument"); return NSX_EARG; }
  const Model& m = h->model;
  if (!m.keep_zero_ops) { set_error("not a template handle (create it with NSX_TEMPLATE)"); return NSX_EARG; }
  const std::vector<HEntry> hl = h_layout(m);
  *n_oph = (int64_t)m.oph_me.size();
  *n_h = (int32_t)hl.size();
  if (h_desc) {
    if (cap < (int32_t)hl.size()) { set_error("h_desc too small"); return NSX_EARG; }
    for (size_t i = 0; i < hl.size(); ++i) {
      const SideFiles& sf = hl[i].side ? m.B : m.A;
      h_desc[3 * i] = hl[i].side;
      h_desc[3 * i + 1] = 2 * hl[i].j + (sf.minJ2 & 1);
      h_desc[3 * i + 2] = hl[i].d;
    }
  }
  return NSX_OK;
}

int nsx_ens_values(const nsx_hv* h, double* values) {
  if (!h || !values) { set_error("null argument"); return NSX_EARG; }
  const Model& m = h->model;
  if (!m.keep_zero_ops) { set_error("not a template handle (create it with NSX_TEMPLATE)"); return NSX_EARG; }
  size_t k = 0;
  for (double v : m.oph_me) values[k++] = v;
  for (const HEntry& he : h_layout(m)) {
    const std::vector<double>& H = (he.side ? m.B : m.A).H[he.j];
    std::memcpy(values + k, H.data(), sizeof(double) * H.size());
    k += H.size();
  }
  return NSX_OK;
}

int nsx_ens_create(nsx_hv* const* tm, int32_t n_t, int32_t S, int32_t P, const double* theta,
                   const double* const* basis, int precision, const nsx_tune_result* cfg, nsx_batch** out) {
  if (!out || !tm || !theta || !basis || n_t <= 0 || S <= 0 || P <= 0) {
    set_error("bad ensemble arguments");
    return NSX_EARG;
  }
  *out = nullptr;
  if (precision < NSX_FP64 || precision > NSX_BF16X3) { set_error("unknown precision"); return NSX_EARG; }
  if (S > kMaxSamples) { set_error("more than 65528 samples in one ensemble"); return NSX_ECAP; }
  std::vector<const Model*> tms;
  for (int t = 0; t < n_t; ++t) {
    if (!tm[t] || !basis[t]) { set_error("null template or basis"); return NSX_EARG; }
    const Model& m = tm[t]->model;
    if (!m.keep_zero_ops) { set_error("not a template handle (create it with NSX_TEMPLATE)"); return NSX_EARG; }
    if (m.separable) { set_error("linear ensembles take .oph interactions, not .dop"); return NSX_ECAP; }
    tms.push_back(&m);
  }
  std::unique_ptr<nsx_batch> b(new nsx_batch());
  b->precision = precision;
  // value columns: every template's .oph section, then the H store
  std::vector<int64_t> opcol(n_t), klen(n_t);
  int64_t n_o = 0;
  for (int t = 0; t < n_t; ++t) {
    opcol[t] = n_o;
    n_o += (int64_t)tms[t]->oph_me.size();
    klen[t] = layout_len(*tms[t]);
  }
  if (n_o > (int64_t)INT32_MAX) { set_error("too many .oph values"); return NSX_ECAP; }
  // templates that read the same interaction must be given the same .oph basis
  {
    std::map<std::string, int> first;
    for (int t = 0; t < n_t; ++t) {
      auto f = first.find(tms[t]->intfile);
      if (f == first.end()) { first[tms[t]->intfile] = t; continue; }
      const int u = f->second;
      const size_t no = tms[t]->oph_me.size();
      if (no != tms[u]->oph_me.size()) { set_error("templates share an interaction but not its keys"); return NSX_EARG; }
      for (int i = 0; i < P; ++i)
        if (std::memcmp(basis[t] + (size_t)i * klen[t], basis[u] + (size_t)i * klen[u], sizeof(double) * no) != 0) {
          set_error("templates that read the same interaction were given different .oph basis values");
          return NSX_EARG;
        }
    }
  }
  HostPlan& hp = b->hplan;
  hp.cfg = cfg ? plan_config(*cfg) : PlanConfig();
  // the stacked stages: on the VALU, on the matrix cores, or fused on the
  // matrix cores (whose untransposed Bhat copy k_ens_bn makes per sample; the
  // like-particle part stays in k_reduce, the H store being per sample)
  hp.cfg.stacked = cfg && (cfg->stacked == 3 || cfg->stacked == 4) ? cfg->stacked : 1;
  hp.cfg.deterministic = 0;
  hp.cfg.big_rows1 = precision == NSX_BF16X3 ? (hp.cfg.big_rows1_req ? hp.cfg.big_rows1_req : 2 * kBigTile) : kBigTile;
  // test hook, as for batches: a smaller block side above which the large-block tiles run
  if (const char* v = std::getenv("NSX_BIG_MIN")) hp.cfg.big_min = std::min(kSmallMax, std::max(1, std::atoi(v)));
  hp.ens = true;
  for (int t = 0; t < n_t; ++t) {
    hp.ens_t = t;
    hp.ens_opcol = (int32_t)opcol[t];
    plan_add(hp, *tms[t]);
  }
  if (hp.cfg.stacked == 4 && !build_mxf(hp, tms)) hp.cfg.stacked = 3;
  if (hp.cfg.stacked == 3)
    build_mstk(hp, tms);
  else if (hp.cfg.stacked == 1)
    build_stacked(hp, tms);
  build_big(hp, tms);
  if (!hp.cfg.stacked || (!hp.items.empty() && hp.g_tile2.empty())) {
    set_error("linear ensembles run on the stacked path and the large-block tiles: a template has a-terms "
              "beyond a stacked chunk or a block above the tiles' limit");
    return NSX_ECAP;
  }
  EnsData& E = b->ens;
  E.S = S;
  E.P = P;
  E.n_o = n_o;
  E.b_t = (int64_t)hp.Bh.size();
  E.h_t = (int64_t)hp.H.size();
  E.K = n_o + E.h_t;
  E.n_fold = (int)hp.ens_fold.size();
  // the concatenated basis (P x K, row-major)
  std::vector<std::vector<HEntry>> hl(n_t);
  std::vector<std::map<std::pair<int, int>, int64_t>> hpos(n_t);  // (side, j) -> column in the template layout
  for (int t = 0; t < n_t; ++t) {
    hl[t] = h_layout(*tms[t]);
    int64_t c = (int64_t)tms[t]->oph_me.size();
    for (const HEntry& he : hl[t]) {
      hpos[t][{he.side, he.j}] = c;
      c += int64_t(he.d) * he.d;
    }
  }
  std::vector<double> cat((size_t)P * E.K, 0.0);
  for (int t = 0; t < n_t; ++t) {
    const size_t no = tms[t]->oph_me.size();
    for (int i = 0; i < P; ++i)
      std::memcpy(cat.data() + (size_t)i * E.K + opcol[t], basis[t] + (size_t)i * klen[t], sizeof(double) * no);
  }
  for (const EnsHBlk& hb : hp.ens_h) {
    const int64_t src = hpos[hb.t].at({hb.side, hb.j});
    for (int i = 0; i < P; ++i)
      std::memcpy(cat.data() + (size_t)i * E.K + n_o + hb.dst, basis[hb.t] + (size_t)i * klen[hb.t] + src,
                  sizeof(double) * hb.n);
  }
  for (const EnsHBlk& hb : hp.ens_h_alias) {  // shared side files: the same basis values
    const int64_t src = hpos[hb.t].at({hb.side, hb.j});
    for (int i = 0; i < P; ++i)
      if (std::memcmp(cat.data() + (size_t)i * E.K + n_o + hb.dst, basis[hb.t] + (size_t)i * klen[hb.t] + src,
                      sizeof(double) * hb.n) != 0) {
        set_error("templates that read the same side files were given different H basis values");
        return NSX_EARG;
      }
  }
  // instances: the templates in order, sample-major
  const int64_t N_t = hp.N, x_t = hp.x_size, xg_t = hp.xg_size;
  const std::vector<int64_t> vb_t = hp.vbase;
  hp.vbase.clear();
  for (int s = 0; s < S; ++s)
    for (int t = 0; t < n_t; ++t) {
      b->models.push_back(tms[t]);
      hp.vbase.push_back((int64_t)s * N_t + vb_t[t]);
    }
  if (!hp.vmask.empty()) {  // window masks of the templates, per sample
    const std::vector<uint8_t> vm_t = hp.vmask;
    hp.vmask.clear();
    for (int s = 0; s < S; ++s) hp.vmask.insert(hp.vmask.end(), vm_t.begin(), vm_t.end());
  }
  hp.N = (int64_t)S * N_t;
  hp.x_size = (int64_t)S * x_t;
  hp.xg_size = (int64_t)S * xg_t;
  hp.flops_ref *= S;
  hp.flops_like *= S;
  hp.bytes_alg *= S;
  std::string e;
  if (precision != NSX_FP64) {
    // the operators are made per sample on the device: no panels
    if (!(e = build_lowp(hp, precision, b->lowp, false)).empty()) { set_error(e); return NSX_EGPU; }
    const size_t esz = precision == NSX_FP32 ? sizeof(float) : sizeof(Bf16x2);
    (void)hipFree(b->lowp.Bh);
    (void)hipFree(b->lowp.H);
    b->lowp.Bh = b->lowp.H = nullptr;
    if (E.b_t > 0) ENS_HIP(hipMalloc(&b->lowp.Bh, esz * (size_t)S * E.b_t));
    if (E.h_t > 0) ENS_HIP(hipMalloc(&b->lowp.H, esz * (size_t)S * E.h_t));
  }
  if (!(e = upload_plan(hp, b->dplan, precision)).empty()) { set_error(e); return NSX_EGPU; }
  DevicePlan& d = b->dplan;
  if (cfg) d.stk_nt = cfg->threads == 128 || cfg->threads == 64 ? cfg->threads : 256;
  (void)hipFree(d.Bh);
  (void)hipFree(d.H);
  d.Bh = d.H = nullptr;
  if (E.b_t > 0) {
    ENS_HIP(hipMalloc(&d.Bh, sizeof(double) * (size_t)S * E.b_t));
    ENS_HIP(hipMemset(d.Bh, 0, sizeof(double) * (size_t)S * E.b_t));
  }
  if (E.h_t > 0) ENS_HIP(hipMalloc(&d.H, sizeof(double) * (size_t)S * E.h_t));
  if (d.mxf && !hp.bn_src.empty()) {  // every sample's untransposed Bhat copy
    E.bn_t = (int64_t)hp.Bn.size();
    E.n_bn = (int)hp.bn_src.size();
    (void)hipFree(d.Bn);
    d.Bn = nullptr;
    ENS_HIP(hipMalloc(&d.Bn, sizeof(double) * (size_t)S * E.bn_t));
    ENS_HIP(hipMemset(d.Bn, 0, sizeof(double) * (size_t)S * E.bn_t));
    if (precision != NSX_FP64) {
      const size_t esz = precision == NSX_FP32 ? sizeof(float) : sizeof(Bf16x2);
      (void)hipFree(b->lowp.Bn);
      b->lowp.Bn = nullptr;
      ENS_HIP(hipMalloc(&b->lowp.Bn, esz * (size_t)S * E.bn_t));
    }
    std::vector<int64_t> src;
    for (const auto& q : hp.bn_src) src.insert(src.end(), q.begin(), q.end());
    if (!(e = up(E.bn_src, src.data(), src.size())).empty()) { set_error(e); return NSX_EGPU; }
  }
  d.ens_s = S;
  d.es.v = N_t;
  d.es.x = x_t;
  d.es.b = E.b_t;
  d.es.h = E.h_t;
  d.es.xg = xg_t;
  d.es.bn = E.bn_t;
  d.es.n = S;
  if (!(e = up(E.theta, theta, (size_t)S * P)).empty()) { set_error(e); return NSX_EGPU; }
  if (!(e = up(E.basis, cat.data(), cat.size())).empty()) { set_error(e); return NSX_EGPU; }
  if (!(e = up(E.bsrc, hp.ens_bsrc.data(), hp.ens_bsrc.size())).empty()) { set_error(e); return NSX_EGPU; }
  if (!(e = up(E.fold, hp.ens_fold.data(), hp.ens_fold.size())).empty()) { set_error(e); return NSX_EGPU; }
  if (!(e = up(E.pair, hp.ens_pair.data(), hp.ens_pair.size())).empty()) { set_error(e); return NSX_EGPU; }
  if (E.K > 0) ENS_HIP(hipMalloc(&E.values, sizeof(double) * (size_t)S * E.K));
  ENS_HIP(hipStreamCreate(&b->stream));
  const int rc = produce(b.get());
  if (rc != NSX_OK) return rc;
  *out = b.release();
  return NSX_OK;
}

int nsx_ens_update(nsx_batch* b, const double* theta) {
  if (!b || !theta) { set_error("null argument"); return NSX_EARG; }
  if (b->ens.S == 0) { set_error("not a linear ensemble"); return NSX_EARG; }
  ENS_HIP(hipStreamSynchronize(b->stream));
  ENS_HIP(hipMemcpy(b->ens.theta, theta, sizeof(double) * (size_t)b->ens.S * b->ens.P, hipMemcpyHostToDevice));
  return produce(b);
}

int nsx_ens_samples(const nsx_batch* b, int32_t* S, int32_t* n_t) {
  if (!b || !S || !n_t) { set_error("null argument"); return NSX_EARG; }
  *S = b->ens.S;
  *n_t = b->ens.S > 0 ? (int32_t)(b->models.size() / b->ens.S) : (int32_t)b->models.size();
  return NSX_OK;
}

}  // extern "C"
