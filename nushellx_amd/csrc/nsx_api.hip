// C ABI (include/nsx_hv.h): handles, device plans, host/device apply, batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nsx_hv.h"
#include "nsx_api_internal.hpp"
#include "nsx_parallel.hpp"

namespace nsx {

static thread_local std::string g_err;
void set_error(const std::string& e) { g_err = e; }

static int code_of(const std::string& e) {
  if (e.rfind("cannot open", 0) == 0) return NSX_EIO;
  if (e.find("not consistent") != std::string::npos) return NSX_ETAG;
  if (e.find("not supported") != std::string::npos) return NSX_ECAP;
  return NSX_EFORMAT;
}

template <typename T>
static std::string up(T*& dst, const std::vector<T>& src) {
  dst = nullptr;
  if (src.empty()) return "";
  if (hipMalloc(&dst, sizeof(T) * src.size()) != hipSuccess) return "hipMalloc failed";
  if (hipMemcpy(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice) != hipSuccess)
    return "hipMemcpy failed";
  return "";
}

std::string upload_plan(const HostPlan& hp, DevicePlan& d, int precision) {
  // the product buffers hold the plan's accumulation type: FP64, or FP32 for
  // the FP32 and BF16 hi/lo modes
  const size_t xbytes = precision == NSX_FP64 ? sizeof(double) : sizeof(float);
  static std::atomic<uint64_t> generation{0};
  d.gen = ++generation;
  std::string e;
  if (!(e = up(d.items2, hp.items2)).empty()) return e;
  if (!(e = up(d.phases, hp.phases)).empty()) return e;
  if (!(e = up(d.pairs, hp.pairs)).empty()) return e;
  d.n_items2 = (int)hp.items2.size();
  d.lds_y = hp.lds_y;
  d.lds_buf = hp.lds_buf;
  d.max_units = hp.max_units;
  d.lds2_bytes = sizeof(double) * (size_t)(2 * hp.lds_buf + hp.lds_y);
  if (d.lds2_bytes > 64 * 1024)
    for (int i = 0; i < num_staged(); ++i)
      if (hipFuncSetAttribute(k_pn_staged_ptr(i), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)d.lds2_bytes) != hipSuccess)
        return "LDS request too large (staged kernel)";
  if (const char* v = std::getenv("NSX_STAGED_NT")) d.staged_nt = std::atoi(v);
  if (const char* v = std::getenv("NSX_STAGED_MODE")) d.staged_mode = std::atoi(v);
  if (const char* v = std::getenv("NSX_ATOMIC")) d.atomic = std::atoi(v);
  d.stacked = hp.cfg.stacked == 1;
  if (hp.cfg.stacked == 2 && !hp.f_task.empty()) {  // fused column path
    if (!(e = up(d.f_task, hp.f_task)).empty()) return e;
    if (!(e = up(d.f_blk, hp.f_blk)).empty()) return e;
    if (!(e = up(d.f_round, hp.f_round)).empty()) return e;
    if (!(e = up(d.f_seg, hp.f_seg)).empty()) return e;
    if (!(e = up(d.f_x, hp.f_x)).empty()) return e;
    if (!(e = up(d.f_cons, hp.f_cons)).empty()) return e;
    if (!(e = up(d.f_ref, hp.f_ref)).empty()) return e;
    if (!(e = up(d.f_blocks, hp.f_blocks)).empty()) return e;
    if (!(e = up(d.f_part_list, hp.f_part_list)).empty()) return e;
    if (hp.f_part_size > 0 && hipMalloc(&d.f_part, sizeof(double) * (size_t)hp.f_part_size) != hipSuccess)
      return "hipMalloc failed (column partials)";
    d.n_ftask = (int)hp.f_task.size();
    d.n_fblocks = (int)hp.f_blocks.size();
    d.f_lds = hp.f_lds;
    d.f_u = hp.f_units <= 2 * kFuseThreads ? 2 : hp.f_units <= 4 * kFuseThreads ? 4 : 8;
    const size_t lds = (precision == NSX_FP64 ? sizeof(double) : sizeof(float)) * (size_t)hp.f_lds;
    if (lds > 48 * 1024)
      if (hipFuncSetAttribute(k_fuse_ptr(precision, d.f_u), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds) != hipSuccess)
        return "LDS request too large (fused kernel)";
    d.fused = 1;
  }
  if (hp.cfg.stacked == 4 && !hp.f_mtask.empty()) {  // fused matrix-core path
    if (!(e = up(d.f_mtask, hp.f_mtask)).empty()) return e;
    if (!(e = up(d.f_mstep, hp.f_mstep)).empty()) return e;
    if (!(e = up(d.f_mu, hp.f_mu)).empty()) return e;
    if (!(e = up(d.Bn, hp.Bn)).empty()) return e;
    d.n_mftask = (int)hp.f_mtask.size();
    for (int x = 0; x < 9; ++x) d.mxf_xcd.s[x] = hp.mxf_xcd[size_t(x)];
    if (hp.mxf_det) {  // deterministic: per-task partial tiles, summed per W tile in a fixed order
      if (!(e = up(d.f_dtile, hp.f_dtile)).empty()) return e;
      if (!(e = up(d.f_dpart, hp.f_dpart)).empty()) return e;
      const size_t np = (size_t)hp.f_mtask.size() * (size_t)hp.mxf_slots * 256;
      if (hipMalloc(&d.f_dpsum, sizeof(double) * std::max<size_t>(np, 1)) != hipSuccess)
        return "hipMalloc failed (deterministic partials)";
      d.n_dtile = (int)hp.f_dtile.size();
      d.mxf_det = 1;
    }
    d.mx_ks = hp.mx_ks;
    d.mxf_slots = hp.mxf_slots;
    d.mxf_like = hp.mxf_like ? 1 : 0;
    d.mxf = 1;
    d.atomic = 1;
  }
  if (hp.cfg.stacked == 3 && !hp.m_task2.empty()) {  // matrix-core stacked path
    if (!(e = up(d.m_unit1, hp.m_unit1)).empty()) return e;
    if (!(e = up(d.s_term2, hp.s_term2)).empty()) return e;
    if (!(e = up(d.s_xref, hp.s_xref)).empty()) return e;
    if (!(e = up(d.m_task1, hp.m_task1)).empty()) return e;
    if (!(e = up(d.m_task2, hp.m_task2)).empty()) return e;
    d.n_mtask1 = (int)hp.m_task1.size();
    d.n_mtask2 = (int)hp.m_task2.size();
    d.mx_ks = hp.mx_ks;
    d.s_lds1 = hp.s_lds1;
    if (hp.x_size > 0 && hipMalloc(&d.X, xbytes * (size_t)hp.x_size) != hipSuccess)
      return "hipMalloc failed (stacked X buffer)";
    d.mstk = 1;
    d.atomic = 1;
  }
  if (d.stacked) {
    if (!(e = up(d.s_chunk1, hp.s_chunk1)).empty()) return e;
    if (!(e = up(d.s_unit1, hp.s_unit1)).empty()) return e;
    if (!(e = up(d.s_chunk2, hp.s_chunk2)).empty()) return e;
    if (!(e = up(d.s_term2, hp.s_term2)).empty()) return e;
    if (!(e = up(d.s_xref, hp.s_xref)).empty()) return e;
    d.n_chunk1 = (int)hp.s_chunk1.size();
    d.n_chunk2 = (int)hp.s_chunk2.size();
    d.s_lds1 = hp.s_lds1;
    d.s_lds2 = hp.s_lds2;
    if (hp.x_size > 0 && hipMalloc(&d.X, xbytes * (size_t)hp.x_size) != hipSuccess)
      return "hipMalloc failed (stacked X buffer)";
    const size_t lds = sizeof(double) * (size_t)std::max(hp.s_lds1, hp.s_lds2);
    if (lds > 64 * 1024 - 4096)
      for (int prec = 0; prec < 3; ++prec)
        for (int nt : {64, 128, 256})
          for (int st = 1; st <= 2; ++st)
            if (hipFuncSetAttribute(k_stk_ptr(st, prec, nt), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds) != hipSuccess)
              return "LDS request too large (stacked kernels)";
    d.atomic = 1;
  }
  if (!hp.g_tile2.empty()) {  // large blocks: tiled GEMM path, which accumulates into W atomically
    if (!(e = up(d.g_task1, hp.g_task1)).empty()) return e;
    if (!(e = up(d.g_tile1, hp.g_tile1)).empty()) return e;
    if (!(e = up(d.g_tile2, hp.g_tile2)).empty()) return e;
    if (!(e = up(d.g_tile3, hp.g_tile3)).empty()) return e;
    if (!(e = up(d.g_term2, hp.g_term2)).empty()) return e;
    if (!(e = up(d.g_xref, hp.g_xref)).empty()) return e;
    d.n_gtile1 = (int)hp.g_tile1.size();
    d.n_gtile2 = (int)hp.g_tile2.size();
    d.n_gtile3 = (int)hp.g_tile3.size();
    d.big_rows1 = hp.cfg.big_rows1;
    if (hp.xg_size > 0 && hipMalloc(&d.XG, xbytes * (size_t)hp.xg_size) != hipSuccess)
      return "hipMalloc failed (large-block X buffer)";
    // X tiles that vanish identically are never computed: they stay zero
    if (hp.xg_size > 0 && hipMemset(d.XG, 0, xbytes * (size_t)hp.xg_size) != hipSuccess)
      return "hipMemset failed (large-block X buffer)";
    d.atomic = 1;
  }
  if (!(e = up(d.items, hp.items)).empty()) return e;
  if (!(e = up(d.aterms, hp.aterms)).empty()) return e;
  if (!(e = up(d.rterms, hp.rterms)).empty()) return e;
  if (!(e = up(d.blocks, hp.blocks)).empty()) return e;
  if (!(e = up(d.part_list, hp.part_list)).empty()) return e;
  if (!(e = up(d.Ad, hp.Ad)).empty()) return e;
  if (!(e = up(d.Bh, hp.Bh)).empty()) return e;
  if (!(e = up(d.H, hp.H)).empty()) return e;
  if (hp.part_size > 0 && hipMalloc(&d.part, sizeof(double) * hp.part_size) != hipSuccess)
    return "hipMalloc failed (partials)";
  d.n_items = (int)hp.items.size();
  d.n_blocks = (int)hp.blocks.size();
  d.max_lds = std::max(hp.max_lds, 1);
  d.N = hp.N;
  if (!hp.vmask.empty()) {  // truncated instances: the pn part reads a masked copy of V
    if (!(e = up(d.vmask, hp.vmask)).empty()) return e;
    if (hipMalloc(&d.Vm, sizeof(double) * (size_t)std::max<int64_t>(hp.N, 1)) != hipSuccess)
      return "hipMalloc failed (masked vector)";
  }
  if (sizeof(double) * (size_t)d.max_lds > 64 * 1024)
    for (int prec = 0; prec < 3; ++prec)
      for (int at = 0; at < 2; ++at)
        if (hipFuncSetAttribute(k_pn_partial_ptr(prec, at != 0), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * d.max_lds)) != hipSuccess)
          return "LDS request too large";
  return "";
}

void free_plan(DevicePlan& d) {
  (void)hipFree(d.items2); (void)hipFree(d.phases); (void)hipFree(d.pairs);
  (void)hipFree(d.items); (void)hipFree(d.aterms); (void)hipFree(d.rterms); (void)hipFree(d.blocks);
  (void)hipFree(d.part_list); (void)hipFree(d.Ad); (void)hipFree(d.Bh); (void)hipFree(d.H); (void)hipFree(d.part);
  (void)hipFree(d.s_chunk1); (void)hipFree(d.s_unit1); (void)hipFree(d.s_chunk2); (void)hipFree(d.s_term2);
  (void)hipFree(d.s_xref); (void)hipFree(d.X);
  (void)hipFree(d.m_task1); (void)hipFree(d.m_unit1); (void)hipFree(d.m_task2);
  (void)hipFree(d.f_mtask); (void)hipFree(d.f_mstep); (void)hipFree(d.f_mu); (void)hipFree(d.Bn);
  (void)hipFree(d.f_dtile); (void)hipFree(d.f_dpart); (void)hipFree(d.f_dpsum);
  (void)hipFree(d.g_task1); (void)hipFree(d.g_tile1); (void)hipFree(d.g_tile2); (void)hipFree(d.g_term2);
  (void)hipFree(d.g_tile3);
  (void)hipFree(d.g_xref); (void)hipFree(d.XG);
  (void)hipFree(d.f_task); (void)hipFree(d.f_blk); (void)hipFree(d.f_round); (void)hipFree(d.f_seg);
  (void)hipFree(d.f_x); (void)hipFree(d.f_cons); (void)hipFree(d.f_ref); (void)hipFree(d.f_blocks);
  (void)hipFree(d.f_part_list); (void)hipFree(d.f_part);
  (void)hipFree(d.vmask); (void)hipFree(d.Vm);
  d = DevicePlan();
}

}  // namespace nsx

using namespace nsx;

static std::string trimmed(const char* s) {
  std::string t = s ? s : "";
  size_t a = t.find_first_not_of(" \t");
  size_t b = t.find_last_not_of(" \t");
  if (a == std::string::npos) return "";
  return t.substr(a, b - a + 1);
}

#define NSX_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) {                                                  \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e));         \
      return NSX_EGPU;                                                       \
    }                                                                        \
  } while (0)

__global__ __launch_bounds__(256) void k_widen(const float* __restrict__ x, double* __restrict__ y, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}

__global__ __launch_bounds__(256) void k_narrow(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

extern "C" {

const char* nsx_last_error(void) { return nsx::g_err.c_str(); }

}  // extern "C"

static int create_impl(const char* nus_prefix, const char* opfile_a, const char* opfile_b,
                       const char* intfile, const char* dopfile, int precision, nsx_hv** out) {
  if (!out) { set_error("null output handle"); return NSX_EARG; }
  if (!nus_prefix || !opfile_a || !opfile_b || !intfile) { set_error("null file name"); return NSX_EARG; }
  *out = nullptr;
  const bool host_only = (precision & NSX_HOST_ONLY) != 0;
  const bool templ = (precision & NSX_TEMPLATE) != 0;
  precision &= ~(NSX_HOST_ONLY | NSX_TEMPLATE);
  if (templ && dopfile) { set_error("templates of linear ensembles take .oph interactions"); return NSX_EARG; }
  if (precision < NSX_FP64 || precision > NSX_BF16X3) { set_error("unknown precision"); return NSX_EARG; }
  std::unique_ptr<nsx_hv> h(new nsx_hv());
  h->precision = precision;
  h->host_only = host_only;
  std::string e = build_model(trimmed(nus_prefix), trimmed(opfile_a), trimmed(opfile_b),
                              trimmed(intfile), h->model, dopfile ? trimmed(dopfile) : std::string(), templ);
  if (!e.empty()) { set_error(e); return code_of(e); }
  if (host_only) {  // operators stay on the host until a batch plans them
    *out = h.release();
    return NSX_OK;
  }
  plan_models({&h->model}, h->hplan, precision);
  if (precision != NSX_FP64) {
    e = build_lowp(h->hplan, precision, h->lowp);
    if (!e.empty()) { set_error(e); return NSX_EGPU; }
  }
  e = upload_plan(h->hplan, h->dplan, precision);
  if (!e.empty()) { set_error(e); return NSX_EGPU; }
  NSX_HIP(hipStreamCreate(&h->stream));
  const int64_t N = std::max<int64_t>(h->model.N, 1);
  NSX_HIP(hipMalloc(&h->d_in, sizeof(double) * N));
  NSX_HIP(hipMalloc(&h->d_out, sizeof(double) * N));
  NSX_HIP(hipMalloc(&h->d_in32, sizeof(float) * N));
  NSX_HIP(hipMalloc(&h->d_out32, sizeof(float) * N));
  *out = h.release();
  return NSX_OK;
}

extern "C" {

int nsx_hv_create(const char* nus_prefix, const char* opfile_a, const char* opfile_b,
                  const char* intfile, int precision, nsx_hv** out) {
  return create_impl(nus_prefix, opfile_a, opfile_b, intfile, nullptr, precision, out);
}

int nsx_hv_create_sep(const char* nus_prefix, const char* opfile_a, const char* opfile_b,
                      const char* intfile, const char* dopfile, int precision, nsx_hv** out) {
  if (!dopfile || trimmed(dopfile).empty()) {
    if (out) *out = nullptr;
    set_error("empty .dop name");
    return NSX_EARG;
  }
  return create_impl(nus_prefix, opfile_a, opfile_b, intfile, dopfile, precision, out);
}

void nsx_hv_destroy(nsx_hv* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_trl_work(h->trl);
  free_plan(h->dplan);
  free_lowp(h->lowp);
  (void)hipFree(h->d_in);
  (void)hipFree(h->d_out);
  (void)hipFree(h->d_in32);
  (void)hipFree(h->d_out32);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int nsx_hv_dims(const nsx_hv* h, int64_t* N, int32_t* basd) {
  if (!h) { set_error("null handle"); return NSX_EARG; }
  if (N) *N = h->model.N;
  if (basd) *basd = h->model.basd;
  return NSX_OK;
}

int nsx_hv_block(const nsx_hv* h, int32_t b, int32_t* ja2, int32_t* jb2, int64_t* pos, int64_t* dim) {
  if (!h || b < 0 || b >= h->model.basd) { set_error("bad block index"); return NSX_EARG; }
  *ja2 = h->model.bas_a[b];
  *jb2 = h->model.bas_b[b];
  *pos = h->model.posnv[b];
  *dim = int64_t(h->model.da[b]) * h->model.db[b];
  return NSX_OK;
}

int nsx_hv_lcoef(const nsx_hv* h, uint8_t* mask) {
  if (!h || !mask) { set_error("null argument"); return NSX_EARG; }
  std::memcpy(mask, h->model.lcoef.data(), h->model.lcoef.size());
  return NSX_OK;
}

int nsx_hv_trunc(const nsx_hv* h, uint8_t* mask) {
  if (!h || !mask) { set_error("null argument"); return NSX_EARG; }
  const std::vector<uint8_t>& m = h->model.truncate ? h->model.trunc : h->model.lcoef;
  std::memcpy(mask, m.data(), m.size());
  return h->model.truncate ? 1 : 0;
}

int nsx_hv_work(const nsx_hv* h, double* flops, double* flops_like, int64_t* bytes) {
  if (!h) { set_error("null handle"); return NSX_EARG; }
  if (h->host_only) {  // no plan: work from the model, bytes unknown until planned
    if (flops) *flops = h->model.flops_ref;
    if (flops_like) *flops_like = h->model.flops_like;
    if (bytes) *bytes = 0;
    return NSX_OK;
  }
  if (flops) *flops = h->hplan.flops_ref;
  if (flops_like) *flops_like = h->hplan.flops_like;
  if (bytes) *bytes = h->hplan.bytes_alg;
  return NSX_OK;
}

int nsx_hv_apply_dev(nsx_hv* h, const double* d_in, double* d_out, void* stream) {
  if (!h || (!d_in && h->model.N) || (!d_out && h->model.N)) { set_error("null argument"); return NSX_EARG; }
  if (h->host_only) { set_error("host-only handle: apply it through a batch"); return NSX_EARG; }
  if (d_in == d_out && h->model.N) { set_error("in-place H*v is not supported"); return NSX_EARG; }
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  hipError_t e = launch_hv(h->dplan, h->lowp, h->precision, d_in, d_out, s);
  if (e != hipSuccess) { set_error(std::string("HIP error: ") + hipGetErrorString(e)); return NSX_EGPU; }
  return NSX_OK;
}

int nsx_hv_apply_f64(nsx_hv* h, const double* v_in, double* v_out) {
  if (!h || (h->model.N && (!v_in || !v_out))) { set_error("null argument"); return NSX_EARG; }
  if (h->host_only) { set_error("host-only handle: apply it through a batch"); return NSX_EARG; }
  const size_t bytes = sizeof(double) * h->model.N;
  if (!bytes) return NSX_OK;
  NSX_HIP(hipMemcpyAsync(h->d_in, v_in, bytes, hipMemcpyHostToDevice, h->stream));
  int rc = nsx_hv_apply_dev(h, h->d_in, h->d_out, h->stream);
  if (rc) return rc;
  NSX_HIP(hipMemcpyAsync(v_out, h->d_out, bytes, hipMemcpyDeviceToHost, h->stream));
  NSX_HIP(hipStreamSynchronize(h->stream));
  return NSX_OK;
}

// The REAL*4 drop-in: the 4-byte vectors cross PCIe as they are and are
// widened to FP64 (and the result narrowed) on the device.
int nsx_hv_apply_f32(nsx_hv* h, const float* v_in, float* v_out) {
  if (!h || (h->model.N && (!v_in || !v_out))) { set_error("null argument"); return NSX_EARG; }
  if (h->host_only) { set_error("host-only handle: apply it through a batch"); return NSX_EARG; }
  const int64_t N = h->model.N;
  if (!N) return NSX_OK;
  const dim3 g((unsigned)((N + 255) / 256));
  NSX_HIP(hipMemcpyAsync(h->d_in32, v_in, sizeof(float) * N, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_widen, g, dim3(256), 0, h->stream, h->d_in32, h->d_in, N);
  NSX_HIP(hipGetLastError());
  int rc = nsx_hv_apply_dev(h, h->d_in, h->d_out, h->stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_narrow, g, dim3(256), 0, h->stream, h->d_out, h->d_out32, N);
  NSX_HIP(hipGetLastError());
  NSX_HIP(hipMemcpyAsync(v_out, h->d_out32, sizeof(float) * N, hipMemcpyDeviceToHost, h->stream));
  NSX_HIP(hipStreamSynchronize(h->stream));
  return NSX_OK;
}

int nsx_batch_create(nsx_hv* const* hs, int32_t n, int precision, nsx_batch** out) {
  return nsx_batch_create_config(hs, n, precision, nullptr, out);
}

int nsx_batch_create_config(nsx_hv* const* hs, int32_t n, int precision, const nsx_tune_result* cfg,
                            nsx_batch** out) {
  if (!out || !hs || n <= 0) { set_error("bad batch arguments"); return NSX_EARG; }
  *out = nullptr;
  std::unique_ptr<nsx_batch> b(new nsx_batch());
  if (precision < NSX_FP64 || precision > NSX_BF16X3) { set_error("unknown precision"); return NSX_EARG; }
  b->precision = precision;
  if (cfg) b->hplan.cfg = plan_config(*cfg);
  for (int i = 0; i < n; ++i) {
    if (!hs[i]) { set_error("null handle in batch"); return NSX_EARG; }
    b->models.push_back(&hs[i]->model);
  }
  plan_models(b->models, b->hplan, precision);
  std::string e;
  if (precision != NSX_FP64 && !(e = build_lowp(b->hplan, precision, b->lowp)).empty()) {
    set_error(e);
    return NSX_EGPU;
  }
  if (!(e = upload_plan(b->hplan, b->dplan, precision)).empty()) { set_error(e); return NSX_EGPU; }
  if (cfg) {
    b->dplan.staged_nt = cfg->threads;
    b->dplan.staged_mode = cfg->mode;
    b->dplan.atomic = cfg->stacked ? 1 : cfg->atomic;
    b->dplan.stk_nt = cfg->threads;
  }
  NSX_HIP(hipStreamCreate(&b->stream));
  *out = b.release();
  return NSX_OK;
}

void nsx_batch_destroy(nsx_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_trl_work(b->trl);
  free_plan(b->dplan);
  free_lowp(b->lowp);
  free_ens(b->ens);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int nsx_batch_dims(const nsx_batch* b, int64_t* N) {
  if (!b || !N) { set_error("null argument"); return NSX_EARG; }
  *N = b->hplan.N;
  return NSX_OK;
}

int nsx_batch_offset(const nsx_batch* b, int32_t i, int64_t* off) {
  if (!b || !off || i < 0 || i >= (int)b->hplan.vbase.size()) { set_error("bad instance index"); return NSX_EARG; }
  *off = b->hplan.vbase[i];
  return NSX_OK;
}

int nsx_batch_work(const nsx_batch* b, double* flops, double* flops_like, int64_t* bytes) {
  if (!b) { set_error("null batch"); return NSX_EARG; }
  if (flops) *flops = b->hplan.flops_ref;
  if (flops_like) *flops_like = b->hplan.flops_like;
  if (bytes) *bytes = b->hplan.bytes_alg;
  return NSX_OK;
}

int nsx_batch_apply_dev(nsx_batch* b, const double* d_in, double* d_out, void* stream) {
  if (!b || !d_in || !d_out || d_in == d_out) { set_error("bad batch apply arguments"); return NSX_EARG; }
  hipStream_t s = stream ? (hipStream_t)stream : b->stream;
  hipError_t e = launch_hv(b->dplan, b->lowp, b->precision, d_in, d_out, s);
  if (e != hipSuccess) { set_error(std::string("HIP error: ") + hipGetErrorString(e)); return NSX_EGPU; }
  return NSX_OK;
}

}  // extern "C"

extern "C" {

int nsx_dev_alloc(int64_t bytes, void** ptr) {
  if (!ptr || bytes < 0) { set_error("bad allocation request"); return NSX_EARG; }
  *ptr = nullptr;
  NSX_HIP(hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 8)));
  return NSX_OK;
}
int nsx_dev_free(void* ptr) {
  NSX_HIP(hipFree(ptr));
  return NSX_OK;
}
int nsx_dev_upload(void* dst, const void* src, int64_t bytes) {
  if (bytes > 0) NSX_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return NSX_OK;
}
int nsx_dev_download(void* dst, const void* src, int64_t bytes) {
  if (bytes > 0) NSX_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return NSX_OK;
}
int nsx_dev_sync(void) {
  NSX_HIP(hipDeviceSynchronize());
  return NSX_OK;
}

int nsx_batch_path(const nsx_batch* b, int32_t* path) {
  if (!b || !path) { set_error("null argument"); return NSX_EARG; }
  *path = b->dplan.fused ? 2 : b->dplan.stacked ? 1 : b->dplan.mstk ? 3 : b->dplan.mxf ? 8 : 0;
  if (b->dplan.mxf && b->dplan.mxf_like) *path |= 16;  // the like-particle part rides in the fused path
  // deterministic: the fused column path or the fused matrix-core path with
  // partials for the small blocks, and unsplit large tiles
  if ((b->dplan.fused && b->hplan.cfg.deterministic) || b->dplan.mxf_det) *path |= 4;
  return NSX_OK;
}

int nsx_batch_time(nsx_batch* b, const double* d_in, double* d_out, int32_t warmup, int32_t iters,
                   double* ms_hv, double* ms_pn, double* ms_reduce) {
  if (!b || iters <= 0) { set_error("bad timing arguments"); return NSX_EARG; }
  hipStream_t s = b->stream;
  for (int i = 0; i < warmup; ++i) {
    int rc = nsx_batch_apply_dev(b, d_in, d_out, s);
    if (rc) return rc;
  }
  // one event pair around `iters` back-to-back launches of each part: the
  // average launch time without per-launch event overhead
  hipEvent_t e0, e1;
  NSX_HIP(hipEventCreate(&e0));
  NSX_HIP(hipEventCreate(&e1));
  auto loop = [&](int part, double* out) -> int {
    NSX_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < iters; ++i) {
      hipError_t e = part == 0   ? launch_hv(b->dplan, b->lowp, b->precision, d_in, d_out, s)
                     : part == 1 ? launch_pn(b->dplan, b->lowp, b->precision, d_in, d_out, s)
                                 : launch_reduce(b->dplan, b->lowp, b->precision, d_in, d_out, s);
      if (e != hipSuccess) { set_error(hipGetErrorString(e)); return NSX_EGPU; }
    }
    NSX_HIP(hipEventRecord(e1, s));
    NSX_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    NSX_HIP(hipEventElapsedTime(&t, e0, e1));
    if (out) *out = t / iters;
    return NSX_OK;
  };
  int rc = loop(0, ms_hv);
  // the pn and reduce parts read the converted input the last H*v left
  if (!rc) rc = loop(1, ms_pn);
  if (!rc) rc = loop(2, ms_reduce);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // extern "C"

// plain streaming copy, 16 bytes per lane: the reference point that scales the
// FETCH_SIZE / WRITE_SIZE counters to bytes, and a measured HBM rate
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     int64_t n) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

extern "C" {

int nsx_dev_copy(void* dst, const void* src, int64_t bytes, int32_t iters, double* ms) {
  if (!dst || !src || bytes <= 0 || bytes % 16 || iters <= 0) { set_error("bad copy arguments"); return NSX_EARG; }
  const int64_t n = bytes / 16;
  const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  hipEvent_t e0, e1;
  NSX_HIP(hipEventCreate(&e0));
  NSX_HIP(hipEventCreate(&e1));
  NSX_HIP(hipEventRecord(e0, nullptr));
  for (int i = 0; i < iters; ++i)
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, nullptr, (const uint4*)src, (uint4*)dst, n);
  NSX_HIP(hipGetLastError());
  NSX_HIP(hipEventRecord(e1, nullptr));
  NSX_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  NSX_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t / iters;
  return NSX_OK;
}

int nsx_device_info(char* name, int32_t len, int32_t* cus) {
  int dev = 0;
  NSX_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  NSX_HIP(hipGetDeviceProperties(&p, dev));
  if (name && len > 0) {
    std::strncpy(name, p.gcnArchName, (size_t)len - 1);
    name[len - 1] = 0;
  }
  if (cus) *cus = p.multiProcessorCount;
  return NSX_OK;
}

}  // extern "C"

namespace nsx {

PlanConfig plan_config(const nsx_tune_result& r) {
  PlanConfig pc;
  pc.fat_cap = std::min(std::max(r.fat_cap, 256), kFatCap);
  pc.max_phases = std::max(r.max_phases, 1);
  pc.stacked = r.stacked >= 2 && r.stacked <= 4 ? r.stacked : r.stacked ? 1 : 0;
  if (pc.stacked == 3) {  // matrix-core stacked path: a-terms per stage-2 workgroup, products per stage-1 wavefront
    if (r.stk_cap > 0) pc.mx_terms2 = r.stk_cap;
    if (r.stk_work > 0) pc.mx_units1 = r.stk_work;
  } else if (pc.stacked == 4) {  // fused matrix-core path: steps per wavefront, output slots per task
    if (r.stk_cap > 0) pc.mxf_steps = r.stk_cap;
    if (r.stk_work == 2) pc.mxf_slots = 2;
  } else {
    if (r.stk_cap > 0) pc.stk_cap = r.stk_cap;
    if (r.stk_work > 0) pc.stk_work = r.stk_work;
  }
  if (r.f_rounds > 0) pc.f_rounds = r.f_rounds;
  pc.deterministic = r.deterministic ? 1 : 0;
  if (r.big_rows1 == kBigTile || r.big_rows1 == 2 * kBigTile) pc.big_rows1_req = r.big_rows1;
  if (r.big_kchunk > 0) pc.big_kchunk = r.big_kchunk;
  return pc;
}

// Plans every model on its own in parallel, then merges the plans in order
// (same result as plan_add of each model into hp in turn).
void plan_models(const std::vector<const Model*>& models, HostPlan& hp, int precision) {
  const int n = (int)models.size();
  // BF16 stage-1 tiles are bound by operand traffic rather than the matrix
  // cores: 128-row tiles halve the re-reads of each Bhat^T panel
  // 128-row stage-1 tiles exist in BF16 hi/lo only
  hp.cfg.big_rows1 = precision == NSX_BF16X3 ? (hp.cfg.big_rows1_req ? hp.cfg.big_rows1_req : 2 * kBigTile) : kBigTile;
  // test hook: a smaller block side above which the large-block GEMM path runs,
  // so that small models exercise it
  if (const char* v = std::getenv("NSX_BIG_MIN")) hp.cfg.big_min = std::min(kSmallMax, std::max(1, std::atoi(v)));
  if (const char* v = std::getenv("NSX_DETERMINISTIC")) hp.cfg.deterministic = std::atoi(v) != 0;
  // deterministic: the fused matrix-core path with partials (the fused column
  // path where a tuning result asks for it, or a slab does not fit the tiles)
  auto det_column_shape = [&]() {  // the fused column shape measured fastest on sd 4p4n J=0..8
    hp.cfg.stacked = 2;
    hp.cfg.stk_cap = 6144;
    hp.cfg.stk_work = 250000;
    hp.cfg.f_rounds = 1;
  };
  if (hp.cfg.deterministic) {
    if (hp.cfg.stacked != 2 && hp.cfg.stacked != 4) {
      hp.cfg.stacked = 4;
      hp.cfg.mxf_steps = 2;
      hp.cfg.mxf_slots = kMfSlots;
    }
    hp.cfg.big_kchunk = 1 << 30;
  }
  auto finish = [&]() {
    if (hp.cfg.stacked == 1) build_stacked(hp, models);
    if (hp.cfg.stacked == 3) build_mstk(hp, models);
    if (hp.cfg.stacked == 4 && !build_mxf(hp, models)) {  // a slab of another shape: the stacked path
      if (hp.cfg.deterministic) {
        det_column_shape();
      } else {
        hp.cfg.stacked = 1;
        build_stacked(hp, models);
      }
    }
    build_big(hp, models);
    // fused column path, or the staged kernel where a column exceeds its limits
    if (hp.cfg.stacked == 2 && !build_fused(hp, models, kFuseThreads)) hp.cfg.stacked = 0;
  };
  if (n == 1) {
    plan_add(hp, *models[0]);
    finish();
    return;
  }
  std::vector<HostPlan> local(n);
  parallel_for(n, [&](int i) {
    local[i].cfg = hp.cfg;
    plan_add(local[i], *models[i]);
  }, 1);
  for (int i = 0; i < n; ++i) {
    plan_merge(hp, local[i]);
    local[i] = HostPlan();  // release as we go
  }
  finish();
}

std::string replan(const std::vector<const Model*>& models, const PlanConfig& cfg, int precision,
                   HostPlan& hp, DevicePlan& d, LowpPlan& lp) {
  const int nt = d.staged_nt, mode = d.staged_mode, at = d.atomic, snt = d.stk_nt;
  free_plan(d);
  free_lowp(lp);
  hp = HostPlan();
  hp.cfg = cfg;
  plan_models(models, hp, precision);
  std::string e;
  if (precision != NSX_FP64 && !(e = build_lowp(hp, precision, lp)).empty()) return e;
  if (!(e = upload_plan(hp, d, precision)).empty()) return e;
  d.staged_nt = nt;
  d.staged_mode = mode;
  d.atomic = hp.cfg.stacked ? 1 : at;
  d.stk_nt = snt;
  return "";
}

// Times one H*v (ms) over `iters` launches after 2 warm-up launches.
static double time_hv(const DevicePlan& d, const LowpPlan& lp, int precision, const double* in,
                      double* out, hipStream_t s, int iters) {
  for (int i = 0; i < 2; ++i) {
    hipError_t e = launch_hv(d, lp, precision, in, out, s);
    if (e != hipSuccess) return -1.0;
  }
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1.0;
  (void)hipEventRecord(a, s);
  for (int i = 0; i < iters; ++i) {
    hipError_t e = launch_hv(d, lp, precision, in, out, s);
    if (e != hipSuccess) return -1.0;
  }
  (void)hipEventRecord(b, s);
  (void)hipEventSynchronize(b);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / iters;
}

static int autotune(const std::vector<const Model*>& models, int precision, HostPlan& hp, DevicePlan& d,
                    LowpPlan& lp, hipStream_t s, int iters, nsx_tune_result* best_out) {
  const int64_t N = std::max<int64_t>(hp.N, 1);
  double *in = nullptr, *out = nullptr;
  if (hipMalloc(&in, sizeof(double) * N) != hipSuccess || hipMalloc(&out, sizeof(double) * N) != hipSuccess) {
    set_error("hipMalloc failed (autotune)");
    return NSX_EGPU;
  }
  std::vector<double> h(N);
  for (int64_t i = 0; i < N; ++i) h[i] = std::sin(0.37 * double(i + 1));
  (void)hipMemcpy(in, h.data(), sizeof(double) * N, hipMemcpyHostToDevice);
  nsx_tune_result best{0, 0, 0, 0, 0, 1e300, 0, 0, 0, 0, 0, 0, 0};
  const int caps[] = {768, 1536, 2560, kFatCap};
  const int chunks[] = {1, 2, 4, 8};
  const int nts[] = {64, 128, 256};
  for (int cap : caps)
    for (int mp : chunks) {
      PlanConfig cfg;
      cfg.fat_cap = cap;
      cfg.max_phases = mp;
      std::string e = replan(models, cfg, precision, hp, d, lp);
      if (!e.empty()) { set_error(e); (void)hipFree(in); (void)hipFree(out); return NSX_EGPU; }
      for (int nt : nts)
        for (int mode = 0; mode < 2; ++mode)
          for (int at = 0; at < 2; ++at) {
            d.staged_nt = nt;
            d.staged_mode = mode;
            d.atomic = at;
            const double ms = time_hv(d, lp, precision, in, out, s, iters);
            if (ms > 0.0 && ms < best.ms) best = {cap, mp, nt, mode, at, ms, 0, 0, 0, 0, 0, 0, 0};
          }
    }
  // stacked two-stage path: chunk capacity (LDS doubles) x work units x threads
  for (int scap : {2048, 4096})
    for (int swork : {256, 512, 1024}) {
      PlanConfig cfg;
      cfg.stacked = 1;
      cfg.stk_cap = scap;
      cfg.stk_work = swork;
      std::string e = replan(models, cfg, precision, hp, d, lp);
      if (!e.empty()) { set_error(e); (void)hipFree(in); (void)hipFree(out); return NSX_EGPU; }
      for (int nt : {128, 256}) {
        d.stk_nt = nt;
        d.atomic = 1;
        const double ms = time_hv(d, lp, precision, in, out, s, iters);
        if (ms > 0.0 && ms < best.ms) best = {kFatCap, kMaxPhases, nt, 0, 1, ms, 1, scap, swork, 0, 0, 0, 0};
      }
    }
  // the stacked stages on the matrix cores: products per stage-1 wavefront x
  // a-terms per stage-2 workgroup
  for (int u1 : {2, 4})
    for (int t2 : {8, 16}) {
      PlanConfig cfg;
      cfg.stacked = 3;
      cfg.mx_units1 = u1;
      cfg.mx_terms2 = t2;
      std::string e = replan(models, cfg, precision, hp, d, lp);
      if (!e.empty()) { set_error(e); (void)hipFree(in); (void)hipFree(out); return NSX_EGPU; }
      if (!d.mstk) break;  // no small blocks
      d.atomic = 1;
      const double ms = time_hv(d, lp, precision, in, out, s, iters);
      if (ms > 0.0 && ms < best.ms) best = {kFatCap, kMaxPhases, 256, 0, 1, ms, 3, t2, u1, 0, 0, 0, 0};
    }
  // the stacked stages fused on the matrix cores (X and Y in registers):
  // steps per wavefront
  for (int sl : {2, kMfSlots})
    for (int fs : {1, 2, 4}) {
      PlanConfig cfg;
      cfg.stacked = 4;
      cfg.mxf_steps = fs;
      cfg.mxf_slots = sl;
      std::string e = replan(models, cfg, precision, hp, d, lp);
      if (!e.empty()) { set_error(e); (void)hipFree(in); (void)hipFree(out); return NSX_EGPU; }
      if (!d.mxf) break;
      d.atomic = 1;
      const double ms = time_hv(d, lp, precision, in, out, s, iters);
      if (ms > 0.0 && ms < best.ms) best = {kFatCap, kMaxPhases, 256, 0, 1, ms, 4, fs, sl, 0, 0, 0, 0};
    }
  // fused column path: LDS of a round x work of a task
  for (int fcap : {3072, 4096, 6144})
    for (int frounds : {1, 2, 4}) {
      PlanConfig cfg;
      cfg.stacked = 2;
      cfg.stk_cap = fcap;
      cfg.stk_work = 250000;
      cfg.f_rounds = frounds;
      std::string e = replan(models, cfg, precision, hp, d, lp);
      if (!e.empty()) { set_error(e); (void)hipFree(in); (void)hipFree(out); return NSX_EGPU; }
      if (!d.fused) break;  // a column exceeds the fused kernel's limits
      const double ms = time_hv(d, lp, precision, in, out, s, iters);
      if (ms > 0.0 && ms < best.ms)
        best = {kFatCap, kMaxPhases, kFuseThreads, 0, 1, ms, 2, fcap, 250000, frounds, 0, 0, 0};
    }
  if (best.ms >= 1e300) {
    set_error("autotune: no configuration ran");
    (void)hipFree(in);
    (void)hipFree(out);
    return NSX_EGPU;
  }
  const PlanConfig cfg = plan_config(best);
  d.staged_nt = best.threads;
  d.staged_mode = best.mode;
  d.atomic = best.atomic;
  d.stk_nt = best.threads;
  std::string e = replan(models, cfg, precision, hp, d, lp);
  if (!e.empty()) { set_error(e); (void)hipFree(in); (void)hipFree(out); return NSX_EGPU; }
  // large-block tiles, on the chosen small-block shape: stage-2 K chunk and
  // (BF16 hi/lo) stage-1 rows
  if (d.n_gtile2 > 0 && !hp.cfg.deterministic) {
    std::vector<int> rows = {kBigTile};
    if (precision == NSX_BF16X3) rows.push_back(2 * kBigTile);
    nsx_tune_result bb = best;
    for (int r1 : rows)
      for (int kc : {32, 64, 128, 256}) {
        nsx_tune_result t = best;
        t.big_rows1 = r1;
        t.big_kchunk = kc;
        e = replan(models, plan_config(t), precision, hp, d, lp);
        if (!e.empty()) { set_error(e); (void)hipFree(in); (void)hipFree(out); return NSX_EGPU; }
        t.ms = time_hv(d, lp, precision, in, out, s, iters);
        if (t.ms > 0.0 && (bb.big_kchunk == 0 || t.ms < bb.ms)) bb = t;
      }
    best = bb;
    e = replan(models, plan_config(best), precision, hp, d, lp);
    if (!e.empty()) { set_error(e); (void)hipFree(in); (void)hipFree(out); return NSX_EGPU; }
  }
  (void)hipFree(in);
  (void)hipFree(out);
  if (best_out) *best_out = best;
  return NSX_OK;
}

}  // namespace nsx

extern "C" {

int nsx_hv_autotune(nsx_hv* h, int32_t iters, nsx_tune_result* best) {
  if (!h) { set_error("null handle"); return NSX_EARG; }
  std::vector<const Model*> ms{&h->model};
  return autotune(ms, h->precision, h->hplan, h->dplan, h->lowp, h->stream, iters > 0 ? iters : 10, best);
}

int nsx_batch_autotune(nsx_batch* b, int32_t iters, nsx_tune_result* best) {
  if (!b) { set_error("null batch"); return NSX_EARG; }
  if (b->ens.S > 0) { set_error("linear ensembles take their configuration at creation"); return NSX_EARG; }
  return autotune(b->models, b->precision, b->hplan, b->dplan, b->lowp, b->stream, iters > 0 ? iters : 10, best);
}

int nsx_batch_set_config(nsx_batch* b, const nsx_tune_result* cfg) {
  if (!b || !cfg) { set_error("null argument"); return NSX_EARG; }
  if (b->ens.S > 0) { set_error("linear ensembles take their configuration at creation"); return NSX_EARG; }
  const PlanConfig pc = plan_config(*cfg);
  b->dplan.staged_nt = cfg->threads;
  b->dplan.staged_mode = cfg->mode;
  b->dplan.atomic = cfg->stacked ? 1 : cfg->atomic;
  b->dplan.stk_nt = cfg->threads;
  std::string e = replan(b->models, pc, b->precision, b->hplan, b->dplan, b->lowp);
  if (!e.empty()) { set_error(e); return NSX_EGPU; }
  return NSX_OK;
}

}  // extern "C"
