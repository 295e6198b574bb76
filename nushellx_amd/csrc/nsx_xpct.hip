// Matrix elements of an operator between states: NuXpct's matrixs
// (src/NuXpct/NuCoreXE.f90:1441-1448), matrixc(i,j) = dot_product(H v_i, v_j).
//
// The states are uploaded once, H is applied to each on the handle's stream
// (the same plan and kernels as every other H*v), and the nvec x nvec dots run
// in one launch: one workgroup per (i, j), FP64 partial sums per lane, a
// wavefront shuffle reduction and one LDS step across the four wavefronts.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "nsx_api_internal.hpp"

namespace nsx {

constexpr int kGramThreads = 256;

__global__ void __launch_bounds__(kGramThreads) k_gram(const double* __restrict__ W, const double* __restrict__ V,
                                                       int64_t N, int nvec, double* __restrict__ M) {
  const int i = blockIdx.x % nvec, j = blockIdx.x / nvec;
  const double* w = W + int64_t(i) * N;
  const double* v = V + int64_t(j) * N;
  double s = 0.0;
  for (int64_t k = threadIdx.x; k < N; k += kGramThreads) s += w[k] * v[k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double part[kGramThreads / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kGramThreads / 64; ++q) t += part[q];
    M[int64_t(j) * nvec + i] = t;
  }
}

// Scratch of nsx_hv_expect_noalloc: the states, their images under H and the
// Gram matrix of one call, in a file-scope device array (8 MiB on every GPU the
// library runs on, held until the process ends).  In host code `scratch` names
// a host-side stand-in, not device memory: the kernels and copies take the
// address hipGetSymbolAddress returns for the current device.
constexpr int64_t kScratchDoubles = int64_t(1) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
static hipError_t scratch_ptr(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

}  // namespace nsx

using namespace nsx;

namespace {

struct DevBuf {
  double* p = nullptr;
  ~DevBuf() { (void)hipFree(p); }
};

int hip_fail(hipError_t e) {
  set_error(std::string("HIP error: ") + hipGetErrorString(e));
  return NSX_EGPU;
}

}  // namespace

extern "C" int nsx_hv_expect(nsx_hv* h, int32_t nvec, const double* vecs, double* M) {
  if (!h || !vecs || !M || nvec < 1) { set_error("bad arguments"); return NSX_EARG; }
  if (h->host_only) { set_error("host-only handle: expectation values need a device plan"); return NSX_EARG; }
  const int64_t N = h->model.N;
  if (N == 0) {
    for (int64_t q = 0; q < int64_t(nvec) * nvec; ++q) M[q] = 0.0;
    return NSX_OK;
  }
  if (int64_t(nvec) * nvec > (int64_t(1) << 31) - 1) { set_error("too many states"); return NSX_EARG; }
  DevBuf V, W, Md;
  hipError_t e;
  const size_t vbytes = sizeof(double) * size_t(N) * nvec;
  if ((e = hipMalloc(&V.p, vbytes)) != hipSuccess) return hip_fail(e);
  if ((e = hipMalloc(&W.p, vbytes)) != hipSuccess) return hip_fail(e);
  if ((e = hipMalloc(&Md.p, sizeof(double) * size_t(nvec) * nvec)) != hipSuccess) return hip_fail(e);
  if ((e = hipMemcpyAsync(V.p, vecs, vbytes, hipMemcpyHostToDevice, h->stream)) != hipSuccess) return hip_fail(e);
  for (int i = 0; i < nvec; ++i) {
    const int rc = nsx_hv_apply_dev(h, V.p + int64_t(i) * N, W.p + int64_t(i) * N, h->stream);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_gram, dim3(nvec * nvec), dim3(kGramThreads), 0, h->stream, W.p, V.p, N, nvec, Md.p);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e);
  if ((e = hipMemcpyAsync(M, Md.p, sizeof(double) * size_t(nvec) * nvec, hipMemcpyDeviceToHost, h->stream)) !=
      hipSuccess)
    return hip_fail(e);
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(e);
  return NSX_OK;
}

// nsx_hv_expect without device allocation: the states (N x nvec), their images
// H v_i (N x nvec) and M (nvec x nvec) live in the library's static scratch, so
// 2 N nvec + nvec^2 must not exceed kScratchDoubles.  Every call on a GPU uses
// that GPU's one scratch array: calls that overlap on one GPU (two host
// threads, or two handles' streams) race and must be serialised by the caller.
// Each call overwrites all of the scratch it reads (states copied in, H v and
// M written in full), so nothing carries over from the previous call.
extern "C" int nsx_hv_expect_noalloc(nsx_hv* h, int32_t nvec, const double* vecs, double* M) {
  if (!h || !vecs || !M || nvec < 1) { set_error("bad arguments"); return NSX_EARG; }
  if (h->host_only) { set_error("host-only handle: expectation values need a device plan"); return NSX_EARG; }
  const int64_t N = h->model.N;
  if (N == 0) {
    for (int64_t q = 0; q < int64_t(nvec) * nvec; ++q) M[q] = 0.0;
    return NSX_OK;
  }
  const int64_t need = 2 * N * int64_t(nvec) + int64_t(nvec) * nvec;
  if (need > kScratchDoubles) {
    set_error("nsx_hv_expect_noalloc: 2*N*nvec + nvec^2 = " + std::to_string(need) +
              " doubles exceeds the static scratch of " + std::to_string(kScratchDoubles) +
              " doubles (8 MiB); use nsx_hv_expect");
    return NSX_ECAP;
  }
  double* base = nullptr;
  hipError_t e = scratch_ptr(&base);
  if (e != hipSuccess) return hip_fail(e);
  double* V = base;
  double* W = base + N * nvec;
  double* Md = base + 2 * N * nvec;
  if ((e = hipMemcpyAsync(V, vecs, sizeof(double) * size_t(N) * nvec, hipMemcpyHostToDevice, h->stream)) !=
      hipSuccess)
    return hip_fail(e);
  for (int i = 0; i < nvec; ++i) {
    const int rc = nsx_hv_apply_dev(h, V + int64_t(i) * N, W + int64_t(i) * N, h->stream);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_gram, dim3(nvec * nvec), dim3(kGramThreads), 0, h->stream, W, V, N, nvec, Md);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e);
  if ((e = hipMemcpyAsync(M, Md, sizeof(double) * size_t(nvec) * nvec, hipMemcpyDeviceToHost, h->stream)) !=
      hipSuccess)
    return hip_fail(e);
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(e);
  return NSX_OK;
}
