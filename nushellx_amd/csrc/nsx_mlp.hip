// Learned warm start: a multilayer perceptron trained and evaluated on the GPU.
//
// The network maps interaction parameters (single-particle energies and JT
// two-body matrix elements, standardised by the caller) to a Lanczos start
// vector in the partition-block basis.  Layers: n_in -> n_hidden (SiLU) ->
// ... (n_layers hidden layers, SiLU) -> n_out (linear).  Weights are kept in
// FP32 as the optimiser's master copy and rounded to BF16 (round to nearest
// even) for every forward and backward product; activations, gradients and all
// accumulation are FP32.  Loss: L = 1/(2 n) sum_rows ||y_pred - y||^2.  Adam
// updates the master weights.  There is no counterpart in the reference (its
// Lanczos always starts from a random vector, src/NuTRL/NuCoreX.f90:299-312).
//
// Parameter layout (host views, get/set/grad): for each layer l in order,
// W_l as out_l x in_l row-major, then b_l (out_l).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/nsx_hv.h"
#include "nsx_api_internal.hpp"

namespace nsx {

__host__ __device__ inline float bf16_to_f32(uint16_t h) { return bits_f32(uint32_t(h) << 16); }

template <typename T>
__device__ inline float ldf(const T* p, int64_t i);
template <>
__device__ inline float ldf<float>(const float* p, int64_t i) { return p[i]; }
template <>
__device__ inline float ldf<uint16_t>(const uint16_t* p, int64_t i) { return bf16_to_f32(p[i]); }

__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float silu_grad(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

constexpr int kTile = 64, kKc = 16, kGT = 256;

typedef float f32x4v __attribute__((ext_vector_type(4)));

// C[m, n] (+)= sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn].
// 64 x 64 output tile per block, K in chunks of 16 through LDS (the load order
// follows whichever stride is unit), multiplied on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32): each of the four wavefronts holds a 32 x 32
// quarter as 2 x 2 blocks of 16 x 16; lane l supplies A(row l % 16, k l / 16)
// and B(k l / 16, column l % 16) from LDS and holds C(4 (l / 16) + v, l % 16).
// BF16 weights are widened to FP32 (exactly) while staging, so the products
// are those of the FP32 formulation.
// Epilogue: + bias[n] (optional), store to C (ldc); with Zact, also store the
// activation (SiLU when act) to Zact.  accumulate: C += result.
// Split K (part != null): block z of the grid's third dimension sums K range
// [z kchunk, (z+1) kchunk) and stores its raw M x N partial at part + z M N;
// k_gemm_reduce then sums the partials in order and applies the epilogue.
template <typename TA, typename TB>
__global__ __launch_bounds__(kGT) void k_gemm(int M, int N, int K, const TA* __restrict__ A, int64_t sam,
                                              int64_t sak, const TB* __restrict__ B, int64_t sbk, int64_t sbn,
                                              float* __restrict__ C, int64_t ldc, int accumulate,
                                              const float* __restrict__ bias, int act, float* __restrict__ Zact,
                                              int kchunk, float* __restrict__ part) {
  __shared__ float As[kKc][kTile + 1];
  __shared__ float Bs[kKc][kTile + 1];
  const int tm = blockIdx.y * kTile, tn = blockIdx.x * kTile;
  if (part) {  // this block's K range
    const int kb = (int)blockIdx.z * kchunk;
    A += (int64_t)kb * sak;
    B += (int64_t)kb * sbk;
    K = min(K - kb, kchunk);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4v acc[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[r][c] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
  const bool a_k_fast = sak == 1, b_n_fast = sbn == 1;
  for (int k0 = 0; k0 < K; k0 += kKc) {
    for (int e = threadIdx.x; e < kKc * kTile; e += kGT) {
      int mm, kk;
      if (a_k_fast) { mm = e / kKc; kk = e % kKc; } else { kk = e / kTile; mm = e % kTile; }
      const int m = tm + mm, k = k0 + kk;
      As[kk][mm] = (m < M && k < K) ? ldf(A, (int64_t)m * sam + (int64_t)k * sak) : 0.0f;
      int nn;
      if (b_n_fast) { kk = e / kTile; nn = e % kTile; } else { nn = e / kKc; kk = e % kKc; }
      const int n = tn + nn, k2 = k0 + kk;
      Bs[kk][nn] = (n < N && k2 < K) ? ldf(B, (int64_t)k2 * sbk + (int64_t)n * sbn) : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kKc; ks += 4) {
      const int kk = ks + lane / 16;
      float a[2], b[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) a[r] = As[kk][wm + 16 * r + lane % 16];
#pragma unroll
      for (int c = 0; c < 2; ++c) b[c] = Bs[kk][wn + 16 * c + lane % 16];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], b[c], acc[r][c], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = tm + wm + 16 * r + 4 * (lane / 16) + v, n = tn + wn + 16 * c + lane % 16;
        if (m >= M || n >= N) continue;
        float val = acc[r][c][v];
        if (part) {
          part[((int64_t)blockIdx.z * M + m) * N + n] = val;
          continue;
        }
        if (bias) val += bias[n];
        float* cp = C + (int64_t)m * ldc + n;
        if (accumulate) val += *cp;
        *cp = val;
        if (Zact) Zact[(int64_t)m * ldc + n] = act ? silu(val) : val;
      }
}

// Sum of the split-K partials (in split order) with k_gemm's epilogue
__global__ __launch_bounds__(256) void k_gemm_reduce(int M, int N, int splits, const float* __restrict__ part,
                                                     float* __restrict__ C, int64_t ldc, int accumulate,
                                                     const float* __restrict__ bias, int act,
                                                     float* __restrict__ Zact) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t mn = (int64_t)M * N;
  if (i >= mn) return;
  const int m = (int)(i / N), n = (int)(i - (int64_t)m * N);
  float v = 0.0f;
  for (int z = 0; z < splits; ++z) v += part[z * mn + i];
  if (bias) v += bias[n];
  float* c = C + (int64_t)m * ldc + n;
  if (accumulate) v += *c;
  *c = v;
  if (Zact) Zact[(int64_t)m * ldc + n] = act ? silu(v) : v;
}

// db[n] = sum_m dZ[m * ld + n]
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dZ, int M, int N, int64_t ld,
                                                float* __restrict__ db) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int m = 0; m < M; ++m) s += dZ[(int64_t)m * ld + n];
  db[n] = s;
}

// dZ = dA * silu'(Z), elementwise
__global__ __launch_bounds__(256) void k_silu_back(float* __restrict__ dZ, const float* __restrict__ Z, int64_t n) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < n) dZ[i] *= silu_grad(Z[i]);
}

// dY = (Yp - Y) / rows; per-block partial sums of (Yp - Y)^2
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Yp, const float* __restrict__ Y,
                                             float* __restrict__ dY, int64_t n, float inv_rows,
                                             double* __restrict__ partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float d = Yp[i] - Y[i];
    dY[i] = d * inv_rows;
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam on the master weights; refreshes the BF16 copy of weight entries
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m1, float* __restrict__ m2,
                                              uint16_t* __restrict__ wbf, int64_t n, float lr, float b1, float b2,
                                              float c1, float c2, float eps) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float x = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  w[i] = x;
  if (wbf) wbf[i] = (uint16_t)bf16_rne(x);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wbf,
                                                 int64_t n) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < n) wbf[i] = (uint16_t)bf16_rne(w[i]);
}

// rows idx[r] of src (ld cols) -> dst row r
__global__ __launch_bounds__(256) void k_gather_rows(const float* __restrict__ src, const int* __restrict__ idx,
                                                     float* __restrict__ dst, int rows, int64_t cols) {
  const int r = blockIdx.y;
  if (r >= rows) return;
  const float* s = src + (int64_t)idx[r] * cols;
  float* d = dst + (int64_t)r * cols;
  for (int64_t c = blockIdx.x * 256ll + threadIdx.x; c < cols; c += (int64_t)gridDim.x * 256) d[c] = s[c];
}

}  // namespace nsx

using namespace nsx;

struct nsx_mlp {
  std::vector<int> width;            // n_in, hidden..., n_out
  std::vector<int64_t> w_off, b_off;  // offsets into the flat parameter array
  int64_t n_params = 0;
  float* p = nullptr;      // master parameters (FP32)
  float* g = nullptr;      // gradients
  float* m1 = nullptr;     // Adam moments
  float* m2 = nullptr;
  uint16_t* pbf = nullptr;  // BF16 copy (weights; bias entries unused)
  int64_t step = 0;
  // activations for `cap` rows: Z_l (pre-activation) and A_l (post) per layer
  int cap = 0;
  std::vector<float*> Z, A;
  float* dA = nullptr;   // scratch gradients, cap x max width
  float* dB = nullptr;
  double* partial = nullptr;
  float* X = nullptr;    // input rows (cap x n_in)
  float* Y = nullptr;    // target rows (cap x n_out)
  hipStream_t s = nullptr;
  float* ws = nullptr;   // split-K partials of the long-K products
  size_t ws_elems = 0;
  int layers() const { return (int)width.size() - 1; }
};

namespace {

#define MLP_HIP(call)                                                            \
  do {                                                                           \
    hipError_t _e = (call);                                                      \
    if (_e != hipSuccess) {                                                      \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e));             \
      return NSX_EGPU;                                                           \
    }                                                                            \
  } while (0)

int free_rows(nsx_mlp* h) {
  for (float* x : h->Z) (void)hipFree(x);
  for (float* x : h->A) (void)hipFree(x);
  h->Z.clear();
  h->A.clear();
  (void)hipFree(h->dA);
  (void)hipFree(h->dB);
  (void)hipFree(h->partial);
  (void)hipFree(h->X);
  (void)hipFree(h->Y);
  h->dA = h->dB = h->X = h->Y = nullptr;
  h->partial = nullptr;
  h->cap = 0;
  return NSX_OK;
}

// make room for `rows` rows of activations
int ensure_rows(nsx_mlp* h, int rows) {
  if (rows <= h->cap) return NSX_OK;
  free_rows(h);
  const int L = h->layers();
  const int wmax = *std::max_element(h->width.begin(), h->width.end());
  h->Z.assign(L, nullptr);
  h->A.assign(L, nullptr);
  for (int l = 0; l < L; ++l) {
    MLP_HIP(hipMalloc(&h->Z[l], sizeof(float) * (size_t)rows * h->width[l + 1]));
    MLP_HIP(hipMalloc(&h->A[l], sizeof(float) * (size_t)rows * h->width[l + 1]));
  }
  MLP_HIP(hipMalloc(&h->dA, sizeof(float) * (size_t)rows * wmax));
  MLP_HIP(hipMalloc(&h->dB, sizeof(float) * (size_t)rows * wmax));
  MLP_HIP(hipMalloc(&h->partial, sizeof(double) * 1024));
  MLP_HIP(hipMalloc(&h->X, sizeof(float) * (size_t)rows * h->width[0]));
  MLP_HIP(hipMalloc(&h->Y, sizeof(float) * (size_t)rows * h->width.back()));
  h->cap = rows;
  return NSX_OK;
}

// Split-K sizing: few output tiles over a long K (the input gradient of the
// output layer: 32 rows x 512 columns over n_out) would leave the GPU idle
// with one block per tile, so K is cut into chunks of at least kSplitK, up to
// kMaxSplits, each block storing a partial.
constexpr int kSplitK = 512, kMaxSplits = 512, kSplitTiles = 128;

// The workspace the split-K partials of an M x N product need (elements);
// 0 when the product runs without splitting.
int64_t split_elems(int M, int N, int K, int* splits, int* kchunk) {
  const int64_t tiles = (int64_t)((N + kTile - 1) / kTile) * ((M + kTile - 1) / kTile);
  *splits = 1;
  *kchunk = K;
  if (tiles >= kSplitTiles || K < 2 * kSplitK) return 0;
  int S = (int)std::min<int64_t>((K + kSplitK - 1) / kSplitK, kMaxSplits);
  const int kc = (K + S - 1) / S;
  S = (K + kc - 1) / kc;
  *splits = S;
  *kchunk = kc;
  return (int64_t)S * M * N;
}

template <typename TA, typename TB>
int gemm(nsx_mlp* h, int M, int N, int K, const TA* A, int64_t sam, int64_t sak, const TB* B, int64_t sbk,
         int64_t sbn, float* C, int64_t ldc, int accumulate, const float* bias, int act, float* Zact) {
  int S, kc;
  const int64_t need = split_elems(M, N, K, &S, &kc);
  if (need > 0) {
    if ((size_t)need > h->ws_elems) {
      MLP_HIP(hipStreamSynchronize(h->s));
      (void)hipFree(h->ws);
      h->ws = nullptr;
      h->ws_elems = 0;
      MLP_HIP(hipMalloc(&h->ws, sizeof(float) * (size_t)need));
      h->ws_elems = (size_t)need;
    }
    const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile, S);
    hipLaunchKernelGGL((k_gemm<TA, TB>), grid, dim3(kGT), 0, h->s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc,
                       accumulate, bias, act, Zact, kc, h->ws);
    const int64_t mn = (int64_t)M * N;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, h->s, M, N, S, h->ws, C, ldc,
                       accumulate, bias, act, Zact);
  } else {
    const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
    hipLaunchKernelGGL((k_gemm<TA, TB>), grid, dim3(kGT), 0, h->s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc,
                       accumulate, bias, act, Zact, K, (float*)nullptr);
  }
  MLP_HIP(hipGetLastError());
  return NSX_OK;
}

// forward over `rows` rows already in h->X; output in h->A[L-1]
int forward(nsx_mlp* h, int rows) {
  const int L = h->layers();
  const float* in = h->X;
  for (int l = 0; l < L; ++l) {
    const int ni = h->width[l], no = h->width[l + 1];
    // Z = in . W^T + b ; A = silu(Z) (hidden) or Z (output)
    const int rc = gemm<float, uint16_t>(h, rows, no, ni, in, ni, 1, h->pbf + h->w_off[l], 1, ni, h->Z[l], no, 0,
                                         h->p + h->b_off[l], l + 1 < L ? 1 : 0, h->A[l]);
    if (rc) return rc;
    in = h->A[l];
  }
  return NSX_OK;
}

// backward from dY (in h->dA, rows x n_out) into h->g; needs forward() first
int backward(nsx_mlp* h, int rows) {
  const int L = h->layers();
  float* d = h->dA;  // dZ of layer l
  float* e = h->dB;
  for (int l = L - 1; l >= 0; --l) {
    const int ni = h->width[l], no = h->width[l + 1];
    const float* in = l > 0 ? h->A[l - 1] : h->X;
    // dW[o, i] = sum_r dZ[r, o] in[r, i]
    int rc = gemm<float, float>(h, no, ni, rows, d, 1, no, in, ni, 1, h->g + h->w_off[l], ni, 0, nullptr, 0, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_colsum, dim3((no + 255) / 256), dim3(256), 0, h->s, d, rows, no, (int64_t)no,
                       h->g + h->b_off[l]);
    if (l == 0) break;
    // dA_prev[r, i] = sum_o dZ[r, o] W[o, i]; then dZ_prev = dA_prev * silu'(Z_prev)
    rc = gemm<float, uint16_t>(h, rows, ni, no, d, no, 1, h->pbf + h->w_off[l], ni, 1, e, ni, 0, nullptr, 0, nullptr);
    if (rc) return rc;
    const int64_t n = (int64_t)rows * ni;
    hipLaunchKernelGGL(k_silu_back, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->s, e, h->Z[l - 1], n);
    std::swap(d, e);
  }
  return NSX_OK;
}

// loss and dY for rows in h->X / h->Y; returns sum of squares through *sq
int loss_grad(nsx_mlp* h, int rows, double* sq) {
  const int L = h->layers();
  const int64_t n = (int64_t)rows * h->width.back();
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, h->s, h->A[L - 1], h->Y, h->dA, n, 1.0f / rows, h->partial);
  if (sq) {
    std::vector<double> part(blocks);
    MLP_HIP(hipMemcpyAsync(part.data(), h->partial, sizeof(double) * blocks, hipMemcpyDeviceToHost, h->s));
    MLP_HIP(hipStreamSynchronize(h->s));
    double t = 0.0;
    for (double x : part) t += x;
    *sq = t;
  }
  return NSX_OK;
}

int refresh_bf16(nsx_mlp* h) {
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((h->n_params + 255) / 256)), dim3(256), 0, h->s, h->p, h->pbf,
                     h->n_params);
  MLP_HIP(hipGetLastError());
  return NSX_OK;
}

}  // namespace

extern "C" {

int nsx_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out, uint64_t seed,
                   nsx_mlp** out) {
  if (!out || n_in < 1 || n_hidden < 1 || n_layers < 0 || n_out < 1) { set_error("bad network shape"); return NSX_EARG; }
  *out = nullptr;
  auto* h = new nsx_mlp();
  h->width.push_back(n_in);
  for (int l = 0; l < n_layers; ++l) h->width.push_back(n_hidden);
  h->width.push_back(n_out);
  int64_t off = 0;
  for (int l = 0; l < h->layers(); ++l) {
    h->w_off.push_back(off);
    off += (int64_t)h->width[l] * h->width[l + 1];
    h->b_off.push_back(off);
    off += h->width[l + 1];
  }
  h->n_params = off;
  // init: W ~ N(0, 1/fan_in) for hidden layers, N(0, 0.01/fan_in) for the output layer, b = 0
  std::vector<float> p(off, 0.0f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < h->layers(); ++l) {
    const double sd = std::sqrt((l + 1 < h->layers() ? 1.0 : 0.01) / h->width[l]);
    std::normal_distribution<double> nd(0.0, sd);
    const int64_t n = (int64_t)h->width[l] * h->width[l + 1];
    for (int64_t i = 0; i < n; ++i) p[h->w_off[l] + i] = (float)nd(rng);
  }
  auto fail = [&](hipError_t e) {
    set_error(std::string("HIP error: ") + hipGetErrorString(e));
    (void)hipFree(h->p); (void)hipFree(h->g); (void)hipFree(h->m1); (void)hipFree(h->m2); (void)hipFree(h->pbf);
    delete h;
    return NSX_EGPU;
  };
  hipError_t e;
  if ((e = hipStreamCreate(&h->s)) != hipSuccess) { delete h; set_error(hipGetErrorString(e)); return NSX_EGPU; }
  if ((e = hipMalloc(&h->p, sizeof(float) * off)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&h->g, sizeof(float) * off)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&h->m1, sizeof(float) * off)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&h->m2, sizeof(float) * off)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&h->pbf, sizeof(uint16_t) * off)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(h->p, p.data(), sizeof(float) * off, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  if ((e = hipMemset(h->m1, 0, sizeof(float) * off)) != hipSuccess) return fail(e);
  if ((e = hipMemset(h->m2, 0, sizeof(float) * off)) != hipSuccess) return fail(e);
  if (refresh_bf16(h) != NSX_OK) return NSX_EGPU;
  *out = h;
  return NSX_OK;
}

void nsx_mlp_destroy(nsx_mlp* h) {
  if (!h) return;
  if (h->s) (void)hipStreamSynchronize(h->s);
  free_rows(h);
  (void)hipFree(h->ws);
  (void)hipFree(h->p);
  (void)hipFree(h->g);
  (void)hipFree(h->m1);
  (void)hipFree(h->m2);
  (void)hipFree(h->pbf);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
}

int nsx_mlp_num_params(const nsx_mlp* h, int64_t* n) {
  if (!h || !n) { set_error("null argument"); return NSX_EARG; }
  *n = h->n_params;
  return NSX_OK;
}

int nsx_mlp_get_params(nsx_mlp* h, float* p) {
  if (!h || !p) { set_error("null argument"); return NSX_EARG; }
  MLP_HIP(hipMemcpyAsync(p, h->p, sizeof(float) * h->n_params, hipMemcpyDeviceToHost, h->s));
  MLP_HIP(hipStreamSynchronize(h->s));
  return NSX_OK;
}

int nsx_mlp_set_params(nsx_mlp* h, const float* p) {
  if (!h || !p) { set_error("null argument"); return NSX_EARG; }
  MLP_HIP(hipMemcpyAsync(h->p, p, sizeof(float) * h->n_params, hipMemcpyHostToDevice, h->s));
  MLP_HIP(hipMemsetAsync(h->m1, 0, sizeof(float) * h->n_params, h->s));
  MLP_HIP(hipMemsetAsync(h->m2, 0, sizeof(float) * h->n_params, h->s));
  h->step = 0;
  int rc = refresh_bf16(h);
  if (rc) return rc;
  MLP_HIP(hipStreamSynchronize(h->s));
  return NSX_OK;
}

int nsx_mlp_grad(nsx_mlp* h, const float* X, const float* Y, int32_t rows, float* grad, double* loss) {
  if (!h || !X || !Y || rows < 1) { set_error("bad arguments"); return NSX_EARG; }
  int rc = ensure_rows(h, rows);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(h->X, X, sizeof(float) * (size_t)rows * h->width[0], hipMemcpyHostToDevice, h->s));
  MLP_HIP(hipMemcpyAsync(h->Y, Y, sizeof(float) * (size_t)rows * h->width.back(), hipMemcpyHostToDevice, h->s));
  if ((rc = forward(h, rows))) return rc;
  double sq = 0.0;
  rc = loss_grad(h, rows, &sq);
  if (rc) return rc;
  if ((rc = backward(h, rows))) return rc;
  MLP_HIP(hipGetLastError());
  if (grad) MLP_HIP(hipMemcpyAsync(grad, h->g, sizeof(float) * h->n_params, hipMemcpyDeviceToHost, h->s));
  MLP_HIP(hipStreamSynchronize(h->s));
  if (loss) *loss = 0.5 * sq / rows;
  return NSX_OK;
}

int nsx_mlp_train(nsx_mlp* h, const float* X, const float* Y, int32_t rows, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, float* epoch_loss) {
  if (!h || !X || !Y || rows < 1 || epochs < 0 || batch < 1) { set_error("bad arguments"); return NSX_EARG; }
  batch = std::min(batch, rows);
  int rc = ensure_rows(h, batch);
  if (rc) return rc;
  const int64_t nin = h->width[0], nout = h->width.back();
  float *dX = nullptr, *dY = nullptr;
  int* didx = nullptr;
  auto cleanup = [&]() { (void)hipFree(dX); (void)hipFree(dY); (void)hipFree(didx); };
#define MLP_TRY(call)                                                            \
  do {                                                                           \
    hipError_t _e = (call);                                                      \
    if (_e != hipSuccess) {                                                      \
      cleanup();                                                                 \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e));             \
      return NSX_EGPU;                                                           \
    }                                                                            \
  } while (0)
  MLP_TRY(hipMalloc(&dX, sizeof(float) * (size_t)rows * nin));
  MLP_TRY(hipMalloc(&dY, sizeof(float) * (size_t)rows * nout));
  MLP_TRY(hipMalloc(&didx, sizeof(int) * (size_t)rows));
  MLP_TRY(hipMemcpyAsync(dX, X, sizeof(float) * (size_t)rows * nin, hipMemcpyHostToDevice, h->s));
  MLP_TRY(hipMemcpyAsync(dY, Y, sizeof(float) * (size_t)rows * nout, hipMemcpyHostToDevice, h->s));
  std::vector<int> perm(rows);
  for (int i = 0; i < rows; ++i) perm[i] = i;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(perm.begin(), perm.end(), rng);
    MLP_TRY(hipMemcpyAsync(didx, perm.data(), sizeof(int) * rows, hipMemcpyHostToDevice, h->s));
    // the host permutation buffer is rewritten next epoch: wait for the copy
    MLP_TRY(hipStreamSynchronize(h->s));
    double ep_sq = 0.0;
    for (int r0 = 0; r0 < rows; r0 += batch) {
      const int nb = std::min(batch, rows - r0);
      hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)std::min<int64_t>((nin + 255) / 256, 64), nb), dim3(256), 0,
                         h->s, dX, didx + r0, h->X, nb, nin);
      hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)std::min<int64_t>((nout + 255) / 256, 64), nb), dim3(256), 0,
                         h->s, dY, didx + r0, h->Y, nb, nout);
      if ((rc = forward(h, nb))) { cleanup(); return rc; }
      double sq = 0.0;
      rc = loss_grad(h, nb, epoch_loss ? &sq : nullptr);
      if (rc) { cleanup(); return rc; }
      ep_sq += sq;
      if ((rc = backward(h, nb))) { cleanup(); return rc; }
      ++h->step;
      const float c1 = 1.0f - std::pow(b1, (float)h->step), c2 = 1.0f - std::pow(b2, (float)h->step);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)((h->n_params + 255) / 256)), dim3(256), 0, h->s, h->p, h->g, h->m1,
                         h->m2, h->pbf, h->n_params, lr, b1, b2, c1, c2, eps);
      MLP_TRY(hipGetLastError());
    }
    if (epoch_loss) epoch_loss[ep] = (float)(0.5 * ep_sq / rows);
  }
  MLP_TRY(hipStreamSynchronize(h->s));
  cleanup();
#undef MLP_TRY
  return NSX_OK;
}

int nsx_mlp_predict(nsx_mlp* h, const float* X, int32_t rows, float* Y) {
  if (!h || !X || !Y || rows < 1) { set_error("bad arguments"); return NSX_EARG; }
  int rc = ensure_rows(h, rows);
  if (rc) return rc;
  MLP_HIP(hipMemcpyAsync(h->X, X, sizeof(float) * (size_t)rows * h->width[0], hipMemcpyHostToDevice, h->s));
  if ((rc = forward(h, rows))) return rc;
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipMemcpyAsync(Y, h->A[h->layers() - 1], sizeof(float) * (size_t)rows * h->width.back(),
                         hipMemcpyDeviceToHost, h->s));
  MLP_HIP(hipStreamSynchronize(h->s));
  return NSX_OK;
}

}  // extern "C"
