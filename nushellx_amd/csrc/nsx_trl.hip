// Thick-restart Lanczos (TRL) for the lowest eigenpairs of the pn J-scheme H.
//
// Replaces TRL.Lnz_OP (src/NuTRL/NuCoreX.f90:240-542).  The reference keeps the
// Lanczos vectors on disk in REAL*4 and re-reads them every step (lnzort,
// :544-582); here the basis Q (N x (m+1), FP64) stays in HBM and the full
// reorthogonalisation is classical Gram-Schmidt applied twice (two GEMV pairs
// per step).  After m steps the projected matrix (diagonal Ritz block, arrow
// row, tridiagonal tail; :426-438,472-481) is diagonalised on the host in FP64,
// and the lowest `keep` Ritz vectors are formed on the device (Q S).  A pair is
// converged when its residual |beta_m s_m,i| <= tol.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/nsx_hv.h"
#include "nsx_api_internal.hpp"

namespace nsx {

constexpr int kVT = 256;       // threads per block of the vector kernels
constexpr int kDotChunks = 64; // row chunks per column in the dot kernels

// partial[c * kDotChunks + chunk] = sum over the chunk's rows of Q[:, c] * w
__global__ __launch_bounds__(kVT) void k_dot_partial(const double* __restrict__ Q, int64_t ldq,
                                                     const double* __restrict__ w, int64_t n,
                                                     double* __restrict__ partial) {
  const int c = blockIdx.x, chunk = blockIdx.y;
  const int64_t per = (n + kDotChunks - 1) / kDotChunks;
  const int64_t r0 = chunk * per, r1 = min(n, r0 + per);
  const double* q = Q + ldq * c;
  double s = 0.0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += kVT) s = fma(q[i], w[i], s);
  __shared__ double red[kVT];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kVT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[c * kDotChunks + chunk] = red[0];
}

__global__ void k_dot_final(const double* __restrict__ partial, int ncols, double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  double s = 0.0;
  for (int k = 0; k < kDotChunks; ++k) s += partial[c * kDotChunks + k];
  out[c] = s;
}

// w -= Q[:, 0:ncols] c
__global__ __launch_bounds__(kVT) void k_sub_qc(double* __restrict__ w, const double* __restrict__ Q,
                                                int64_t ldq, const double* __restrict__ c, int ncols,
                                                int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)kVT + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < ncols; ++k) s = fma(Q[i + ldq * k], c[k], s);
  w[i] -= s;
}

// y = alpha * x
__global__ __launch_bounds__(kVT) void k_scale(double* __restrict__ y, const double* __restrict__ x,
                                               double alpha, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)kVT + threadIdx.x;
  if (i < n) y[i] = alpha * x[i];
}

// Y[:, 0:k] = Q[:, 0:m] S[0:m, 0:k] (S column-major, ld m)
__global__ __launch_bounds__(kVT) void k_ritz(double* __restrict__ Y, const double* __restrict__ Q,
                                              int64_t ldq, const double* __restrict__ S, int m, int k,
                                              int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)kVT + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < k; ++c) {
    double s = 0.0;
    for (int j = 0; j < m; ++j) s = fma(Q[i + ldq * j], S[j + (int64_t)m * c], s);
    Y[i + ldq * c] = s;
  }
}

// Cyclic Jacobi eigensolver for a small dense symmetric matrix (row-major n x n).
// Eigenvalues ascending in w, eigenvectors as columns of V (column-major n x n).
static void jacobi_eig(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
  V.assign(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) V[size_t(i) * n + i] = 1.0;
  auto a = [&](int i, int j) -> double& { return A[size_t(i) * n + j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        tot += a(i, j) * a(i, j);
        if (i != j) off += a(i, j) * a(i, j);
      }
    if (off <= 1e-30 * std::max(tot, 1e-300)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[size_t(p) * n + k], vkq = V[size_t(q) * n + k];
          V[size_t(p) * n + k] = c * vkp - s * vkq;
          V[size_t(q) * n + k] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  w.resize(n);
  std::vector<double> V2(size_t(n) * n);
  for (int c = 0; c < n; ++c) {
    w[c] = a(idx[c], idx[c]);
    std::memcpy(&V2[size_t(c) * n], &V[size_t(idx[c]) * n], sizeof(double) * n);
  }
  V.swap(V2);
}

struct TrlWork {
  double *Q = nullptr, *w = nullptr, *c = nullptr, *partial = nullptr, *S = nullptr, *Y = nullptr;
  ~TrlWork() {
    (void)hipFree(Q); (void)hipFree(w); (void)hipFree(c); (void)hipFree(partial);
    (void)hipFree(S); (void)hipFree(Y);
  }
};

}  // namespace nsx

using namespace nsx;

#define TRL_HIP(call)                                                     \
  do {                                                                    \
    hipError_t _e = (call);                                               \
    if (_e != hipSuccess) {                                               \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e));      \
      return NSX_EGPU;                                                    \
    }                                                                     \
  } while (0)

extern "C" int nsx_trl_solve(nsx_hv* h, const nsx_trl_opts* opt, const double* start, double* evals,
                             double* evecs, double* residuals, nsx_trl_stats* stats) {
  if (!h || !opt || !evals) { set_error("null argument"); return NSX_EARG; }
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t N = h->model.N;
  const int nev = opt->nev;
  if (nev < 1 || N < 1) { set_error("nothing to solve"); return NSX_EARG; }
  int64_t nstates = 0;
  for (uint8_t b : h->model.lcoef) nstates += b;
  if (nev > nstates) { set_error("more eigenvalues requested than basis states"); return NSX_EARG; }
  const int m = (int)std::min<int64_t>(std::max(opt->max_basis, nev + 2), nstates);
  const int keep = std::min(std::max(opt->keep, nev), std::max(m - 2, 1));
  const double tol = opt->tol > 0 ? opt->tol : 1e-6;
  const int max_iter = opt->max_iter > 0 ? opt->max_iter : 5000;
  hipStream_t s = h->stream;
  TrlWork W;
  const int64_t ldq = N;
  TRL_HIP(hipMalloc(&W.Q, sizeof(double) * N * (m + 1)));
  TRL_HIP(hipMalloc(&W.w, sizeof(double) * N));
  TRL_HIP(hipMalloc(&W.c, sizeof(double) * (m + 1)));
  TRL_HIP(hipMalloc(&W.partial, sizeof(double) * (m + 1) * kDotChunks));
  TRL_HIP(hipMalloc(&W.S, sizeof(double) * (size_t)m * m));
  TRL_HIP(hipMalloc(&W.Y, sizeof(double) * N * (m + 1)));
  const dim3 gv((unsigned)((N + kVT - 1) / kVT));

  // projections c = Q[:, 0:nc]^T x, returned on the host
  std::vector<double> hc(m + 1);
  auto project = [&](const double* x, int nc) -> hipError_t {
    hipLaunchKernelGGL(k_dot_partial, dim3(nc, kDotChunks), dim3(kVT), 0, s, W.Q, ldq, x, N, W.partial);
    hipLaunchKernelGGL(k_dot_final, dim3((nc + 63) / 64), dim3(64), 0, s, W.partial, nc, W.c);
    hipError_t e = hipMemcpyAsync(hc.data(), W.c, sizeof(double) * nc, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  };
  auto norm2 = [&](const double* x) -> double {
    hipLaunchKernelGGL(k_dot_partial, dim3(1, kDotChunks), dim3(kVT), 0, s, x, ldq, x, N, W.partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(64), 0, s, W.partial, 1, W.c);
    double v = 0.0;
    (void)hipMemcpyAsync(&v, W.c, sizeof(double), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    return v;
  };

  // start vector: given, or 0.5 - U[0,1) on the parity mask (as :299-312, own RNG)
  std::vector<double> v0(N, 0.0);
  std::mt19937_64 rng(opt->seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  auto random_fill = [&](std::vector<double>& v) {
    for (int64_t i = 0; i < N; ++i) v[i] = h->model.lcoef[i] ? 0.5 - U(rng) : 0.0;
  };
  if (start) {
    for (int64_t i = 0; i < N; ++i) v0[i] = h->model.lcoef[i] ? start[i] : 0.0;
  } else {
    random_fill(v0);
  }
  TRL_HIP(hipMemcpyAsync(W.w, v0.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
  double nrm = std::sqrt(norm2(W.w));
  if (!(nrm > 0.0) || !std::isfinite(nrm)) {
    random_fill(v0);
    TRL_HIP(hipMemcpyAsync(W.w, v0.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    nrm = std::sqrt(norm2(W.w));
  }
  hipLaunchKernelGGL(k_scale, gv, dim3(kVT), 0, s, W.Q, W.w, 1.0 / nrm, N);

  // projected matrix T ((m+1) x (m+1), row-major) and bookkeeping
  std::vector<double> T(size_t(m + 1) * (m + 1), 0.0);
  auto t = [&](int i, int j) -> double& { return T[size_t(i) * (m + 1) + j]; };
  int k0 = 0, matvecs = 0, restarts = 0, converged = 0;
  std::vector<double> theta, Svec, res(nev, 0.0);
  double beta_last = 0.0;
  bool done = false;
  while (!done) {
    int j_end = m;
    for (int j = k0; j < m; ++j) {
      int rc = nsx_hv_apply_dev(h, W.Q + ldq * j, W.w, s);
      if (rc) return rc;
      ++matvecs;
      // CGS2 against q_0..q_j
      TRL_HIP(project(W.w, j + 1));
      double alpha = hc[j];
      std::vector<double> c1(hc.begin(), hc.begin() + j + 1);
      TRL_HIP(hipMemcpyAsync(W.c, c1.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_sub_qc, gv, dim3(kVT), 0, s, W.w, W.Q, ldq, W.c, j + 1, N);
      TRL_HIP(project(W.w, j + 1));
      alpha += hc[j];
      TRL_HIP(hipMemcpyAsync(W.c, hc.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_sub_qc, gv, dim3(kVT), 0, s, W.w, W.Q, ldq, W.c, j + 1, N);
      t(j, j) = alpha;
      double beta = std::sqrt(norm2(W.w));
      if (!std::isfinite(beta) || !std::isfinite(alpha)) { set_error("non-finite Lanczos coefficient"); return NSX_EGPU; }
      if (j + 1 >= nstates) { beta_last = 0.0; j_end = j + 1; break; }
      if (beta < 1e-7 * std::max(1.0, std::fabs(alpha))) {
        // invariant subspace: continue from a fresh random direction, coupling 0 (:368-382)
        random_fill(v0);
        TRL_HIP(hipMemcpyAsync(W.w, v0.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
        for (int pass = 0; pass < 2; ++pass) {
          TRL_HIP(project(W.w, j + 1));
          TRL_HIP(hipMemcpyAsync(W.c, hc.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, s));
          hipLaunchKernelGGL(k_sub_qc, gv, dim3(kVT), 0, s, W.w, W.Q, ldq, W.c, j + 1, N);
        }
        const double nb = std::sqrt(norm2(W.w));
        hipLaunchKernelGGL(k_scale, gv, dim3(kVT), 0, s, W.Q + ldq * (j + 1), W.w, 1.0 / nb, N);
        beta = 0.0;
      } else {
        hipLaunchKernelGGL(k_scale, gv, dim3(kVT), 0, s, W.Q + ldq * (j + 1), W.w, 1.0 / beta, N);
      }
      if (j + 1 < m + 1) { t(j, j + 1) = beta; t(j + 1, j) = beta; }
      beta_last = beta;
    }
    const int mm = j_end;
    std::vector<double> Tm(size_t(mm) * mm);
    for (int i = 0; i < mm; ++i)
      for (int j = 0; j < mm; ++j) Tm[size_t(i) * mm + j] = t(i, j);
    jacobi_eig(Tm, mm, theta, Svec);
    converged = 0;
    for (int i = 0; i < nev; ++i) {
      res[i] = std::fabs(beta_last * Svec[size_t(i) * mm + (mm - 1)]);
      if (res[i] <= tol) ++converged;
    }
    if (converged == nev || mm < m || matvecs >= max_iter) {
      done = true;
      // Ritz vectors of the wanted pairs
      if (evecs) {
        std::vector<double> Sk(size_t(mm) * nev);
        for (int c = 0; c < nev; ++c)
          for (int j = 0; j < mm; ++j) Sk[size_t(c) * mm + j] = Svec[size_t(c) * mm + j];
        TRL_HIP(hipMemcpyAsync(W.S, Sk.data(), sizeof(double) * Sk.size(), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_ritz, gv, dim3(kVT), 0, s, W.Y, W.Q, ldq, W.S, mm, nev, N);
        TRL_HIP(hipMemcpyAsync(evecs, W.Y, sizeof(double) * N * nev, hipMemcpyDeviceToHost, s));
        TRL_HIP(hipStreamSynchronize(s));
      }
      break;
    }
    // thick restart: keep the lowest `keep` Ritz vectors and the residual direction q_m
    const int kk = std::min(keep, mm - 1);
    std::vector<double> Sk(size_t(mm) * kk);
    for (int c = 0; c < kk; ++c)
      for (int j = 0; j < mm; ++j) Sk[size_t(c) * mm + j] = Svec[size_t(c) * mm + j];
    TRL_HIP(hipMemcpyAsync(W.S, Sk.data(), sizeof(double) * Sk.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_ritz, gv, dim3(kVT), 0, s, W.Y, W.Q, ldq, W.S, mm, kk, N);
    TRL_HIP(hipMemcpyAsync(W.Q, W.Y, sizeof(double) * N * kk, hipMemcpyDeviceToDevice, s));
    TRL_HIP(hipMemcpyAsync(W.Q + ldq * kk, W.Q + ldq * mm, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < kk; ++i) {
      t(i, i) = theta[i];
      const double b = beta_last * Svec[size_t(i) * mm + (mm - 1)];
      t(i, kk) = b;
      t(kk, i) = b;
    }
    k0 = kk;
    ++restarts;
  }
  for (int i = 0; i < nev; ++i) {
    evals[i] = theta[i];
    if (residuals) residuals[i] = res[i];
  }
  if (stats) {
    stats->matvecs = matvecs;
    stats->restarts = restarts;
    stats->converged = converged;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return NSX_OK;
}
