// Thick-restart Lanczos (TRL) for the lowest eigenpairs, for one instance or a
// whole batch of instances advanced in lockstep (one batched H*v per step).
//
// Replaces TRL.Lnz_OP (src/NuTRL/NuCoreX.f90:240-542).  The reference keeps the
// Lanczos vectors on disk in REAL*4 and re-reads them every step (lnzort,
// :544-582); here the basis Q (N_total x (m+1), FP64) stays in HBM and the full
// reorthogonalisation is classical Gram-Schmidt applied twice with segmented
// (per-instance) reductions.  A restart cycle of m steps runs without host
// synchronisation: alpha/beta of every step are recorded on the device and read
// back once per cycle.  On the host, per instance, the projected matrix
// (diagonal Ritz block, arrow row, tridiagonal tail; :426-438,472-481) is
// diagonalised in FP64, residuals |beta_m s_m,i| decide convergence, and the
// lowest `keep` Ritz vectors are formed on the device for the restart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nsx_hv.h"
#include "nsx_api_internal.hpp"
#include "nsx_eig.hpp"
#include "nsx_parallel.hpp"

namespace nsx {

constexpr int kVT = 256;

// c[i * ldc + col] = <Q[:, col], w> over the rows of instance i; grid (B, ncols)
__global__ __launch_bounds__(kVT) void k_seg_dot(const double* __restrict__ Q, int64_t ldq, int col0,
                                                 const double* __restrict__ w, const int64_t* __restrict__ seg,
                                                 double* __restrict__ c, int ldc) {
  const int i = blockIdx.x, col = col0 + blockIdx.y;
  const int64_t r0 = seg[i], r1 = seg[i + 1];
  const double* q = Q + ldq * col;
  // rounds of four rows per thread (four independent chains, every load of a
  // round in flight together); the last round is guarded instead of serialised
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int64_t base = r0 + threadIdx.x; base < r1; base += 4 * kVT) {
    double qa[4], wa[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t r = base + (int64_t)u * kVT;
      qa[u] = r < r1 ? q[r] : 0.0;
      wa[u] = r < r1 ? w[r] : 0.0;
    }
    s0 = fma(qa[0], wa[0], s0);
    s1 = fma(qa[1], wa[1], s1);
    s2 = fma(qa[2], wa[2], s2);
    s3 = fma(qa[3], wa[3], s3);
  }
  double s = (s0 + s1) + (s2 + s3);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double red[kVT / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kVT / 64; ++k) t += red[k];
    c[(int64_t)i * ldc + col] = t;
  }
}

// Long instances: k_seg_dot runs over row chunks of at most kDotChunk rows
// (a segment table of chunks, partial dots per chunk), and this kernel sums the
// chunks of instance i in a fixed order: c[i, col] = sum_{ch in [cfirst[i], cfirst[i+1])} part[ch, col]
constexpr int64_t kDotChunk = 16384;
__global__ __launch_bounds__(kVT) void k_dot_fin(const double* __restrict__ part, const int* __restrict__ cfirst,
                                                 int ncols, double* __restrict__ c, int ldc) {
  const int i = blockIdx.x;
  const int ch0 = cfirst[i], ch1 = cfirst[i + 1];
  for (int col = threadIdx.x; col < ncols; col += kVT) {
    // eight chunks per round, their loads in flight together
    double t[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int ch = ch0; ch < ch1; ch += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] += ch + u < ch1 ? part[(int64_t)(ch + u) * ldc + col] : 0.0;
    }
    c[(int64_t)i * ldc + col] = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
  }
}

// Subtraction passes: a workgroup takes 64 rows and its four wavefronts split
// the columns (wave p sums columns p, p+4, ...), so a thread's chain of loads
// is a quarter as long and four times as many workgroups share the GPU; the
// four partial sums of a row meet in LDS.
constexpr int kSubRows = kVT / 4;

// sum_{col < ncols, col = p mod 4} Q[r, col] c[inst(r), col] of this thread's
// (row, column class); every thread of the block calls it
__device__ __forceinline__ double sub_partial(const double* __restrict__ Q, int64_t ldq, int ncols,
                                              const double* __restrict__ c, int ldc, const int* __restrict__ inst,
                                              int64_t n, int64_t r, int p) {
  // eight columns of the class per round, every load issued before the first
  // product, so a round costs one memory latency
  double s0 = 0.0, s1 = 0.0;
  if (r < n) {
    const double* ci = c + (int64_t)inst[r] * ldc;
    const double* qr = Q + r;
    for (int k = p; k < ncols; k += 32) {
      double qv[8], cv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kk = k + 4 * u;
        qv[u] = kk < ncols ? qr[ldq * kk] : 0.0;
        cv[u] = kk < ncols ? ci[kk] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        s0 = fma(qv[u], cv[u], s0);
        s1 = fma(qv[u + 1], cv[u + 1], s1);
      }
    }
  }
  return s0 + s1;
}

// w[r] -= sum_{col < ncols} Q[r, col] c[inst(r), col]; grid ceil(n / 64)
__global__ __launch_bounds__(kVT) void k_seg_sub(double* __restrict__ w, const double* __restrict__ Q,
                                                 int64_t ldq, int ncols, const double* __restrict__ c, int ldc,
                                                 const int* __restrict__ inst, int64_t n) {
  __shared__ double part[4][kSubRows];
  const int lane = threadIdx.x % kSubRows, p = threadIdx.x / kSubRows;
  const int64_t r = blockIdx.x * (int64_t)kSubRows + lane;
  part[p][lane] = sub_partial(Q, ldq, ncols, c, ldc, inst, n, r, p);
  __syncthreads();
  if (p == 0 && r < n) w[r] -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// Norm accumulators: kNrmSlots per instance, a workgroup adding into slot
// (block % kNrmSlots), so that the workgroups of one long instance do not all
// contend for one address; k_seg_finish sums the slots.
constexpr int kNrmSlots = 16;

// Second Gram-Schmidt pass fused with the norm: w[r] -= sum_k Q[r, k] c[inst(r), k],
// and per instance sum w^2 accumulated into the instance's nrm slots (a wavefront reduction per
// instance present in the block's rows, one FP64 atomic add each).
__global__ __launch_bounds__(kVT) void k_seg_sub_norm(double* __restrict__ w, const double* __restrict__ Q,
                                                      int64_t ldq, int ncols, const double* __restrict__ c, int ldc,
                                                      const int* __restrict__ inst, int64_t n,
                                                      double* __restrict__ nrm) {
  __shared__ double part[4][kSubRows];
  const int lane = threadIdx.x % kSubRows, p = threadIdx.x / kSubRows;
  const int64_t r0 = blockIdx.x * (int64_t)kSubRows;
  const int64_t r = r0 + lane;
  part[p][lane] = sub_partial(Q, ldq, ncols, c, ldc, inst, n, r, p);
  __syncthreads();
  if (p != 0) return;  // wavefront 0 finishes the rows (kSubRows = 64 lanes)
  double sq = 0.0;
  int my = -1;
  if (r < n) {
    my = inst[r];
    const double x = w[r] - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
    w[r] = x;
    sq = x * x;
  }
  const int i_lo = inst[r0];
  const int i_hi = inst[min(r0 + kSubRows, n) - 1];
  for (int i = i_lo; i <= i_hi; ++i) {  // instances are contiguous: usually one per block
    double v = (my == i) ? sq : 0.0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) unsafeAtomicAdd(nrm + (int64_t)i * kNrmSlots + blockIdx.x % kNrmSlots, v);
  }
}

// Finishes step j: beta = sqrt(sum of the instance's nrm slots), alpha = c1[j] + c2[j]; the row that
// starts an instance records alpha/beta, marks a breakdown (dim[i] = j + 1: an
// invariant subspace was reached) and clears the other norm slot for step
// j + 1; every row writes q_{j+1} = w / beta (0 once the instance stopped)
// and clears w, the output of the next step's H*v (which then needs no
// clearing kernel of its own).
// A row reading dim after the starting row updated it computes the same zero
// scale, since the update happens only on breakdown.
__global__ __launch_bounds__(kVT) void k_seg_finish(double* __restrict__ q, double* __restrict__ w,
                                                    const int64_t* __restrict__ seg, const int* __restrict__ inst,
                                                    int64_t n, const double* __restrict__ c1,
                                                    const double* __restrict__ c2, int ldc, int j,
                                                    const double* __restrict__ nrm, double* __restrict__ nrm_next,
                                                    double* __restrict__ alpha_h, double* __restrict__ beta_h, int B,
                                                    int* __restrict__ dim) {
  const int64_t r = blockIdx.x * (int64_t)kVT + threadIdx.x;
  if (r >= n) return;
  const int i = inst[r];
  const double a = c1[(int64_t)i * ldc + j] + c2[(int64_t)i * ldc + j];
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < kNrmSlots; ++q) ss += nrm[(int64_t)i * kNrmSlots + q];
  const double b = sqrt(ss);
  const bool live = dim[i] < 0;
  const bool breakdown = !(b > 1e-10 * fmax(1.0, fabs(a)));
  const bool ok = live && !breakdown;
  if (r == seg[i]) {
    alpha_h[(int64_t)j * B + i] = a;
    beta_h[(int64_t)j * B + i] = ok ? b : 0.0;
    if (live && breakdown) dim[i] = j + 1;
#pragma unroll
    for (int q = 0; q < kNrmSlots; ++q) nrm_next[(int64_t)i * kNrmSlots + q] = 0.0;
  }
  q[r] = ok ? w[r] / b : 0.0;
  w[r] = 0.0;
}

// truncated bases: W = 0 outside the N-hbar-omega window after every H*v, as
// Lnz_OP's `where (.not.trunc) coefL = 0` (src/NuTRL/NuCoreX.f90:307-308, 320-321, 343-344)
__global__ __launch_bounds__(kVT) void k_mask(double* __restrict__ w, const uint8_t* __restrict__ keep, int64_t n) {
  const int64_t r = blockIdx.x * (int64_t)kVT + threadIdx.x;
  if (r < n && !keep[r]) w[r] = 0.0;
}

__global__ __launch_bounds__(kVT) void k_fill0(double* __restrict__ x, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)kVT + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// Y[r, col] = sum_j Q[r, j] S[inst(r)][j][col], S per instance (mm x kk, column-major)
__global__ __launch_bounds__(kVT) void k_seg_ritz(double* __restrict__ Y, const double* __restrict__ Q,
                                                  int64_t ldq, const double* __restrict__ S, int mm, int kk,
                                                  const int* __restrict__ inst, int64_t n) {
  const int64_t r = blockIdx.x * (int64_t)kVT + threadIdx.x;
  if (r >= n) return;
  const double* Si = S + (int64_t)inst[r] * mm * kk;
  for (int c = 0; c < kk; ++c) {
    double s = 0.0;
    for (int j = 0; j < mm; ++j) s = fma(Q[r + ldq * j], Si[j + (int64_t)mm * c], s);
    Y[r + ldq * c] = s;
  }
}

// untyped device pointer with typed views
struct DPtr {
  void* p;
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// Lanczos workspace kept by a handle or batch between solves, so that the
// captured graphs (which hold its addresses) are replayed by later solves too.
struct TrlWork {
  int64_t N = 0;
  int B = 0, m = 0, keep = 0, nev = 0, prec = -1;
  uint64_t plan_gen = 0;
  void *Q = nullptr, *W = nullptr, *Y = nullptr, *inst_of = nullptr, *seg = nullptr, *c1 = nullptr,
       *c2 = nullptr, *ah = nullptr, *bh = nullptr, *scale = nullptr, *dim = nullptr, *S = nullptr,
       *wmask = nullptr, *cseg = nullptr, *cfirst = nullptr, *cpart = nullptr;  // chunked dots
  double *hha = nullptr, *hhb = nullptr;  // pinned alpha / beta histories
  int* hdim[2] = {nullptr, nullptr};      // pinned dimension flags, one per block parity
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<std::pair<std::pair<int, int>, hipGraphExec_t>> graphs;
  void clear_graphs() {
    for (auto& x : graphs) (void)hipGraphExecDestroy(x.second);
    graphs.clear();
  }
  void release() {
    clear_graphs();
    for (void* p : {Q, W, Y, inst_of, seg, c1, c2, ah, bh, scale, dim, S, wmask, cseg, cfirst, cpart}) (void)hipFree(p);
    Q = W = Y = inst_of = seg = c1 = c2 = ah = bh = scale = dim = S = wmask = cseg = cfirst = cpart = nullptr;
    (void)hipHostFree(hha);
    (void)hipHostFree(hhb);
    (void)hipHostFree(hdim[0]);
    (void)hipHostFree(hdim[1]);
    hha = hhb = nullptr;
    hdim[0] = hdim[1] = nullptr;
    for (auto& e : ev)
      if (e) { (void)hipEventDestroy(e); e = nullptr; }
    N = 0;
    B = m = keep = nev = 0;
  }
  ~TrlWork() { release(); }
};

void free_trl_work(TrlWork* w) { delete w; }

struct Instance {
  int64_t off, n;
  const std::vector<uint8_t>* keep;  // states the solve spans: the parity mask, or the window mask
  int64_t nstates;
  bool masked;                       // truncated basis: W is masked after every H*v
};

// The instance of one model: its kept states and their count.
static Instance instance_of(const Model& m, int64_t off) {
  const std::vector<uint8_t>* keep = m.truncate ? &m.trunc : &m.lcoef;
  int64_t nst = 0;
  for (uint8_t b : *keep) nst += b;
  return Instance{off, m.N, keep, nst, m.truncate};
}

// Lockstep TRL over instances sharing one device plan.  Outputs per instance i:
// evals[i*nev..], residuals[i*nev..], converged_at[i] (matvecs when all nev
// pairs first met tol, -1 if never); evecs (optional): N x nev column-major.
static int trl_run(const DevicePlan& d, const LowpPlan& lp, int prec, hipStream_t s, TrlWork*& work,
                   const std::vector<Instance>& inst, const nsx_trl_opts& opt, const double* start,
                   double* evals, double* evecs, double* residuals, int32_t* converged_at,
                   nsx_trl_stats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  const int B = (int)inst.size();
  const int nev = opt.nev;
  int64_t N = 0, min_states = INT64_MAX;
  for (const auto& x : inst) { N = std::max(N, x.off + x.n); min_states = std::min(min_states, x.nstates); }
  if (nev < 1 || N < 1) { set_error("nothing to solve"); return NSX_EARG; }
  if (nev > min_states) { set_error("more eigenvalues requested than basis states of an instance"); return NSX_EARG; }
  const int m = std::max(opt.max_basis, nev + 2);
  const int keep = std::min(std::max(opt.keep, nev), m - 2);
  const double tol = opt.tol > 0 ? opt.tol : 1e-6;
  const int max_iter = opt.max_iter > 0 ? opt.max_iter : 5000;
  const int check = opt.check_every > 0 ? opt.check_every : 8;
  const int64_t ldq = N;
  const int ldc = m + 1;

  // row chunks of the dot products: every instance split into pieces of at most
  // kDotChunk rows (at most B + N / kDotChunk of them)
  const int64_t max_chunks = B + N / kDotChunk + 1;
  if (!work) work = new TrlWork();
  TrlWork& ws = *work;
  const bool fresh = !(ws.N == N && ws.B == B && ws.m == m && ws.keep == keep && ws.nev == nev);
  if (fresh) {
    ws.release();
    auto dm = [&](void*& p, size_t bytes) { return hipMalloc(&p, bytes) == hipSuccess; };
    bool ok = dm(ws.Q, sizeof(double) * N * (m + 1)) && dm(ws.W, sizeof(double) * N) &&
              dm(ws.Y, sizeof(double) * N * std::max(keep, nev)) && dm(ws.inst_of, sizeof(int) * N) &&
              dm(ws.seg, sizeof(int64_t) * (B + 1)) && dm(ws.c1, sizeof(double) * B * ldc) &&
              dm(ws.c2, sizeof(double) * B * ldc) && dm(ws.ah, sizeof(double) * (m + 1) * B) &&
              dm(ws.bh, sizeof(double) * (m + 1) * B) && dm(ws.scale, sizeof(double) * 2 * B * kNrmSlots) &&
              dm(ws.dim, sizeof(int) * B) && dm(ws.S, sizeof(double) * (size_t)B * m * std::max(keep, nev)) &&
              dm(ws.wmask, (size_t)N) && dm(ws.cseg, sizeof(int64_t) * (max_chunks + 1)) &&
              dm(ws.cfirst, sizeof(int) * (B + 1)) && dm(ws.cpart, sizeof(double) * max_chunks * ldc) &&
              hipHostMalloc((void**)&ws.hha, sizeof(double) * (m + 1) * B) == hipSuccess &&
              hipHostMalloc((void**)&ws.hhb, sizeof(double) * (m + 1) * B) == hipSuccess &&
              hipHostMalloc((void**)&ws.hdim[0], sizeof(int) * B) == hipSuccess &&
              hipHostMalloc((void**)&ws.hdim[1], sizeof(int) * B) == hipSuccess &&
              hipEventCreateWithFlags(&ws.ev[0], hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&ws.ev[1], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      ws.release();
      set_error("allocation failed (Lanczos workspace)");
      return NSX_EGPU;
    }
    ws.N = N; ws.B = B; ws.m = m; ws.keep = keep; ws.nev = nev;
  }
  if (ws.plan_gen != d.gen || ws.prec != prec) {  // graphs hold plan pointers and kernel choices
    ws.clear_graphs();
    ws.plan_gen = d.gen;
    ws.prec = prec;
  }
  const DPtr Q{ws.Q}, W{ws.W}, Y{ws.Y}, inst_of{ws.inst_of}, seg{ws.seg}, c1{ws.c1}, c2{ws.c2}, ah{ws.ah},
      bh{ws.bh}, scale{ws.scale}, dim{ws.dim}, S{ws.S}, kmask{ws.wmask}, cseg{ws.cseg}, cfirst{ws.cfirst},
      cpart{ws.cpart};
#define TRL_HIP(call)                                                          \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess) {                                                    \
      set_error(std::string("HIP error: ") + hipGetErrorString(_e));           \
      return NSX_EGPU;                                                         \
    }                                                                          \
  } while (0)

  // segment table and row -> instance map (instances are contiguous and ordered)
  std::vector<int64_t> hseg(B + 1);
  std::vector<int> hinst(N, 0);
  for (int i = 0; i < B; ++i) {
    hseg[i] = inst[i].off;
    for (int64_t r = inst[i].off; r < inst[i].off + inst[i].n; ++r) hinst[r] = i;
  }
  hseg[B] = inst[B - 1].off + inst[B - 1].n;
  for (int i = 0; i + 1 < B; ++i)
    if (inst[i].off + inst[i].n != inst[i + 1].off) { set_error("instances must be contiguous"); return NSX_EARG; }
  TRL_HIP(hipMemcpy(seg.p, hseg.data(), sizeof(int64_t) * (B + 1), hipMemcpyHostToDevice));
  TRL_HIP(hipMemcpy(inst_of.p, hinst.data(), sizeof(int) * N, hipMemcpyHostToDevice));
  // an instance longer than kDotChunk rows would leave one workgroup per column
  // streaming it alone: its dots then run over row chunks
  bool chunked = false;
  for (const auto& x : inst) chunked = chunked || x.n > kDotChunk;
  int nchunks = 0;
  if (chunked) {
    std::vector<int64_t> hcs;
    std::vector<int> hcf(B + 1, 0);
    for (int i = 0; i < B; ++i) {
      hcf[i] = (int)hcs.size();
      for (int64_t r = inst[i].off; r < inst[i].off + inst[i].n; r += kDotChunk) hcs.push_back(r);
    }
    hcf[B] = (int)hcs.size();
    nchunks = (int)hcs.size();
    hcs.push_back(hseg[B]);
    if (nchunks > max_chunks) { set_error("internal: dot chunk table overflow"); return NSX_EGPU; }
    TRL_HIP(hipMemcpy(cseg.p, hcs.data(), sizeof(int64_t) * hcs.size(), hipMemcpyHostToDevice));
    TRL_HIP(hipMemcpy(cfirst.p, hcf.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice));
  }
  // window masks of truncated instances (others keep every row: H conserves parity)
  bool any_masked = false;
  for (const auto& x : inst) any_masked = any_masked || x.masked;
  if (any_masked) {
    std::vector<uint8_t> hk(N, 1);
    for (const auto& x : inst)
      if (x.masked)
        for (int64_t r = 0; r < x.n; ++r) hk[x.off + r] = (*x.keep)[r];
    TRL_HIP(hipMemcpy(kmask.p, hk.data(), (size_t)N, hipMemcpyHostToDevice));
  }
  const dim3 gv((unsigned)((N + kVT - 1) / kVT));
  const dim3 gsub((unsigned)((N + kSubRows - 1) / kSubRows));

  // start vectors (given, or 0.5 - U[0,1) on the parity mask as :299-312 with an own RNG)
  {
    std::vector<double> v0(N, 0.0);
    std::mt19937_64 rng(opt.seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (int i = 0; i < B; ++i) {
      double nrm = 0.0;
      for (int64_t r = 0; r < inst[i].n; ++r) {
        const bool on = (*inst[i].keep)[r] != 0;
        double x = on ? (start ? start[inst[i].off + r] : 0.5 - U(rng)) : 0.0;
        if (!std::isfinite(x)) x = 0.0;
        v0[inst[i].off + r] = x;
        nrm += x * x;
      }
      if (!(nrm > 1e-300)) {  // unusable warm start: fall back to a random start
        nrm = 0.0;
        for (int64_t r = 0; r < inst[i].n; ++r) {
          const double x = (*inst[i].keep)[r] ? 0.5 - U(rng) : 0.0;
          v0[inst[i].off + r] = x;
          nrm += x * x;
        }
      }
      const double sc = 1.0 / std::sqrt(nrm);
      for (int64_t r = 0; r < inst[i].n; ++r) v0[inst[i].off + r] *= sc;
    }
    TRL_HIP(hipMemcpy(Q.p, v0.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  }

  // c[i, col] = Q[:, col] . W over instance i's rows, col < ncols
  auto dots = [&](int ncols, double* c) {
    if (!chunked) {
      hipLaunchKernelGGL(k_seg_dot, dim3(B, ncols), dim3(kVT), 0, s, Q.as<double>(), ldq, 0, W.as<double>(),
                         seg.as<int64_t>(), c, ldc);
      return;
    }
    hipLaunchKernelGGL(k_seg_dot, dim3(nchunks, ncols), dim3(kVT), 0, s, Q.as<double>(), ldq, 0, W.as<double>(),
                       cseg.as<int64_t>(), cpart.as<double>(), ldc);
    hipLaunchKernelGGL(k_dot_fin, dim3(B), dim3(kVT), 0, s, cpart.as<double>(), cfirst.as<int>(), ncols, c, ldc);
  };
  // Lanczos steps [j0, j1) on the stream: H*v, two Gram-Schmidt passes, normalise
  auto launch_steps = [&](int j0, int j1) -> hipError_t {
    // norm accumulator of the block's first step and its H*v output (later
    // steps: both cleared by k_seg_finish)
    hipLaunchKernelGGL(k_fill0, dim3((B * kNrmSlots + kVT - 1) / kVT), dim3(kVT), 0, s,
                       scale.as<double>() + (size_t)(j0 & 1) * B * kNrmSlots, (int64_t)B * kNrmSlots);
    hipLaunchKernelGGL(k_fill0, gv, dim3(kVT), 0, s, W.as<double>(), N);
    for (int j = j0; j < j1; ++j) {
      hipError_t e = launch_hv(d, lp, prec, Q.as<double>() + ldq * j, W.as<double>(), s, true);
      if (e != hipSuccess) return e;
      if (any_masked) hipLaunchKernelGGL(k_mask, gv, dim3(kVT), 0, s, W.as<double>(), kmask.as<uint8_t>(), N);
      dots(j + 1, c1.as<double>());
      hipLaunchKernelGGL(k_seg_sub, gsub, dim3(kVT), 0, s, W.as<double>(), Q.as<double>(), ldq, j + 1,
                         c1.as<double>(), ldc, inst_of.as<int>(), N);
      dots(j + 1, c2.as<double>());
      double* nrm_j = scale.as<double>() + (size_t)(j & 1) * B * kNrmSlots;
      double* nrm_n = scale.as<double>() + (size_t)((j + 1) & 1) * B * kNrmSlots;
      hipLaunchKernelGGL(k_seg_sub_norm, gsub, dim3(kVT), 0, s, W.as<double>(), Q.as<double>(), ldq, j + 1,
                         c2.as<double>(), ldc, inst_of.as<int>(), N, nrm_j);
      hipLaunchKernelGGL(k_seg_finish, gv, dim3(kVT), 0, s, Q.as<double>() + ldq * (j + 1), W.as<double>(),
                         seg.as<int64_t>(), inst_of.as<int>(), N, c1.as<double>(), c2.as<double>(), ldc, j, nrm_j,
                         nrm_n, ah.as<double>(), bh.as<double>(), B, dim.as<int>());
    }
    return hipGetLastError();
  };
  // captured blocks of steps, keyed by (j0, j1), kept in the workspace: the
  // addresses they hold stay fixed while the workspace and plan do
  auto run_steps = [&](int j0, int j1) -> hipError_t {
    if (!opt.graph) return launch_steps(j0, j1);
    for (auto& x : ws.graphs)
      if (x.first == std::make_pair(j0, j1)) return hipGraphLaunch(x.second, s);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    e = launch_steps(j0, j1);
    hipError_t e2 = hipStreamEndCapture(s, &gr);
    if (e != hipSuccess) { if (gr) (void)hipGraphDestroy(gr); return e; }
    if (e2 != hipSuccess) return e2;
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return e;
    ws.graphs.push_back({{j0, j1}, ge});
    return hipGraphLaunch(ge, s);
  };

  // per-instance state.  T: projected matrix ((m+1) x (m+1), row-major) of the
  // current cycle; rows [0, kk) hold the kept Ritz values and the arrow column.
  std::vector<std::vector<double>> T(B, std::vector<double>(size_t(m + 1) * (m + 1), 0.0));
  // done: the Krylov space of the instance is exhausted (its pairs are exact);
  // placed: its Ritz vectors already sit in Q[:, 0:nev] (set by a restart).
  std::vector<int> done(B, 0), placed(B, 0), hdim(B, -1), mm_of(B, 0);
  std::vector<double> theta_all(size_t(B) * nev, 0.0), res_all(size_t(B) * nev, 0.0);
  std::vector<std::vector<double>> Sv(B);  // full eigenvectors where computed (mm x mm, col-major)
  for (int i = 0; i < B; ++i) converged_at[i] = -1;
  int matvecs = 0, restarts = 0, cycle_start = 0;
  int matvecs_run = 0;  // H*v launched, including a block queued past the final test
  const int kk = std::min(keep, m - 1);

  // checkpoint at every thick restart: the restart state (kept Ritz vectors and
  // residual vector, projected matrices, per-instance flags), FP64, so that a
  // later solve with the same options resumes exactly where this one stopped
  const std::string ckpt = opt.checkpoint ? std::string(opt.checkpoint) : std::string();
  const char kMagic[8] = {'N', 'S', 'X', 'T', 'R', 'L', '0', '1'};
  auto save_ckpt = [&]() -> int {
    std::vector<double> hq(size_t(N) * (kk + 1));
    TRL_HIP(hipMemcpyAsync(hq.data(), Q.p, sizeof(double) * hq.size(), hipMemcpyDeviceToHost, s));
    TRL_HIP(hipStreamSynchronize(s));
    const std::string tmp = ckpt + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) { set_error("cannot open " + tmp + " for writing"); return NSX_EIO; }
    const int32_t hdr[6] = {B, m, kk, nev, matvecs, restarts};
    const int64_t n64 = N;
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 6, f) == 6 && std::fwrite(&n64, 8, 1, f) == 1 &&
              std::fwrite(done.data(), 4, B, f) == (size_t)B && std::fwrite(placed.data(), 4, B, f) == (size_t)B &&
              std::fwrite(converged_at, 4, B, f) == (size_t)B &&
              std::fwrite(theta_all.data(), 8, theta_all.size(), f) == theta_all.size() &&
              std::fwrite(res_all.data(), 8, res_all.size(), f) == res_all.size();
    for (int i = 0; ok && i < B; ++i) ok = std::fwrite(T[i].data(), 8, T[i].size(), f) == T[i].size();
    ok = ok && std::fwrite(hq.data(), 8, hq.size(), f) == hq.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), ckpt.c_str()) != 0) { set_error("write failed: " + ckpt); return NSX_EIO; }
    return NSX_OK;
  };
  if (!ckpt.empty()) {
    if (FILE* f = std::fopen(ckpt.c_str(), "rb")) {
      char mg[8];
      int32_t hdr[6];
      int64_t n64 = 0;
      bool ok = std::fread(mg, 1, 8, f) == 8 && std::memcmp(mg, kMagic, 8) == 0 && std::fread(hdr, 4, 6, f) == 6 &&
                std::fread(&n64, 8, 1, f) == 1;
      if (!ok || hdr[0] != B || hdr[1] != m || hdr[2] != kk || hdr[3] != nev || n64 != N) {
        std::fclose(f);
        set_error("checkpoint " + ckpt + " does not match this solve (instances, basis, keep, nev or size)");
        return NSX_EFORMAT;
      }
      std::vector<double> hq(size_t(N) * (kk + 1));
      ok = std::fread(done.data(), 4, B, f) == (size_t)B && std::fread(placed.data(), 4, B, f) == (size_t)B &&
           std::fread(converged_at, 4, B, f) == (size_t)B &&
           std::fread(theta_all.data(), 8, theta_all.size(), f) == theta_all.size() &&
           std::fread(res_all.data(), 8, res_all.size(), f) == res_all.size();
      for (int i = 0; ok && i < B; ++i) ok = std::fread(T[i].data(), 8, T[i].size(), f) == T[i].size();
      ok = ok && std::fread(hq.data(), 8, hq.size(), f) == hq.size();
      std::fclose(f);
      if (!ok) { set_error("checkpoint " + ckpt + " is truncated"); return NSX_EFORMAT; }
      TRL_HIP(hipMemcpy(Q.p, hq.data(), sizeof(double) * hq.size(), hipMemcpyHostToDevice));
      matvecs = matvecs_run = hdr[4];
      restarts = hdr[5];
      cycle_start = kk;
    }
  }
  double* const ha = ws.hha;  // pinned: filled by asynchronous copies
  double* const hb = ws.hhb;
  std::vector<std::vector<double>> LC(B, std::vector<double>(m + 1, 0.0));  // last eigenvector rows
  double host_s = 0.0;  // wall time of the host projected-problem work

  // Rayleigh-Ritz of instance i on its current mm x mm matrix: eigenvalues into
  // Th[i], eigenvectors into Sv[i] when full, else only their last components.
  std::vector<std::vector<double>> Th(B, std::vector<double>(m + 1, 0.0));
  auto solve = [&](int i, int mm, bool full) -> bool {
    std::vector<double> Tm(size_t(mm) * mm);
    for (int a = 0; a < mm; ++a)
      for (int b = 0; b < mm; ++b) Tm[size_t(a) * mm + b] = T[i][size_t(a) * (m + 1) + b];
    mm_of[i] = mm;
    if (!full) return sym_eig(Tm.data(), mm, false, Th[i].data(), LC[i].data());
    Sv[i].assign(size_t(mm) * mm, 0.0);
    if (!sym_eig(Tm.data(), mm, true, Th[i].data(), Sv[i].data())) return false;
    for (int c = 0; c < mm; ++c) LC[i][c] = Sv[i][size_t(c) * mm + mm - 1];
    return true;
  };
  // record the wanted pairs' values and residuals |beta_last s_{mm-1,q}| after solve()
  // standard Lanczos (Lnzs, src/NuTRL/NuCoreX.f90:595-751): no restarts, and a
  // pair counts as converged when its Ritz value moved by <= tol since the
  // previous test; the lowest min(nev + 1, dimension) values are watched (:719-737)
  const bool standard = opt.standard != 0;
  const int nwatch = nev + 1;
  std::vector<std::vector<double>> prev(B, std::vector<double>(nwatch, NAN));
  auto record = [&](int i, double beta_last) {
    int conv = 0;
    for (int q = 0; q < nev; ++q) {
      // a Krylov space that closed after mm_of[i] < nev steps holds only mm_of[i] pairs
      const bool have = q < mm_of[i];
      theta_all[size_t(i) * nev + q] = have ? Th[i][q] : NAN;
      const double r = have ? std::fabs(beta_last * LC[i][q]) : NAN;
      res_all[size_t(i) * nev + q] = r;
      conv += r <= tol;
    }
    if (standard) {
      const int nw = std::min(nwatch, mm_of[i]);
      bool stable = true;
      for (int q = 0; q < nw; ++q) {
        stable = stable && std::fabs(Th[i][q] - prev[i][q]) <= tol;  // NaN (no previous test) fails
        prev[i][q] = Th[i][q];
      }
      conv = stable ? nev : 0;
    }
    if (conv == nev && converged_at[i] < 0) converged_at[i] = matvecs;
  };

  bool finished = false;
  while (!finished) {
    // ---- one restart cycle: steps [cycle_start, m), tested every `check` steps
    for (int i = 0; i < B; ++i) hdim[i] = done[i] ? 0 : -1;  // 0: frozen, -1: growing
    TRL_HIP(hipMemcpyAsync(dim.p, hdim.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    // Blocks of `check` steps.  The next block of the cycle is queued before the
    // host tests the current one, so the projected-problem work on the host
    // overlaps the GPU; its results only ever extend the basis.
    auto queue_block = [&](int b0, int b1, int slot) -> int {
      TRL_HIP(run_steps(b0, b1));
      TRL_HIP(hipMemcpyAsync(ha + (size_t)b0 * B, ah.as<double>() + (size_t)b0 * B,
                             sizeof(double) * (b1 - b0) * B, hipMemcpyDeviceToHost, s));
      TRL_HIP(hipMemcpyAsync(hb + (size_t)b0 * B, bh.as<double>() + (size_t)b0 * B,
                             sizeof(double) * (b1 - b0) * B, hipMemcpyDeviceToHost, s));
      TRL_HIP(hipMemcpyAsync(ws.hdim[slot], dim.p, sizeof(int) * B, hipMemcpyDeviceToHost, s));
      TRL_HIP(hipEventRecord(ws.ev[slot], s));
      matvecs_run += b1 - b0;
      return NSX_OK;
    };
    bool all_conv = false;
    int j0 = cycle_start, j1 = std::min(m, cycle_start + check), slot = 0;
    if (int rc = queue_block(j0, j1, slot)) return rc;
    while (true) {
      const bool spec = j1 < m && matvecs_run < max_iter;
      const int k0 = j1, k1 = std::min(m, j1 + check);
      if (spec)
        if (int rc = queue_block(k0, k1, slot ^ 1)) return rc;
      TRL_HIP(hipEventSynchronize(ws.ev[slot]));
      matvecs += j1 - j0;
      const int* hd = ws.hdim[slot];
      for (int x = j0 * B; x < j1 * B; ++x)
        if (!std::isfinite(ha[x]) || !std::isfinite(hb[x])) { set_error("non-finite Lanczos coefficient"); return NSX_EGPU; }
      std::atomic<int> failed(0);
      const auto th0 = std::chrono::steady_clock::now();
      parallel_for(B, [&](int i) {
        if (done[i]) return;
        auto t = [&](int a, int b) -> double& { return T[i][size_t(a) * (m + 1) + b]; };
        const bool broke = hd[i] > 0;
        const int last = broke ? hd[i] : j1;
        // alpha_j on the diagonal; beta_{j-1} couples j-1 and j, also across the
        // boundary of the previous block (not across the restart: the arrow does)
        for (int j = j0; j < last; ++j) {
          t(j, j) = ha[(size_t)j * B + i];
          if (j > cycle_start) t(j - 1, j) = t(j, j - 1) = hb[(size_t)(j - 1) * B + i];
        }
        const double beta_last = broke ? 0.0 : hb[(size_t)(last - 1) * B + i];
        if (!solve(i, last, broke)) { failed = 1; return; }
        record(i, beta_last);
        if (broke) {  // invariant subspace: eigenpairs exact, Ritz vectors from Sv
          done[i] = 1;
          // fewer pairs than wanted: the missing ones stay NaN and unconverged
          if (converged_at[i] < 0 && last >= nev) converged_at[i] = matvecs;
        }
      }, 2);
      host_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - th0).count();
      if (failed) { set_error("projected eigenproblem did not converge"); return NSX_EGPU; }
      all_conv = true;  // every instance converged, or exhausted (done) with nothing left to find
      for (int i = 0; i < B; ++i) all_conv = all_conv && (converged_at[i] >= 0 || done[i]);
      if (all_conv || matvecs >= max_iter || !spec) break;
      j0 = k0;
      j1 = k1;
      slot ^= 1;
    }
    if (all_conv || matvecs >= max_iter || standard) { finished = true; break; }  // standard: basis full

    // ---- thick restart of every instance: Q[:, 0:kk] = Q S, Q[:, kk] = q_m
    std::vector<double> Sdev(size_t(B) * m * kk, 0.0);
    std::atomic<int> failed(0);
    const auto th0 = std::chrono::steady_clock::now();
    parallel_for(B, [&](int i) {
      double* Si = &Sdev[(size_t)i * m * kk];
      if (placed[i]) {  // frozen instance whose vectors are already Q[:, 0:kk]
        for (int c = 0; c < kk; ++c) Si[(size_t)c * m + c] = 1.0;
        return;
      }
      if (done[i]) {  // vectors from the solve at the breakdown; frozen from now on
        const int mm = mm_of[i];
        for (int c = 0; c < std::min(kk, mm); ++c)
          for (int j = 0; j < mm; ++j) Si[(size_t)c * m + j] = Sv[i][size_t(c) * mm + j];
        placed[i] = 1;
        return;
      }
      auto t = [&](int a, int b) -> double& { return T[i][size_t(a) * (m + 1) + b]; };
      const double beta_last = hb[(size_t)(m - 1) * B + i];
      if (!solve(i, m, true)) { failed = 1; return; }
      for (int c = 0; c < kk; ++c)
        for (int j = 0; j < m; ++j) Si[(size_t)c * m + j] = Sv[i][size_t(c) * m + j];
      std::fill(T[i].begin(), T[i].end(), 0.0);
      for (int c = 0; c < kk; ++c) {
        t(c, c) = Th[i][c];
        const double b = beta_last * Sv[i][size_t(c) * m + (m - 1)];
        t(c, kk) = b;
        t(kk, c) = b;
      }
    }, 1);  // full eigenvectors of every instance: one thread each
    host_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - th0).count();
    if (failed) { set_error("projected eigenproblem did not converge"); return NSX_EGPU; }
    TRL_HIP(hipMemcpyAsync(S.p, Sdev.data(), sizeof(double) * Sdev.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_seg_ritz, gv, dim3(kVT), 0, s, Y.as<double>(), Q.as<double>(), ldq, S.as<double>(), m,
                       kk, inst_of.as<int>(), N);
    TRL_HIP(hipGetLastError());
    TRL_HIP(hipMemcpyAsync(Q.p, Y.p, sizeof(double) * N * kk, hipMemcpyDeviceToDevice, s));
    TRL_HIP(hipMemcpyAsync(Q.as<double>() + ldq * kk, Q.as<double>() + ldq * m, sizeof(double) * N,
                           hipMemcpyDeviceToDevice, s));
    cycle_start = kk;
    ++restarts;
    if (!ckpt.empty()) {
      const int rc = save_ckpt();
      if (rc != NSX_OK) return rc;
    }
  }
  for (int i = 0; i < B; ++i)
    for (int q = 0; q < nev; ++q) {
      evals[size_t(i) * nev + q] = theta_all[size_t(i) * nev + q];
      if (residuals) residuals[size_t(i) * nev + q] = res_all[size_t(i) * nev + q];
    }
  if (evecs) {
    // Ritz vectors of the wanted pairs of every instance from the current basis
    int mmax = nev;
    for (int i = 0; i < B; ++i) mmax = std::max(mmax, mm_of[i]);
    std::vector<double> Sk(size_t(B) * mmax * nev, 0.0);
    std::atomic<int> failed(0);
    parallel_for(B, [&](int i) {
      double* Si = &Sk[(size_t)i * mmax * nev];
      if (placed[i]) {
        for (int c = 0; c < nev; ++c) Si[(size_t)c * mmax + c] = 1.0;
        return;
      }
      const int mm = mm_of[i];
      if (!done[i] && !solve(i, mm, true)) { failed = 1; return; }  // the last test formed only the last row
      for (int c = 0; c < std::min(nev, mm); ++c)
        for (int j = 0; j < mm; ++j) Si[(size_t)c * mmax + j] = Sv[i][size_t(c) * mm + j];
    }, 1);
    if (failed) { set_error("projected eigenproblem did not converge"); return NSX_EGPU; }
    TRL_HIP(hipMemcpyAsync(S.p, Sk.data(), sizeof(double) * Sk.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_seg_ritz, gv, dim3(kVT), 0, s, Y.as<double>(), Q.as<double>(), ldq, S.as<double>(), mmax,
                       nev, inst_of.as<int>(), N);
    TRL_HIP(hipGetLastError());
    TRL_HIP(hipMemcpyAsync(evecs, Y.p, sizeof(double) * N * nev, hipMemcpyDeviceToHost, s));
    TRL_HIP(hipStreamSynchronize(s));
  }
  if (stats) {
    int conv = 0;
    for (int i = 0; i < B; ++i) conv += converged_at[i] >= 0;
    stats->matvecs = matvecs;  // H*v up to the final test (one queued block may have run past it)
    stats->restarts = restarts;
    stats->converged = conv;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats->host_seconds = host_s;
  }
#undef TRL_HIP
  return NSX_OK;
}

}  // namespace nsx

using namespace nsx;

extern "C" int nsx_trl_solve(nsx_hv* h, const nsx_trl_opts* opt, const double* start, double* evals,
                             double* evecs, double* residuals, nsx_trl_stats* stats) {
  if (!h || !opt || !evals) { set_error("null argument"); return NSX_EARG; }
  if (h->host_only) { set_error("host-only handle: solve it through a batch"); return NSX_EARG; }
  std::vector<Instance> inst{instance_of(h->model, 0)};
  nsx_trl_opts o = *opt;
  o.max_basis = (int32_t)std::min<int64_t>(o.max_basis, inst[0].nstates);
  int32_t conv_at = -1;
  int rc = trl_run(h->dplan, h->lowp, h->precision, h->stream, h->trl, inst, o, start, evals, evecs, residuals,
                   &conv_at, stats);
  if (rc == NSX_OK && stats) stats->converged = conv_at >= 0 ? opt->nev : 0;
  if (rc == NSX_OK && opt->nev > 0 && std::isnan(evals[opt->nev - 1])) {
    set_error("Krylov space closed with fewer than nev eigenpairs (the start vector lies in a small invariant "
              "subspace, e.g. an eigenvector, or H has few distinct eigenvalues); the missing ones are NaN");
    return NSX_ECAP;
  }
  return rc;
}

extern "C" int nsx_batch_trl_solve(nsx_batch* b, const nsx_trl_opts* opt, const double* start, double* evals,
                                   double* evecs, double* residuals, int32_t* converged_at, nsx_trl_stats* stats) {
  if (!b || !opt || !evals || !converged_at) { set_error("null argument"); return NSX_EARG; }
  std::vector<Instance> inst;
  int64_t min_st = INT64_MAX;
  for (size_t i = 0; i < b->models.size(); ++i) {
    inst.push_back(instance_of(*b->models[i], b->hplan.vbase[i]));
    min_st = std::min(min_st, inst.back().nstates);
  }
  nsx_trl_opts o = *opt;
  o.max_basis = (int32_t)std::min<int64_t>(o.max_basis, std::max<int64_t>(min_st, o.nev + 2));
  return trl_run(b->dplan, b->lowp, b->precision, b->stream, b->trl, inst, o, start, evals, evecs, residuals,
                 converged_at, stats);
}
