// HIP kernels of the pn J-scheme H*v for gfx950.
//
// Math per output block L = (Ja, Jb) (column-major d(Ja) x d(Jb)):
//   W_L = sum_items(L) sum_{a-terms} A_k Y_k        Y_k = sum_R r_R V_R Bhat_R^T
//       + H_a V_L + V_L H_b^T
// The block contractions are templated on the operand storage type T:
//   double   FP64 operands, FP64 accumulation (the plain drop-in)
//   float    FP32 operands, FP32 accumulation
//   Bf16x2   BF16 hi/lo pairs x = hi + lo, a*b = ah*bh + ah*bl + al*bh, FP32 accumulation
// W is always written in FP64.  FP64 vector FMA runs at the FP64 matrix-core
// rate on MI355X (78.6 TFLOP/s both), so the ragged, tiny (<= 33 x 33 in the sd
// shell) contractions use VALU FMAs with 4 output columns per thread; blocks
// with a side above 40 (pf shell, up to 615) run as 64 x 64 matrix-core tiles
// (k_big0/1/2: FP64 MFMA in FP64, FP32 MFMA in FP32; k_big*_bf: BF16 MFMA for
// BF16 hi/lo pairs).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "nsx_kernels.hpp"

namespace nsx {

// ---------------------------------------------------------------- precision traits
template <typename T>
struct Prec;
template <>
struct Prec<double> {
  using acc = double;
  __device__ static double fma3(double a, double b, double s) { return fma(a, b, s); }
  __device__ static double from(double x) { return x; }
};
template <>
struct Prec<float> {
  using acc = float;
  __device__ static float fma3(float a, float b, float s) { return fmaf(a, b, s); }
  __device__ static float from(float x) { return x; }
};
__device__ __forceinline__ float bf_hi(Bf16x2 x) { return __uint_as_float(x.bits & 0xffff0000u); }
__device__ __forceinline__ float bf_lo(Bf16x2 x) { return __uint_as_float(x.bits << 16); }
template <>
struct Prec<Bf16x2> {
  using acc = float;
  // split product: the lo*lo term is below the pair's resolution and is dropped
  __device__ static float fma3(Bf16x2 a, Bf16x2 b, float s) {
    const float ah = bf_hi(a), al = bf_lo(a), bh = bf_hi(b), bl = bf_lo(b);
    s = fmaf(al, bh, s);
    s = fmaf(ah, bl, s);
    return fmaf(ah, bh, s);
  }
  __device__ static Bf16x2 from(float x) { return bf16x2_split(x); }
};

// FP64 vector -> working precision (input of every reduced-precision H*v)
template <typename T>
__global__ __launch_bounds__(256) void k_convert(const double* __restrict__ x, T* __restrict__ y, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i < n) y[i] = Prec<T>::from((typename Prec<T>::acc)x[i]);
}

// The pn part's input on truncated plans: V masked to the window states
// (mask 0 elsewhere), in the working precision
template <typename T>
__global__ __launch_bounds__(256) void k_mask_in(const double* __restrict__ x, const uint8_t* __restrict__ mask,
                                                 T* __restrict__ y, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (i < n) y[i] = Prec<T>::from((typename Prec<T>::acc)(mask[i] ? x[i] : 0.0));
}

// ---------------------------------------------------------------- direct kernel (large blocks)
template <typename T, bool ATOMIC>
__global__ __launch_bounds__(kThreads) void k_pn_partial(
    const DItem* __restrict__ items, const DATerm* __restrict__ aterms,
    const DRTerm* __restrict__ rterms, const T* __restrict__ V, const T* __restrict__ Ad,
    const T* __restrict__ Bh, double* __restrict__ part, double* __restrict__ W) {
  using Acc = typename Prec<T>::acc;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Y = reinterpret_cast<T*>(smem_raw);
  const DItem it = items[blockIdx.x];
  const int da = it.da, dar = it.dar, db = it.db, nc = it.nc, j0 = it.j0;
  const int nW = da * nc, nY = dar * nc;
  Acc w[kMaxE];
#pragma unroll
  for (int q = 0; q < kMaxE; ++q) w[q] = Acc(0);
  for (int a = it.a_begin; a < it.a_end; ++a) {
    const DATerm Tm = aterms[a];
    for (int e = threadIdx.x; e < nY; e += kThreads) {
      const int i = e % dar, jj = e / dar;
      const int j = j0 + jj;
      Acc y = Acc(0);
      for (int u = Tm.r_begin; u < Tm.r_end; ++u) {
        const DRTerm R = rterms[u];
        const T* v = V + R.v_off + i;
        const T* b = Bh + R.b_off + (int64_t)R.dbr * j;  // Bhat^T column j
        Acc s = Acc(0);
        for (int k = 0; k < R.dbr; ++k) s = Prec<T>::fma3(v[(int64_t)dar * k], b[k], s);
        y += Acc(R.r) * s;
      }
      Y[e] = Prec<T>::from(y);
    }
    __syncthreads();
    const T* A = Ad + Tm.a_off;
#pragma unroll
    for (int q = 0; q < kMaxE; ++q) {
      const int e = threadIdx.x + kThreads * q;
      if (e < nW) {
        const int i = e % da, jj = e / da;
        Acc s = Acc(0);
        for (int k = 0; k < dar; ++k) s = Prec<T>::fma3(A[i + (int64_t)da * k], Y[k + dar * jj], s);
        w[q] += s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < kMaxE; ++q) {
    const int e = threadIdx.x + kThreads * q;
    if (e < nW) {
      const int i = e % da, jj = e / da;
      const int64_t ij = i + (int64_t)da * (j0 + jj);
      if (ATOMIC) unsafeAtomicAdd(W + it.w_off + ij, (double)w[q]);
      else part[it.part_off + ij] = (double)w[q];
    }
  }
}

// ---------------------------------------------------------------- staged kernel (sd-size blocks)
typedef __attribute__((address_space(3))) void lds_void_t;

// Copies n dwords src[0..n) to LDS dst[0..n) with direct global->LDS loads: the
// data bypasses VGPRs and lands at (wave-uniform base) + lane * 4 bytes.
template <int NT>
__device__ __forceinline__ void dma_run(const uint32_t* __restrict__ src, uint32_t* dst, int n) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int base = wave * 64; base < n; base += NT) {
    const int e = base + lane;
    if (e < n) __builtin_amdgcn_global_load_lds(src + e, (lds_void_t*)(dst + base), 4, 0, 0);
  }
}

// One workgroup (NT threads) runs a chunk of an item's phases.  A phase's
// operand runs (V_R blocks, Bhat^T blocks, A) are staged in LDS, either copied
// through registers at the start of the phase (MODE 0, single buffer) or
// fetched one phase ahead by direct global->LDS loads into a second buffer
// (MODE 1) so the copy of phase p+1 overlaps the arithmetic of phase p.  Y
// (dar x db) and W (da x db) accumulate in registers, MU (row, 4-column) units
// per thread: one LDS read of V (or A) feeds 4 FMAs against broadcast reads of
// Bhat^T (or Y).
template <typename T, int NT, int MU, int MODE, bool ATOMIC>
__global__ __launch_bounds__(NT) void k_pn_staged(
    const DItem2* __restrict__ items, const DPhase* __restrict__ phases,
    const DPair* __restrict__ pairs, const T* __restrict__ V, const T* __restrict__ Ad,
    const T* __restrict__ Bt, double* __restrict__ part, double* __restrict__ W, int lds_buf) {
  using Acc = typename Prec<T>::acc;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  T* bufs[2] = {sm, sm + (MODE == 1 ? lds_buf : 0)};
  T* sY = sm + (MODE == 1 ? 2 : 1) * lds_buf;
  const DItem2 it = items[blockIdx.x];
  const int da = it.da, dar = it.dar, db = it.db;
  const int ncb = (db + kColBlk - 1) / kColBlk;
  const int nuY = dar * ncb, nuW = da * ncb;
  const int tid = threadIdx.x;
  constexpr int kDw = sizeof(T) / 4;  // dwords per element
  Acc w[MU][kColBlk], y[MU][kColBlk];
#pragma unroll
  for (int q = 0; q < MU; ++q)
#pragma unroll
    for (int c = 0; c < kColBlk; ++c) { w[q][c] = Acc(0); y[q][c] = Acc(0); }

  auto issue = [&](const DPhase& P, T* dst) {
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
    dma_run<NT>(reinterpret_cast<const uint32_t*>(V + P.v_off), d, kDw * P.nv);
    dma_run<NT>(reinterpret_cast<const uint32_t*>(Bt + P.b_off), d + kDw * P.nv, kDw * P.nb);
    if (P.na > 0)
      dma_run<NT>(reinterpret_cast<const uint32_t*>(Ad + P.a_off), d + kDw * (P.nv + P.nb), kDw * P.na);
  };

  if (MODE == 1 && it.ph_begin < it.ph_end) issue(phases[it.ph_begin], bufs[0]);
  for (int p = it.ph_begin; p < it.ph_end; ++p) {
    const DPhase P = phases[p];
    T* buf = bufs[(p - it.ph_begin) & 1];
    if (MODE == 1) {
      __builtin_amdgcn_s_waitcnt(0);  // this wave's loads of phase p have landed
      __syncthreads();                // ... and every other wave's; phase p-1 is done
      if (p + 1 < it.ph_end) issue(phases[p + 1], bufs[(p + 1 - it.ph_begin) & 1]);
    } else {
      const int nvb = P.nv + P.nb, nt = nvb + P.na;
      __syncthreads();  // the previous phase has finished reading buf and sY
#pragma unroll 8
      for (int e = tid; e < nt; e += NT)
        buf[e] = e < P.nv ? V[P.v_off + e] : (e < nvb ? Bt[P.b_off + (e - P.nv)] : Ad[P.a_off + (e - nvb)]);
      __syncthreads();
    }
    // Y += r V_R Bhat^T for every pair of the phase
    for (int pr = P.pair_begin; pr < P.pair_end; ++pr) {
      const DPair Q = pairs[pr];
      const T* sV = buf + Q.v_lds;
      const T* sB = buf + Q.b_lds;
      const int dbr = Q.dbr;
#pragma unroll
      for (int q = 0; q < MU; ++q) {
        const int u = tid + NT * q;
        if (u < nuY) {
          const int i = u % dar, j0 = (u / dar) * kColBlk;
          const T* bc[kColBlk];
#pragma unroll
          for (int c = 0; c < kColBlk; ++c) bc[c] = sB + dbr * min(j0 + c, db - 1);
          Acc s0[kColBlk], s1[kColBlk];
#pragma unroll
          for (int c = 0; c < kColBlk; ++c) { s0[c] = Acc(0); s1[c] = Acc(0); }
          int k = 0;
          for (; k + 1 < dbr; k += 2) {
            const T v0 = sV[i + dar * k], v1 = sV[i + dar * (k + 1)];
#pragma unroll
            for (int c = 0; c < kColBlk; ++c) {
              s0[c] = Prec<T>::fma3(v0, bc[c][k], s0[c]);
              s1[c] = Prec<T>::fma3(v1, bc[c][k + 1], s1[c]);
            }
          }
          if (k < dbr) {
            const T v0 = sV[i + dar * k];
#pragma unroll
            for (int c = 0; c < kColBlk; ++c) s0[c] = Prec<T>::fma3(v0, bc[c][k], s0[c]);
          }
#pragma unroll
          for (int c = 0; c < kColBlk; ++c) y[q][c] += Acc(Q.r) * (s0[c] + s1[c]);
        }
      }
    }
    if (P.na > 0) {
      // publish Y, then W += A Y with A at the end of the phase's buffer
#pragma unroll
      for (int q = 0; q < MU; ++q) {
        const int u = tid + NT * q;
        if (u < nuY) {
          const int i = u % dar, j0 = (u / dar) * kColBlk;
#pragma unroll
          for (int c = 0; c < kColBlk; ++c)
            if (j0 + c < db) sY[i + dar * (j0 + c)] = Prec<T>::from(y[q][c]);
        }
#pragma unroll
        for (int c = 0; c < kColBlk; ++c) y[q][c] = Acc(0);
      }
      __syncthreads();
      const T* sA = buf + P.nv + P.nb;
#pragma unroll
      for (int q = 0; q < MU; ++q) {
        const int u = tid + NT * q;
        if (u < nuW) {
          const int i = u % da, j0 = (u / da) * kColBlk;
          const T* yc[kColBlk];
#pragma unroll
          for (int c = 0; c < kColBlk; ++c) yc[c] = sY + dar * min(j0 + c, db - 1);
          Acc s0[kColBlk], s1[kColBlk];
#pragma unroll
          for (int c = 0; c < kColBlk; ++c) { s0[c] = Acc(0); s1[c] = Acc(0); }
          int k = 0;
          for (; k + 1 < dar; k += 2) {
            const T a0 = sA[i + da * k], a1 = sA[i + da * (k + 1)];
#pragma unroll
            for (int c = 0; c < kColBlk; ++c) {
              s0[c] = Prec<T>::fma3(a0, yc[c][k], s0[c]);
              s1[c] = Prec<T>::fma3(a1, yc[c][k + 1], s1[c]);
            }
          }
          if (k < dar) {
            const T a0 = sA[i + da * k];
#pragma unroll
            for (int c = 0; c < kColBlk; ++c) s0[c] = Prec<T>::fma3(a0, yc[c][k], s0[c]);
          }
#pragma unroll
          for (int c = 0; c < kColBlk; ++c) w[q][c] += s0[c] + s1[c];
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < MU; ++q) {
    const int u = tid + NT * q;
    if (u < nuW) {
      const int i = u % da, j0 = (u / da) * kColBlk;
#pragma unroll
      for (int c = 0; c < kColBlk; ++c)
        if (j0 + c < db) {
          const int64_t ij = i + (int64_t)da * (j0 + c);
          if (ATOMIC) unsafeAtomicAdd(W + it.w_off + ij, (double)w[q][c]);
          else part[it.part_off + ij] = (double)w[q][c];
        }
    }
  }
}

// ---------------------------------------------------------------- like-particle / reduce
// W_L = (sum of the pn partials of L, when use_partials) + H_a V_L + V_L H_b^T.
// In atomic mode this runs first (use_partials = 0) and the pn kernels add to W.
// A whole block whose V_L, H_a and H_b fit in `cap` elements of LDS is staged
// there (packed: V_L, then H_a, then H_b) and contracted from LDS; otherwise
// the operands are read from global memory.
template <typename T, int NT>
__device__ __forceinline__ void reduce_block(DBlock b, const int64_t* __restrict__ part_list,
                                             const double* __restrict__ part, const T* __restrict__ V,
                                             const T* __restrict__ H, double* __restrict__ W, int use_partials,
                                             T* sm, int cap, int skip_large) {
  using Acc = typename Prec<T>::acc;
  if (skip_large && (b.da > kSmallMax || b.db > kSmallMax)) return;  // the large-block GEMM tiles do it
  if (!use_partials) b.p_end = b.p_begin;
  const int da = b.da, db = b.db, nW = da * b.nc;
  const T* v = V + b.v_off;
  const int nv = da * db;
  if (b.j0 == 0 && b.nc == db && nv + da * da + db * db <= cap) {
    T* sv = sm;
    T* sha = sm + nv;
    T* shb = sha + da * da;
    for (int e = threadIdx.x; e < nv; e += NT) sv[e] = v[e];
    if (b.ha_off >= 0)
      for (int e = threadIdx.x; e < da * da; e += NT) sha[e] = H[b.ha_off + e];
    if (b.hb_off >= 0)
      for (int e = threadIdx.x; e < db * db; e += NT) shb[e] = H[b.hb_off + e];
    __syncthreads();
    for (int e = threadIdx.x; e < nv; e += NT) {
      const int i = e % da, j = e / da;
      double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
      int p = b.p_begin;
      for (; p + 3 < b.p_end; p += 4) {
        w0 += part[part_list[p] + e];
        w1 += part[part_list[p + 1] + e];
        w2 += part[part_list[p + 2] + e];
        w3 += part[part_list[p + 3] + e];
      }
      for (; p < b.p_end; ++p) w0 += part[part_list[p] + e];
      double w = (w0 + w1) + (w2 + w3);
      if (b.ha_off >= 0) {
        Acc s = Acc(0);
        for (int k = 0; k < da; ++k) s = Prec<T>::fma3(sha[i + da * k], sv[k + da * j], s);
        w += (double)s;
      }
      if (b.hb_off >= 0) {
        Acc s = Acc(0);
        for (int k = 0; k < db; ++k) s = Prec<T>::fma3(sv[i + da * k], shb[j + db * k], s);
        w += (double)s;
      }
      W[b.v_off + e] = w;
    }
    return;
  }
  for (int e = threadIdx.x; e < nW; e += NT) {
    const int i = e % da, j = b.j0 + e / da;
    const int64_t ij = i + (int64_t)da * j;
    double w = 0.0;
    for (int p = b.p_begin; p < b.p_end; ++p) w += part[part_list[p] + ij];
    if (b.ha_off >= 0) {
      const T* Ha = H + b.ha_off;
      Acc s = Acc(0);
      for (int k = 0; k < da; ++k) s = Prec<T>::fma3(Ha[i + (int64_t)da * k], v[k + (int64_t)da * j], s);
      w += (double)s;
    }
    if (b.hb_off >= 0) {
      const T* Hb = H + b.hb_off;
      Acc s = Acc(0);
      for (int k = 0; k < db; ++k) s = Prec<T>::fma3(v[i + (int64_t)da * k], Hb[j + (int64_t)db * k], s);
      w += (double)s;
    }
    W[b.v_off + ij] = w;
  }
}

constexpr int kReduceLds = 3 * kSmallMax * kSmallMax;

// W = 0 before the fused matrix-core path adds every term (pn and like-particle)
__global__ __launch_bounds__(kThreads) void k_zero_w(double* __restrict__ W, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) W[i] = 0.0;
}

// Linear ensembles (atomic mode, no partials): blockIdx.y is the sample, whose
// vectors and H store sit at the strides es (zero for one sample).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_reduce(
    const DBlock* __restrict__ blocks, const int64_t* __restrict__ part_list,
    const double* __restrict__ part, const T* __restrict__ V, const T* __restrict__ H,
    double* __restrict__ W, int use_partials, int skip_large, EnsStride es) {
  __shared__ T sm[kReduceLds];
  const int64_t smp = blockIdx.y;
  V += smp * es.v;
  W += smp * es.v;
  H += smp * es.h;
  reduce_block<T, kThreads>(blocks[blockIdx.x], part_list, part, V, H, W, use_partials, sm, kReduceLds,
                            skip_large);
}

// ---------------------------------------------------------------- stacked two-stage path
// Stage 1: one workgroup per chunk of (V_R, Bhat^T) pairs sharing one Bhat^T
// block.  Bhat^T and the chunk's V_R blocks are staged in LDS with one round of
// loads; each work unit is (row i, 4 columns) of one pair's X = V_R Bhat^T,
// written to the X buffer in the accumulation type.
// Index of the last segment whose start (read through `start`) is <= e, by
// binary search over n sorted starts.
template <typename F>
__device__ __forceinline__ int seg_of(int e, int n, F start) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start(mid) <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

constexpr int kStkBatch = 8;  // loads each thread keeps in flight when staging

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Map them so that XCD x takes the contiguous chunk range [x*per, (x+1)*per):
// neighbouring chunks share operands, which then stay in one XCD's L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_chunk(int b, int n) {
  const int per = (n + kXcds - 1) / kXcds;
  return (b % kXcds) * per + b / kXcds;
}
// Runs of g consecutive tiles on one XCD, the runs dealt round-robin over the
// XCDs: workgroup b (the (b / 8)-th on XCD b % 8) takes tile
// ((b / 8) / g * 8 + b % 8) g + (b / 8) % g.  Neighbouring tiles that share
// operands then share that XCD's L2, and the XCDs stay within g tiles of an
// even load.  The grid is the tile count rounded up to 8 g (tiles beyond it exit).
__device__ __forceinline__ int xcd_group(int b, int g) {
  const int loc = b / kXcds;
  return ((loc / g) * kXcds + b % kXcds) * g + loc % g;
}

// Linear ensembles launch (workgroups per sample) x (samples rounded up to 8):
// the dispatch index L = blockIdx.y * gridDim.x + blockIdx.x runs on XCD L % 8,
// and is mapped so that every workgroup of a sample runs on one XCD, in order
// (sample s on XCD s % 8), which keeps the sample's vectors and operators in
// that XCD's L2.  Returns the workgroup index within the sample, or -1 for
// the padding samples.
__device__ __forceinline__ int ens_block(const EnsStride& es, int64_t& smp) {
  const int gx = gridDim.x;
  const int64_t L = (int64_t)blockIdx.y * gx + blockIdx.x;
  const int64_t r = L / kXcds;
  smp = (r / gx) * kXcds + L % kXcds;
  return smp < es.n ? (int)(r % gx) : -1;
}

// Occupancy targets: both stages fit 8 workgroups of LDS per CU (2048-double
// staging buffers) and, told so, the compiler keeps them within 64 VGPRs
// without spills (stage 1: 60 instead of 114).
// The like-particle blocks (W_L = H_a V_L + V_L H_b^T, which initialise W for
// stage 2's atomics) ride in the same launch: workgroups [0, g_red) take one
// block each, the rest take stage-1 chunks.  g_red is a multiple of 8, so the
// XCD mapping of the chunks is unchanged.
// V feeds the like-particle blocks, Vp (V masked to the N-hbar-omega window on
// truncated instances, else V) the stage-1 chunks.
// Linear ensembles run every sample through the same tables: blockIdx.y is the
// sample, whose vectors, X, Bhat^T and H stores sit at fixed strides (es; all
// zero for one sample).  Grids along x are multiples of 8, so the XCD of a
// workgroup still follows blockIdx.x.
template <typename T, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_stk1(
    const SChunk1* __restrict__ chunks, int n_chunks, const SUnit1* __restrict__ units, const T* __restrict__ V,
    const T* __restrict__ Vp, const T* __restrict__ Bt, typename Prec<T>::acc* __restrict__ X, const DBlock* __restrict__ blocks,
    int n_blocks, int g_red, const T* __restrict__ H, double* __restrict__ W, int lds_cap, int skip_large,
    EnsStride es) {
  using Acc = typename Prec<T>::acc;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  __shared__ SUnit1 su[kStkMeta];
  int bx = blockIdx.x;
  if (es.n > 1) {
    int64_t smp;
    bx = ens_block(es, smp);
    if (bx < 0) return;
    V += smp * es.v;
    Vp += smp * es.v;
    W += smp * es.v;
    Bt += smp * es.b;
    H += smp * es.h;
    X += smp * es.x;
  }
  if (bx < g_red) {
    if (bx < n_blocks) reduce_block<T, NT>(blocks[bx], nullptr, nullptr, V, H, W, 0, sm, lds_cap, skip_large);
    return;
  }
  const int ci = es.n > 1 ? bx - g_red : xcd_chunk(bx - g_red, n_chunks);
  if (ci >= n_chunks) return;
  const SChunk1 c = chunks[ci];
  const int nu = c.u_end - c.u_begin;
  const int tid = threadIdx.x;
  for (int t = tid; t < nu; t += NT) su[t] = units[c.u_begin + t];
  __syncthreads();
  // stage Bhat^T then every V_R block: LDS holds their concatenation, so the
  // LDS index is the element index; kStkBatch independent loads per thread
  const int nb = c.dbr * c.db, ntot = c.lds;
  for (int e0 = tid; e0 < ntot; e0 += NT * kStkBatch) {
    T r[kStkBatch];
#pragma unroll
    for (int q = 0; q < kStkBatch; ++q) {
      const int e = e0 + q * NT;
      if (e < ntot) {
        if (e < nb) {
          r[q] = Bt[c.b_off + e];
        } else {
          const int u = seg_of(e, nu, [&](int x) { return su[x].lds; });
          r[q] = Vp[su[u].v_off + (e - su[u].lds)];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kStkBatch; ++q)
      if (e0 + q * NT < ntot) sm[e0 + q * NT] = r[q];
  }
  __syncthreads();
  const int dbr = c.dbr, db = c.db;
  int u = 0;
  for (int w = tid; w < c.work; w += NT) {
    while (u + 1 < nu && su[u + 1].work0 <= w) ++u;
    const int dar = su[u].dar, lw = w - su[u].work0;
    const int i = lw % dar, j0 = (lw / dar) * kColBlk;
    const T* sV = sm + su[u].lds;
    const T* bc[kColBlk];
#pragma unroll
    for (int q = 0; q < kColBlk; ++q) bc[q] = sm + dbr * min(j0 + q, db - 1);
    Acc s0[kColBlk], s1[kColBlk];
#pragma unroll
    for (int q = 0; q < kColBlk; ++q) { s0[q] = Acc(0); s1[q] = Acc(0); }
    int k = 0;
    for (; k + 1 < dbr; k += 2) {
      const T v0 = sV[i + dar * k], v1 = sV[i + dar * (k + 1)];
#pragma unroll
      for (int q = 0; q < kColBlk; ++q) {
        s0[q] = Prec<T>::fma3(v0, bc[q][k], s0[q]);
        s1[q] = Prec<T>::fma3(v1, bc[q][k + 1], s1[q]);
      }
    }
    if (k < dbr) {
      const T v0 = sV[i + dar * k];
#pragma unroll
      for (int q = 0; q < kColBlk; ++q) s0[q] = Prec<T>::fma3(v0, bc[q][k], s0[q]);
    }
    Acc* x = X + su[u].x_off + i;
#pragma unroll
    for (int q = 0; q < kColBlk; ++q)
      if (j0 + q < db) x[(int64_t)dar * (j0 + q)] = s0[q] + s1[q];
  }
}

// Stage 2: one workgroup per chunk of the a-terms of one left block L.  Every
// term's A and Y = sum_u r_u X_u are staged in LDS in one round (X references
// staged first); each work unit (row i, 4 columns) of W_L sums A_k Y_k over the
// chunk's terms in registers and is added into W once, with FP64 atomics.
template <typename T, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_stk2(const SChunk2* __restrict__ chunks, int n_chunks,
                                             const STerm2* __restrict__ terms, const SXRef* __restrict__ xrefs,
                                             const T* __restrict__ Ad, const typename Prec<T>::acc* __restrict__ X,
                                             double* __restrict__ W, EnsStride es) {
  using Acc = typename Prec<T>::acc;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  __shared__ STerm2 st[kStkTerms];
  __shared__ SXRef sx[kStkXref];
  int ci;
  if (es.n > 1) {
    int64_t smp;
    ci = ens_block(es, smp);
    if (ci < 0) return;
    X += smp * es.x;
    W += smp * es.v;
  } else {
    ci = xcd_chunk(blockIdx.x, n_chunks);
  }
  if (ci >= n_chunks) return;
  const SChunk2 c = chunks[ci];
  const int nt = c.t_end - c.t_begin;
  const int tid = threadIdx.x;
  for (int t = tid; t < nt; t += NT) st[t] = terms[c.t_begin + t];
  const int x0 = c.x_begin, nx = c.x_end - x0;  // issued with the term loads
  for (int x = tid; x < nx; x += NT) sx[x] = xrefs[x0 + x];
  __syncthreads();
  // staging: A elements are plain copies; a Y element sums r X over its term's
  // references.  The X loads of all kStkBatch elements are issued in rounds of
  // kStkRefs references before any is used, so one round costs one latency.
  constexpr int kStkRefs = 2, kStk2Batch = 4;  // 8 X loads in flight: 60 VGPRs, 8 waves per SIMD
  const int da = c.da, db = c.db, ntot = c.lds;
  for (int e0 = tid; e0 < ntot; e0 += NT * kStk2Batch) {
    Acc y[kStk2Batch];
    T a[kStk2Batch];
    int le[kStk2Batch], xb[kStk2Batch], xe[kStk2Batch];
    bool isa[kStk2Batch];
    int nref = 0;
#pragma unroll
    for (int q = 0; q < kStk2Batch; ++q) {
      const int e = e0 + q * NT;
      y[q] = Acc(0);
      isa[q] = false;
      le[q] = 0;
      xb[q] = xe[q] = 0;
      if (e < ntot) {
        const int t = seg_of(e, nt, [&](int x) { return st[x].lds; });
        const int l = e - st[t].lds, na = da * st[t].dar;
        if (l < na) {
          isa[q] = true;
          a[q] = Ad[st[t].a_off + l];
        } else {
          le[q] = l - na;
          xb[q] = st[t].x_begin - x0;
          xe[q] = st[t].x_end - x0;
          nref = max(nref, xe[q] - xb[q]);
        }
      }
    }
    for (int r0 = 0; r0 < nref; r0 += kStkRefs) {
      Acc xv[kStk2Batch][kStkRefs];
#pragma unroll
      for (int q = 0; q < kStk2Batch; ++q)
#pragma unroll
        for (int u = 0; u < kStkRefs; ++u) {
          const int x = xb[q] + r0 + u;
          xv[q][u] = x < xe[q] ? X[sx[x].x_off + le[q]] : Acc(0);
        }
#pragma unroll
      for (int q = 0; q < kStk2Batch; ++q)
#pragma unroll
        for (int u = 0; u < kStkRefs; ++u) {
          const int x = xb[q] + r0 + u;
          if (x < xe[q]) y[q] += Acc(sx[x].r) * xv[q][u];
        }
    }
#pragma unroll
    for (int q = 0; q < kStk2Batch; ++q)
      if (e0 + q * NT < ntot) sm[e0 + q * NT] = isa[q] ? a[q] : Prec<T>::from(y[q]);
  }
  __syncthreads();
  // W_L = [A_1 .. A_T] [Y_1; ..; Y_T]: a long inner dimension over few outputs,
  // so the terms are split over `nsplit` thread groups (split-K).  With
  // nsplit > 1 every thread has at most one work unit: its partial waits in
  // registers until all threads are done with the staged operands, then the
  // partials take over the staging buffer (which holds NT * kColBlk Acc) and
  // are summed per unit.
  const int ncb = (db + kColBlk - 1) / kColBlk, units = da * ncb;
  int nsplit = 1;
  while (nsplit < 8 && nsplit * 2 * units <= NT && nsplit * 2 <= nt) nsplit *= 2;
  // sum over the terms t = sp, sp + nsplit, ... of A_t(i, :) Y_t(:, j0 .. j0+3)
  auto contract = [&](int unit, int sp, Acc (&acc)[kColBlk]) {
    const int i = unit % da, j0 = (unit / da) * kColBlk;
#pragma unroll
    for (int q = 0; q < kColBlk; ++q) acc[q] = Acc(0);
    for (int t = sp; t < nt; t += nsplit) {
      const int dar = st[t].dar;
      const T* sA = sm + st[t].lds;
      const T* sY = sA + da * dar;
      const T* yc[kColBlk];
#pragma unroll
      for (int q = 0; q < kColBlk; ++q) yc[q] = sY + dar * min(j0 + q, db - 1);
      for (int k = 0; k < dar; ++k) {
        const T a0 = sA[i + da * k];
#pragma unroll
        for (int q = 0; q < kColBlk; ++q) acc[q] = Prec<T>::fma3(a0, yc[q][k], acc[q]);
      }
    }
  };
  if (nsplit == 1) {
    for (int unit = tid; unit < units; unit += NT) {
      Acc acc[kColBlk];
      contract(unit, 0, acc);
      const int i = unit % da, j0 = (unit / da) * kColBlk;
      double* wp = W + c.w_off + i;
#pragma unroll
      for (int q = 0; q < kColBlk; ++q)
        if (j0 + q < db) unsafeAtomicAdd(wp + (int64_t)da * (j0 + q), (double)acc[q]);
    }
    return;
  }
  const bool active = tid < units * nsplit;
  Acc acc[kColBlk];
  if (active) contract(tid % units, tid / units, acc);
  Acc* red = reinterpret_cast<Acc*>(smem_raw);
  __syncthreads();
  if (active) {
#pragma unroll
    for (int q = 0; q < kColBlk; ++q) red[(size_t)tid * kColBlk + q] = acc[q];
  }
  __syncthreads();
  for (int unit = tid; unit < units; unit += NT) {
    const int i = unit % da, j0 = (unit / da) * kColBlk;
    double* wp = W + c.w_off + i;
#pragma unroll
    for (int q = 0; q < kColBlk; ++q) {
      if (j0 + q >= db) continue;
      Acc v = Acc(0);
      for (int sp = 0; sp < nsplit; ++sp) v += red[(size_t)(sp * units + unit) * kColBlk + q];
      unsafeAtomicAdd(wp + (int64_t)da * (j0 + q), (double)v);
    }
  }
}

// ---------------------------------------------------------------- fused column path
// One workgroup per task: the output column Jb of one instance for a range of
// its slabs, in rounds.  Per round r:
//   X   X_x = V_R Bhat_R^T, units (row, 4 columns) of every product; meanwhile
//       the loads of round r+1's tables (products, consumers, references,
//       segments) are in flight, and they are stored to the other table buffer;
//   Y   Y_c = sum_f r_f X_f per consumer, into the space of V_R / Bhat^T;
//       meanwhile round r+1's operands (its V_R, Bhat^T and A blocks) are
//       loaded into registers, kFPre per thread;
//   W   every thread adds A_c Y_c of the round's consumers of its own output
//       units' block, in plan order, to register accumulators;
// then round r+1's operands go from the registers to LDS.  No round waits on a
// global load that was not issued a phase earlier.  The column's outputs are
// owned by the threads (at most U units each) for the whole task and written
// once as its partial: no atomics, and the sum order is the plan's, so the
// H*v is bitwise reproducible.
constexpr int kFPre = 16;  // operand elements of the next round a thread holds in registers

template <typename T>
__device__ __forceinline__ const T* fuse_src(const FSeg& g, const T* V, const T* Ad, const T* Bh) {
  return (g.kind == 0 ? V : g.kind == 1 ? Bh : Ad) + g.src;
}

template <typename T, int NT, int U>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(U <= 2 ? 3 : 2))) void k_fuse(const FTask* __restrict__ tasks, int n_tasks,
                                             const FBlk* __restrict__ blks, const FRound* __restrict__ rounds,
                                             const FSeg* __restrict__ segs, const FX* __restrict__ fxs,
                                             const FCons* __restrict__ fcons, const FRef* __restrict__ frefs,
                                             const T* __restrict__ V, const T* __restrict__ Ad,
                                             const T* __restrict__ Bh, double* __restrict__ part) {
  using Acc = typename Prec<T>::acc;
  static_assert(sizeof(T) == sizeof(Acc), "LDS regions are shared between operand and product types");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  Acc* smx = reinterpret_cast<Acc*>(smem_raw);
  __shared__ FRound srs[kFMaxRounds];
  __shared__ __align__(16) FX sx[2][kFMaxX];
  __shared__ __align__(16) FCons sc[2][kFMaxC];
  __shared__ double sfr[2][kFMaxRef];
  __shared__ int16_t sfx[2][kFMaxRef];
  __shared__ __align__(16) FSeg sg[2][kFMaxSeg];
  const int ti = xcd_chunk((int)blockIdx.x, n_tasks);
  if (ti >= n_tasks) return;
  __shared__ FBlk sb[kFMaxBlk];
  const FTask tk = tasks[ti];
  const int tid = threadIdx.x;
  const int db = tk.db, nr = tk.r_end - tk.r_begin, nblk = tk.b_end - tk.b_begin;
  for (int k = tid; k < nr; k += NT) srs[k] = rounds[tk.r_begin + k];
  if (tid < nblk) sb[tid] = blks[tk.b_begin + tid];
  __syncthreads();
  // this thread's output units of the column: block (task-relative) and unit
  // within the block (row lu % da, columns 4 (lu / da) ..)
  int ub[U], ul[U];
  Acc acc[U][kColBlk];
#pragma unroll
  for (int q = 0; q < U; ++q) {
    const int u = tid + NT * q;
    ub[q] = -1;
    ul[q] = 0;
    if (u < tk.units) {
      const int b = seg_of(u, nblk, [&](int x) { return sb[x].unit0; });
      ub[q] = b;
      ul[q] = u - sb[b].unit0;
    }
#pragma unroll
    for (int c = 0; c < kColBlk; ++c) acc[q][c] = Acc(0);
  }
  // round tables: at most one entry of each table per thread (NT >= the table
  // limits), moved as raw words so that they stay in registers
  static_assert(sizeof(FX) == 12 && sizeof(FCons) == 32 && sizeof(FRef) == 16 && sizeof(FSeg) == 16,
                "round table layouts");
  static_assert(NT >= kFMaxX && NT >= kFMaxC && NT >= kFMaxRef && NT >= kFMaxSeg, "one table entry per thread");
  uint32_t tx0 = 0, tx1 = 0, tx2 = 0;
  uint4 tc = make_uint4(0, 0, 0, 0), tf = make_uint4(0, 0, 0, 0), tg = make_uint4(0, 0, 0, 0);
  uint32_t tc4 = 0;  // FCons.w0
  auto tables_load = [&](const FRound& Q) {
    if (tid < Q.x_end - Q.x_begin) {
      const uint32_t* w = reinterpret_cast<const uint32_t*>(fxs + Q.x_begin + tid);
      tx0 = w[0];
      tx1 = w[1];
      tx2 = w[2];
    }
    if (tid < Q.c_end - Q.c_begin) {
      tc = *reinterpret_cast<const uint4*>(fcons + Q.c_begin + tid);
      tc4 = reinterpret_cast<const uint32_t*>(fcons + Q.c_begin + tid)[4];
    }
    if (tid < Q.f_end - Q.f_begin) tf = *reinterpret_cast<const uint4*>(frefs + Q.f_begin + tid);
    if (tid < Q.s_end - Q.s_begin) tg = *reinterpret_cast<const uint4*>(segs + Q.s_begin + tid);
  };
  auto tables_store = [&](const FRound& Q, int b) {
    if (tid < Q.x_end - Q.x_begin) {
      uint32_t* w = reinterpret_cast<uint32_t*>(&sx[b][tid]);
      w[0] = tx0;
      w[1] = tx1;
      w[2] = tx2;
    }
    if (tid < Q.c_end - Q.c_begin) {
      *reinterpret_cast<uint4*>(&sc[b][tid]) = tc;
      reinterpret_cast<uint32_t*>(&sc[b][tid])[4] = tc4;
    }
    if (tid < Q.f_end - Q.f_begin) {  // FRef: r (8 bytes), then x_lds
      sfr[b][tid] = __hiloint2double((int)tf.y, (int)tf.x);
      sfx[b][tid] = (int16_t)tf.z;
    }
    if (tid < Q.s_end - Q.s_begin) *reinterpret_cast<uint4*>(&sg[b][tid]) = tg;
  };
  // element e of a round's staging order lives in segment k: a binary search
  // for a thread's first element (k < 0), then upward steps (its elements are
  // NT apart, a few segments at most)
  auto seg_at = [&](int b, int ns, int e, int& k) {
    if (k < 0) { k = seg_of(e, ns, [&](int x) { return (int)sg[b][x].pre; }); return; }
    while (k + 1 < ns && sg[b][k + 1].pre <= e) ++k;
  };
  tables_load(srs[0]);
  tables_store(srs[0], 0);
  __syncthreads();
  {  // round 0's operands, loaded in batches of kFPre with every load of a batch in flight
    const FRound& Q = srs[0];
    const int ns = Q.s_end - Q.s_begin;
    int k = -1;
    for (int e0 = tid; e0 < Q.stage; e0 += NT * kFPre) {
      T v[kFPre];
      int kk = k;  // the stores below repeat the loads' segment walk
#pragma unroll
      for (int q = 0; q < kFPre; ++q) {
        const int e = e0 + NT * q;
        if (e < Q.stage) {
          seg_at(0, ns, e, kk);
          v[q] = fuse_src(sg[0][kk], V, Ad, Bh)[e - sg[0][kk].pre];
        }
      }
#pragma unroll
      for (int q = 0; q < kFPre; ++q) {
        const int e = e0 + NT * q;
        if (e < Q.stage) {
          seg_at(0, ns, e, k);
          sm[sg[0][k].lds + (e - sg[0][k].pre)] = v[q];
        }
      }
    }
  }
  __syncthreads();
  for (int r = 0; r < nr; ++r) {
    const FRound& R = srs[r];
    const int b = r & 1, nb = b ^ 1;
    const bool more = r + 1 < nr;
    const int nx = R.x_end - R.x_begin, nc = R.c_end - R.c_begin;
    if (more) tables_load(srs[r + 1]);
    // X = V_R Bhat^T: unit (row i, columns j0 .. j0+3) of product x
    {
      int x = tid < R.xwork ? seg_of(tid, nx, [&](int i) { return (int)sx[b][i].work0; }) : 0;
      for (int w = tid; w < R.xwork; w += NT) {
        while (x + 1 < nx && sx[b][x + 1].work0 <= w) ++x;
        const FX X = sx[b][x];
        const int dar = X.dar, dbr = X.dbr;
        const int lw = w - X.work0, i = lw % dar, j0 = (lw / dar) * kColBlk;
        const T* sV = sm + X.v_lds + i;
        const T* bc[kColBlk];
#pragma unroll
        for (int q = 0; q < kColBlk; ++q) bc[q] = sm + X.b_lds + dbr * min(j0 + q, db - 1);
        Acc s0[kColBlk], s1[kColBlk];
#pragma unroll
        for (int q = 0; q < kColBlk; ++q) { s0[q] = Acc(0); s1[q] = Acc(0); }
        int k = 0;
        for (; k + 1 < dbr; k += 2) {
          const T v0 = sV[dar * k], v1 = sV[dar * (k + 1)];
#pragma unroll
          for (int q = 0; q < kColBlk; ++q) {
            s0[q] = Prec<T>::fma3(v0, bc[q][k], s0[q]);
            s1[q] = Prec<T>::fma3(v1, bc[q][k + 1], s1[q]);
          }
        }
        if (k < dbr) {
          const T v0 = sV[dar * k];
#pragma unroll
          for (int q = 0; q < kColBlk; ++q) s0[q] = Prec<T>::fma3(v0, bc[q][k], s0[q]);
        }
        Acc* xo = smx + X.x_lds + i;
#pragma unroll
        for (int q = 0; q < kColBlk; ++q)
          if (j0 + q < db) xo[dar * (j0 + q)] = s0[q] + s1[q];
      }
    }
    if (more) tables_store(srs[r + 1], nb);
    __syncthreads();
    // round r+1's operands into registers while this round forms Y and W
    T pre[kFPre];
    const int nstage = more ? srs[r + 1].stage : 0;
    const int ns1 = more ? srs[r + 1].s_end - srs[r + 1].s_begin : 0;
    {
      int k = -1;
#pragma unroll
      for (int q = 0; q < kFPre; ++q) {
        const int e = tid + NT * q;
        if (e < nstage) {
          seg_at(nb, ns1, e, k);
          pre[q] = fuse_src(sg[nb][k], V, Ad, Bh)[e - sg[nb][k].pre];
        }
      }
    }
    // Y_c = sum_f r_f X_f, element by element (the X of one consumer share a layout)
    {
      int c = tid < R.ywork ? seg_of(tid, nc, [&](int i) { return sc[b][i].y0; }) : 0;
      for (int e = tid; e < R.ywork; e += NT) {
        while (c + 1 < nc && sc[b][c + 1].y0 <= e) ++c;
        const FCons C = sc[b][c];
        const int le = e - C.y0;
        Acc y = Acc(0);
        for (int f = C.f_begin; f < C.f_end; ++f) y += Acc(sfr[b][f]) * smx[sfx[b][f] + le];
        sm[C.y_lds + le] = Prec<T>::from(y);
      }
    }
    __syncthreads();
    // W items (consumer c, row, 4 columns) of A_c Y_c, dealt over all threads:
    // each writes its four sums over the X space (dead once Y is formed)
    Acc* wp = smx + R.wp_lds;
    {
      int c = tid < R.wwork ? seg_of(tid, nc, [&](int i) { return sc[b][i].w0; }) : 0;
      for (int it = tid; it < R.wwork; it += NT) {
        while (c + 1 < nc && sc[b][c + 1].w0 <= it) ++c;
        const int da = sc[b][c].da, dar = sc[b][c].dar, lu = it - sc[b][c].w0;
        const int i = lu % da, j0 = (lu / da) * kColBlk;
        const T* sA = sm + sc[b][c].a_lds + i;
        const T* yc[kColBlk];
#pragma unroll
        for (int j = 0; j < kColBlk; ++j) yc[j] = sm + sc[b][c].y_lds + dar * min(j0 + j, db - 1);
        Acc s0[kColBlk], s1[kColBlk];
#pragma unroll
        for (int j = 0; j < kColBlk; ++j) { s0[j] = Acc(0); s1[j] = Acc(0); }
        int k = 0;
        for (; k + 1 < dar; k += 2) {
          const T a0 = sA[da * k], a1 = sA[da * (k + 1)];
          T y0[kColBlk], y1[kColBlk];
#pragma unroll
          for (int j = 0; j < kColBlk; ++j) { y0[j] = yc[j][k]; y1[j] = yc[j][k + 1]; }
#pragma unroll
          for (int j = 0; j < kColBlk; ++j) {
            s0[j] = Prec<T>::fma3(a0, y0[j], s0[j]);
            s1[j] = Prec<T>::fma3(a1, y1[j], s1[j]);
          }
        }
        if (k < dar) {
          const T a0 = sA[da * k];
#pragma unroll
          for (int j = 0; j < kColBlk; ++j) s0[j] = Prec<T>::fma3(a0, yc[j][k], s0[j]);
        }
#pragma unroll
        for (int j = 0; j < kColBlk; ++j) wp[kColBlk * it + j] = s0[j] + s1[j];
      }
    }
    __syncthreads();
    // every owner adds its units' item sums, consumer by consumer in plan order
#pragma unroll
    for (int q = 0; q < U; ++q) {
      if (ub[q] < 0) continue;
      const int c0 = R.cb[ub[q]], c1 = R.cb[ub[q] + 1];
      for (int c = c0; c < c1; ++c) {
        const Acc* w = wp + kColBlk * (sc[b][c].w0 + ul[q]);
#pragma unroll
        for (int j = 0; j < kColBlk; ++j) acc[q][j] += w[j];
      }
    }
    if (!more) break;
    __syncthreads();
    {  // round r+1's operands to LDS: the prefetched ones, then any beyond kFPre per thread
      int k = -1;
#pragma unroll
      for (int q = 0; q < kFPre; ++q) {
        const int e = tid + NT * q;
        if (e < nstage) {
          seg_at(nb, ns1, e, k);
          sm[sg[nb][k].lds + (e - sg[nb][k].pre)] = pre[q];
        }
      }
      for (int e = tid + NT * kFPre; e < nstage; e += NT) {
        seg_at(nb, ns1, e, k);
        sm[sg[nb][k].lds + (e - sg[nb][k].pre)] = fuse_src(sg[nb][k], V, Ad, Bh)[e - sg[nb][k].pre];
      }
    }
    __syncthreads();
  }
  // the column partial: every owned unit, touched or not
#pragma unroll
  for (int q = 0; q < U; ++q) {
    if (ub[q] < 0) continue;
    const FBlk B = sb[ub[q]];
    const int i = ul[q] % B.da, j0 = (ul[q] / B.da) * kColBlk;
    double* po = part + tk.part_off + B.wc_off + i;
#pragma unroll
    for (int j = 0; j < kColBlk; ++j)
      if (j0 + j < db) po[(int64_t)B.da * (j0 + j)] = (double)acc[q][j];
  }
}

// ---------------------------------------------------------------- large-block GEMM path
// 64 x 64 output tiles on the matrix cores: FP64 (v_mfma_f64_16x16x4f64) for
// FP64 plans, FP32 (v_mfma_f32_16x16x4f32) for FP32 plans.  (The BigIn<Bf16x2>
// expansion of a hi/lo pair to its FP32 value hi + lo, exact in FP32, keeps
// these templates complete for BF16 operands; BF16 plans run the k_big*_bf
// kernels below on the BF16 matrix cores instead.)  In both, lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16);
// it holds D(l / 16 + 4 v, l % 16) in FP64 and D(4 (l / 16) + v, l % 16) in FP32.
// Four wavefronts take 32 x 32 quarters as 2 x 2 MFMA blocks.  The product is
// formed transposed, D = C^T, so that a lane's outputs run down a column of the
// column-major result and stores (or atomic adds) coalesce.  K advances 16 at
// a time: both operand tiles go through LDS (k-major rows of 64, padded) and the
// next tiles are loaded into registers while the current ones are multiplied.
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kBigK = kBigStep;       // K step (rows of each LDS tile)
constexpr int kBigQ = kBigK / 4;      // staged elements per thread and operand (256 threads, 64 x kBigK)
constexpr int kBigJS = 256 / kBigK;   // column stride of the k-major staging map
constexpr int kBigLd = kBigTile + 1;  // padded LDS row

template <typename C>
struct BigMma;
template <>
struct BigMma<double> {
  using vec = f64x4;
  __device__ static vec mma(double a, double b, vec c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int v) { return lane / 16 + 4 * v; }
};
template <>
struct BigMma<float> {
  using vec = f32x4;
  __device__ static vec mma(float a, float b, vec c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int v) { return 4 * (lane / 16) + v; }
};
// storage type -> compute type of the tiles
template <typename T>
struct BigIn;
template <>
struct BigIn<double> {
  using C = double;
  __device__ static double get(double x) { return x; }
};
template <>
struct BigIn<float> {
  using C = float;
  __device__ static float get(float x) { return x; }
};
template <>
struct BigIn<Bf16x2> {
  using C = float;
  __device__ static float get(Bf16x2 x) { return bf_hi(x) + bf_lo(x); }
};

// lowest set bit of the 128-bit step mask (m0, m1), cleared; -1 when empty
// (selects on values, so that the two words stay in registers)
__device__ __forceinline__ int pop_step(uint64_t& m0, uint64_t& m1) {
  const bool lo = m0 != 0;
  const uint64_t m = lo ? m0 : m1;
  if (!m) return -1;
  const int b = __builtin_ctzll(m);
  const uint64_t rest = m & (m - 1);
  m0 = lo ? rest : m0;
  m1 = lo ? m1 : rest;
  return lo ? b : 64 + b;
}

// one K step of the tile: sP holds the row operand of D (the columns j of C) as
// [k][j], sQ the column operand (the rows i of C) as [k][i]
template <typename C>
__device__ __forceinline__ void big_mma(const C* sP, const C* sQ, int wj, int wi, int lane,
                                        typename BigMma<C>::vec (&acc)[2][2]) {
#pragma unroll
  for (int ks = 0; ks < kBigK; ks += 4) {
    C a[2], b[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) a[r] = sP[(ks + lane / 16) * kBigLd + wj + r * 16 + lane % 16];
#pragma unroll
    for (int c = 0; c < 2; ++c) b[c] = sQ[(ks + lane / 16) * kBigLd + wi + c * 16 + lane % 16];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[r][c] = BigMma<C>::mma(a[r], b[c], acc[r][c]);
  }
}

template <typename C, int CB = 2>
__device__ __forceinline__ void big_zero(typename BigMma<C>::vec (&acc)[2][CB]) {
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[r][c] = typename BigMma<C>::vec{C(0), C(0), C(0), C(0)};
}

// G1: X = V_R Bhat^T for one 64 x 64 tile of one pair (V_R m x k, Bhat^T k x n)
template <typename T>
__global__ __launch_bounds__(256) void k_big1(const GTile1* __restrict__ tiles,
                                              const GTask1* __restrict__ tasks, const T* __restrict__ V,
                                              const T* __restrict__ Bt, typename BigIn<T>::C* __restrict__ X,
                                              EnsStride es) {
  using C = typename BigIn<T>::C;
  __shared__ C sP[2][kBigK * kBigLd], sQ[2][kBigK * kBigLd];  // double-buffered: one barrier per K step
  const GTile1 t = tiles[blockIdx.x];
  const GTask1 g = tasks[t.task];
  {  // linear ensembles: blockIdx.y is the sample (strides zero otherwise)
    const int64_t smp = blockIdx.y;
    V += smp * es.v;
    Bt += smp * es.b;
    X += smp * es.xg;
  }
  const int m = g.m, n = g.n, K = g.k, i0 = t.i0, j0 = t.j0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wj = (wave & 1) * 32, wi = (wave >> 1) * 32;
  typename BigMma<C>::vec acc[2][2];
  big_zero<C>(acc);
  // staging: V tile element (i = tid % 64, kk = tid / 64 + 4q), Bhat^T tile
  // element (kk = tid % 16, j = tid / 16 + 16q), zero outside the matrices
  C rv[kBigQ], rb[kBigQ];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) {
      const int i = i0 + tid % 64, kk = k0 + tid / 64 + 4 * q;
      rv[q] = (i < m && kk < K) ? BigIn<T>::get(V[g.v_off + i + (int64_t)m * kk]) : C(0);
    }
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) {
      const int kk = k0 + tid % kBigK, j = j0 + tid / kBigK + kBigJS * q;
      rb[q] = (kk < K && j < n) ? BigIn<T>::get(Bt[g.b_off + kk + (int64_t)K * j]) : C(0);
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) sQ[buf][(tid / 64 + 4 * q) * kBigLd + tid % 64] = rv[q];
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) sP[buf][(tid % kBigK) * kBigLd + tid / kBigK + kBigJS * q] = rb[q];
  };
  // only the K steps where this tile's Bhat^T columns are not all zero.  Step
  // s is multiplied from buffer cur while step s+1 is loaded and then staged
  // into the other buffer, whose last readers finished before the previous barrier
  uint64_t m0 = t.kmask[0], m1 = t.kmask[1];
  int step = pop_step(m0, m1), cur = 0;
  if (step >= 0) {
    load(step * kBigK);
    stage(0);
    __syncthreads();
  }
  while (step >= 0) {
    step = pop_step(m0, m1);
    if (step >= 0) load(step * kBigK);
    big_mma<C>(sP[cur], sQ[cur], wj, wi, lane, acc);
    if (step >= 0) stage(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + c * 16 + lane % 16, j = j0 + wj + r * 16 + BigMma<C>::row(lane, v);
        if (i < m && j < n) X[g.x_off + i + (int64_t)big_ldx(m) * j] = acc[r][c][v];
      }
}

// G2: one 64 x 64 tile of W_L += sum_k A_k Y_k over a chunk of a-terms, with
// Y_k = sum_u r_u X_u formed while staging; the tile is added into W atomically
template <typename T>
__global__ __launch_bounds__(256) void k_big2(const GTile2* __restrict__ tiles,
                                              const GTerm2* __restrict__ terms, const SXRef* __restrict__ xrefs,
                                              const T* __restrict__ Ad, const typename BigIn<T>::C* __restrict__ X,
                                              double* __restrict__ W, EnsStride es) {
  using C = typename BigIn<T>::C;
  __shared__ C sP[2][kBigK * kBigLd], sQ[2][kBigK * kBigLd];  // double-buffered: one barrier per K step
  const GTile2 t = tiles[blockIdx.x];
  const int da = t.da, db = t.db, i0 = t.i0, j0 = t.j0;
  {  // linear ensembles: blockIdx.y is the sample (strides zero otherwise)
    const int64_t smp = blockIdx.y;
    X += smp * es.xg;
    W += smp * es.v;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wj = (wave & 1) * 32, wi = (wave >> 1) * 32;
  typename BigMma<C>::vec acc[2][2];
  big_zero<C>(acc);
  C ra[kBigQ], ry[kBigQ];
  // staging: A tile element (i = tid % 64, kk = tid / 64 + 4q), Y tile element
  // (kk = tid % kBigK, j = tid / kBigK + kBigJS q)
  auto load = [&](const GTerm2& tm, int k0) {
    const int dar = tm.dar, ldx = big_ldx(dar);
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) {
      const int i = i0 + tid % 64, kk = k0 + tid / 64 + 4 * q;
      ra[q] = (i < da && kk < dar) ? BigIn<T>::get(Ad[tm.a_off + i + (int64_t)da * kk]) : C(0);
    }
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) ry[q] = C(0);
    for (int x = tm.x_begin; x < tm.x_end; ++x) {
      const SXRef xr = xrefs[x];
      const C r = C(xr.r);
#pragma unroll
      for (int q = 0; q < kBigQ; ++q) {
        const int kk = k0 + tid % kBigK, j = j0 + tid / kBigK + kBigJS * q;
        if (kk < dar && j < db) ry[q] = fma(r, X[xr.x_off + kk + (int64_t)ldx * j], ry[q]);
      }
    }
  };
  // (term, K step) pairs in order, only the steps where the term's A rows of
  // this tile are not all zero
  int tc = t.t_begin;
  GTerm2 tm = terms[tc];
  uint64_t m0 = tm.kmask[0], m1 = tm.kmask[1];
  int step = pop_step(m0, m1);
  while (step < 0 && ++tc < t.t_end) {
    tm = terms[tc];
    m0 = tm.kmask[0];
    m1 = tm.kmask[1];
    step = pop_step(m0, m1);
  }
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) sQ[buf][(tid / 64 + 4 * q) * kBigLd + tid % 64] = ra[q];
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) sP[buf][(tid % kBigK) * kBigLd + tid / kBigK + kBigJS * q] = ry[q];
  };
  int cur = 0;
  if (step >= 0) {
    load(tm, step * kBigK);
    stage(0);
    __syncthreads();
  }
  while (step >= 0) {
    // next (term, step) pair, loaded while this step's products run and then
    // staged into the other buffer
    step = pop_step(m0, m1);
    while (step < 0 && ++tc < t.t_end) {
      tm = terms[tc];
      m0 = tm.kmask[0];
      m1 = tm.kmask[1];
      step = pop_step(m0, m1);
    }
    if (step >= 0) load(tm, step * kBigK);
    big_mma<C>(sP[cur], sQ[cur], wj, wi, lane, acc);
    if (step >= 0) stage(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + c * 16 + lane % 16, j = j0 + wj + r * 16 + BigMma<C>::row(lane, v);
        if (i < da && j < db) unsafeAtomicAdd(W + t.w_off + i + (int64_t)da * j, (double)acc[r][c][v]);
      }
}

// G0: W_L = H_a V_L + V_L H_b^T for one 64 x 64 tile of a large block, stored
// (it runs before the pn kernels add into W).  Two K phases: H_a (da x da) times
// V_L (da x db), then V_L times H_b^T, where H_b is symmetric: H_b^T(k, j) = H_b(k, j).
template <typename T>
__global__ __launch_bounds__(256) void k_big0(const GTile3* __restrict__ tiles, const T* __restrict__ V,
                                              const T* __restrict__ H, double* __restrict__ W, EnsStride es) {
  using C = typename BigIn<T>::C;
  __shared__ C sP[kBigK * kBigLd], sQ[kBigK * kBigLd];
  const GTile3 t = tiles[blockIdx.x];
  const int da = t.da, db = t.db, i0 = t.i0, j0 = t.j0;
  {  // linear ensembles: blockIdx.y is the sample (strides zero otherwise)
    const int64_t smp = blockIdx.y;
    V += smp * es.v;
    H += smp * es.h;
    W += smp * es.v;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wj = (wave & 1) * 32, wi = (wave >> 1) * 32;
  typename BigMma<C>::vec acc[2][2];
  big_zero<C>(acc);
  const T* v = V + t.v_off;
  // phase 0: A = H_a (ld da), Y = V_L (ld da), K = da; phase 1: A = V_L (ld da), Y = H_b (ld db), K = db
  for (int ph = 0; ph < 2; ++ph) {
    const int64_t hoff = ph == 0 ? t.ha_off : t.hb_off;
    if (hoff < 0) continue;
    const T* Am = ph == 0 ? H + hoff : v;
    const T* Ym = ph == 0 ? v : H + hoff;
    const int K = ph == 0 ? da : db, ldy = ph == 0 ? da : db;
    for (int k0 = 0; k0 < K; k0 += kBigK) {
      C ra[kBigQ], ry[kBigQ];
#pragma unroll
      for (int q = 0; q < kBigQ; ++q) {
        const int i = i0 + tid % 64, kk = k0 + tid / 64 + 4 * q;
        ra[q] = (i < da && kk < K) ? BigIn<T>::get(Am[i + (int64_t)da * kk]) : C(0);
      }
#pragma unroll
      for (int q = 0; q < kBigQ; ++q) {
        const int kk = k0 + tid % kBigK, j = j0 + tid / kBigK + kBigJS * q;
        ry[q] = (kk < K && j < db) ? BigIn<T>::get(Ym[kk + (int64_t)ldy * j]) : C(0);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kBigQ; ++q) sQ[(tid / 64 + 4 * q) * kBigLd + tid % 64] = ra[q];
#pragma unroll
      for (int q = 0; q < kBigQ; ++q) sP[(tid % kBigK) * kBigLd + tid / kBigK + kBigJS * q] = ry[q];
      __syncthreads();
      big_mma<C>(sP, sQ, wj, wi, lane, acc);
    }
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + wi + c * 16 + lane % 16, j = j0 + wj + r * 16 + BigMma<C>::row(lane, q);
        if (i < da && j < db) W[t.v_off + i + (int64_t)da * j] = (double)acc[r][c][q];
      }
}

// ---------------------------------------------------------------- large blocks, BF16 hi/lo
// The same tiles for BF16 hi/lo operands on the BF16 matrix cores
// (v_mfma_f32_16x16x16bf16_1k, FP32 accumulation): per 16-deep K step and
// 16 x 16 block, three products ah*bh + ah*bl + al*bh (the lo*lo term is below
// the pair's resolution), one instruction each.  Lane l supplies A(l % 16,
// 4 (l / 16) + t) and B(4 (l / 16) + t, l % 16), t = 0..3, as four packed
// BF16, and holds D(4 (l / 16) + v, l % 16).  LDS keeps the hi and lo planes
// of both operands as [row][k], 16 K values per row padded to 20 (40 bytes),
// so that a lane's four K values are one conflict-free 8-byte read.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
constexpr int kBfLd = kBigK + 4;
constexpr int kBfPlane = kBigTile * kBfLd;  // one operand plane of one buffer (uint16)

template <int TM = kBigTile>  // TM: rows of C (the Q operand) per tile
struct alignas(16) BfTile {  // one buffer: hi / lo planes of the row operand (P, [j][k]) and column operand (Q, [i][k])
  uint16_t ph[kBfPlane], pl[kBfPlane], qh[TM * kBfLd], ql[TM * kBfLd];
};

__device__ __forceinline__ void bf_put(uint16_t* hi, uint16_t* lo, int idx, Bf16x2 x) {
  hi[idx] = (uint16_t)(x.bits >> 16);
  lo[idx] = (uint16_t)(x.bits & 0xffffu);
}

// a wavefront's 32 x (16 CB) block of D
template <int TM, int CB = 2>
__device__ __forceinline__ void big_mma_bf(const BfTile<TM>& b, int wj, int wi, int lane, f32x4 (&acc)[2][CB]) {
  const int kq = 4 * (lane / 16);
  bf16x4 ah[2], al[2], bh[CB], bl[CB];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int o = (wj + r * 16 + lane % 16) * kBfLd + kq;
    ah[r] = *reinterpret_cast<const bf16x4*>(b.ph + o);
    al[r] = *reinterpret_cast<const bf16x4*>(b.pl + o);
  }
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    const int o = (wi + c * 16 + lane % 16) * kBfLd + kq;
    bh[c] = *reinterpret_cast<const bf16x4*>(b.qh + o);
    bl[c] = *reinterpret_cast<const bf16x4*>(b.ql + o);
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[r], bh[c], acc[r][c], 0, 0, 0);
      acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[r], bl[c], acc[r][c], 0, 0, 0);
      acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[r], bh[c], acc[r][c], 0, 0, 0);
    }
}

// G1 in BF16 hi/lo: X = V_R Bhat^T for one TM x 64 tile (TM = 64 or 128), X in FP32
template <int TM>
__global__ __launch_bounds__(256) void k_big1_bf(const GTile1* __restrict__ tiles, const GTask1* __restrict__ tasks,
                                                 const Bf16x2* __restrict__ V, const Bf16x2* __restrict__ Bt,
                                                 float* __restrict__ X, EnsStride es) {
  constexpr int CB = TM / 32, QV = TM * kBigK / 256, VS = 256 / TM;
  __shared__ BfTile<TM> sb[2];  // double-buffered: one barrier per K step
  const GTile1 t = tiles[blockIdx.x];
  const GTask1 g = tasks[t.task];
  {  // linear ensembles: blockIdx.y is the sample (strides zero otherwise)
    const int64_t smp = blockIdx.y;
    V += smp * es.v;
    Bt += smp * es.b;
    X += smp * es.xg;
  }
  const int m = g.m, n = g.n, K = g.k, i0 = t.i0, j0 = t.j0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wj = (wave & 1) * 32, wi = (wave >> 1) * (TM / 2);
  f32x4 acc[2][CB];
  big_zero<float, CB>(acc);
  // staging: V tile element (i = tid % TM, kk = tid / TM + VS q) into Q[i][kk],
  // Bhat^T tile element (kk = tid % 16, j = tid / 16 + 16q) into P[j][kk]
  Bf16x2 rv[QV], rb[kBigQ];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < QV; ++q) {
      const int i = i0 + tid % TM, kk = k0 + tid / TM + VS * q;
      rv[q] = (i < m && kk < K) ? V[g.v_off + i + (int64_t)m * kk] : Bf16x2{0u};
    }
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) {
      const int kk = k0 + tid % kBigK, j = j0 + tid / kBigK + kBigJS * q;
      rb[q] = (kk < K && j < n) ? Bt[g.b_off + kk + (int64_t)K * j] : Bf16x2{0u};
    }
  };
  auto stage = [&](BfTile<TM>& b) {
#pragma unroll
    for (int q = 0; q < QV; ++q) bf_put(b.qh, b.ql, (tid % TM) * kBfLd + tid / TM + VS * q, rv[q]);
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) bf_put(b.ph, b.pl, (tid / kBigK + kBigJS * q) * kBfLd + tid % kBigK, rb[q]);
  };
  uint64_t m0 = t.kmask[0], m1 = t.kmask[1];
  int step = pop_step(m0, m1), cur = 0;
  if (step >= 0) {
    load(step * kBigK);
    stage(sb[0]);
    __syncthreads();
  }
  while (step >= 0) {
    step = pop_step(m0, m1);
    if (step >= 0) load(step * kBigK);
    big_mma_bf<TM, CB>(sb[cur], wj, wi, lane, acc);
    if (step >= 0) stage(sb[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + c * 16 + lane % 16, j = j0 + wj + r * 16 + 4 * (lane / 16) + v;
        if (i < m && j < n) X[g.x_off + i + (int64_t)big_ldx(m) * j] = acc[r][c][v];
      }
}

// G2 in BF16 hi/lo: W_L += sum_k A_k Y_k, Y_k = sum_u r_u X_u formed in FP32
// and split into a hi/lo pair while staging
__global__ __launch_bounds__(256) void k_big2_bf(const GTile2* __restrict__ tiles, const GTerm2* __restrict__ terms,
                                                 const SXRef* __restrict__ xrefs, const Bf16x2* __restrict__ Ad,
                                                 const float* __restrict__ X, double* __restrict__ W, EnsStride es) {
  __shared__ BfTile<> sb[2];
  const GTile2 t = tiles[blockIdx.x];
  const int da = t.da, db = t.db, i0 = t.i0, j0 = t.j0;
  {  // linear ensembles: blockIdx.y is the sample (strides zero otherwise)
    const int64_t smp = blockIdx.y;
    X += smp * es.xg;
    W += smp * es.v;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wj = (wave & 1) * 32, wi = (wave >> 1) * 32;
  f32x4 acc[2][2];
  big_zero<float>(acc);
  Bf16x2 ra[kBigQ];
  float ry[kBigQ];
  auto load = [&](const GTerm2& tm, int k0) {
    const int dar = tm.dar, ldx = big_ldx(dar);
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) {
      const int i = i0 + tid % 64, kk = k0 + tid / 64 + 4 * q;
      ra[q] = (i < da && kk < dar) ? Ad[tm.a_off + i + (int64_t)da * kk] : Bf16x2{0u};
    }
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) ry[q] = 0.0f;
    for (int x = tm.x_begin; x < tm.x_end; ++x) {
      const SXRef xr = xrefs[x];
      const float r = (float)xr.r;
#pragma unroll
      for (int q = 0; q < kBigQ; ++q) {
        const int kk = k0 + tid % kBigK, j = j0 + tid / kBigK + kBigJS * q;
        if (kk < dar && j < db) ry[q] = fmaf(r, X[xr.x_off + kk + (int64_t)ldx * j], ry[q]);
      }
    }
  };
  auto stage = [&](BfTile<>& b) {
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) bf_put(b.qh, b.ql, (tid % 64) * kBfLd + tid / 64 + 4 * q, ra[q]);
#pragma unroll
    for (int q = 0; q < kBigQ; ++q)
      bf_put(b.ph, b.pl, (tid / kBigK + kBigJS * q) * kBfLd + tid % kBigK, bf16x2_split(ry[q]));
  };
  int tc = t.t_begin;
  GTerm2 tm = terms[tc];
  uint64_t m0 = tm.kmask[0], m1 = tm.kmask[1];
  int step = pop_step(m0, m1);
  while (step < 0 && ++tc < t.t_end) {
    tm = terms[tc];
    m0 = tm.kmask[0];
    m1 = tm.kmask[1];
    step = pop_step(m0, m1);
  }
  int cur = 0;
  if (step >= 0) {
    load(tm, step * kBigK);
    stage(sb[0]);
    __syncthreads();
  }
  while (step >= 0) {
    step = pop_step(m0, m1);
    while (step < 0 && ++tc < t.t_end) {
      tm = terms[tc];
      m0 = tm.kmask[0];
      m1 = tm.kmask[1];
      step = pop_step(m0, m1);
    }
    if (step >= 0) load(tm, step * kBigK);
    big_mma_bf<kBigTile>(sb[cur], wj, wi, lane, acc);
    if (step >= 0) stage(sb[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + c * 16 + lane % 16, j = j0 + wj + r * 16 + 4 * (lane / 16) + v;
        if (i < da && j < db) unsafeAtomicAdd(W + t.w_off + i + (int64_t)da * j, (double)acc[r][c][v]);
      }
}

// Operand loads through a buffer resource on a wavefront-uniform base: one
// 32-bit byte offset per lane, and a lane outside its block passes an offset
// beyond the resource's range (kOobOff), which returns zero without touching
// memory.  Word 3 is the gfx9 raw-buffer format (32-bit data).
constexpr uint32_t kOobOff = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mk_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7fffffff, 0x00020000);
}
template <typename X>
struct BufLd;
// soff: a wavefront-uniform byte offset added in the scalar offset field (a
// lane passing kOobOff stays out of range whatever soff is)
template <>
struct BufLd<double> {
  __device__ static double ld(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff = 0) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
  }
};
template <>
struct BufLd<float> {
  __device__ static float ld(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff = 0) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
  }
};
template <>
struct BufLd<Bf16x2> {
  __device__ static Bf16x2 ld(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff = 0) {
    return Bf16x2{__builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0)};
  }
};
// byte offset of element idx, or kOobOff when the lane is outside its block
template <typename X>
__device__ __forceinline__ uint32_t boff(bool ok, int idx) {
  return ok ? (uint32_t)idx * (uint32_t)sizeof(X) : kOobOff;
}

// ---------------------------------------------------------------- large blocks, BF16 panels
// The same tiles with every operand read as 16-byte panel units (PanelBlk):
// a unit is four K values of one row, hi halves then lo halves, exactly what
// one lane of the BF16 matrix-core instruction takes for one 16-deep K step
// (hi, then lo, four packed BF16 each).  Staging a K step of a 64-row operand
// is one 16-byte load and one 16-byte LDS store per thread, in place of four
// 4-byte loads and eight 2-byte stores; the tiles are bound by that staging,
// not by the matrix cores.  Operands are read through buffer resources (a
// lane outside its block passes kOobOff and reads zero).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
// the hi (first two dwords) and lo (last two) halves of a unit as MFMA operands
// (built from the named components: a bit cast of a swizzled subvector such as
// u.zw reads the wrong dwords with this compiler)
__device__ __forceinline__ bf16x4 unit_hi(u32x4 u) { return __builtin_bit_cast(bf16x4, u32x2{u.x, u.y}); }
__device__ __forceinline__ bf16x4 unit_lo(u32x4 u) { return __builtin_bit_cast(bf16x4, u32x2{u.z, u.w}); }

// One panel unit per thread, from column-major Bf16x2 blocks.
__global__ __launch_bounds__(256) void k_panel_bf(const PanelChunk* __restrict__ chunks,
                                                  const PanelBlk* __restrict__ blks,
                                                  const Bf16x2* __restrict__ src, uint4* __restrict__ dst) {
  const PanelChunk c = chunks[blockIdx.x];
  const PanelBlk b = blks[c.blk];
  const int nx = b.kcols ? b.rows : b.cols, K = b.kcols ? b.cols : b.rows;
  const int u = c.u0 + (int)threadIdx.x;  // = 4 (s nx + x) + qq: neighbouring threads write neighbouring units
  if (u >= nx * b.ldq) return;
  const int x = (u % (4 * nx)) / 4, q = 4 * (u / (4 * nx)) + u % 4;
  uint32_t h[4], l[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int kk = 4 * q + t;
    const int64_t e = b.kcols ? x + (int64_t)b.rows * kk : kk + (int64_t)b.rows * x;
    const uint32_t v = kk < K ? src[b.src + e].bits : 0u;
    h[t] = v >> 16;
    l[t] = v & 0xffffu;
  }
  dst[b.dst + u] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), l[0] | (l[1] << 16), l[2] | (l[3] << 16));
}

// byte offset of unit (row x, quad qq) of K step s in a panel of nx rows, or
// kOobOff when the row is outside the block (nx <= 2048 rows, K <= 2048: a
// panel is at most 16 MB)
__device__ __forceinline__ uint32_t panel_off(bool ok, int nx, int s, int x, int qq) {
  return ok ? 16u * (uint32_t)(4 * (s * nx + x) + qq) : kOobOff;
}
__device__ __forceinline__ u32x4 ld_unit(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

// LDS tiles of panel units, [row][quad] with the quad swizzled by the row:
// unit (row, g) at 4 row + (g ^ (((row >> 1) ^ (row >> 2)) & 3)).  A lane of
// an MFMA reads unit (16-row block + lane % 16, lane / 16) with one 16-byte
// read; the swizzle spreads the 16 rows of one quad over distinct banks (for
// both 8- and 16-lane read groups), and a wavefront's staging store (16 rows
// x 4 quads) is 1 KB of contiguous LDS.
template <int TM = kBigTile>  // TM: rows of the column operand (Q)
struct alignas(16) BfUTile {
  u32x4 p[kBigTile * 4], q[TM * 4];
};
__device__ __forceinline__ int unit_slot(int row, int g) { return 4 * row + (g ^ (((row >> 1) ^ (row >> 2)) & 3)); }

// a wavefront's 32 x (16 CB) block of D from a unit tile (as big_mma_bf)
template <int TM, int CB = 2>
__device__ __forceinline__ void big_mma_bfu(const BfUTile<TM>& b, int wj, int wi, int lane, f32x4 (&acc)[2][CB]) {
  const int g = lane / 16, r16 = lane % 16;
  bf16x4 ah[2], al[2], bh[CB], bl[CB];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const u32x4 u = b.p[unit_slot(wj + r * 16 + r16, g)];
    ah[r] = unit_hi(u);
    al[r] = unit_lo(u);
  }
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    const u32x4 u = b.q[unit_slot(wi + c * 16 + r16, g)];
    bh[c] = unit_hi(u);
    bl[c] = unit_lo(u);
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[r], bh[c], acc[r][c], 0, 0, 0);
      acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[r], bl[c], acc[r][c], 0, 0, 0);
      acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[r], bh[c], acc[r][c], 0, 0, 0);
    }
}

// G1 on panels: X = V_R Bhat^T for one TM x 64 tile.  Staging map: unit u of
// a K step is (row u / 4, quad u % 4); the V_R tile has TM rows (TM / 64
// units per thread), the Bhat^T tile 64 (one unit per thread).
template <int TM>
__global__ __launch_bounds__(256) void k_big1_bfp(const GTile1* __restrict__ tiles, const GTask1* __restrict__ tasks,
                                                  const uint4* __restrict__ Vt, const uint4* __restrict__ Bp,
                                                  float* __restrict__ X, int n_tiles, int grp) {
  constexpr int CB = TM / 32, QV = TM / 64;
  __shared__ BfUTile<TM> sb[2];  // double-buffered: one barrier per K step
  const int ti = grp > 0 ? xcd_group((int)blockIdx.x, grp) : (int)blockIdx.x;
  if (ti >= n_tiles) return;
  const GTile1 t = tiles[ti];
  const GTask1 g = tasks[t.task];
  const int m = g.m, n = g.n, i0 = t.i0, j0 = t.j0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wj = (wave & 1) * 32, wi = (wave >> 1) * (TM / 2);
  f32x4 acc[2][CB];
  big_zero<float, CB>(acc);
  const __amdgpu_buffer_rsrc_t rvs = mk_rsrc(Vt + g.v_off), rbs = mk_rsrc(Bp + g.b_off);
  u32x4 rv[QV], rb;
  auto load = [&](int s) {
#pragma unroll
    for (int q = 0; q < QV; ++q) {
      const int u = tid + 256 * q, i = i0 + u / 4;
      rv[q] = ld_unit(rvs, panel_off(i < m, m, s, i, u % 4));
    }
    const int j = j0 + tid / 4;
    rb = ld_unit(rbs, panel_off(j < n, n, s, j, tid % 4));
  };
  auto stage = [&](BfUTile<TM>& b) {
#pragma unroll
    for (int q = 0; q < QV; ++q) {
      const int u = tid + 256 * q;
      b.q[unit_slot(u / 4, u % 4)] = rv[q];
    }
    b.p[unit_slot(tid / 4, tid % 4)] = rb;
  };
  uint64_t m0 = t.kmask[0], m1 = t.kmask[1];
  int step = pop_step(m0, m1), cur = 0;
  if (step >= 0) {
    load(step);
    stage(sb[0]);
    __syncthreads();
  }
  while (step >= 0) {
    step = pop_step(m0, m1);
    if (step >= 0) load(step);
    big_mma_bfu<TM, CB>(sb[cur], wj, wi, lane, acc);
    if (step >= 0) stage(sb[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + c * 16 + lane % 16, j = j0 + wj + r * 16 + 4 * (lane / 16) + v;
        if (i < m && j < n) X[g.x_off + i + (int64_t)big_ldx(m) * j] = acc[r][c][v];
      }
}

// G2 on panels: W_L += sum_k A_k Y_k.  The A tile is one panel unit per
// thread; Y = sum_u r_u X_u is formed as four consecutive K values of one
// column (one 16-byte load per X reference: X columns are padded with zeros
// to whole K steps and 64-byte aligned) and split into hi/lo units.
__global__ __launch_bounds__(256) void k_big2_bfp(const GTile2* __restrict__ tiles, const GTerm2* __restrict__ terms,
                                                  const SXRef* __restrict__ xrefs, const uint4* __restrict__ Ap,
                                                  const float* __restrict__ X, double* __restrict__ W, int n_tiles,
                                                  int grp) {
  __shared__ BfUTile<> sb[2];
  const int ti = grp > 0 ? xcd_group((int)blockIdx.x, grp) : (int)blockIdx.x;
  if (ti >= n_tiles) return;
  const GTile2 t = tiles[ti];
  const int da = t.da, db = t.db, i0 = t.i0, j0 = t.j0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wj = (wave & 1) * 32, wi = (wave >> 1) * 32;
  const int ro = tid / 4, qo = tid % 4;  // staging map: row (A: i, Y: j) and K quad
  f32x4 acc[2][2];
  big_zero<float>(acc);
  u32x4 ra;
  float4 ry;
  auto load = [&](const GTerm2& tm, int s) {
    const int i = i0 + ro, j = j0 + ro;
    ra = ld_unit(mk_rsrc(Ap + tm.a_off), panel_off(i < da, da, s, i, qo));
    ry = make_float4(0.f, 0.f, 0.f, 0.f);
    if (j < db) {
      const int64_t col = kBigK * s + 4 * qo + (int64_t)big_ldx(tm.dar) * j;
      for (int x = tm.x_begin; x < tm.x_end; ++x) {
        const SXRef xr = xrefs[x];
        const float r = (float)xr.r;
        const float4 xv = *reinterpret_cast<const float4*>(X + xr.x_off + col);
        ry.x = fmaf(r, xv.x, ry.x);
        ry.y = fmaf(r, xv.y, ry.y);
        ry.z = fmaf(r, xv.z, ry.z);
        ry.w = fmaf(r, xv.w, ry.w);
      }
    }
  };
  auto stage = [&](BfUTile<>& b) {
    const int o = unit_slot(ro, qo);
    b.q[o] = ra;
    const uint32_t s0 = bf16x2_split(ry.x).bits, s1 = bf16x2_split(ry.y).bits;
    const uint32_t s2 = bf16x2_split(ry.z).bits, s3 = bf16x2_split(ry.w).bits;
    b.p[o] = u32x4{(s0 >> 16) | (s1 & 0xffff0000u), (s2 >> 16) | (s3 & 0xffff0000u), (s0 & 0xffffu) | (s1 << 16),
                   (s2 & 0xffffu) | (s3 << 16)};
  };
  int tc = t.t_begin;
  GTerm2 tm = terms[tc];
  uint64_t m0 = tm.kmask[0], m1 = tm.kmask[1];
  int step = pop_step(m0, m1);
  while (step < 0 && ++tc < t.t_end) {
    tm = terms[tc];
    m0 = tm.kmask[0];
    m1 = tm.kmask[1];
    step = pop_step(m0, m1);
  }
  int cur = 0;
  if (step >= 0) {
    load(tm, step);
    stage(sb[0]);
    __syncthreads();
  }
  while (step >= 0) {
    step = pop_step(m0, m1);
    while (step < 0 && ++tc < t.t_end) {
      tm = terms[tc];
      m0 = tm.kmask[0];
      m1 = tm.kmask[1];
      step = pop_step(m0, m1);
    }
    if (step >= 0) load(tm, step);
    big_mma_bfu<kBigTile>(sb[cur], wj, wi, lane, acc);
    if (step >= 0) stage(sb[cur ^ 1]);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + c * 16 + lane % 16, j = j0 + wj + r * 16 + 4 * (lane / 16) + v;
        if (i < da && j < db) unsafeAtomicAdd(W + t.w_off + i + (int64_t)da * j, (double)acc[r][c][v]);
      }
}

// G0 in BF16 hi/lo: W_L = H_a V_L + V_L H_b^T for one 64 x 64 tile, stored
__global__ __launch_bounds__(256) void k_big0_bf(const GTile3* __restrict__ tiles, const Bf16x2* __restrict__ V,
                                                 const Bf16x2* __restrict__ H, double* __restrict__ W, EnsStride es) {
  __shared__ BfTile<> sb[1];
  const GTile3 t = tiles[blockIdx.x];
  const int da = t.da, db = t.db, i0 = t.i0, j0 = t.j0;
  {  // linear ensembles: blockIdx.y is the sample (strides zero otherwise)
    const int64_t smp = blockIdx.y;
    V += smp * es.v;
    H += smp * es.h;
    W += smp * es.v;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wj = (wave & 1) * 32, wi = (wave >> 1) * 32;
  f32x4 acc[2][2];
  big_zero<float>(acc);
  const Bf16x2* v = V + t.v_off;
  for (int ph = 0; ph < 2; ++ph) {
    const int64_t hoff = ph == 0 ? t.ha_off : t.hb_off;
    if (hoff < 0) continue;
    const Bf16x2* Am = ph == 0 ? H + hoff : v;
    const Bf16x2* Ym = ph == 0 ? v : H + hoff;
    const int K = ph == 0 ? da : db, ldy = ph == 0 ? da : db;
    for (int k0 = 0; k0 < K; k0 += kBigK) {
      Bf16x2 ra[kBigQ], ry[kBigQ];
#pragma unroll
      for (int q = 0; q < kBigQ; ++q) {
        const int i = i0 + tid % 64, kk = k0 + tid / 64 + 4 * q;
        ra[q] = (i < da && kk < K) ? Am[i + (int64_t)da * kk] : Bf16x2{0u};
      }
#pragma unroll
      for (int q = 0; q < kBigQ; ++q) {
        const int kk = k0 + tid % kBigK, j = j0 + tid / kBigK + kBigJS * q;
        ry[q] = (kk < K && j < db) ? Ym[kk + (int64_t)ldy * j] : Bf16x2{0u};
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kBigQ; ++q) bf_put(sb[0].qh, sb[0].ql, (tid % 64) * kBfLd + tid / 64 + 4 * q, ra[q]);
#pragma unroll
      for (int q = 0; q < kBigQ; ++q)
        bf_put(sb[0].ph, sb[0].pl, (tid / kBigK + kBigJS * q) * kBfLd + tid % kBigK, ry[q]);
      __syncthreads();
      big_mma_bf<kBigTile>(sb[0], wj, wi, lane, acc);
    }
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + wi + c * 16 + lane % 16, j = j0 + wj + r * 16 + 4 * (lane / 16) + q;
        if (i < da && j < db) W[t.v_off + i + (int64_t)da * j] = (double)acc[r][c][q];
      }
}

const void* k_stk_ptr(int stage, int prec, int nt) {
#define NSX_STK(T)                                                                                  \
  if (stage == 1) return nt == 64 ? (const void*)k_stk1<T, 64> : nt == 128 ? (const void*)k_stk1<T, 128> \
                                                               : (const void*)k_stk1<T, 256>;         \
  return nt == 64 ? (const void*)k_stk2<T, 64> : nt == 128 ? (const void*)k_stk2<T, 128> : (const void*)k_stk2<T, 256>;
  switch (prec) {
    case 0: { NSX_STK(double) }
    case 1: { NSX_STK(float) }
    default: { NSX_STK(Bf16x2) }
  }
#undef NSX_STK
}

const void* k_fuse_ptr(int prec, int u) {
#define NSX_FUSE(T) \
  return u <= 2 ? (const void*)k_fuse<T, kFuseThreads, 2> : u <= 4 ? (const void*)k_fuse<T, kFuseThreads, 4> \
                                                               : (const void*)k_fuse<T, kFuseThreads, 8>;
  switch (prec) {
    case 0: { NSX_FUSE(double) }
    case 1: { NSX_FUSE(float) }
    default: { NSX_FUSE(Bf16x2) }
  }
#undef NSX_FUSE
}

// ---------------------------------------------------------------- matrix-core stacked path
// The stacked stages on 16 x 16 x 4 matrix-core steps, one tile per
// wavefront, with every operand loaded straight into the registers the
// instruction reads (lane l holds row l % 16, K l / 16 of the first operand and
// K l / 16, column l % 16 of the second): the small blocks need no LDS staging,
// no per-element index search and no barrier.  Task, unit and term records are
// wavefront-uniform (scalar loads).  The number of K steps NK of a product or
// term is a template parameter of its body (dispatched on the uniform inner
// size), so a body's loads are straight-line code: a lane outside its block
// reads the block's first element (always in range) and uses zero, and all of
// the loads of a step are in flight together.

// calls f(std::integral_constant<int, NK>) for NK = n, 1 <= n <= MAXN
template <int N, int MAXN, typename F>
__device__ __forceinline__ void mx_dispatch(int n, F&& f) {
  if constexpr (N <= MAXN) {
    if (n == N) {
      f(std::integral_constant<int, N>{});
      return;
    }
    mx_dispatch<N + 1, MAXN>(n, f);
  }
}

// Stage 1 (one wavefront per MTask1; workgroups [0, g_red) take the
// like-particle blocks as in k_stk1): X_u^T[j][i] = (V_R Bhat^T)[i][j] for the
// task's 16 columns and each of its row tiles (MUnit1), the next tile's V
// operand loaded while the current one is multiplied and stored.
template <typename T, int KS>
__global__ __launch_bounds__(256) void k_mx1(const MTask1* __restrict__ tasks, int n_tasks,
                                             const MUnit1* __restrict__ units, const T* __restrict__ V,
                                             const T* __restrict__ Vp, const T* __restrict__ Bt,
                                             typename BigIn<T>::C* __restrict__ X, const DBlock* __restrict__ blocks,
                                             int n_blocks, int g_red, const T* __restrict__ H,
                                             double* __restrict__ W, int lds_cap, int skip_large, EnsStride es) {
  using C = typename BigIn<T>::C;
  using M = BigMma<C>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int bx = blockIdx.x;
  if (es.n > 1) {
    int64_t smp;
    bx = ens_block(es, smp);
    if (bx < 0) return;
    V += smp * es.v;
    Vp += smp * es.v;
    W += smp * es.v;
    Bt += smp * es.b;
    H += smp * es.h;
    X += smp * es.x;
  }
  if (bx < g_red) {
    if (bx < n_blocks)
      reduce_block<T, 256>(blocks[bx], nullptr, nullptr, V, H, W, 0, reinterpret_cast<T*>(smem_raw), lds_cap,
                           skip_large);
    return;
  }
  const int n_wg = (n_tasks + 3) / 4;
  const int wg = es.n > 1 ? bx - g_red : xcd_chunk(bx - g_red, n_wg);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63, ti = wg * 4 + wave;
  if (ti >= n_tasks) return;  // whole wavefronts; no barrier follows
  const MTask1 t = tasks[ti];
  const int kq = lane >> 4, l16 = lane & 15;
  const int dbr = t.dbr, db = t.db, j = t.j0 + l16;
  const bool jok = j < db;
  mx_dispatch<1, KS>((dbr + 3) >> 2, [&](auto nk) {
    constexpr int NK = decltype(nk)::value;
    C b[NK];
    const T* B = Bt + t.b_off;
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) {
      const int k = 4 * ks + kq;
      const bool ok = k < dbr && jok;
      const C x = BigIn<T>::get(B[ok ? k + dbr * j : 0]);
      b[ks] = ok ? x : C(0);
    }
    auto load_a = [&](const MUnit1& s, C (&a)[NK]) {
      const T* Vs = Vp + s.v_off;
      const bool iok = l16 < s.rows;
#pragma unroll
      for (int ks = 0; ks < NK; ++ks) {
        const int k = 4 * ks + kq;
        const bool ok = k < dbr && iok;
        const C x = BigIn<T>::get(Vs[ok ? l16 + s.ld * k : 0]);
        a[ks] = ok ? x : C(0);
      }
    };
    MUnit1 cur = units[t.u_begin];
    C a[NK];
    load_a(cur, a);
    for (int u = t.u_begin; u < t.u_end; ++u) {
      const bool more = u + 1 < t.u_end;
      const MUnit1 nxt = units[more ? u + 1 : u];
      C an[NK];
      load_a(nxt, an);  // the last one re-reads the current tile
      typename M::vec acc = {C(0), C(0), C(0), C(0)};
#pragma unroll
      for (int ks = 0; ks < NK; ++ks) acc = M::mma(a[ks], b[ks], acc);
      C* Xs = X + cur.x_off;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = M::row(lane, v);
        if (r < cur.rows && jok) Xs[j + db * r] = acc[v];
      }
#pragma unroll
      for (int ks = 0; ks < NK; ++ks) a[ks] = an[ks];
      cur = nxt;
    }
  });
}

// Stage 2 (one workgroup per MTask2): the transposed tile
// W_L^T[j][i] += sum_k Y_k^T[j][:] A_k^T[:][i], Y_k^T[j][k'] = sum_u r_u X_u^T[j][k'],
// the wavefronts taking terms t_begin + w, + w + 4, ...; their four partial
// tiles are summed in LDS and the sum is added into W with FP64 atomics.
// The X references of a term are read RU at a time, all their K steps' loads
// in flight together with A's.
template <typename T, int KS, int RU>
__global__ __launch_bounds__(256) void k_mx2(const MTask2* __restrict__ tasks, int n_tasks,
                                             const STerm2* __restrict__ terms, const SXRef* __restrict__ xrefs,
                                             const T* __restrict__ Ad, const typename BigIn<T>::C* __restrict__ X,
                                             double* __restrict__ W, EnsStride es) {
  using C = typename BigIn<T>::C;
  using M = BigMma<C>;
  __shared__ C red[4][4][64];
  int ti;
  if (es.n > 1) {
    int64_t smp;
    ti = ens_block(es, smp);
    if (ti < 0) return;
    X += smp * es.x;
    W += smp * es.v;
  } else {
    ti = xcd_chunk(blockIdx.x, n_tasks);
  }
  if (ti >= n_tasks) return;  // whole workgroups
  const MTask2 t = tasks[ti];
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int kq = lane >> 4, l16 = lane & 15;
  const int da = t.da, db = t.db, j = t.j0 + l16, i = t.i0 + l16;
  const bool jok = j < db, iok = i < da;
  typename M::vec acc = {C(0), C(0), C(0), C(0)};
  for (int tt = t.t_begin + wave; tt < t.t_end; tt += 4) {
    const STerm2 s = terms[tt];
    const int dar = s.dar;
    mx_dispatch<1, KS>((dar + 3) >> 2, [&](auto nk) {
      constexpr int NK = decltype(nk)::value;
      C a[NK], y[NK];
      const T* A = Ad + s.a_off;
#pragma unroll
      for (int ks = 0; ks < NK; ++ks) {
        const int k = 4 * ks + kq;
        const bool ok = k < dar && iok;
        const C x = BigIn<T>::get(A[ok ? i + da * k : 0]);
        a[ks] = ok ? x : C(0);
        y[ks] = C(0);
      }
      for (int x0 = s.x_begin; x0 < s.x_end; x0 += RU) {
        const C* Xq[RU];
        C r[RU], xv[RU][NK];
#pragma unroll
        for (int q = 0; q < RU; ++q) {
          const bool ok = x0 + q < s.x_end;
          const SXRef xr = xrefs[ok ? x0 + q : x0];
          Xq[q] = X + xr.x_off;
          r[q] = ok ? C(xr.r) : C(0);
        }
#pragma unroll
        for (int q = 0; q < RU; ++q)
#pragma unroll
          for (int ks = 0; ks < NK; ++ks) {
            const int k = 4 * ks + kq;
            const bool ok = k < dar && jok;
            const C x = Xq[q][ok ? j + db * k : 0];
            xv[q][ks] = ok ? x : C(0);
          }
#pragma unroll
        for (int q = 0; q < RU; ++q)
#pragma unroll
          for (int ks = 0; ks < NK; ++ks) y[ks] += r[q] * xv[q][ks];
      }
#pragma unroll
      for (int ks = 0; ks < NK; ++ks) acc = M::mma(y[ks], a[ks], acc);
    });
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) red[wave][v][lane] = acc[v];
  __syncthreads();
  {  // thread (wave w, lane) sums element v = w of the four partial tiles
    const int v = wave;
    const C sum = (red[0][v][lane] + red[1][v][lane]) + (red[2][v][lane] + red[3][v][lane]);
    const int jr = t.j0 + M::row(lane, v);
    if (jr < db && iok) unsafeAtomicAdd(W + t.w_off + i + (int64_t)da * jr, (double)sum);
  }
}

// ---------------------------------------------------------------- fused matrix-core path
// One workgroup per MFTask (one 16-row tile of Ja', one 16-column tile of Jb,
// up to kMfSlots output tiles of the (instance, Ja', Jb) group); wavefront w
// takes steps [p[w], p[w + 1]).  Per step (slab): every product tile
// X_u = V_u Bhat_u^T on the matrix cores, Y_s += r_su X_u in registers, then
// W_s^T += Y_s^T A_s^T with Y_s's result registers as the first operand (the
// result register q of a 16 x 16 x 4 step holds, in lane l, row M::row(l, q)
// and column l % 16, i.e. K step q of the next product when A is read at
// those rows).  X and Y never leave the registers; the slots' tiles are summed
// over the four wavefronts in LDS and added into W with FP64 atomics (the
// like-particle part is stored before, by k_reduce).
// Occupancy: the tiles are bound by the latency of dependent operand loads,
// so with up to 6 K steps per product the register budget is capped for 4
// wavefronts per SIMD with 4 slots (6 dwords spilled; the compiler's own
// choice gives 3: sd 4p4n FP64 H*v 30.6 -> 27.5 us) and 5 with 2 (none).
// With 10 K steps the caps stay where the registers fit without spills.
template <int KS, int SL>
constexpr int mxf_waves() { return KS <= 6 ? (SL == 2 ? 5 : 4) : (SL == 2 ? 4 : 3); }
template <typename T, int KS, int SL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(mxf_waves<KS, SL>(), 8))) void k_mxf(const MFTask* __restrict__ tasks, int n_tasks,
                                             const MFStep* __restrict__ steps, const MFU* __restrict__ us,
                                             const T* __restrict__ Vp, const T* __restrict__ Ad,
                                             const T* __restrict__ Bt, double* __restrict__ W, EnsStride es,
                                             XcdSplit xs, double* __restrict__ P) {
  using C = typename BigIn<T>::C;
  using M = BigMma<C>;
  using vec = typename M::vec;
  __shared__ C red[4][SL][4][64];
  int ti;
  if (es.n > 1) {
    int64_t smp;
    ti = ens_block(es, smp);
    if (ti < 0) return;
    Vp += smp * es.v;
    W += smp * es.v;
    Bt += smp * es.bn;
    if (ti >= n_tasks) return;  // whole workgroups
  } else {  // workgroup b runs on XCD b % 8: the (b / 8)-th task of that XCD's range
    const int x = (int)(blockIdx.x % kXcds);
    ti = xs.s[x] + (int)(blockIdx.x / kXcds);
    if (ti >= xs.s[x + 1]) return;
  }
  const MFTask t = tasks[ti];
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int kq = lane >> 4, l16 = lane & 15;
  const bool aok = l16 < t.rows_a, jok = l16 < t.cols_j;
  vec w[SL];
#pragma unroll
  for (int q = 0; q < SL; ++q) w[q] = vec{C(0), C(0), C(0), C(0)};
  // the wavefront's step range, kept in scalar registers
  const int st_begin = __builtin_amdgcn_readfirstlane(wave == 0 ? t.p[0] : wave == 1 ? t.p[1] : wave == 2 ? t.p[2] : t.p[3]);
  const int st_end = __builtin_amdgcn_readfirstlane(wave == 0 ? t.p[1] : wave == 1 ? t.p[2] : wave == 2 ? t.p[3] : t.p[4]);
  // K steps of the W products that can hold rows of the Ja' tile: in the FP64
  // result layout register q holds rows 4q..4q+3, in the FP32 one rows q, q+4, ..
  const int nka = sizeof(C) == 8 ? (t.rows_a + 3) >> 2 : min(4, t.rows_a);
  // FP64 A operands: K step ks reads row kp = kq + 4 ks, so a lane's offsets
  // are one base per slot plus the uniform 4 da ks rows (scalar offset); only
  // the last step has rows beyond rows_a (lanes kq >= rows_a - 4 (nka - 1)).
  // Two registers per slot instead of one per slot and step.
  uint32_t abase[SL], alast[SL];
  if constexpr (sizeof(C) == 8) {
    const bool lok = kq < t.rows_a - 4 * (nka - 1);
#pragma unroll
    for (int q = 0; q < SL; ++q) {
      abase[q] = boff<T>(l16 < t.rows_i[q], l16 + t.da[q] * kq);
      alast[q] = lok ? abase[q] : kOobOff;
    }
  }
  mx_dispatch<1, 4>(nka, [&](auto nka_c) {
  constexpr int NKA = decltype(nka_c)::value;
  for (int st = st_begin; st < st_end; ++st) {
    const MFStep S = steps[st];
    vec y[SL];
#pragma unroll
    for (int q = 0; q < SL; ++q) y[q] = vec{C(0), C(0), C(0), C(0)};
    MFU U = us[S.u_begin];  // the next record is requested a product ahead
    for (int u = S.u_begin; u < S.u_end; ++u) {
      const MFU Un = us[u + 1 < S.u_end ? u + 1 : u];
      const int dbr = U.dbr;
      vec x = {C(0), C(0), C(0), C(0)};
      mx_dispatch<1, KS>((dbr + 3) >> 2, [&](auto nk) {
        constexpr int NK = decltype(nk)::value;
        C a[NK], b[NK];
        const __amdgpu_buffer_rsrc_t rv = mk_rsrc(Vp + U.v_off), rb = mk_rsrc(Bt + U.b_off);
#pragma unroll
        for (int ks = 0; ks < NK; ++ks) {
          const int k = 4 * ks + kq;
          a[ks] = BigIn<T>::get(BufLd<T>::ld(rv, boff<T>(k < dbr && aok, l16 + t.ld_v * k)));
          b[ks] = BigIn<T>::get(BufLd<T>::ld(rb, boff<T>(k < dbr && jok, l16 + t.db * k)));
        }
#pragma unroll
        for (int ks = 0; ks < NK; ++ks) x = M::mma(a[ks], b[ks], x);
      });
#pragma unroll
      for (int q = 0; q < SL; ++q) {
        const C r = C(U.r[q]);
#pragma unroll
        for (int v = 0; v < 4; ++v) y[q][v] += r * x[v];
      }
      U = Un;
    }
#pragma unroll
    for (int q = 0; q < SL; ++q) {
      if (S.a_off[q] < 0) continue;  // slot not in this slab (uniform)
      const __amdgpu_buffer_rsrc_t ra = mk_rsrc(Ad + S.a_off[q]);
      C av[NKA];
      if constexpr (sizeof(C) == 8) {
#pragma unroll
        for (int ks = 0; ks < NKA; ++ks)
          av[ks] = BigIn<T>::get(BufLd<T>::ld(ra, ks + 1 < NKA ? abase[q] : alast[q],
                                              (uint32_t)(4 * ks * t.da[q]) * (uint32_t)sizeof(T)));
      } else {
        const bool iok = l16 < t.rows_i[q];
#pragma unroll
        for (int ks = 0; ks < NKA; ++ks) {
          const int kp = M::row(lane, ks);
          av[ks] = BigIn<T>::get(BufLd<T>::ld(ra, boff<T>(kp < t.rows_a && iok, l16 + t.da[q] * kp)));
        }
      }
#pragma unroll
      for (int ks = 0; ks < NKA; ++ks) w[q] = M::mma(y[q][ks], av[ks], w[q]);
    }
  }
  });
#pragma unroll
  for (int q = 0; q < SL; ++q)
#pragma unroll
    for (int v = 0; v < 4; ++v) red[wave][q][v][lane] = w[q][v];
  __syncthreads();
  // thread (wave w, lane) sums register v = w of every slot over the wavefronts
  const int v = wave;
  const int jr = M::row(lane, v);
#pragma unroll
  for (int q = 0; q < SL; ++q) {
    if (t.w_off[q] < 0) continue;
    const C sum = (red[0][q][v][lane] + red[1][q][v][lane]) + (red[2][q][v][lane] + red[3][q][v][lane]);
    if (P)  // deterministic: the task's partial tile, element row + 16 col
      P[((int64_t)ti * SL + q) * 256 + l16 + 16 * jr] = (double)sum;
    else if (l16 < t.rows_i[q] && jr < t.cols_j)
      unsafeAtomicAdd(W + t.w_off[q] + l16 + (int64_t)t.da[q] * jr, (double)sum);
  }
}

// Deterministic fused path: one workgroup per W tile adds the tile's partial
// tiles onto W (one addend onto the stored value).  The sum has a fixed
// order: partial number n of the tile's list goes to accumulator n % 4, the
// accumulators meet as (s0 + s1) + (s2 + s3); eight partials are loaded at a
// time, so a thread's loads are in flight together.
__global__ __launch_bounds__(256) void k_mxf_sum(const MFDTile* __restrict__ tiles, const int32_t* __restrict__ parts,
                                                 const double* __restrict__ P, double* __restrict__ W) {
  const MFDTile t = tiles[blockIdx.x];
  const int e = threadIdx.x, r = e & 15, c = e >> 4;
  if (r >= t.rows || c >= t.cols) return;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  int k = t.c_begin;
  for (; k + 8 <= t.c_end; k += 8) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = P[(int64_t)parts[k + u] * 256 + e];
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u & 3] += x[u];
  }
  for (int u = 0; k < t.c_end; ++k, ++u) s[u & 3] += P[(int64_t)parts[k] * 256 + e];
  W[t.w_off + r + (int64_t)t.da * c] += (s[0] + s[1]) + (s[2] + s[3]);
}

// ---------------------------------------------------------------- instance tables
struct StagedInst {
  int prec, nt, mu, mode, atomic;
  const void* fn;
};
#define NSX_INST(T, P, NT, MU, MODE, AT) {P, NT, MU, MODE, AT, (const void*)k_pn_staged<T, NT, MU, MODE, AT>}
#define NSX_INSTS(T, P, MODE, AT)                                                                    \
  NSX_INST(T, P, 64, 1, MODE, AT), NSX_INST(T, P, 64, 2, MODE, AT), NSX_INST(T, P, 64, 4, MODE, AT), \
      NSX_INST(T, P, 64, 7, MODE, AT), NSX_INST(T, P, 128, 1, MODE, AT),                             \
      NSX_INST(T, P, 128, 2, MODE, AT), NSX_INST(T, P, 128, 4, MODE, AT),                            \
      NSX_INST(T, P, 256, 1, MODE, AT), NSX_INST(T, P, 256, 2, MODE, AT)
#define NSX_ALL(T, P) \
  NSX_INSTS(T, P, 0, false), NSX_INSTS(T, P, 1, false), NSX_INSTS(T, P, 0, true), NSX_INSTS(T, P, 1, true)
static const StagedInst kStaged[] = {NSX_ALL(double, 0), NSX_ALL(float, 1), NSX_ALL(Bf16x2, 2)};
#undef NSX_ALL
#undef NSX_INSTS
#undef NSX_INST
constexpr int kNumStaged = sizeof(kStaged) / sizeof(kStaged[0]);

// smallest instance with the requested precision, threads and mode covering max_units
static int pick_staged(int prec, int nt, int mode, bool atomic, int max_units) {
  int best = -1;
  for (int i = 0; i < kNumStaged; ++i) {
    const StagedInst& s = kStaged[i];
    if (s.prec != prec || s.nt != nt || s.mode != mode || s.atomic != (int)atomic ||
        s.nt * s.mu < max_units)
      continue;
    if (best < 0 || s.mu < kStaged[best].mu) best = i;
  }
  return best;
}

int num_staged() { return kNumStaged; }
const void* k_pn_staged_ptr(int i) { return kStaged[i].fn; }
const void* k_pn_partial_ptr(int prec, bool atomic) {
  switch (prec) {
    case 0: return atomic ? (const void*)k_pn_partial<double, true> : (const void*)k_pn_partial<double, false>;
    case 1: return atomic ? (const void*)k_pn_partial<float, true> : (const void*)k_pn_partial<float, false>;
    default: return atomic ? (const void*)k_pn_partial<Bf16x2, true> : (const void*)k_pn_partial<Bf16x2, false>;
  }
}

// ---------------------------------------------------------------- launchers
// FP64 plans in atomic mode run their large blocks on the GEMM tiles: the
// like-particle part (k_big0, before the pn kernels) and the pn part (k_big1, k_big2)
static bool big_like_active(const DevicePlan& d) { return d.atomic && d.n_gtile3 > 0; }

// V: the like-particle input (read by the stacked stage-1 launch, which carries
// those blocks); Vp: the pn input (V masked on truncated plans)
template <typename T>
static hipError_t pn_T(const DevicePlan& d, int prec, const T* V, const T* Vp, const T* Ad, const T* Bh, const T* H,
                       double* W, hipStream_t s, const T* Bnp = nullptr, const LowpPlan* lp = nullptr) {
  const bool at = d.atomic != 0;
  if (d.fused) {  // column partials of the small blocks (summed by the reduce)
    if (d.n_ftask > 0) {
      void* a[] = {(void*)&d.f_task, (void*)&d.n_ftask, (void*)&d.f_blk, (void*)&d.f_round, (void*)&d.f_seg,
                   (void*)&d.f_x,    (void*)&d.f_cons,  (void*)&d.f_ref, (void*)&Vp,        (void*)&Ad,
                   (void*)&Bh,       (void*)&d.f_part};
      const unsigned g = (unsigned)(((d.n_ftask + kXcds - 1) / kXcds) * kXcds);
      hipError_t e = hipLaunchKernel(k_fuse_ptr(prec, d.f_u), dim3(g), dim3(kFuseThreads), a,
                                     sizeof(T) * (size_t)d.f_lds, s);
      if (e != hipSuccess) return e;
    }
  } else if (d.stacked) {  // two-stage path for the small blocks (always accumulates into W)
    if (!at) return hipErrorInvalidConfiguration;
    using Acc = typename Prec<T>::acc;
    Acc* X = reinterpret_cast<Acc*>(d.X);
    // stage 1 with the like-particle blocks (the reduce of this H*v)
    const int g_red = ((d.n_blocks + kXcds - 1) / kXcds) * kXcds;
    const int g1 = ((d.n_chunk1 + kXcds - 1) / kXcds) * kXcds;
    if (g_red + g1 > 0) {
      const T* Hs = H;
      const int cap = d.s_lds1;
      const int skip = big_like_active(d) ? 1 : 0;
      void* a1[] = {(void*)&d.s_chunk1, (void*)&d.n_chunk1, (void*)&d.s_unit1, (void*)&V,   (void*)&Vp,
                    (void*)&Bh,         (void*)&X,          (void*)&d.blocks,  (void*)&d.n_blocks,
                    (void*)&g_red,      (void*)&Hs,         (void*)&W,         (void*)&cap, (void*)&skip,
                    (void*)&d.es};
      const unsigned gy = d.ens_s > 1 ? (unsigned)((d.ens_s + kXcds - 1) / kXcds * kXcds) : 1u;
      hipError_t e = hipLaunchKernel(k_stk_ptr(1, prec, d.stk_nt), dim3((unsigned)(g_red + g1), gy),
                                     dim3(d.stk_nt), a1, sizeof(T) * (size_t)d.s_lds1, s);
      if (e != hipSuccess) return e;
    }
    if (d.n_chunk2 > 0) {
      const Acc* Xc = X;
      void* a2[] = {(void*)&d.s_chunk2, (void*)&d.n_chunk2, (void*)&d.s_term2, (void*)&d.s_xref,
                    (void*)&Ad,         (void*)&Xc,         (void*)&W,          (void*)&d.es};
      const unsigned g2 = (unsigned)(((d.n_chunk2 + kXcds - 1) / kXcds) * kXcds);
      const size_t lds2 = std::max(sizeof(T) * (size_t)d.s_lds2, sizeof(Acc) * (size_t)d.stk_nt * kColBlk);
      const unsigned gy = d.ens_s > 1 ? (unsigned)((d.ens_s + kXcds - 1) / kXcds * kXcds) : 1u;
      hipError_t e = hipLaunchKernel(k_stk_ptr(2, prec, d.stk_nt), dim3(g2, gy), dim3(d.stk_nt), a2, lds2, s);
      if (e != hipSuccess) return e;
    }
  } else if (d.mxf) {  // fused matrix-core path for the small blocks (adds into W)
    if (!at) return hipErrorInvalidConfiguration;
    const T* Bn = Bnp;
    if (d.n_mftask > 0) {
      int most = 0;  // the largest XCD range sets the grid (ensembles: tasks per sample)
      for (int x = 0; x < kXcds; ++x) most = std::max(most, d.mxf_xcd.s[x + 1] - d.mxf_xcd.s[x]);
      const unsigned g = d.ens_s > 1 ? (unsigned)(((d.n_mftask + kXcds - 1) / kXcds) * kXcds)
                                     : (unsigned)(kXcds * most);
      const unsigned gy = d.ens_s > 1 ? (unsigned)((d.ens_s + kXcds - 1) / kXcds * kXcds) : 1u;
#define NSX_MXF(KS, SL)                                                                                   \
  hipLaunchKernelGGL((k_mxf<T, KS, SL>), dim3(g, gy), dim3(256), 0, s, d.f_mtask, d.n_mftask, d.f_mstep, d.f_mu, \
                     Vp, Ad, Bn, W, d.es, d.mxf_xcd, d.mxf_det ? d.f_dpsum : nullptr)
      if (d.mx_ks <= 6) {
        if (d.mxf_slots == 2) NSX_MXF(6, 2); else NSX_MXF(6, kMfSlots);
      } else {
        if (d.mxf_slots == 2) NSX_MXF(10, 2); else NSX_MXF(10, kMfSlots);
      }
#undef NSX_MXF
      if (d.mxf_det && d.n_dtile > 0)
        hipLaunchKernelGGL(k_mxf_sum, dim3((unsigned)d.n_dtile), dim3(256), 0, s, d.f_dtile, d.f_dpart,
                           (const double*)d.f_dpsum, W);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  } else if (d.mstk) {  // matrix-core stacked path for the small blocks (accumulates into W)
    if (!at) return hipErrorInvalidConfiguration;
    using Cx = typename BigIn<T>::C;
    Cx* X = static_cast<Cx*>(d.X);
    const int g_red = ((d.n_blocks + kXcds - 1) / kXcds) * kXcds;
    const int g1 = (((d.n_mtask1 + 3) / 4 + kXcds - 1) / kXcds) * kXcds;
    const unsigned gy = d.ens_s > 1 ? (unsigned)((d.ens_s + kXcds - 1) / kXcds * kXcds) : 1u;
    const int skip = big_like_active(d) ? 1 : 0, cap = d.s_lds1;
    const size_t lds1 = sizeof(T) * (size_t)cap;
    if (g_red + g1 > 0) {
      if (d.mx_ks <= 6)
        hipLaunchKernelGGL((k_mx1<T, 6>), dim3((unsigned)(g_red + g1), gy), dim3(256), lds1, s, d.m_task1,
                           d.n_mtask1, d.m_unit1, V, Vp, Bh, X, d.blocks, d.n_blocks, g_red, H, W, cap, skip, d.es);
      else
        hipLaunchKernelGGL((k_mx1<T, 10>), dim3((unsigned)(g_red + g1), gy), dim3(256), lds1, s, d.m_task1,
                           d.n_mtask1, d.m_unit1, V, Vp, Bh, X, d.blocks, d.n_blocks, g_red, H, W, cap, skip, d.es);
    }
    if (d.n_mtask2 > 0) {
      const unsigned g2 = (unsigned)(((d.n_mtask2 + kXcds - 1) / kXcds) * kXcds);
      if (d.mx_ks <= 6)
        hipLaunchKernelGGL((k_mx2<T, 6, 2>), dim3(g2, gy), dim3(256), 0, s, d.m_task2, d.n_mtask2, d.s_term2,
                           d.s_xref, Ad, (const Cx*)X, W, d.es);
      else
        hipLaunchKernelGGL((k_mx2<T, 10, 2>), dim3(g2, gy), dim3(256), 0, s, d.m_task2, d.n_mtask2, d.s_term2,
                           d.s_xref, Ad, (const Cx*)X, W, d.es);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  } else if (d.n_items2 > 0) {
    const int k = pick_staged(prec, d.staged_nt, d.staged_mode, at, d.max_units);
    if (k < 0) return hipErrorInvalidConfiguration;
    const size_t lds = sizeof(T) * (size_t)((d.staged_mode == 1 ? 2 : 1) * d.lds_buf + d.lds_y);
    void* args[] = {(void*)&d.items2, (void*)&d.phases, (void*)&d.pairs, (void*)&Vp,
                    (void*)&Ad,       (void*)&Bh,       (void*)&d.part,  (void*)&W,
                    (void*)&d.lds_buf};
    hipError_t e = hipLaunchKernel(kStaged[k].fn, dim3(d.n_items2), dim3(d.staged_nt), args, lds, s);
    if (e != hipSuccess) return e;
  }
  if (at && d.n_gtile2 > 0) {  // large blocks on the matrix cores
    // the plan's precision is fixed, so XG only ever holds X in this compute type
    using C = typename BigIn<T>::C;
    C* XG = static_cast<C*>(d.XG);
    if constexpr (std::is_same<T, Bf16x2>::value) {  // BF16 pairs on the BF16 matrix cores
      if (lp && lp->Bp && d.ens_s == 1) {  // operands as panels
        const int g1 = lp->xcd_grp1, g2 = lp->xcd_grp2;
        auto grid = [](int n, int g) { return (unsigned)(g > 0 ? (n + kXcds * g - 1) / (kXcds * g) * (kXcds * g) : n); };
        if (d.n_gtile1 > 0) {
          const dim3 gr(grid(d.n_gtile1, g1));
          if (d.big_rows1 == 2 * kBigTile)
            hipLaunchKernelGGL(k_big1_bfp<2 * kBigTile>, gr, dim3(256), 0, s, d.g_tile1, lp->g_task1p, lp->Vt,
                               lp->Bp, XG, d.n_gtile1, g1);
          else
            hipLaunchKernelGGL(k_big1_bfp<kBigTile>, gr, dim3(256), 0, s, d.g_tile1, lp->g_task1p, lp->Vt, lp->Bp,
                               XG, d.n_gtile1, g1);
        }
        hipLaunchKernelGGL(k_big2_bfp, dim3(grid(d.n_gtile2, g2)), dim3(256), 0, s, d.g_tile2, lp->g_term2p,
                           d.g_xref, lp->Ap, (const C*)XG, W, d.n_gtile2, g2);
        return hipGetLastError();
      }
      if (d.n_gtile1 > 0) {
        if (d.big_rows1 == 2 * kBigTile)
          hipLaunchKernelGGL(k_big1_bf<2 * kBigTile>, dim3((unsigned)d.n_gtile1, (unsigned)d.ens_s), dim3(256), 0, s,
                             d.g_tile1, d.g_task1, Vp, Bh, XG, d.es);
        else
          hipLaunchKernelGGL(k_big1_bf<kBigTile>, dim3((unsigned)d.n_gtile1, (unsigned)d.ens_s), dim3(256), 0, s,
                             d.g_tile1, d.g_task1, Vp, Bh, XG, d.es);
      }
      hipLaunchKernelGGL(k_big2_bf, dim3((unsigned)d.n_gtile2, (unsigned)d.ens_s), dim3(256), 0, s, d.g_tile2,
                         d.g_term2, d.g_xref, Ad, (const C*)XG, W, d.es);
    } else {
      if (d.n_gtile1 > 0)
        hipLaunchKernelGGL(k_big1<T>, dim3((unsigned)d.n_gtile1, (unsigned)d.ens_s), dim3(256), 0, s, d.g_tile1,
                           d.g_task1, Vp, Bh, XG, d.es);
      hipLaunchKernelGGL(k_big2<T>, dim3((unsigned)d.n_gtile2, (unsigned)d.ens_s), dim3(256), 0, s, d.g_tile2,
                         d.g_term2, d.g_xref, Ad, (const C*)XG, W, d.es);
    }
    return hipGetLastError();
  }
  if (d.n_items > 0) {
    const size_t lds = sizeof(T) * (size_t)std::max(d.max_lds, 1);
    void* args[] = {(void*)&d.items, (void*)&d.aterms, (void*)&d.rterms, (void*)&Vp,
                    (void*)&Ad,      (void*)&Bh,       (void*)&d.part,   (void*)&W};
    hipError_t e = hipLaunchKernel(k_pn_partial_ptr(prec, at), dim3(d.n_items), dim3(kThreads), args, lds, s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t reduce_T(const DevicePlan& d, int prec, const T* V, const T* H, double* W, hipStream_t s,
                           bool w_zero) {
  const bool big = big_like_active(d);
  if (d.fused) {  // W = column partials + like-particle part, stored (large blocks: k_big0)
    if (d.n_fblocks > 0)
      hipLaunchKernelGGL(k_reduce<T>, dim3(d.n_fblocks), dim3(kThreads), 0, s, d.f_blocks, d.f_part_list,
                         d.f_part, V, H, W, 1, big ? 1 : 0, EnsStride());
  }
  // the stacked stage-1 launch carries the small like-particle blocks
  const bool like_in_pn = d.mxf && d.mxf_like;  // the fused matrix-core path adds the like-particle part
  if (like_in_pn && d.N > 0 && !w_zero)
    hipLaunchKernelGGL(k_zero_w, dim3((unsigned)((d.N + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, W, d.N);
  if (!d.stacked && !d.mstk && !d.fused && !like_in_pn && d.n_blocks > 0)
    hipLaunchKernelGGL(k_reduce<T>, dim3(d.n_blocks, (unsigned)d.ens_s), dim3(kThreads), 0, s, d.blocks,
                       d.part_list, d.part, V, H, W, d.atomic ? 0 : 1, big ? 1 : 0, d.es);
  if (big) {
    if constexpr (std::is_same<T, Bf16x2>::value)
      hipLaunchKernelGGL(k_big0_bf, dim3((unsigned)d.n_gtile3, (unsigned)d.ens_s), dim3(256), 0, s, d.g_tile3, V, H,
                         W, d.es);
    else
      hipLaunchKernelGGL(k_big0<T>, dim3((unsigned)d.n_gtile3, (unsigned)d.ens_s), dim3(256), 0, s, d.g_tile3, V, H,
                         W, d.es);
  }
  return hipGetLastError();
}

// the pn input of a plan: the masked copy made by launch_convert on truncated plans
static const void* pn_input(const DevicePlan& d, const LowpPlan& lp, int prec, const double* V) {
  if (prec == 0) return d.vmask ? (const void*)d.Vm : (const void*)V;
  return d.vmask ? lp.vin_m : lp.vin;
}

hipError_t launch_pn(const DevicePlan& d, const LowpPlan& lp, int prec, const double* V, double* W,
                     hipStream_t s) {
  const void* Vp = pn_input(d, lp, prec, V);
  switch (prec) {
    case 0: return pn_T<double>(d, 0, V, (const double*)Vp, d.Ad, d.Bh, d.H, W, s, d.Bn);
    case 1:
      return pn_T<float>(d, 1, (const float*)lp.vin, (const float*)Vp, (const float*)lp.Ad, (const float*)lp.Bh,
                         (const float*)lp.H, W, s, (const float*)lp.Bn);
    default:
      return pn_T<Bf16x2>(d, 2, (const Bf16x2*)lp.vin, (const Bf16x2*)Vp, (const Bf16x2*)lp.Ad,
                          (const Bf16x2*)lp.Bh, (const Bf16x2*)lp.H, W, s, (const Bf16x2*)lp.Bn, &lp);
  }
}

hipError_t launch_reduce(const DevicePlan& d, const LowpPlan& lp, int prec, const double* V, double* W,
                         hipStream_t s, bool w_zero) {
  switch (prec) {
    case 0: return reduce_T<double>(d, 0, V, d.H, W, s, w_zero);
    case 1: return reduce_T<float>(d, 1, (const float*)lp.vin, (const float*)lp.H, W, s, w_zero);
    default: return reduce_T<Bf16x2>(d, 2, (const Bf16x2*)lp.vin, (const Bf16x2*)lp.H, W, s, w_zero);
  }
}

hipError_t launch_panels(const PanelChunk* chunks, int n_chunks, const PanelBlk* blks, const void* src, uint4* dst,
                         hipStream_t s) {
  if (n_chunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_panel_bf, dim3((unsigned)n_chunks), dim3(256), 0, s, chunks, blks, (const Bf16x2*)src, dst);
  return hipGetLastError();
}

hipError_t convert_to(int prec, const double* x, void* y, int64_t n, hipStream_t s) {
  if (prec == 0 || n == 0) return hipSuccess;
  const dim3 g((unsigned)((n + 255) / 256));
  if (prec == 1)
    hipLaunchKernelGGL(k_convert<float>, g, dim3(256), 0, s, x, (float*)y, n);
  else
    hipLaunchKernelGGL(k_convert<Bf16x2>, g, dim3(256), 0, s, x, (Bf16x2*)y, n);
  return hipGetLastError();
}

hipError_t launch_convert(const DevicePlan& d, const LowpPlan& lp, int prec, const double* V, hipStream_t s) {
  if (d.N == 0) return hipSuccess;
  const dim3 g((unsigned)((d.N + 255) / 256));
  if (d.vmask) {  // the pn part's masked input
    if (prec == 0)
      hipLaunchKernelGGL(k_mask_in<double>, g, dim3(256), 0, s, V, d.vmask, d.Vm, d.N);
    else if (prec == 1)
      hipLaunchKernelGGL(k_mask_in<float>, g, dim3(256), 0, s, V, d.vmask, (float*)lp.vin_m, d.N);
    else
      hipLaunchKernelGGL(k_mask_in<Bf16x2>, g, dim3(256), 0, s, V, d.vmask, (Bf16x2*)lp.vin_m, d.N);
  }
  if (prec == 0) return hipGetLastError();
  if (prec == 1)
    hipLaunchKernelGGL(k_convert<float>, g, dim3(256), 0, s, V, (float*)lp.vin, d.N);
  else
    hipLaunchKernelGGL(k_convert<Bf16x2>, g, dim3(256), 0, s, V, (Bf16x2*)lp.vin, d.N);
  if (prec == 2 && lp.n_vt_chunk > 0)  // the large-block tiles' V_R panels, from the pn input
    hipLaunchKernelGGL(k_panel_bf, dim3((unsigned)lp.n_vt_chunk), dim3(256), 0, s, lp.vt_chunk, lp.vt_blk,
                       (const Bf16x2*)(d.vmask ? lp.vin_m : lp.vin), lp.Vt);
  return hipGetLastError();
}

// One H*v: deterministic mode runs pn partials then the reduce; atomic mode
// writes the like-particle part first and lets the pn kernels add into W.
// fused plans: k_big0 (large blocks' like-particle part, stored), k_fuse (column
// partials), k_reduce (small blocks stored), then k_big1 / k_big2 add the
// large-block terms
template <typename T>
static hipError_t fused_hv_T(const DevicePlan& d, int prec, const T* V, const T* Vp, const T* Ad, const T* Bh,
                             const T* H, double* W, hipStream_t s, const LowpPlan* lp) {
  const bool big = big_like_active(d);
  if (big) {
    if constexpr (std::is_same<T, Bf16x2>::value)
      hipLaunchKernelGGL(k_big0_bf, dim3((unsigned)d.n_gtile3, (unsigned)d.ens_s), dim3(256), 0, s, d.g_tile3, V, H,
                         W, d.es);
    else
      hipLaunchKernelGGL(k_big0<T>, dim3((unsigned)d.n_gtile3, (unsigned)d.ens_s), dim3(256), 0, s, d.g_tile3, V, H,
                         W, d.es);
  }
  DevicePlan small = d;   // k_fuse alone (the large-block items run on the tiles below)
  small.n_gtile2 = 0;
  small.n_items = 0;
  hipError_t e = pn_T<T>(small, prec, V, Vp, Ad, Bh, H, W, s);
  if (e != hipSuccess) return e;
  if (d.n_fblocks > 0)
    hipLaunchKernelGGL(k_reduce<T>, dim3(d.n_fblocks), dim3(kThreads), 0, s, d.f_blocks, d.f_part_list, d.f_part,
                       V, H, W, 1, big ? 1 : 0, EnsStride());
  if (d.n_gtile2 > 0) {
    DevicePlan lg = d;    // the large-block tiles alone
    lg.n_ftask = 0;
    return pn_T<T>(lg, prec, V, Vp, Ad, Bh, H, W, s, nullptr, lp);
  }
  return hipGetLastError();
}

hipError_t launch_hv(const DevicePlan& d, const LowpPlan& lp, int prec, const double* V, double* W,
                     hipStream_t s, bool w_zero) {
  hipError_t e = launch_convert(d, lp, prec, V, s);
  if (e != hipSuccess) return e;
  if (d.fused) {
    const void* Vp = pn_input(d, lp, prec, V);
    switch (prec) {
      case 0: return fused_hv_T<double>(d, 0, V, (const double*)Vp, d.Ad, d.Bh, d.H, W, s, &lp);
      case 1:
        return fused_hv_T<float>(d, 1, (const float*)lp.vin, (const float*)Vp, (const float*)lp.Ad,
                                 (const float*)lp.Bh, (const float*)lp.H, W, s, &lp);
      default:
        return fused_hv_T<Bf16x2>(d, 2, (const Bf16x2*)lp.vin, (const Bf16x2*)Vp, (const Bf16x2*)lp.Ad,
                                  (const Bf16x2*)lp.Bh, (const Bf16x2*)lp.H, W, s, &lp);
    }
  }
  if (d.atomic) {
    e = launch_reduce(d, lp, prec, V, W, s, w_zero);
    if (e != hipSuccess) return e;
    return launch_pn(d, lp, prec, V, W, s);
  }
  e = launch_pn(d, lp, prec, V, W, s);
  if (e != hipSuccess) return e;
  return launch_reduce(d, lp, prec, V, W, s);
}

}  // namespace nsx
