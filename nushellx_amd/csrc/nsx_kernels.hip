// HIP kernels of the pn J-scheme H*v (FP64 path) for gfx950.
//
// Math per output block L = (Ja, Jb) (column-major d(Ja) x d(Jb)):
//   W_L = sum_items(L) sum_{a-terms} A_k Y_k        Y_k = sum_R r_R V_R Bhat_R^T
//       + H_a V_L + V_L H_b^T
// Kernel 1 owns one (item, column tile): Y lives in LDS, the A Y accumulation in
// registers; partials go to a private buffer.  Kernel 2 owns one (block, column
// tile): deterministic partial sum plus the like-particle products.  FP64 vector
// FMA runs at the FP64 matrix-core rate on MI355X (78.6 TFLOP/s both), so these
// ragged, tiny (<= 33 x 33 in the sd shell) contractions use VALU FMAs.
#include <hip/hip_runtime.h>

#include "nsx_kernels.hpp"

namespace nsx {

template <bool ATOMIC>
__global__ __launch_bounds__(kThreads) void k_pn_partial(
    const DItem* __restrict__ items, const DATerm* __restrict__ aterms,
    const DRTerm* __restrict__ rterms, const double* __restrict__ V,
    const double* __restrict__ Ad, const double* __restrict__ Bh, double* __restrict__ part,
    double* __restrict__ W) {
  extern __shared__ double Y[];
  const DItem it = items[blockIdx.x];
  const int da = it.da, dar = it.dar, db = it.db, nc = it.nc, j0 = it.j0;
  const int nW = da * nc, nY = dar * nc;
  double w[kMaxE];
#pragma unroll
  for (int q = 0; q < kMaxE; ++q) w[q] = 0.0;
  for (int a = it.a_begin; a < it.a_end; ++a) {
    const DATerm T = aterms[a];
    for (int e = threadIdx.x; e < nY; e += kThreads) {
      const int i = e % dar, jj = e / dar;
      const int j = j0 + jj;
      double y = 0.0;
      for (int u = T.r_begin; u < T.r_end; ++u) {
        const DRTerm R = rterms[u];
        const double* v = V + R.v_off + i;
        const double* b = Bh + R.b_off + (int64_t)R.dbr * j;  // Bhat^T column j
        double s = 0.0;
        for (int k = 0; k < R.dbr; ++k) s = fma(v[(int64_t)dar * k], b[k], s);
        y = fma(R.r, s, y);
      }
      Y[e] = y;
    }
    __syncthreads();
    const double* A = Ad + T.a_off;
#pragma unroll
    for (int q = 0; q < kMaxE; ++q) {
      const int e = threadIdx.x + kThreads * q;
      if (e < nW) {
        const int i = e % da, jj = e / da;
        double s = 0.0;
        for (int k = 0; k < dar; ++k) s = fma(A[i + (int64_t)da * k], Y[k + dar * jj], s);
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
      if (ATOMIC) unsafeAtomicAdd(W + it.w_off + ij, w[q]);
      else part[it.part_off + ij] = w[q];
    }
  }
}

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

// Staged pn kernel for blocks with all sides <= kSmallMax.  One workgroup (NT
// threads) runs a chunk of an item's phases.  A phase's operand runs (V_R
// blocks, Bhat^T blocks, A) are staged in LDS, either copied through registers
// at the start of the phase (MODE 0, single buffer) or fetched one phase ahead
// by direct global->LDS loads into a second buffer (MODE 1) so the copy of
// phase p+1 overlaps the arithmetic of phase p.  Y (dar x db) and W (da x db)
// accumulate in registers, MU (row, 4-column) units per thread: one LDS read of
// V (or A) feeds 4 FMAs against broadcast reads of Bhat^T (or Y).
template <int NT, int MU, int MODE, bool ATOMIC>
__global__ __launch_bounds__(NT) void k_pn_staged(
    const DItem2* __restrict__ items, const DPhase* __restrict__ phases,
    const DPair* __restrict__ pairs, const double* __restrict__ V, const double* __restrict__ Ad,
    const double* __restrict__ Bt, double* __restrict__ part, double* __restrict__ W, int lds_buf) {
  extern __shared__ double sm[];
  double* bufs[2] = {sm, sm + (MODE == 1 ? lds_buf : 0)};
  double* sY = sm + (MODE == 1 ? 2 : 1) * lds_buf;
  const DItem2 it = items[blockIdx.x];
  const int da = it.da, dar = it.dar, db = it.db;
  const int ncb = (db + kColBlk - 1) / kColBlk;
  const int nuY = dar * ncb, nuW = da * ncb;
  const int tid = threadIdx.x;
  double w[MU][kColBlk], y[MU][kColBlk];
#pragma unroll
  for (int q = 0; q < MU; ++q)
#pragma unroll
    for (int c = 0; c < kColBlk; ++c) { w[q][c] = 0.0; y[q][c] = 0.0; }

  auto issue = [&](const DPhase& P, double* dst) {
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
    dma_run<NT>(reinterpret_cast<const uint32_t*>(V + P.v_off), d, 2 * P.nv);
    dma_run<NT>(reinterpret_cast<const uint32_t*>(Bt + P.b_off), d + 2 * P.nv, 2 * P.nb);
    if (P.na > 0)
      dma_run<NT>(reinterpret_cast<const uint32_t*>(Ad + P.a_off), d + 2 * (P.nv + P.nb), 2 * P.na);
  };

  if (MODE == 1 && it.ph_begin < it.ph_end) issue(phases[it.ph_begin], bufs[0]);
  for (int p = it.ph_begin; p < it.ph_end; ++p) {
    const DPhase P = phases[p];
    double* buf = bufs[(p - it.ph_begin) & 1];
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
      const double* sV = buf + Q.v_lds;
      const double* sB = buf + Q.b_lds;
      const int dbr = Q.dbr;
#pragma unroll
      for (int q = 0; q < MU; ++q) {
        const int u = tid + NT * q;
        if (u < nuY) {
          const int i = u % dar, j0 = (u / dar) * kColBlk;
          const double* bc[kColBlk];
#pragma unroll
          for (int c = 0; c < kColBlk; ++c) bc[c] = sB + dbr * min(j0 + c, db - 1);
          double s0[kColBlk] = {0.0, 0.0, 0.0, 0.0}, s1[kColBlk] = {0.0, 0.0, 0.0, 0.0};
          int k = 0;
          for (; k + 1 < dbr; k += 2) {
            const double v0 = sV[i + dar * k], v1 = sV[i + dar * (k + 1)];
#pragma unroll
            for (int c = 0; c < kColBlk; ++c) {
              s0[c] = fma(v0, bc[c][k], s0[c]);
              s1[c] = fma(v1, bc[c][k + 1], s1[c]);
            }
          }
          if (k < dbr) {
            const double v0 = sV[i + dar * k];
#pragma unroll
            for (int c = 0; c < kColBlk; ++c) s0[c] = fma(v0, bc[c][k], s0[c]);
          }
#pragma unroll
          for (int c = 0; c < kColBlk; ++c) y[q][c] = fma(Q.r, s0[c] + s1[c], y[q][c]);
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
            if (j0 + c < db) sY[i + dar * (j0 + c)] = y[q][c];
        }
#pragma unroll
        for (int c = 0; c < kColBlk; ++c) y[q][c] = 0.0;
      }
      __syncthreads();
      const double* sA = buf + P.nv + P.nb;
#pragma unroll
      for (int q = 0; q < MU; ++q) {
        const int u = tid + NT * q;
        if (u < nuW) {
          const int i = u % da, j0 = (u / da) * kColBlk;
          const double* yc[kColBlk];
#pragma unroll
          for (int c = 0; c < kColBlk; ++c) yc[c] = sY + dar * min(j0 + c, db - 1);
          double s0[kColBlk] = {0.0, 0.0, 0.0, 0.0}, s1[kColBlk] = {0.0, 0.0, 0.0, 0.0};
          int k = 0;
          for (; k + 1 < dar; k += 2) {
            const double a0 = sA[i + da * k], a1 = sA[i + da * (k + 1)];
#pragma unroll
            for (int c = 0; c < kColBlk; ++c) {
              s0[c] = fma(a0, yc[c][k], s0[c]);
              s1[c] = fma(a1, yc[c][k + 1], s1[c]);
            }
          }
          if (k < dar) {
            const double a0 = sA[i + da * k];
#pragma unroll
            for (int c = 0; c < kColBlk; ++c) s0[c] = fma(a0, yc[c][k], s0[c]);
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
          if (ATOMIC) unsafeAtomicAdd(W + it.w_off + ij, w[q][c]);
          else part[it.part_off + ij] = w[q][c];
        }
    }
  }
}

// Instances of the staged kernel: (threads, units per thread, load mode).
struct StagedInst {
  int nt, mu, mode, atomic;
  const void* fn;
};
#define NSX_INST(NT, MU, MODE, AT) {NT, MU, MODE, AT, (const void*)k_pn_staged<NT, MU, MODE, AT>}
#define NSX_INSTS(MODE, AT)                                                                 \
  NSX_INST(64, 1, MODE, AT), NSX_INST(64, 2, MODE, AT), NSX_INST(64, 4, MODE, AT),          \
      NSX_INST(64, 7, MODE, AT), NSX_INST(128, 1, MODE, AT), NSX_INST(128, 2, MODE, AT),    \
      NSX_INST(128, 4, MODE, AT), NSX_INST(256, 1, MODE, AT), NSX_INST(256, 2, MODE, AT)
static const StagedInst kStaged[] = {NSX_INSTS(0, false), NSX_INSTS(1, false), NSX_INSTS(0, true),
                                     NSX_INSTS(1, true)};
#undef NSX_INSTS
#undef NSX_INST
constexpr int kNumStaged = sizeof(kStaged) / sizeof(kStaged[0]);

// smallest instance with the requested threads and mode covering max_units
static int pick_staged(int nt, int mode, bool atomic, int max_units) {
  int best = -1;
  for (int i = 0; i < kNumStaged; ++i) {
    const StagedInst& s = kStaged[i];
    if (s.nt != nt || s.mode != mode || s.atomic != (int)atomic || s.nt * s.mu < max_units) continue;
    if (best < 0 || s.mu < kStaged[best].mu) best = i;
  }
  return best;
}

// W_L = (sum of the pn partials of L, when use_partials) + H_a V_L + V_L H_b^T.
// In atomic mode this runs first (use_partials = 0) and the pn kernels add to W.
__global__ __launch_bounds__(kThreads) void k_reduce(
    const DBlock* __restrict__ blocks, const int64_t* __restrict__ part_list,
    const double* __restrict__ part, const double* __restrict__ V, const double* __restrict__ H,
    double* __restrict__ W, int use_partials) {
  DBlock b = blocks[blockIdx.x];
  if (!use_partials) b.p_end = b.p_begin;
  const int da = b.da, db = b.db, nW = da * b.nc;
  const double* v = V + b.v_off;
  if (da <= kSmallMax && db <= kSmallMax && b.j0 == 0 && b.nc == db) {
    // small block: stage V_L, H_a, H_b in LDS, then contract from LDS
    __shared__ double sv[kSmallMax * kSmallMax], sha[kSmallMax * kSmallMax],
        shb[kSmallMax * kSmallMax];
    const int nv = da * db;
    for (int e = threadIdx.x; e < nv; e += kThreads) sv[e] = v[e];
    if (b.ha_off >= 0)
      for (int e = threadIdx.x; e < da * da; e += kThreads) sha[e] = H[b.ha_off + e];
    if (b.hb_off >= 0)
      for (int e = threadIdx.x; e < db * db; e += kThreads) shb[e] = H[b.hb_off + e];
    __syncthreads();
    for (int e = threadIdx.x; e < nv; e += kThreads) {
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
        double s = 0.0;
        for (int k = 0; k < da; ++k) s = fma(sha[i + da * k], sv[k + da * j], s);
        w += s;
      }
      if (b.hb_off >= 0) {
        double s = 0.0;
        for (int k = 0; k < db; ++k) s = fma(sv[i + da * k], shb[j + db * k], s);
        w += s;
      }
      W[b.v_off + e] = w;
    }
    return;
  }
  for (int e = threadIdx.x; e < nW; e += kThreads) {
    const int i = e % da, j = b.j0 + e / da;
    const int64_t ij = i + (int64_t)da * j;
    double w = 0.0;
    for (int p = b.p_begin; p < b.p_end; ++p) w += part[part_list[p] + ij];
    if (b.ha_off >= 0) {
      const double* Ha = H + b.ha_off;
      double s = 0.0;
      for (int k = 0; k < da; ++k) s = fma(Ha[i + (int64_t)da * k], v[k + (int64_t)da * j], s);
      w += s;
    }
    if (b.hb_off >= 0) {
      const double* Hb = H + b.hb_off;
      double s = 0.0;
      for (int k = 0; k < db; ++k) s = fma(v[i + (int64_t)da * k], Hb[j + (int64_t)db * k], s);
      w += s;
    }
    W[b.v_off + ij] = w;
  }
}

hipError_t launch_pn_f64(const DevicePlan& d, const double* V, double* W, hipStream_t s) {
  const bool at = d.atomic != 0;
  if (d.n_items2 > 0) {
    const int k = pick_staged(d.staged_nt, d.staged_mode, at, d.max_units);
    if (k < 0) return hipErrorInvalidConfiguration;
    const size_t lds = sizeof(double) * (size_t)((d.staged_mode == 1 ? 2 : 1) * d.lds_buf + d.lds_y);
    void* args[] = {(void*)&d.items2, (void*)&d.phases, (void*)&d.pairs, (void*)&V,
                    (void*)&d.Ad,     (void*)&d.Bh,     (void*)&d.part,  (void*)&W,
                    (void*)&d.lds_buf};
    hipError_t e = hipLaunchKernel(kStaged[k].fn, dim3(d.n_items2), dim3(d.staged_nt), args, lds, s);
    if (e != hipSuccess) return e;
  }
  if (d.n_items > 0) {
    if (at)
      hipLaunchKernelGGL(k_pn_partial<true>, dim3(d.n_items), dim3(kThreads), d.lds_bytes, s, d.items,
                         d.aterms, d.rterms, V, d.Ad, d.Bh, d.part, W);
    else
      hipLaunchKernelGGL(k_pn_partial<false>, dim3(d.n_items), dim3(kThreads), d.lds_bytes, s, d.items,
                         d.aterms, d.rterms, V, d.Ad, d.Bh, d.part, W);
  }
  return hipGetLastError();
}

hipError_t launch_reduce_f64(const DevicePlan& d, const double* V, double* W, hipStream_t s) {
  if (d.n_blocks > 0) {
    hipLaunchKernelGGL(k_reduce, dim3(d.n_blocks), dim3(kThreads), 0, s, d.blocks, d.part_list,
                       d.part, V, d.H, W, d.atomic ? 0 : 1);
  }
  return hipGetLastError();
}

// One H*v: deterministic mode runs pn partials then the reduce; atomic mode
// writes the like-particle part first and lets the pn kernels add into W.
hipError_t launch_hv_f64(const DevicePlan& d, const double* V, double* W, hipStream_t s) {
  if (d.atomic) {
    hipError_t e = launch_reduce_f64(d, V, W, s);
    if (e != hipSuccess) return e;
    return launch_pn_f64(d, V, W, s);
  }
  hipError_t e = launch_pn_f64(d, V, W, s);
  if (e != hipSuccess) return e;
  return launch_reduce_f64(d, V, W, s);
}

const void* k_pn_partial_ptr() { return (const void*)k_pn_partial<false>; }
const void* k_pn_partial_atomic_ptr() { return (const void*)k_pn_partial<true>; }
// number of staged-kernel instances and their entry points (launch attributes)
int num_staged() { return kNumStaged; }
const void* k_pn_staged_ptr(int i) { return kStaged[i].fn; }

std::string build_lowp(const HostPlan&, int, LowpPlan&) {
  return "reduced-precision H*v is not built yet";
}
void free_lowp(LowpPlan& lp) { lp = LowpPlan(); }
hipError_t launch_pn_lowp(const DevicePlan&, const LowpPlan&, const double*, double*, hipStream_t) {
  return hipErrorNotSupported;
}
hipError_t launch_reduce_lowp(const DevicePlan&, const LowpPlan&, const double*, double*, hipStream_t) {
  return hipErrorNotSupported;
}
hipError_t launch_hv_lowp(const DevicePlan&, const LowpPlan&, const double*, double*, hipStream_t) {
  return hipErrorNotSupported;
}

}  // namespace nsx
