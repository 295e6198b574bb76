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

__global__ __launch_bounds__(kThreads) void k_pn_partial(
    const DItem* __restrict__ items, const DATerm* __restrict__ aterms,
    const DRTerm* __restrict__ rterms, const double* __restrict__ V,
    const double* __restrict__ Ad, const double* __restrict__ Bh, double* __restrict__ part) {
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
        const double* b = Bh + R.b_off + j;
        double s = 0.0;
        for (int k = 0; k < R.dbr; ++k) s = fma(v[(int64_t)dar * k], b[(int64_t)db * k], s);
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
      part[it.part_off + i + (int64_t)da * (j0 + jj)] = w[q];
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_reduce(
    const DBlock* __restrict__ blocks, const int64_t* __restrict__ part_list,
    const double* __restrict__ part, const double* __restrict__ V, const double* __restrict__ H,
    double* __restrict__ W) {
  const DBlock b = blocks[blockIdx.x];
  const int da = b.da, db = b.db, nW = da * b.nc;
  const double* v = V + b.v_off;
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

hipError_t launch_pn_f64(const DevicePlan& d, const double* V, hipStream_t s) {
  if (d.n_items > 0) {
    hipLaunchKernelGGL(k_pn_partial, dim3(d.n_items), dim3(kThreads), d.lds_bytes, s, d.items,
                       d.aterms, d.rterms, V, d.Ad, d.Bh, d.part);
  }
  return hipGetLastError();
}

hipError_t launch_reduce_f64(const DevicePlan& d, const double* V, double* W, hipStream_t s) {
  if (d.n_blocks > 0) {
    hipLaunchKernelGGL(k_reduce, dim3(d.n_blocks), dim3(kThreads), 0, s, d.blocks, d.part_list,
                       d.part, V, d.H, W);
  }
  return hipGetLastError();
}

hipError_t launch_hv_f64(const DevicePlan& d, const double* V, double* W, hipStream_t s) {
  hipError_t e = launch_pn_f64(d, V, s);
  if (e != hipSuccess) return e;
  return launch_reduce_f64(d, V, W, s);
}

const void* k_pn_partial_ptr() { return (const void*)k_pn_partial; }

std::string build_lowp(const HostPlan&, int, LowpPlan&) {
  return "reduced-precision H*v is not built yet";
}
void free_lowp(LowpPlan& lp) { lp = LowpPlan(); }
hipError_t launch_pn_lowp(const DevicePlan&, const LowpPlan&, const double*, hipStream_t) {
  return hipErrorNotSupported;
}
hipError_t launch_reduce_lowp(const DevicePlan&, const LowpPlan&, const double*, double*, hipStream_t) {
  return hipErrorNotSupported;
}
hipError_t launch_hv_lowp(const DevicePlan&, const LowpPlan&, const double*, double*, hipStream_t) {
  return hipErrorNotSupported;
}

}  // namespace nsx
