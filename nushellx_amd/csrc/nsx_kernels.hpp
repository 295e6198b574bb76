// Device-side plan and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "nsx_plan.hpp"

namespace nsx {

struct DevicePlan {
  DItem2* items2 = nullptr;
  DPhase* phases = nullptr;
  DPair* pairs = nullptr;
  int n_items2 = 0, lds_y = 1, lds_buf = 1;
  int max_units = 1;       // staged kernel: largest units per item
  int staged_nt = 128;     // staged kernel threads per workgroup
  int staged_mode = 0;     // 0: copy at phase start, 1: direct LDS loads one phase ahead
  int atomic = 0;          // 1: pn kernels add into W with FP64 atomics (no partials)
  size_t lds2_bytes = 0;
  DItem* items = nullptr;
  DATerm* aterms = nullptr;
  DRTerm* rterms = nullptr;
  DBlock* blocks = nullptr;
  int64_t* part_list = nullptr;
  double *Ad = nullptr, *Bh = nullptr, *H = nullptr, *part = nullptr;
  int n_items = 0, n_blocks = 0;
  size_t lds_bytes = 0;
  int64_t N = 0;
};

// Uploads a host plan; returns "" or an error message.
std::string upload_plan(const HostPlan& hp, DevicePlan& d);
void free_plan(DevicePlan& d);

hipError_t launch_hv_f64(const DevicePlan& d, const double* V, double* W, hipStream_t s);
hipError_t launch_pn_f64(const DevicePlan& d, const double* V, double* W, hipStream_t s);
hipError_t launch_reduce_f64(const DevicePlan& d, const double* V, double* W, hipStream_t s);
const void* k_pn_partial_ptr();
const void* k_pn_partial_atomic_ptr();
const void* k_pn_staged_ptr(int i);
int num_staged();

// Reduced-precision operands of the block contractions (FP32 or BF16 hi/lo).
struct LowpPlan {
  int precision = 0;
  void* Ad = nullptr;   // operand copies in the working precision
  void* Bh = nullptr;
  void* H = nullptr;
  void* vin = nullptr;  // input vector converted per call
};
std::string build_lowp(const HostPlan& hp, int precision, LowpPlan& lp);
void free_lowp(LowpPlan& lp);
hipError_t launch_hv_lowp(const DevicePlan& d, const LowpPlan& lp, const double* V, double* W,
                          hipStream_t s);
hipError_t launch_pn_lowp(const DevicePlan& d, const LowpPlan& lp, const double* V, double* W, hipStream_t s);
hipError_t launch_reduce_lowp(const DevicePlan& d, const LowpPlan& lp, const double* V, double* W,
                              hipStream_t s);

}  // namespace nsx
