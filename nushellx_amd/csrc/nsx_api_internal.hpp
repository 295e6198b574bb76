// Handle types behind the C ABI.
#pragma once
#include <cstdio>
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "nsx_kernels.hpp"
#include "nsx_model.hpp"
#include "nsx_plan.hpp"
#include "../../include/nsx_hv.h"

namespace nsx {
struct TrlWork;  // Lanczos workspace kept between solves (nsx_trl.hip)
void free_trl_work(TrlWork* w);
}  // namespace nsx

struct nsx_hv {
  nsx::Model model;
  nsx::HostPlan hplan;
  nsx::DevicePlan dplan;
  nsx::LowpPlan lowp;
  int precision = 0;
  bool host_only = false;  // no device plan: only usable inside a batch
  nsx::TrlWork* trl = nullptr;
  hipStream_t stream = nullptr;
  double* d_in = nullptr;
  double* d_out = nullptr;
  float* d_in32 = nullptr;   // REAL*4 vectors of the drop-in call (nsx_hv_apply_f32)
  float* d_out32 = nullptr;
};

// Linear ensemble of a batch (nsx_ens_create): per sample s the operator
// values are values[s] = theta[s] . basis (P parameters, K columns: the
// templates' .oph sections, then the H store), and the Bhat^T and H stores of
// every sample are produced from them on the device.
struct EnsData {
  int S = 0, P = 0;
  int64_t K = 0, n_o = 0, b_t = 0, h_t = 0, bn_t = 0;
  int n_fold = 0, n_bn = 0;
  double *theta = nullptr, *basis = nullptr, *values = nullptr, *bsrc = nullptr;
  int64_t* bn_src = nullptr;  // fused matrix-core plans: {Bh offset, Bn offset, dbr, db} per copied block
  nsx::EnsFold* fold = nullptr;
  nsx::EnsPair* pair = nullptr;
};

struct nsx_batch {
  std::vector<const nsx::Model*> models;  // owned by the instance handles
  nsx::HostPlan hplan;
  nsx::DevicePlan dplan;
  nsx::LowpPlan lowp;
  int precision = 0;
  nsx::TrlWork* trl = nullptr;
  hipStream_t stream = nullptr;
  EnsData ens;                            // S = 0: an ordinary batch
};

namespace nsx {
void set_error(const std::string& e);
// Fortran sequential unformatted records (4-byte length markers)
bool write_record(FILE* f, const std::vector<unsigned char>& payload);
bool read_records(const std::string& path, std::vector<std::vector<unsigned char>>& out);
void plan_models(const std::vector<const Model*>& models, HostPlan& hp, int precision);
PlanConfig plan_config(const nsx_tune_result& r);
void free_ens(EnsData& e);
// Rebuilds and uploads the plan of `models` with `cfg`; keeps kernel choices of `d`.
std::string replan(const std::vector<const Model*>& models, const PlanConfig& cfg, int precision,
                   HostPlan& hp, DevicePlan& d, LowpPlan& lp);
}
