// Handle types behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "nsx_kernels.hpp"
#include "nsx_model.hpp"
#include "nsx_plan.hpp"

struct nsx_hv {
  nsx::Model model;
  nsx::HostPlan hplan;
  nsx::DevicePlan dplan;
  nsx::LowpPlan lowp;
  int precision = 0;
  hipStream_t stream = nullptr;
  double* d_in = nullptr;
  double* d_out = nullptr;
};

struct nsx_batch {
  nsx::HostPlan hplan;
  nsx::DevicePlan dplan;
  nsx::LowpPlan lowp;
  int precision = 0;
  hipStream_t stream = nullptr;
};

namespace nsx {
void set_error(const std::string& e);
}
