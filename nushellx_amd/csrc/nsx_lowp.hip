// Reduced-precision operand copies (FP32 or BF16 hi/lo) of a host plan.
#include <hip/hip_runtime.h>

#include <vector>

#include "nsx_kernels.hpp"

namespace nsx {

template <typename T, typename F>
static std::string upload_as(const std::vector<double>& src, void*& dst, F conv) {
  dst = nullptr;
  std::vector<T> tmp(src.size());
  for (size_t i = 0; i < src.size(); ++i) tmp[i] = conv(src[i]);
  if (tmp.empty()) return "";
  if (hipMalloc(&dst, sizeof(T) * tmp.size()) != hipSuccess) return "hipMalloc failed (reduced-precision operands)";
  if (hipMemcpy(dst, tmp.data(), sizeof(T) * tmp.size(), hipMemcpyHostToDevice) != hipSuccess)
    return "hipMemcpy failed (reduced-precision operands)";
  return "";
}

std::string build_lowp(const HostPlan& hp, int precision, LowpPlan& lp) {
  lp = LowpPlan();
  lp.precision = precision;
  std::string e;
  size_t esz;
  if (precision == 1) {
    auto cf = [](double x) { return (float)x; };
    if (!(e = upload_as<float>(hp.Ad, lp.Ad, cf)).empty()) return e;
    if (!(e = upload_as<float>(hp.Bh, lp.Bh, cf)).empty()) return e;
    if (!(e = upload_as<float>(hp.H, lp.H, cf)).empty()) return e;
    if (!(e = upload_as<float>(hp.Bn, lp.Bn, cf)).empty()) return e;
    esz = sizeof(float);
  } else if (precision == 2) {
    auto cb = [](double x) { return bf16x2_split((float)x); };
    if (!(e = upload_as<Bf16x2>(hp.Ad, lp.Ad, cb)).empty()) return e;
    if (!(e = upload_as<Bf16x2>(hp.Bh, lp.Bh, cb)).empty()) return e;
    if (!(e = upload_as<Bf16x2>(hp.H, lp.H, cb)).empty()) return e;
    if (!(e = upload_as<Bf16x2>(hp.Bn, lp.Bn, cb)).empty()) return e;
    esz = sizeof(Bf16x2);
  } else {
    return "";
  }
  if (hipMalloc(&lp.vin, esz * (size_t)(hp.N > 0 ? hp.N : 1)) != hipSuccess) return "hipMalloc failed (vector)";
  if (!hp.vmask.empty() && hipMalloc(&lp.vin_m, esz * (size_t)(hp.N > 0 ? hp.N : 1)) != hipSuccess)
    return "hipMalloc failed (masked vector)";
  return "";
}

void free_lowp(LowpPlan& lp) {
  (void)hipFree(lp.Ad);
  (void)hipFree(lp.Bh);
  (void)hipFree(lp.H);
  (void)hipFree(lp.vin);
  (void)hipFree(lp.vin_m);
  (void)hipFree(lp.Bn);
  lp = LowpPlan();
}

}  // namespace nsx
