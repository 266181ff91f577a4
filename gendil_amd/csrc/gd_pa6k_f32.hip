// Kronecker p-adaptive 6D kernel, FP32.
#include "gd_internal.hpp"
#include "gd_pa6k.hpp"
#include "gd_padapt_api.hpp"

namespace gd {
namespace pa {

template <>
int launch_kron<float>(bool adv, int grid, int it0, int it1, const void* pp, float mass_coef, float adv_coef,
                    const float* x, float* y, hipStream_t s)
{
  constexpr size_t lds = (size_t)pa6::block_words() * sizeof(float);
  static_assert(lds <= 160 * 1024, "Kronecker 6D kernel exceeds the LDS of one CU");
  auto* k = adv ? &pa6::pa6k_kernel<float, true> : &pa6::pa6k_kernel<float, false>;
  static bool attr_set[2] = {false, false};
  if (!attr_set[adv ? 1 : 0]) {
    GD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set[adv ? 1 : 0] = true;
  }
  if (grid <= 0) return GD_OK;
  hipLaunchKernelGGL(k, dim3(grid), dim3(256), lds, s, static_cast<const PParams<float>*>(pp), it0, it1, mass_coef,
                     adv_coef, x, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

}  // namespace pa
}  // namespace gd
