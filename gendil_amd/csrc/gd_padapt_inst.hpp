// Instantiation of the p-adaptive kernels for one precision.
#pragma once

#include "gd_internal.hpp"
#include "gd_padapt.hpp"
#include "gd_padapt_api.hpp"

namespace gd {
namespace pa {

template <typename T, int D, bool ADV, int TPE4, int TPE3, int TPE2>
int launch_one(int grid, int it0, int it1, const void* pp, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s)
{
  constexpr size_t lds = (size_t)lds_words<D, TPE4, TPE3, TPE2>() * sizeof(T);
  static_assert(lds <= 160 * 1024, "p-adaptive kernel exceeds the LDS of one CU");
  auto* k = &padapt_kernel<T, D, ADV, TPE4, TPE3, TPE2>;
  static bool attr_set = false;
  if (!attr_set) {
    GD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    attr_set = true;
  }
  if (grid <= 0) return GD_OK;
  hipLaunchKernelGGL(k, dim3(grid), dim3(kNT), lds, s, static_cast<const PParams<T>*>(pp), it0, it1, mass_coef,
                     adv_coef, x, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

template <typename T, int D, bool ADV>
int launch_dim(int v, int grid, int it0, int it1, const void* pp, T m, T a, const T* x, T* y, hipStream_t s)
{
  if constexpr (D == 6) {
    if (v == 0) return launch_one<T, 6, ADV, 256, 64, 64>(grid, it0, it1, pp, m, a, x, y, s);
    return launch_one<T, 6, ADV, 256, 256, 64>(grid, it0, it1, pp, m, a, x, y, s);
  } else {
    return launch_one<T, D, ADV, 64, 64, 64>(grid, it0, it1, pp, m, a, x, y, s);
  }
}

template <typename T>
int launch_impl(int D, bool adv, int v, int grid, int it0, int it1, const void* pp, T m, T a, const T* x, T* y,
                hipStream_t s)
{
  switch (D) {
    case 3: return adv ? launch_dim<T, 3, true>(v, grid, it0, it1, pp, m, a, x, y, s)
                       : launch_dim<T, 3, false>(v, grid, it0, it1, pp, m, a, x, y, s);
    case 6: return adv ? launch_dim<T, 6, true>(v, grid, it0, it1, pp, m, a, x, y, s)
                       : launch_dim<T, 6, false>(v, grid, it0, it1, pp, m, a, x, y, s);
    default: return fail(GD_ERR_UNSUPPORTED, "p-adaptive apply: dimension not compiled");
  }
}

}  // namespace pa
}  // namespace gd
