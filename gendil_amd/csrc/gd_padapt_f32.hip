// p-adaptive kernels, FP32.
#include "gd_padapt_inst.hpp"

namespace gd {
namespace pa {
template <>
int launch<float>(int D, bool adv, int v, int grid, int it0, int it1, const void* pp, float m, float a,
                   const float* x, float* y, hipStream_t s)
{
  return launch_impl<float>(D, adv, v, grid, it0, it1, pp, m, a, x, y, s);
}
}  // namespace pa
}  // namespace gd
