// p-adaptive kernels, FP64.
#include "gd_padapt_inst.hpp"

namespace gd {
namespace pa {
template <>
int launch<double>(int D, bool adv, int v, int grid, int it0, int it1, const void* pp, double m, double a,
                   const double* x, double* y, hipStream_t s)
{
  return launch_impl<double>(D, adv, v, grid, it0, it1, pp, m, a, x, y, s);
}
}  // namespace pa
}  // namespace gd
