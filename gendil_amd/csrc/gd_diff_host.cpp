// C-ABI SIPDG diffusion apply: argument checks, kernel lookup and launch
// (DiffusionOperator::operator(), diffusionoperator.hpp:469-476).
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "gd_diff.hpp"

namespace gd {
namespace {

const DiffSet* find_diff(int D, int N, int Q)
{
  DiffList l{nullptr, 0};
  if (D == 1 || D == 2) l = diff_kernels_d12();
  else if (D == 3) l = diff_kernels_d3();
  else if (D == 6) l = diff_kernels_d6();
  for (int i = 0; i < l.count; ++i)
    if (l.sets[i].D == D && l.sets[i].N == N && l.sets[i].Q == Q) return &l.sets[i];
  return nullptr;
}

template <typename T>
int diff_apply(const gd_plan* plan, const T* x, T* y, hipStream_t s)
{
  if (!plan) return fail(GD_ERR_INVALID, "null plan");
  if (plan->padapt) return fail(GD_ERR_UNSUPPORTED, "p-adaptive plans have no diffusion apply");
  if (!plan->diff_set) return fail(GD_ERR_INVALID, "diffusion coefficients not set (gd_plan_set_diffusion)");
  if (!x || !y) return fail(GD_ERR_INVALID, "null vector");
  const size_t bytes = (size_t)plan->ndofs * sizeof(T);
  const char* xb = reinterpret_cast<const char*>(x);
  const char* yb = reinterpret_cast<const char*>(y);
  if (xb < yb + bytes && yb < xb + bytes) return fail(GD_ERR_INVALID, "input and output must not overlap");
  const int N = plan->desc.order + 1, Q = plan->desc.num_quad;
  const DiffSet* ks = find_diff(plan->dim, N, Q);
  if (!ks)
    return fail(GD_ERR_UNSUPPORTED, "no diffusion kernel compiled for dim=" + std::to_string(plan->dim) +
                                        " nodes=" + std::to_string(N) + " quad=" + std::to_string(Q));
  const TableSet* ts = find_tables(plan, N, Q);
  if (!ts) return fail(GD_ERR_INVALID, "plan has no tables for its own rule");
  const Params<T> prm = make_params<T>(plan, ts, T(0), T(0));
  // the diffusivity descriptor goes to the kernel by value (host copy of the
  // device descriptor that prm.mu points to)
  if constexpr (std::is_same_v<T, double>) return ks->f64(prm, &plan->fields64_host[plan->dim + 2], x, y, s);
  else return ks->f32(prm, &plan->fields32_host[plan->dim + 2], x, y, s);
}

}  // namespace
}  // namespace gd

int gd_diff_apply_f64(const gd_plan* plan, const double* x, double* y, gd_stream stream)
{
  return gd::diff_apply<double>(plan, x, y, (hipStream_t)stream);
}

int gd_diff_apply_f32(const gd_plan* plan, const float* x, float* y, gd_stream stream)
{
  return gd::diff_apply<float>(plan, x, y, (hipStream_t)stream);
}
