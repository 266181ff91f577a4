// C-ABI Krylov solves: thin wrappers over the header-only solvers in
// include/gendil_amd/krylov.hpp.
#include "gendil_amd/krylov.hpp"

#include "gd_internal.hpp"

namespace {

gendil_amd::SolveOptions to_options(const gd_solve_params* p)
{
  gendil_amd::SolveOptions o;
  if (p) {
    if (p->restart > 0) o.restart = p->restart;
    if (p->max_iters > 0) o.max_iters = p->max_iters;
    if (p->rel_tol > 0) o.rel_tol = p->rel_tol;
    if (p->max_refinements > 0) o.max_refinements = p->max_refinements;
    if (p->inner_rel_tol > 0) o.inner_rel_tol = p->inner_rel_tol;
  }
  return o;
}

void to_result(const gendil_amd::SolveReport& r, gd_solve_result* out)
{
  if (!out) return;
  out->success = r.success ? 1 : 0;
  out->iterations = r.iterations;
  out->refinements = r.cycles;
  out->final_rel_residual = r.final_rel_residual;
  out->seconds = r.seconds;
}

template <typename F>
int guarded(F&& f)
{
  try {
    return f();
  } catch (const gendil_amd::Error& e) {
    gd::set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    gd::set_error(e.what());
    return GD_ERR_INVALID;
  }
}

}  // namespace

extern "C" {

int gd_gmres_f64(const gd_plan* plan, double mass_coef, double adv_coef, const double* b, double* x,
                 const gd_solve_params* params, gd_solve_result* result, gd_stream stream)
{
  if (!plan || !b || !x) return gd::fail(GD_ERR_INVALID, "gd_gmres_f64: null argument");
  return guarded([&] {
    gendil_amd::CombinedOperator K{plan, mass_coef, adv_coef, stream};
    auto rep = gendil_amd::gmres_solve<double>(K, gd_plan_num_dofs(plan), b, x, to_options(params));
    to_result(rep, result);
    return rep.success ? GD_OK : GD_ERR_NOT_CONVERGED;
  });
}

int gd_gmres_mixed(const gd_plan* plan, double mass_coef, double adv_coef, const double* b, double* x,
                   const gd_solve_params* params, gd_solve_result* result, gd_stream stream)
{
  if (!plan || !b || !x) return gd::fail(GD_ERR_INVALID, "gd_gmres_mixed: null argument");
  return guarded([&] {
    gendil_amd::CombinedOperator K{plan, mass_coef, adv_coef, stream};
    auto rep = gendil_amd::gmres_solve<float>(K, gd_plan_num_dofs(plan), b, x, to_options(params));
    to_result(rep, result);
    return rep.success ? GD_OK : GD_ERR_NOT_CONVERGED;
  });
}

int gd_cg_f64(const gd_plan* plan, double mass_coef, double adv_coef, const double* b, double* x,
              const gd_solve_params* params, gd_solve_result* result, gd_stream stream)
{
  if (!plan || !b || !x) return gd::fail(GD_ERR_INVALID, "gd_cg_f64: null argument");
  return guarded([&] {
    gendil_amd::CombinedOperator K{plan, mass_coef, adv_coef, stream};
    auto rep = gendil_amd::cg_solve(K, gd_plan_num_dofs(plan), b, x, to_options(params));
    to_result(rep, result);
    return rep.success ? GD_OK : GD_ERR_NOT_CONVERGED;
  });
}

int gd_cg_mixed(const gd_plan* plan, double mass_coef, double adv_coef, const double* b, double* x,
                const gd_solve_params* params, gd_solve_result* result, gd_stream stream)
{
  if (!plan || !b || !x) return gd::fail(GD_ERR_INVALID, "gd_cg_mixed: null argument");
  return guarded([&] {
    gendil_amd::CombinedOperator K{plan, mass_coef, adv_coef, stream};
    auto rep = gendil_amd::cg_mixed_solve(K, gd_plan_num_dofs(plan), b, x, to_options(params));
    to_result(rep, result);
    return rep.success ? GD_OK : GD_ERR_NOT_CONVERGED;
  });
}

int gd_bicgstab_f64(const gd_plan* plan, double mass_coef, double adv_coef, const double* b, double* x,
                    const gd_solve_params* params, gd_solve_result* result, gd_stream stream)
{
  if (!plan || !b || !x) return gd::fail(GD_ERR_INVALID, "gd_bicgstab_f64: null argument");
  return guarded([&] {
    gendil_amd::CombinedOperator K{plan, mass_coef, adv_coef, stream};
    auto rep = gendil_amd::bicgstab_solve(K, gd_plan_num_dofs(plan), b, x, to_options(params));
    to_result(rep, result);
    return rep.success ? GD_OK : GD_ERR_NOT_CONVERGED;
  });
}

}  // extern "C"
