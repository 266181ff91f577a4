// The GenDiL-side binding (include/gendil_amd/gendil_policy.hpp) compiled
// against GenDiL's own headers.
//
// At compile time: the factories named with Mi355xKernelConfiguration and a
// descriptor return the MI355X operator for every value category of the
// descriptor, while the same factories with a lambda still return GenDiL's
// own operator types, and GMRES_no_alloc / BiCGSTAB_no_alloc on the MI355X
// operator resolve to the binding's device-resident overloads.
//
// At run time (needs the GPU): GenDiL's own SerialKernelConfiguration
// operators and the MI355X operators built from the same GenDiL space, rule
// and vectors give the same results (mass, advection with and without inflow
// data, mass + advection, a space at an L2Restriction shift), and GenDiL's own
// host-loop GMRES_no_alloc and BiCGSTAB_no_alloc on GenDiL's operator agree
// with the binding's device-resident overloads on the MI355X operator.
//
// Built by tests/cpp/Makefile where the GenDiL sources exist; prints one line
// per check and exits non-zero on the first mismatch.
#include <gendil/gendil.hpp>

#include "gendil_amd/gendil_policy.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace gendil;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kDim = 3;
constexpr int kOrder = 3;
constexpr int kQuad = 4;

using GLL3 = decltype(MakeLobattoFiniteElement(FiniteElementOrders<kOrder, kOrder, kOrder>{}));
using Space = decltype(MakeFiniteElementSpace(std::declval<CartesianMesh<3>>(), std::declval<GLL3>()));
using Rule = decltype(MakeIntegrationRule(IntegrationRuleNumPoints<kQuad, kQuad, kQuad>{}));

// the benchmark field: a lambda for GenDiL, a separable descriptor for the binding
auto c2_lambda = [](const std::array<Real, 3>& X, Real (&v)[3]) {
  v[0] = 1.0 + 0.5 * std::sin(2.0 * kPi * X[1]);
  v[1] = 0.5 + 0.5 * std::cos(2.0 * kPi * X[2]);
  v[2] = 0.25 + 0.25 * std::sin(2.0 * kPi * X[0]);
};

gendil_amd::VectorField<3> c2_descriptor()
{
  using gendil_amd::SeparableField;
  return gendil_amd::VectorField<3>{{
      SeparableField::OneAxis(1, [](double y) { return 1.0 + 0.5 * std::sin(2.0 * kPi * y); }),
      SeparableField::OneAxis(2, [](double z) { return 0.5 + 0.5 * std::cos(2.0 * kPi * z); }),
      SeparableField::OneAxis(0, [](double x) { return 0.25 + 0.25 * std::sin(2.0 * kPi * x); }),
  }};
}

// ---- overload resolution, checked by the compiler --------------------------

using Mi = Mi355xKernelConfiguration;
using VF = gendil_amd::VectorField<3>;
using AF = gendil_amd::AffineField<3>;
using SF = gendil_amd::SeparableField;

static_assert(std::is_same_v<decltype(MakeAdvectionOperator<Mi>(std::declval<const Space&>(),
                                                                std::declval<const Rule&>(), std::declval<VF&>())),
                             Mi355xOperator>);
static_assert(std::is_same_v<decltype(MakeAdvectionOperator<Mi>(std::declval<const Space&>(),
                                                                std::declval<const Rule&>(),
                                                                std::declval<const VF&>())),
                             Mi355xOperator>);
static_assert(std::is_same_v<decltype(MakeAdvectionOperator<Mi>(std::declval<const Space&>(),
                                                                std::declval<const Rule&>(), std::declval<VF>())),
                             Mi355xOperator>);
static_assert(std::is_same_v<decltype(MakeAdvectionOperator<Mi>(std::declval<const Space&>(),
                                                                std::declval<const Rule&>(), std::declval<AF&>())),
                             Mi355xOperator>);
static_assert(std::is_same_v<decltype(MakeAdvectionOperator<Mi>(std::declval<const Space&>(),
                                                                std::declval<const Rule&>(), std::declval<VF&>(),
                                                                std::declval<SF&>())),
                             Mi355xOperator>);
static_assert(std::is_same_v<decltype(MakeMassFiniteElementOperator<Mi>(std::declval<const Space&>(),
                                                                        std::declval<const Rule&>(),
                                                                        std::declval<SF&>())),
                             Mi355xOperator>);
static_assert(std::is_same_v<decltype(MakeMassAdvectionOperator<Mi>(std::declval<const Space&>(),
                                                                    std::declval<const Rule&>(), std::declval<VF&>(),
                                                                    std::declval<SF&>())),
                             Mi355xOperator>);
static_assert(std::is_same_v<decltype(MakeDiffusionOperator<Mi>(std::declval<const Space&>(),
                                                                std::declval<const Rule&>(), std::declval<SF&>(),
                                                                -1.0, 16.0)),
                             Mi355xOperator>);
// a lambda still builds GenDiL's own operator
static_assert(!std::is_same_v<decltype(MakeAdvectionOperator<SerialKernelConfiguration>(
                                  std::declval<const Space&>(), std::declval<const Rule&>(), c2_lambda)),
                              Mi355xOperator>);

// the solver calls keep GenDiL's call shape and, on the MI355X operator,
// resolve to the binding's device-resident loops (Mi355xSolverResult), on a
// lambda-built GenDiL operator to GenDiL's own (IterativeSolverResult)
using DotLambda = Real (*)(const Vector&, const Vector&);
static_assert(std::is_same_v<decltype(GMRES_no_alloc(std::declval<const Mi355xOperator&>(), std::declval<Vector&>(),
                                                     std::declval<DotLambda&>(), Integer{}, Integer{}, Real{},
                                                     std::declval<Vector&>(), std::declval<std::vector<Vector>&>(),
                                                     std::declval<Vector&>(), std::declval<std::vector<Real>&>(),
                                                     std::declval<std::vector<Real>&>(),
                                                     std::declval<std::vector<Real>&>(),
                                                     std::declval<std::vector<Real>&>(),
                                                     std::declval<std::vector<Real>&>())),
                             Mi355xSolverResult>);
static_assert(std::is_same_v<decltype(BiCGSTAB_no_alloc(std::declval<const Mi355xOperator&>(),
                                                        std::declval<Vector&>(), std::declval<DotLambda&>(),
                                                        Integer{}, Real{}, std::declval<Vector&>(),
                                                        std::declval<Vector&>(), std::declval<Vector&>(),
                                                        std::declval<Vector&>(), std::declval<Vector&>(),
                                                        std::declval<Vector&>(), std::declval<Vector&>())),
                             Mi355xSolverResult>);

// ---- run-time comparison ----------------------------------------------------

double rel_l2(const Vector& a, const Vector& b, Integer begin, Integer end)
{
  const Real* x = a.ReadHostData();
  const Real* y = b.ReadHostData();
  double num = 0.0, den = 0.0;
  for (Integer i = begin; i < end; ++i) {
    num += (x[i] - y[i]) * (x[i] - y[i]);
    den += y[i] * y[i];
  }
  return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

int failures = 0;

void report(const char* what, double err, double tol)
{
  const bool ok = err <= tol;
  std::printf("%-58s rel l2 %.3e  (tol %.0e) %s\n", what, err, tol, ok ? "ok" : "FAIL");
  std::fflush(stdout);
  failures += ok ? 0 : 1;
}

template <typename RefOp, typename MiOp>
void compare(const char* what, const RefOp& ref, const MiOp& mi, Integer n, Integer begin, Integer end, unsigned seed)
{
  Vector x(n), y_ref(n), y_mi(n);
  FillRandom(x, seed);
  y_ref = 0.0;
  y_mi = 0.0;
  ref(x, y_ref);
  mi(x, y_mi);
  report(what, rel_l2(y_mi, y_ref, begin, end), 1e-12);
}

}  // namespace

int main()
{
  const Integer n = 6;
  const Real h = 1.0 / n;
  // 1) the benchmark operator on a periodic mesh, GenDiL's CartesianMesh<3>
  {
    CartesianMesh<3> mesh({n, n, n}, {h, h, h}, {0.0, 0.0, 0.0}, true);
    auto fes = MakeFiniteElementSpace(mesh, MakeLobattoFiniteElement(FiniteElementOrders<kOrder, kOrder, kOrder>{}));
    auto ir = MakeIntegrationRule(IntegrationRuleNumPoints<kQuad, kQuad, kQuad>{});
    const Integer nd = fes.GetNumberOfFiniteElementDofs();
    auto A_ref = MakeAdvectionOperator<SerialKernelConfiguration>(fes, ir, c2_lambda);
    auto A_mi = MakeAdvectionOperator<Mi>(fes, ir, c2_descriptor());
    compare("advection, periodic, benchmark field", A_ref, A_mi, nd, 0, nd, 12345);

    auto sigma = [](const std::array<Real, 3>& X) -> Real { return 1.0 + X[0] * X[1] + X[2] * X[2]; };
    SF sigma_d;
    sigma_d.terms.push_back(SF::Term{1.0, {}});
    {
      SF::Term t;
      t.factor[0] = [](double x) { return x; };
      t.factor[1] = [](double y) { return y; };
      sigma_d.terms.push_back(t);
    }
    {
      SF::Term t;
      t.factor[2] = [](double z) { return z * z; };
      sigma_d.terms.push_back(t);
    }
    auto M_ref = MakeMassFiniteElementOperator<SerialKernelConfiguration>(fes, ir, sigma);
    auto M_mi = MakeMassFiniteElementOperator<Mi>(fes, ir, sigma_d);
    compare("mass, sigma = 1 + xy + z^2", M_ref, M_mi, nd, 0, nd, 7);

    // SIPDG diffusion with the same field as diffusivity (the harness's own
    // diffusion fixtures cover the other rules and dimensions)
    auto D_ref = MakeDiffusionOperator<SerialKernelConfiguration>(fes, ir, sigma, -1.0, 16.0);
    auto D_mi = MakeDiffusionOperator<Mi>(fes, ir, sigma_d, -1.0, 16.0);
    compare("SIPDG diffusion, mu = 1 + xy + z^2", D_ref, D_mi, nd, 0, nd, 9);

    auto MA_ref = MakeMassAdvectionOperator<SerialKernelConfiguration>(fes, ir, c2_lambda, sigma);
    auto MA_mi = MakeMassAdvectionOperator<Mi>(fes, ir, c2_descriptor(), sigma_d);
    compare("mass + advection (one pass)", MA_ref, MA_mi, nd, 0, nd, 8);

    // GenDiL's own Krylov loops (host dot products) on GenDiL's operator
    // against the binding's device-resident overloads on the MI355X operator:
    // K = M(1/dt) + A on the periodic mesh
    const Real dt = 0.02;
    auto inv_dt = [dt](const std::array<Real, 3>&) -> Real { return 1.0 / dt; };
    auto K_ref = MakeMassAdvectionOperator<SerialKernelConfiguration>(fes, ir, c2_lambda, inv_dt);
    auto K_mi = MakeMassAdvectionOperator<Mi>(fes, ir, c2_descriptor(), SF::Constant(1.0 / dt));
    Vector b(nd);
    FillRandom(b, 99);
    auto dot = [](const Vector& u, const Vector& v) { return Dot(u, v); };
    // the element mass matrices of GLL p = 3 dominate K's spectrum: GMRES(20)
    // needs several hundred iterations to 1e-8
    const Integer restart = 20, max_it = 2000;
    const Real tol = 1e-8;
    {
      std::vector<Vector> V;
      for (Integer i = 0; i <= restart; ++i) V.emplace_back(nd);
      Vector w(nd), x_ref(nd), x_mi(nd);
      std::vector<Real> H((restart + 1) * restart), cs(restart), sn(restart), e1(restart + 1), yv(restart);
      x_ref = 0.0;
      x_mi = 0.0;
      const auto r_ref = GMRES_no_alloc(K_ref, b, dot, max_it, restart, tol, x_ref, V, w, H, cs, sn, e1, yv);
      const auto r_mi = GMRES_no_alloc(K_mi, b, dot, max_it, restart, tol, x_mi, V, w, H, cs, sn, e1, yv);
      static_assert(std::is_same_v<decltype(r_ref), const IterativeSolverResult>);
      static_assert(std::is_same_v<decltype(r_mi), const Mi355xSolverResult>);
      std::printf("GMRES(%d): GenDiL %s in %zu iterations (res %.3e); binding %s in %zu iterations (res %.3e)\n",
                  (int)restart, r_ref.success ? "converged" : "failed", (size_t)r_ref.iterations, r_ref.relative_error,
                  r_mi.success ? "converged" : "failed", (size_t)r_mi.iterations, r_mi.relative_error);
      const long di = (long)r_ref.iterations - (long)r_mi.iterations;
      if (!r_ref.success || !r_mi.success || di > 2 || di < -2) ++failures;
      report("GMRES_no_alloc: binding vs GenDiL iterate", rel_l2(x_mi, x_ref, 0, nd), 1e-8);
    }
    {
      Vector x_ref(nd), x_mi(nd), r(nd), rh(nd), p(nd), v(nd), s(nd), t(nd);
      x_ref = 0.0;
      x_mi = 0.0;
      const auto r_ref = BiCGSTAB_no_alloc(K_ref, b, dot, max_it, tol, x_ref, r, rh, p, v, s, t);
      const auto r_mi = BiCGSTAB_no_alloc(K_mi, b, dot, max_it, tol, x_mi, r, rh, p, v, s, t);
      std::printf("BiCGSTAB: GenDiL %zu iterations (res %.3e); binding %zu iterations (res %.3e)\n",
                  (size_t)r_ref.iterations, r_ref.relative_error, (size_t)r_mi.iterations, r_mi.relative_error);
      // BiCGSTAB's iterates follow the rounding of its inner products (host
      // sums here, fixed-order device reductions there), so the two runs drift
      // apart over hundreds of iterations; both stop at a relative residual
      // of 1e-8, which bounds their difference by that times K's condition.
      // GenDiL's own count is not reproducible on this system: repeated runs
      // of this binary on 16 OpenMP threads stopped after 477, 515 and 614
      // iterations, while the binding's device reductions stop after 456
      // every time.  So both must converge, agree in the iterate, and the
      // binding may not need more than a fifth more iterations than GenDiL.
      const long slack = std::max<long>(2, (long)r_ref.iterations / 5);
      if (!r_ref.success || !r_mi.success || (long)r_mi.iterations > (long)r_ref.iterations + slack) ++failures;
      report("BiCGSTAB_no_alloc: binding vs GenDiL iterate", rel_l2(x_mi, x_ref, 0, nd), 1e-5);
    }
  }
  // 2) inflow data on a non-periodic Cartesian3DMesh, and a second space at an
  //    L2Restriction shift in the same vectors
  {
    Cartesian3DMesh mesh(h, n, n, n);
    auto fe = MakeLobattoFiniteElement(FiniteElementOrders<2, 2, 2>{});
    L2Restriction r0{0};
    auto fes0 = MakeFiniteElementSpace(mesh, fe, r0);
    const Integer n0 = fes0.GetNumberOfFiniteElementDofs();
    L2Restriction r1{n0};
    auto fes1 = MakeFiniteElementSpace(mesh, fe, r1);
    auto ir = MakeIntegrationRule(IntegrationRuleNumPoints<3, 3, 3>{});
    auto g = [](const std::array<Real, 3>& X) -> Real { return 1.0 + X[0] - 0.5 * X[1]; };
    SF g_d;
    g_d.terms.push_back(SF::Term{1.0, {}});
    {
      SF::Term t;
      t.factor[0] = [](double x) { return x; };
      g_d.terms.push_back(t);
    }
    {
      SF::Term t;
      t.scale = -0.5;
      t.factor[1] = [](double y) { return y; };
      g_d.terms.push_back(t);
    }
    auto yzx = [](const std::array<Real, 3>& X, Real (&v)[3]) {
      v[0] = X[1];
      v[1] = X[2];
      v[2] = X[0];
    };
    AF yzx_d;
    yzx_d.A[0][1] = yzx_d.A[1][2] = yzx_d.A[2][0] = 1.0;
    auto A_ref = MakeAdvectionOperator<SerialKernelConfiguration>(fes1, ir, yzx, g);
    auto A_mi = MakeAdvectionOperator<Mi>(fes1, ir, yzx_d, g_d);
    compare("advection, inflow g, Cartesian3DMesh, shifted space", A_ref, A_mi, 2 * n0, n0, 2 * n0, 5);
  }
  std::printf("%s\n", failures ? "FAILED" : "all binding checks passed");
  return failures ? 1 : 0;
}
