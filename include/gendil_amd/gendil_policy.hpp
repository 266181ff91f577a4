// GenDiL-side binding of the MI355X operators: the header a GenDiL maintainer
// adds so that GenDiL's own call sites select the gfx950 kernels by naming a
// kernel policy, e.g.
//
//   auto A = MakeAdvectionOperator< Mi355xKernelConfiguration >( fes, ir, beta );
//   A( x, y );                                      // y overwritten (L2 semantics)
//   GMRES_no_alloc( A, b, dot, max_it, m, tol, x, V, w, H, cs, sn, e1, yv );
//
// It is included after <gendil/gendil.hpp> and compiles against GenDiL's own
// types: the spaces come from MakeFiniteElementSpace over CartesianMesh<D> or
// Cartesian{1,2,3}DMesh (cartesianndmesh.hpp:16-80, cartesian3dmesh.hpp:22-77)
// with uniform-order MakeLobattoFiniteElement / MakeLegendreFiniteElement
// (finiteelement.hpp:103-109), the rules from MakeIntegrationRule over
// IntegrationRuleNumPoints (integrationrule.hpp:80), vectors are gendil::Vector
// (vector.hpp:84), and an L2Restriction shift (restriction.hpp:38) is honoured.
//
// GenDiL's coefficients are device lambdas, which cannot cross the C-ABI.  The
// overloads here take the descriptor types of gendil_amd/operators.hpp BY
// VALUE (VectorField / AffineField for beta, SeparableField for sigma and g),
// so they are more specialised than GenDiL's `Adv&&` / `Sigma&` templates
// (advectionoperator.hpp:1416-1528, massoperator.hpp:270,
// massadvectionoperator.hpp:447) and win overload resolution for every value
// category, while a lambda argument still reaches GenDiL's own operators.
//
// Vector residency: with GENDIL_USE_DEVICE (a HIP build of GenDiL) the
// operators take Vector::ReadDeviceData() / WriteDeviceData() directly; in a
// host-only build of GenDiL they stage the vectors through device buffers of
// their own (PCIe copies around every apply).
#pragma once

#include <gendil/gendil.hpp>

#include <array>
#include <memory>
#include <tuple>
#include <type_traits>

#include "gendil_amd.h"
#include "gendil_amd/krylov.hpp"
#include "gendil_amd/operators.hpp"

namespace gendil {

/// Kernel policy tag that selects the MI355X (gfx950) operators.
struct Mi355xKernelConfiguration {};

namespace mi355x_detail {

// ---- what a GenDiL space and rule are, read from their types ----------------

template <typename Orders>
struct uniform_order;
template <Integer O0, Integer... Os>
struct uniform_order<FiniteElementOrders<O0, Os...>> {
  static constexpr bool value = ((Os == O0) && ...);
  static constexpr int order = (int)O0;
};

template <typename FES>
struct space_traits {
  using shape = typename FES::finite_element_type::shape_functions;
  using sf0 = std::tuple_element_t<0, typename shape::shape_functions_1d_tuple>;
  static constexpr int dim = (int)shape::dim;
  static constexpr int order = uniform_order<typename shape::orders>::order;
  static_assert(uniform_order<typename shape::orders>::value,
                "Mi355xKernelConfiguration: one polynomial order on every axis");
  static constexpr bool lobatto = std::is_same_v<sf0, GaussLobattoLegendreShapeFunctions<(Integer)order>>;
  static constexpr bool legendre = std::is_same_v<sf0, GaussLegendreShapeFunctions<(Integer)order>>;
  static_assert(lobatto || legendre, "Mi355xKernelConfiguration: GLL or GL Lagrange elements");
};

template <typename Points>
struct gauss_points;
template <Integer N0, Integer... Ns>
struct gauss_points<TensorPoints<GaussLegendrePoints<N0>, GaussLegendrePoints<Ns>...>> {
  static constexpr bool uniform = ((Ns == N0) && ...);
  static constexpr int num = (int)N0;
  static constexpr int dim = 1 + (int)sizeof...(Ns);
};

template <typename IR>
struct rule_traits {
  using pts = gauss_points<typename IR::points>;
  static_assert(pts::uniform, "Mi355xKernelConfiguration: one Gauss rule on every axis");
  static constexpr int num_points = pts::num;
  static constexpr int dim = pts::dim;
};

// ---- the Cartesian mesh a GenDiL space is built on ------------------------

template <Integer Dim>
gendil_amd::CartesianMesh<(int)Dim> describe(const CartesianMesh<Dim>& m)
{
  gendil_amd::CartesianMesh<(int)Dim> o;
  for (int d = 0; d < (int)Dim; ++d) {
    o.sizes[d] = (int64_t)m.sizes[d];
    o.h[d] = m.h[d];
    o.origin[d] = m.mesh_origin[d];
    o.periodic[d] = m.is_periodic_boundary[d];
  }
  return o;
}

template <typename Conn>
constexpr bool periodic_connectivity = false;
template <Integer D>
constexpr bool periodic_connectivity<PeriodicCartesianConnectivity<D>> = true;

template <typename Conn>
gendil_amd::CartesianMesh<3> describe(const Cartesian3DMeshBase<Conn>& m)
{
  gendil_amd::CartesianMesh<3> o;
  const double h[3] = {m.h_x, m.h_y, m.h_z};
  for (int d = 0; d < 3; ++d) {
    o.sizes[d] = (int64_t)m.connectivity.sizes[d];
    o.h[d] = h[d];
    o.origin[d] = m.mesh_origin[d];
    o.periodic[d] = periodic_connectivity<Conn>;
  }
  return o;
}

template <typename Conn>
gendil_amd::CartesianMesh<2> describe(const Cartesian2DMeshBase<Conn>& m)
{
  gendil_amd::CartesianMesh<2> o;
  const double h[2] = {m.h_x, m.h_y};
  for (int d = 0; d < 2; ++d) {
    o.sizes[d] = (int64_t)m.connectivity.sizes[d];
    o.h[d] = h[d];
    o.origin[d] = m.mesh_origin[d];
    o.periodic[d] = periodic_connectivity<Conn>;
  }
  return o;
}

template <typename FES>
auto make_space(const FES& fes)
{
  using T = space_traits<FES>;
  const auto mesh = describe(static_cast<const typename FES::mesh_type&>(fes));
  static_assert(std::tuple_size_v<decltype(mesh.sizes)> == (size_t)T::dim, "mesh and element dimensions differ");
  const auto fe = T::lobatto ? gendil_amd::MakeLobattoFiniteElement(T::order)
                             : gendil_amd::MakeLegendreFiniteElement(T::order);
  return gendil_amd::MakeFiniteElementSpace(mesh, fe);
}

template <typename FES>
int64_t shift_of(const FES& fes)
{
  static_assert(std::is_same_v<typename FES::restriction_type, L2Restriction>,
                "Mi355xKernelConfiguration: L2 (discontinuous) spaces only");
  return (int64_t)fes.restriction.shift;
}

}  // namespace mi355x_detail

/// y = mass_coef * M(sigma) x + adv_coef * A(beta, g) x on a gfx950 plan, with
/// GenDiL's operator call shape: operator()(const Vector&, Vector&) overwrites
/// the space's part of `out` (massoperator.hpp:209-216,
/// advectionoperator.hpp:1070-1076).  Moves, does not copy (it owns the plan).
class Mi355xOperator {
 public:
  template <typename FES, typename IR>
  Mi355xOperator(const FES& fes, const IR&, double mass_coef, double adv_coef, bool diffusion = false)
      : mass_coef_(mass_coef), adv_coef_(adv_coef), diffusion_(diffusion), shift_(mi355x_detail::shift_of(fes))
  {
    static_assert(mi355x_detail::space_traits<FES>::dim == mi355x_detail::rule_traits<IR>::dim,
                  "space and rule dimensions differ");
    plan_ = std::make_unique<gendil_amd::Plan>(
        mi355x_detail::make_space(fes), gendil_amd::MakeIntegrationRule(mi355x_detail::rule_traits<IR>::num_points));
    n_ = plan_->num_dofs();
  }

  gendil_amd::Plan& plan() const { return *plan_; }
  double mass_coef() const { return mass_coef_; }
  double adv_coef() const { return adv_coef_; }
  /// The SIPDG diffusion operator rather than mass_coef M + adv_coef A.
  bool diffusion() const { return diffusion_; }
  int64_t num_dofs() const { return n_; }
  int64_t shift() const { return shift_; }
  gd_stream stream() const { return stream_; }
  /// Launch stream for later applies and solves (nullptr = default stream).
  void set_stream(gd_stream s) { stream_ = s; }

  void operator()(const Vector& in, Vector& out) const
  {
    check_size(in);
    check_size(out);
#ifdef GENDIL_USE_DEVICE
    apply(in.ReadDeviceData() + shift_, out.WriteDeviceData() + shift_);
#else
    // GenDiL built without a device: its vectors live in host memory
    stage();
    to_device(in, dx_);
    apply(dx_.data(), dy_.data());
    to_host(dy_, out);
#endif
  }

  /// Runs `solve(b, x)` on device pointers to the space's parts of rhs and x
  /// (staged through device buffers when GenDiL's vectors are host-resident).
  template <typename F>
  void with_device(const Vector& rhs, Vector& x, F&& solve) const
  {
    check_size(rhs);
    check_size(x);
#ifdef GENDIL_USE_DEVICE
    solve(rhs.ReadDeviceData() + shift_, x.ReadWriteDeviceData() + shift_);
#else
    stage();
    to_device(rhs, dx_);
    to_device(x, dy_);
    solve(dx_.data(), dy_.data());
    to_host(dy_, x);
#endif
  }

 private:
  void apply(const double* x, double* y) const
  {
    if (diffusion_) gendil_amd::check(gd_diff_apply_f64(plan_->get(), x, y, stream_), "gd_diff_apply_f64");
    else gendil_amd::check(gd_apply_f64(plan_->get(), mass_coef_, adv_coef_, x, y, stream_), "gd_apply_f64");
  }
  void check_size(const Vector& v) const
  {
    if ((int64_t)v.Size() < shift_ + n_)
      throw gendil_amd::Error(GD_ERR_INVALID, "Mi355xOperator: vector too short for the space");
  }
#ifndef GENDIL_USE_DEVICE
  void stage() const
  {
    if (!dx_.data()) {
      dx_ = gendil_amd::DeviceArray<double>((size_t)n_);
      dy_ = gendil_amd::DeviceArray<double>((size_t)n_);
    }
  }
  void to_device(const Vector& v, gendil_amd::DeviceArray<double>& d) const
  {
    gendil_amd::check(gd_stream_sync(stream_), "gd_stream_sync");
    gendil_amd::check(gd_memcpy_h2d(d.data(), v.ReadHostData() + shift_, (size_t)n_ * sizeof(double)),
                      "gd_memcpy_h2d");
  }
  void to_host(const gendil_amd::DeviceArray<double>& d, Vector& v) const
  {
    gendil_amd::check(gd_stream_sync(stream_), "gd_stream_sync");
    gendil_amd::check(gd_memcpy_d2h(v.ReadWriteHostData() + shift_, d.data(), (size_t)n_ * sizeof(double)),
                      "gd_memcpy_d2h");
  }
#endif

  std::unique_ptr<gendil_amd::Plan> plan_;
  double mass_coef_, adv_coef_;
  bool diffusion_ = false;
  int64_t shift_ = 0, n_ = 0;
  gd_stream stream_ = nullptr;
  mutable gendil_amd::DeviceArray<double> dx_, dy_;
};

// ---- factories: GenDiL's names, selected by the policy tag -----------------

/// MakeMassFiniteElementOperator<Mi355xKernelConfiguration>(fes, ir, sigma)
/// (massoperator.hpp:270-287).
template <typename KernelPolicy, typename FES, typename IR>
  requires std::is_same_v<KernelPolicy, Mi355xKernelConfiguration>
Mi355xOperator MakeMassFiniteElementOperator(const FES& fes, const IR& ir, gendil_amd::SeparableField sigma)
{
  Mi355xOperator op(fes, ir, 1.0, 0.0);
  op.plan().set_sigma(sigma);
  return op;
}

/// MakeAdvectionOperator<Mi355xKernelConfiguration>(fes, ir, adv): no boundary
/// data, boundary faces skipped (advectionoperator.hpp:1411-1428).
template <typename KernelPolicy, typename FES, typename IR, int Dim>
  requires std::is_same_v<KernelPolicy, Mi355xKernelConfiguration>
Mi355xOperator MakeAdvectionOperator(const FES& fes, const IR& ir, gendil_amd::VectorField<Dim> adv)
{
  Mi355xOperator op(fes, ir, 0.0, 1.0);
  op.plan().set_beta(adv);
  op.plan().set_no_boundary_terms();
  return op;
}

template <typename KernelPolicy, typename FES, typename IR, int Dim>
  requires std::is_same_v<KernelPolicy, Mi355xKernelConfiguration>
Mi355xOperator MakeAdvectionOperator(const FES& fes, const IR& ir, gendil_amd::AffineField<Dim> adv)
{
  Mi355xOperator op(fes, ir, 0.0, 1.0);
  op.plan().set_beta(adv);
  op.plan().set_no_boundary_terms();
  return op;
}

/// MakeAdvectionOperator<Mi355xKernelConfiguration>(fes, ir, adv, g): upwind
/// inflow data g on the boundary (advectionoperator.hpp:1471-1493).
template <typename KernelPolicy, typename FES, typename IR, int Dim>
  requires std::is_same_v<KernelPolicy, Mi355xKernelConfiguration>
Mi355xOperator MakeAdvectionOperator(const FES& fes, const IR& ir, gendil_amd::VectorField<Dim> adv,
                                     gendil_amd::SeparableField g)
{
  Mi355xOperator op(fes, ir, 0.0, 1.0);
  op.plan().set_beta(adv);
  op.plan().set_inflow(g);
  return op;
}

template <typename KernelPolicy, typename FES, typename IR, int Dim>
  requires std::is_same_v<KernelPolicy, Mi355xKernelConfiguration>
Mi355xOperator MakeAdvectionOperator(const FES& fes, const IR& ir, gendil_amd::AffineField<Dim> adv,
                                     gendil_amd::SeparableField g)
{
  Mi355xOperator op(fes, ir, 0.0, 1.0);
  op.plan().set_beta(adv);
  op.plan().set_inflow(g);
  return op;
}

// With the policy tag but a coefficient that is not a descriptor (a lambda),
// GenDiL's own templates would build GenDiL's operator on a policy it cannot
// launch.  These overloads have the same parameters as GenDiL's and are more
// constrained, so they are chosen instead, and they stop the build.
template <typename T>
constexpr bool mi355x_descriptor = false;
template <int Dim>
constexpr bool mi355x_descriptor<gendil_amd::VectorField<Dim>> = true;
template <int Dim>
constexpr bool mi355x_descriptor<gendil_amd::AffineField<Dim>> = true;
template <>
constexpr bool mi355x_descriptor<gendil_amd::SeparableField> = true;
template <typename... T>
constexpr bool mi355x_all_descriptors = (mi355x_descriptor<std::remove_cvref_t<T>> && ...);

template <typename KernelPolicy, typename FES, typename IR, typename Adv>
  requires(std::is_same_v<KernelPolicy, Mi355xKernelConfiguration> && !mi355x_all_descriptors<Adv>)
auto MakeAdvectionOperator(const FES&, const IR&, Adv&&)
{
  static_assert(mi355x_all_descriptors<Adv>,
                "Mi355xKernelConfiguration takes beta as a gendil_amd::VectorField or AffineField descriptor: "
                "device lambdas cannot cross the C-ABI");
  return 0;
}

template <typename KernelPolicy, typename FES, typename IR, typename Adv, typename BCType>
  requires(std::is_same_v<KernelPolicy, Mi355xKernelConfiguration> && !mi355x_all_descriptors<Adv, BCType>)
auto MakeAdvectionOperator(const FES&, const IR&, Adv&&, BCType&&)
{
  static_assert(mi355x_all_descriptors<Adv, BCType>,
                "Mi355xKernelConfiguration takes beta and g as gendil_amd descriptors (VectorField / AffineField, "
                "SeparableField): device lambdas cannot cross the C-ABI");
  return 0;
}

template <typename KernelPolicy, typename FES, typename IR, typename Sigma>
  requires(std::is_same_v<KernelPolicy, Mi355xKernelConfiguration> && !mi355x_all_descriptors<Sigma>)
auto MakeMassFiniteElementOperator(const FES&, const IR&, Sigma&)
{
  static_assert(mi355x_all_descriptors<Sigma>,
                "Mi355xKernelConfiguration takes sigma as a gendil_amd::SeparableField descriptor: device lambdas "
                "cannot cross the C-ABI");
  return 0;
}

template <typename KernelPolicy, typename FES, typename IR, typename Adv, typename Sigma>
  requires(std::is_same_v<KernelPolicy, Mi355xKernelConfiguration> && !mi355x_all_descriptors<Adv, Sigma>)
auto MakeMassAdvectionOperator(const FES&, const IR&, Adv&, Sigma&)
{
  static_assert(mi355x_all_descriptors<Adv, Sigma>,
                "Mi355xKernelConfiguration takes beta and sigma as gendil_amd descriptors: device lambdas cannot "
                "cross the C-ABI");
  return 0;
}

/// MakeMassAdvectionOperator<Mi355xKernelConfiguration>(fes, ir, adv, sigma):
/// sigma M + A in one pass, no boundary terms (massadvectionoperator.hpp:
/// 439-461, whose single-lambda face loop skips boundary faces).
template <typename KernelPolicy, typename FES, typename IR, int Dim>
  requires std::is_same_v<KernelPolicy, Mi355xKernelConfiguration>
Mi355xOperator MakeMassAdvectionOperator(const FES& fes, const IR& ir, gendil_amd::VectorField<Dim> adv,
                                         gendil_amd::SeparableField sigma)
{
  Mi355xOperator op(fes, ir, 1.0, 1.0);
  op.plan().set_beta(adv);
  op.plan().set_sigma(sigma);
  op.plan().set_no_boundary_terms();
  return op;
}

/// MakeDiffusionOperator<Mi355xKernelConfiguration>(fes, ir, velocity, sigma,
/// kappa) (diffusionoperator.hpp:519-537): SIPDG with diffusivity `velocity`,
/// symmetry parameter sigma and penalty kappa.
template <typename KernelPolicy, typename FES, typename IR>
  requires std::is_same_v<KernelPolicy, Mi355xKernelConfiguration>
Mi355xOperator MakeDiffusionOperator(const FES& fes, const IR& ir, gendil_amd::SeparableField velocity,
                                     const Real sigma, const Real kappa)
{
  Mi355xOperator op(fes, ir, 0.0, 0.0, true);
  op.plan().set_diffusion(velocity, sigma, kappa);
  return op;
}

template <typename KernelPolicy, typename FES, typename IR, typename Velocity>
  requires(std::is_same_v<KernelPolicy, Mi355xKernelConfiguration> && !mi355x_all_descriptors<Velocity>)
auto MakeDiffusionOperator(const FES&, const IR&, Velocity&, const Real, const Real)
{
  static_assert(mi355x_all_descriptors<Velocity>,
                "Mi355xKernelConfiguration takes the diffusivity as a gendil_amd::SeparableField descriptor: device "
                "lambdas cannot cross the C-ABI");
  return 0;
}

// ---- Krylov solvers: GenDiL's signatures, device-resident loops ------------
// GenDiL's solvers are templates on the operator; these overloads, more
// specialised in the operator, run the library's device-resident loops (no
// host dot products) on the operator's plan.  The host workspaces GenDiL's
// signatures carry are accepted and left unused; `dot` is not called (the
// inner products are FP64 device reductions).  The result is GenDiL's
// IterativeSolverResult under a type of its own (no members added, so
// structured bindings and conversions to IterativeSolverResult work as for
// GenDiL's solvers), which lets callers see which loop ran.
struct Mi355xSolverResult : IterativeSolverResult {};

namespace mi355x_detail {
// the device loops apply mass_coef M + adv_coef A of the plan
inline void require_mass_advection(const Mi355xOperator& A, const char* solver)
{
  if (A.diffusion())
    throw gendil_amd::Error(GD_ERR_UNSUPPORTED, std::string(solver) +
                                                    ": the device loops run the mass/advection apply, not diffusion");
}
}  // namespace mi355x_detail

/// GMRES_no_alloc (gmres.hpp:66-288): restarted GMRES(restart), all FP64.
template <typename RHSType, typename DotType, typename VectorArray, typename VectorT, typename Array>
Mi355xSolverResult GMRES_no_alloc(const Mi355xOperator& A, const RHSType& rhs, const DotType&, const Integer max_iters,
                                     const Integer restart, const Real tol, VectorT& x, VectorArray&, VectorT&,
                                     Array&, Array&, Array&, Array&, Array&)
{
  mi355x_detail::require_mass_advection(A, "GMRES_no_alloc");
  gd_solve_params prm{(int32_t)restart, (int32_t)max_iters, tol, 100000, 0.0};
  gd_solve_result res{};
  A.with_device(rhs, x, [&](const double* b, double* xd) {
    const int rc = gd_gmres_f64(A.plan().get(), A.mass_coef(), A.adv_coef(), b, xd, &prm, &res, A.stream());
    if (rc != GD_OK && rc != GD_ERR_NOT_CONVERGED) gendil_amd::check(rc, "gd_gmres_f64");
  });
  return Mi355xSolverResult{{res.success != 0, (Integer)res.iterations, res.final_rel_residual}};
}

/// BiCGSTAB_no_alloc (bicgstab.hpp:44-157): the reference's stopping rule,
/// early exit, fallback and breakdown guards, device-resident.
template <typename RHSType, typename Dot, typename VectorT>
Mi355xSolverResult BiCGSTAB_no_alloc(const Mi355xOperator& A, const RHSType& rhs, const Dot&, Integer max_iter,
                                        Real tol, VectorT& x, VectorT&, VectorT&, VectorT&, VectorT&, VectorT&,
                                        VectorT&)
{
  mi355x_detail::require_mass_advection(A, "BiCGSTAB_no_alloc");
  gd_solve_params prm{30, (int32_t)max_iter, tol, 1, 0.0};
  gd_solve_result res{};
  A.with_device(rhs, x, [&](const double* b, double* xd) {
    const int rc = gd_bicgstab_f64(A.plan().get(), A.mass_coef(), A.adv_coef(), b, xd, &prm, &res, A.stream());
    if (rc != GD_OK && rc != GD_ERR_NOT_CONVERGED) gendil_amd::check(rc, "gd_bicgstab_f64");
  });
  return Mi355xSolverResult{{res.success != 0, (Integer)res.iterations, res.final_rel_residual}};
}

}  // namespace gendil
