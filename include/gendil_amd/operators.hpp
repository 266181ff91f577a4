// Header-only C++20 operator API over the C-ABI, with GenDiL's call shape:
//
//   auto mesh = gendil_amd::CartesianMesh<3>{{64, 64, 64}, {h, h, h}, {0, 0, 0}, {true, true, true}};
//   auto fes  = gendil_amd::MakeFiniteElementSpace(mesh, gendil_amd::MakeLobattoFiniteElement(3));
//   auto ir   = gendil_amd::MakeIntegrationRule(4);
//   auto A    = gendil_amd::MakeAdvectionOperator(fes, ir, beta);        // advectionoperator.hpp:1416-1528
//   A(x, y);                                                              // advectionoperator.hpp:1070
//
// Factories mirror MakeMassFiniteElementOperator (massoperator.hpp:270),
// MakeAdvectionOperator with and without inflow data (advectionoperator.hpp:
// 1416-1528), MakeMassAdvectionOperator (massadvectionoperator.hpp:447) and
// MakeDiffusionOperator (diffusionoperator.hpp:519-537).
// GenDiL's coefficient lambdas cannot cross the C-ABI, so sigma, beta and g
// are descriptors: sums of products of 1D functions (tabulated here at the
// plan's sample positions), affine fields, or constants.  Vectors live on
// the device; y is overwritten (L2 semantics).  Errors throw
// gendil_amd::Error.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <random>
#include <utility>
#include <vector>

#include "gendil_amd.h"
#include "gendil_amd/krylov.hpp"

namespace gendil_amd {

// ---- discretization ------------------------------------------------------

template <int Dim>
struct CartesianMesh {
  std::array<int64_t, Dim> sizes{};
  std::array<double, Dim> h{};
  std::array<double, Dim> origin{};
  std::array<bool, Dim> periodic{};
  int64_t num_cells() const
  {
    int64_t n = 1;
    for (auto s : sizes) n *= s;
    return n;
  }
};

// [0,1]^Dim with n elements per axis (cartesianndmesh.hpp:16-80)
template <int Dim>
CartesianMesh<Dim> MakeUnitCartesianMesh(int64_t n, bool periodic)
{
  CartesianMesh<Dim> m;
  for (int d = 0; d < Dim; ++d) {
    m.sizes[d] = n;
    m.h[d] = 1.0 / (double)n;
    m.origin[d] = 0.0;
    m.periodic[d] = periodic;
  }
  return m;
}

// n cells of size h from `origin` (Cartesian1DMesh, cartesian1dmesh.hpp)
inline CartesianMesh<1> Cartesian1DMesh(double h, int64_t n, double origin = 0.0, bool periodic = false)
{
  return CartesianMesh<1>{{n}, {h}, {origin}, {periodic}};
}

// nx x ny square cells of size h (Cartesian2DMesh, cartesian2dmesh.hpp)
inline CartesianMesh<2> Cartesian2DMesh(double h, int64_t nx, int64_t ny, std::array<double, 2> origin = {0.0, 0.0},
                                        bool periodic = false)
{
  return CartesianMesh<2>{{nx, ny}, {h, h}, origin, {periodic, periodic}};
}

// nx x ny x nz cubic cells of size h (Cartesian3DMesh(h, nx, ny, nz),
// cartesian3dmesh.hpp:32-47)
inline CartesianMesh<3> Cartesian3DMesh(double h, int64_t nx, int64_t ny, int64_t nz,
                                        std::array<double, 3> origin = {0.0, 0.0, 0.0}, bool periodic = false)
{
  return CartesianMesh<3>{{nx, ny, nz}, {h, h, h}, origin, {periodic, periodic, periodic}};
}

namespace detail {
template <int D0, int D1>
CartesianMesh<D0 + D1> concat(const CartesianMesh<D0>& a, const CartesianMesh<D1>& b)
{
  CartesianMesh<D0 + D1> m;
  for (int d = 0; d < D0; ++d) {
    m.sizes[d] = a.sizes[d];
    m.h[d] = a.h[d];
    m.origin[d] = a.origin[d];
    m.periodic[d] = a.periodic[d];
  }
  for (int d = 0; d < D1; ++d) {
    m.sizes[D0 + d] = b.sizes[d];
    m.h[D0 + d] = b.h[d];
    m.origin[D0 + d] = b.origin[d];
    m.periodic[D0 + d] = b.periodic[d];
  }
  return m;
}
}  // namespace detail

// Tensor product of Cartesian meshes (MakeCartesianProductMesh,
// productmesh.hpp:154-225) as the Cartesian mesh over the concatenated axes.
// The two index alike: the product numbers cell (head, tail) as
// head + cells(head) * tail (productmesh.hpp:91-94), which unrolled over the
// factors is the first-index-fastest order of the concatenated axes; a face
// normal to an axis is the owning factor's face, with its neighbour and
// boundary (productmesh.hpp:104-134); cells are products of the factors'
// cells, so sizes, spacings, origins and periodicity concatenate.
template <int D0>
CartesianMesh<D0> MakeCartesianProductMesh(const CartesianMesh<D0>& m)
{
  return m;
}

template <int D0, int D1, int... Ds>
auto MakeCartesianProductMesh(const CartesianMesh<D0>& m0, const CartesianMesh<D1>& m1,
                              const CartesianMesh<Ds>&... rest)
{
  static_assert(D0 + D1 + (0 + ... + Ds) <= GD_MAX_DIM, "at most GD_MAX_DIM axes");
  return MakeCartesianProductMesh(detail::concat(m0, m1), rest...);
}

struct FiniteElement {
  int order;
  int basis;
};
inline FiniteElement MakeLobattoFiniteElement(int order) { return {order, GD_BASIS_GLL}; }
inline FiniteElement MakeLegendreFiniteElement(int order) { return {order, GD_BASIS_GL}; }

struct IntegrationRule {
  int num_points;
};
inline IntegrationRule MakeIntegrationRule(int num_points) { return {num_points}; }

template <int Dim>
struct FiniteElementSpace {
  CartesianMesh<Dim> mesh;
  FiniteElement fe;
  std::vector<int8_t> elem_order;  // p-adaptive: per-element order (empty = uniform)
  std::vector<int8_t> elem_quad;   // p-adaptive: per-element Gauss points
  bool padaptive() const { return !elem_order.empty(); }
};

template <int Dim>
FiniteElementSpace<Dim> MakeFiniteElementSpace(const CartesianMesh<Dim>& mesh, const FiniteElement& fe)
{
  return FiniteElementSpace<Dim>{mesh, fe, {}, {}};
}

// Per-element orders and rules (no GenDiL counterpart: GenDiL fixes orders per
// space at compile time, tensorshapefunctions.hpp:21).
template <int Dim>
FiniteElementSpace<Dim> MakePAdaptiveSpace(const CartesianMesh<Dim>& mesh, std::vector<int8_t> order,
                                           std::vector<int8_t> quad)
{
  if ((int64_t)order.size() != mesh.num_cells() || order.size() != quad.size())
    throw Error(GD_ERR_INVALID, "MakePAdaptiveSpace: one order and one rule per element");
  int8_t pmax = 0;
  for (auto p : order) pmax = p > pmax ? p : pmax;
  return FiniteElementSpace<Dim>{mesh, FiniteElement{pmax, GD_BASIS_GLL}, std::move(order), std::move(quad)};
}

// ---- coefficients --------------------------------------------------------

// f(X) = sum_r scale_r prod_d f_{r,d}(X_d); an empty factor is 1.
struct SeparableField {
  struct Term {
    double scale = 1.0;
    std::array<std::function<double(double)>, GD_MAX_DIM> factor{};
  };
  std::vector<Term> terms;

  static SeparableField Constant(double c)
  {
    SeparableField f;
    if (c != 0.0) f.terms.push_back(Term{c, {}});
    return f;
  }
  static SeparableField OneAxis(int axis, std::function<double(double)> fn, double scale = 1.0)
  {
    SeparableField f;
    Term t;
    t.scale = scale;
    t.factor[axis] = std::move(fn);
    f.terms.push_back(std::move(t));
    return f;
  }
};

template <int Dim>
struct VectorField {
  std::array<SeparableField, Dim> components;
};

// beta(X) = c + A X
template <int Dim>
struct AffineField {
  std::array<std::array<double, Dim>, Dim> A{};
  std::array<double, Dim> c{};
};

// ---- plan ----------------------------------------------------------------

class Plan {
 public:
  template <int Dim>
  Plan(const FiniteElementSpace<Dim>& fes, const IntegrationRule& ir) : dim_(Dim)
  {
    gd_desc d{};
    d.dim = Dim;
    for (int i = 0; i < Dim; ++i) {
      d.sizes[i] = fes.mesh.sizes[i];
      d.h[i] = fes.mesh.h[i];
      d.origin[i] = fes.mesh.origin[i];
      d.periodic[i] = fes.mesh.periodic[i] ? 1 : 0;
      sizes_[i] = fes.mesh.sizes[i];
    }
    d.order = fes.fe.order;
    d.num_quad = ir.num_points;
    d.basis = fes.fe.basis;
    if (fes.padaptive()) {
      d.elem_order = fes.elem_order.data();
      d.elem_quad = fes.elem_quad.data();
    }
    q_ = ir.num_points;
    check(gd_plan_create(&d, &p_), "gd_plan_create");
  }
  ~Plan()
  {
    if (p_) (void)gd_plan_destroy(p_);
  }
  Plan(const Plan&) = delete;
  Plan& operator=(const Plan&) = delete;

  gd_plan* get() const { return p_; }
  int64_t num_dofs() const { return gd_plan_num_dofs(p_); }

  void set_sigma(const SeparableField& f)
  {
    Tabulated t = tabulate(f);
    t.desc.data = t.data.empty() ? nullptr : t.data.data();
    check(gd_plan_set_sigma(p_, &t.desc), "gd_plan_set_sigma");
  }
  template <int Dim>
  void set_beta(const VectorField<Dim>& beta)
  {
    std::array<Tabulated, GD_MAX_DIM> t;
    std::array<gd_separable, GD_MAX_DIM> arr{};
    for (int d = 0; d < Dim; ++d) {
      t[d] = tabulate(beta.components[d]);
      arr[d] = t[d].desc;
      arr[d].data = t[d].data.empty() ? nullptr : t[d].data.data();
    }
    check(gd_plan_set_beta_separable(p_, arr.data()), "gd_plan_set_beta_separable");
  }
  template <int Dim>
  void set_beta(const AffineField<Dim>& beta)
  {
    std::array<double, Dim * Dim> A{};
    for (int i = 0; i < Dim; ++i)
      for (int j = 0; j < Dim; ++j) A[i * Dim + j] = beta.A[i][j];
    check(gd_plan_set_beta_affine(p_, A.data(), beta.c.data()), "gd_plan_set_beta_affine");
  }
  void set_inflow(const SeparableField& g)
  {
    Tabulated t = tabulate(g);
    t.desc.data = t.data.empty() ? nullptr : t.data.data();
    check(gd_plan_set_bc(p_, GD_BC_INFLOW, &t.desc), "gd_plan_set_bc");
  }
  void set_no_boundary_terms() { check(gd_plan_set_bc(p_, GD_BC_NONE, nullptr), "gd_plan_set_bc"); }
  void set_diffusion(const SeparableField& mu, double sigma, double kappa)
  {
    Tabulated t = tabulate(mu);
    t.desc.data = t.data.empty() ? nullptr : t.data.data();
    check(gd_plan_set_diffusion(p_, &t.desc, sigma, kappa), "gd_plan_set_diffusion");
  }

 private:
  struct Tabulated {
    gd_separable desc{};
    std::vector<double> data;
  };
  // sample every factor at the plan's per-axis positions (gd_plan_axis_points)
  Tabulated tabulate(const SeparableField& f) const
  {
    if ((int)f.terms.size() > GD_MAX_TERMS) throw Error(GD_ERR_INVALID, "at most GD_MAX_TERMS separable terms");
    Tabulated t;
    t.desc.rank = (int32_t)f.terms.size();
    for (int r = 0; r < GD_MAX_TERMS; ++r)
      for (int d = 0; d < GD_MAX_DIM; ++d) t.desc.offset[r][d] = -1;
    std::vector<double> pts;
    for (int r = 0; r < (int)f.terms.size(); ++r) {
      t.desc.scale[r] = f.terms[r].scale;
      for (int d = 0; d < dim_; ++d) {
        if (!f.terms[r].factor[d]) continue;
        pts.resize((size_t)sizes_[d] * (q_ + 2));
        check(gd_plan_axis_points(p_, d, pts.data()), "gd_plan_axis_points");
        t.desc.offset[r][d] = (int64_t)t.data.size();
        for (double x : pts) t.data.push_back(f.terms[r].factor[d](x));
      }
    }
    t.desc.data = t.data.empty() ? nullptr : t.data.data();
    t.desc.data_len = (int64_t)t.data.size();
    return t;
  }

  gd_plan* p_ = nullptr;
  int dim_;
  int q_ = 0;
  std::array<int64_t, GD_MAX_DIM> sizes_{};
};

// ---- device vectors ----------------------------------------------------------

template <typename T>
class Vector {
 public:
  explicit Vector(int64_t n) : n_(n), data_((size_t)n) {}
  int64_t Size() const { return n_; }
  const T* ReadDeviceData() const { return data_.data(); }
  T* WriteDeviceData() { return data_.data(); }
  T* ReadWriteDeviceData() { return data_.data(); }
  void FromHost(const std::vector<T>& h)
  {
    if ((int64_t)h.size() != n_) throw Error(GD_ERR_INVALID, "Vector::FromHost: size mismatch");
    check(gd_memcpy_h2d(data_.data(), h.data(), (size_t)n_ * sizeof(T)), "gd_memcpy_h2d");
  }
  std::vector<T> ToHost() const
  {
    std::vector<T> h((size_t)n_);
    check(gd_device_sync(), "gd_device_sync");
    check(gd_memcpy_d2h(h.data(), data_.data(), (size_t)n_ * sizeof(T)), "gd_memcpy_d2h");
    return h;
  }

 private:
  int64_t n_;
  DeviceArray<T> data_;
};

// GenDiL's FillRandom (vector.hpp:632-645): std::mt19937(seed) driving
// std::uniform_real_distribution<double>(min, max) over the global index.
inline std::vector<double> FillRandomHost(int64_t n, unsigned seed = 12345, double lo = -1.0, double hi = 1.0)
{
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> dist(lo, hi);
  std::vector<double> v((size_t)n);
  for (auto& e : v) e = dist(gen);
  return v;
}

// ---- operators -----------------------------------------------------------------

// y = mass_coef * M(sigma) x + adv_coef * A(beta, g) x  on device vectors.
class Operator {
 public:
  Operator(std::shared_ptr<Plan> plan, double mass_coef, double adv_coef)
      : plan_(std::move(plan)), mass_(mass_coef), adv_(adv_coef)
  {
  }
  void operator()(const double* x, double* y) const
  {
    check(gd_apply_f64(plan_->get(), mass_, adv_, x, y, stream_), "gd_apply_f64");
  }
  void operator()(const float* x, float* y) const
  {
    check(gd_apply_f32(plan_->get(), (float)mass_, (float)adv_, x, y, stream_), "gd_apply_f32");
  }
  template <typename T>
  void operator()(const Vector<T>& x, Vector<T>& y) const
  {
    (*this)(x.ReadDeviceData(), y.WriteDeviceData());
  }
  // z = (mass_coef M + adv_coef A)^T r (FP64; no GenDiL counterpart)
  void Transpose(const Vector<double>& r, Vector<double>& z) const
  {
    check(gd_apply_T_f64(plan_->get(), mass_, adv_, r.ReadDeviceData(), z.WriteDeviceData(), stream_),
          "gd_apply_T_f64");
  }
  int64_t num_dofs() const { return plan_->num_dofs(); }
  const Plan& plan() const { return *plan_; }
  void SetStream(gd_stream s) { stream_ = s; }
  CombinedOperator Combined() const { return CombinedOperator{plan_->get(), mass_, adv_, stream_}; }

 private:
  std::shared_ptr<Plan> plan_;
  double mass_, adv_;
  gd_stream stream_ = nullptr;
};

template <int Dim>
Operator MakeMassFiniteElementOperator(const FiniteElementSpace<Dim>& fes, const IntegrationRule& ir,
                                       const SeparableField& sigma = SeparableField::Constant(1.0))
{
  auto plan = std::make_shared<Plan>(fes, ir);
  plan->set_sigma(sigma);
  return Operator(std::move(plan), 1.0, 0.0);
}

// No boundary data: boundary faces are skipped (faceloop.hpp:66-72).
template <int Dim, typename Beta>
Operator MakeAdvectionOperator(const FiniteElementSpace<Dim>& fes, const IntegrationRule& ir, const Beta& beta)
{
  auto plan = std::make_shared<Plan>(fes, ir);
  plan->set_beta(beta);
  return Operator(std::move(plan), 0.0, 1.0);
}

// Inflow data g: upwind flux with u+ = g on boundary faces (advectionoperator.hpp:391-438).
template <int Dim, typename Beta>
Operator MakeAdvectionOperator(const FiniteElementSpace<Dim>& fes, const IntegrationRule& ir, const Beta& beta,
                               const SeparableField& g)
{
  auto plan = std::make_shared<Plan>(fes, ir);
  plan->set_beta(beta);
  plan->set_inflow(g);
  return Operator(std::move(plan), 0.0, 1.0);
}

// mass_coef * M(sigma) + adv_coef * A(beta) in one pass; GenDiL's fused
// operator (massadvectionoperator.hpp:447) has no boundary terms, as here
// without `g`.
template <int Dim, typename Beta>
Operator MakeMassAdvectionOperator(const FiniteElementSpace<Dim>& fes, const IntegrationRule& ir, const Beta& beta,
                                   const SeparableField& sigma, double mass_coef = 1.0, double adv_coef = 1.0)
{
  auto plan = std::make_shared<Plan>(fes, ir);
  plan->set_sigma(sigma);
  plan->set_beta(beta);
  return Operator(std::move(plan), mass_coef, adv_coef);
}

// ---- SIPDG diffusion ----------------------------------------------------------

// y = A_diff x on device vectors (DiffusionOperator::operator(),
// diffusionoperator.hpp:469-476), y overwritten.
class DiffusionOperator {
 public:
  explicit DiffusionOperator(std::shared_ptr<Plan> plan) : plan_(std::move(plan)) {}
  void operator()(const double* x, double* y) const
  {
    check(gd_diff_apply_f64(plan_->get(), x, y, stream_), "gd_diff_apply_f64");
  }
  void operator()(const float* x, float* y) const
  {
    check(gd_diff_apply_f32(plan_->get(), x, y, stream_), "gd_diff_apply_f32");
  }
  template <typename T>
  void operator()(const Vector<T>& x, Vector<T>& y) const
  {
    (*this)(x.ReadDeviceData(), y.WriteDeviceData());
  }
  int64_t num_dofs() const { return plan_->num_dofs(); }
  void SetStream(gd_stream s) { stream_ = s; }

 private:
  std::shared_ptr<Plan> plan_;
  gd_stream stream_ = nullptr;
};

// MakeDiffusionOperator(fes, int_rule, velocity, sigma, kappa)
// (diffusionoperator.hpp:519-537): `velocity` is the diffusivity, sigma the
// symmetry parameter (-1 symmetric, 0 incomplete, +1 non-symmetric interior
// penalty), kappa the penalty factor.
template <int Dim>
DiffusionOperator MakeDiffusionOperator(const FiniteElementSpace<Dim>& fes, const IntegrationRule& ir,
                                        const SeparableField& velocity, double sigma, double kappa)
{
  auto plan = std::make_shared<Plan>(fes, ir);
  plan->set_diffusion(velocity, sigma, kappa);
  return DiffusionOperator(std::move(plan));
}

// ---- face-centric operators (accumulate into y) --------------------------------

// The interior faces of a Cartesian mesh, every axis, no periodic wrap
// (MakeCartesianInteriorFaceConnectivity, cartesianglobalfaceconnectivity.hpp:89-108).
template <int Dim>
struct CartesianInteriorFaceConnectivity {
  std::array<int64_t, Dim> sizes{};
};
template <int Dim>
CartesianInteriorFaceConnectivity<Dim> MakeCartesianInteriorFaceConnectivity(const std::array<int64_t, Dim>& sizes)
{
  return {sizes};
}
// Face `Face` of the minus mesh glued to the opposite face of the plus mesh
// (cartesianintermeshfaceconnectivity.hpp:36-180); equal tangential sizes.
template <int Dim, int Face>
struct CartesianIntermeshFaceConnectivity {
  std::array<int64_t, Dim> sizes_minus{}, sizes_plus{};
};
// Coarse (minus) against a mesh with an integer number of fine cells per
// coarse cell along each tangential axis (plus)
// (nonconformingcartesianintermeshfaceconnectivity.hpp:73-243).
template <int Dim, int Face>
struct NonconformingCartesianIntermeshFaceConnectivity {
  std::array<int64_t, Dim> sizes_minus{}, sizes_plus{};
};

// y += the upwind fluxes of the face mesh, each face once: + into the minus
// cell, - into the plus cell (AdvectionLocalFaceOperator,
// advectionoperator.hpp:478-548, 596-676).  Two-space operators read and
// write the two spaces' parts of one vector, at offsets shift_minus and
// shift_plus (GenDiL's L2Restriction{shift}).
class FaceOperator {
 public:
  enum Kind { kInterior, kIntermesh, kRefined };
  FaceOperator(std::shared_ptr<Plan> minus, std::shared_ptr<Plan> plus, Kind kind, int face, int64_t shift_minus,
               int64_t shift_plus)
      : minus_(std::move(minus)), plus_(std::move(plus)), kind_(kind), face_(face), sm_(shift_minus), sp_(shift_plus)
  {
  }
  void operator()(const Vector<double>& x, Vector<double>& y) const
  {
    const double* xd = x.ReadDeviceData();
    double* yd = y.ReadWriteDeviceData();
    if (kind_ == kInterior) {
      if (x.Size() < sm_ + minus_->num_dofs() || y.Size() < sm_ + minus_->num_dofs())
        throw Error(GD_ERR_INVALID, "FaceOperator: vector too short");
      check(gd_adv_face_apply_f64(minus_->get(), xd + sm_, yd + sm_, stream_), "gd_adv_face_apply_f64");
      return;
    }
    const int64_t need = std::max(sm_ + minus_->num_dofs(), sp_ + plus_->num_dofs());
    if (x.Size() < need || y.Size() < need) throw Error(GD_ERR_INVALID, "FaceOperator: vector too short");
    if (kind_ == kIntermesh)
      check(gd_adv_intermesh_apply_f64(minus_->get(), plus_->get(), face_, xd + sm_, xd + sp_, yd + sm_, yd + sp_,
                                       stream_),
            "gd_adv_intermesh_apply_f64");
    else
      check(gd_adv_hmesh_apply_f64(minus_->get(), plus_->get(), face_, xd + sm_, xd + sp_, yd + sm_, yd + sp_,
                                   stream_),
            "gd_adv_hmesh_apply_f64");
  }
  void SetStream(gd_stream s) { stream_ = s; }

 private:
  std::shared_ptr<Plan> minus_, plus_;
  Kind kind_;
  int face_;
  int64_t sm_, sp_;
  gd_stream stream_ = nullptr;
};

// MakeAdvectionFaceOperator(fes, face_mesh, ir, adv) (advectionoperator.hpp:1546-1566)
template <int Dim, typename Beta>
FaceOperator MakeAdvectionFaceOperator(const FiniteElementSpace<Dim>& fes,
                                       const CartesianInteriorFaceConnectivity<Dim>& conn, const IntegrationRule& ir,
                                       const Beta& beta)
{
  if (conn.sizes != fes.mesh.sizes) throw Error(GD_ERR_INVALID, "face connectivity sizes differ from the mesh");
  auto plan = std::make_shared<Plan>(fes, ir);
  plan->set_beta(beta);
  return FaceOperator(plan, plan, FaceOperator::kInterior, 0, 0, 0);
}

namespace detail {
template <int Dim, typename Beta>
FaceOperator two_space_face_operator(const FiniteElementSpace<Dim>& fm, const FiniteElementSpace<Dim>& fp,
                                     const std::array<int64_t, Dim>& sm, const std::array<int64_t, Dim>& sp,
                                     FaceOperator::Kind kind, int face, const IntegrationRule& ir, const Beta& beta,
                                     int64_t shift_minus, int64_t shift_plus)
{
  if (sm != fm.mesh.sizes || sp != fp.mesh.sizes)
    throw Error(GD_ERR_INVALID, "face connectivity sizes differ from the meshes");
  auto pm = std::make_shared<Plan>(fm, ir);
  auto pp = std::make_shared<Plan>(fp, ir);
  pm->set_beta(beta);
  const int64_t sp_default = shift_plus < 0 ? pm->num_dofs() : shift_plus;
  return FaceOperator(pm, pp, kind, face, shift_minus, sp_default);
}
}  // namespace detail

// MakeAdvectionFaceOperator(fes_minus, fes_plus, face_mesh, ir, adv)
// (advectionoperator.hpp:1586-1607); by default the vector is [minus | plus].
template <int Dim, int Face, typename Beta>
FaceOperator MakeAdvectionFaceOperator(const FiniteElementSpace<Dim>& fm, const FiniteElementSpace<Dim>& fp,
                                       const CartesianIntermeshFaceConnectivity<Dim, Face>& conn,
                                       const IntegrationRule& ir, const Beta& beta, int64_t shift_minus = 0,
                                       int64_t shift_plus = -1)
{
  return detail::two_space_face_operator<Dim>(fm, fp, conn.sizes_minus, conn.sizes_plus, FaceOperator::kIntermesh, Face,
                                         ir, beta, shift_minus, shift_plus);
}
template <int Dim, int Face, typename Beta>
FaceOperator MakeAdvectionFaceOperator(const FiniteElementSpace<Dim>& fm, const FiniteElementSpace<Dim>& fp,
                                       const NonconformingCartesianIntermeshFaceConnectivity<Dim, Face>& conn,
                                       const IntegrationRule& ir, const Beta& beta, int64_t shift_minus = 0,
                                       int64_t shift_plus = -1)
{
  return detail::two_space_face_operator<Dim>(fm, fp, conn.sizes_minus, conn.sizes_plus, FaceOperator::kRefined, Face,
                                         ir, beta, shift_minus, shift_plus);
}

// ---- solvers on Operator -------------------------------------------------------

// Restarted GMRES (GMRES_no_alloc semantics, gmres.hpp:74-288); x holds x0.
inline SolveReport GMRES(const Operator& A, const Vector<double>& b, Vector<double>& x, const SolveOptions& opt = {})
{
  return gmres_solve<double>(A.Combined(), b.Size(), b.ReadDeviceData(), x.ReadWriteDeviceData(), opt);
}

// FP32 operator apply and basis inside each cycle, FP64 refinement.
inline SolveReport GMRESMixed(const Operator& A, const Vector<double>& b, Vector<double>& x,
                              const SolveOptions& opt = {})
{
  return gmres_solve<float>(A.Combined(), b.Size(), b.ReadDeviceData(), x.ReadWriteDeviceData(), opt);
}

// Unpreconditioned CG (conjugategradient.hpp:217) for SPD operators.
inline SolveReport ConjugateGradient(const Operator& A, const Vector<double>& b, Vector<double>& x,
                                     const SolveOptions& opt = {})
{
  return cg_solve(A.Combined(), b.Size(), b.ReadDeviceData(), x.ReadWriteDeviceData(), opt);
}

}  // namespace gendil_amd
